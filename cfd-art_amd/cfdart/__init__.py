"""Python bindings (ctypes) of libcfdart.so, the MI355X MUSCL-Hancock sweep library.

The product is the C-ABI library (include/cfdart.h, include/cfdart_host.h) and its C++ driver;
this module is the thin plumbing that tests and the benchmark use to call it. It never falls
back to a CPU implementation: if the library or a GPU is missing, calls raise.
"""
from ._lib import (  # noqa: F401
    CfdartError, Desc, DiscParams, BubbleParams, Grid, Batch, time_tile, lib, library_path,
    EULER, ALLAIRE, TRANSMISSIVE, REFLECTIVE, MINMOD, VANLEER, GARRICK, PVRS,
    NAIVE, STRANG, ALTERNATING, MUSCL_HANCOCK, GODUNOV, Y_TRANSPOSED, Y_TRANSPOSED_NARROW, ARITH_EXACT, ARITH_FAST, ncomp, make_desc, random_discs, shocked_bubbles, planar,
    prim_to_cons, device_count, state_checksum, test_case,
)
from . import autotune, configs, serve  # noqa: F401
