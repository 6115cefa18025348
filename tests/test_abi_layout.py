"""CPU: the ctypes plumbing agrees with the C headers. A small C program compiled against
include/*.h prints every struct's size and field offsets and the enumerators and selector
constants; the ctypes mirrors in cfdart/_lib.py and cfdart/serve.py must match them exactly
(a field appended to a header without its mirror would shift every later field)."""
import ctypes as C
import os
import subprocess
import tempfile

import pytest

import cfdart
from cfdart import serve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "cfdart_desc": cfdart.Desc,
    "cfdart_disc_params": cfdart.DiscParams,
    "cfdart_bubble_params": cfdart.BubbleParams,
    "cfdart_request": serve.Request,
    "cfdart_serve_opts": serve.ServeOpts,
    "cfdart_request_result": serve.RequestResult,
    "cfdart_serve_summary": serve.ServeSummary,
}

CONSTANTS = {
    "CFDART_EULER": cfdart.EULER, "CFDART_ALLAIRE": cfdart.ALLAIRE,
    "CFDART_TRANSMISSIVE": cfdart.TRANSMISSIVE, "CFDART_REFLECTIVE": cfdart.REFLECTIVE,
    "CFDART_MINMOD": cfdart.MINMOD, "CFDART_VANLEER": cfdart.VANLEER,
    "CFDART_GARRICK": cfdart.GARRICK, "CFDART_PVRS": cfdart.PVRS,
    "CFDART_NAIVE": cfdart.NAIVE, "CFDART_STRANG": cfdart.STRANG, "CFDART_ALTERNATING": cfdart.ALTERNATING,
    "CFDART_MUSCL_HANCOCK": cfdart.MUSCL_HANCOCK, "CFDART_GODUNOV": cfdart.GODUNOV,
    "CFDART_Y_TRANSPOSED": cfdart.Y_TRANSPOSED, "CFDART_Y_TRANSPOSED_NARROW": cfdart.Y_TRANSPOSED_NARROW,
    "CFDART_ARITH_EXACT": cfdart.ARITH_EXACT, "CFDART_ARITH_FAST": cfdart.ARITH_FAST,
}


@pytest.fixture(scope="module")
def layout():
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cfdart.h"', '#include "cfdart_host.h"',
             '#include "cfdart_batch.h"', '#include "cfdart_serve.h"', 'int main(void) {']
    for cname, py in STRUCTS.items():
        lines.append(f'    printf("size {cname} %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'    printf("off {cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    for name in CONSTANTS:
        lines.append(f'    printf("const {name} %d\\n", (int)({name}));')
    lines += ['    return 0;', '}']
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        with open(src, "w") as f:
            f.write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-std=c11", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return [ln.split() for ln in out.splitlines()]


def test_struct_sizes_and_offsets(layout):
    for row in layout:
        if row[0] == "size":
            assert C.sizeof(STRUCTS[row[1]]) == int(row[2]), row
        elif row[0] == "off":
            assert getattr(STRUCTS[row[1]], row[2]).offset == int(row[3]), row


def test_constants(layout):
    seen = {row[1]: int(row[2]) for row in layout if row[0] == "const"}
    assert seen == CONSTANTS
