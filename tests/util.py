"""Comparison helpers shared by the tests."""
import numpy as np


def interior(a, ng=2):
    return a[ng:-ng, ng:-ng, :]


def bitwise_equal(a, b):
    """Bitwise equality, with any NaN equal to any NaN (NaN sign/payload is not portable)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    same_bits = a.view(np.uint64) == b.view(np.uint64)
    return bool(np.all((na & nb) | (~na & ~nb & same_bits)))


def field_errors(a, b):
    """Per-component relative L1 and Linf errors of a against b (interior arrays)."""
    out = []
    for k in range(a.shape[-1]):
        x, y = a[..., k], b[..., k]
        den1 = np.sum(np.abs(y))
        den_inf = np.max(np.abs(y))
        l1 = np.sum(np.abs(x - y)) / den1 if den1 > 0 else float(np.sum(np.abs(x - y)))
        li = np.max(np.abs(x - y)) / den_inf if den_inf > 0 else float(np.max(np.abs(x - y)))
        out.append((float(l1), float(li)))
    return out


def n_bit_diffs(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    na, nb = np.isnan(a), np.isnan(b)
    same = (na & nb) | (~na & ~nb & (a.view(np.uint64) == b.view(np.uint64)))
    return int(np.size(same) - np.count_nonzero(same))
