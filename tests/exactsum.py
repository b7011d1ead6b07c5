"""Exact residual evaluation for the parity tests (checker only): every
product of two FP64 numbers is written as the exact sum of four FP64 products
of Dekker halves, and each row's terms are added by math.fsum, which returns
the correctly rounded sum."""
import math

import numpy as np


def _split(a):
    """Dekker's split: a = hi + lo with 26-bit halves, so that products of two
    halves are exact in FP64."""
    c = a * 134217729.0  # 2^27 + 1
    hi = c - (c - a)
    return hi, a - hi


def exact_residual(A, b, x, xl=None):
    """||b - A (x + xl)|| / ||b|| from the oracle CSC (xl optional), each row's
    residual rounded once from its exact value: every product a_ij x_j is the
    exact sum of four FP64 products of split halves, and math.fsum adds a
    row's terms exactly."""
    import math
    cp, ri, va = A.export()
    cols = np.repeat(np.arange(len(cp) - 1), np.diff(cp.astype(np.int64)))
    ah, al = _split(va)
    parts = []
    for xv in (x, xl) if xl is not None else (x,):
        h, l = _split(xv[cols])
        parts += [-(ah * h), -(ah * l), -(al * h), -(al * l)]
    terms = np.stack(parts, axis=1)
    order = np.argsort(ri, kind="stable")
    rows = ri.astype(np.int64)[order]
    terms = terms[order]
    starts = np.searchsorted(rows, np.arange(len(b)))
    ends = np.searchsorted(rows, np.arange(len(b)), side="right")
    r = np.array([math.fsum([b[i]] + list(terms[starts[i]:ends[i]].ravel()))
                  for i in range(len(b))])
    rl = r.astype(np.longdouble)
    return float(np.sqrt((rl * rl).sum()) / np.sqrt((b.astype(np.longdouble) ** 2).sum()))
