"""Shared helpers of the test suite (comparison metrics, level runners)."""
import numpy as np


def maxrel_per_var(a, b):
    """Per-variable max |a-b| / max|b| (the variable's own scale), over the leading axis."""
    out = []
    for n in range(a.shape[0]):
        scale = np.max(np.abs(b[n]))
        if scale == 0.0:
            scale = 1.0
        out.append(float(np.max(np.abs(a[n] - b[n])) / scale))
    return np.array(out)


def l1rel_per_var(a, b):
    out = []
    for n in range(a.shape[0]):
        den = np.sum(np.abs(b[n]))
        if den == 0.0:
            den = 1.0
        out.append(float(np.sum(np.abs(a[n] - b[n])) / den))
    return np.array(out)


def cellrel_per_var(a, b, floor=1e-8):
    """Per-variable max over cells of |a-b| / max(|b|, floor * max|b|): the relative difference of
    every cell against its own value, with cells below `floor` of the variable's scale measured
    against that floor (a variable that crosses zero has no cell-relative error at the crossing)."""
    out = []
    for n in range(a.shape[0]):
        scale = np.max(np.abs(b[n]))
        if scale == 0.0:
            scale = 1.0
        den = np.maximum(np.abs(b[n]), floor * scale)
        out.append(float(np.max(np.abs(a[n] - b[n]) / den)))
    return np.array(out)
