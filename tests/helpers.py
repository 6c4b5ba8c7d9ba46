"""Shared comparison helpers for the parity tests."""
import numpy as np


def rel_err(a, b, floor):
    """max |a-b| / max(|b|, floor), elementwise."""
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def row_norm_err(a, b, n):
    """Largest per-row normwise relative error of a batch of column-major n x n blocks: for each block and row i,
    max_j |a_ij - b_ij| / max_j |b_ij|."""
    A = np.asarray(a).reshape(-1, n, n)   # [block, col, row]
    B = np.asarray(b).reshape(-1, n, n)
    num = np.max(np.abs(A - B), axis=1)
    den = np.max(np.abs(B), axis=1)
    return float(np.max(num / np.maximum(den, 1e-300)))
