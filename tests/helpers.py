"""Shared helpers of the test-suite (norm-wise relative errors, small seeded problems)."""
import numpy as np


def rel_err(a, b):
    """‖a − b‖∞ / ‖b‖∞ (norm-wise; entries near zero do not blow the ratio up)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def per_item_rel_err(a, b):
    """max over the leading axis of ‖a_i − b_i‖∞ / ‖b_i‖∞."""
    a = np.asarray(a, dtype=np.float64).reshape(a.shape[0], -1)
    b = np.asarray(b, dtype=np.float64).reshape(b.shape[0], -1)
    den = np.maximum(np.abs(b).max(axis=1), 1e-300)
    return float((np.abs(a - b).max(axis=1) / den).max())
