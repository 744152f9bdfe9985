"""Shared helpers of the test suite: building the same graph for the engine
and the oracle and comparing their results."""
import numpy as np

import fastlbp_amd as fl
import oracle


def make_list(type_, head, body_off, body, p, q, clamped, m0, m1):
    return fl.FactorList(type_, head, body_off, body, p, q, clamped, m0, m1)


def oracle_clamp(flist, labels, values):
    return oracle.clamp(flist, labels, values, make_list)


def same_bits(a, b):
    """Bit-identical doubles; NaNs must sit at the same positions but their
    payload and sign may differ (x86 produces a negative NaN for 0/0, the
    GPU a positive one; no later arithmetic can tell them apart)."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a[~na].view(np.uint64), b[~nb].view(np.uint64))


def random_alarm_clamps(rng, alarms, max_n):
    n = int(rng.integers(1, max_n + 1))
    idx = rng.choice(len(alarms), size=min(n, len(alarms)), replace=False)
    return alarms[idx], rng.integers(0, 2, size=idx.shape[0]).astype(np.uint8)
