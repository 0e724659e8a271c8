"""Helpers shared by the GPU parity tests."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_tbin(name):
    return np.fromfile(os.path.join(GOLDEN, name), dtype=np.uint32)


def g1_fixture():
    s = load_tbin("test_scmul_samples.tbin").reshape(-1, 8)
    n = s.shape[0] // 3
    return s[:n].copy(), s[n:].reshape(n, 16).copy(), load_tbin("test_scmul_reduction_soln.tbin")


def g2_fixture():
    s = load_tbin("test2_scmul_samples.tbin").reshape(-1, 8)
    n = s.shape[0] // 5
    return s[:n].copy(), s[n:].reshape(n, 32).copy(), load_tbin("test2_scmul_reduction_soln.tbin")
