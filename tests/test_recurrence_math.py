"""The product's FP64 per-pair algebra (fastfuncstuff_amd/csrc/reml_math.h),
compiled for the host, against the dense-Cholesky oracle.

The innovations recurrence must reproduce L^-1 (and L^-T) of the dense
run-block-diagonal Cholesky factor and its log-determinant to FP64 rounding,
with and without censor gaps and run restarts (the reference's own invariant,
tests/test_glm_whiten.py:30-50, is whitener == dense L^-1 to 1e-10)."""

import ctypes
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import reml_oracle as O

ROOT = Path(__file__).resolve().parents[1]
F64 = torch.float64


@pytest.fixture(scope="module")
def rm(tmp_path_factory):
    out = tmp_path_factory.mktemp("rm") / "librm.so"
    src = ROOT / "tests" / "native" / "reml_math_host.cpp"
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-o", str(out), str(src)], check=True)
    lib = ctypes.CDLL(str(out))
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int32)
    lib.rm_pair_coefficients.restype = ctypes.c_double
    lib.rm_pair_coefficients.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int, ip, ctypes.c_int, dp, dp, dp]
    lib.rm_whiten_forward.argtypes = [ctypes.c_int, dp, dp, dp, ip, dp, dp]
    lib.rm_whiten_adjoint.argtypes = [ctypes.c_int, dp, dp, dp, ip, dp, dp]
    lib.rm_cholesky.argtypes = [ctypes.c_int, dp]
    lib.rm_cholesky_inverse.argtypes = [ctypes.c_int, dp, dp, dp]
    lib.rm_gauss_jordan.argtypes = [ctypes.c_int, dp, dp, dp]
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double if a.dtype == np.float64 else ctypes.c_int32))


def gaps(n, run_starts, tau):
    g = np.zeros(n, np.int32)
    for s, e in O.run_bounds(n, run_starts):
        if tau is None:
            g[s:e - 1] = 1
        else:
            g[s:e - 1] = np.diff(np.asarray(tau[s:e]))
    return g


def dense_Linv_via_recurrence(rm, a, b, n, run_starts, tau, white=0):
    g = gaps(n, run_starts, tau)
    s2, c, al = np.zeros(n), np.zeros(n), np.zeros(n)
    ld = rm.rm_pair_coefficients(a, b, n, _p(g), white, _p(s2), _p(c), _p(al))
    M = np.zeros((n, n))
    MT = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        w = np.zeros(n)
        rm.rm_whiten_forward(n, _p(s2), _p(c), _p(al), _p(g), _p(e), _p(w))
        M[:, j] = w
        rm.rm_whiten_adjoint(n, _p(s2), _p(c), _p(al), _p(g), _p(e), _p(w))
        MT[:, j] = w
    return M, MT, ld


CASES = [
    ("single", [60], False),
    ("runs", [40, 25, 31], False),
    ("censored", [40, 60, 33], True),
]


@pytest.mark.parametrize("name,lens,censor", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("a,b", [(0.0, 0.0), (0.6, 0.0), (0.0, -0.45), (0.8, 0.8), (0.75, -0.3), (-0.5, 0.7),
                                 (0.3, 0.25), (0.79, -0.79)])
def test_recurrence_equals_dense_cholesky(rm, name, lens, censor, a, b):
    rng = np.random.default_rng(len(lens) * 7 + int(censor))
    n = sum(lens)
    starts = [sum(lens[:i]) for i in range(len(lens))]
    tau = None
    if censor:
        tau = np.concatenate([np.sort(rng.choice(m + 6, m, replace=False)) for m in lens])
    if O.arma_lambda(a, b) < 0:
        pytest.skip("lambda < 0 pairs are filtered out")
    Linv, ld_ref = O.whitener(a, b, n, starts, None if tau is None else torch.from_numpy(tau), F64)
    M, MT, ld = dense_Linv_via_recurrence(rm, a, b, n, starts, tau)
    np.testing.assert_allclose(M, Linv.numpy(), rtol=0, atol=1e-10)
    np.testing.assert_allclose(MT, Linv.numpy().T, rtol=0, atol=1e-10)
    assert abs(ld - float(ld_ref)) < 1e-9


def test_white_pairs_are_exact_identity(rm):
    n = 50
    M, MT, ld = dense_Linv_via_recurrence(rm, 0.35, -0.35, n, [0], None, white=1)
    assert np.array_equal(M, np.eye(n)) and np.array_equal(MT, np.eye(n))
    assert ld == 2.0 * n * np.log(1.0 + 1e-10)


def test_small_dense_linear_algebra(rm):
    rng = np.random.default_rng(1)
    p = 9
    B = rng.standard_normal((p, p))
    A = B @ B.T + p * np.eye(p)
    L = A.copy()
    assert rm.rm_cholesky(p, _p(L)) == 0
    np.testing.assert_allclose(L, np.linalg.cholesky(A), rtol=0, atol=1e-12)
    Ai = np.zeros((p, p))
    tmp = np.zeros((p, p))
    rm.rm_cholesky_inverse(p, _p(L), _p(Ai), _p(tmp))
    np.testing.assert_allclose(Ai, np.linalg.inv(A), rtol=1e-10, atol=1e-13)
    inv = np.zeros((p, p))
    work = np.zeros((p, p))
    assert rm.rm_gauss_jordan(p, _p(A), _p(inv), _p(work)) == 0
    np.testing.assert_allclose(inv, np.linalg.inv(A), rtol=1e-10, atol=1e-13)
    S = np.ones((3, 3))
    assert rm.rm_cholesky(3, _p(S)) == -1
