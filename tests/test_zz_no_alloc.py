"""The allocation-free GLS entry point (ffs_reml_gls_noalloc) through the C
ABI against ffs_reml_gls with a caller workspace: identical outputs (bitwise:
the same kernels on the same grouping rule), on two calls in a row with the
second on fewer voxels (the shared device scratch keeps the first call's
contents), on the summary-statistics path and the series path; sizes past
the scratch capacity are refused."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _outputs(grid, V, dev, series):
    f32 = dict(dtype=torch.float32, device=dev)
    o = {"betas": torch.full((V, grid.p), -7.0, **f32), "tstats": torch.full((V, grid.p), -7.0, **f32),
         "fstats": torch.full((V,), -7.0, **f32), "sigma2": torch.full((V,), -7.0, **f32),
         "r2": torch.full((V,), -7.0, **f32)}
    o["resid_w"] = torch.full((V, grid.T), -7.0, dtype=torch.float64, device=dev) if series else None
    return o


def _call(lib, grid, Y, scale, best, o, work, st):
    from fastfuncstuff_amd import _lib

    args = [grid.handle, _lib.ptr(Y), Y.stride(0), _lib.ptr(scale), _lib.ptr(best), Y.shape[0],
            _lib.ptr(o["betas"]), _lib.ptr(o["tstats"]), _lib.ptr(o["fstats"]), _lib.ptr(o["sigma2"]),
            _lib.ptr(o["r2"]), _lib.ptr(o["resid_w"]), grid.T, None, None]
    if work is None:
        return lib.ffs_reml_gls_noalloc(*args, st)
    return lib.ffs_reml_gls(*args, _lib.ptr(work), st)


def _same(a, b):
    return bool(torch.equal(torch.nan_to_num(a, nan=12345.0), torch.nan_to_num(b, nan=12345.0)))


@pytest.mark.parametrize("series", [False, True])
def test_gls_noalloc_matches_gls(series):
    from fastfuncstuff_amd import _lib, synth
    from fastfuncstuff_amd.glm.arma import RemlGrid

    dev = torch.device("cuda", 0)
    lib = _lib.load()
    st = _lib.stream_handle(dev)
    wl = synth.gpu_config(V=6000, T=200, seed=7, noise_seed=8)
    grid = RemlGrid(torch.from_numpy(wl.X), torch.from_numpy(wl.a_grid), torch.from_numpy(wl.b_grid), dev,
                    run_starts=wl.run_starts, tau=torch.from_numpy(wl.tau))
    Yall = torch.from_numpy(wl.Y).to(dev)
    rng = np.random.default_rng(3)
    for V in (6000, 1700):  # second call smaller: stale scratch contents must not leak
        print(f"[noalloc] series={series} V={V}", flush=True)
        Y = Yall[:V].contiguous()
        _, scale, _ = grid.prep(Y, normalize=True)
        idx = rng.integers(0, grid.K, V).astype(np.int32)
        idx[::97] = 17  # one large group spanning waves
        idx[5] = -1     # invalid pair: NaN outputs
        best = torch.from_numpy(idx).to(dev)
        n = ctypes.c_int64()
        _lib.check(lib.ffs_reml_gls_workspace(grid.handle, V, ctypes.byref(n)), "gls_workspace")
        work = torch.empty(int(n.value), dtype=torch.int32, device=dev)
        ref = _outputs(grid, V, dev, series)
        got = _outputs(grid, V, dev, series)
        _lib.check(_call(lib, grid, Y, scale, best, ref, work, st), "gls")
        _lib.check(_call(lib, grid, Y, scale, best, got, None, st), "gls_noalloc")
        torch.cuda.synchronize(dev)
        for k in ref:
            if ref[k] is not None:
                assert _same(got[k], ref[k]), (k, V)
        assert bool(torch.isnan(got["betas"][5]).all())


def test_gls_noalloc_refuses_past_capacity():
    from fastfuncstuff_amd import _lib
    from fastfuncstuff_amd.glm.arma import RemlGrid

    dev = torch.device("cuda", 0)
    lib = _lib.load()
    T = 50
    X = torch.ones(T, 1, dtype=torch.float64)
    grid = RemlGrid(X, torch.linspace(-0.8, 0.8, 5), torch.linspace(-0.8, 0.8, 5), dev)
    V = 3_000_000  # more than the scratch can group; refused before any pointer is read
    rc = lib.ffs_reml_gls_noalloc(grid.handle, None, T, None, None, V, None, None, None, None, None, None, T, None,
                                  None, _lib.stream_handle(dev))
    assert rc == -3
    assert "maximum" in _lib.last_error()
