"""Precision sweep of the projection arithmetic (fp32 / split fp16 x3 / split
bf16 x3 / fp64, the last with the recurrence energy in fp64 too) against the
FP64 oracle, on the plain T=200, p=6 design with per-voxel
task SNR log-uniform in [0.1, 10] (2 000 voxels, 33 x 33 grid).

Per mode: REML score error vs FP64 (|gpu - oracle| <= atol + rtol |score|;
fp64: 1e-7 + 1e-10 |score| against the oracle on the same scaled series),
argmin agreement with the FP64 global argmin (mismatches only at FP64 near-ties
<= tie), and beta / t error of the FP64 GLS at the selected pair vs the FP64
oracle refit at the FP64 argmin on the voxels where the pairs agree
(rel <= 1e-5).  The table is printed (pytest -s) for the record.
"""

import numpy as np
import pytest
import torch

from oracle import reml_oracle as O

pytestmark = pytest.mark.gpu
F64 = torch.float64
TOL = {"fp32": (5e-4, 2e-6, 2e-3), "fp16x3": (5e-4, 2e-6, 2e-3), "bf16x3": (2e-2, 2e-5, 2e-2)}


@pytest.fixture(scope="module")
def dev():
    from fastfuncstuff_amd import _lib

    _lib.load()
    return torch.device("cuda", 0)


def test_precision_sweep(dev):
    from fastfuncstuff_amd import synth
    from fastfuncstuff_amd.glm.arma import RemlGrid

    print("[precision] sweep", flush=True)
    wl = synth.plain_config(V=2000, T=200, seed=3, snr_range=(0.1, 10.0))
    X = torch.from_numpy(wl.X)
    ag, bg = torch.from_numpy(wl.a_grid), torch.from_numpy(wl.b_grid)
    g64 = O.precompute_grid(X.double(), ag, bg, None, None, F64)
    Yd = torch.from_numpy(wl.Y).double()
    s = Yd.std(dim=1).clamp(min=1e-8)
    S64 = O.score_surface(X.double(), Yd / s[:, None], g64)
    k64, _ = O.select_exhaustive(S64)
    V = S64.shape[0]
    b_ref, t_ref, _, _, _, _, _ = O.gls_refit(X.double(), Yd / s[:, None], g64, k64, dtype=F64)
    b_ref = b_ref * s[:, None]
    rows = []
    for mode, (atol, rtol, tie) in TOL.items():
        grid = RemlGrid(X, ag, bg, dev)
        grid.set_projection(mode)
        Y = torch.from_numpy(wl.Y).to(dev)
        Yt, scale, _ = grid.prep(Y)
        S = grid.surface(Yt, V)
        err = (S.cpu().double() - S64).abs()
        assert bool((err <= atol + rtol * S64.abs()).all()), f"{mode}: max score error {float(err.max()):.3e}"
        best, _ = grid.select(S, "exhaustive")
        k = best.cpu().long()
        same = k == k64
        gap = (S64[torch.arange(V), k] - S64[torch.arange(V), k64]).abs()
        assert bool((gap[~same] <= tie).all()), f"{mode}: mismatch beyond near-ties"
        out = grid.gls(Y, scale, best)
        db = ((out["betas"].cpu().double() - b_ref).abs() / torch.maximum(b_ref.abs(), 1e-3 * s[:, None]))[same]
        dt = ((out["tstats"].cpu().double() - t_ref).abs() / t_ref.abs().clamp(min=1))[same]
        assert float(db.max()) <= 1e-5 and float(dt.max()) <= 1e-5
        rows.append((mode, float(err.max()), float(err.median()), float(same.double().mean()), int((~same).sum()),
                     float(db.max()), float(dt.max())))
    # fp64: the float64 surface (recurrence energy and projection in float64)
    grid = RemlGrid(X, ag, bg, dev)
    Y = torch.from_numpy(wl.Y).to(dev)
    _, scale, _ = grid.prep(Y)
    S = grid.surface_f64(Y, scale)
    # the series is divided by the engine's float32 scale: the oracle scores the same series
    S64g = O.score_surface(X.double(), Yd / scale.cpu().double()[:, None], g64)
    err = (S.cpu() - S64g).abs()
    assert bool((err <= 1e-7 + 1e-10 * S64g.abs()).all()), f"fp64: max score error {float(err.max()):.3e}"
    best, _ = grid.select_f64(S, "exhaustive")
    k = best.cpu().long()
    same = k == k64
    gap = (S64[torch.arange(V), k] - S64[torch.arange(V), k64]).abs()
    assert bool((gap[~same] <= 1e-7).all()), "fp64: mismatch beyond near-ties"
    out = grid.gls(Y, scale, best)
    db = ((out["betas"].cpu().double() - b_ref).abs() / torch.maximum(b_ref.abs(), 1e-3 * s[:, None]))[same]
    dt = ((out["tstats"].cpu().double() - t_ref).abs() / t_ref.abs().clamp(min=1))[same]
    assert float(db.max()) <= 1e-5 and float(dt.max()) <= 1e-5
    rows.append(("fp64", float(err.max()), float(err.median()), float(same.double().mean()), int((~same).sum()),
                 float(db.max()), float(dt.max())))
    print("\nmode     max|dS|    median|dS|  argmin=FP64  n_mismatch  beta_rel_max  t_rel_max")
    for r in rows:
        print(f"{r[0]:8s} {r[1]:.3e}  {r[2]:.3e}   {r[3]:.5f}     {r[4]:4d}        {r[5]:.2e}      {r[6]:.2e}")


def test_use_double_is_the_float64_search(dev):
    """use_double: float64 search on the raw series (no normalisation or
    residualisation), as the FP64 oracle; scores within 1e-6 + 1e-10 |score|
    (a series at level 500 leaves ~1e-7 of float64 cancellation in
    |y_w|^2 - |Q'y_w|^2 on both sides), hierarchical pair identical except
    near-ties <= 1e-6, betas / t within 1e-5; float32 and float64 input give
    the same pairs."""
    from fastfuncstuff_amd import synth
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[precision] use_double", flush=True)
    wl = synth.gpu_config(V=1500, T=160, seed=9, noise_seed=10)
    X = torch.from_numpy(wl.X)
    ag, bg = torch.from_numpy(wl.a_grid), torch.from_numpy(wl.b_grid)
    tau = torch.from_numpy(wl.tau)
    Y32 = torch.from_numpy(wl.Y) + 500.0  # a large mean: no residualisation in float64
    res = fit_glm_arma11(Y32, X, 2.0, a_grid=ag, b_grid=bg, device=dev, run_starts=wl.run_starts, tau=tau,
                         use_double=True, save_profile_likelihoods=True)
    ref = O.fit_arma11(Y32.double(), X, ag, bg, wl.run_starts, tau, dtype=F64, selector="hierarchical")
    assert res.betas.dtype == torch.float64 and res.reml_likelihood.dtype == torch.float64
    err = (res.reml_lklhd_surface - ref.surface).abs()
    assert bool((err <= 1e-6 + 1e-10 * ref.surface.abs()).all()), f"max score error {float(err.max()):.3e}"
    diff = res.pair_index != ref.pair_index
    print(f"[precision] use_double: {int(diff.sum())} pair mismatches of {Y32.shape[0]}", flush=True)
    V = Y32.shape[0]
    gap = (ref.surface[torch.arange(V), res.pair_index] - ref.surface[torch.arange(V), ref.pair_index]).abs()
    assert bool((gap[diff] <= 1e-6).all())
    ok = ~diff
    sc = Y32.double().std(dim=1)
    db = (res.betas - ref.betas).abs() / torch.maximum(ref.betas.abs(), 1e-3 * sc[:, None])
    assert float(db[ok].max()) <= 1e-5
    assert float(((res.tstats - ref.tstats).abs() / ref.tstats.abs().clamp(min=1))[ok].max()) <= 1e-5
    lk = ref.reml_likelihood
    assert bool(((res.reml_likelihood - lk).abs()[ok] <= 1e-6 + 1e-10 * lk.abs()[ok]).all())
    res64 = fit_glm_arma11(Y32.double(), X, 2.0, a_grid=ag, b_grid=bg, device=dev, run_starts=wl.run_starts, tau=tau,
                           use_double=True)
    assert bool((res64.pair_index == res.pair_index).all())
