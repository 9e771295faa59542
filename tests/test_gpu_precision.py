"""Precision sweep of the projection arithmetic (fp32 / split fp16 x3 / split
bf16 x3) against the FP64 oracle, on the plain T=200, p=6 design with per-voxel
task SNR log-uniform in [0.1, 10] (2 000 voxels, 33 x 33 grid).

Per mode: REML score error vs FP64 (|gpu - oracle| <= atol + rtol |score|),
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
    print("\nmode     max|dS|    median|dS|  argmin=FP64  n_mismatch  beta_rel_max  t_rel_max")
    for r in rows:
        print(f"{r[0]:8s} {r[1]:.3e}  {r[2]:.3e}   {r[3]:.5f}     {r[4]:4d}        {r[5]:.2e}      {r[6]:.2e}")
