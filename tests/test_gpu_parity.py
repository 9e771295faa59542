"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle.

Tolerances (stated here, used below):
* REML score: |gpu - oracle| <= 5e-4 + 2e-6 |score|  (the GPU surface is float32,
  as the reference's is; fp32 rounding of a ~10^3 score is ~1e-4);
* selected pair: identical to the oracle's argmin except where the oracle's own
  FP64 scores of the two pairs differ by <= 2e-3 (a near-tie, counted);
* log-determinants: 1e-9 absolute;
* GLS at the selected pair: betas rel <= 1e-5 (of max(|beta|, 1e-3 * sd)),
  t rel <= 1e-5, F / sigma2 rel <= 1e-5, R^2 abs <= 1e-6.
"""

import numpy as np
import pytest
import torch

from oracle import reml_oracle as O

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _case(name):
    rng = np.random.default_rng({"plain": 1, "censored": 2, "afni": 3, "ragged": 4, "wide": 5, "wide60": 6,
                                 "wide68": 7, "wide90": 8, "seg120": 9, "seg300": 10, "seg520": 11,
                                 "segcens": 12}[name])
    if name in ("plain", "ragged", "wide", "wide60", "wide68", "wide90", "seg120", "seg300", "seg520"):
        # design widths 40 / 60 / 68 / 90 cover the f64-MFMA prep's 3 to 6 column
        # tiles; 120 / 300 / 520 the segmented path (2 / 5 / 9 segments of 64
        # columns, one to three column groups of the prep)
        T = {"wide60": 140, "wide68": 150, "wide90": 180, "seg120": 220, "seg300": 420, "seg520": 600}.get(name, 100)
        p = {"plain": 6, "ragged": 6, "wide": 40, "wide60": 60, "wide68": 68, "wide90": 90, "seg120": 120,
             "seg300": 300, "seg520": 520}[name]
        V = {"plain": 256, "ragged": 333, "wide": 130, "seg520": 70}.get(name, 100)
        X = np.column_stack([np.ones(T), np.linspace(-1, 1, T), rng.standard_normal((T, p - 2))])
        run_starts, tau = None, None
        ag = bg = torch.linspace(-0.8, 0.8, {"seg300": 17, "seg520": 9}.get(name, 33))
    elif name in ("censored", "segcens"):
        # segcens: two runs with censored volumes at p = 110 (segmented path)
        lens, p, V = ([70, 80], 12, 200) if name == "censored" else ([140, 150], 110, 120)
        full = sum(lens)
        good = np.sort(rng.choice(full, full - 8, replace=False))
        run_starts, tau = O.build_censor_run_info([0, lens[0]], full, good.tolist())
        T = len(good)
        X = np.column_stack([np.ones(T), rng.standard_normal((T, p - 1))])
        ag = bg = torch.linspace(-0.8, 0.8, 33)
    else:  # AFNI default grid, two equal runs
        T, p, V = 120, 4, 192
        X = np.column_stack([np.ones(T), rng.standard_normal((T, p - 1))])
        run_starts, tau = [0, 60], None
        ag, bg = O.default_grids()
    a = rng.uniform(0.1, 0.8, V)
    b = rng.uniform(-0.5, 0.3, V)
    e = np.zeros((V, T))
    eps_prev = rng.standard_normal(V)
    for t in range(T):
        eps = rng.standard_normal(V)
        e[:, t] = (a * e[:, t - 1] if t else 0) + eps + b * eps_prev
        eps_prev = eps
    Y = (rng.standard_normal((V, p)) @ X.T * 0.5 + 5.0 + e).astype(np.float32)
    return X, Y, ag, bg, run_starts, tau


def _oracle_on_normalised(X, Y, grid):
    Yd = torch.from_numpy(Y).double()
    s = Yd.std(dim=1).clamp(min=1e-8)
    return O.score_surface(torch.from_numpy(X), Yd / s[:, None], grid), s


@pytest.fixture(scope="module")
def dev():
    from fastfuncstuff_amd import _lib

    _lib.load()
    return torch.device("cuda", 0)


@pytest.mark.parametrize("name", ["plain", "censored", "afni", "ragged", "wide", "wide60", "wide68", "wide90",
                                  "seg120", "seg300", "seg520", "segcens"])
def test_surface_selection_and_gls_match_oracle(dev, name):
    from fastfuncstuff_amd.glm.arma import RemlGrid

    print(f"[parity] case={name}", flush=True)
    X, Y, ag, bg, run_starts, tau = _case(name)
    grid_o = O.precompute_grid(torch.from_numpy(X), ag, bg, run_starts, tau, F64)
    grid = RemlGrid(torch.from_numpy(X), ag, bg, dev, run_starts=run_starts, tau=tau)
    assert grid.param_list == [(float(a), float(b)) for a, b in grid_o.params.float().tolist()]
    np.testing.assert_allclose(grid.logdet_R, grid_o.logdet_R.numpy(), rtol=0, atol=1e-9)
    np.testing.assert_allclose(grid.logdet_XRX, grid_o.logdet_XRX.numpy(), rtol=0, atol=1e-9)

    Yd = torch.from_numpy(Y).to(dev)
    V = Y.shape[0]
    Yt, scale, _ = grid.prep(Yd, normalize=True)
    S = grid.surface(Yt, V).cpu().double()
    S64, s_ref = _oracle_on_normalised(X, Y, grid_o)
    np.testing.assert_allclose(scale.cpu().double().numpy(), s_ref.numpy(), rtol=1e-6)
    err = (S - S64).abs()
    tol = 5e-4 + 2e-6 * S64.abs()
    assert bool((err <= tol).all()), f"max score error {float(err.max()):.3e}"

    for hier in (False, True):
        best, score = grid.select(S.float().to(dev), hierarchical=hier)
        best = best.cpu().long()
        sel = O.select_hierarchical if hier else (lambda s, p: O.select_exhaustive(s))
        k_ref, _ = sel(S64.float(), grid_o.params) if hier else O.select_exhaustive(S64.float())
        diff = best != k_ref
        gap = (S64[torch.arange(V), best] - S64[torch.arange(V), k_ref]).abs()
        assert bool((gap[diff] <= 2e-3).all()), f"hier={hier}: {int(diff.sum())} mismatches beyond near-ties"
        assert int(diff.sum()) <= max(1, V // 100)

    best, _ = grid.select(S.float().to(dev), hierarchical=True)
    idx = best.cpu().long()
    Yn = torch.from_numpy(Y).double() / s_ref[:, None]
    ref = O.gls_refit(torch.from_numpy(X), Yn, grid_o, idx, dtype=F64)
    # summary statistics only (f64 MFMA path for p <= 32), then with the
    # whitened residual series (per-lane path)
    _check_gls(grid.gls(Yd, scale, best), ref, s_ref)
    out = grid.gls(Yd, scale, best, want_resid_w=True)
    _check_gls(out, ref, s_ref)
    rw_ref = ref[6] * s_ref[:, None]
    assert float((out["resid_w"].cpu() - rw_ref).abs().max()) <= 1e-6 * float(rw_ref.abs().max())


def _check_gls(out, ref, s_ref):
    betas, tst, fst, s2, r2 = ref[0] * s_ref[:, None], ref[1], ref[2], ref[3] * s_ref**2, ref[4]
    gb = out["betas"].cpu().double()
    denom = torch.maximum(betas.abs(), 1e-3 * s_ref[:, None])
    assert float(((gb - betas).abs() / denom).max()) <= 1e-5
    gt = out["tstats"].cpu().double()
    assert float(((gt - tst).abs() / torch.maximum(tst.abs(), torch.ones_like(tst))).max()) <= 1e-5
    assert float(((out["fstats"].cpu().double() - fst).abs() / fst.abs().clamp(min=1)).max()) <= 1e-5
    assert float(((out["sigma2"].cpu().double() - s2).abs() / s2).max()) <= 1e-5
    assert float((out["r2"].cpu().double() - r2).abs().max()) <= 1e-6


def test_white_pairs_tie_bitwise(dev):
    from fastfuncstuff_amd.glm.arma import RemlGrid

    X, Y, ag, bg, rs, tau = _case("plain")
    grid = RemlGrid(torch.from_numpy(X), ag, bg, dev)
    lam = O.arma_lambda(grid.params[:, 0], grid.params[:, 1])
    white = (lam == 0).nonzero().flatten()
    assert white.numel() == 33
    Yt, _, _ = grid.prep(torch.from_numpy(Y).to(dev))
    S = grid.surface(Yt, Y.shape[0]).cpu()
    Sw = S[:, white]
    assert bool((Sw == Sw[:, :1]).all())


def test_single_voxel_and_ljung_box(dev):
    from fastfuncstuff_amd.glm.arma import RemlGrid, ljung_box_device, ljung_box_max_lag

    X, Y, ag, bg, rs, tau = _case("censored")
    grid = RemlGrid(torch.from_numpy(X), ag, bg, dev, run_starts=rs, tau=tau)
    Yd = torch.from_numpy(Y[:1]).to(dev)
    Yt, scale, _ = grid.prep(Yd)
    S = grid.surface(Yt, 1)
    best, _ = grid.select(S, hierarchical=True)
    out = grid.gls(Yd, scale, best, want_resid_w=True)
    T = X.shape[0]
    lag = ljung_box_max_lag(T, X.shape[1], min(rs[1], T - rs[1]))
    from fastfuncstuff_amd.glm.arma import build_ljung_box_tau

    lt = build_ljung_box_tau(T, rs, tau)
    lb = ljung_box_device(out["resid_w"], lag, lt).cpu().double()
    ref = O.ljung_box(out["resid_w"].cpu(), lag, lt)
    assert float((lb - ref).abs().max()) <= 1e-4 * max(1.0, float(ref.abs().max()))


# Split 16-bit projections (hi + lo halves, three MFMA products, fp32 accumulation).
# Score tolerance |gpu - oracle| <= atol + rtol |score| per mode; argmin mismatches
# only where the oracle's FP64 scores of the two pairs are within `tie`.
SPLIT_TOL = {"fp16x3": dict(atol=5e-4, rtol=2e-6, tie=2e-3), "bf16x3": dict(atol=2e-2, rtol=2e-5, tie=2e-2)}


@pytest.mark.parametrize("mode", ["fp16x3", "bf16x3"])
@pytest.mark.parametrize("name", ["plain", "censored", "afni", "wide", "seg120", "seg300", "segcens"])
def test_split_projection_matches_oracle(dev, name, mode):
    from fastfuncstuff_amd.glm.arma import RemlGrid

    print(f"[parity] split case={name} mode={mode}", flush=True)
    X, Y, ag, bg, run_starts, tau = _case(name)
    grid_o = O.precompute_grid(torch.from_numpy(X), ag, bg, run_starts, tau, F64)
    grid = RemlGrid(torch.from_numpy(X), ag, bg, dev, run_starts=run_starts, tau=tau)
    grid.set_projection(mode)
    V = Y.shape[0]
    Yt, _, _ = grid.prep(torch.from_numpy(Y).to(dev), normalize=True)
    S = grid.surface(Yt, V).cpu().double()
    S64, _ = _oracle_on_normalised(X, Y, grid_o)
    tol = SPLIT_TOL[mode]
    err = (S - S64).abs()
    print(f"[parity] {mode} {name}: max score error {float(err.max()):.3e}", flush=True)
    assert bool((err <= tol["atol"] + tol["rtol"] * S64.abs()).all()), f"max score error {float(err.max()):.3e}"
    best, _ = grid.select(S.float().to(dev), hierarchical=False)
    best = best.cpu().long()
    k_ref, _ = O.select_exhaustive(S64.float())
    diff = best != k_ref
    gap = (S64[torch.arange(V), best] - S64[torch.arange(V), k_ref]).abs()
    assert bool((gap[diff] <= tol["tie"]).all()), f"{int(diff.sum())} mismatches beyond near-ties"
    with pytest.raises(ValueError):
        grid.set_projection("fp8")


def test_gls_groups_spanning_waves(dev):
    """Pair groups larger than one wave, one-voxel groups and invalid pairs in
    one call: every voxel matches the oracle's refit at its pair."""
    from fastfuncstuff_amd.glm.arma import RemlGrid

    print("[parity] gls grouping", flush=True)
    X, Y, ag, bg, run_starts, tau = _case("censored")
    grid_o = O.precompute_grid(torch.from_numpy(X), ag, bg, run_starts, tau, F64)
    grid = RemlGrid(torch.from_numpy(X), ag, bg, dev, run_starts=run_starts, tau=tau)
    V = Y.shape[0]
    rng = np.random.default_rng(9)
    idx = np.where(rng.random(V) < 0.7, 17, rng.integers(0, grid.K, V))  # one group of ~140 voxels
    idx[5] = grid.K + 3  # invalid pair -> NaN statistics
    idx[6] = -1
    best = torch.from_numpy(idx.astype(np.int32)).to(dev)
    Yd = torch.from_numpy(Y).to(dev)
    _, scale, _ = grid.prep(Yd, normalize=True)
    out = grid.gls(Yd, scale, best)
    s_ref = torch.from_numpy(Y).double().std(dim=1).clamp(min=1e-8)
    ok = torch.ones(V, dtype=torch.bool)
    ok[5] = ok[6] = False
    ref = O.gls_refit(torch.from_numpy(X), torch.from_numpy(Y).double()[ok] / s_ref[ok, None], grid_o,
                      torch.from_numpy(idx)[ok].long(), dtype=F64)
    _check_gls({k: v[ok.to(dev)] for k, v in out.items() if v is not None}, ref, s_ref[ok])
    for v in (5, 6):
        assert bool(torch.isnan(out["betas"][v]).all()) and bool(torch.isnan(out["sigma2"][v]))


@pytest.mark.parametrize("name", ["plain", "censored", "afni"])
def test_selectors_exact_on_device_surface(dev, name):
    """The three selection rules on the engine's own float32 surface give
    exactly the oracle's pair for every voxel (same inputs, same rules)."""
    from fastfuncstuff_amd.glm.arma import RemlGrid

    print(f"[parity] selectors case={name}", flush=True)
    X, Y, ag, bg, run_starts, tau = _case(name)
    grid = RemlGrid(torch.from_numpy(X), ag, bg, dev, run_starts=run_starts, tau=tau)
    Yt, _, _ = grid.prep(torch.from_numpy(Y).to(dev))
    S = grid.surface(Yt, Y.shape[0])
    Sc = S.cpu()
    params = grid.params.float()
    for mode, ref in (("exhaustive", lambda: O.select_exhaustive(Sc)),
                      ("hierarchical", lambda: O.select_hierarchical(Sc, params)),
                      ("subset", lambda: O.search_hierarchical_subset(Sc, params))):
        best, score = grid.select(S, mode)
        k_ref, s_ref = ref()
        assert bool((best.cpu().long() == k_ref.long()).all()), mode
        assert bool((score.cpu() == s_ref.float()).all()), mode
