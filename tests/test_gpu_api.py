"""GPU tests of the public search variants against the oracle restatements.

Tolerances: compute_reml_likelihood rel 1e-9 (FP64 both sides); the grid
searches compute their surface in float32: selected pair equal to the oracle's
except near-ties (FP64 score gap <= 2e-3), scores |diff| <= 5e-4 + 2e-6 |score|.
"""

import numpy as np
import pytest
import torch

from oracle import reml_oracle as O

pytestmark = pytest.mark.gpu
F64 = torch.float64


@pytest.fixture(scope="module")
def dev():
    from fastfuncstuff_amd import _lib

    _lib.load()
    return torch.device("cuda", 0)


def _voxels(rng, T, X, V, mean=3.0):
    e = np.zeros((V, T))
    a = rng.uniform(0.1, 0.8, V)
    b = rng.uniform(-0.5, 0.3, V)
    ep = rng.standard_normal(V)
    for t in range(T):
        eps = rng.standard_normal(V)
        e[:, t] = (a * e[:, t - 1] if t else 0) + eps + b * ep
        ep = eps
    return (rng.standard_normal((V, X.shape[1])) @ X.T * 0.5 + mean + e).astype(np.float32)


def _check_search(k, s, k_ref, S_ref):
    V = S_ref.shape[0]
    diff = k != k_ref
    gap = (S_ref[torch.arange(V), k] - S_ref[torch.arange(V), k_ref]).abs()
    assert bool((gap[diff] <= 2e-3).all()), f"{int(diff.sum())} mismatches beyond near-ties"
    ref = S_ref[torch.arange(V), k]
    assert bool(((s.double() - ref).abs() <= 5e-4 + 2e-6 * ref.abs()).all())


def test_compute_reml_likelihood(dev):
    from fastfuncstuff_amd.glm.arma import compute_reml_likelihood

    print("[api] dense likelihood", flush=True)
    rng = np.random.default_rng(1)
    T, p = 90, 4
    X = np.column_stack([np.ones(T), rng.standard_normal((T, p - 1))])
    y = rng.standard_normal(T)
    for R in (O.covariance(0.6, -0.3, T, [0, 40], None, F64), torch.eye(T, dtype=F64) * 2.0):
        got = compute_reml_likelihood(torch.from_numpy(X), torch.from_numpy(y), R, device=dev)
        want = O.reml_likelihood_dense(torch.from_numpy(X), torch.from_numpy(y), R)
        assert abs(got - want) <= 1e-9 * abs(want)
    Y2 = rng.standard_normal((T, 3))
    R = O.covariance(0.3, 0.2, T, dtype=F64)
    got = compute_reml_likelihood(torch.from_numpy(X), torch.from_numpy(Y2), R, device=dev)
    assert abs(got - O.reml_likelihood_dense(torch.from_numpy(X), torch.from_numpy(Y2), R)) <= 1e-9 * abs(got)
    bad = torch.diag(torch.cat([torch.ones(T - 1, dtype=F64), -torch.ones(1, dtype=F64)]))
    assert compute_reml_likelihood(torch.from_numpy(X), torch.from_numpy(y), bad, device=dev) == 1e10


def test_ridge_grid_searches(dev):
    from fastfuncstuff_amd.glm.arma import reml_grid_search, reml_grid_search_batched, valid_grid_pairs

    print("[api] ridge searches", flush=True)
    rng = np.random.default_rng(2)
    T, V = 60, 12
    X = np.column_stack([np.ones(T), np.linspace(-1, 1, T), rng.standard_normal(T)])
    Y = _voxels(rng, T, X, V)
    ag, bg = torch.linspace(0.1, 0.8, 8), torch.linspace(-0.4, 0.3, 8)
    params, lik = reml_grid_search_batched(torch.from_numpy(Y), torch.from_numpy(X), ag, bg, dev)
    pairs = valid_grid_pairs(ag, bg)
    k_ref, s_ref = O.ridge_grid_search(X, Y, pairs)
    # oracle surface for the near-tie check
    S_ref = torch.stack([torch.tensor([O.reml_likelihood_dense(X, Y[v], O.covariance(float(a), float(b), T,
                                                                                      dtype=F64))
                                       for a, b in pairs.double().tolist()]) for v in range(V)])
    k = torch.tensor([next(i for i, pr in enumerate(pairs.tolist()) if pr == prm) for prm in params.tolist()])
    _check_search(k, lik, k_ref, S_ref)
    a, b, l1 = reml_grid_search(torch.from_numpy(X), torch.from_numpy(Y[0]), ag, bg, dev)
    assert (a, b) == tuple(params[0].tolist()) and abs(l1 - float(lik[0])) <= 1e-6 * abs(l1)
    # raw series at a large level (the search sees the data unnormalised):
    # the score must not lose its digits to the mean
    Yo = (Y + 500.0).astype(np.float32)
    params, lik = reml_grid_search_batched(torch.from_numpy(Yo), torch.from_numpy(X), ag, bg, dev)
    k_ref, _ = O.ridge_grid_search(X, Yo, pairs)
    S_ref = torch.stack([torch.tensor([O.reml_likelihood_dense(X, torch.from_numpy(Yo[v]).double(),
                                                               O.covariance(float(a), float(b), T, dtype=F64))
                                       for a, b in pairs.double().tolist()]) for v in range(V)])
    k = torch.tensor([next(i for i, pr in enumerate(pairs.tolist()) if pr == prm) for prm in params.tolist()])
    _check_search(k, lik, k_ref, S_ref)


def test_batch_grid_search_priority_order_and_background(dev):
    from fastfuncstuff_amd.glm.arma import RemlGrid, _PRIORITY_PARAMS, batch_reml_grid_search

    print("[api] batch search", flush=True)
    rng = np.random.default_rng(3)
    T, V = 100, 200
    X = np.column_stack([np.ones(T), rng.standard_normal((T, 2))])
    Y = _voxels(rng, T, X, V)
    Y[7] = 5.0  # constant series: background
    ag, bg = torch.linspace(0, 0.8, 9), torch.linspace(-0.8, 0.8, 17)
    params, lik = batch_reml_grid_search(torch.from_numpy(X), torch.from_numpy(Y).T, ag, bg, dev)
    grid = RemlGrid(torch.from_numpy(X), ag, bg, dev)
    g64 = O.precompute_grid(torch.from_numpy(X).double(), ag, bg, None, None, F64)
    S_ref = O.score_surface(torch.from_numpy(X).double(), torch.from_numpy(Y).double(), g64)
    order = [k for k, pr in enumerate(grid.param_list) if pr in _PRIORITY_PARAMS]
    order += [k for k in range(grid.K) if k not in order]
    k_ref, _ = O.select_in_order(S_ref.float(), order)
    pl = grid.param_list
    k = torch.tensor([pl.index(tuple(pr)) for pr in params.cpu().tolist()])
    live = torch.ones(V, dtype=torch.bool)
    live[7] = False
    _check_search(k[live], lik.cpu()[live], k_ref[live], S_ref[live])
    assert params[7].tolist() == [0.0, 0.0] and float(lik[7]) == -1.0


def test_estimate_arma11_per_voxel_matches_oracle(dev):
    """design/flobs.py:2750-2940 on the engine: two runs of different length,
    per-run polort (block-diagonal Legendre drift), run starts at the run
    boundary, default grids; the exhaustive argmin equals the oracle's on every
    voxel whose best and second-best oracle scores differ by more than 1e-3,
    and the chosen pair's oracle score is within 1e-5 (relative) of the
    oracle minimum on every voxel (white-noise surfaces hold near-ties)."""
    from fastfuncstuff_amd.design.flobs import estimate_arma11_per_voxel, legendre_polynomials

    print("[api] estimate_arma11_per_voxel", flush=True)
    rng = np.random.default_rng(8)
    n1, n2, V = 90, 70, 120
    t1, t2 = rng.standard_normal((n1, 2)), rng.standard_normal((n2, 2))
    d1 = rng.standard_normal((V, n1)) + (t1 @ rng.standard_normal((2, V))).T
    d2 = rng.standard_normal((V, n2)) + (t2 @ rng.standard_normal((2, V))).T
    got = estimate_arma11_per_voxel([torch.from_numpy(d1).float(), torch.from_numpy(d2).float()],
                                    [torch.from_numpy(t1), torch.from_numpy(t2)], [1, 2], device=dev,
                                    verbose=False)
    assert got.shape == (V, 2) and got.dtype == np.float32
    Z = np.zeros((n1 + n2, 5))
    Z[:n1, :2] = legendre_polynomials(n1, 1)
    Z[n1:, 2:] = legendre_polynomials(n2, 2)
    X = torch.from_numpy(np.concatenate([np.concatenate([t1, t2]), Z], axis=1)).float().double()
    Y = torch.from_numpy(np.concatenate([d1, d2], axis=1)).float().double()
    ag, bg = O.default_grids()
    g = O.precompute_grid(X, ag, bg, run_starts=[0, n1])
    S = O.score_surface(X, Y, g)
    k, _ = O.select_exhaustive(S)
    srt = S.sort(dim=1).values
    clear = (srt[:, 1] - srt[:, 0]) > 1e-3
    ref = g.params[k].float().numpy()
    assert clear.sum() > V // 4
    assert np.array_equal(got[clear.numpy()], ref[clear.numpy()])
    # everywhere: the engine's pair scores within 1e-5 (relative) of the oracle minimum
    lut = {(float(a), float(b)): i for i, (a, b) in enumerate(g.params.tolist())}
    kg = torch.tensor([lut[(float(a), float(b))] for a, b in got.tolist()])
    sg = S.gather(1, kg[:, None])[:, 0]
    assert float(((sg - srt[:, 0]).abs() / srt[:, 0].abs().clamp(min=1.0)).max()) <= 1e-5


def test_whiteners_match_oracle(dev):
    """glm/whiten.py whiteners on the engine's recurrence: dense L^-1 of a
    censored two-run timeline vs the oracle's Cholesky inverse (1e-10), rows /
    cols / stack_cols consistent with it, log det R per pair (1e-9), the
    autocorr grid's pairs equal the search grid's (zero inserted), and an
    invalid pair raising as in the reference."""
    from fastfuncstuff_amd.glm.arma import RemlGrid, precompute_autocorr_grid
    from fastfuncstuff_amd.glm.whiten import arma11_whitener, arma11_whiteners, stack_cols

    print("[api] whiteners", flush=True)
    n = 40
    tau = torch.tensor(list(range(15)) + list(range(17, 42)))
    rs = [0, 15]
    w = arma11_whitener(0.6, 0.2, n, run_starts=rs, tau=tau, device=dev, dtype=torch.float64)
    Li, ld = O.whitener(float(torch.tensor(0.6, dtype=torch.float64)), 0.2, n, run_starts=rs, tau=tau)
    D = w.dense().cpu()
    assert float((D - Li).abs().max()) <= 1e-10 * float(Li.abs().max())
    rng = np.random.default_rng(3)
    M = torch.from_numpy(rng.standard_normal((n, 3)))
    torch.testing.assert_close(w.cols(M.to(dev)).cpu(), Li @ M, rtol=1e-10, atol=1e-10)
    torch.testing.assert_close(w.rows(M.T.contiguous().to(dev)).cpu(), M.T @ Li.T, rtol=1e-10, atol=1e-10)
    g = torch.linspace(-0.8, 0.8, 5)
    ws, lds, plist = arma11_whiteners(g, g, n, run_starts=rs, tau=tau, device=dev, dtype=torch.float64)
    for k in (0, len(plist) - 1):
        a, b = plist[k]
        Lk, ldk = O.whitener(a, b, n, run_starts=rs, tau=tau)
        assert abs(float(lds[k]) - float(ldk)) <= 1e-9 * max(1.0, abs(float(ldk)))
    S = stack_cols(ws[:2], M.to(dev)).cpu()
    assert S.shape == (2, n, 3)
    cache = precompute_autocorr_grid(n, g, g, dev, run_starts=rs, tau=tau)
    assert cache["param_list"] == RemlGrid(torch.ones(n, 1), g, g, dev, run_starts=rs, tau=tau).param_list
    with pytest.raises(ValueError):
        arma11_whitener(0.2, -0.6, n, device=dev)


def test_prewhiten_ljung_box_and_comparison_helpers(dev):
    """prewhiten_with_arma11 vs the oracle's dense whitening (1e-10); the
    Ljung-Box helper on fp32 and fp64 residuals equals the oracle's statistic;
    the -dsort pass with want_ljung_box fills it; compare_ols_vs_arma11 runs."""
    from fastfuncstuff_amd.glm import compare_ols_vs_arma11, compute_ljung_box_statistic, prewhiten_with_arma11
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[api] prewhiten / ljung-box / compare", flush=True)
    rng = np.random.default_rng(12)
    T, V = 60, 40
    X = torch.from_numpy(np.column_stack([np.ones(T), np.linspace(-1, 1, T), rng.standard_normal(T)]))
    Y = torch.from_numpy(rng.standard_normal((T, V)))
    Xw, Yw, L = prewhiten_with_arma11(X, Y, 0.5, 0.1, run_starts=[0, 25])
    Li, _ = O.whitener(0.5, 0.1, T, run_starts=[0, 25])  # (a, b) taken in fp64, as the covariance builder takes them
    torch.testing.assert_close(Xw, Li @ X, rtol=1e-10, atol=1e-10)
    torch.testing.assert_close(Yw, Li @ Y, rtol=1e-10, atol=1e-10)
    R = L.double() @ L.double().T
    torch.testing.assert_close(torch.linalg.inv(R), Li.T @ Li, rtol=1e-8, atol=1e-8)
    res = Yw.T.contiguous()
    ref = O.ljung_box(res, None, None).numpy()
    for r in (res, res.float()):
        got = compute_ljung_box_statistic(r)
        assert np.allclose(got, ref, rtol=1e-4, atol=1e-4)
    Yv = (Y.T.numpy() + 0.3 * X[:, 1].numpy()).astype(np.float32)
    D = rng.standard_normal((V, T)).astype(np.float32)
    ag = bg = torch.linspace(-0.8, 0.8, 9)
    fit = fit_glm_arma11(Yv, X.numpy(), 2.0, a_grid=ag, b_grid=bg, device=dev, dsort=D, want_ljung_box=True)
    assert fit.ljung_box is not None and bool(torch.isfinite(fit.ljung_box).all())
    out = compare_ols_vs_arma11(Yv, X.numpy(), 2.0, device=dev)
    assert set(out) >= {"ols", "arma11", "tstat_ratio", "r2_improvement", "summary"}


def test_staged_upload_equals_plain_copy(dev):
    """Host arrays of 16 MB and more go to the device through two pinned
    staging buffers (chunks of 32 MB); the device tensor equals torch's own
    copy for float32 and float64 sources, a partial last chunk, a second call
    reusing the buffers, a non-contiguous source and a small array."""
    from fastfuncstuff_amd import _lib

    print("[upload] staged host -> device copies", flush=True)
    rng = np.random.default_rng(5)
    cases = [rng.standard_normal((70_001, 300)).astype(np.float32),  # 3 chunks, the last partial
             rng.standard_normal((20_000, 400)),  # float64 source: converted on the host
             rng.standard_normal((30_000, 400)).astype(np.float32)[:, ::2],  # non-contiguous
             rng.standard_normal((10, 10)).astype(np.float32)]
    for a in cases + cases[:1]:
        got = _lib.upload_f32(a, dev)
        ref = torch.from_numpy(np.ascontiguousarray(a)).to(dev, torch.float32)
        torch.cuda.synchronize()
        assert got.dtype == torch.float32 and got.is_contiguous() and got.shape == ref.shape
        assert torch.equal(got, ref)
