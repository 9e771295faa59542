"""GPU tests of the lambda-band search (learned path, low-confidence and
window-edge voxels).

* Band membership from the library equals a restatement here: pairs with
  |lambda_k - (-1 + i h)| <= w, lambda = (1 + ab)(a + b) / (1 + 2ab + b^2) in
  FP64; edge flag = a lattice 4-neighbour that is a pair outside the band.
* Band scores are bitwise equal to the full-grid surface (each projection
  mode), the band argmin is the first minimum of the surface restricted to the
  band (lambda == 0 pairs mapped to the first of them), need = the edge flag of
  that pair, and voxels not flagged (or with non-finite proposals) are left
  untouched.
* End to end: with bands the learned search sends fewer voxels to the full
  grid and still agrees with the global argmin on >= 99% of voxels.
"""

import ctypes
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

W, H = 0.14, 0.02


@pytest.fixture(scope="module")
def dev():
    from fastfuncstuff_amd import _lib

    _lib.load()
    return torch.device("cuda", 0)


def _lam(a, b):
    return (1.0 + a * b) * (a + b) / (1.0 + 2.0 * a * b + b * b)


def _grid(dev, V, T=120, seed=11, noise_seed=None, extra=0):
    from fastfuncstuff_amd import synth
    from fastfuncstuff_amd.glm.arma import RemlGrid

    wl = synth.plain_config(V=V, T=T, seed=seed, noise_seed=noise_seed)
    X = wl.X
    if extra:  # a wider design: random columns beside the plain one
        X = np.concatenate([X, np.random.default_rng(99).standard_normal((X.shape[0], extra))], 1)
    grid = RemlGrid(torch.from_numpy(X), torch.from_numpy(wl.a_grid), torch.from_numpy(wl.b_grid), dev)
    return grid, torch.from_numpy(wl.Y).to(dev)


def _bands_restated(grid):
    ab = np.array(grid.param_list, dtype=np.float64)
    lam = _lam(ab[:, 0], ab[:, 1])
    nb = int(math.floor(2.0 / H + 0.5)) + 1
    h = float(np.float32(H))  # the fp32 step the library uses
    lut = np.asarray(grid.lut)
    pos = {int(lut[x, y]): (x, y) for x in range(lut.shape[0]) for y in range(lut.shape[1]) if lut[x, y] >= 0}
    bands, edges = [], []
    for i in range(nb):
        c = -1.0 + i * h
        inb = np.abs(lam - c) <= W
        ks = np.nonzero(inb)[0]
        e = []
        for k in ks:
            x, y = pos[int(k)]
            out = False
            for xx, yy in ((x - 1, y), (x + 1, y), (x, y - 1), (x, y + 1)):
                if 0 <= xx < lut.shape[0] and 0 <= yy < lut.shape[1] and lut[xx, yy] >= 0 and not inb[lut[xx, yy]]:
                    out = True
            e.append(int(out))
        bands.append(ks)
        edges.append(np.array(e, dtype=np.int64))
    return bands, edges


def test_band_tables_match_restatement(dev):
    print("[band] tables", flush=True)
    grid, _ = _grid(dev, 64)
    grid.set_bands(W, H)
    pairs, cnt, edge = grid.band_info()
    bands, edges = _bands_restated(grid)
    assert pairs.shape[0] == len(bands)
    for i, (ks, e) in enumerate(zip(bands, edges)):
        n = int(cnt[i])
        assert n == len(ks), i
        assert np.array_equal(pairs[i, :n].numpy(), ks), i
        assert bool((pairs[i, n:] == -1).all())
        assert np.array_equal(edge[i, :n].numpy(), e), i
    grid.set_bands(0.0, H)
    assert grid.band_info() is None


@pytest.mark.parametrize("mode,extra", [("fp32", 0), ("fp16x3", 0), ("bf16x3", 0), ("fp16x3", 34), ("bf16x3", 62)])
def test_band_scores_bitwise_equal_full_grid(dev, mode, extra):
    """Band scores equal the full-grid surface bitwise; also for p = 40 and 68
    (split halves from the f64-MFMA prep for 32 < p <= 96)."""
    from fastfuncstuff_amd import _lib

    print(f"[band] bitwise mode={mode} extra={extra}", flush=True)
    lib = _lib.load()
    grid, Y = _grid(dev, 900, extra=extra)
    grid.set_projection(mode)
    grid.set_bands(W, H)
    V, K = Y.shape[0], grid.K
    Yt, scale_ref, _ = grid.prep(Y)
    S = grid.surface(Yt, V).cpu()
    rng = np.random.default_rng(5)
    # proposals over and beyond the lattice box, some non-finite
    a = rng.uniform(-1.2, 1.2, V)
    b = rng.uniform(-1.2, 1.2, V)
    a[::37] = np.nan
    b[5::41] = np.inf
    props = torch.tensor(np.stack([a, b, np.zeros(V)], 1), dtype=torch.float32, device=dev)
    need_in = torch.from_numpy(rng.choice([0, 1, 2], V, p=[0.2, 0.4, 0.4]).astype(np.int32)).to(dev)
    need = need_in.clone()
    best = torch.full((V,), -7, dtype=torch.int32, device=dev)
    score = torch.full((V,), 123.0, dtype=torch.float32, device=dev)
    scale = torch.full((V,), -1.0, dtype=torch.float32, device=dev)
    nb = ctypes.c_int64()
    _lib.check(lib.ffs_reml_band_workspace(grid.handle, V, ctypes.byref(nb)), "ws")
    ws = torch.zeros(int(nb.value), dtype=torch.uint8, device=dev)
    _lib.check(lib.ffs_reml_band_search(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(props), _lib.ptr(best),
                                        _lib.ptr(score), _lib.ptr(need), _lib.ptr(scale), None, _lib.ptr(ws),
                                        ws.numel(), _lib.stream_handle(dev)), "band_search")
    best, score, need, scale = best.cpu().long(), score.cpu(), need.cpu(), scale.cpu()
    need_in = need_in.cpu()
    bands, edges = _bands_restated(grid)
    white = [k for k, (pa, pb) in enumerate(grid.param_list) if (np.float32(pa) + np.float32(pb)) == 0]
    a_ax, b_ax = np.array(grid.a_axis, dtype=np.float64), np.array(grid.b_axis, dtype=np.float64)
    n_band = 0
    for v in range(V):
        ah, bh = np.float32(a[v]), np.float32(b[v])
        if int(need_in[v]) == 0 or not (np.isfinite(ah) and np.isfinite(bh)):
            assert int(best[v]) == -7 and score[v].item() == 123.0 and int(need[v]) == int(need_in[v]), v
            continue
        ac = float(min(max(ah, np.float32(a_ax.min())), np.float32(a_ax.max())))
        bc = float(min(max(bh, np.float32(b_ax.min())), np.float32(b_ax.max())))
        lam = min(max(_lam(ac, bc), -1.0), 1.0)
        i = min(max(int(np.rint((lam + 1.0) / float(np.float32(H)))), 0), len(bands) - 1)
        ks = bands[i]
        if len(ks) == 0:
            assert int(best[v]) == -7 and int(need[v]) == int(need_in[v]), v
            continue
        n_band += 1
        sv = S[v, torch.from_numpy(ks)]
        j = int(torch.argmin(sv))  # first minimum in slot (= pair) order
        k = int(ks[j])
        assert score[v].item() == S[v, k].item(), v  # bitwise
        if k in white:
            k = white[0]
        assert int(best[v]) == k, v
        assert int(need[v]) == int(edges[i][j]), v
        assert scale[v].item() == scale_ref[v].item(), v
    assert n_band > V // 3


def test_band_search_refused_without_bands(dev):
    from fastfuncstuff_amd import _lib

    grid, Y = _grid(dev, 64)
    lib = _lib.load()
    rc = lib.ffs_reml_band_energy(grid.handle, 64, None, 0, _lib.stream_handle(dev))
    assert rc == -4  # no bands on this plan


def test_learned_search_with_bands(dev):
    from fastfuncstuff_amd.learned import LocalSearcher, ProposalMLP, calibrate_conf_min, features, full_grid_labels

    print("[band] learned end to end", flush=True)
    grid, Ytr = _grid(dev, 100_000, T=300, seed=21)
    _, Yte = _grid(dev, 20_000, T=300, seed=21, noise_seed=22)
    _, Yva = _grid(dev, 20_000, T=300, seed=21, noise_seed=23)
    grid.set_projection("fp16x3")
    lab, _ = full_grid_labels(grid, Ytr)
    mlp = ProposalMLP(dev, seed=0)
    mlp.train(features(grid, Ytr), lab, epochs=12, batch=2048, lr=3e-3, log_every=50)
    _, best_va = full_grid_labels(grid, Yva)
    calibrate_conf_min(mlp, grid, Yva, best_va, target=0.999)
    _, best_te = full_grid_labels(grid, Yte)
    res0 = LocalSearcher(grid, mlp, Yte.shape[0]).run(Yte)
    grid.set_bands(W, H)
    stats = {}
    res1 = LocalSearcher(grid, mlp, Yte.shape[0]).run(Yte, stats=stats)
    a0 = float((res0.best == best_te).double().mean())
    a1 = float((res1.best == best_te).double().mean())
    print(f"[band] agreement {a0:.5f} -> {a1:.5f}, full grid {res0.n_fallback} -> {res1.n_fallback}, "
          f"band voxels {stats.get('band')}", flush=True)
    assert a1 >= 0.99
    assert res1.n_fallback < res0.n_fallback


@pytest.mark.parametrize("mode,use_coef,extra", [("fp32", False, 0), ("fp16x3", False, 0), ("fp16x3", True, 0),
                                                ("bf16x3", False, 62)])
def test_subset_search_matches_exhaustive(dev, mode, use_coef, extra):
    """Full-grid search of a device-counted subset: the first `cap` listed
    voxels get the exhaustive first minimum of the full surface (bitwise
    score); the others are untouched.  use_coef: the residualisation
    coefficients and scales come from ffs_reml_prep_features instead.
    extra = 62: a p = 68 design (gathered columns through the f64-MFMA prep
    for 32 < p <= 96 with a device-side column count)."""
    from fastfuncstuff_amd import _lib

    print(f"[band] subset search mode={mode} coef={use_coef} extra={extra}", flush=True)
    lib = _lib.load()
    grid, Y = _grid(dev, 900, extra=extra)
    grid.set_projection(mode)
    V = Y.shape[0]
    Yt, scale_ref, _ = grid.prep(Y)
    S = grid.surface(Yt, V)
    b_ref, s_ref = grid.select(S, "exhaustive")
    b_ref, s_ref = b_ref.cpu(), s_ref.cpu()
    g = torch.Generator().manual_seed(3)
    idx = torch.randperm(V, generator=g)[:300].to(torch.int32)
    cap = 256
    idx_d = torch.full((V,), -5, dtype=torch.int32, device=dev)
    idx_d[:300] = idx.to(dev)
    count = torch.tensor([300], dtype=torch.int32, device=dev)
    best = torch.full((V,), -7, dtype=torch.int32, device=dev)
    score = torch.full((V,), 123.0, dtype=torch.float32, device=dev)
    scale = torch.full((V,), -1.0, dtype=torch.float32, device=dev)
    coef = None
    if use_coef:
        stride = ctypes.c_int32()
        _lib.check(lib.ffs_reml_plan_coef_stride(grid.handle, ctypes.byref(stride)), "stride")
        assert stride.value == 16
        coef = torch.full((V, stride.value), float("nan"), dtype=torch.float64, device=dev)
        feats = torch.empty(V, 10, dtype=torch.float32, device=dev)
        _lib.check(lib.ffs_reml_prep_features(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(scale), 1,
                                              _lib.ptr(feats), _lib.ptr(coef), _lib.stream_handle(dev)), "features")
        assert torch.equal(scale.cpu(), scale_ref.cpu())
        assert bool(torch.isfinite(coef).all())
    nb = ctypes.c_int64()
    _lib.check(lib.ffs_reml_subset_workspace(grid.handle, cap, ctypes.byref(nb)), "ws")
    ws = torch.empty(int(nb.value), dtype=torch.uint8, device=dev)
    _lib.check(lib.ffs_reml_subset_search(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(idx_d), _lib.ptr(count),
                                          cap, _lib.ptr(best), _lib.ptr(score), _lib.ptr(scale),
                                          _lib.ptr(coef) if use_coef else None, _lib.ptr(ws), ws.numel(),
                                          _lib.stream_handle(dev)), "subset")
    best, score, scale = best.cpu(), score.cpu(), scale.cpu()
    done = idx[:cap].long()
    assert torch.equal(best[done], b_ref[done])
    assert torch.equal(score[done], s_ref[done])  # bitwise
    assert torch.equal(scale[done], scale_ref.cpu()[done])
    rest = torch.ones(V, dtype=torch.bool)
    rest[done] = False
    assert bool((best[rest] == -7).all()) and bool((score[rest] == 123.0).all())
    if not use_coef:
        assert bool((scale[rest] == -1.0).all())
