"""GPU tests of the learned-proposal path.

* MLP forward / gradient / Adam kernels vs the FP64 oracle (oracle/mlp_oracle.py):
  forward rtol 1e-5; loss rtol 1e-5; gradient |g - g64| <= 1e-3 |g64| + 1e-4 max|g64|;
  Adam rtol 1e-6.
* Residual-autocorrelation features vs oracle.residual_acf: |d| <= 2e-5.
* Window scores are bitwise equal to the full-grid surface in each projection
  mode, and the window argmin equals the argmin of the surface restricted to
  the window.
* Training on the full-grid labels beats an untrained network: fewer pair
  evaluations per voxel and >= 99% agreement with the global argmin.
"""

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from fastfuncstuff_amd import _lib

    _lib.load()
    return torch.device("cuda", 0)


@pytest.mark.parametrize("T", [80, 200, 400])
def test_features_match_oracle(dev, T):
    """Residual autocorrelation features (the time range of each 64-voxel group
    is split over the waves of its block) vs the FP64 oracle: |d| <= 2e-5."""
    from fastfuncstuff_amd import synth
    from fastfuncstuff_amd.glm.arma import RemlGrid
    from fastfuncstuff_amd.learned import features

    wl = synth.plain_config(V=300, T=T, seed=11)
    grid = RemlGrid(torch.from_numpy(wl.X), torch.from_numpy(wl.a_grid), torch.from_numpy(wl.b_grid), dev)
    got = features(grid, torch.from_numpy(wl.Y).to(dev)).cpu().numpy().astype(np.float64)
    ref = M.residual_acf(wl.Y, wl.X, M.N_IN)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 2e-5


def test_mlp_forward_grad_adam_match_oracle(dev):
    from fastfuncstuff_amd.learned import ProposalMLP

    rng = np.random.default_rng(3)
    theta = M.init_params(4) + 0.05 * rng.standard_normal(M.N_PARAMS)
    n, B = 1000, 257
    X = rng.uniform(-0.6, 0.95, (n, M.N_IN))
    tgt = rng.uniform(-0.8, 0.8, (n, 2))
    idx = rng.choice(n, B, replace=False).astype(np.int32)
    mlp = ProposalMLP(dev, params=theta.astype(np.float32))
    theta32 = theta.astype(np.float32).astype(np.float64)
    Xd = torch.from_numpy(X.astype(np.float32)).to(dev)
    Td = torch.from_numpy(tgt.astype(np.float32)).to(dev)
    Id = torch.from_numpy(idx).to(dev)
    X32 = X.astype(np.float32).astype(np.float64)
    T32 = tgt.astype(np.float32).astype(np.float64)

    out = mlp.forward(Xd, Id).cpu().double().numpy()
    ref, _ = M.forward(theta32, X32[idx])
    np.testing.assert_allclose(out, ref, rtol=1e-5, atol=1e-5)

    loss, grad = mlp.loss_and_grad(Xd, Td, Id)
    l_ref, g_ref = M.loss_and_grad(theta32, X32[idx], T32[idx])
    assert abs(float(loss.item()) - l_ref) <= 1e-5 * max(1.0, abs(l_ref))
    g = grad.cpu().double().numpy()
    assert np.all(np.abs(g - g_ref) <= 1e-3 * np.abs(g_ref) + 1e-4 * np.abs(g_ref).max())

    # one Adam step from zero moments
    from fastfuncstuff_amd import _lib

    lib = _lib.load()
    m = torch.zeros_like(mlp.params)
    v = torch.zeros_like(mlp.params)
    g32 = grad.clone()
    _lib.check(lib.ffs_mlp_adam(_lib.ptr(mlp.params), _lib.ptr(g32), _lib.ptr(m), _lib.ptr(v), mlp.n_params, 1e-3,
                                0.9, 0.999, 1e-8, 1, _lib.stream_handle(dev)), "adam")
    th_ref, _, _ = M.adam_step(theta32, g32.cpu().double().numpy(), np.zeros(M.N_PARAMS), np.zeros(M.N_PARAMS), 1,
                               lr=1e-3)
    np.testing.assert_allclose(mlp.params.cpu().double().numpy(), th_ref, rtol=1e-6, atol=1e-7)


def test_train_steps_equal_grad_and_adam_steps(dev):
    """ffs_mlp_train_steps (an epoch of steps launched from the library, the
    reduction and Adam fused) is bitwise the sequence ffs_mlp_grad +
    ffs_mlp_adam at the same learning-rate schedule, losses included."""
    from fastfuncstuff_amd import _lib

    lib = _lib.load()
    st = _lib.stream_handle(dev)
    rng = np.random.default_rng(5)
    n, B, steps, total = 3000, 500, 5, 9
    theta = (M.init_params(6) + 0.05 * rng.standard_normal(M.N_PARAMS)).astype(np.float32)
    X = torch.from_numpy(rng.uniform(-0.6, 0.95, (n, M.N_IN)).astype(np.float32)).to(dev)
    Tg = torch.from_numpy(rng.uniform(-0.8, 0.8, (n, 2)).astype(np.float32)).to(dev)
    perm = torch.from_numpy(rng.permutation(n).astype(np.int32)).to(dev)
    ws = torch.empty(int(lib.ffs_mlp_workspace_floats(B)), dtype=torch.float32, device=dev)
    lr, lr_final, step0 = 3e-3, 1e-4, 2

    def fresh():
        p = torch.from_numpy(theta.copy()).to(dev)
        return p, torch.zeros_like(p), torch.zeros_like(p)

    p1, m1, v1 = fresh()
    losses = torch.empty(steps, dtype=torch.float32, device=dev)
    _lib.check(lib.ffs_mlp_train_steps(_lib.ptr(p1), _lib.ptr(m1), _lib.ptr(v1), _lib.ptr(X), _lib.ptr(Tg),
                                       _lib.ptr(perm), B, steps, step0, total, lr, lr_final, 0.9, 0.999, 1e-8,
                                       _lib.ptr(losses), _lib.ptr(ws), st), "train_steps")
    p2, m2, v2 = fresh()
    g = torch.empty_like(p2)
    loss = torch.empty(1, dtype=torch.float32, device=dev)
    ref_losses = []
    for i in range(steps):
        step = step0 + i + 1
        lr_t = float(np.float32(lr) * np.float32(np.float32(lr_final) / np.float32(lr))
                     ** np.float32((step - 1) / (total - 1)))
        _lib.check(lib.ffs_mlp_grad(_lib.ptr(p2), _lib.ptr(X), _lib.ptr(Tg), _lib.ptr(perm[i * B:(i + 1) * B]), B,
                                    _lib.ptr(g), _lib.ptr(loss), _lib.ptr(ws), st), "grad")
        ref_losses.append(float(loss.item()))
        _lib.check(lib.ffs_mlp_adam(_lib.ptr(p2), _lib.ptr(g), _lib.ptr(m2), _lib.ptr(v2), M.N_PARAMS, lr_t, 0.9,
                                    0.999, 1e-8, step, st), "adam")
    assert losses.cpu().tolist() == ref_losses
    # the learning rate is formed in float32 on both sides (powf vs numpy
    # float32 power may differ in the last bit): parameters to 1 ulp-level
    torch.testing.assert_close(p1, p2, rtol=2e-6, atol=1e-8)
    torch.testing.assert_close(m1, m2, rtol=0, atol=0)


def _grid_and_data(dev, V, T=120, seed=0, noise_seed=None):
    from fastfuncstuff_amd import synth
    from fastfuncstuff_amd.glm.arma import RemlGrid

    wl = synth.plain_config(V=V, T=T, seed=seed, noise_seed=noise_seed)
    grid = RemlGrid(torch.from_numpy(wl.X), torch.from_numpy(wl.a_grid), torch.from_numpy(wl.b_grid), dev)
    return grid, torch.from_numpy(wl.Y).to(dev), wl


@pytest.mark.parametrize("mode,use_coef", [("fp32", False), ("fp16x3", False), ("bf16x3", False),
                                           ("fp32", True), ("fp16x3", True)])
def test_window_scores_bitwise_equal_full_grid(dev, mode, use_coef):
    """use_coef: the window preparation takes each voxel's residualisation
    coefficients and scale from ffs_reml_prep_features (no first pass)."""
    from fastfuncstuff_amd.learned import LocalSearcher, ProposalMLP

    print(f"[learned] window bitwise mode={mode} coef={use_coef}", flush=True)
    grid, Y, _ = _grid_and_data(dev, 700, seed=11)
    grid.set_projection(mode)
    V = Y.shape[0]
    Yt, _, _ = grid.prep(Y)
    S = grid.surface(Yt, V).cpu()
    # proposals: exact lattice points, spread over the grid (some clamped at edges)
    rng = np.random.default_rng(1)
    ai = rng.integers(0, len(grid.a_axis), V)
    bi = rng.integers(0, len(grid.b_axis), V)
    props = torch.tensor(np.stack([np.array(grid.a_axis)[ai], np.array(grid.b_axis)[bi], np.full(V, 50.0)], 1),
                         dtype=torch.float32, device=dev)

    class Fixed(ProposalMLP):
        def forward(self, feats, idx=None):
            return props

    mlp = Fixed(dev)
    ls = LocalSearcher(grid, mlp, V)
    from fastfuncstuff_amd import _lib

    lib = _lib.load()
    best = torch.empty(V, dtype=torch.int32, device=dev)
    score = torch.empty(V, dtype=torch.float32, device=dev)
    need = torch.empty(V, dtype=torch.int32, device=dev)
    scale = torch.empty(V, dtype=torch.float32, device=dev)
    coef = None
    if use_coef:
        assert ls.coef is not None
        feats = torch.empty(V, 10, dtype=torch.float32, device=dev)
        _lib.check(lib.ffs_reml_prep_features(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(scale), 1,
                                              _lib.ptr(feats), _lib.ptr(ls.coef), _lib.stream_handle(dev)), "feats")
        coef = _lib.ptr(ls.coef)
    _lib.check(lib.ffs_reml_local_search(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(props), -1e30,
                                         _lib.ptr(best), _lib.ptr(score), _lib.ptr(need), _lib.ptr(scale), coef,
                                         _lib.ptr(ls.ws), ls.ws.numel(), _lib.stream_handle(dev)), "local")
    best, score = best.cpu().long(), score.cpu()
    lut = grid.lut
    na, nb = lut.shape
    wp, wc, _ = grid.window_info()
    white = [k for k, (a, b) in enumerate(grid.param_list) if (np.float32(a) + np.float32(b)) == 0]
    for v in range(V):
        ca = min(max(ai[v], 2), na - 3)
        cb = min(max(bi[v], 2), nb - 3)
        c = ca * nb + cb
        ks = wp[c, :int(wc[c])].tolist()
        if not ks:  # window entirely in the lambda < 0 region: no local optimum
            assert int(best[v]) == -1 and score[v].item() == float("inf")
            continue
        sv = S[v, ks]
        j = int(torch.argmin(sv))  # first minimum in slot (= pair) order
        k = ks[j]
        if k in white:
            k = white[0]
        assert int(best[v]) == k, v
        assert score[v].item() == S[v, ks[j]].item()  # bitwise


def _window_restatement(grid):
    """Windows of the learned search restated: the valid pairs of the 5 x 5
    lattice cells around each cell, in pair order; edge = the slot is on a
    window side that the lattice continues past."""
    lut = np.asarray(grid.lut)
    na, nb = lut.shape
    out = []
    for ca in range(na):
        for cb in range(nb):
            ks, edge = [], []
            for x in range(ca - 2, ca + 3):
                for y in range(cb - 2, cb + 3):
                    if 0 <= x < na and 0 <= y < nb and lut[x, y] >= 0:
                        ks.append(int(lut[x, y]))
                        edge.append(int((x == ca - 2 and x > 0) or (x == ca + 2 and x < na - 1)
                                        or (y == cb - 2 and y > 0) or (y == cb + 2 and y < nb - 1)))
            out.append((ks, edge))
    return out


def test_window_tables_match_restatement(dev):
    """The plan's window table equals the restated 5 x 5 windows; every
    window holds at most 25 pairs, in pair order, and contains its centre
    when the centre is a valid pair."""
    grid, _, _ = _grid_and_data(dev, 64, seed=11)
    wp, wc, we = grid.window_info()
    ref = _window_restatement(grid)
    lut = np.asarray(grid.lut)
    nb = lut.shape[1]
    assert wp.shape[0] == len(ref)
    for c, (ks, edge) in enumerate(ref):
        n = int(wc[c])
        assert wp[c, :n].tolist() == ks, c
        assert we[c, :n].tolist() == edge, c
        assert n <= 25 and ks == sorted(ks)
        if lut[c // nb, c % nb] >= 0:
            assert int(lut[c // nb, c % nb]) in ks


def test_trained_network_gains_over_untrained(dev):
    from fastfuncstuff_amd.learned import LocalSearcher, ProposalMLP, features, full_grid_labels

    grid, Ytr, _ = _grid_and_data(dev, 100_000, T=300, seed=21)
    _, Yte, _ = _grid_and_data(dev, 20_000, T=300, seed=21, noise_seed=22)
    lab, _ = full_grid_labels(grid, Ytr)
    f = features(grid, Ytr)
    _, best_te = full_grid_labels(grid, Yte)
    trained = ProposalMLP(dev, seed=0)
    log = trained.train(f, lab, epochs=12, batch=2048, lr=3e-3, log_every=20)
    assert log.losses[-1][1] < log.losses[0][1]
    untrained = ProposalMLP(dev, seed=0)
    K = grid.K
    evals = {}
    for name, net in (("trained", trained), ("untrained", untrained)):
        res = LocalSearcher(grid, net, Yte.shape[0]).run(Yte)
        agree = float((res.best == best_te).double().mean())
        frac = res.n_fallback / Yte.shape[0]
        evals[name] = 25 + frac * K
        print(f"[learned] {name}: agreement {agree:.5f}, fallback {frac:.4f}, pair evals/voxel {evals[name]:.1f}")
        assert agree >= 0.99
    assert evals["trained"] <= 0.6 * evals["untrained"]
