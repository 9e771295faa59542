"""GPU tests of designs wider than 96 columns through the drop-in API (the
segmented path: each pair's columns in 64-column segments of the projection
operand, FP64 prep, one wave per voxel group in the GLS).

Bars (as in test_gpu_parity): selected pair = the FP64 oracle's hierarchical
choice except FP64 near-ties (<= 2e-3); at the engine's pairs, against the
FP64 oracle refit: betas rel <= 1e-5 (of max(|beta|, 1e-3 sd)), t rel <= 1e-5,
F / sigma2 rel <= 1e-5, R^2 abs <= 1e-6, whitened residuals, residuals and
fitted values <= 1e-5 sd, Ljung-Box rel <= 1e-5.
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


def _case(seed=12, T=240, p=128, V=160):
    rng = np.random.default_rng(seed)
    X = np.column_stack([np.ones(T), np.linspace(-1, 1, T), rng.standard_normal((T, p - 2))])
    a = rng.uniform(0.1, 0.8, V)
    b = rng.uniform(-0.5, 0.3, V)
    e = np.zeros((V, T))
    eps_prev = rng.standard_normal(V)
    for t in range(T):
        eps = rng.standard_normal(V)
        e[:, t] = (a * e[:, t - 1] if t else 0) + eps + b * eps_prev
        eps_prev = eps
    Y = (rng.standard_normal((V, p)) @ X.T * 0.5 + 5.0 + e).astype(np.float32)
    return X, Y


def test_wide_design_fit_api_matches_oracle(dev):
    """fit_glm_arma11 on a 128-column design with a task subset (F over 32
    columns), residual series and Ljung-Box: the default selector (the
    hierarchical descent, on the full surface for these widths) and every
    output against the FP64 oracle; a refit from the selected (a, b) as
    precomputed_arma_params (a pair-list plan) gives the same statistics; the
    learned selector refuses the width."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[wide] fit_glm_arma11 p=128", flush=True)
    X, Y = _case()
    T, p = X.shape
    V = Y.shape[0]
    ag = bg = torch.linspace(-0.8, 0.8, 17)
    task = list(range(2, 34))
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, task_indices=task, want_residuals=True,
                         want_predicted=True, want_ljung_box=True)
    X64, Y64 = torch.from_numpy(X), torch.from_numpy(Y).double()
    gr = O.precompute_grid(X64, ag, bg, None, None, F64)
    s = Y64.std(dim=1).clamp(min=1e-8)
    Yn = Y64 / s[:, None]
    S64 = O.score_surface(X64, Yn, gr)
    k_ref, _ = O.select_hierarchical(S64.float(), gr.params)
    idx = res.pair_index.long()
    ar = torch.arange(V)
    diff = idx != k_ref
    gap = (S64[ar, idx] - S64[ar, k_ref]).abs()
    assert bool((gap[diff] <= 2e-3).all()), f"{int(diff.sum())} mismatches beyond near-ties"
    b, t, f, s2, r2, _, rw = O.gls_refit(X64, Yn, gr, idx, task_indices=task)
    sc = s[:, None]
    bt = b[:, task] * sc
    assert float(((res.betas.double() - bt).abs() / torch.maximum(bt.abs(), 1e-3 * sc)).max()) <= 1e-5
    tt = t[:, task]
    assert float(((res.tstats.double() - tt).abs() / tt.abs().clamp(min=1)).max()) <= 1e-5
    assert float(((res.fstats.double() - f).abs() / f.abs().clamp(min=1)).max()) <= 1e-5
    assert float(((res.sigma2.double() - s2 * s**2).abs() / (s2 * s**2)).max()) <= 1e-5
    assert float((res.r2.double() - r2).abs().max()) <= 1e-6
    pred = b @ X64.T
    assert float(((res.residuals_whitened.double() - rw * sc).abs() / sc).max()) <= 1e-5
    assert float(((res.predicted.double() - pred * sc).abs() / sc).max()) <= 1e-5
    assert float(((res.residuals.double() - (Yn - pred) * sc).abs() / sc).max()) <= 1e-5
    lb = O.ljung_box(rw, O.ljung_box_max_lag(T, p, T))
    assert float(((res.ljung_box.double() - lb).abs() / lb.abs().clamp(min=1)).max()) <= 1e-5

    res2 = fit_glm_arma11(Y, X, 2.0, device=dev, precomputed_arma_params=res.arma_params, task_indices=task)
    for name in ("betas", "tstats", "fstats", "sigma2"):
        got, want = getattr(res2, name).double(), getattr(res, name).double()
        assert float(((got - want).abs() / want.abs().clamp(min=1e-3)).max()) <= 1e-5, name

    with pytest.raises(NotImplementedError):
        fit_glm_arma11(Y[:8], X, 2.0, a_grid=ag, b_grid=bg, device=dev, selector="learned")


def test_wide_design_use_double_is_the_float64_search(dev):
    """use_double on a 128-column design: the float64 search on the raw
    series (the FP64 surface over 32-column chunks of each pair's table)
    against the FP64 oracle, with the bars of test_gpu_precision: scores within
    1e-6 + 1e-10 |score|, hierarchical pair identical except near-ties <= 1e-6,
    betas / t within 1e-5."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[wide] use_double p=128", flush=True)
    X, Y = _case(seed=13, V=64)
    V = Y.shape[0]
    ag = bg = torch.linspace(-0.8, 0.8, 9)
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, use_double=True, save_profile_likelihoods=True)
    ref = O.fit_arma11(torch.from_numpy(Y).double(), torch.from_numpy(X), ag, bg, dtype=F64, selector="hierarchical")
    err = (res.reml_lklhd_surface - ref.surface).abs()
    assert bool((err <= 1e-6 + 1e-10 * ref.surface.abs()).all()), f"max score error {float(err.max()):.3e}"
    diff = res.pair_index != ref.pair_index
    gap = (ref.surface[torch.arange(V), res.pair_index] - ref.surface[torch.arange(V), ref.pair_index]).abs()
    assert bool((gap[diff] <= 1e-6).all())
    ok = ~diff
    sc = torch.from_numpy(Y).double().std(dim=1)
    db = (res.betas - ref.betas).abs() / torch.maximum(ref.betas.abs(), 1e-3 * sc[:, None])
    assert float(db[ok].max()) <= 1e-5
    assert float(((res.tstats - ref.tstats).abs() / ref.tstats.abs().clamp(min=1))[ok].max()) <= 1e-5
