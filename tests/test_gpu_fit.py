"""GPU tests of the fit API (fit_glm_arma11 / fit_glm_arma11_grouped) against
the oracle's restatement of the reference fit.

Tolerances: selected pair equal to the float32 reference-semantics fit's except
near-ties (score gap <= 2e-3); at the engine's pairs, against the FP64 oracle
refit: betas rel <= 1e-5 (of max(|beta|, 1e-3 sd)), t rel <= 1e-5; partial /
semi-partial R^2 abs <= 1e-4 (they come from the float32 t-statistics, and the
"task" mode divides by 1 - sum(nuisance) >= 0.01).
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


def _data(rng, T, X, V):
    a = rng.uniform(0.1, 0.8, V)
    b = rng.uniform(-0.5, 0.3, V)
    e = np.zeros((V, T))
    eps_prev = rng.standard_normal(V)
    for t in range(T):
        eps = rng.standard_normal(V)
        e[:, t] = (a * e[:, t - 1] if t else 0) + eps + b * eps_prev
        eps_prev = eps
    return (rng.standard_normal((V, X.shape[1])) @ X.T * 0.5 + 20.0 + e).astype(np.float32)


def _design(rng, T, shift=0):
    box = np.zeros(T)
    for s in range(5 + shift, T, 23):
        box[s:s + 6] = 1.0
    return np.column_stack([box, np.roll(box, 3 + shift), np.ones(T), np.linspace(-1, 1, T),
                            rng.standard_normal(T) * 0.1])


def _compare(res, ref, Y, X, ag, bg, task=None):
    """Selection vs the float32 reference-semantics fit (near-ties allowed);
    betas / t vs the FP64 oracle refit at the engine's pairs.  Returns the
    FP64 refit's full t-statistics and R^2."""
    V = Y.shape[0]
    idx = res.pair_index.long()
    same = idx == ref.pair_index.long()
    gap = ref.surface[torch.arange(V), idx] - ref.surface[torch.arange(V), ref.pair_index.long()]
    assert bool((gap[~same].abs() <= 2e-3).all()), f"{int((~same).sum())} mismatches beyond near-ties"
    s = torch.from_numpy(Y).double().std(dim=1).clamp(min=1e-8)
    g64 = O.precompute_grid(torch.from_numpy(X).double(), ag, bg, None, None, F64)
    b64, t64, _, _, r2, _, _ = O.gls_refit(torch.from_numpy(X).double(), torch.from_numpy(Y).double() / s[:, None],
                                           g64, idx, dtype=F64)
    b64 = b64 * s[:, None]
    bt, tt = (b64[:, task], t64[:, task]) if task else (b64, t64)
    denom = torch.maximum(bt.abs(), 1e-3 * s[:, None])
    assert float(((res.betas.double() - bt).abs() / denom).max()) <= 1e-5
    assert float(((res.tstats.double() - tt).abs() / tt.abs().clamp(min=1)).max()) <= 1e-5
    return t64, r2


def test_fit_partial_r2_matches_oracle(dev):
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[fit] partial r2", flush=True)
    rng = np.random.default_rng(5)
    T, V = 110, 240
    X = _design(rng, T)
    Y = _data(rng, T, X, V)
    ag = bg = torch.linspace(-0.8, 0.8, 17)
    task = [0, 1]
    for pm, sm in (("full", "full"), ("task", "task")):
        res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, task_indices=task, want_r2_partial=True,
                             r2_partial_mode=pm, want_r2_semipartial=True, r2_semipartial_mode=sm)
        ref = O.fit_arma11(torch.from_numpy(Y), torch.from_numpy(X).float(), ag, bg, dtype=torch.float32,
                           task_indices=task)
        t_full, r2 = _compare(res, ref, Y, X, ag, bg, task)
        # oracle partial R^2 from the FP64 refit's full t-statistics and R^2
        part, nu, semi, sn = O.partial_r2(t_full, r2.double(), res.dof, task, pm, sm)
        for got, want in ((res.r2_partial, part), (res.r2_partial_nuisance, nu), (res.r2_semipartial, semi),
                          (res.r2_semipartial_nuisance, sn)):
            assert float((got.double() - want).abs().max()) <= 1e-4


def test_grouped_fit_matches_per_group_oracle(dev):
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11_grouped

    print("[fit] grouped", flush=True)
    rng = np.random.default_rng(6)
    T, V = 100, 300
    designs = {g: _design(rng, T, shift=g) for g in (0, 3, 7)}
    gidx = rng.choice([0, 3, 7], V, p=[0.6, 0.3, 0.1])
    Y = np.empty((V, T), dtype=np.float32)
    for g, X in designs.items():
        m = gidx == g
        Y[m] = _data(rng, T, X, int(m.sum()))
    ag = bg = torch.linspace(-0.8, 0.8, 17)
    res = fit_glm_arma11_grouped(Y, {g: torch.from_numpy(X) for g, X in designs.items()}, torch.from_numpy(gidx),
                                 2.0, group_label="HRF", a_grid=ag, b_grid=bg, device=dev)
    assert res.betas.shape == (V, 5) and res.sigma2.shape == (V,)
    for g, X in designs.items():
        m = gidx == g
        ref = O.fit_arma11(torch.from_numpy(Y[m]), torch.from_numpy(X).float(), ag, bg, dtype=torch.float32)

        class Sub:
            pass

        sub = Sub()
        for name in ("pair_index", "betas", "tstats"):
            setattr(sub, name, getattr(res, name)[torch.from_numpy(m)])
        _compare(sub, ref, Y[m], X, ag, bg)
    with pytest.raises(KeyError):
        fit_glm_arma11_grouped(Y, {0: designs[0]}, torch.from_numpy(gidx), 2.0, device=dev)
    with pytest.raises(ValueError):
        fit_glm_arma11_grouped(Y, designs, torch.from_numpy(gidx[:-1]), 2.0, device=dev)


def test_rvar_write_load_refit(dev, tmp_path):
    """-Rvar round trip (glm/arma.py:6181-6454): the (a, b) written by
    save_arma_rvar and read back by load_arma_params reproduce the fit when
    passed as precomputed_arma_params; the -Rlklhd surface file holds the
    fit's surface with one labelled sub-brick per pair."""
    from fastfuncstuff_amd import outputs
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[fit] rvar round trip", flush=True)
    rng = np.random.default_rng(8)
    T, V = 90, 60
    X = _design(rng, T)
    Y = _data(rng, T, X, V)
    ag = bg = torch.linspace(-0.8, 0.8, 17)
    res = fit_glm_arma11(Y.reshape(3, 4, 5, T), X, 2.0, a_grid=ag, b_grid=bg, device=dev, want_ljung_box=True,
                         save_profile_likelihoods=True)
    f = outputs.save_arma_rvar(res, tmp_path / "rvar.nii.gz", volume_shape=(3, 4, 5))
    img = outputs.read_nifti(f)
    assert img.labels == ["a", "b", "lam", "StDev", "-LogLik", "LjungBox"]
    assert img.stataux[5] == (6, (float(res.ljung_box_dof),))
    ab = outputs.load_arma_params(f)
    assert np.array_equal(ab, res.arma_params.numpy())
    res2 = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, precomputed_arma_params=ab)
    b1, b2 = res.betas.reshape(V, -1).double(), res2.betas.double()
    assert float(((b2 - b1).abs() / b1.abs().clamp(min=1e-3)).max()) <= 1e-6
    lk = outputs.read_nifti(outputs.save_reml_likelihood_surface(res, tmp_path / "lk.nii", volume_shape=(3, 4, 5)))
    assert len(lk.labels) == res.reml_lklhd_surface.shape[1]
    assert np.array_equal(lk.data.reshape(V, -1), res.reml_lklhd_surface.float().numpy())


def test_glt_contrasts_match_oracle(dev):
    """In-loop single-row GLTs (glm/arma.py:4201-4253, 5421-5509) against the
    FP64 oracle at the engine's pairs: contrast estimates rel <= 1e-5 (of
    max(|cb|, 1e-3 sd)), t rel <= 1e-5, partial / semi-partial R^2 <= 1e-4;
    multi-row GLTs raise as in the reference."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[fit] glt", flush=True)
    rng = np.random.default_rng(9)
    T, V = 100, 200
    X = _design(rng, T)
    Y = _data(rng, T, X, V)
    ag = bg = torch.linspace(-0.8, 0.8, 17)
    C = [np.array([1.0, -1.0, 0, 0, 0]), np.array([[0.5, 0.5, 0, 0, 0]])]
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, glt_labels=["diff", "mean"],
                         glt_matrices=C, want_r2_partial=True, want_r2_semipartial=True)
    assert res.contrast_labels == ["diff", "mean"] and res.contrast_betas.shape == (V, 2)
    idx = res.pair_index.long()
    s = torch.from_numpy(Y).double().std(dim=1).clamp(min=1e-8)
    g64 = O.precompute_grid(torch.from_numpy(X).double(), ag, bg, None, None, F64)
    Yn = torch.from_numpy(Y).double() / s[:, None]
    b64, _, _, s2, r2, _, _ = O.gls_refit(torch.from_numpy(X).double(), Yn, g64, idx, dtype=F64)
    Cm = torch.tensor(np.stack([C[0], C[1][0]]))
    cb, ct, part, semi = O.glt_contrasts(torch.from_numpy(X).double(), g64, idx, b64, s2, Cm, T - 5, r2)
    cb = cb * s[:, None]
    den = torch.maximum(cb.abs(), 1e-3 * s[:, None])
    assert float(((res.contrast_betas.double() - cb).abs() / den).max()) <= 1e-5
    assert float(((res.contrast_tstats.double() - ct).abs() / ct.abs().clamp(min=1)).max()) <= 1e-5
    assert float((res.contrast_r2_partial.double() - part).abs().max()) <= 1e-4
    assert float((res.contrast_r2_semipartial.double() - semi).abs().max()) <= 1e-4
    with pytest.raises(NotImplementedError):
        fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, glt_labels=["f"],
                       glt_matrices=[np.eye(5)[:2]])


def test_column_sets_match_oracle(dev):
    """-nuisance_summary column sets (glm/arma.py:3543-3585, 4190-4196,
    5522-5527, 5845-5848) against the FP64 oracle at the engine's pairs: Wald F
    rel <= 1e-5, fitted-signal SD and one-column coefficients rel <= 1e-5 (of
    max(|x|, 1e-3 sd)); NaN where the reference leaves NaN."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[fit] column sets", flush=True)
    rng = np.random.default_rng(13)
    T, V = 100, 200
    X = _design(rng, T)
    Y = _data(rng, T, X, V)
    ag = bg = torch.linspace(-0.8, 0.8, 17)
    sets = [[0], [1, 2], [2, 3, 4], [0, 9]]
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, column_sets=sets, task_indices=[0, 1])
    assert res.set_fstats.shape == (V, 4) and res.betas.shape == (V, 2)
    idx = res.pair_index.long()
    s = torch.from_numpy(Y).double().std(dim=1).clamp(min=1e-8)
    g64 = O.precompute_grid(torch.from_numpy(X).double(), ag, bg, None, None, F64)
    Yn = torch.from_numpy(Y).double() / s[:, None]
    b64, _, _, s2, _, _, _ = O.gls_refit(torch.from_numpy(X).double(), Yn, g64, idx, dtype=F64)
    F, sd, coef = O.column_set_stats(torch.from_numpy(X).double(), g64, idx, b64, s2, sets)
    sd, coef = sd * s[:, None], coef * s[:, None]
    ok = torch.tensor([True, True, True, False])
    assert float(((res.set_fstats.double()[:, ok] - F[:, ok]).abs() / F[:, ok].abs().clamp(min=1)).max()) <= 1e-5
    den = torch.maximum(sd[:, ok].abs(), 1e-3 * s[:, None])
    assert float(((res.set_sd.double()[:, ok] - sd[:, ok]).abs() / den).max()) <= 1e-5
    den0 = torch.maximum(coef[:, 0].abs(), 1e-3 * s)
    assert float(((res.set_coef.double()[:, 0] - coef[:, 0]).abs() / den0).max()) <= 1e-5
    assert bool(torch.isnan(res.set_fstats[:, 3]).all()) and bool(torch.isnan(res.set_coef[:, 1:]).all())


def test_dsort_pass_matches_oracle(dev):
    """-dsort voxel-wise regressors (glm/arma.py:3248-3539): the final GLS per
    voxel on [X | D_v] at the pair selected without them, whitened on the
    device, against the FP64 oracle at the engine's pairs: base and dsort
    betas rel <= 1e-5 (of max(|b|, 1e-3 sd)), t and F rel <= 1e-5, sigma2 rel
    <= 1e-6, R^2 <= 1e-6; dof drops by the dsort count; the search (pairs) is
    the one without dsort."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[fit] dsort", flush=True)
    rng = np.random.default_rng(17)
    T, V, q = 100, 150, 2
    X = _design(rng, T)
    Y = _data(rng, T, X, V)
    D = rng.standard_normal((V, q, T)).astype(np.float32)
    Y = Y + 0.5 * D[:, 0, :]
    ag = bg = torch.linspace(-0.8, 0.8, 17)
    base = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev)
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, dsort=D, dsort_labels=["wm", "csf"],
                         task_indices=[0, 1], column_sets=[[0], [5], [2, 6]])
    assert torch.equal(res.pair_index, base.pair_index)
    assert res.dof == T - 5 - q and res.dsort_labels == ["wm", "csf"]
    idx = res.pair_index.long()
    s = torch.from_numpy(Y).double().std(dim=1).clamp(min=1e-8)
    g64 = O.precompute_grid(torch.from_numpy(X).double(), ag, bg, None, None, F64)
    Yn = torch.from_numpy(Y).double() / s[:, None]
    b, t, s2, r2, F, _, _, Ai = O.dsort_gls(torch.from_numpy(X).double(), torch.from_numpy(D).double(), Yn, g64,
                                            idx, task_indices=[0, 1])
    bu = b * s[:, None]
    den = torch.maximum(bu.abs(), 1e-3 * s[:, None])
    assert float(((res.betas.double() - bu[:, :2]).abs() / den[:, :2]).max()) <= 1e-5
    assert float(((res.dsort_betas.double() - bu[:, 5:]).abs() / den[:, 5:]).max()) <= 1e-5
    rel = lambda a, r: float(((a.double() - r).abs() / r.abs().clamp(min=1)).max())
    assert rel(res.tstats, t[:, :2]) <= 1e-5 and rel(res.dsort_tstats, t[:, 5:]) <= 1e-5
    assert rel(res.fstats, F) <= 1e-5
    assert float(((res.sigma2.double() - s2 * s**2).abs() / (s2 * s**2)).max()) <= 1e-6
    assert float((res.r2.double() - r2).abs().max()) <= 1e-6
    # the one-column set of a dsort column: F = t^2 of that column
    assert rel(res.set_fstats[:, 1], t[:, 5] ** 2) <= 1e-4


def test_dsort_series_outputs_match_oracle(dev):
    """Series outputs of the batched -dsort solve (glm/arma.py:3418-3430,
    3522-3528): whitened residuals, residuals and prediction of the extended
    fit in original units vs the FP64 oracle (abs <= 1e-5 of the voxel sd),
    and the Ljung-Box statistic recomputed on the extended fit's whitened
    residuals (rel <= 1e-5 against the oracle's on the oracle residuals)."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[fit] dsort series outputs", flush=True)
    rng = np.random.default_rng(18)
    T, V, q = 120, 100, 3
    X = _design(rng, T)
    Y = _data(rng, T, X, V)
    D = rng.standard_normal((V, q, T)).astype(np.float32)
    Y = Y + 0.4 * D[:, 1, :]
    ag = bg = torch.linspace(-0.8, 0.8, 9)
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, dsort=D, want_residuals=True,
                         want_predicted=True, want_ljung_box=True)
    idx = res.pair_index.long()
    s = torch.from_numpy(Y).double().std(dim=1).clamp(min=1e-8)
    g64 = O.precompute_grid(torch.from_numpy(X).double(), ag, bg, None, None, F64)
    Yn = torch.from_numpy(Y).double() / s[:, None]
    b, t, s2, r2, F, rw, pred, _ = O.dsort_gls(torch.from_numpy(X).double(), torch.from_numpy(D).double(), Yn, g64,
                                               idx)
    sc = s[:, None]
    assert float(((res.residuals_whitened.double() - rw * sc).abs() / sc).max()) <= 1e-5
    assert float(((res.predicted.double() - pred * sc).abs() / sc).max()) <= 1e-5
    assert float(((res.residuals.double() - (Yn - pred) * sc).abs() / sc).max()) <= 1e-5
    lb = O.ljung_box(rw, O.ljung_box_max_lag(T, X.shape[1], T))
    assert float(((res.ljung_box.double() - lb).abs() / lb.abs().clamp(min=1)).max()) <= 1e-5


def test_ols_baseline_matches_oracle(dev):
    """want_ols (glm/arma.py:4288-4323 -> glm/core.py fit_glm): the OLS fit
    that runs alongside, on the engine's GLS kernel at the white pair, vs the
    FP64 OLS oracle: betas / stderr / t / F rel <= 1e-5, sigma2 rel <= 1e-6,
    R^2 and r2_run <= 1e-6, GLT estimate and t rel <= 1e-5; a 2-D list
    (multi-run) input raises as unsupported."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11
    from fastfuncstuff_amd.glm.core import fit_glm

    print("[fit] ols baseline", flush=True)
    rng = np.random.default_rng(21)
    T, V = 100, 200
    X = _design(rng, T)
    Y = _data(rng, T, X, V)
    ag = bg = torch.linspace(-0.8, 0.8, 9)
    C = [np.array([1.0, -1.0, 0, 0, 0])]
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, want_ols=True, want_ols_residuals=True,
                         task_indices=[0, 1], glt_labels=["d"], glt_matrices=C)
    o = res.ols_results
    ref = O.ols_fit(torch.from_numpy(X).double(), torch.from_numpy(Y).double(), task_indices=[0, 1],
                    C=torch.tensor(np.stack(C)))
    rel = lambda a, r: float(((a.double() - r).abs() / r.abs().clamp(min=1e-3)).max())
    assert o.dof == T - 5 and o.betas.shape == (V, 2)
    assert rel(o.betas, ref["betas"]) <= 1e-5 and rel(o.stderr, ref["stderr"]) <= 1e-5
    assert rel(o.tstats, ref["tstats"]) <= 1e-5 and rel(o.fstats, ref["fstats"]) <= 1e-5
    assert float(((o.sigma2.double() - ref["sigma2"]).abs() / ref["sigma2"]).max()) <= 1e-6
    assert float((o.r2.double() - ref["r2"]).abs().max()) <= 1e-6
    assert float((o.r2_run[:, 0].double() - ref["r2_run"]).abs().max()) <= 1e-6
    assert float((o.meanvol.double() - ref["meanvol"]).abs().max()) <= 1e-4
    assert rel(o.contrast_betas, ref["contrast_betas"]) <= 1e-5
    assert rel(o.contrast_tstats, ref["contrast_tstats"]) <= 1e-5
    assert float((o.residuals.double() - ref["resid"]).abs().max()) <= 1e-4
    with pytest.raises(NotImplementedError):
        fit_glm([Y, Y], [X, X], 2.0, device=dev)


def test_dsort_nods_snapshot_and_run_bounds(dev):
    """want_dsort_nods keeps the no-dsort fit (equal to a fit without dsort);
    the compact per-run form (dsort_run_bounds) equals the explicitly expanded
    zero-padded per-run columns (glm/arma.py:3284-3303, 3389-3398, 5780-5825);
    default labels dsort#k."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[fit] dsort nods / run bounds", flush=True)
    rng = np.random.default_rng(23)
    T, V = 100, 120
    X = _design(rng, T)
    Y = _data(rng, T, X, V)
    Dc = rng.standard_normal((V, 1, T)).astype(np.float32)
    Y = Y + 0.4 * Dc[:, 0, :]
    ag = bg = torch.linspace(-0.8, 0.8, 9)
    base = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev)
    rb = [0, 60, T]
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, dsort=Dc, dsort_run_bounds=rb,
                         want_dsort_nods=True)
    assert torch.equal(res.nods_results.betas, base.betas) and res.nods_results.dof == T - 5
    assert res.dsort_labels == ["dsort#0", "dsort#1"] and res.dof == T - 5 - 2
    De = np.zeros((V, 2, T), dtype=np.float32)
    De[:, 0, :60] = Dc[:, 0, :60]
    De[:, 1, 60:] = Dc[:, 0, 60:]
    ref = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, dsort=De)
    assert torch.equal(res.betas, ref.betas) and torch.equal(res.dsort_betas, ref.dsort_betas)
    assert torch.equal(res.sigma2, ref.sigma2)


def test_grouped_dsort_matches_per_group_fits(dev):
    """Grouped fit with voxel-wise regressors: dsort is sliced per group and the
    merged outputs equal the per-group fits (glm/arma.py:2931-2940, 3035-3036)."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11, fit_glm_arma11_grouped

    print("[fit] grouped dsort", flush=True)
    rng = np.random.default_rng(29)
    T, V = 80, 90
    X0, X1 = _design(rng, T), _design(rng, T, shift=2)
    grp = rng.integers(0, 2, V)
    Y = _data(rng, T, X0, V)
    D = rng.standard_normal((V, T)).astype(np.float32)
    ag = bg = torch.linspace(-0.8, 0.8, 9)
    res = fit_glm_arma11_grouped(Y, {0: X0, 1: X1}, grp, 2.0, a_grid=ag, b_grid=bg, device=dev, dsort=D,
                                 want_dsort_nods=True)
    for g, Xg in ((0, X0), (1, X1)):
        m = grp == g
        ref = fit_glm_arma11(Y[m], Xg, 2.0, a_grid=ag, b_grid=bg, device=dev, dsort=D[m], want_dsort_nods=True)
        assert torch.equal(res.betas[torch.from_numpy(m)], ref.betas)
        assert torch.equal(res.dsort_betas[torch.from_numpy(m)], ref.dsort_betas)
        assert torch.equal(res.nods_results.betas[torch.from_numpy(m)], ref.nods_results.betas)


def test_dsort_wide_design_matches_oracle(dev):
    """-dsort with a 40-column design (wider than the f64-MFMA GLS tiles; the
    per-lane GLS and the recurrence tables serve it): base and dsort betas rel
    <= 1e-5 of max(|b|, 1e-3 sd), t rel <= 1e-5, sigma2 rel <= 1e-6 vs the
    FP64 oracle at the engine's pairs."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[fit] dsort p=40", flush=True)
    rng = np.random.default_rng(23)
    T, V, q = 160, 64, 1
    base = _design(rng, T)
    extra = np.column_stack([np.cos(np.pi * (j + 1) * np.arange(T) / T) for j in range(35)])
    X = np.column_stack([base, extra])
    assert X.shape[1] == 40
    Y = _data(rng, T, base, V)
    D = rng.standard_normal((V, q, T)).astype(np.float32)
    Y = Y + 0.5 * D[:, 0, :]
    ag = bg = torch.linspace(-0.8, 0.8, 9)
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, dsort=D, task_indices=[0, 1])
    assert res.dof == T - 40 - q
    idx = res.pair_index.long()
    s = torch.from_numpy(Y).double().std(dim=1).clamp(min=1e-8)
    g64 = O.precompute_grid(torch.from_numpy(X).double(), ag, bg, None, None, F64)
    Yn = torch.from_numpy(Y).double() / s[:, None]
    b, t, s2, _, _, _, _, _ = O.dsort_gls(torch.from_numpy(X).double(), torch.from_numpy(D).double(), Yn, g64, idx,
                                          task_indices=[0, 1])
    bu = b * s[:, None]
    den = torch.maximum(bu.abs(), 1e-3 * s[:, None])
    assert float(((res.betas.double() - bu[:, :2]).abs() / den[:, :2]).max()) <= 1e-5
    assert float(((res.dsort_betas.double() - bu[:, 40:]).abs() / den[:, 40:]).max()) <= 1e-5
    rel = lambda a, r: float(((a.double() - r).abs() / r.abs().clamp(min=1)).max())
    assert rel(res.tstats, t[:, :2]) <= 1e-5 and rel(res.dsort_tstats, t[:, 40:]) <= 1e-5
    assert float(((res.sigma2.double() - s2 * s**2).abs() / (s2 * s**2)).max()) <= 1e-6


def test_global_estimation_matches_oracle(dev):
    """estimate_per_voxel=False (glm/arma.py:5603-5780): the pair is the ridge
    search of the mean series (== oracle.ridge_grid_search), every voxel gets
    it; GLS vs oracle.global_fit on the std-normalised data: betas rel <= 1e-5
    of max(|b|, 1e-3 sd), t / F rel <= 1e-5, sigma2 rel <= 1e-6, R^2 <= 1e-6,
    Ljung-Box of the whitened residuals, residual series; a task list shorter
    than the design is refused."""
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[fit] global estimation", flush=True)
    rng = np.random.default_rng(31)
    T, V = 120, 90
    X = _design(rng, T)
    Y = _data(rng, T, X, V)
    ag = bg = torch.linspace(-0.8, 0.8, 9)
    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, estimate_per_voxel=False, want_residuals=True,
                         want_predicted=True, want_ljung_box=True, want_r2_partial=True, column_sets=[[0, 1], [3]],
                         run_starts=[0, 70])
    Yt = torch.from_numpy(Y)
    pairs = O.valid_pairs(ag, bg)
    k, _ = O.ridge_grid_search(torch.from_numpy(X).float().double(), Yt.float().mean(0).double()[None], pairs,
                               run_starts=[0, 70])
    a, b = pairs[int(k[0])].tolist()
    assert torch.all(res.arma_params[:, 0] == a) and torch.all(res.arma_params[:, 1] == b)
    s = Yt.float().std(dim=1).clamp(min=1e-8).double()
    g = O.global_fit(torch.from_numpy(X).float().double(), Yt.float().double() / s[:, None], a, b, run_starts=[0, 70])
    bu = g["betas"] * s[:, None]
    den = torch.maximum(bu.abs(), 1e-3 * s[:, None])
    assert float(((res.betas.double() - bu).abs() / den).max()) <= 1e-5
    rel = lambda x, r: float(((x.double() - r).abs() / r.abs().clamp(min=1)).max())
    assert rel(res.tstats, g["tstats"]) <= 1e-5 and rel(res.fstats, g["fstats"]) <= 1e-5
    s2 = g["sigma2"] * s**2
    assert float(((res.sigma2.double() - s2).abs() / s2).max()) <= 1e-6
    assert float((res.r2.double() - g["r2"]).abs().max()) <= 1e-6
    assert res.dof == T - X.shape[1]
    t2 = g["tstats"] ** 2
    assert float((res.r2_partial.double() - t2 / (t2 + res.dof)).abs().max()) <= 1e-5
    rw = g["resid_w"] * s[:, None]
    assert float((res.residuals_whitened.double() - rw).abs().max()) <= 1e-4 * float(s.max())
    assert float((res.residuals.double() - g["resid"] * s[:, None]).abs().max()) <= 1e-4 * float(s.max())
    lb = O.ljung_box(g["resid_w"], O.ljung_box_max_lag(T, X.shape[1], 50), O.ljung_box_tau(T, [0, 70]))
    assert np.allclose(res.ljung_box.numpy(), lb.numpy(), rtol=1e-4, atol=1e-4)
    assert rel(res.set_fstats[:, 1], g["tstats"][:, 3] ** 2) <= 1e-4
    with pytest.raises(ValueError):
        fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, estimate_per_voxel=False, task_indices=[0, 1])


def test_pair_list_plan_continuous_params(dev):
    """Per-voxel (a, b) that are all distinct (as loaded from a smoothed -Rvar):
    the plan is a pair list whose memory is linear in the pairs, the GLS at
    each voxel's own whitener matches the FP64 oracle, and the search entry
    points refuse the plan."""
    from fastfuncstuff_amd._lib import RemlLibraryError
    from fastfuncstuff_amd.glm.arma import RemlGrid, fit_glm_arma11

    print("[fit] pair-list plan, 5000 distinct pairs", flush=True)
    rng = np.random.default_rng(77)
    T, p, V = 120, 5, 5000
    X = np.column_stack([np.ones(T), np.linspace(-1, 1, T), rng.standard_normal((T, p - 2))])
    a = rng.uniform(0.1, 0.8, V)
    b = rng.uniform(-0.5, 0.3, V)
    Y = (rng.standard_normal((V, p)) @ X.T + rng.standard_normal((V, T))).astype(np.float32)
    pp = torch.from_numpy(np.stack([a, b], 1))
    res = fit_glm_arma11(Y, X, 2.0, precomputed_arma_params=pp, device=dev)
    sel = list(range(0, V, 125))
    Linv = torch.stack([O.whitener(float(a[v]), float(b[v]), T)[0] for v in sel])
    g = O.Grid(pp[sel], Linv, None, None, None, None)
    Yd = torch.from_numpy(Y).double()[sel]
    ref = O.gls_refit(torch.from_numpy(X), Yd, g, torch.arange(len(sel)), dtype=F64)
    s = Yd.std(dim=1)
    denom = torch.maximum(ref[0].abs(), 1e-3 * s[:, None])
    assert float(((res.betas[sel].double() - ref[0]).abs() / denom).max()) <= 1e-5
    assert float(((res.tstats[sel].double() - ref[1]).abs() / ref[1].abs().clamp(min=1)).max()) <= 1e-5
    assert float(((res.sigma2[sel].double() - ref[3]).abs() / ref[3]).max()) <= 1e-5
    grid = RemlGrid(torch.from_numpy(X), None, None, dev, pairs=pp[:64], lattice=False)
    Yt, _, _ = grid.prep(torch.from_numpy(Y[:64]).to(dev))
    with pytest.raises(RemlLibraryError):
        grid.surface(Yt, 64)


@pytest.mark.parametrize("p", [6, 40, 68])
def test_grouped_engine_matches_per_group_fits(dev, p):
    """The grouped engine (one launch per stage over all groups, HIP graph)
    gives the per-group fits' results: scores within 2e-4, pairs on >= 99.9%
    of voxels (float32 ties aside), betas / t / F / sigma2 / R^2 within 1e-6
    relative where the pairs agree; the autotuner reports its candidates and
    choice; graph replays with new data in the same buffers reproduce a fresh
    fit of that data."""
    from fastfuncstuff_amd.glm.arma import RemlGrid, RemlGroup, fit_glm_arma11_grouped

    print(f"[fit] grouped engine p={p}", flush=True)
    rng = np.random.default_rng(100 + p)
    T, V, G = 150, 1200, 4
    ag = bg = torch.linspace(-0.8, 0.8, 17)
    designs = {}
    for h in range(G):
        designs[h + 10] = np.column_stack([np.ones(T), np.linspace(-1, 1, T), rng.standard_normal((T, p - 2))])
    groups = rng.choice(np.arange(10, 10 + G), V, p=[0.5, 0.3, 0.15, 0.05])
    e = np.zeros((V, T))
    a = rng.uniform(0.1, 0.8, V)
    for t in range(T):
        e[:, t] = (a * e[:, t - 1] if t else 0) + rng.standard_normal(V)
    sig = np.stack([rng.standard_normal(p) @ designs[h].T for h in groups]) * 0.3
    Y = torch.from_numpy((sig + 2.0 + e).astype(np.float32))
    eng = fit_glm_arma11_grouped(Y, designs, torch.from_numpy(groups), 2.0, a_grid=ag, b_grid=bg, device=dev,
                                 autotune=True, save_profile_likelihoods=True)
    loop = fit_glm_arma11_grouped(Y, designs, torch.from_numpy(groups), 2.0, a_grid=ag, b_grid=bg, device=dev,
                                  grouped_engine=False, save_profile_likelihoods=True)
    tun = eng.grouped_engine["tuning"]
    print(f"[fit] tuning {tun}", flush=True)
    assert tun["chosen"]["ms"] > 0 and len(tun["candidates"]) >= 3
    # the tuner may pick the in-kernel split projection: scores agree to fp32
    # rounding, pairs except float32 ties
    dS = (eng.reml_lklhd_surface.double() - loop.reml_lklhd_surface.double()).abs()
    assert float(dS.max()) <= 2e-4
    same = eng.pair_index == loop.pair_index
    assert float(same.double().mean()) >= 0.999
    for name in ("betas", "tstats", "fstats", "sigma2", "r2"):
        x, y = getattr(eng, name).double()[same], getattr(loop, name).double()[same]
        assert float(((x - y).abs() / y.abs().clamp(min=1e-3)).max()) <= 1e-6, name
    # replays on new data in the same buffer equal a fresh engine fit
    grids = RemlGrid.create_many([torch.from_numpy(designs[h]) for h in sorted(designs)], ag, bg, dev)
    for g in grids:
        g.set_projection("fp16x3")
    gv = np.searchsorted(sorted(designs), groups).astype(np.int32)
    engine = RemlGroup(grids, gv)
    Yd = Y.to(dev)
    first = {k: v.clone() for k, v in engine.fit(Yd).items()}
    Yd.mul_(1.5).add_(0.25)
    second = {k: v.clone() for k, v in engine.fit(Yd).items()}
    fresh = RemlGroup(grids, gv).fit(Yd)
    assert bool((second["best"] == fresh["best"]).all()) and bool((second["betas"] == fresh["betas"]).all())
    assert not bool((second["betas"] == first["betas"]).all())
    engine.close()


def _deficient_design(rng, T, p):
    """Column 2 duplicates column 1 and column p - 2 is column 0 + 2 column 3
    (two dependent columns, rank p - 2); the others are smooth plus noise."""
    t = np.linspace(0, 1, T)
    base = np.column_stack([np.sin((j + 1) * np.pi * t) + 0.3 * rng.standard_normal(T) for j in range(p - 2)])
    cols = [base[:, 0], base[:, 1], base[:, 1].copy()] + [base[:, j] for j in range(2, p - 3)]
    cols += [base[:, 0] + 2.0 * base[:, 2], base[:, p - 3]]
    return np.column_stack(cols)


@pytest.mark.parametrize("p", [6, 40, 110])
def test_rank_deficient_design_matches_oracle(dev, p):
    """A design with a duplicated and a combined column (rank p - 2) builds
    instead of raising (glm/arma.py:1493-1500 falls back to QR): rank p - 2;
    log dets equal the oracle's rank-r tables (the QR branch with its
    rounding-picked directions removed) to 1e-8; the full-grid surface is
    within 5e-4 + 4e-6 |S| + 2e-6 (T - p) E / rss of the FP64 oracle surface
    on those tables (the last term is float32 cancellation in rss = E - P:
    the float32 search at the (0.8, 0.8) corner of the grid is 3.2e-3 off at
    |S| = 844 on this data with the full-rank reduced design too); at the
    engine's pairs the betas are the minimum-norm solution (the pinv fallback,
    glm/arma.py:3410-3416) and t, F, sigma^2, R^2 match the FP64 oracle refit
    (betas / t rel <= 1e-5, F / sigma^2 rel <= 1e-5, R^2 rel <= 1e-6).  The
    grouped engine with this design and a full-rank one gives the same for
    the deficient group (p = 6 uses the f64 MFMA GLS, p = 40 the wide GLS,
    p = 110 the segmented path: fp32 projection over 64-column segments, the
    one-wave-per-voxel GLS, and the per-group loop for the grouped fit)."""
    from fastfuncstuff_amd.glm.arma import RemlGrid, fit_glm_arma11, fit_glm_arma11_grouped

    print(f"[fit] rank-deficient design p={p}", flush=True)
    rng = np.random.default_rng(40 + p)
    T, V = 150, 300
    X = _deficient_design(rng, T, p)
    Y = _data(rng, T, X, V)
    ag = bg = torch.linspace(-0.8, 0.8, 9)
    X64, Y64 = torch.from_numpy(X), torch.from_numpy(Y).double()
    gr = O.precompute_grid(X64, ag, bg, rank_fallback="rank")
    grid = RemlGrid(X, ag, bg, dev)
    assert grid.rank == p - 2
    ld_ref = gr.logdet_XRX.numpy()
    assert np.abs(grid.logdet_XRX - ld_ref).max() <= 1e-8 * max(1.0, np.abs(ld_ref).max())
    grid.close()

    s = Y64.std(dim=1).clamp(min=1e-8)
    Yn = Y64 / s[:, None]
    S64 = O.score_surface(X64, Yn, gr)
    # conditioning of rss = E - P: E / rss per (voxel, pair), E = |L^-1 r|^2 of
    # the residualised series.  A relative float32 error d in E and P moves the
    # score by (T - p) d E / rss; on these smooth designs E / rss reaches 70-80
    # at the (0.8, 0.8) corner (P = 0.987 E), where the measured error is
    # <= 1.4e-6 (T - p) E / rss for p = 6, 40 and 110
    Yr = Yn - (Yn @ O.design_pinv(X64).T) @ X64.T
    E64 = torch.stack([((Yr @ gr.Linv[k].T) ** 2).sum(1) for k in range(gr.params.shape[0])], 1)
    rss64 = torch.exp((S64 - gr.logdet_R - gr.logdet_XRX) / (T - p))
    cancel = (T - p) * 2e-6 * E64 / rss64

    def check(res, rows):
        S = res.reml_lklhd_surface.double() if res.reml_lklhd_surface is not None else None
        if S is not None:
            err = (S - S64[rows]).abs()
            print(f"[fit] rank-deficient p={p}: surface error vs FP64 max {float(err.max()):.3e}", flush=True)
            assert bool((err <= 5e-4 + 4e-6 * S64[rows].abs() + cancel[rows]).all()), float(err.max())
        idx = res.pair_index.long()
        b, t, f, s2, r2, _, _ = O.gls_refit(X64, Yn[rows], gr, idx, rank_fallback="pinv")
        sc = s[rows]
        b = b * sc[:, None]
        den = torch.maximum(b.abs(), 1e-3 * sc[:, None])
        assert float(((res.betas.double() - b).abs() / den).max()) <= 1e-5
        assert float(((res.tstats.double() - t).abs() / t.abs().clamp(min=1)).max()) <= 1e-5
        assert float(((res.fstats.double() - f).abs() / f.abs().clamp(min=1)).max()) <= 1e-5
        assert float(((res.sigma2.double() - s2 * sc**2).abs() / (s2 * sc**2)).max()) <= 1e-5
        # R^2 is far below 0 here (no intercept column): float32 output, relative bar
        assert float(((res.r2.double() - r2).abs() / r2.abs().clamp(min=1)).max()) <= 1e-6

    res = fit_glm_arma11(Y, X, 2.0, a_grid=ag, b_grid=bg, device=dev, save_profile_likelihoods=True)
    check(res, torch.arange(V))
    # grouped engine: this design for voxels with gid 1, a full-rank one for gid 0
    gid = rng.integers(0, 2, V)
    Xf = rng.standard_normal((T, p))
    res_g = fit_glm_arma11_grouped(Y, {0: Xf, 1: X}, torch.from_numpy(gid), 2.0, a_grid=ag, b_grid=bg, device=dev,
                                   save_profile_likelihoods=True)
    m = torch.from_numpy(gid == 1)

    class Sub:
        pass

    sub = Sub()
    for name in ("pair_index", "betas", "tstats", "fstats", "sigma2", "r2", "reml_lklhd_surface"):
        val = getattr(res_g, name)
        setattr(sub, name, None if val is None else val[m])
    check(sub, m.nonzero().flatten())
