"""OLS GLM (``fit_glm``) for the baseline the REML fit runs alongside.

The reference's ``fit_glm_arma11(want_ols=True)`` (glm/arma.py:4288-4379)
calls ``fit_glm`` (glm/core.py:522-1280) on the same data and design with
``max_poly_degree=-1`` (the design is complete) and a single run.  That call is
what this module provides: ordinary least squares is the GLS of the white pair
(a, b) = (0, 0), so the fit runs on the engine's FP64 GLS kernel with a
one-pair plan; the statistics the reference derives from it (standard errors,
the task-column F, per-run R^2, partial R^2, single-row GLTs) follow its
formulas.  Polynomial baselines, extra regressors and multi-run lists are not
part of that call and raise ``NotImplementedError``.
"""

from __future__ import annotations

import torch

from .. import _lib


class GLMResults:
    """glm/core.py:40-140: the fields of the OLS results the engine fills."""

    def __init__(self):
        self.betas = None  # (n_voxels, n_task) task-column coefficients
        self.r2 = None
        self.r2_partial = None
        self.r2_partial_nuisance = None
        self.r2_semipartial = None
        self.r2_semipartial_nuisance = None
        self.r2_run = None
        self.residuals = None
        self.predicted = None
        self.meanvol = None
        self.tstats = None
        self.stderr = None
        self.sigma2 = None
        self.fstats = None
        self.dof = None
        self.xtx_inv = None
        self.original_shape = None
        self.contrast_labels = None
        self.contrast_betas = None
        self.contrast_tstats = None


def fit_glm(data, design, tr: float, max_poly_degree=-1, extra_regressors=None, want_residuals: bool = False,
            want_predicted: bool = False, want_r2_run: bool = True, want_r2_partial: bool = False,
            r2_partial_mode: str = "full", want_r2_semipartial: bool = False, r2_semipartial_mode: str = "full",
            device=None, chunk_size: int | None = None, verbose: bool = False, preload_data_to_device: bool = True,
            use_double: bool = False, glt_labels=None, glt_matrices=None, task_indices=None, debug_memory=False,
            debug_design=False) -> GLMResults:
    """glm/core.py:522-1280 for one run with a complete design (the call of
    glm/arma.py:4302-4323).  Per voxel, with task columns S (``task_indices``,
    all columns by default) and dof = T - p: betas_S; sigma2 = max(RSS / dof, 0);
    stderr = sqrt(sigma2 diag((X'X)^-1)_S); t = b / (stderr + 1e-10); F = t^2
    for one task column, else b_S' ((X'X)^-1_SS)^-1 b_S / (|S| sigma2 + 1e-10);
    R^2 = CoD; r2_run = 1 - RSS_run / (SS_tot + 1e-10) clamped to [0, 1], with
    the prediction of the task columns only unless ``want_predicted`` (then the
    full prediction), as the reference computes it; meanvol = mean over time;
    partial R^2 = t^2 / (t^2 + dof) over all columns, split into task /
    nuisance ("task" mode divides by max(1 - sum(nuisance), 0.01)), semi-partial
    = partial * max(1 - R^2, 0) with the same split; GLT: c'b and
    c'b / (sqrt(sigma2 c'(X'X)^-1 c) + 1e-10)."""
    from .arma import RemlGrid

    if isinstance(data, list) or isinstance(design, list):
        raise NotImplementedError("multi-run lists: the OLS baseline fits one run")
    if max_poly_degree not in (-1,) or extra_regressors is not None:
        raise NotImplementedError("polynomial baselines / extra regressors: the OLS baseline takes a complete design")
    dev = _lib.require_device(device)
    data = torch.as_tensor(data)
    original_shape = None
    if data.ndim == 4:
        original_shape = tuple(data.shape[:3])
        data = data.reshape(-1, data.shape[-1])
    elif data.ndim != 2:
        raise ValueError(f"Data must be 2D (n_voxels, n_timepoints) or 4D (nx, ny, nz, nt), got {data.ndim}D")
    X = torch.as_tensor(design, dtype=torch.float64)
    V, T = data.shape
    if X.shape[0] != T:
        raise ValueError(f"Timepoints mismatch: Data has {T}, Design has {X.shape[0]}")
    p = X.shape[1]
    task = list(task_indices) if task_indices is not None else list(range(p))
    if not task:
        raise ValueError("No task regressors detected; cannot compute GLM statistics")
    dof = max(1, T - p)
    f64 = torch.float64
    out_dtype = torch.float64 if use_double else torch.float32
    grid = RemlGrid(X, None, None, dev, pairs=torch.zeros(1, 2, dtype=f64), task_indices=task)
    Ai = grid.ainv()[0].to(dev)
    ti = torch.as_tensor(task, dtype=torch.long, device=dev)
    diag_full = torch.diagonal(Ai)
    diag_task = diag_full.index_select(0, ti)
    Ai_task_inv = torch.linalg.inv(Ai.index_select(0, ti).index_select(1, ti))
    Xd = X.to(dev)
    nuis = sorted(set(range(p)) - set(task))
    C = None
    if glt_labels and glt_matrices:
        rows = []
        for m in glt_matrices:
            c = torch.as_tensor(m, dtype=f64)
            if c.ndim == 2 and c.shape[0] == 1:
                c = c[0]
            if c.ndim != 1:
                raise NotImplementedError(
                    f"Multi-row GLT contrasts (F-tests) not yet supported in OLS. "
                    f"Got shape {tuple(c.shape)}, expected (n_regressors,) or (1, n_regressors)")
            rows.append(c)
        C = torch.stack(rows).to(dev)
        cvar = torch.einsum("cr,rs,cs->c", C, Ai, C)
    res = GLMResults()
    res.original_shape = original_shape
    res.dof = dof
    res.xtx_inv = Ai.cpu()
    nT = len(task)
    res.betas = torch.zeros(V, nT, dtype=out_dtype)
    res.tstats = torch.zeros(V, nT, dtype=out_dtype)
    res.stderr = torch.zeros(V, nT, dtype=out_dtype)
    res.sigma2 = torch.zeros(V, dtype=out_dtype)
    res.fstats = torch.zeros(V, dtype=out_dtype)
    res.r2 = torch.zeros(V, dtype=out_dtype)
    res.meanvol = torch.zeros(V, dtype=out_dtype)
    if want_r2_run:
        res.r2_run = torch.zeros(V, 1, dtype=out_dtype)
    if want_residuals:
        res.residuals = torch.zeros(V, T, dtype=out_dtype)
    if want_predicted:
        res.predicted = torch.zeros(V, T, dtype=out_dtype)
    if C is not None:
        res.contrast_labels = list(glt_labels)
        res.contrast_betas = torch.zeros(V, C.shape[0], dtype=out_dtype)
        res.contrast_tstats = torch.zeros(V, C.shape[0], dtype=out_dtype)
    nb = chunk_size or min(V, 1 << 20)
    with torch.cuda.device(dev):
        for b0 in range(0, V, nb):
            b1 = min(V, b0 + nb)
            Yb = data[b0:b1].to(device=dev, dtype=torch.float32).contiguous()
            best = torch.zeros(b1 - b0, dtype=torch.int32, device=dev)
            g = grid.gls(Yb, None, best, want_residuals=want_residuals, want_predicted=want_predicted)
            b = g["betas"].double()
            sig = g["sigma2"].double().clamp(min=0.0)
            bt = b.index_select(1, ti)
            se = torch.sqrt(sig[:, None] * diag_task[None, :])
            t = bt / (se + 1e-10)
            if nT == 1:
                F = t[:, 0] ** 2
            else:
                F = torch.einsum("bi,ij,bj->b", bt, Ai_task_inv, bt) / (nT * sig + 1e-10)
            r2 = g["r2"].double()
            Yd = Yb.double()
            res.betas[b0:b1] = bt.cpu().to(out_dtype)
            res.tstats[b0:b1] = t.cpu().to(out_dtype)
            res.stderr[b0:b1] = se.cpu().to(out_dtype)
            res.sigma2[b0:b1] = sig.cpu().to(out_dtype)
            res.fstats[b0:b1] = F.cpu().to(out_dtype)
            res.r2[b0:b1] = r2.cpu().to(out_dtype)
            res.meanvol[b0:b1] = Yd.mean(1).cpu().to(out_dtype)
            if want_r2_run:
                pred = g["predicted"].double() if want_predicted else bt @ Xd.index_select(1, ti).T
                ss_tot = (Yd - Yd.mean(1, keepdim=True)).pow(2).sum(1)
                rr = 1 - (Yd - pred).pow(2).sum(1) / (ss_tot + 1e-10)
                res.r2_run[b0:b1, 0] = rr.clamp(0, 1).cpu().to(out_dtype)
            if want_residuals:
                res.residuals[b0:b1] = g["residuals"].cpu().to(out_dtype)
            if want_predicted:
                res.predicted[b0:b1] = g["predicted"].cpu().to(out_dtype)
            if want_r2_partial or want_r2_semipartial:
                tf = b / (torch.sqrt(sig[:, None] * diag_full[None, :]) + 1e-10)
                full = tf**2 / (tf**2 + dof)
                pt = full.index_select(1, ti)
                pn = full[:, nuis] if nuis else None
                if want_r2_partial:
                    val = pt
                    if r2_partial_mode == "task" and pn is not None:
                        val = pt / torch.clamp(1.0 - pn.sum(1, keepdim=True), min=0.01)
                    _store(res, "r2_partial", b0, b1, V, val, out_dtype)
                    if pn is not None:
                        _store(res, "r2_partial_nuisance", b0, b1, V, pn, out_dtype)
                if want_r2_semipartial:
                    rem = torch.clamp(1.0 - r2, min=0.0)[:, None]
                    st, sn = pt * rem, (pn * rem if pn is not None else None)
                    val = st
                    if r2_semipartial_mode == "task" and sn is not None:
                        val = st / torch.clamp(1.0 - sn.sum(1, keepdim=True), min=0.01)
                    _store(res, "r2_semipartial", b0, b1, V, val, out_dtype)
                    if sn is not None:
                        _store(res, "r2_semipartial_nuisance", b0, b1, V, sn, out_dtype)
            if C is not None:
                cb = b @ C.T
                ct = cb / (torch.sqrt((sig[:, None] * cvar[None, :]).clamp(min=0.0)) + 1e-10)
                res.contrast_betas[b0:b1] = cb.cpu().to(out_dtype)
                res.contrast_tstats[b0:b1] = ct.cpu().to(out_dtype)
    return res


def _store(res, name, b0, b1, V, val, dtype):
    if getattr(res, name) is None:
        setattr(res, name, torch.zeros((V,) + tuple(val.shape[1:]), dtype=dtype))
    getattr(res, name)[b0:b1] = val.cpu().to(dtype)
