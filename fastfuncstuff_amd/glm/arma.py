"""ARMA(1,1) REML GLM on MI355X: the ``fastfuncstuff.glm.arma`` fit API.

Host-side mirror of the reference interface (fastfuncstuff/glm/arma.py):
``fit_glm_arma11`` -> ``ARMA11Results``, ``precompute_reml_grid``,
``search_voxels_precomputed_grid``, ``select_arma_params_hierarchical_from_surface``
and the grid / lambda / censoring helpers, with the same argument meaning and
error behaviour.  All per-voxel arithmetic runs in ``libffsreml.so``
(hand-written gfx950 kernels); this module only shapes inputs, owns the device
buffers and collects results on the host, as the reference does.
"""

from __future__ import annotations

import ctypes
import math
import warnings
from collections.abc import Mapping
from typing import Any

import numpy as np
import torch

from .. import _lib

# glm/arma.py:234-243 (AFNI 3dREMLfit -MAXa 0.8 -MAXb 0.8 -Grid 3)
DEFAULT_ARMA_A_GRID = (0.0, 0.8, 9)
DEFAULT_ARMA_B_GRID = (-0.8, 0.8, 17)

# glm/arma.py:173-200
VOXEL_SCATTER_ATTRS = [
    "betas", "tstats", "r2", "r2_partial", "r2_partial_nuisance", "r2_semipartial",
    "r2_semipartial_nuisance", "arma_params", "arma_lambda", "reml_likelihood", "sigma2", "fstats",
    "residuals", "residuals_whitened", "predicted", "contrast_betas", "contrast_tstats",
    "contrast_fstats", "contrast_r2_partial", "contrast_r2_semipartial", "dsort_betas", "dsort_tstats",
    "ljung_box", "set_fstats", "set_sd", "set_coef",
]

# Widest design of the learned searches (windows and lambda bands), the subset
# search and the grouped engine; wider
# designs (up to 768 columns) use the full-grid surface, with the projection
# over 64-column segments of each pair.
WIDE_P = 96

# Runs are offset by this much in the Ljung-Box pseudo-time (glm/arma.py:5969).
_LJ_RUN_SEPARATION = 66666


# ---------------------------------------------------------------------------
# Grid helpers (host logic, float32 semantics of the reference)
# ---------------------------------------------------------------------------


def get_default_arma_grids(device: torch.device | None = None) -> tuple[torch.Tensor, torch.Tensor]:
    """glm/arma.py:484-513: a = linspace(0, 0.8, 9), b = linspace(-0.8, 0.8, 17)."""
    return (torch.linspace(*DEFAULT_ARMA_A_GRID, device=device),
            torch.linspace(*DEFAULT_ARMA_B_GRID, device=device))


def ensure_zero_in_grid(a_grid: torch.Tensor, b_grid: torch.Tensor, tolerance: float = 1e-6):
    """glm/arma.py:516-582: snap |g| < tolerance to exactly 0, or insert 0 and sort."""

    def fix(g: torch.Tensor) -> torch.Tensor:
        near = torch.abs(g) < tolerance
        if torch.any(near):
            g = g.clone()
            g[near] = 0.0
            return g
        return torch.sort(torch.cat([torch.zeros(1, dtype=g.dtype, device=g.device), g]))[0]

    return fix(a_grid), fix(b_grid)


def _compute_arma11_lambda(a, b):
    """glm/arma.py:695-729: lambda = (a+b)(1+ab)/(1+b^2+2ab)."""
    return (a + b) * (1 + a * b) / (1 + b**2 + 2 * a * b)


def compute_arma_lambda(a: float, b: float) -> float:
    """glm/arma.py:434-462."""
    return float(_compute_arma11_lambda(a, b))


def _grid_tensor(g, name: str) -> torch.Tensor:
    t = torch.as_tensor(g)
    if t.ndim != 1 or t.numel() == 0:
        raise ValueError(f"{name} must be a non-empty 1-D grid")
    return t.detach().cpu().to(torch.float32) if not t.dtype.is_floating_point else t.detach().cpu()


def valid_grid_pairs(a_grid: torch.Tensor, b_grid: torch.Tensor) -> torch.Tensor:
    """The reference's validity filter (glm/arma.py:999-1048), in the grid dtype:
    meshgrid 'ij' (a-major, b-minor), |a|,|b| < 1, |1-a^2| >= 1e-6, gamma0 > 0,
    lambda >= 0.  Returns the (K, 2) valid pairs in search order."""
    am, bm = torch.meshgrid(a_grid, b_grid, indexing="ij")
    a, b = am.reshape(-1), bm.reshape(-1)
    ok = (torch.abs(a) < 1.0) & (torch.abs(b) < 1.0)
    denom = 1 - a**2
    ok &= torch.abs(denom) >= 1e-6
    ok &= (1 + b**2 + 2 * a * b) / (denom + 1e-10) > 0
    a, b = a[ok], b[ok]
    keep = _compute_arma11_lambda(a, b) >= 0
    return torch.stack([a[keep], b[keep]], dim=1)


def estimate_valid_grid_pairs(a_grid: torch.Tensor, b_grid: torch.Tensor) -> int:
    """glm/arma.py:291-338."""
    return int(valid_grid_pairs(_grid_tensor(a_grid, "a_grid"), _grid_tensor(b_grid, "b_grid")).shape[0])


def build_censor_run_info(run_starts, n_total: int, good_list=None):
    """glm/arma.py:752-820: run starts in retained-index space + within-run tau."""
    starts = run_starts.tolist() if isinstance(run_starts, torch.Tensor) else list(run_starts)
    if good_list is None:
        return [int(s) for s in starts], None
    retained = torch.as_tensor(good_list.tolist() if isinstance(good_list, torch.Tensor) else list(good_list),
                               dtype=torch.long).flatten()
    st = torch.tensor(starts, dtype=torch.long)
    run_id = torch.bucketize(retained, st, right=True) - 1
    tau = retained - st[run_id]
    out: list[int] = []
    for r in range(len(starts)):
        pos = (run_id == r).nonzero(as_tuple=False)
        if pos.numel() > 0:
            out.append(int(pos[0].item()))
    return (out or [0]), tau


def run_bounds(n: int, run_starts=None) -> list[tuple[int, int]]:
    """glm/whiten.py:27-38."""
    if run_starts is None:
        return [(0, n)]
    s = sorted({int(x) for x in (run_starts.tolist() if isinstance(run_starts, torch.Tensor) else run_starts)})
    if not s or s[0] != 0:
        s = [0, *s]
    e = [*s[1:], n]
    return [(x, y) for x, y in zip(s, e) if y > x]


def _run_block_mask(run_starts, n: int, device=None) -> torch.Tensor:
    """glm/arma.py:732-749: (n, n) bool, True within a run."""
    rid = torch.zeros(n, dtype=torch.long)
    for r, (s, e) in enumerate(run_bounds(n, run_starts)):
        rid[s:e] = r
    return (rid[None, :] == rid[:, None]).to(device)


def build_arma11_covariance_batch(a_grid, b_grid, n: int, device=None, dtype: torch.dtype = torch.float32,
                                  run_starts=None, tau=None):
    """glm/arma.py:939-1103 (without the AFNI band): the dense correlation
    matrices R[i, j] = corr[|tau_i - tau_j|], corr = [1, lam, lam a, lam a^2, ...],
    zeroed across runs, of every valid pair of the grid (the search's filter and
    order).  Returns (R (K, n, n), params (K, 2), param_list).  A utility for
    callers that need R itself; the engine whitens by recurrence instead."""
    a_grid, b_grid = _grid_tensor(a_grid, "a_grid"), _grid_tensor(b_grid, "b_grid")
    pairs = valid_grid_pairs(a_grid.to(dtype), b_grid.to(dtype))
    if pairs.shape[0] == 0:
        return torch.empty(0, n, n, device=device), torch.empty(0, 2, device=device), []
    a, b = pairs[:, 0], pairs[:, 1]
    lam = _compute_arma11_lambda(a, b).to(dtype)
    t = torch.arange(n) if tau is None else torch.as_tensor(tau, dtype=torch.long).flatten()
    dist = torch.abs(t[None, :] - t[:, None])
    max_lag = int(dist.max().item()) if n > 0 else 0
    corr = torch.zeros(pairs.shape[0], max_lag + 1, dtype=dtype)
    corr[:, 0] = 1.0
    if max_lag >= 1:
        corr[:, 1] = lam
        if max_lag >= 2:
            ex = torch.arange(1, max_lag, dtype=dtype)
            corr[:, 2:] = lam[:, None] * a[:, None].to(dtype) ** ex[None, :]
    R = corr[:, dist]
    if run_starts is not None:
        R = R * _run_block_mask(run_starts, n).to(dtype)
    R = R.contiguous().to(device)
    params = pairs.to(device)
    return R, params, [(float(x), float(y)) for x, y in pairs.tolist()]


def build_arma11_covariance(a: float, b: float, n: int, device=None, dtype: torch.dtype = torch.float32,
                            run_starts=None, tau=None):
    """glm/arma.py:823-936: the dense R of one (a, b), or None when it is not a
    valid ARMA(1,1) correlation."""
    R, _, plist = build_arma11_covariance_batch(torch.tensor([a], dtype=dtype), torch.tensor([b], dtype=dtype), n,
                                                device, dtype, run_starts=run_starts, tau=tau)
    return R[0] if plist else None


def timeline_gaps(n: int, run_starts=None, tau=None) -> np.ndarray:
    """Per retained index: tau[t+1]-tau[t] inside a run, 0 at each run's last point."""
    gap = np.zeros(n, dtype=np.int32)
    t_arr = None if tau is None else np.asarray(torch.as_tensor(tau).flatten().tolist(), dtype=np.int64)
    if t_arr is not None and t_arr.shape[0] != n:
        raise ValueError(f"tau must have one entry per retained timepoint ({n}), got {t_arr.shape[0]}")
    for s, e in run_bounds(n, run_starts):
        if t_arr is None:
            gap[s:e - 1] = 1
        else:
            d = np.diff(t_arr[s:e])
            if np.any(d <= 0):
                raise ValueError("tau must be strictly increasing within each run")
            gap[s:e - 1] = d
        gap[e - 1] = 0
    return gap


class _RemlMode:
    """Run-wide REML mode of glm/arma.py:59-92 (``set_afni_mode``).

    Only the default mode is built: the exact dense ARMA(1,1) covariance, the
    default grid filter and hierarchical top level.  The AFNI-faithful mode
    changes the covariance itself (R truncated to AFNI's bmax off-diagonals,
    glm/arma.py:918-926, 1075-1089), which the two-term innovations recurrence
    of every kernel does not represent, so a search in that mode raises rather
    than returning default-mode results under the AFNI-mode name."""

    afni_faithful: bool = False
    corcut: float = 1e-4


_REML_MODE = _RemlMode()


def set_afni_mode(enabled: bool, corcut: float = 1e-4) -> None:
    """glm/arma.py:85-92.  ``enabled=False`` (the reference's default) is
    accepted; ``enabled=True`` raises NotImplementedError (banded R is not
    built)."""
    if enabled:
        raise NotImplementedError(
            "AFNI-faithful REML mode (banded R, corcut grid filter, AFNI ltop) is not built; "
            "only the default exact-covariance mode is available")
    _REML_MODE.afni_faithful = False
    _REML_MODE.corcut = float(corcut)


def _check_reml_mode() -> None:
    """Refuse a search when the mode object was switched to AFNI-faithful by
    hand (set_afni_mode itself refuses it)."""
    if _REML_MODE.afni_faithful:
        raise NotImplementedError(
            "AFNI-faithful REML mode (banded R, corcut grid filter, AFNI ltop) is not built; "
            "set _REML_MODE.afni_faithful = False")


def hierarchical_ltop(n_a: int, n_b: int) -> int:
    """glm/arma.py:2290-2296 (default, non-AFNI mode)."""
    return max(0, min(n_a, n_b).bit_length() - 2)


try:
    from threadpoolctl import threadpool_limits as _threadpool_limits
except ImportError:  # optional: only bounds the BLAS pool around small SVDs
    _threadpool_limits = None


def design_pinv(X: np.ndarray, tol_dtype=np.float32) -> np.ndarray:
    """glm/core.py:248-272: rank-truncated SVD pseudo-inverse in float64 with the
    design dtype's tolerance.  Returns pinv (p x T)."""
    Xd = np.asarray(X, dtype=np.float64)
    if _threadpool_limits is not None:
        # a T x p SVD is too small for a threaded BLAS: one thread is several
        # times faster than a contended pool
        with _threadpool_limits(1):
            U, S, Vh = np.linalg.svd(Xd, full_matrices=False)
    else:
        U, S, Vh = np.linalg.svd(Xd, full_matrices=False)
    tol = max(Xd.shape) * np.finfo(tol_dtype).eps * (S[0] if S.size else 0.0)
    r = int((S > tol).sum())
    return (Vh[:r].T / S[:r]) @ U[:, :r].T


# ---------------------------------------------------------------------------
# Precomputed grid = a device plan
# ---------------------------------------------------------------------------


class RemlGrid(Mapping):
    """Per-pair tables for one design (replaces the dict of
    ``precompute_reml_grid``, glm/arma.py:1371-1531).

    Maps ``(a, b)`` to ``{'a', 'b', 'logdet_Rcorr', 'logdet_XwTXw'}`` like the
    reference's dict; the whiteners, whitened designs and projection operands
    live on the device inside the plan."""

    def __init__(self, X, a_grid, b_grid, device, run_starts=None, tau=None, task_indices=None, pairs=None,
                 ridge: float = 0.0, ensure_zero: bool = True, lattice: bool | None = None):
        args = self._setup_host(X, a_grid, b_grid, device, run_starts, tau, task_indices, pairs, ridge, ensure_zero,
                                lattice)
        lib = _lib.load()
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            if self.ridge > 0:
                rc = lib.ffs_reml_plan_create_ridge(
                    ctypes.byref(handle), self.T, self.p, self.X64.ctypes.data, args["pinvT"].ctypes.data,
                    self.gap.ctypes.data, self.K, args["ab"].ctypes.data, args["white"].ctypes.data,
                    args["n_a"], args["n_b"], self.lut.ctypes.data, args["za"], args["zb"], self.ltop, self.ridge,
                    _lib.stream_handle(self.device))
            else:
                task = args["task"]
                rc = lib.ffs_reml_plan_create(
                    ctypes.byref(handle), self.T, self.p, self.X64.ctypes.data, args["pinvT"].ctypes.data,
                    self.gap.ctypes.data, self.K, args["ab"].ctypes.data, args["white"].ctypes.data,
                    args["n_a"], args["n_b"], self.lut.ctypes.data, args["za"], args["zb"], self.ltop,
                    int(task.size), task.ctypes.data if task.size else None, _lib.stream_handle(self.device))
        _lib.check(rc, "precompute_reml_grid")
        self._attach(handle)

    @classmethod
    def create_many(cls, designs, a_grid, b_grid, device, run_starts=None, tau=None, task_indices=None):
        """One plan per design (a list of (T, p) arrays of one shape) on one
        timeline and (a, b) grid, the per-pair tables of all designs built in
        one device launch (ffs_reml_plan_create_many).  Returns the list of
        RemlGrid objects."""
        grids = [cls.__new__(cls) for _ in designs]
        setups = [g._setup_host(X, a_grid, b_grid, device, run_starts, tau, task_indices, None, 0.0, True, None)
                  for g, X in zip(grids, designs)]
        g0, a0 = grids[0], setups[0]
        if any(g.X64.shape != g0.X64.shape for g in grids):
            raise ValueError("create_many: every design must have the same shape")
        Xall = np.ascontiguousarray(np.stack([g.X64 for g in grids]))
        Pall = np.ascontiguousarray(np.stack([a["pinvT"] for a in setups]))
        G = len(grids)
        handles = (ctypes.c_void_p * G)()
        task = a0["task"]
        lib = _lib.load()
        with torch.cuda.device(g0.device):
            rc = lib.ffs_reml_plan_create_many(
                ctypes.cast(handles, ctypes.c_void_p), G, g0.T, g0.p, Xall.ctypes.data, Pall.ctypes.data,
                g0.gap.ctypes.data, g0.K, a0["ab"].ctypes.data, a0["white"].ctypes.data, a0["n_a"], a0["n_b"],
                g0.lut.ctypes.data, a0["za"], a0["zb"], g0.ltop, int(task.size),
                task.ctypes.data if task.size else None, _lib.stream_handle(g0.device))
        _lib.check(rc, "precompute_reml_grid (grouped designs)")
        for g, h in zip(grids, handles):
            g._attach(ctypes.c_void_p(h))
        return grids

    def _setup_host(self, X, a_grid, b_grid, device, run_starts, tau, task_indices, pairs, ridge, ensure_zero,
                    lattice) -> dict:
        """Host-side tables of the plan (pairs, lattice, timeline, pinv);
        returns the remaining plan-creation arguments."""
        _check_reml_mode()
        self.device = _lib.require_device(device)
        X = torch.as_tensor(X)
        if X.ndim != 2:
            raise ValueError("design must be 2-D (n_timepoints, n_regressors)")
        self.X64 = np.ascontiguousarray(X.detach().cpu().to(torch.float64).numpy())
        self.T, self.p = self.X64.shape
        # ``lattice`` False: a pair list (e.g. per-voxel parameters loaded from
        # -Rvar) for the GLS only; no (a, b) lattice, window or search tables,
        # so the plan grows linearly with the number of pairs
        self.lattice = (pairs is None) if lattice is None else bool(lattice)
        if pairs is None:
            a_grid, b_grid = _grid_tensor(a_grid, "a_grid"), _grid_tensor(b_grid, "b_grid")
            if ensure_zero:
                a_grid, b_grid = ensure_zero_in_grid(a_grid, b_grid)
            pairs = valid_grid_pairs(a_grid, b_grid)
        self.params = pairs.to(torch.float32) if pairs.dtype != torch.float64 else pairs
        if self.params.shape[0] == 0:
            raise ValueError("No valid (a,b) parameters in grid")
        self.K = int(self.params.shape[0])
        a_np = self.params[:, 0].double().numpy()
        b_np = self.params[:, 1].double().numpy()
        lam32 = _compute_arma11_lambda(self.params[:, 0], self.params[:, 1])
        white = (lam32 == 0).to(torch.int32).numpy()
        if self.lattice:
            self.a_axis = sorted(set(a_np.tolist()))
            self.b_axis = sorted(set(b_np.tolist()))
        else:
            if ridge > 0:
                raise ValueError("a ridge plan needs an (a, b) lattice")
            self.a_axis, self.b_axis = [], []
        n_a, n_b = len(self.a_axis), len(self.b_axis)
        ai = {v: i for i, v in enumerate(self.a_axis)}
        bi = {v: i for i, v in enumerate(self.b_axis)}
        lut = -np.ones((n_a, n_b), dtype=np.int32)
        if self.lattice:
            for k in range(self.K):
                lut[ai[a_np[k]], bi[b_np[k]]] = k
        self.lut = lut
        self.ridge = float(ridge)
        if self.ridge > 0 and task_indices is not None:
            raise ValueError("a ridge plan is search-only (no task subset)")
        za = next((i for i, v in enumerate(self.a_axis) if abs(v) < 1e-6), None)
        zb = next((i for i, v in enumerate(self.b_axis) if abs(v) < 1e-6), None)
        self.has_zero = za is not None and zb is not None and lut[za, zb] >= 0
        za = 0 if za is None else za
        zb = 0 if zb is None else zb
        self.zero = (za, zb)
        self.ltop = hierarchical_ltop(n_a, n_b)
        self.gap = timeline_gaps(self.T, run_starts, tau)
        self.run_starts = run_starts
        self.tau = tau
        pinvT = np.ascontiguousarray(design_pinv(self.X64).T)
        ab = np.ascontiguousarray(np.stack([a_np, b_np], 1))
        task = np.asarray(task_indices if task_indices is not None else [], dtype=np.int32)
        self.task_indices = None if task_indices is None else list(task_indices)
        return {"pinvT": pinvT, "ab": ab, "white": np.ascontiguousarray(white), "task": task, "n_a": n_a, "n_b": n_b,
                "za": za, "zb": zb}

    def _attach(self, handle) -> None:
        """Take ownership of a created plan and read its dimensions and log-dets."""
        lib = _lib.load()
        self._handle = handle
        tp, vp, _ = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(lib.ffs_reml_plan_dims(handle, ctypes.byref(tp), ctypes.byref(vp), None), "plan_dims")
        self.Tpad, self.vox_pad = tp.value, vp.value
        ldR = np.zeros(self.K)
        ldX = np.zeros(self.K)
        _lib.check(lib.ffs_reml_plan_logdets(handle, ldR.ctypes.data, ldX.ctypes.data), "plan_logdets")
        self.logdet_R, self.logdet_XRX = ldR, ldX
        # column rank of the design: below p the plan holds the rank-r tables
        # (the reference's QR fallback, glm/arma.py:1493-1508) and the GLS
        # reports the minimum-norm beta
        self.rank = int(lib.ffs_reml_plan_rank(handle))
        self.param_list = [(float(a), float(b)) for a, b in self.params.tolist()]
        self.projection = "fp32"
        self.bands = None

    # Mapping protocol (the reference's dict keyed by (a, b))
    def __getitem__(self, key):
        k = self.param_list.index((float(key[0]), float(key[1])))
        return {"a": self.param_list[k][0], "b": self.param_list[k][1],
                "logdet_Rcorr": torch.tensor(self.logdet_R[k]), "logdet_XwTXw": torch.tensor(self.logdet_XRX[k])}

    def __iter__(self):
        return iter(self.param_list)

    def __len__(self):
        return self.K

    @property
    def handle(self):
        return self._handle

    def close(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            _lib.load().ffs_reml_plan_destroy(h)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- device helpers -------------------------------------------------
    def vpad(self, V: int) -> int:
        return (V + self.vox_pad - 1) // self.vox_pad * self.vox_pad

    def prep(self, Y: torch.Tensor, normalize: bool = True, want_features: bool = False,
             rows: torch.Tensor | None = None, scale: torch.Tensor | None = None, residualize: bool = True):
        """Normalise + residualise Y (V, T) float32 on the device.  Returns
        (Yt (Tpad, ncols_pad) time-major residual, scale (V,), features (V, 10) | None).
        ``rows`` (int32, optional) selects the voxel of each output column."""
        V = Y.shape[0]
        ncols = V if rows is None else rows.shape[0]
        Vp = self.vpad(ncols)
        n_rows = ctypes.c_int32()
        _lib.check(_lib.load().ffs_reml_plan_yt_rows(self._handle, ctypes.byref(n_rows)), "plan_yt_rows")
        # rows Tpad.. hold the residual's split halves in the 16-bit projection modes
        Yt = torch.empty(n_rows.value, Vp, dtype=torch.float32, device=self.device)
        Yt.ffs_projection = self.projection
        if scale is None:
            scale = torch.empty(V, dtype=torch.float32, device=self.device)
        feats = torch.empty(V, 10, dtype=torch.float32, device=self.device) if want_features else None
        lib = _lib.load()
        flags = int(bool(normalize)) | (0 if residualize else 2)
        _lib.check(lib.ffs_reml_prep(self._handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(rows), ncols, _lib.ptr(Yt),
                                     Vp, _lib.ptr(scale), flags, _lib.ptr(feats),
                                     _lib.stream_handle(self.device)), "prep")
        return Yt, scale, feats

    def prep_features(self, Y: torch.Tensor, normalize: bool = True):
        """Per-voxel scale and lag-1..10 residual autocorrelations only (no
        time-major residual is written).  Returns (None, scale, feats)."""
        V = Y.shape[0]
        scale = torch.empty(V, dtype=torch.float32, device=self.device)
        feats = torch.empty(V, 10, dtype=torch.float32, device=self.device)
        _lib.check(_lib.load().ffs_reml_prep(self._handle, _lib.ptr(Y), Y.stride(0), V, None, V, None, 0,
                                             _lib.ptr(scale), int(normalize), _lib.ptr(feats),
                                             _lib.stream_handle(self.device)), "prep_features")
        return None, scale, feats

    def surface(self, Yt: torch.Tensor, V: int, out: torch.Tensor | None = None,
                n_chunks: int | None = None) -> torch.Tensor:
        """(V, K) REML score surface: recurrence energy then MFMA projection.
        ``n_chunks`` > 1 pipelines voxel chunks over two streams (measured no
        faster at T=400, V=200k: the energy kernel fills every CU)."""
        if getattr(Yt, "ffs_projection", self.projection) != self.projection:
            raise ValueError(f"residual prepared for projection {Yt.ffs_projection!r}, grid is in {self.projection!r}; "
                             "call prep() again after set_projection()")
        S = out if out is not None else torch.empty(V, self.K, dtype=torch.float32, device=self.device)
        if n_chunks is None:
            n_chunks = 1
        _lib.check(_lib.load().ffs_reml_surface(self._handle, _lib.ptr(Yt), Yt.stride(0), V, _lib.ptr(S), S.stride(0),
                                                int(n_chunks), _lib.stream_handle(self.device)), "surface")
        return S

    def surface_f64(self, Y: torch.Tensor, scale: torch.Tensor | None = None,
                    out: torch.Tensor | None = None) -> torch.Tensor:
        """(V, K) float64 REML surface of Y (V, T; float32 or float64 on the
        device) divided by ``scale`` (optional): float64 recurrence energy and
        projection, no residualisation (the reference's float64 search)."""
        if Y.dtype not in (torch.float32, torch.float64) or Y.ndim != 2 or Y.stride(1) != 1:
            raise ValueError("surface_f64 needs a row-contiguous (V, T) float32 or float64 tensor")
        V = Y.shape[0]
        S = out if out is not None else torch.empty(V, self.K, dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().ffs_reml_surface_f64(self._handle, _lib.ptr(Y), int(Y.dtype == torch.float64),
                                                    Y.stride(0), V, _lib.ptr(scale), _lib.ptr(S), S.stride(0),
                                                    _lib.stream_handle(self.device)), "surface_f64")
        return S

    def select_f64(self, S: torch.Tensor, selector: str):
        """``select`` on a float64 surface: (pair index int32, score float64)."""
        if selector not in self.SELECTORS:
            raise ValueError(f"selector must be one of {sorted(self.SELECTORS)}, got {selector!r}")
        V = S.shape[0]
        best = torch.empty(V, dtype=torch.int32, device=self.device)
        score = torch.empty(V, dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().ffs_reml_select_f64(self._handle, _lib.ptr(S), S.stride(0), V,
                                                   self.SELECTORS[selector], _lib.ptr(best), _lib.ptr(score),
                                                   _lib.stream_handle(self.device)), "select_f64")
        return best, score

    PROJECTIONS = {"fp32": 0, "bf16x3": 1, "fp16x3": 2}

    def set_projection(self, mode: str):
        """Arithmetic of the projection GEMM: ``"fp32"`` (fp32 MFMA),
        ``"fp16x3"`` or ``"bf16x3"`` (hi/lo split operands, three 16-bit MFMA
        products, fp32 accumulation)."""
        if mode not in self.PROJECTIONS:
            raise ValueError(f"projection must be one of {sorted(self.PROJECTIONS)}, got {mode!r}")
        _lib.check(_lib.load().ffs_reml_plan_set_projection(self._handle, self.PROJECTIONS[mode]), "set_projection")
        self.projection = mode

    def set_bands(self, half_width: float = 0.10, bin_width: float = 0.02):
        """Lag-1 correlation bands for the learned search: band i holds the
        pairs whose lambda lies within ``half_width`` of ``-1 + i * bin_width``.
        ``half_width <= 0`` removes them."""
        lib = _lib.load()
        _lib.check(lib.ffs_reml_plan_set_bands(self._handle, float(half_width), float(bin_width),
                                               _lib.stream_handle(self.device)), "set_bands")
        self.bands = (float(half_width), float(bin_width)) if half_width > 0 and bin_width > 0 else None

    def band_info(self):
        """(pairs, counts, edge) of the bands as CPU int32 tensors
        (n_bands x slots, n_bands, n_bands x slots), or None without bands."""
        lib = _lib.load()
        nb, sb = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(lib.ffs_reml_plan_band_info(self._handle, ctypes.byref(nb), ctypes.byref(sb), None, None, None),
                   "band_info")
        if nb.value <= 0:
            return None
        pairs = torch.empty(nb.value, sb.value, dtype=torch.int32)
        edge = torch.empty(nb.value, sb.value, dtype=torch.int32)
        cnt = torch.empty(nb.value, dtype=torch.int32)
        _lib.check(lib.ffs_reml_plan_band_info(self._handle, None, None, _lib.ptr(pairs), _lib.ptr(cnt),
                                               _lib.ptr(edge)), "band_info")
        return pairs, cnt, edge

    def window_info(self):
        """(pairs, counts, edge) of the learned search's windows, one per
        lattice cell ai * n_b + bi, as CPU int32 tensors (n_cells x slots,
        n_cells, n_cells x slots)."""
        lib = _lib.load()
        nw, sw = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(lib.ffs_reml_plan_window_info(self._handle, ctypes.byref(nw), ctypes.byref(sw), None, None, None),
                   "window_info")
        pairs = torch.empty(nw.value, sw.value, dtype=torch.int32)
        edge = torch.empty(nw.value, sw.value, dtype=torch.int32)
        cnt = torch.empty(nw.value, dtype=torch.int32)
        _lib.check(lib.ffs_reml_plan_window_info(self._handle, None, None, _lib.ptr(pairs), _lib.ptr(cnt),
                                                 _lib.ptr(edge)), "window_info")
        return pairs, cnt, edge

    def ainv(self) -> torch.Tensor:
        """(X'R^-1X)^-1 per pair, (K, p, p) float64 on the CPU."""
        out = torch.empty(self.K, self.p, self.p, dtype=torch.float64)
        _lib.check(_lib.load().ffs_reml_plan_ainv(self._handle, _lib.ptr(out)), "plan_ainv")
        return out

    SELECTORS = {"exhaustive": 0, "hierarchical": 1, "subset": 2}

    def select(self, S: torch.Tensor, hierarchical: bool | str):
        """Pair index + score per voxel.  ``hierarchical``: True / "hierarchical"
        (descent on the surface, glm/arma.py:2224-2322), False / "exhaustive"
        (global first minimum), or "subset" (the subset descent of
        glm/arma.py:2021-2221)."""
        if isinstance(hierarchical, str):
            if hierarchical not in self.SELECTORS:
                raise ValueError(f"selector must be one of {sorted(self.SELECTORS)}, got {hierarchical!r}")
            mode = self.SELECTORS[hierarchical]
        else:
            mode = int(bool(hierarchical))
        V = S.shape[0]
        best = torch.empty(V, dtype=torch.int32, device=self.device)
        score = torch.empty(V, dtype=torch.float32, device=self.device)
        _lib.check(_lib.load().ffs_reml_select(self._handle, _lib.ptr(S), S.stride(0), V, mode,
                                               _lib.ptr(best), _lib.ptr(score), _lib.stream_handle(self.device)),
                   "select")
        return best, score

    def hier_search(self, Y: torch.Tensor):
        """The hierarchical descent (glm/arma.py:2224-2322) on demand: the pairs
        and scores that ``select(surface(prep(Y)), hierarchical=True)`` gives,
        evaluated on the lattice points the descent visits only (about 45 per
        voxel; same scores bitwise).  Y: (V, T) float32 on the device.  Returns
        (best, score, scale)."""
        lib = _lib.load()
        st = _lib.stream_handle(self.device)
        if not getattr(self, "_hier_built", False):
            _lib.check(lib.ffs_reml_plan_build_hier(self._handle, st), "plan_build_hier")
            self._hier_built = True
        V = Y.shape[0]
        nb = ctypes.c_int64()
        _lib.check(lib.ffs_reml_hier_workspace(self._handle, V, ctypes.byref(nb)), "hier_workspace")
        ws = torch.empty(max(int(nb.value), 1), dtype=torch.uint8, device=self.device)
        best = torch.empty(V, dtype=torch.int32, device=self.device)
        score = torch.empty(V, dtype=torch.float32, device=self.device)
        scale = torch.empty(V, dtype=torch.float32, device=self.device)
        _lib.check(lib.ffs_reml_hier_search(self._handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(best),
                                            _lib.ptr(score), _lib.ptr(scale), _lib.ptr(ws), ws.numel(), st),
                   "hier_search")
        return best, score, scale

    def gls(self, Y: torch.Tensor, scale: torch.Tensor | None, best: torch.Tensor, want_resid_w: bool = False,
            want_residuals: bool = False, want_predicted: bool = False):
        V = Y.shape[0]
        dev = self.device
        f32 = dict(dtype=torch.float32, device=dev)
        out = {"betas": torch.empty(V, self.p, **f32), "tstats": torch.empty(V, self.p, **f32),
               "fstats": torch.empty(V, **f32), "sigma2": torch.empty(V, **f32), "r2": torch.empty(V, **f32)}
        rw = torch.empty(V, self.T, dtype=torch.float64, device=dev) if want_resid_w else None
        res = torch.empty(V, self.T, **f32) if want_residuals else None
        pred = torch.empty(V, self.T, **f32) if want_predicted else None
        n_work = ctypes.c_int64()
        _lib.check(_lib.load().ffs_reml_gls_workspace(self._handle, V, ctypes.byref(n_work)), "gls_workspace")
        work = torch.empty(int(n_work.value), dtype=torch.int32, device=dev)
        _lib.check(_lib.load().ffs_reml_gls(
            self._handle, _lib.ptr(Y), Y.stride(0), _lib.ptr(scale), _lib.ptr(best), V, _lib.ptr(out["betas"]),
            _lib.ptr(out["tstats"]), _lib.ptr(out["fstats"]), _lib.ptr(out["sigma2"]), _lib.ptr(out["r2"]),
            _lib.ptr(rw), self.T, _lib.ptr(res), _lib.ptr(pred), _lib.ptr(work), _lib.stream_handle(dev)), "gls")
        out["resid_w"], out["residuals"], out["predicted"] = rw, res, pred
        return out


def precompute_autocorr_grid(n_timepoints: int, a_grid, b_grid, device=None, *, verbose: bool = False,
                             cholesky_on_cpu: bool = True, dtype: torch.dtype = torch.float32, run_starts=None,
                             tau=None) -> dict:
    """glm/arma.py:1306-1368: the design-independent half of the grid — one
    whitener per valid (a, b) (after zero insertion) and log det R — as
    ``{"param_list", "whiteners", "logdet_Rcorr_batch", "dtype"}``.  The
    whiteners are recurrence handles on the device (``cholesky_on_cpu`` has
    nothing to choose); ``precompute_reml_grid`` accepts the dict as
    ``autocorr_cache`` (its tables are built on the device per design either way)."""
    from .whiten import arma11_whiteners

    a_grid, b_grid = ensure_zero_in_grid(_grid_tensor(a_grid, "a_grid"), _grid_tensor(b_grid, "b_grid"))
    ws, logdet, plist = arma11_whiteners(a_grid.to(dtype), b_grid.to(dtype), n_timepoints, run_starts=run_starts,
                                         tau=tau, device=device, dtype=dtype)
    if verbose:
        print(f"  {len(plist)} whiteners (recurrence) for {len(a_grid) * len(b_grid)} grid cells", flush=True)
    return {"param_list": plist, "whiteners": ws, "logdet_Rcorr_batch": logdet, "dtype": dtype}


def precompute_reml_grid(X, n_timepoints, a_grid, b_grid, device=None, verbose=False, cholesky_on_cpu=True,
                         dtype=torch.float32, debug_memory=False, use_qr=False, run_starts=None,
                         autocorr_cache=None, tau=None, task_indices=None) -> RemlGrid:
    """glm/arma.py:1371-1531: build the per-pair tables for design X."""
    X = torch.as_tensor(X)
    if X.shape[0] != n_timepoints:
        raise ValueError(f"Timepoints mismatch: Data has {n_timepoints}, Design has {X.shape[0]}")
    grid = RemlGrid(X, a_grid, b_grid, device, run_starts=run_starts, tau=tau, task_indices=task_indices)
    if verbose and grid.rank < grid.p:
        print(f"  design has rank {grid.rank} < {grid.p} columns: rank-deficient tables (QR-fallback log-det, "
              "minimum-norm GLS)")
    return grid


def _to_device_f32(Y, device) -> torch.Tensor:
    return _lib.upload_f32(Y, device)


def search_voxels_precomputed_grid(X, Y_batch, precomputed_grid: RemlGrid, device=None, verbose=False,
                                   enable_timing=False, return_profile=False):
    """glm/arma.py:1886-2018: exhaustive search (strict '<', first pair wins ties).

    Returns (best_params (V,2), best_lik (V,)) and, with return_profile, the
    (V, K) surface and the param list.  Outputs stay on the device."""
    grid = precomputed_grid
    Y = _to_device_f32(Y_batch, grid.device)
    if Y.ndim != 2 or Y.shape[1] != grid.T:
        raise ValueError(f"Timepoints mismatch: Data has {Y.shape[-1]}, Design has {grid.T}")
    V = Y.shape[0]
    Yt, _, _ = grid.prep(Y, normalize=False)
    S = grid.surface(Yt, V)
    best, score = grid.select(S, hierarchical=False)
    params = grid.params.to(grid.device)[best.long()]
    if return_profile:
        return params, score, S, list(grid.param_list)
    return params, score


def search_voxels_precomputed_grid_hierarchical(X, Y_batch, precomputed_grid: RemlGrid, device=None,
                                                verbose=False):
    """glm/arma.py:2021-2221: per-voxel subset descent (start (0,0), lattice
    points that are multiples of 2^lev inside +/-2^lev of the running best,
    sequential strict '<').  The scores of the visited pairs are the
    surface's, so the descent runs on the device surface.  Returns
    (best_params (V, 2), best_lik (V,)) on the device."""
    grid = precomputed_grid
    Y = _to_device_f32(Y_batch, grid.device)
    if Y.ndim != 2 or Y.shape[1] != grid.T:
        raise ValueError(f"Timepoints mismatch: Data has {Y.shape[-1]}, Design has {grid.T}")
    V = Y.shape[0]
    Yt, _, _ = grid.prep(Y, normalize=False)
    S = grid.surface(Yt, V)
    best, score = grid.select(S, "subset")
    return grid.params.to(grid.device)[best.long()], score


def select_arma_params_hierarchical_from_surface(surface: torch.Tensor, param_list, device=None,
                                                 grid: RemlGrid | None = None):
    """glm/arma.py:2224-2322 on a (V, K) surface.  ``grid`` (optional) supplies
    the lattice; otherwise it is rebuilt from ``param_list``."""
    _check_reml_mode()
    if grid is None or list(grid.param_list) != list(param_list):
        raise ValueError("select_arma_params_hierarchical_from_surface needs the RemlGrid the surface came from")
    S = surface.to(device=grid.device, dtype=torch.float32).contiguous()
    best, score = grid.select(S, hierarchical=True)
    return grid.params.to(grid.device)[best.long()], score


# ---------------------------------------------------------------------------
# Public search variants (glm/arma.py:1106-1303, 1574-1883, 2577-2833)
# ---------------------------------------------------------------------------


def compute_reml_likelihood(X: torch.Tensor, Y: torch.Tensor, R: torch.Tensor, device=None) -> float:
    """glm/arma.py:1106-1186: REML score of one dense covariance R (T x T):
    2 sum log(diag chol R + 1e-10) + log det(X_w'X_w + 1e-6 I)
    + (T - p) log(||Y_w - X_w beta||^2 + 1e-10), beta solved with the same
    ridge; 1e10 when R is not positive definite.  Y: (T,) or (T, m)."""
    dev = _lib.require_device(device if device is not None else (X.device if X.is_cuda else None))
    f64 = dict(dtype=torch.float64, device=dev)
    Xd = torch.as_tensor(X).to(**f64).contiguous()
    Yd = torch.as_tensor(Y).to(**f64)
    Yd = (Yd[:, None] if Yd.ndim == 1 else Yd).contiguous()
    Rd = torch.as_tensor(R).to(**f64).contiguous()
    T, p = Xd.shape
    m = Yd.shape[1]
    if Yd.shape[0] != T or tuple(Rd.shape) != (T, T):
        raise ValueError(f"Timepoints mismatch: X has {T}, Y has {Yd.shape[0]}, R is {tuple(Rd.shape)}")
    n = T * T + T * (p + m) + p * p + p * m
    work = torch.empty(n, **f64)
    out = torch.empty(1, **f64)
    _lib.check(_lib.load().ffs_reml_dense_likelihood(_lib.ptr(Xd), _lib.ptr(Yd), m, _lib.ptr(Rd), T, p,
                                                      _lib.ptr(work), n, _lib.ptr(out), _lib.stream_handle(dev)),
               "compute_reml_likelihood")
    return float(out.item())


def _default_axes(a_grid, b_grid):
    """The reference's grid defaulting (glm/arma.py:1253-1266)."""
    da, db = get_default_arma_grids()
    a = da if a_grid is None else _grid_tensor(a_grid, "a_grid")
    b = db if b_grid is None else _grid_tensor(b_grid, "b_grid")
    return a, b


def reml_grid_search_batched(data, design, a_grid, b_grid, device=None, verbose=False, dtype=torch.float32,
                             grid_chunk_size=1, voxel_batch_size=None, y_scale=None, run_starts=None, tau=None):
    """glm/arma.py:2577-2833: exhaustive search over the valid pairs of the given
    grids (no zero insertion) on the raw data (optionally / y_scale) with the
    ridge score (X_w'X_w + 1e-6 I in the log
    determinant and in beta); sequential strict '<' in pair order; voxels that
    never improve keep (0, 0) and +inf.  Returns CPU (best_params, best_lik).
    The chunking arguments only shape memory use in the reference and are
    accepted for compatibility."""
    dev = _lib.require_device(device)
    Y = torch.as_tensor(data)
    if y_scale is not None:
        Y = Y / torch.as_tensor(y_scale).to(Y.dtype)[:, None]
    X = torch.as_tensor(design)
    V = Y.shape[0]
    pairs = valid_grid_pairs(_grid_tensor(a_grid, "a_grid"), _grid_tensor(b_grid, "b_grid"))
    best_params = torch.zeros(V, 2, dtype=dtype)
    best_lik = torch.full((V,), float("inf"), dtype=dtype)
    if pairs.shape[0] == 0:
        return best_params, best_lik
    grid = RemlGrid(X, None, None, dev, run_starts=run_starts, tau=tau, pairs=pairs, ridge=1e-6, lattice=True)
    Yd = _to_device_f32(Y, dev)
    # The score is computed as |y_w|^2 - |projection|^2 in fp32; on raw data
    # with a large mean that difference loses most of its digits.  Removing the
    # OLS fit first changes the ridge residual by O(1e-6 |c|) for y -> y - X c
    # (the reference's own float32 residual carries more error than that).
    Yt, _, _ = grid.prep(Yd, normalize=False, residualize=True)
    S = grid.surface(Yt, V)
    best, score = grid.select(S, "exhaustive")
    score = score.cpu()
    ok = torch.isfinite(score)
    best_params[ok] = grid.params[best.cpu().long()[ok]].to(dtype)
    best_lik[ok] = score[ok].to(dtype)
    return best_params, best_lik


def reml_grid_search(X, Y, a_grid=None, b_grid=None, device=None, run_starts=None, tau=None):
    """glm/arma.py:1189-1303: single-voxel exhaustive search with the dense-R
    score of compute_reml_likelihood (ridge 1e-6, raw data), a-outer / b-inner,
    strict '<'; (0.5, 0.0, inf) when no pair is valid.  Returns (a, b, lik)."""
    a_g, b_g = _default_axes(a_grid, b_grid)
    Yv = torch.as_tensor(Y).reshape(1, -1)
    params, lik = reml_grid_search_batched(Yv, X, a_g, b_g, device, run_starts=run_starts, tau=tau)
    if not torch.isfinite(lik[0]):
        return 0.5, 0.0, float("inf")
    return float(params[0, 0]), float(params[0, 1]), float(lik[0])


# glm/arma.py:1754-1765: pairs evaluated first by batch_reml_grid_search
_PRIORITY_PARAMS = [(0.0, 0.3), (0.0, 0.2), (0.0, 0.4), (0.0, 0.1), (0.0, 0.5), (0.8, -0.6), (0.1, 0.3),
                    (0.8, -0.5), (0.8, -0.4), (0.0, 0.0)]


def batch_reml_grid_search(X, Y_batch, a_grid=None, b_grid=None, device=None, precomputed=None,
                           dtype=torch.float32, enable_early_stopping=False, run_starts=None, tau=None):
    """glm/arma.py:1574-1883: exhaustive search of Y_batch (T x V), residualised
    on X (no normalisation), zero inserted into both grids, with the pairs of
    the reference's priority list evaluated first (so exact ties go to the
    first pair in that order; the list is matched by exact float equality as
    in the reference); voxels with residual variance < 1e-6 get (0, 0) and
    likelihood -1.  Returns (best_params (V, 2), best_lik (V,)) on the device."""
    if enable_early_stopping:
        raise NotImplementedError("chunk-level early stopping is not supported (it skips pairs)")
    if dtype != torch.float32:
        raise NotImplementedError("the engine searches in float32")
    dev = _lib.require_device(device)
    a_g, b_g = _default_axes(a_grid, b_grid)
    grid = precomputed if isinstance(precomputed, RemlGrid) else RemlGrid(X, a_g, b_g, dev, run_starts=run_starts,
                                                                          tau=tau)
    Y = _to_device_f32(torch.as_tensor(Y_batch).T, dev)
    V = Y.shape[0]
    Yt, _, _ = grid.prep(Y, normalize=False)
    S = grid.surface(Yt, V)
    order = [k for k, pr in enumerate(grid.param_list) if pr in _PRIORITY_PARAMS]
    order += [k for k in range(grid.K) if k not in set(order)]
    perm = torch.tensor(order, dtype=torch.long, device=dev)
    Sp = S[:, perm].contiguous()
    j, score = grid.select(Sp, "exhaustive")
    best = perm[j.long()]
    params = grid.params.to(device=dev, dtype=torch.float32)[best]
    var = Yt[:grid.T, :V].var(dim=0)
    bg = var < 1e-6
    params[bg] = 0.0
    score = torch.where(bg, torch.full_like(score, -1.0), score)
    return params, score


# ---------------------------------------------------------------------------
# Ljung-Box helpers (glm/arma.py:5972-6049) and the device call
# ---------------------------------------------------------------------------


def ljung_box_max_lag(n_time: int, n_regressors: int = 0, min_run: int | None = None) -> int:
    """glm/arma.py:5972-6000."""
    if min_run is None:
        min_run = n_time
    if min_run < 2:
        return 0
    hh = n_regressors + 2 + min(min_run // 8, int(round(3.0 * math.log(min_run))))
    return min(hh, min_run // 2)


def _resolve_ljung_box_lag(n_time: int, max_lag) -> int:
    """glm/arma.py:6003-6014."""
    hh = int(max_lag) if max_lag else 0
    if hh < 2 or hh > n_time // 2:
        hh = 2 + min(n_time // 8, int(round(3.0 * math.log(n_time))))
        hh = min(hh, n_time // 2)
    return hh


def build_ljung_box_tau(n_time: int, run_starts=None, tau=None, device=None):
    """glm/arma.py:6017-6049."""
    starts = (run_starts.tolist() if isinstance(run_starts, torch.Tensor)
              else (list(run_starts) if run_starts is not None else []))
    starts = sorted({int(s) for s in starts if 0 <= int(s) < n_time}) or [0]
    if tau is None and len(starts) == 1 and starts[0] == 0:
        return None
    st = torch.tensor(starts, dtype=torch.long)
    idx = torch.arange(n_time, dtype=torch.long)
    rid = torch.bucketize(idx, st, right=True) - 1
    within = torch.as_tensor(tau).to(torch.long).flatten().cpu() if tau is not None else idx - st[rid]
    out = within + _LJ_RUN_SEPARATION * rid
    return out.to(device) if device is not None else out


def ljung_box_bins(n: int, max_lag, tau=None):
    """Host binning of lag pairs by time lag (the reference bins on the host too,
    glm/arma.py:6103-6123).  Returns (hh, bin_off, pairs, wbin)."""
    hh = _resolve_ljung_box_lag(n, max_lag)
    t = np.arange(n) if tau is None else np.asarray(torch.as_tensor(tau).tolist(), dtype=np.int64)
    by_bin: list[list[tuple[int, int]]] = [[] for _ in range(hh + 1)]
    for kk in range(1, hh + 1):
        dj = t[kk:] - t[: n - kk]
        for j in np.nonzero((dj > 0) & (dj <= hh))[0]:
            by_bin[int(dj[j])].append((int(j), int(j + kk)))
    off = np.zeros(hh + 2, dtype=np.int32)
    flat: list[int] = []
    for d in range(1, hh + 1):
        off[d] = len(flat) // 2
        for i, j in by_bin[d]:
            flat += [i, j]
        off[d + 1] = len(flat) // 2
    nj = np.array([len(b) for b in by_bin], dtype=np.float64)
    w = np.where(nj > 1, 1.0 / np.maximum(nj, 1), 0.0)
    w[0] = 0.0
    return hh, off, np.asarray(flat, dtype=np.int32), w


def ljung_box_device(resid_w: torch.Tensor, max_lag=None, tau=None) -> torch.Tensor:
    """Ljung-Box of (V, n) whitened residuals on the device (float32 out); the
    kernel reads fp64 rows, so other dtypes / layouts are converted first."""
    if resid_w.dtype != torch.float64 or resid_w.stride(1) != 1:
        resid_w = resid_w.to(torch.float64).contiguous()
    V, n = resid_w.shape
    dev = resid_w.device
    if n < 10:
        return torch.zeros(V, dtype=torch.float32, device=dev)
    hh, off, pairs, w = ljung_box_bins(n, max_lag, tau)
    if hh < 1:
        return torch.zeros(V, dtype=torch.float32, device=dev)
    d_off = torch.from_numpy(off).to(dev)
    d_pairs = torch.from_numpy(pairs if pairs.size else np.zeros(2, np.int32)).to(dev)
    d_w = torch.from_numpy(w).to(dev)
    out = torch.empty(V, dtype=torch.float32, device=dev)
    _lib.check(_lib.load().ffs_reml_ljung_box(_lib.ptr(resid_w), resid_w.stride(0), V, n, hh, _lib.ptr(d_off),
                                              _lib.ptr(d_pairs), _lib.ptr(d_w), _lib.ptr(out),
                                              _lib.stream_handle(dev)), "ljung_box")
    return out


# ---------------------------------------------------------------------------
# Results container + fit
# ---------------------------------------------------------------------------


class ARMA11Results:
    """Container for ARMA(1,1) GLM results (glm/arma.py:585-692): same fields."""

    def __init__(self):
        self.betas = None
        self.tstats = None
        self.r2 = None
        self.r2_partial = None
        self.r2_partial_nuisance = None
        self.r2_semipartial = None
        self.r2_semipartial_nuisance = None
        self.arma_params = None
        self.arma_lambda = None
        self.reml_likelihood = None
        self.residuals = None
        self.predicted = None
        self.residuals_whitened = None
        self.ljung_box = None
        self.ljung_box_dof = None
        self.sigma2 = None
        self.var_betas = None
        self.original_shape = None
        self.fstats = None
        self.set_fstats = None
        self.set_sd = None
        self.set_coef = None
        self.nuisance_sets = None
        self.voxel_mean = None
        self.dof = None
        self.tr = None
        self.voxel_mask = None
        self.full_shape = None
        self.fitted_column_indices = None
        self.n_regressors_full = None
        self.contrast_labels = None
        self.contrast_betas = None
        self.contrast_tstats = None
        self.contrast_fstats = None
        self.contrast_r2_partial = None
        self.contrast_r2_semipartial = None
        self.affine = None
        self.ols_results = None
        self.dsort_betas = None
        self.dsort_tstats = None
        self.dsort_labels = None
        self.nods_results = None
        self.reml_lklhd_surface = None
        self.reml_surface_params = None
        # selected pair index per voxel (search order), kept for diagnostics
        self.pair_index = None


def fit_glm_arma11(data, design, tr, a_grid=None, b_grid=None, estimate_per_voxel=True, batch_size=None,
                   want_residuals=False, want_predicted=False, want_ljung_box=False, want_r2_partial=False,
                   r2_partial_mode="full", want_r2_semipartial=False, r2_semipartial_mode="full",
                   precomputed_arma_params=None, device=None, verbose=False, use_double=False,
                   force_exhaustive_search=False, task_indices=None, save_profile_likelihoods=False,
                   run_starts=None, tau=None, grid: RemlGrid | None = None, projection: str = "fp16x3",
                   selector: str | None = None, proposal=None, hierarchical_search: str = "descent",
                   **kwargs) -> ARMA11Results:
    """glm/arma.py:3650-5878 ``fit_glm_arma11`` on the MI355X engine.

    Per-voxel REML search over the (a, b) grid (hierarchical descent by default,
    global argmin with ``force_exhaustive_search``), then an FP64 GLS refit at
    the selected pair.  float32 mode divides each voxel by its std before the
    fit and unscales betas / sigma2 / residuals afterwards (glm/arma.py:3947-3954,
    5827-5862); ``use_double`` runs the reference's float64 search instead: raw
    series (no normalisation, no residualisation), float64 surface
    (``RemlGrid.surface_f64``) and selection, float64 outputs; the GLS reads
    the float32 samples (exact for float32 data).  Outputs are CPU tensors.
    ``projection`` ("fp16x3" default, "fp32", "bf16x3") is the arithmetic of
    the surface GEMM (an engine option; the reference computes it in fp32;
    split fp16 keeps ~22-bit operands with fp32 accumulation and meets the
    same score / argmin bars).  ``selector`` picks the search rule:
    "hierarchical" (default, the reference's GPU path), "subset" (the
    reference's CPU default, glm/arma.py:2021-2221), "exhaustive" (same as
    ``force_exhaustive_search``) or "learned" (no counterpart in the
    reference: the proposal network's 5 x 5 window, then a lambda band, then
    the full grid for what is left; its target is the exhaustive global
    argmin).  ``proposal`` (learned only): a ``ProposalMLP``, a path written by
    ``ProposalMLP.save``, or None to train one for this design first
    (``learned.train_proposal``, a few seconds).  ``hierarchical_search``
    (engine option for the hierarchical selector): "descent" (default) scores
    only the lattice points the descent visits (``RemlGrid.hier_search``, the
    same pairs and scores as the surface descent), "surface" computes the
    full surface first; the surface is computed whenever it is an output
    (``save_profile_likelihoods``).
    """
    if hierarchical_search not in ("descent", "surface"):
        raise ValueError(f"hierarchical_search must be 'descent' or 'surface', got {hierarchical_search!r}")
    dev = _lib.require_device(device)
    if selector == "learned":
        if force_exhaustive_search:
            raise ValueError("selector='learned' and force_exhaustive_search are exclusive")
        if use_double:
            raise NotImplementedError("the learned search scores normalised float32 data; use_double is not supported")
        if save_profile_likelihoods:
            raise ValueError("selector='learned' evaluates no full surface; save_profile_likelihoods needs a "
                             "full-grid selector")
    elif selector not in (None, "hierarchical", "subset", "exhaustive"):
        raise ValueError(f"selector must be 'hierarchical', 'subset', 'exhaustive' or 'learned', got {selector!r}")
    out_dtype = torch.float64 if use_double else torch.float32
    data = torch.as_tensor(data)
    design = torch.as_tensor(design)
    original_shape = None
    if data.ndim == 4:
        original_shape = tuple(data.shape[:3])
        data = data.reshape(-1, data.shape[-1])
    elif data.ndim != 2:
        raise ValueError(f"Data must be 2D or 4D, got {data.ndim}D")
    V, T = data.shape
    if design.shape[0] != T:
        raise ValueError(f"Timepoints mismatch: Data has {T}, Design has {design.shape[0]}")
    p = design.shape[1]
    if not estimate_per_voxel:
        unsupported = [k for k, on in (("precomputed_arma_params", precomputed_arma_params is not None),
                                       ("want_r2_semipartial", want_r2_semipartial),
                                       ("save_profile_likelihoods", save_profile_likelihoods),
                                       ("glt_matrices", bool(kwargs.get("glt_matrices"))),
                                       ("dsort", kwargs.get("dsort") is not None),
                                       ("want_ols", bool(kwargs.get("want_ols")))) if on]
        if unsupported:
            raise NotImplementedError(f"estimate_per_voxel=False with {', '.join(unsupported)}")
        return _fit_global(data, design, tr, a_grid, b_grid, dev, original_shape=original_shape,
                           task_indices=task_indices, want_residuals=want_residuals, want_predicted=want_predicted,
                           want_ljung_box=want_ljung_box, want_r2_partial=want_r2_partial,
                           r2_partial_mode=r2_partial_mode, use_double=use_double, run_starts=run_starts, tau=tau,
                           column_sets=kwargs.get("column_sets"), batch_size=batch_size)

    res = ARMA11Results()
    res.original_shape = original_shape
    res.fitted_column_indices = task_indices
    res.n_regressors_full = p
    res.dof = max(1, T - p)
    res.tr = tr

    if precomputed_arma_params is not None:
        pp = torch.as_tensor(precomputed_arma_params).detach().cpu()
        if pp.shape != (V, 2):
            raise ValueError(f"precomputed_arma_params must have shape ({V}, 2), got {tuple(pp.shape)}")
        uniq, inv = torch.unique(pp.to(torch.float64), dim=0, return_inverse=True)
        bad = (uniq.abs() >= 1).any(dim=1)
        if bool(bad.any()):
            a, b = uniq[bad][0].tolist()
            raise ValueError(f"Invalid ARMA(1,1) parameters: a={a}, b={b}")
        grid = RemlGrid(design, None, None, dev, run_starts=run_starts, tau=tau, task_indices=task_indices,
                        pairs=uniq.to(pp.dtype if pp.dtype.is_floating_point else torch.float64), lattice=False)
        fixed_best = inv.to(torch.int32)
    else:
        if a_grid is None or b_grid is None:
            da, db = get_default_arma_grids()
            a_grid = da if a_grid is None else a_grid
            b_grid = db if b_grid is None else b_grid
        if grid is None:
            grid = RemlGrid(design, a_grid, b_grid, dev, run_starts=run_starts, tau=tau, task_indices=task_indices)
        if grid.projection != projection:
            grid.set_projection(projection)
        fixed_best = None
    if p > WIDE_P and selector == "learned":
        raise NotImplementedError(f"the learned selector needs p <= {WIDE_P} design columns (p = {p})")

    searcher = None
    if selector == "learned" and fixed_best is None:
        from ..learned import LocalSearcher, ProposalMLP, train_proposal

        if grid.bands is None:
            grid.set_bands(0.10, 0.02)
        if proposal is None:
            mlp = train_proposal(grid)
        elif isinstance(proposal, ProposalMLP):
            mlp = proposal
        else:
            mlp = ProposalMLP.load(proposal, dev)
        res_searchers: dict = {}

        def searcher(Yb):
            ls = res_searchers.get(Yb.shape[0])
            if ls is None:
                res_searchers.clear()
                ls = res_searchers[Yb.shape[0]] = LocalSearcher(grid, mlp, Yb.shape[0])
            r = ls.run(Yb)
            return r.best, r.score, r.scale

    K = grid.K
    nb = batch_size or min(V, 1 << 20)
    f = torch.float32
    res.betas = torch.zeros(V, len(task_indices) if task_indices else p, dtype=out_dtype)
    res.tstats = torch.zeros_like(res.betas)
    res.r2 = torch.zeros(V, dtype=torch.float32)
    res.arma_params = torch.zeros(V, 2, dtype=out_dtype)
    res.arma_lambda = torch.zeros(V, dtype=out_dtype)
    res.reml_likelihood = torch.zeros(V, dtype=out_dtype)
    res.sigma2 = torch.zeros(V, dtype=out_dtype)
    res.fstats = torch.zeros(V, dtype=out_dtype)
    res.pair_index = torch.zeros(V, dtype=torch.int64)
    if want_residuals:
        res.residuals = torch.zeros(V, T, dtype=out_dtype)
        res.residuals_whitened = torch.zeros(V, T, dtype=out_dtype)
    if want_predicted:
        res.predicted = torch.zeros(V, T, dtype=out_dtype)
    surface_acc = (torch.empty(V, K, dtype=out_dtype) if save_profile_likelihoods and fixed_best is None
                   else None)
    lj_lag = lj_tau = None
    if want_ljung_box:
        starts = sorted({int(s) for s in (run_starts.tolist() if isinstance(run_starts, torch.Tensor)
                                          else run_starts)} if run_starts is not None else {0})
        min_run = (min(b - a for a, b in zip(starts, starts[1:] + [T])) if len(starts) > 1 else T)
        lj_lag = ljung_box_max_lag(T, p, min_run)
        lj_tau = build_ljung_box_tau(T, run_starts, tau)
        res.ljung_box = torch.zeros(V, dtype=torch.float32)
        res.ljung_box_dof = max(1, _resolve_ljung_box_lag(T, lj_lag) - 2)
    glt_labels, glt_matrices = kwargs.get("glt_labels"), kwargs.get("glt_matrices")
    C = Qc = None
    if glt_labels and glt_matrices is not None:
        # glm/arma.py:4201-4253: single-row contrasts over the full design
        rows = []
        for m in glt_matrices:
            c = torch.as_tensor(m, dtype=torch.float64)
            if c.ndim == 2 and c.shape[0] == 1:
                c = c[0]
            if c.ndim != 1:
                raise NotImplementedError(
                    f"Multi-row GLT contrasts (F-tests) not yet supported. "
                    f"Got shape {tuple(c.shape)}, expected (n_regressors,) or (1, n_regressors)")
            if c.shape[0] != p:
                raise ValueError(f"GLT has {c.shape[0]} columns, design has {p}")
            rows.append(c)
        C = torch.stack(rows)
        # Var(c'beta) / sigma^2 = c' (X'R^-1X)^-1 c per pair (glm/arma.py:5030-5043, 5454-5460)
        Qc = torch.einsum("cr,krs,cs->kc", C, grid.ainv(), C).to(dev)
        C = C.to(dev)
        n_c = C.shape[0]
        res.contrast_labels = list(glt_labels)
        res.contrast_betas = torch.zeros(V, n_c, dtype=out_dtype)
        res.contrast_tstats = torch.zeros(V, n_c, dtype=out_dtype)
        if want_r2_partial:
            res.contrast_r2_partial = torch.zeros(V, n_c, dtype=out_dtype)
        if want_r2_semipartial:
            res.contrast_r2_semipartial = torch.zeros(V, n_c, dtype=out_dtype)
    column_sets = kwargs.get("column_sets")
    set_mats = None
    if column_sets:
        # glm/arma.py:3543-3562: per set, (A_SS + 1e-8 I)^-1 of every pair (A =
        # (X'R^-1X)^-1) and the covariance of its centred design columns; None
        # for a set naming a column the design does not have (a -dsort column)
        Ai = grid.ainv()
        Xd = torch.as_tensor(design, dtype=torch.float64)
        Xc = Xd - Xd.mean(0, keepdim=True)
        set_mats = []
        for cols in column_sets:
            keep = [int(c) for c in cols if int(c) < p]
            if len(keep) != len(cols) or not keep:
                set_mats.append(None)
                continue
            ix = torch.as_tensor(keep, dtype=torch.long)
            A = Ai[:, ix][:, :, ix]
            Minv = torch.linalg.inv(A + 1e-8 * torch.eye(len(keep), dtype=torch.float64))
            Cs = Xc[:, ix].T @ Xc[:, ix] / Xd.shape[0]
            set_mats.append((ix.to(dev), Minv.to(dev), Cs.to(dev)))
        res.set_fstats = torch.full((V, len(column_sets)), float("nan"), dtype=out_dtype)
        res.set_sd = torch.full_like(res.set_fstats, float("nan"))
        res.set_coef = torch.full_like(res.set_fstats, float("nan"))
    dsort = kwargs.get("dsort")
    scale_all = None
    if dsort is not None:
        dsort = torch.as_tensor(dsort)
        if dsort.ndim == 2:
            dsort = dsort.unsqueeze(1)
        if dsort.shape[0] != V or dsort.shape[2] != T:
            raise ValueError(f"dsort must have shape ({V}, n_dsort, {T}), got {tuple(dsort.shape)}")
        dsort = _dsort_constant_guard(dsort)
        scale_all = torch.ones(V, dtype=torch.float64)
    params_dev = grid.params.to(dev)
    # the hierarchical descent on demand: same pairs and scores as the descent
    # on the full surface, when no surface is kept
    use_descent = (fixed_best is None and searcher is None and not use_double and surface_acc is None
                   and not force_exhaustive_search and selector in (None, "hierarchical")
                   and hierarchical_search == "descent")
    with torch.cuda.device(dev):
        for b0 in range(0, V, nb):
            b1 = min(V, b0 + nb)
            Yb = _to_device_f32(data[b0:b1], dev)
            Vb = b1 - b0
            if searcher is not None:
                best, score, scale = searcher(Yb)
            elif fixed_best is None and use_double:
                # the reference's float64 search: raw series, float64 surface
                # and selection (glm/arma.py:3869)
                src = data[b0:b1]
                Yd = (src.to(dev, torch.float64) if src.dtype == torch.float64 else Yb).contiguous()
                S = grid.surface_f64(Yd)
                best, score = grid.select_f64(S, "exhaustive" if force_exhaustive_search else (selector or "hierarchical"))
                scale = None
                if surface_acc is not None:
                    surface_acc[b0:b1].copy_(S)
                del S, Yd
            elif fixed_best is None and use_descent:
                best, score, scale = grid.hier_search(Yb)
            elif fixed_best is None:
                Yt, scale, _ = grid.prep(Yb, normalize=not use_double)
                S = grid.surface(Yt, Vb)
                del Yt
                best, score = grid.select(S, "exhaustive" if force_exhaustive_search else (selector or "hierarchical"))
                if surface_acc is not None:
                    surface_acc[b0:b1].copy_(S)
                del S
            else:
                best = fixed_best[b0:b1].to(dev)
                score = torch.zeros(Vb, dtype=f, device=dev)
                scale = None
                if not use_double:
                    _, scale, _ = grid.prep(Yb, normalize=True)
            if scale_all is not None and scale is not None and not use_double:
                scale_all[b0:b1].copy_(scale.double())
            g = grid.gls(Yb, None if use_double else scale, best, want_resid_w=want_ljung_box or want_residuals,
                         want_residuals=want_residuals, want_predicted=want_predicted)
            ti = task_indices
            res.betas[b0:b1].copy_(g["betas"][:, ti] if ti else g["betas"])
            res.tstats[b0:b1].copy_(g["tstats"][:, ti] if ti else g["tstats"])
            res.fstats[b0:b1].copy_(g["fstats"])
            res.sigma2[b0:b1].copy_(g["sigma2"])
            res.r2[b0:b1].copy_(g["r2"])
            sel = params_dev[best.long()]
            res.arma_params[b0:b1].copy_(sel)
            res.reml_likelihood[b0:b1].copy_(score)
            res.pair_index[b0:b1].copy_(best)
            if want_residuals:
                res.residuals[b0:b1].copy_(g["residuals"])
                res.residuals_whitened[b0:b1].copy_(g["resid_w"])
            if want_predicted:
                res.predicted[b0:b1].copy_(g["predicted"])
            if want_ljung_box:
                res.ljung_box[b0:b1].copy_(ljung_box_device(g["resid_w"], lj_lag, lj_tau))
            if C is not None:
                # contrast estimates and t in the normalised units of the fit
                # (glm/arma.py:5429-5471), estimates unscaled like betas (:5835)
                sb = scale.double() if scale is not None else torch.ones(Vb, dtype=torch.float64, device=dev)
                cb = (g["betas"].double() / sb[:, None]) @ C.T
                var = (g["sigma2"].double() / sb**2)[:, None] * Qc[best.long()]
                ct = cb / (var.clamp(min=0.0).sqrt() + 1e-10)
                res.contrast_betas[b0:b1].copy_(cb * sb[:, None])
                res.contrast_tstats[b0:b1].copy_(ct)
                if want_r2_partial or want_r2_semipartial:
                    cp = ct**2 / (ct**2 + res.dof)
                    if want_r2_partial:
                        res.contrast_r2_partial[b0:b1].copy_(cp)
                    if want_r2_semipartial:
                        rem = (1.0 - g["r2"].double()).clamp(min=0.0)[:, None]
                        res.contrast_r2_semipartial[b0:b1].copy_(cp * rem)
            if set_mats is not None:
                # glm/arma.py:3564-3585 in the normalised units of the fit; SD and
                # coef unscaled like betas (:5845-5848)
                sb = scale.double() if scale is not None else torch.ones(Vb, dtype=torch.float64, device=dev)
                bn = g["betas"].double() / sb[:, None]
                s2n = g["sigma2"].double() / sb**2
                for gi, m in enumerate(set_mats):
                    if m is None:
                        continue
                    ix, Minv, Cs = m
                    bs = bn.index_select(1, ix)
                    Fv = torch.einsum("bi,bij,bj->b", bs, Minv[best.long()], bs) / (ix.numel() * s2n + 1e-30)
                    sdv = torch.einsum("bi,ij,bj->b", bs, Cs, bs).clamp(min=0).sqrt() * sb
                    res.set_fstats[b0:b1, gi].copy_(Fv)
                    res.set_sd[b0:b1, gi].copy_(sdv)
                    if ix.numel() == 1:
                        res.set_coef[b0:b1, gi].copy_(bs[:, 0] * sb)
            if want_r2_partial or want_r2_semipartial:
                part = _partial_r2(g["tstats"].to(out_dtype), g["r2"], res.dof, p, task_indices, want_r2_partial,
                                   r2_partial_mode, want_r2_semipartial, r2_semipartial_mode)
                for name, val in part.items():
                    if val is None:
                        continue
                    if getattr(res, name) is None:
                        setattr(res, name, torch.zeros((V,) + tuple(val.shape[1:]), dtype=val.dtype))
                    getattr(res, name)[b0:b1].copy_(val)
    if kwargs.get("want_ols"):
        # glm/arma.py:4288-4323: the OLS baseline on the same data and design
        from .core import fit_glm

        res.ols_results = fit_glm(data, design, tr, device=dev, use_double=use_double, max_poly_degree=-1,
                                  want_residuals=bool(kwargs.get("want_ols_residuals")),
                                  want_predicted=bool(kwargs.get("want_ols_predicted")),
                                  want_r2_partial=want_r2_partial, r2_partial_mode=r2_partial_mode,
                                  want_r2_semipartial=want_r2_semipartial, r2_semipartial_mode=r2_semipartial_mode,
                                  glt_labels=glt_labels, glt_matrices=glt_matrices, task_indices=task_indices)
        res.ols_results.original_shape = original_shape
    a, b = res.arma_params[:, 0], res.arma_params[:, 1]
    res.arma_lambda = (b + a) * (1 + a * b) / (1 + 2 * a * b + b**2 + 1e-10)  # glm/arma.py:4769-4777
    if dsort is not None:
        # glm/arma.py:5780-5825: the base fit is the no-dsort fit; snapshot it
        # when asked, then redo the final GLS per voxel with the extended design
        nods = None
        if kwargs.get("want_dsort_nods"):
            nods = ARMA11Results()
            for attr, val in vars(res).items():
                setattr(nods, attr, val.clone() if isinstance(val, torch.Tensor) else val)
        run_bounds = kwargs.get("dsort_run_bounds")
        _dsort_pass(grid, data, design, dsort, res, scale_all, dev, out_dtype, task_indices=task_indices,
                    labels=kwargs.get("dsort_labels"), C=C, column_sets=column_sets,
                    want_r2_partial=want_r2_partial, r2_partial_mode=r2_partial_mode,
                    want_r2_semipartial=want_r2_semipartial, r2_semipartial_mode=r2_semipartial_mode,
                    want_residuals=want_residuals, want_predicted=want_predicted,
                    lj=(lj_lag, lj_tau) if want_ljung_box else None, batch=nb, run_bounds=run_bounds)
        res.nods_results = nods
    if fixed_best is not None:
        res.reml_likelihood = None
    if surface_acc is not None:
        res.reml_lklhd_surface = surface_acc
        res.reml_surface_params = list(grid.param_list)
    return res


def _fit_global(data, design, tr, a_grid, b_grid, dev, *, original_shape, task_indices, want_residuals, want_predicted,
                want_ljung_box, want_r2_partial, r2_partial_mode, use_double, run_starts, tau, column_sets,
                batch_size) -> ARMA11Results:
    """glm/arma.py:5603-5780, ``estimate_per_voxel=False``: one (a, b) for every
    voxel, found by ``reml_grid_search`` on the mean time series of the raw
    data, then one GLS with the shared whitener: the design and each voxel's
    (std-normalised) series go through the pair's recurrence on the device
    (ffs_reml_whiten_cols) and the solve uses A = X_w'X_w + 1e-6 I in fp64.
    Per voxel: beta over all p columns, sigma2 = |r_w|^2 / dof, t = beta /
    (sqrt(sigma2 diag A^-1) + 1e-10), F = beta_S' (sigma2 A^-1_SS + 1e-8 I)^-1
    beta_S / |S|, R^2 = CoD against X beta, partial R^2, column sets with A^-1,
    Ljung-Box on r_w; betas / sigma2 / series unscaled as in the per-voxel fit.
    The reference stores the full beta row in this mode, so ``task_indices``
    must name every design column (a shorter list fails there too)."""
    from .whiten import arma11_whitener

    f64 = torch.float64
    out_dtype = f64 if use_double else torch.float32
    V, T = data.shape
    p = design.shape[1]
    if task_indices and len(task_indices) != p:
        raise ValueError(f"estimate_per_voxel=False stores all {p} design columns; task_indices names "
                         f"{len(task_indices)} of them")
    Yc = data.to(out_dtype)
    Xc = design.to(out_dtype)
    a_opt, b_opt, lik = reml_grid_search(Xc, Yc.mean(dim=0), a_grid, b_grid, dev, run_starts=run_starts, tau=tau)
    res = ARMA11Results()
    res.original_shape = original_shape
    res.fitted_column_indices = task_indices
    res.n_regressors_full = p
    dof = max(1, T - p)
    res.dof = dof
    res.tr = tr
    res.arma_params = torch.empty(V, 2, dtype=out_dtype)
    res.arma_params[:, 0] = a_opt
    res.arma_params[:, 1] = b_opt
    res.reml_likelihood = torch.full((V,), float(lik), dtype=out_dtype)
    res.arma_lambda = torch.full((V,), (b_opt + a_opt) * (1 + a_opt * b_opt) / (1 + 2 * a_opt * b_opt + b_opt**2 + 1e-10),
                                 dtype=out_dtype)
    res.pair_index = torch.zeros(V, dtype=torch.int64)
    w = arma11_whitener(a_opt, b_opt, T, run_starts=run_starts, tau=tau, device=dev, dtype=f64)
    Xd = Xc.to(dev, f64)
    Xw = w.cols(Xd)
    Ai = torch.linalg.inv(Xw.T @ Xw + 1e-6 * torch.eye(p, dtype=f64, device=dev))
    S = list(task_indices) if task_indices else list(range(p))
    ti = torch.as_tensor(S, dtype=torch.long, device=dev)
    A_SS = Ai.index_select(0, ti).index_select(1, ti)
    eye_S = torch.eye(len(S), dtype=f64, device=dev)
    dAi = torch.diagonal(Ai)
    res.betas = torch.zeros(V, p, dtype=out_dtype)
    res.tstats = torch.zeros(V, p, dtype=out_dtype)
    res.sigma2 = torch.zeros(V, dtype=out_dtype)
    res.fstats = torch.zeros(V, dtype=out_dtype)
    res.r2 = torch.zeros(V, dtype=torch.float32)
    if want_residuals:
        res.residuals = torch.zeros(V, T, dtype=out_dtype)
        res.residuals_whitened = torch.zeros(V, T, dtype=out_dtype)
    if want_predicted:
        res.predicted = torch.zeros(V, T, dtype=out_dtype)
    lj = None
    if want_ljung_box:
        starts = sorted({int(x) for x in (run_starts.tolist() if isinstance(run_starts, torch.Tensor)
                                          else run_starts)} if run_starts is not None else {0})
        min_run = (min(e - b for b, e in zip(starts, starts[1:] + [T])) if len(starts) > 1 else T)
        lj = (ljung_box_max_lag(T, p, min_run), build_ljung_box_tau(T, run_starts, tau))
        res.ljung_box = torch.zeros(V, dtype=torch.float32)
        res.ljung_box_dof = max(1, _resolve_ljung_box_lag(T, lj[0]) - 2)
    set_mats = None
    if column_sets:
        Xcen = Xd - Xd.mean(0, keepdim=True)
        set_mats = []
        for cols in column_sets:
            keep = [int(c) for c in cols if int(c) < p]
            if len(keep) != len(cols) or not keep:
                set_mats.append(None)
                continue
            ix = torch.as_tensor(keep, dtype=torch.long, device=dev)
            Minv = torch.linalg.inv(Ai.index_select(0, ix).index_select(1, ix)
                                    + 1e-8 * torch.eye(len(keep), dtype=f64, device=dev))
            set_mats.append((ix, Minv, Xcen[:, ix].T @ Xcen[:, ix] / T))
        res.set_fstats = torch.full((V, len(column_sets)), float("nan"), dtype=out_dtype)
        res.set_sd = torch.full_like(res.set_fstats, float("nan"))
        res.set_coef = torch.full_like(res.set_fstats, float("nan"))
    scale = None if use_double else Yc.std(dim=1).clamp(min=1e-8)
    nb = max(1, min(batch_size or (1 << 18), (1 << 28) // (T * 8 * 4)))
    for b0 in range(0, V, nb):
        b1 = min(V, b0 + nb)
        Y = Yc[b0:b1].to(dev, f64)
        sc = torch.ones(b1 - b0, 1, dtype=f64, device=dev)
        if scale is not None:
            sc = scale[b0:b1].to(dev, f64)[:, None]
            Y = Y / sc
        Yw = w.rows(Y)
        beta = (Yw @ Xw) @ Ai
        rw = Yw - beta @ Xw.T
        pred = beta @ Xd.T
        sig = rw.pow(2).sum(1) / dof
        t = beta / (torch.sqrt(sig[:, None] * dAi[None, :]) + 1e-10)
        bS = beta.index_select(1, ti)
        F = torch.einsum("bi,bi->b", bS, torch.linalg.solve(sig[:, None, None] * A_SS + 1e-8 * eye_S,
                                                              bS.unsqueeze(2)).squeeze(2)) / len(S)
        ss_tot = (Y - Y.mean(1, keepdim=True)).pow(2).sum(1)
        r2 = torch.where(ss_tot <= 1e-10, torch.zeros_like(ss_tot),
                         (1.0 - (Y - pred).pow(2).sum(1) / ss_tot.clamp(min=1e-10)).clamp(max=1.0))
        res.betas[b0:b1].copy_(beta * sc)
        res.tstats[b0:b1].copy_(t)
        res.sigma2[b0:b1].copy_(sig * sc[:, 0] ** 2)
        res.fstats[b0:b1].copy_(F)
        res.r2[b0:b1].copy_(r2)
        if set_mats is not None:
            for gi, m in enumerate(set_mats):
                if m is None:
                    continue
                ix, Minv, Cs = m
                bs = beta.index_select(1, ix)
                res.set_fstats[b0:b1, gi] = (torch.einsum("bi,ij,bj->b", bs, Minv, bs)
                                             / (ix.numel() * sig + 1e-30)).cpu().to(out_dtype)
                res.set_sd[b0:b1, gi] = (torch.einsum("bi,ij,bj->b", bs, Cs, bs).clamp(min=0).sqrt()
                                         * sc[:, 0]).cpu().to(out_dtype)
                if ix.numel() == 1:
                    res.set_coef[b0:b1, gi].copy_(bs[:, 0] * sc[:, 0])
        if want_r2_partial:
            part = _partial_r2(t.to(out_dtype), r2.float(), dof, p, task_indices, True, r2_partial_mode, False,
                               "full")
            for name, val in part.items():
                if val is None:
                    continue
                if getattr(res, name) is None:
                    setattr(res, name, torch.zeros((V,) + tuple(val.shape[1:]), dtype=val.dtype))
                getattr(res, name)[b0:b1].copy_(val)
        if lj is not None:
            res.ljung_box[b0:b1].copy_(ljung_box_device(rw, lj[0], lj[1]))
        if want_residuals:
            res.residuals[b0:b1].copy_(((Y - pred) * sc))
            res.residuals_whitened[b0:b1].copy_(rw * sc)
        if want_predicted:
            res.predicted[b0:b1].copy_(pred * sc)
    return res


def _dsort_constant_guard(dsort: torch.Tensor) -> torch.Tensor:
    """glm/arma.py:3207-3245: a -dsort voxel series that is constant and nonzero
    through time (std < 1e-8, mean |x| > 1e-8) is replaced by its dataset's mean
    series over the other voxels (all voxels when none is left), with a warning."""
    stds = dsort.std(dim=2)
    means = dsort.abs().mean(dim=2)
    bad = (stds < 1e-8) & (means > 1e-8)
    n_bad = int(bad.sum().item())
    if n_bad == 0:
        return dsort
    dsort = dsort.clone()
    for d in range(dsort.shape[1]):
        bad_d = bad[:, d]
        if not bool(bad_d.any()):
            continue
        good_d = ~bad_d
        mean_ts = dsort[good_d, d, :].mean(dim=0) if bool(good_d.any()) else dsort[:, d, :].mean(dim=0)
        dsort[bad_d, d, :] = mean_ts
    warnings.warn(f"-dsort: {n_bad} voxel timeseries were ~constant through time and were replaced by their "
                  "dataset mean (matches 3dREMLfit).", stacklevel=3)
    return dsort


def _dsort_pass(grid: RemlGrid, data, design, dsort, res, scale_all, dev, out_dtype, *, task_indices, labels, C,
                column_sets, want_r2_partial, r2_partial_mode, want_r2_semipartial, r2_semipartial_mode,
                want_residuals, want_predicted, lj, batch, run_bounds=None) -> None:
    """glm/arma.py:3248-3539: the final GLS redone per voxel on [X | dsort_v]
    at the (a, b) selected without dsort.  The voxel's columns (design, its
    dsort series, its normalised data) are whitened on the device by the
    plan's recurrence of its pair (ffs_reml_whiten_cols), then solved per voxel
    in fp64 (A + 1e-8 I by Cholesky) by the batched solve kernel
    (ffs_reml_dsort_gls, one block per voxel); voxels whose Cholesky fails take
    the reference's pinv(A) solve (torch on the device).  Overwrites betas / tstats / sigma2 / r2 /
    fstats / dof (and the GLT, partial R^2, column-set, Ljung-Box, residual and
    prediction outputs), fills dsort_betas / dsort_tstats; everything computed
    in the normalised units of the fit and unscaled like betas (:5827-5848).
    ``run_bounds`` ([t0, t1, ..., T]): ``dsort`` holds one run-concatenated
    series per set and each set becomes one zero-padded column per run
    (run-major within the set, :3284-3303, 3389-3398), expanded per batch."""
    lib = _lib.load()
    st = _lib.stream_handle(dev)
    f64 = torch.float64
    V, n_sets, T = dsort.shape
    n_runs = len(run_bounds) - 1 if run_bounds is not None else 1
    q = n_sets * n_runs
    Xd = torch.as_tensor(design, dtype=f64).to(dev)
    p = Xd.shape[1]
    pe = p + q
    dof = max(1, T - pe)
    res.dof = dof
    res.dsort_labels = list(labels) if labels else [f"dsort#{k}" for k in range(q)]
    res.dsort_betas = torch.zeros(V, q, dtype=out_dtype)
    res.dsort_tstats = torch.zeros(V, q, dtype=out_dtype)
    fci = list(task_indices) if task_indices else list(range(p))
    fci_d = torch.as_tensor(fci, dtype=torch.long, device=dev)
    fci_h = np.ascontiguousarray(fci, dtype=np.int32)
    eye_f = torch.eye(len(fci), dtype=f64, device=dev)
    want_ai = bool(column_sets) or C is not None
    want_rw = lj is not None or want_residuals
    ws1 = ctypes.c_int64()
    _lib.check(lib.ffs_reml_dsort_workspace(grid.handle, 1, q, len(fci), ctypes.byref(ws1)), "dsort_workspace")
    per_voxel = 8 * (ws1.value + (q + 3) * T + 2 * pe + (pe * pe if want_ai else 0))
    nb = max(1, min(batch, max(1, (1 << 28) // per_voxel)))
    for b0 in range(0, V, nb):
        b1 = min(V, b0 + nb)
        B = b1 - b0
        s = scale_all[b0:b1].to(dev)
        Y = torch.as_tensor(data[b0:b1]).to(dev, f64) / s[:, None]
        if run_bounds is not None:
            Dc = dsort[b0:b1].to(dev, f64)
            D = torch.zeros(B, q, T, dtype=f64, device=dev)
            for si in range(n_sets):
                for r in range(n_runs):
                    t0, t1 = int(run_bounds[r]), int(run_bounds[r + 1])
                    D[:, si * n_runs + r, t0:t1] = Dc[:, si, t0:t1]
        else:
            D = dsort[b0:b1].to(dev, f64)
        D = D.contiguous()
        Y = Y.contiguous()
        pair = res.pair_index[b0:b1].to(dev, torch.int32).contiguous()
        beta = torch.empty(B, pe, dtype=f64, device=dev)
        tst = torch.empty_like(beta)
        sig, Fv, r2 = (torch.empty(B, dtype=f64, device=dev) for _ in range(3))
        info = torch.empty(B, dtype=torch.int32, device=dev)
        Ai = torch.empty(B, pe, pe, dtype=f64, device=dev) if want_ai else None
        rw = torch.empty(B, T, dtype=f64, device=dev) if want_rw else None
        pred = torch.empty(B, T, dtype=f64, device=dev) if (want_predicted or want_residuals) else None
        work = torch.empty(B * ws1.value, dtype=f64, device=dev)
        _lib.check(lib.ffs_reml_dsort_gls(grid.handle, _lib.ptr(pair), B, q, _lib.ptr(D), _lib.ptr(Y),
                                          fci_h.ctypes.data, len(fci), _lib.ptr(beta), _lib.ptr(tst), _lib.ptr(sig),
                                          _lib.ptr(Fv), _lib.ptr(r2), _lib.ptr(Ai), _lib.ptr(rw), None,
                                          _lib.ptr(pred), _lib.ptr(info), _lib.ptr(work), B * ws1.value, st),
                   "dsort_gls")
        del work
        bad = (info == 1).nonzero().flatten()
        if bad.numel():
            # Cholesky of A + 1e-8 I failed: the reference's pinv(A) solve (glm/arma.py:3410-3416)
            _dsort_pinv(grid, lib, st, Xd, D[bad], Y[bad], pair[bad], fci_d, eye_f, dof,
                        beta, tst, sig, Fv, r2, Ai, rw, pred, bad)
        if bool((info == 2).any()):
            nan = info == 2
            for t_ in (beta, tst):
                t_[nan] = float("nan")
            for t_ in (sig, Fv, r2):
                t_[nan] = float("nan")
        sc = s[:, None]
        bb, tb = beta[:, :p], tst[:, :p]
        res.betas[b0:b1].copy_(((bb[:, task_indices] if task_indices else bb) * sc))
        res.tstats[b0:b1].copy_(tb[:, task_indices] if task_indices else tb)
        res.dsort_betas[b0:b1].copy_(beta[:, p:] * sc)
        res.dsort_tstats[b0:b1].copy_(tst[:, p:])
        res.sigma2[b0:b1].copy_(sig * s**2)
        res.fstats[b0:b1].copy_(Fv)
        res.r2[b0:b1].copy_(r2)
        if column_sets:
            # per voxel: the extended design's (X'X)^-1 block (glm/arma.py:3461-3480)
            Xo = torch.cat([Xd.unsqueeze(0).expand(B, T, p), D.transpose(1, 2)], 2)
            Xo = Xo - Xo.mean(1, keepdim=True)
            for gi, cset in enumerate(column_sets):
                ix = torch.as_tensor([int(c) for c in cset], dtype=torch.long, device=dev)
                Ag = Ai.index_select(1, ix).index_select(2, ix)
                bg = beta.index_select(1, ix)
                eye_g = torch.eye(ix.numel(), dtype=f64, device=dev)
                qg = torch.einsum("bi,bi->b", bg, torch.linalg.solve(Ag + 1e-8 * eye_g, bg.unsqueeze(2)).squeeze(2))
                res.set_fstats[b0:b1, gi].copy_((qg / (ix.numel() * sig + 1e-30)))
                sigt = torch.einsum("bti,bi->bt", Xo.index_select(2, ix), bg)
                res.set_sd[b0:b1, gi].copy_((sigt.std(dim=1, correction=0) * s))
                if ix.numel() == 1:
                    res.set_coef[b0:b1, gi].copy_(bg[:, 0] * s)
        if want_r2_partial or want_r2_semipartial:
            part = _partial_r2(tb.to(out_dtype), r2.float(), dof, p, task_indices, want_r2_partial, r2_partial_mode,
                               want_r2_semipartial, r2_semipartial_mode)
            for name, val in part.items():
                if val is None:
                    continue
                if getattr(res, name) is None:
                    setattr(res, name, torch.zeros((V,) + tuple(val.shape[1:]), dtype=val.dtype))
                getattr(res, name)[b0:b1].copy_(val)
        if C is not None:
            # GLTs on the base-design block (dsort betas cannot enter a GLT)
            Cd = C.to(dev, f64)
            cb = bb @ Cd.T
            cv = sig[:, None] * torch.einsum("cr,brs,cs->bc", Cd, Ai[:, :p, :p], Cd)
            ct = cb / (cv.clamp(min=0.0).sqrt() + 1e-10)
            res.contrast_betas[b0:b1].copy_(cb * sc)
            res.contrast_tstats[b0:b1].copy_(ct)
            if want_r2_partial and res.contrast_r2_partial is not None:
                res.contrast_r2_partial[b0:b1].copy_((ct**2 / (ct**2 + dof)))
            if want_r2_semipartial and res.contrast_r2_semipartial is not None:
                rem = (1.0 - r2).clamp(min=0.0)[:, None]
                res.contrast_r2_semipartial[b0:b1].copy_((ct**2 / (ct**2 + dof) * rem))
        if lj is not None and res.ljung_box is not None:
            res.ljung_box[b0:b1].copy_(ljung_box_device(rw, lj[0], lj[1]))
        if want_residuals:
            res.residuals[b0:b1].copy_(((Y - pred) * sc))
            res.residuals_whitened[b0:b1].copy_(rw * sc)
        if want_predicted:
            res.predicted[b0:b1].copy_(pred * sc)


def _dsort_pinv(grid, lib, st, Xd, D, Y, pair, fci_d, eye_f, dof, beta, tst, sig, Fv, r2, Ai, rw, pred, rows):
    """-dsort voxels whose Cholesky of A + 1e-8 I failed: the reference's pinv
    branch (glm/arma.py:3410-3416), A^+ = pinv(A) of A = E_w'E_w without the
    1e-8 term, beta = A^+ E_w'y_w, the other outputs as the batched kernel
    defines them.  The reference switches a whole sub-batch to pinv when one
    voxel fails; per voxel here (the full-rank voxels' solutions differ by the
    1e-8 ridge only)."""
    f64 = torch.float64
    B, q, T = D.shape
    p = Xd.shape[1]
    pe = p + q
    cols = torch.cat([Xd.T.unsqueeze(0).expand(B, p, T), D, Y.unsqueeze(1)], dim=1).contiguous()
    W = torch.empty_like(cols)
    _lib.check(lib.ffs_reml_whiten_cols(grid.handle, _lib.ptr(pair.contiguous()), B, pe + 1, _lib.ptr(cols), T,
                                        _lib.ptr(W), st), "whiten_cols")
    Xw = W[:, :pe, :].transpose(1, 2)  # (B, T, pe)
    yw = W[:, pe, :]
    Ap = torch.linalg.pinv(Xw.transpose(1, 2) @ Xw)
    b = (Ap @ (Xw.transpose(1, 2) @ yw.unsqueeze(2))).squeeze(2)
    r_w = yw - (Xw @ b.unsqueeze(2)).squeeze(2)
    s2 = r_w.pow(2).sum(1) / dof
    pr = b[:, :p] @ Xd.T + torch.einsum("bq,bqt->bt", b[:, p:], D)
    ss_res = (Y - pr).pow(2).sum(1)
    ss_tot = (Y - Y.mean(1, keepdim=True)).pow(2).sum(1)
    bt = b.index_select(1, fci_d)
    At = Ap.index_select(1, fci_d).index_select(2, fci_d)
    quad = torch.einsum("bi,bij,bj->b", bt, torch.linalg.inv(At + 1e-8 * eye_f), bt)
    beta[rows] = b
    tst[rows] = b / (torch.sqrt(s2[:, None] * torch.diagonal(Ap, dim1=1, dim2=2)) + 1e-10)
    sig[rows] = s2
    Fv[rows] = quad / (s2 * fci_d.numel() + 1e-30)
    r2[rows] = torch.where(ss_tot <= 1e-10, torch.zeros_like(ss_tot),
                           (1.0 - ss_res / ss_tot.clamp(min=1e-10)).clamp(max=1.0)).to(f64)
    if Ai is not None:
        Ai[rows] = Ap
    if rw is not None:
        rw[rows] = r_w
    if pred is not None:
        pred[rows] = pr


def _partial_r2(t_full, r2, dof, p, task_indices, want_partial, partial_mode, want_semi, semi_mode) -> dict:
    """glm/arma.py:5054-5175 on one batch: partial R^2 = t^2 / (t^2 + dof) per
    regressor (all p t-statistics), split into task / nuisance columns; mode
    "task" divides the task values by max(1 - sum(nuisance), 0.01); semi-partial
    = partial * max(1 - R^2, 0), same split and rescaling."""
    t2 = t_full**2
    full = t2 / (t2 + dof)
    if task_indices is not None:
        nuis = sorted(set(range(p)) - set(task_indices))
        task_part = full[:, list(task_indices)]
        nuis_part = full[:, nuis] if nuis else None
    else:
        task_part, nuis_part = full, None
    out = {}
    if want_partial:
        if partial_mode == "task" and nuis_part is not None:
            out["r2_partial"] = task_part / torch.clamp(1.0 - nuis_part.sum(dim=1, keepdim=True), min=0.01)
        else:
            out["r2_partial"] = task_part
        out["r2_partial_nuisance"] = nuis_part
    if want_semi:
        remain = torch.clamp(1.0 - r2.to(t_full.dtype).unsqueeze(1), min=0.0)
        semi_task = task_part * remain
        semi_nuis = nuis_part * remain if nuis_part is not None else None
        if semi_mode == "task" and semi_nuis is not None:
            out["r2_semipartial"] = semi_task / torch.clamp(1.0 - semi_nuis.sum(dim=1, keepdim=True), min=0.01)
        else:
            out["r2_semipartial"] = semi_task
        out["r2_semipartial_nuisance"] = semi_nuis
    return out


class RemlGroup:
    """Voxel groups with their own designs on one timeline and (a, b) grid,
    fitted with one launch per stage over all groups (ffs_reml_group_*): voxels
    laid out group by group (each padded to 256 columns), per-design tables
    picked per block, the stages captured into one HIP graph and replayed.
    ``grids`` (from ``RemlGrid.create_many``) must outlive the group."""

    def __init__(self, grids: list, group_of_voxel):
        self.grids = list(grids)
        g0 = self.grids[0]
        self.device, self.K, self.T, self.p = g0.device, g0.K, g0.T, g0.p
        gv = np.ascontiguousarray(np.asarray(group_of_voxel, dtype=np.int32))
        self.V = int(gv.shape[0])
        handles = (ctypes.c_void_p * len(self.grids))(*[g.handle.value for g in self.grids])
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.load().ffs_reml_group_create(ctypes.byref(h), len(self.grids),
                                                         ctypes.cast(handles, ctypes.c_void_p), self.V,
                                                         gv.ctypes.data, _lib.stream_handle(self.device)),
                       "group_create")
        self._handle = h
        self.tuning = None
        self._out = None

    def _outputs(self):
        if self._out is None:
            f32 = dict(dtype=torch.float32, device=self.device)
            V, p = self.V, self.p
            self._out = {"best": torch.empty(V, dtype=torch.int32, device=self.device),
                         "score": torch.empty(V, **f32), "scale": torch.empty(V, **f32),
                         "betas": torch.empty(V, p, **f32), "tstats": torch.empty(V, p, **f32),
                         "fstats": torch.empty(V, **f32), "sigma2": torch.empty(V, **f32), "r2": torch.empty(V, **f32)}
        return self._out

    def fit(self, Y: torch.Tensor, selector: str = "hierarchical", surface: torch.Tensor | None = None) -> dict:
        """All stages for Y (V, T) float32 on the device; returns device tensors
        (reused across calls) best, score, scale, betas, tstats, fstats,
        sigma2, r2; ``surface`` (optional (V, K) float32) receives the scores."""
        if Y.shape != (self.V, self.T) or Y.dtype != torch.float32 or Y.stride(1) != 1:
            raise ValueError(f"Y must be a row-contiguous ({self.V}, {self.T}) float32 device tensor")
        o = self._outputs()
        mode = RemlGrid.SELECTORS[selector]
        _lib.check(_lib.load().ffs_reml_group_fit(
            self._handle, _lib.ptr(Y), Y.stride(0), mode, _lib.ptr(o["best"]), _lib.ptr(o["score"]),
            _lib.ptr(o["scale"]), _lib.ptr(o["betas"]), _lib.ptr(o["tstats"]), _lib.ptr(o["fstats"]),
            _lib.ptr(o["sigma2"]), _lib.ptr(o["r2"]), _lib.ptr(surface), 0 if surface is None else surface.stride(0),
            _lib.stream_handle(self.device)), "group_fit")
        return o

    def tune(self, Y: torch.Tensor, selector: str = "hierarchical", reps: int = 3) -> dict:
        """Autotune the launch configuration on Y (see ffs_reml_group_tune);
        returns and keeps the report (candidates with ms per fit, the choice)."""
        import json

        o = self._outputs()
        buf = ctypes.create_string_buffer(4096)
        _lib.check(_lib.load().ffs_reml_group_tune(
            self._handle, _lib.ptr(Y), Y.stride(0), RemlGrid.SELECTORS[selector], _lib.ptr(o["best"]),
            _lib.ptr(o["score"]), _lib.ptr(o["scale"]), _lib.ptr(o["betas"]), _lib.ptr(o["tstats"]),
            _lib.ptr(o["fstats"]), _lib.ptr(o["sigma2"]), _lib.ptr(o["r2"]), int(reps), buf, len(buf),
            _lib.stream_handle(self.device)), "group_tune")
        self.tuning = json.loads(buf.value.decode())
        return self.tuning

    def close(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            _lib.load().ffs_reml_group_destroy(h)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# fit_glm_arma11 options the grouped engine computes (others take the per-group loop)
_GROUPED_ENGINE_KW = {"a_grid", "b_grid", "device", "run_starts", "tau", "task_indices", "projection",
                      "force_exhaustive_search", "selector", "save_profile_likelihoods", "batch_size", "verbose",
                      "estimate_per_voxel", "autotune"}


def _grouped_engine_fit(data_t, designs_by_group, groups, group_indices, kw) -> ARMA11Results:
    """fit_glm_arma11_grouped on the grouped engine: one plan per design (tables
    of all designs in one launch), one graph of single launches over all
    groups, outputs in voxel order."""
    dev = _lib.require_device(kw.get("device"))
    V = data_t.shape[0]
    X0 = torch.as_tensor(designs_by_group[groups[0]])
    T, p = X0.shape
    if data_t.shape[1] != T:
        raise ValueError(f"Timepoints mismatch: Data has {data_t.shape[1]}, Design has {T}")
    ti = kw.get("task_indices")
    grids = RemlGrid.create_many([torch.as_tensor(designs_by_group[h]) for h in groups], kw["a_grid"], kw["b_grid"],
                                 dev, run_starts=kw.get("run_starts"), tau=kw.get("tau"), task_indices=ti)
    proj = kw.get("projection", "fp16x3")
    for g in grids:
        g.set_projection(proj)
    gid = {h: i for i, h in enumerate(groups)}
    gv = np.asarray([gid[int(h)] for h in group_indices.tolist()], dtype=np.int32)
    sel = "exhaustive" if kw.get("force_exhaustive_search") else (kw.get("selector") or "hierarchical")
    if sel not in RemlGrid.SELECTORS:
        raise ValueError(f"selector must be one of {sorted(RemlGrid.SELECTORS)}, got {sel!r}")
    engine = RemlGroup(grids, gv)
    Y = _to_device_f32(data_t, dev)
    tuning = engine.tune(Y, sel) if kw.get("autotune") else None
    S = torch.empty(V, grids[0].K, dtype=torch.float32, device=dev) if kw.get("save_profile_likelihoods") else None
    o = engine.fit(Y, sel, surface=S)
    res = ARMA11Results()
    res.dof = max(1, T - p)
    res.tr = kw.get("tr")
    res.fitted_column_indices = ti
    res.n_regressors_full = p
    ix = torch.as_tensor(ti, dtype=torch.long, device=dev) if ti else None
    res.betas = (o["betas"].index_select(1, ix) if ix is not None else o["betas"]).cpu()
    res.tstats = (o["tstats"].index_select(1, ix) if ix is not None else o["tstats"]).cpu()
    res.fstats, res.sigma2, res.r2 = o["fstats"].cpu(), o["sigma2"].cpu(), o["r2"].cpu()
    best = o["best"].long()
    params = grids[0].params.to(dev, torch.float32)
    res.arma_params = params[best.clamp(min=0)].cpu()
    res.reml_likelihood = o["score"].cpu()
    res.pair_index = best.cpu()
    a, b = res.arma_params[:, 0], res.arma_params[:, 1]
    res.arma_lambda = (b + a) * (1 + a * b) / (1 + 2 * a * b + b**2 + 1e-10)  # glm/arma.py:4769-4777
    if S is not None:
        res.reml_lklhd_surface = S.cpu()
        res.reml_surface_params = list(grids[0].param_list)
    res.grouped_engine = {"groups": len(groups), "tuning": tuning}
    engine.close()
    return res


def fit_glm_arma11_grouped(data, designs_by_group: dict, group_indices, tr: float, group_label: str = "group",
                           **fit_kwargs) -> ARMA11Results:
    """glm/arma.py:2836-3183: REML with one design per voxel group (per-voxel
    HRF, slicewise baseline, ...).  Each group is fitted on its voxel subset
    with its own plan (the design-dependent tables), on the same (a, b) grid,
    and the per-voxel outputs are scattered back to the full voxel axis
    (``VOXEL_SCATTER_ATTRS``, zeros elsewhere), as the reference merges them.
    ``precomputed_arma_params`` is sliced per group.

    Without options that need the per-group fit (series outputs, Ljung-Box,
    partial R^2, GLTs, column sets, -dsort, the OLS baseline, precomputed
    parameters, use_double, the learned selector) the fit runs on the grouped
    engine (``RemlGroup``): the tables of every design in one launch, then one
    launch per stage over all groups, captured in a HIP graph;
    ``autotune=True`` times its launch configurations first and reports the
    choice in ``result.grouped_engine``.  ``grouped_engine=False`` keeps the
    per-group fits."""
    group_indices = torch.as_tensor(group_indices).long().cpu()
    data_t = torch.as_tensor(data)
    n_voxels = data_t.shape[0]
    if group_indices.shape[0] != n_voxels:
        raise ValueError(f"group_indices length {group_indices.shape[0]} != data n_voxels {n_voxels}")
    groups = sorted(int(h) for h in group_indices.unique().tolist())
    missing = [h for h in groups if h not in designs_by_group]
    if missing:
        raise KeyError(f"{group_label} ids {missing} present in voxels but missing from designs_by_group "
                       f"(have {sorted(designs_by_group.keys())})")
    first = torch.as_tensor(designs_by_group[groups[0]])
    shape = tuple(first.shape)
    for h in groups[1:]:
        if tuple(torch.as_tensor(designs_by_group[h]).shape) != shape:
            raise ValueError(f"designs_by_group[{h}] shape {tuple(torch.as_tensor(designs_by_group[h]).shape)} "
                             f"!= {shape} — per-{group_label} designs must share dims")
    kw = dict(fit_kwargs)
    kw.pop("spatial_metadata", None)
    pre = kw.pop("precomputed_arma_params", None)
    if pre is not None:
        pre = torch.as_tensor(pre).reshape(-1, 2)
        if pre.shape[0] != n_voxels:
            raise ValueError(f"precomputed_arma_params has {pre.shape[0]} rows but data has {n_voxels} voxels")
    # glm/arma.py:2931-2940, 3035-3036: voxel-wise regressors sliced per group
    dsort_full = kw.pop("dsort", None)
    if dsort_full is not None:
        dsort_full = torch.as_tensor(dsort_full)
        if dsort_full.ndim == 2:
            dsort_full = dsort_full.unsqueeze(1)
        if dsort_full.shape[0] != n_voxels:
            raise ValueError(f"dsort has {dsort_full.shape[0]} voxels but data has {n_voxels}")
    if kw.get("a_grid") is None or kw.get("b_grid") is None:
        da, db = get_default_arma_grids()
        kw.setdefault("a_grid", da)
        kw.setdefault("b_grid", db)
        kw["a_grid"] = da if kw["a_grid"] is None else kw["a_grid"]
        kw["b_grid"] = db if kw["b_grid"] is None else kw["b_grid"]
    # one launch per stage over all groups unless an option needs the per-group fit
    use_engine = kw.pop("grouped_engine", True)
    if (use_engine and pre is None and dsort_full is None and set(kw) <= _GROUPED_ENGINE_KW
            and kw.get("estimate_per_voxel", True) and shape[1] <= WIDE_P and n_voxels > 0
            and kw.get("selector") != "learned"):
        return _grouped_engine_fit(data_t, designs_by_group, groups, group_indices, dict(kw, tr=tr))
    kw.pop("autotune", None)
    per, masks = {}, {}
    order = torch.argsort(group_indices, stable=True)
    counts = torch.bincount(group_indices - group_indices.min()) if n_voxels else torch.zeros(0, dtype=torch.long)
    g0 = int(group_indices.min()) if n_voxels else 0
    offs = torch.cumsum(counts, 0) - counts
    for h in groups:
        # voxel ids of the group in ascending order (the order a boolean mask gives)
        m = order[int(offs[h - g0]):int(offs[h - g0]) + int(counts[h - g0])]
        sub_kw = dict(kw)
        if pre is not None:
            sub_kw["precomputed_arma_params"] = pre.index_select(0, m)
        if dsort_full is not None:
            sub_kw["dsort"] = dsort_full.index_select(0, m)
        per[h] = fit_glm_arma11(data_t.index_select(0, m.to(data_t.device)), designs_by_group[h], tr, **sub_kw)
        masks[h] = m
    merged = ARMA11Results()
    tpl = per[groups[0]]
    for name in ("dof", "tr", "fitted_column_indices", "n_regressors_full", "contrast_labels", "ljung_box_dof",
                 "dsort_labels"):
        setattr(merged, name, getattr(tpl, name))
    for attr in VOXEL_SCATTER_ATTRS + ["pair_index"]:
        ref = next((getattr(per[h], attr) for h in groups if getattr(per[h], attr, None) is not None), None)
        if ref is None:
            continue
        full = torch.zeros((n_voxels,) + tuple(ref.shape[1:]), dtype=ref.dtype)
        for h in groups:
            v = getattr(per[h], attr, None)
            if v is not None:
                full.index_copy_(0, masks[h], v.cpu().to(full.dtype))
        setattr(merged, attr, full)
    nods = [h for h in groups if getattr(per[h], "nods_results", None) is not None]
    if nods:
        # glm/arma.py:3087-3115: the per-group no-dsort snapshots scattered likewise
        nm = ARMA11Results()
        tpl_n = per[nods[0]].nods_results
        for name in ("dof", "tr", "fitted_column_indices", "n_regressors_full", "contrast_labels",
                     "ljung_box_dof"):
            setattr(nm, name, getattr(tpl_n, name, None))
        for attr in VOXEL_SCATTER_ATTRS + ["pair_index"]:
            ref = getattr(tpl_n, attr, None)
            if ref is None:
                continue
            full = torch.zeros((n_voxels,) + tuple(ref.shape[1:]), dtype=ref.dtype)
            for h in nods:
                full.index_copy_(0, masks[h], getattr(per[h].nods_results, attr).cpu().to(full.dtype))
            setattr(nm, attr, full)
        merged.nods_results = nm
    ols = [h for h in groups if getattr(per[h], "ols_results", None) is not None]
    if ols:
        # glm/arma.py:3139-3175: the per-group OLS results scattered to the voxel axis
        from .core import GLMResults

        om = GLMResults()
        tpl_o = per[ols[0]].ols_results
        om.dof, om.contrast_labels = tpl_o.dof, tpl_o.contrast_labels
        for attr in ("betas", "tstats", "stderr", "sigma2", "fstats", "r2", "meanvol", "r2_run", "residuals",
                     "predicted", "r2_partial", "r2_partial_nuisance", "r2_semipartial", "r2_semipartial_nuisance",
                     "contrast_betas", "contrast_tstats"):
            ref = getattr(tpl_o, attr, None)
            if ref is None:
                continue
            full = torch.zeros((n_voxels,) + tuple(ref.shape[1:]), dtype=ref.dtype)
            for h in ols:
                full.index_copy_(0, masks[h], getattr(per[h].ols_results, attr).cpu().to(full.dtype))
            setattr(om, attr, full)
        merged.ols_results = om
    surf = next((h for h in groups if per[h].reml_lklhd_surface is not None), None)
    if surf is not None:
        merged.reml_surface_params = per[surf].reml_surface_params
        ref = per[surf].reml_lklhd_surface
        full = torch.zeros((n_voxels, ref.shape[1]), dtype=ref.dtype)
        for h in groups:
            if per[h].reml_lklhd_surface is not None:
                full.index_copy_(0, masks[h], per[h].reml_lklhd_surface.cpu().to(full.dtype))
        merged.reml_lklhd_surface = full
    return merged


def fit_glm_arma11_perhrf(data, designs_by_hrf: dict, hrf_indices, tr: float, **fit_kwargs) -> ARMA11Results:
    """glm/arma.py:3186-3204: per-voxel-HRF alias of the grouped fit."""
    return fit_glm_arma11_grouped(data, designs_by_hrf, hrf_indices, tr, group_label="HRF", **fit_kwargs)


def compute_ljung_box_statistic(residuals, max_lag=None, tau=None) -> np.ndarray:
    """glm/arma.py:6144-6178: Ljung-Box statistic (h - 2 DOF, 0 = not
    computable) of prewhitened residuals (V, T), on the device."""
    dev = _lib.require_device(None)
    r = torch.as_tensor(residuals).to(dev, torch.float64).contiguous()
    return ljung_box_device(r, max_lag, tau).cpu().numpy().astype(np.float32)


def prewhiten_with_arma11(X, Y, a: float, b: float, run_starts=None, tau=None):
    """glm/arma.py:2326-2400: (L^-1 X, L^-1 Y, L) with R = L L' the ARMA(1,1)
    correlation of the timeline.  L^-1 is applied by the engine's recurrence;
    L (dense lower Cholesky factor, on X's device) is returned for callers that
    solve with it, as the reference returns it."""
    from .whiten import arma11_whitener

    X = torch.as_tensor(X)
    Y = torch.as_tensor(Y)
    n = X.shape[0]
    dtype = X.dtype if X.dtype.is_floating_point else torch.float32
    R = build_arma11_covariance(a, b, n, torch.device("cpu"), torch.float64, run_starts=run_starts, tau=tau)
    if R is None:
        raise ValueError(f"Invalid ARMA(1,1) parameters: a={a}, b={b}")
    L = torch.linalg.cholesky(R).to(device=X.device, dtype=dtype)
    w = arma11_whitener(a, b, n, run_starts=run_starts, tau=tau, device=None, dtype=torch.float64)
    return w.cols(X), w.cols(Y), L


def compare_ols_vs_arma11(data, design, tr: float, device=None) -> dict:
    """glm/arma.py:5881-5966: OLS and ARMA(1,1) fits of the same data, the
    mean |t| ratio (ARMA / OLS), the mean R^2 difference and a printed summary."""
    from .core import fit_glm

    dev = _lib.require_device(device)
    ols = fit_glm(data, design, tr, device=dev, verbose=False)
    arma = fit_glm_arma11(data, design, tr, device=dev)
    t_ols, t_arma = ols.tstats.abs().mean(), arma.tstats.abs().mean()
    tstat_ratio = t_arma / (t_ols + 1e-10)
    r2_improvement = arma.r2.mean() - ols.r2.mean()
    a_mean, b_mean = arma.arma_params[:, 0].mean(), arma.arma_params[:, 1].mean()
    rows = [
        ("", "OLS", "ARMA(1,1)", "ARMA vs OLS"),
        ("mean R^2", f"{ols.r2.mean():.4f}", f"{arma.r2.mean():.4f}", f"{r2_improvement:+.4f}"),
        ("mean |t|", f"{t_ols:.3f}", f"{t_arma:.3f}", f"x{tstat_ratio:.3f}"),
    ]
    summary = "\n".join(f"{r[0]:<10}{r[1]:>10}{r[2]:>12}{r[3]:>14}" for r in rows)
    summary += f"\nARMA(1,1) mean (a, b) = ({a_mean:.3f}, {b_mean:.3f}), mean lag-1 correlation {arma.arma_lambda.mean():.3f}"
    print(summary)
    return {"ols": ols, "arma11": arma, "tstat_ratio": float(tstat_ratio), "r2_improvement": float(r2_improvement),
            "summary": summary}


def save_arma_rvar(*args, **kwargs):
    """glm/arma.py:6181-6377 (re-exported from fastfuncstuff_amd.outputs)."""
    from ..outputs import save_arma_rvar as _save

    return _save(*args, **kwargs)


def load_arma_params(*args, **kwargs):
    """glm/arma.py:6380-6454 (re-exported from fastfuncstuff_amd.outputs)."""
    from ..outputs import load_arma_params as _load

    return _load(*args, **kwargs)
