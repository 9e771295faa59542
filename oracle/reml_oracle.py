"""CPU oracle for the ARMA(1,1) REML search + GLS refit (test infrastructure only).

This module is the *checker*: a from-scratch torch-CPU restatement of the
reference algorithm in ``fastfuncstuff/glm/arma.py`` and ``glm/whiten.py``.
It builds the dense run-block-diagonal correlation matrix R for every grid
pair, factors it with a dense Cholesky, whitens with the dense L^-1 and scores
every voxel exactly as the reference does.  It is deliberately slow and
simple.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it; the product path never does.

Two precisions share one code path:

* ``dtype=torch.float64`` -- the FP64 oracle (the reference's ``use_double``
  branch: no std normalisation, no residualisation, float64 everywhere);
* ``dtype=torch.float32`` -- the "reference-semantics" float32 restatement
  (std normalisation, float32 residualisation, float32 GEMMs with float64
  sums), which is also the timed CPU baseline.

Pinning (what checks each function, see ``tests/test_oracle_pinning.py``):

* grids / pair filter / lambda: known answers from the reference's tests
  (lambda(0.5,0)=0.5, lambda(0,0)=0, the 0.5/0.2 value), pair counts
  (561 pairs and 33 exact lambda=0 ties on the 33x33 grid) checked by brute
  force enumeration in exact rational arithmetic;
* ``build_censor_run_info``: the reference test's example
  ``([0,5],10,[0,1,3,4,5,6,8,9]) -> ([0,4],[0,1,3,4,0,1,3,4])``;
* covariance: AR(1) closed form R[i,j]=a^|i-j|, MA(1) tridiagonal, censor
  lag-2 gap = lambda*a, exact-zero cross-run blocks, symmetry / PD;
* REML score: AR(1) closed-form inverse and determinant (independent of any
  Cholesky), MA(1) eigen-closed-form determinant, (a,b)=(0,0) == OLS;
* GLS: (0,0) betas == numpy lstsq OLS; whitened GLS == generalised least
  squares solved with an explicit R^-1 (numpy.linalg.solve);
* hierarchical selector: the reference's own bound (<=0.2% disagreement
  between the surface descent and the subset-GEMM descent) plus a brute-force
  per-voxel restatement on small grids;
* Ljung-Box: zero residuals -> 0, AR(0.7) residuals -> large, and a brute
  force double loop over lag pairs for censored tau.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

# ---------------------------------------------------------------------------
# Grids, lambda, validity filter
# ---------------------------------------------------------------------------

DEFAULT_ARMA_A_GRID = (0.0, 0.8, 9)  # glm/arma.py:234-238
DEFAULT_ARMA_B_GRID = (-0.8, 0.8, 17)  # glm/arma.py:239-243


def default_grids() -> tuple[torch.Tensor, torch.Tensor]:
    """glm/arma.py:484-513 ``get_default_arma_grids`` (float32 linspace)."""
    return torch.linspace(*DEFAULT_ARMA_A_GRID), torch.linspace(*DEFAULT_ARMA_B_GRID)


def ensure_zero_in_grid(a_grid, b_grid, tolerance: float = 1e-6):
    """glm/arma.py:516-582: snap |g|<tol to 0, or insert 0 and sort."""

    def fix(g: torch.Tensor) -> torch.Tensor:
        near = torch.abs(g) < tolerance
        if torch.any(near):
            g = g.clone()
            g[near] = 0.0
            return g
        return torch.sort(torch.cat([torch.zeros(1, dtype=g.dtype), g]))[0]

    return fix(a_grid), fix(b_grid)


def arma_lambda(a, b):
    """glm/arma.py:695-729 ``_compute_arma11_lambda`` (no epsilon)."""
    return (a + b) * (1 + a * b) / (1 + b**2 + 2 * a * b)


def valid_pairs(a_grid: torch.Tensor, b_grid: torch.Tensor) -> torch.Tensor:
    """glm/arma.py:999-1048: meshgrid 'ij' (a-major, b-minor), keep
    |a|,|b|<1, |1-a^2|>=1e-6, gamma0>0, lambda>=0.  Arithmetic in the grid
    dtype (float32 in the default run).  Returns (K, 2) in grid dtype."""
    am, bm = torch.meshgrid(a_grid, b_grid, indexing="ij")
    a, b = am.reshape(-1), bm.reshape(-1)
    ok = (torch.abs(a) < 1.0) & (torch.abs(b) < 1.0)
    denom = 1 - a**2
    ok &= torch.abs(denom) >= 1e-6
    gamma0 = (1 + b**2 + 2 * a * b) / (denom + 1e-10)
    ok &= gamma0 > 0
    a, b = a[ok], b[ok]
    lam = arma_lambda(a, b)
    keep = lam >= 0
    return torch.stack([a[keep], b[keep]], dim=1)


# ---------------------------------------------------------------------------
# Runs, censoring, covariance
# ---------------------------------------------------------------------------


def build_censor_run_info(run_starts, n_total, good_list=None):
    """glm/arma.py:752-820."""
    starts = [int(s) for s in run_starts]
    if good_list is None:
        return starts, None
    retained = torch.as_tensor(list(good_list), dtype=torch.long).flatten()
    st = torch.tensor(starts, dtype=torch.long)
    run_id = torch.bucketize(retained, st, right=True) - 1
    tau = retained - st[run_id]
    rs = []
    for r in range(len(starts)):
        pos = (run_id == r).nonzero(as_tuple=False)
        if pos.numel() > 0:
            rs.append(int(pos[0].item()))
    return (rs or [0]), tau


def run_bounds(n: int, run_starts=None) -> list[tuple[int, int]]:
    """glm/whiten.py:27-38."""
    if run_starts is None:
        return [(0, n)]
    s = sorted({int(x) for x in run_starts})
    if not s or s[0] != 0:
        s = [0, *s]
    e = [*s[1:], n]
    return [(x, y) for x, y in zip(s, e) if y > x]


def covariance(a: float, b: float, n: int, run_starts=None, tau=None, dtype=torch.float64):
    """glm/arma.py:823-936 ``build_arma11_covariance`` (dense, exact, no AFNI band).

    R[i,j] = corr[|tau_i - tau_j|] with corr=[1, lam, lam*a, lam*a^2, ...],
    zeroed across run boundaries."""
    lam = arma_lambda(a, b)
    t = torch.arange(n) if tau is None else torch.as_tensor(tau, dtype=torch.long).flatten()
    dist = torch.abs(t[None, :] - t[:, None])
    max_lag = int(dist.max().item()) if n > 0 else 0
    corr = torch.zeros(max_lag + 1, dtype=dtype)
    corr[0] = 1.0
    if max_lag >= 1:
        corr[1] = lam
        if max_lag >= 2:
            # glm/arma.py:1072-1075: powers a^(1..max_lag-1)
            ex = torch.arange(1, max_lag, dtype=dtype)
            corr[2:] = lam * torch.as_tensor(a, dtype=dtype) ** ex
    R = corr[dist]
    if run_starts is not None:
        starts = torch.as_tensor(list(run_starts), dtype=torch.long)
        rid = torch.bucketize(torch.arange(n), starts, right=True) - 1
        R = R * (rid[None, :] == rid[:, None]).to(dtype)
    return R


def whitener(a, b, n, run_starts=None, tau=None, dtype=torch.float64):
    """glm/whiten.py:197-271 ``arma11_whiteners`` restated densely.

    Returns (L_inv (n,n) block-diagonal, logdet_R) with
    logdet_R = 2*sum(log(diag L + 1e-10)) (glm/whiten.py:224)."""
    Linv = torch.zeros(n, n, dtype=dtype)
    logdet = torch.zeros((), dtype=dtype)
    for s, e in run_bounds(n, run_starts):
        sub_tau = None if tau is None else torch.as_tensor(tau)[s:e]
        R = covariance(a, b, e - s, None, sub_tau, dtype)
        L = torch.linalg.cholesky(R)
        logdet = logdet + 2 * torch.log(torch.diagonal(L) + 1e-10).sum()
        Linv[s:e, s:e] = torch.linalg.solve_triangular(L, torch.eye(e - s, dtype=dtype), upper=False)
    return Linv, logdet


# ---------------------------------------------------------------------------
# Grid precompute + search
# ---------------------------------------------------------------------------


@dataclass
class Grid:
    params: torch.Tensor  # (K, 2) grid dtype values (float32 values, stored as given)
    Linv: torch.Tensor  # (K, T, T)
    Xw: torch.Tensor  # (K, T, p)
    Q: torch.Tensor  # (K, T, p)
    logdet_R: torch.Tensor  # (K,)
    logdet_XRX: torch.Tensor  # (K,)
    dtype: torch.dtype = torch.float64
    extra: dict = field(default_factory=dict)

    @property
    def param_list(self):
        return [(float(a), float(b)) for a, b in self.params.tolist()]


def dependent_columns(X) -> list[int]:
    """Columns of X that lie in the span of the columns before them: j such
    that rank(X[:, :j+1]) == rank(X[:, :j]), each rank from the singular
    values in float64 with the default relative tolerance (max(T, j+1) eps
    sigma_max).  Pinned by construction in the tests (designs with known
    duplicated and combined columns)."""
    Xd = torch.as_tensor(X).to(torch.float64)
    dep, prev = [], 0
    for j in range(Xd.shape[1]):
        r = int(torch.linalg.matrix_rank(Xd[:, :j + 1]))
        if r == prev:
            dep.append(j)
        prev = r
    return dep


def precompute_grid(X, a_grid, b_grid, run_starts=None, tau=None, dtype=torch.float64,
                    autocorr: Grid | None = None, rank_fallback: str = "qr") -> Grid:
    """glm/arma.py:1306-1531 ``precompute_reml_grid`` (Cholesky-QR branch).

    X_w = L^-1 X;  L_xtx = chol(X_w'X_w);  Q = X_w L_xtx^-T;
    logdet(X'R^-1X) = 2 sum log(|diag L_xtx| + 1e-10)  (glm/arma.py:1485-1506).
    ``autocorr``: a grid of the same pairs, timeline and dtype whose whiteners
    and log det R are reused (the reference's ``autocorr_cache``,
    glm/arma.py:1306-1368, shared by the groups of a grouped fit).

    ``rank_fallback`` for a design with dependent columns: "qr" runs the
    reference's branch as written (Cholesky, and QR of X_w for every pair when
    any Cholesky fails, glm/arma.py:1493-1500; ``extra["used_qr"]`` says which
    ran).  "rank" is that QR branch in exact arithmetic: for a column in the
    span of the columns before it Householder leaves R_jj = 0, so log det
    X'R^-1X carries 2 log(1e-10) for it, and Q's column j is undetermined
    (rounding picks some unit vector orthogonal to the columns before it); it
    is taken as zero here, so Q spans exactly col(X_w) (the rank-r tables).
    """
    X = torch.as_tensor(X).to(dtype)
    T, p = X.shape
    a_grid, b_grid = ensure_zero_in_grid(torch.as_tensor(a_grid), torch.as_tensor(b_grid))
    params = valid_pairs(a_grid.to(dtype) if dtype == torch.float64 else a_grid,
                         b_grid.to(dtype) if dtype == torch.float64 else b_grid)
    K = params.shape[0]
    if autocorr is not None:
        if (autocorr.dtype != dtype or autocorr.Linv.shape[1] != T
                or not torch.equal(autocorr.params.to(params.dtype), params)):
            raise ValueError("autocorr grid does not match these pairs, timeline length or dtype")
        Linv, ldR = autocorr.Linv, autocorr.logdet_R
    else:
        Linv = torch.empty(K, T, T, dtype=dtype)
        ldR = torch.empty(K, dtype=dtype)
        for k in range(K):
            a, b = float(params[k, 0]), float(params[k, 1])
            Linv[k], ldR[k] = whitener(a, b, T, run_starts, tau, dtype)
    Xw = Linv @ X
    dep = dependent_columns(X) if rank_fallback == "rank" else []
    if dep:
        ind = [j for j in range(p) if j not in dep]
        Xr = Xw[:, :, ind]
        Lr = torch.linalg.cholesky(Xr.transpose(1, 2) @ Xr)
        Q = torch.zeros_like(Xw)
        Q[:, :, ind] = torch.linalg.solve_triangular(Lr, Xr.transpose(1, 2), upper=False).transpose(1, 2)
        ldX = (2 * torch.log(torch.abs(torch.diagonal(Lr, dim1=1, dim2=2)) + 1e-10).sum(1)
               + 2 * len(dep) * math.log(1e-10))
        return Grid(params, Linv, Xw, Q, ldR, ldX, dtype, extra={"dependent": dep, "used_qr": False})
    used_qr, Rq = False, None
    try:
        Lx = torch.linalg.cholesky(Xw.transpose(1, 2) @ Xw)
        Q = torch.linalg.solve_triangular(Lx, Xw.transpose(1, 2), upper=False).transpose(1, 2)
        diag = torch.diagonal(Lx, dim1=1, dim2=2)
    except torch.linalg.LinAlgError:  # glm/arma.py:1497-1500
        used_qr = True
        Q, Rq = torch.linalg.qr(Xw)
        diag = torch.diagonal(Rq, dim1=1, dim2=2)
    ldX = 2 * torch.log(torch.abs(diag) + 1e-10).sum(1)
    return Grid(params, Linv, Xw, Q, ldR, ldX, dtype, extra={"used_qr": used_qr, "qr_R": Rq})


def design_pinv(X: torch.Tensor) -> torch.Tensor:
    """glm/core.py:248-272: rank-truncated SVD pinv in float64, cast back."""
    Xd = X.to(torch.float64)
    U, S, Vh = torch.linalg.svd(Xd, full_matrices=False)
    tol = max(Xd.shape) * torch.finfo(X.dtype).eps * S[0]
    r = int((S > tol).sum())
    return ((Vh[:r].T / S[:r]) @ U[:, :r].T).to(X.dtype)


def residualize(X, Y):
    """glm/arma.py:1554-1571: r = Y - (Y pinv^T) X^T in Y's dtype; float64 is
    returned untouched."""
    if Y.dtype == torch.float64 or X.shape[1] == 0:
        return Y
    Xd = X.to(Y.dtype)
    pinv = design_pinv(Xd)
    return Y - (Y @ pinv.T) @ Xd.T


def score_surface(X, Y, grid: Grid) -> torch.Tensor:
    """glm/arma.py:1886-2018 inner loop: returns the (V, K) likelihood surface
    in Y's dtype (float64 sums, cast back, glm/arma.py:1993-2004)."""
    T, p = X.shape
    Yr = residualize(X, Y)
    V, K = Y.shape[0], grid.params.shape[0]
    surf = torch.empty(V, K, dtype=Y.dtype)
    acc = torch.float64
    for k in range(K):
        Yw = Yr @ grid.Linv[k].to(Y.dtype).T
        QtY = grid.Q[k].to(Y.dtype).T @ Yw.T
        rss = Yw.pow(2).sum(1, dtype=acc) - QtY.pow(2).sum(0, dtype=acc)
        lik = grid.logdet_R[k].to(acc) + grid.logdet_XRX[k].to(acc) + (T - p) * torch.log(rss + 1e-10)
        surf[:, k] = lik.to(Y.dtype)
    return surf


def select_exhaustive(surface: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Strict '<' update in pair order (glm/arma.py:2006-2010): the first pair
    attaining the minimum wins; NaN never improves on +inf."""
    V, K = surface.shape
    best = torch.full((V,), float("inf"), dtype=surface.dtype)
    idx = torch.zeros(V, dtype=torch.long)
    has = torch.zeros(V, dtype=torch.bool)
    for k in range(K):
        imp = surface[:, k] < best
        best = torch.where(imp, surface[:, k], best)
        idx = torch.where(imp, torch.full_like(idx, k), idx)
        has |= imp
    return idx, best


def _lattice(params):
    a_sorted = sorted({float(a) for a, _ in params.tolist()})
    b_sorted = sorted({float(b) for _, b in params.tolist()})
    ai = [a_sorted.index(float(a)) for a, _ in params.tolist()]
    bi = [b_sorted.index(float(b)) for _, b in params.tolist()]
    return a_sorted, b_sorted, ai, bi


def hier_ltop(n_a: int, n_b: int, afni_faithful: bool = False) -> int:
    """glm/arma.py:2290-2296."""
    if afni_faithful:
        pna = (n_a - 1).bit_length() - 1
        pnb = (n_b - 1).bit_length() - 1
        return max(0, min(pna, pnb) - 2)
    return max(0, min(n_a, n_b).bit_length() - 2)


def select_hierarchical(surface: torch.Tensor, params: torch.Tensor, afni_faithful=False, trace=False):
    """glm/arma.py:2224-2322 ``select_arma_params_hierarchical_from_surface``,
    restated per voxel with plain loops.  Returns (pair index, score), and with
    ``trace`` also, per voxel, the pair index held after each level (ltop
    first).  A level whose candidates include a NaN takes NaN as its minimum
    (``Tensor.min`` propagates NaN, glm/arma.py:2299, 2311), so it never
    improves the running best."""
    a_sorted, b_sorted, ai, bi = _lattice(params)
    n_a, n_b = len(a_sorted), len(b_sorted)
    lut = -np.ones((n_a, n_b), dtype=np.int64)
    for k, (x, y) in enumerate(zip(ai, bi)):
        lut[x, y] = k
    za = next(i for i, v in enumerate(a_sorted) if abs(v) < 1e-6)
    zb = next(i for i, v in enumerate(b_sorted) if abs(v) < 1e-6)
    ltop = hier_ltop(n_a, n_b, afni_faithful)
    S = surface.numpy()
    V = S.shape[0]
    inf = np.array(np.inf, dtype=S.dtype)

    def val(v, x, y):
        k = lut[x, y]
        return S[v, k] if k >= 0 else inf

    out_k = np.zeros(V, dtype=np.int64)
    out_s = np.zeros(V, dtype=S.dtype)
    levels = []
    for v in range(V):
        ba, bb = za, zb
        best = val(v, ba, bb)
        path = []
        for lev in range(ltop, -1, -1):
            step = 1 << lev
            if lev == ltop:
                ac = list(range(0, n_a, step))
                bc = list(range(0, n_b, step))
            else:
                ac = [min(max(ba + o, 0), n_a - 1) for o in (-step, 0, step)]
                bc = [min(max(bb + o, 0), n_b - 1) for o in (-step, 0, step)]
            # candidates flattened b-outer / a-inner, first minimum wins; a NaN
            # candidate makes the level's minimum NaN
            mn, sel = None, None
            for y in bc:
                for x in ac:
                    s = val(v, x, y)
                    if np.isnan(s):
                        mn, sel = s, (x, y)
                        break
                    if mn is None or s < mn:
                        mn, sel = s, (x, y)
                if mn is not None and np.isnan(mn):
                    break
            if mn < best:
                best = mn
                ba, bb = sel
            path.append(int(lut[ba, bb]))
        out_k[v] = lut[ba, bb]
        out_s[v] = best
        levels.append(path)
    if trace:
        return torch.from_numpy(out_k), torch.from_numpy(out_s), levels
    return torch.from_numpy(out_k), torch.from_numpy(out_s)


def search_hierarchical_subset(surface: torch.Tensor, params: torch.Tensor, afni_faithful=False):
    """glm/arma.py:2021-2221 (CPU default) restated on a precomputed surface:
    start at (0,0), then per level evaluate the lattice points inside each
    voxel's own +/-step window in b-outer / a-inner order with a sequential
    strict '<' update (so the best point moves *during* a level's sweep, unlike
    the surface selector which takes one argmin per level)."""
    a_sorted, b_sorted, ai, bi = _lattice(params)
    n_a, n_b = len(a_sorted), len(b_sorted)
    lut = -np.ones((n_a, n_b), dtype=np.int64)
    for k, (x, y) in enumerate(zip(ai, bi)):
        lut[x, y] = k
    za = next(i for i, v in enumerate(a_sorted) if abs(v) < 1e-6)
    zb = next(i for i, v in enumerate(b_sorted) if abs(v) < 1e-6)
    ltop = hier_ltop(n_a, n_b, afni_faithful)
    S = surface.numpy()
    V = S.shape[0]
    out_k = np.zeros(V, dtype=np.int64)
    out_s = np.zeros(V, dtype=S.dtype)
    for v in range(V):
        ba, bb = za, zb
        best = np.array(np.inf, dtype=S.dtype)
        k0 = lut[za, zb]
        if S[v, k0] < best:
            best = S[v, k0]
        for lev in range(ltop, -1, -1):
            step = 1 << lev
            if lev == ltop:
                alo, ahi, blo, bhi = 0, n_a - 1, 0, n_b - 1
            else:
                alo, ahi = max(ba - step, 0), min(ba + step, n_a - 1)
                blo, bhi = max(bb - step, 0), min(bb + step, n_b - 1)
            for y in range(0, n_b, step):
                if not (blo <= y <= bhi):
                    continue
                for x in range(0, n_a, step):
                    k = lut[x, y]
                    if k < 0 or not (alo <= x <= ahi):
                        continue
                    if S[v, k] < best:
                        best = S[v, k]
                        ba, bb = x, y
        out_k[v] = lut[ba, bb]
        out_s[v] = best
    return torch.from_numpy(out_k), torch.from_numpy(out_s)


def search_subset_gemm(X, Y, grid: Grid, afni_faithful=False):
    """glm/arma.py:2021-2221 ``search_voxels_precomputed_grid_hierarchical``
    (the reference's CPU default) with its own work: per level, the lattice
    points that are multiples of the step, each scored (dense whitening GEMM,
    float64 sums) only for the voxels whose window (fixed at the level's
    start) holds it, with the sequential strict '<' update in b-outer / a-inner
    order, after the (0, 0) evaluation of every voxel.  Returns (pair index,
    score).  Same rule as ``search_hierarchical_subset`` on a full surface."""
    T, p = X.shape
    Yr = residualize(X, Y)
    a_sorted, b_sorted, ai_of, bi_of = _lattice(grid.params)
    n_a, n_b = len(a_sorted), len(b_sorted)
    lut = -np.ones((n_a, n_b), dtype=np.int64)
    for k, (x, y) in enumerate(zip(ai_of, bi_of)):
        lut[x, y] = k
    za = next(i for i, v in enumerate(a_sorted) if abs(v) < 1e-6)
    zb = next(i for i, v in enumerate(b_sorted) if abs(v) < 1e-6)
    V = Y.shape[0]
    acc = torch.float64
    ba = torch.full((V,), za, dtype=torch.long)
    bb = torch.full((V,), zb, dtype=torch.long)
    best = torch.full((V,), float("inf"), dtype=Y.dtype)

    def evaluate(x, y, idx):
        k = int(lut[x, y])
        if k < 0 or idx.numel() == 0:
            return
        Yw = Yr.index_select(0, idx) @ grid.Linv[k].to(Y.dtype).T
        QtY = Yw @ grid.Q[k].to(Y.dtype)
        rss = Yw.pow(2).sum(1, dtype=acc) - QtY.pow(2).sum(1, dtype=acc)
        lik = (grid.logdet_R[k].to(acc) + grid.logdet_XRX[k].to(acc) + (T - p) * torch.log(rss + 1e-10)).to(Y.dtype)
        up = lik < best.index_select(0, idx)
        sel = idx[up]
        best[sel] = lik[up]
        ba[sel] = x
        bb[sel] = y

    evaluate(za, zb, torch.arange(V))
    for lev in range(hier_ltop(n_a, n_b, afni_faithful), -1, -1):
        step = 1 << lev
        if lev == hier_ltop(n_a, n_b, afni_faithful):
            alo = torch.zeros(V, dtype=torch.long)
            ahi = torch.full((V,), n_a - 1, dtype=torch.long)
            blo = torch.zeros(V, dtype=torch.long)
            bhi = torch.full((V,), n_b - 1, dtype=torch.long)
        else:
            alo, ahi = (ba - step).clamp(0, n_a - 1), (ba + step).clamp(0, n_a - 1)
            blo, bhi = (bb - step).clamp(0, n_b - 1), (bb + step).clamp(0, n_b - 1)
        for y in range(0, n_b, step):
            in_b = (blo <= y) & (y <= bhi)
            for x in range(0, n_a, step):
                evaluate(x, y, (in_b & (alo <= x) & (x <= ahi)).nonzero().flatten())
    return torch.from_numpy(lut[ba.numpy(), bb.numpy()]), best


# ---------------------------------------------------------------------------
# GLS refit and statistics
# ---------------------------------------------------------------------------


def cod_r2(y_true: torch.Tensor, y_pred: torch.Tensor) -> torch.Tensor:
    """glm/xval.py:563-616: 1 - SS_res/max(SS_tot,1e-10); 0 if SS_tot<=1e-10; <=1."""
    ss_res = ((y_true - y_pred) ** 2).sum(1)
    ss_tot = ((y_true - y_true.mean(1, keepdim=True)) ** 2).sum(1)
    deg = ss_tot <= 1e-10
    r2 = 1.0 - ss_res / torch.clamp(ss_tot, min=1e-10)
    return torch.where(deg, torch.zeros_like(r2), torch.clamp(r2, max=1.0))


def ljung_box_max_lag(n_time: int, n_regressors: int = 0, min_run: int | None = None) -> int:
    """glm/arma.py:5972-6000."""
    if min_run is None:
        min_run = n_time
    if min_run < 2:
        return 0
    hh = n_regressors + 2 + min(min_run // 8, int(round(3.0 * math.log(min_run))))
    return min(hh, min_run // 2)


def resolve_ljung_box_lag(n_time: int, max_lag) -> int:
    """glm/arma.py:6003-6014."""
    hh = int(max_lag) if max_lag else 0
    if hh < 2 or hh > n_time // 2:
        hh = 2 + min(n_time // 8, int(round(3.0 * math.log(n_time))))
        hh = min(hh, n_time // 2)
    return hh


def ljung_box_tau(n_time: int, run_starts=None, tau=None):
    """glm/arma.py:6017-6049 (runs offset by 66666)."""
    starts = sorted({int(s) for s in (run_starts or []) if 0 <= int(s) < n_time}) or [0]
    if tau is None and len(starts) == 1 and starts[0] == 0:
        return None
    st = torch.tensor(starts, dtype=torch.long)
    idx = torch.arange(n_time, dtype=torch.long)
    rid = torch.bucketize(idx, st, right=True) - 1
    within = torch.as_tensor(tau, dtype=torch.long).flatten() if tau is not None else idx - st[rid]
    return within + 66666 * rid


def ljung_box(resid: torch.Tensor, max_lag=None, tau=None) -> torch.Tensor:
    """glm/arma.py:6059-6141 restated with explicit lag-pair loops (float64)."""
    r = resid.to(torch.float64)
    V, n = r.shape
    out = torch.zeros(V, dtype=torch.float64)
    if n < 10:
        return out
    hh = resolve_ljung_box_lag(n, max_lag)
    if hh < 1:
        return out
    sum0 = (r * r).sum(1)
    ok = sum0 >= 1e-10
    sum0 = torch.where(ok, sum0, torch.ones_like(sum0))
    sumk = torch.zeros(V, hh + 1, dtype=torch.float64)
    nj = np.zeros(hh + 1, dtype=np.int64)
    t = np.arange(n) if tau is None else np.asarray(torch.as_tensor(tau).tolist(), dtype=np.int64)
    for kk in range(1, hh + 1):
        dj = t[kk:] - t[: n - kk]
        for j in np.nonzero((dj > 0) & (dj <= hh))[0]:
            sumk[:, dj[j]] += r[:, j] * r[:, j + kk]
            nj[dj[j]] += 1
    good = nj > 1
    good[0] = False
    if not good.any():
        return out
    gi = np.nonzero(good)[0]
    ck = sumk[:, gi] / sum0[:, None]
    g = (ck**2 / torch.from_numpy(nj[gi].astype(np.float64))).sum(1) * (n * (n + 2.0))
    return torch.where(ok, g.clamp(max=1.0e10), torch.zeros_like(g))


@dataclass
class OracleFit:
    params: torch.Tensor  # (V, 2) selected (a, b)
    pair_index: torch.Tensor  # (V,)
    reml_likelihood: torch.Tensor  # (V,)
    surface: torch.Tensor  # (V, K)
    betas: torch.Tensor
    tstats: torch.Tensor
    fstats: torch.Tensor
    sigma2: torch.Tensor
    r2: torch.Tensor
    arma_lambda: torch.Tensor
    ljung_box: torch.Tensor | None
    residuals_whitened: torch.Tensor | None
    grid: Grid
    scale: torch.Tensor | None


def gls_refit(X, Y, grid: Grid, pair_index, task_indices=None, dtype=torch.float64,
              want_ljung_box=False, lj_max_lag=None, lj_tau=None, rank_fallback: str = "inv"):
    """glm/arma.py:4784-5601, X'X branch (use_qr=False), restated per voxel group.

    A = X_w'X_w; A^-1 via Cholesky; beta = (y_w X_w) A^-T; sigma2 =
    sum(resid_w^2)/dof (float64 sum); t = beta/(sqrt(sigma2 diag A^-1)+1e-10);
    F = b_task' inv(A^-1_task + 1e-8 I) b_task / (sigma2 p_task + 1e-10), or
    b' A b / (sigma2 p + 1e-10) with no task subset; R^2 = CoD(y, X beta).
    ``rank_fallback``: "inv" is the group path's fallback when the Cholesky
    fails (glm/arma.py:4864-4873); "pinv" takes A^+ (torch.linalg.pinv,
    hermitian) whenever X has dependent columns, the -dsort fallback
    (glm/arma.py:3410-3416): beta is then the minimum-norm solution."""
    rank_def = rank_fallback == "pinv" and bool(dependent_columns(X))
    X = torch.as_tensor(X).to(dtype)
    Y = Y.to(dtype)
    T, p = X.shape
    V = Y.shape[0]
    dof = max(1, T - p)
    betas = torch.zeros(V, p, dtype=dtype)
    tst = torch.zeros(V, p, dtype=dtype)
    fst = torch.zeros(V, dtype=dtype)
    s2 = torch.zeros(V, dtype=dtype)
    r2 = torch.zeros(V, dtype=dtype)
    lb = torch.zeros(V, dtype=torch.float64) if want_ljung_box else None
    rw = torch.zeros(V, T, dtype=dtype)
    for k in torch.unique(pair_index).tolist():
        m = (pair_index == k).nonzero().flatten()
        Linv = grid.Linv[k].to(dtype)
        Xw = Linv @ X
        A = Xw.T @ Xw
        if rank_def:
            Ainv = torch.linalg.pinv(A, hermitian=True)
        else:
            try:
                Ainv = torch.cholesky_inverse(torch.linalg.cholesky(A))
            except torch.linalg.LinAlgError:
                Ainv = torch.linalg.inv(A)
        yw = Y[m] @ Linv.T
        b = (yw @ Xw) @ Ainv.T
        res = yw - b @ Xw.T
        sig = (res.pow(2).sum(1, dtype=torch.float64) / dof).to(dtype)
        se = torch.sqrt(sig[:, None] * torch.diagonal(Ainv)[None, :])
        t = b / (se + 1e-10)
        if task_indices is not None:
            ti = torch.as_tensor(task_indices)
            At = Ainv[ti][:, ti]
            Ati = torch.linalg.inv(At + 1e-8 * torch.eye(len(ti), dtype=dtype))
            bt = b[:, ti]
            F = torch.einsum("bi,ij,bj->b", bt, Ati, bt) / (sig * len(ti) + 1e-10)
        else:
            F = torch.einsum("bi,ij,bj->b", b, A, b) / (sig * p + 1e-10)
        betas[m], tst[m], fst[m], s2[m] = b, t, F, sig
        r2[m] = cod_r2(Y[m], b @ X.T)
        rw[m] = res
        if want_ljung_box:
            lb[m] = ljung_box(res, lj_max_lag, lj_tau)
    return betas, tst, fst, s2, r2, lb, rw


def glt_contrasts(X, grid: Grid, pair_index, betas, sigma2, C, dof, r2=None, dtype=torch.float64):
    """glm/arma.py:5421-5509: single-row GLTs per voxel.  contrast beta = C beta;
    Var = sigma2 c' (X_w'X_w)^-1 c with X_w = L^-1 X of the voxel's pair;
    t = cb / (sqrt(max(Var, 0)) + 1e-10); partial R^2 = t^2 / (t^2 + dof);
    semi-partial = partial * max(1 - R^2, 0)."""
    X = torch.as_tensor(X).to(dtype)
    C = torch.as_tensor(C).to(dtype)
    cb = betas.to(dtype) @ C.T
    var = torch.zeros_like(cb)
    for k in torch.unique(pair_index).tolist():
        m = (pair_index == k).nonzero().flatten()
        Xw = grid.Linv[k].to(dtype) @ X
        Ainv = torch.linalg.inv(Xw.T @ Xw)
        var[m] = sigma2.to(dtype)[m, None] * torch.einsum("cr,rs,cs->c", C, Ainv, C)[None, :]
    t = cb / (var.clamp(min=0.0).sqrt() + 1e-10)
    part = t**2 / (t**2 + dof)
    semi = part * (1.0 - r2.to(dtype)).clamp(min=0.0)[:, None] if r2 is not None else None
    return cb, t, part, semi


def column_set_stats(X, grid: Grid, pair_index, betas, sigma2, column_sets, dtype=torch.float64):
    """glm/arma.py:3543-3585 (-nuisance_summary column sets) per voxel, in the
    units of betas / sigma2: for a set S of q columns, F = b_S' (A_SS + 1e-8 I)^-1
    b_S / (q sigma2 + 1e-30) with A = (X_w'X_w)^-1 of the voxel's pair (X_w =
    L^-1 X), SD = sqrt(max(b_S' C_S b_S, 0)) with C_S the covariance over time
    of the centred unwhitened columns (divisor T), coef = b for a one-column set.
    NaN for a set naming a column past the design (and coef for wider sets)."""
    X = torch.as_tensor(X).to(dtype)
    T, p = X.shape
    V, G = betas.shape[0], len(column_sets)
    F = torch.full((V, G), float("nan"), dtype=dtype)
    sd = torch.full_like(F, float("nan"))
    coef = torch.full_like(F, float("nan"))
    Xc = X - X.mean(0, keepdim=True)
    b_all = betas.to(dtype)
    s2 = sigma2.to(dtype)
    inv_by_pair = {}
    for k in torch.unique(pair_index).tolist():
        Xw = grid.Linv[k].to(dtype) @ X
        inv_by_pair[k] = torch.linalg.inv(Xw.T @ Xw)
    for g, cols in enumerate(column_sets):
        if not cols or any(c >= p for c in cols):
            continue
        ix = torch.as_tensor(cols, dtype=torch.long)
        q = len(cols)
        Cs = Xc[:, ix].T @ Xc[:, ix] / T
        b = b_all[:, ix]
        for k, Ainv in inv_by_pair.items():
            m = (pair_index == k).nonzero().flatten()
            Minv = torch.linalg.inv(Ainv[ix][:, ix] + 1e-8 * torch.eye(q, dtype=dtype))
            F[m, g] = torch.einsum("bi,ij,bj->b", b[m], Minv, b[m]) / (q * s2[m] + 1e-30)
        sd[:, g] = torch.einsum("bi,ij,bj->b", b, Cs, b).clamp(min=0).sqrt()
        if q == 1:
            coef[:, g] = b[:, 0]
    return F, sd, coef


def dsort_gls(X, D, Y, grid: Grid, pair_index, task_indices=None, dtype=torch.float64):
    """glm/arma.py:3248-3460 (-dsort pass) restated per voxel: extended design
    [X | D_v'] (D (V, q, T)), whitened by the voxel's pair (dense L^-1);
    A = X_ext,w' X_ext,w + 1e-8 I, beta = A^-1 X_ext,w' y_w, A^-1 by Cholesky;
    sigma2 = |resid_w|^2 / (T - p - q); t = beta / (sqrt(sigma2 diag A^-1) + 1e-10);
    R^2 = CoD(y, X beta_base + D' beta_dsort); F over the task columns (all
    base columns without a task subset) = b' (A^-1_SS + 1e-8 I)^-1 b /
    (sigma2 p_f + 1e-30).  Returns (betas (V, p+q), tstats, sigma2, r2, F,
    resid_w, pred, A^-1 (V, p+q, p+q))."""
    X = torch.as_tensor(X).to(dtype)
    D = torch.as_tensor(D).to(dtype)
    Y = torch.as_tensor(Y).to(dtype)
    T, p = X.shape
    V, q = D.shape[0], D.shape[1]
    pe = p + q
    dof = max(1, T - pe)
    fci = list(task_indices) if task_indices is not None else list(range(p))
    out_b = torch.zeros(V, pe, dtype=dtype)
    out_t = torch.zeros_like(out_b)
    out_s2 = torch.zeros(V, dtype=dtype)
    out_r2 = torch.zeros(V, dtype=dtype)
    out_F = torch.zeros(V, dtype=dtype)
    out_rw = torch.zeros(V, T, dtype=dtype)
    out_pred = torch.zeros(V, T, dtype=dtype)
    out_Ai = torch.zeros(V, pe, pe, dtype=dtype)
    for v in range(V):
        L = grid.Linv[int(pair_index[v])].to(dtype)
        Xe = torch.cat([X, D[v].T], dim=1)
        Xw, yw = L @ Xe, L @ Y[v]
        A = Xw.T @ Xw + 1e-8 * torch.eye(pe, dtype=dtype)
        Ai = torch.cholesky_inverse(torch.linalg.cholesky(A))
        b = Ai @ (Xw.T @ yw)
        rw = yw - Xw @ b
        sig = float(rw.pow(2).sum()) / dof
        t = b / (torch.sqrt(sig * torch.diagonal(Ai)) + 1e-10)
        pred = Xe @ b
        ix = torch.as_tensor(fci)
        At = Ai[ix][:, ix]
        bt = b[ix]
        Ati = torch.linalg.inv(At + 1e-8 * torch.eye(len(fci), dtype=dtype))
        out_b[v], out_t[v], out_s2[v] = b, t, sig
        out_F[v] = (bt @ Ati @ bt) / (sig * len(fci) + 1e-30)
        out_r2[v] = cod_r2(Y[v][None], pred[None])[0]
        out_rw[v], out_pred[v], out_Ai[v] = rw, pred, Ai
    return out_b, out_t, out_s2, out_r2, out_F, out_rw, out_pred, out_Ai


def ols_fit(X, Y, task_indices=None, C=None, dtype=torch.float64):
    """glm/core.py:293-385, 880-1110 (one run, complete design): OLS betas by
    the normal equations; sigma2 = max(RSS / (T - p), 0); stderr and t on the
    task columns; F = t^2 for one task column, else b_S' ((X'X)^-1_SS)^-1 b_S /
    (|S| sigma2 + 1e-10); R^2 = CoD; r2_run from the task-column prediction;
    GLT c'b and its t.  Returns a dict of (V, ...) tensors."""
    X = torch.as_tensor(X).to(dtype)
    Y = torch.as_tensor(Y).to(dtype)
    T, p = X.shape
    task = list(task_indices) if task_indices is not None else list(range(p))
    dof = max(1, T - p)
    Ai = torch.linalg.inv(X.T @ X)
    b = Y @ X @ Ai.T
    resid = Y - b @ X.T
    sig = (resid.pow(2).sum(1) / dof).clamp(min=0.0)
    ti = torch.as_tensor(task)
    bt = b[:, ti]
    se = torch.sqrt(sig[:, None] * torch.diagonal(Ai)[ti][None, :])
    t = bt / (se + 1e-10)
    if len(task) == 1:
        F = t[:, 0] ** 2
    else:
        F = torch.einsum("bi,ij,bj->b", bt, torch.linalg.inv(Ai[ti][:, ti]), bt) / (len(task) * sig + 1e-10)
    ss_tot = (Y - Y.mean(1, keepdim=True)).pow(2).sum(1)
    r2 = cod_r2(Y, b @ X.T)
    pred_t = bt @ X[:, ti].T
    r2_run = (1 - (Y - pred_t).pow(2).sum(1) / (ss_tot + 1e-10)).clamp(0, 1)
    out = {"betas": bt, "betas_full": b, "tstats": t, "stderr": se, "sigma2": sig, "fstats": F, "r2": r2,
           "r2_run": r2_run, "meanvol": Y.mean(1), "resid": resid}
    if C is not None:
        C = torch.as_tensor(C).to(dtype)
        cb = b @ C.T
        cv = sig[:, None] * torch.einsum("cr,rs,cs->c", C, Ai, C)[None, :]
        out["contrast_betas"], out["contrast_tstats"] = cb, cb / (cv.clamp(min=0.0).sqrt() + 1e-10)
    return out


def fit_arma11(data, design, a_grid=None, b_grid=None, run_starts=None, tau=None,
               dtype=torch.float64, selector="hierarchical", task_indices=None,
               want_ljung_box=False, grid: Grid | None = None) -> OracleFit:
    """glm/arma.py:3650-5878 ``fit_glm_arma11`` (per-voxel estimation, no
    precomputed params, dense whitening) restated.

    float32: Y is divided by its unbiased std clamped at 1e-8 (glm/arma.py:3947-3954),
    the search residualises in float32, betas / sigma2 are unscaled at the end
    (glm/arma.py:5827-5862).  float64: no scaling (``use_double``)."""
    Y = torch.as_tensor(data).to(dtype)
    X = torch.as_tensor(design).to(dtype)
    T, p = X.shape
    if Y.shape[1] != T:
        raise ValueError(f"Timepoints mismatch: Data has {Y.shape[1]}, Design has {T}")
    if a_grid is None or b_grid is None:
        da, db = default_grids()
        a_grid = da if a_grid is None else torch.as_tensor(a_grid)
        b_grid = db if b_grid is None else torch.as_tensor(b_grid)
    if grid is None:
        grid = precompute_grid(X, a_grid, b_grid, run_starts, tau, dtype)
    scale = None
    Yn = Y
    if dtype != torch.float64:
        scale = Y.std(dim=1).clamp(min=1e-8)
        Yn = Y / scale[:, None]
    surf = score_surface(X, Yn, grid)
    if selector == "exhaustive":
        idx, best = select_exhaustive(surf)
    elif selector == "hierarchical":
        idx, best = select_hierarchical(surf, grid.params)
    elif selector == "subset":
        idx, best = search_hierarchical_subset(surf, grid.params)
    else:
        raise ValueError(selector)
    params = grid.params[idx].to(dtype)
    a, b = params[:, 0], params[:, 1]
    lam = (b + a) * (1 + a * b) / (1 + 2 * a * b + b**2 + 1e-10)  # glm/arma.py:4769-4777
    lj_lag = lj_tau = None
    if want_ljung_box:
        starts = sorted({int(s) for s in (run_starts or [0])})
        mins = T if len(starts) <= 1 else min(
            e - s for s, e in zip(starts, starts[1:] + [T]))
        lj_lag = ljung_box_max_lag(T, p, mins)
        lj_tau = ljung_box_tau(T, run_starts, tau)
    betas, tst, fst, s2, r2, lb, rw = gls_refit(X, Yn, grid, idx, task_indices, dtype,
                                                 want_ljung_box, lj_lag, lj_tau)
    if scale is not None:
        betas = betas * scale[:, None]
        s2 = s2 * scale**2
        rw = rw * scale[:, None]
    if task_indices is not None:
        betas, tst = betas[:, task_indices], tst[:, task_indices]
    return OracleFit(params, idx, best, surf, betas, tst, fst, s2, r2.to(torch.float32),
                     lam, lb, rw, grid, scale)


def partial_r2(t_full, r2, dof, task_indices=None, partial_mode="full", semi_mode="full"):
    """glm/arma.py:5054-5175 restated: partial R^2 = t^2 / (t^2 + dof) for every
    regressor; with task_indices the task / nuisance columns are split and mode
    "task" divides the task values by clamp(1 - sum(nuisance), min=0.01);
    semi-partial = partial * clamp(1 - R^2, min=0) with the same split / mode.
    Returns (r2_partial, r2_partial_nuisance, r2_semipartial, r2_semipartial_nuisance)."""
    t_full = torch.as_tensor(t_full)
    p = t_full.shape[1]
    t2 = t_full**2
    full = t2 / (t2 + dof)
    if task_indices is not None:
        nuis = sorted(set(range(p)) - set(task_indices))
        task = full[:, list(task_indices)]
        nu = full[:, nuis] if nuis else None
    else:
        task, nu = full, None
    part = task / torch.clamp(1.0 - nu.sum(1, keepdim=True), min=0.01) if (partial_mode == "task" and nu is not None) \
        else task
    remain = torch.clamp(1.0 - torch.as_tensor(r2).to(t_full.dtype)[:, None], min=0.0)
    st = task * remain
    sn = nu * remain if nu is not None else None
    semi = st / torch.clamp(1.0 - sn.sum(1, keepdim=True), min=0.01) if (semi_mode == "task" and sn is not None) \
        else st
    return part, nu, semi, sn



def reml_likelihood_dense(X, Y, R):
    """glm/arma.py:1106-1186 ``compute_reml_likelihood`` restated in FP64:
    L = chol(R); 2 sum log(diag L + 1e-10) + log det(X_w'X_w + 1e-6 I)
    + (T - p) log(||Y_w - X_w beta||^2 + 1e-10), beta = (X_w'X_w + 1e-6 I)^-1 X_w'Y_w."""
    X = torch.as_tensor(X, dtype=torch.float64)
    Y = torch.as_tensor(Y, dtype=torch.float64)
    Y = Y[:, None] if Y.ndim == 1 else Y
    R = torch.as_tensor(R, dtype=torch.float64)
    T, p = X.shape
    L = torch.linalg.cholesky(R)
    term1 = 2 * torch.sum(torch.log(torch.diag(L) + 1e-10))
    Xw = torch.linalg.solve_triangular(L, X, upper=False)
    Yw = torch.linalg.solve_triangular(L, Y, upper=False)
    A = Xw.T @ Xw + 1e-6 * torch.eye(p, dtype=torch.float64)
    sign, ld = torch.linalg.slogdet(A)
    term2 = ld if sign > 0 else torch.tensor(1e10, dtype=torch.float64)
    beta = torch.linalg.solve(A, Xw.T @ Yw)
    rss = torch.sum((Yw - Xw @ beta) ** 2)
    return float(term1 + term2 + (T - p) * torch.log(rss + 1e-10))


def ridge_grid_search(X, Y, params, run_starts=None, tau=None):
    """glm/arma.py:2577-2833 ``reml_grid_search_batched`` restated per voxel:
    dense ARMA(1,1) covariance per pair (run blocks, tau gaps), the ridge score
    of reml_likelihood_dense on the raw data, sequential strict '<' in pair
    order.  Returns (pair index, score), index -1 / score inf when nothing
    improves on +inf."""
    Y = torch.as_tensor(Y, dtype=torch.float64)
    V, T = Y.shape
    best_k = torch.full((V,), -1, dtype=torch.long)
    best_s = torch.full((V,), float("inf"), dtype=torch.float64)
    for k, (a, b) in enumerate(torch.as_tensor(params, dtype=torch.float64).tolist()):
        R = covariance(a, b, T, run_starts, tau, torch.float64)
        for v in range(V):
            s = reml_likelihood_dense(X, Y[v], R)
            if s < best_s[v]:
                best_s[v] = s
                best_k[v] = k
    return best_k, best_s


def select_in_order(surface: torch.Tensor, order):
    """Exhaustive first-minimum selection visiting the pairs in ``order``
    (glm/arma.py:1754-1776 evaluation order of batch_reml_grid_search)."""
    S = surface[:, list(order)]
    j = torch.argmin(S, dim=1)  # first minimum in the visiting order
    return torch.as_tensor(list(order))[j], S[torch.arange(S.shape[0]), j]


def global_fit(X, Y, a, b, run_starts=None, tau=None, task_indices=None, dtype=torch.float64):
    """glm/arma.py:5603-5780 (``estimate_per_voxel=False``) after the search:
    one whitener L^-1 of (a, b) for every voxel, A = X_w'X_w + 1e-6 I, beta =
    A^-1 X_w'y_w, sigma2 = |y_w - X_w beta|^2 / (T - p), t = beta / (sqrt(sigma2
    diag A^-1) + 1e-10) over all p columns, F = beta_S' (sigma2 A^-1_SS +
    1e-8 I)^-1 beta_S / |S| (S = task columns, all when None), R^2 = CoD of y
    against X beta.  Y (V, T) in the units the fit sees.  Returns a dict."""
    X = torch.as_tensor(X).to(dtype)
    Y = torch.as_tensor(Y).to(dtype)
    T, p = X.shape
    Li, _ = whitener(a, b, T, run_starts, tau, dtype)
    Xw = Li @ X
    Yw = Y @ Li.T
    Ai = torch.linalg.inv(Xw.T @ Xw + 1e-6 * torch.eye(p, dtype=dtype))
    beta = Yw @ Xw @ Ai
    rw = Yw - beta @ Xw.T
    dof = max(1, T - p)
    sig = rw.pow(2).sum(1) / dof
    t = beta / (torch.sqrt(sig[:, None] * torch.diagonal(Ai)[None, :]) + 1e-10)
    S = list(task_indices) if task_indices else list(range(p))
    ti = torch.as_tensor(S)
    F = torch.empty(Y.shape[0], dtype=dtype)
    for v in range(Y.shape[0]):
        M = sig[v] * Ai[ti][:, ti] + 1e-8 * torch.eye(len(S), dtype=dtype)
        bs = beta[v, ti]
        F[v] = bs @ torch.linalg.inv(M) @ bs / len(S)
    pred = beta @ X.T
    return {"betas": beta, "tstats": t, "sigma2": sig, "fstats": F, "r2": cod_r2(Y, pred), "resid_w": rw,
            "resid": Y - pred, "pred": pred, "Ainv": Ai, "dof": dof}
