"""Seeded synthetic fMRI workloads for the ARMA(1,1) REML path.

Designs follow the reference's inputs: a double-gamma HRF (gamma pdf shape 6
minus 1/6 of shape 16, unit scale, fastfuncstuff/design/hrf.py:354-372)
convolved with boxcars, Legendre polynomial baselines on [-1, 1]
(design/builder.py:105-140), AR(1) random-walk motion-like columns; noise is
stationary ARMA(1,1) with unit marginal variance.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy import stats


def double_gamma_hrf(tr: float, peak: float = 6.0, under: float = 16.0, ratio: float = 1.0 / 6.0,
                     duration: float = 32.0, dt: float = 0.1) -> tuple[np.ndarray, float]:
    """High-resolution double-gamma kernel (sampled every ``dt`` s) and its dt."""
    t = np.arange(0.0, duration, dt)
    h = stats.gamma.pdf(t, peak) - ratio * stats.gamma.pdf(t, under)
    return h / np.abs(h).max(), dt


def boxcar_regressor(onsets: np.ndarray, dur: float, n: int, tr: float, hrf=None) -> np.ndarray:
    """Boxcars at ``onsets`` (s) of length ``dur`` (s), convolved with the HRF at
    high resolution and sampled every TR, scaled to unit peak."""
    h, dt = hrf if hrf is not None else double_gamma_hrf(tr)
    nh = int(round(n * tr / dt))
    s = np.zeros(nh)
    for o in onsets:
        i0 = int(round(o / dt))
        s[i0:min(nh, i0 + max(1, int(round(dur / dt))))] = 1.0
    x = np.convolve(s, h)[:nh][:: int(round(tr / dt))][:n]
    m = np.abs(x).max()
    return x / m if m > 0 else x


def legendre_baseline(n: int, order: int) -> np.ndarray:
    """Legendre polynomials P_0..P_order on [-1, 1], unnormalised (AFNI -polort)."""
    x = np.linspace(-1.0, 1.0, n)
    return np.stack([np.polynomial.legendre.Legendre.basis(k)(x) for k in range(order + 1)], axis=1)


def arma11_noise(rng: np.random.Generator, a: np.ndarray, b: np.ndarray, T: int, df: np.ndarray | None = None,
                 dtype=np.float32) -> np.ndarray:
    """Stationary ARMA(1,1) x_t = a x_{t-1} + e_t + b e_{t-1}, unit marginal variance.

    ``df`` (optional, per voxel, inf = Gaussian) draws Student-t innovations
    scaled to unit variance.  Returns (V, T)."""
    V = a.shape[0]
    sig2 = (1 - a**2) / (1 + 2 * a * b + b**2)
    sd = np.sqrt(sig2)

    def innov(size):
        z = rng.standard_normal(size)
        if df is not None:
            heavy = np.isfinite(df)
            if heavy.any():
                d = df[heavy][:, None] if z.ndim == 2 else df[heavy]
                tz = rng.standard_t(np.broadcast_to(d, z[heavy].shape))
                z[heavy] = tz / np.sqrt(d / (d - 2.0))
        return z

    out = np.empty((V, T), dtype=np.float64)
    e_prev = innov(V) * sd
    # stationary start: Var(x)=1, Cov(x_0, e_0) = sig2
    x = e_prev + np.sqrt(np.maximum(1.0 - sig2, 0.0)) * rng.standard_normal(V)
    out[:, 0] = x
    for t in range(1, T):
        e = innov(V) * sd
        x = a * x + e + b * e_prev
        out[:, t] = x
        e_prev = e
    return out.astype(dtype)


def arma11_noise_torch(gen, a, b, T: int):
    """arma11_noise on the device with a torch.Generator (same model)."""
    import torch

    V = a.shape[0]
    sig2 = (1 - a * a) / (1 + 2 * a * b + b * b)
    sd = torch.sqrt(sig2)
    out = torch.empty(V, T, dtype=torch.float32, device=a.device)
    e_prev = torch.randn(V, generator=gen, device=a.device) * sd
    x = e_prev + torch.sqrt(torch.clamp(1.0 - sig2, min=0.0)) * torch.randn(V, generator=gen, device=a.device)
    out[:, 0] = x
    z = torch.randn(T - 1, V, generator=gen, device=a.device) * sd
    for t in range(1, T):
        e = z[t - 1]
        x = a * x + e + b * e_prev
        out[:, t] = x
        e_prev = e
    return out


@dataclass
class Workload:
    name: str
    Y: np.ndarray  # (V, T_retained) float32
    X: np.ndarray  # (T_retained, p) float64
    tr: float
    run_starts: list | None
    tau: np.ndarray | None
    a_true: np.ndarray
    b_true: np.ndarray
    betas: np.ndarray
    a_grid: np.ndarray
    b_grid: np.ndarray
    task_indices: list | None = None


def config_grid() -> tuple[np.ndarray, np.ndarray]:
    """a, b in [-0.8, 0.8] step 0.05 (33 x 33), float32 linspace."""
    import torch

    g = torch.linspace(-0.8, 0.8, 33).numpy()
    return g.copy(), g.copy()


def plain_config(V: int = 20_000, T: int = 200, tr: float = 2.0, seed: int = 1, snr_range=None,
                 noise_seed: int | None = None) -> Workload:
    """Two task boxcar regressors (20 random-onset 10 s blocks each) + Legendre
    0-3 baseline (p=6), ARMA a~U(0.1,0.8), b~U(-0.5,0.3), betas N(0,1).
    ``snr_range=(lo, hi)``: per-voxel task amplitude log-uniform in that range."""
    rng = np.random.default_rng(seed)
    hrf = double_gamma_hrf(tr)
    dur = T * tr
    cols = []
    for _ in range(2):
        on = np.sort(rng.uniform(0, dur - 10.0, 20))
        cols.append(boxcar_regressor(on, 10.0, T, tr, hrf))
    X = np.concatenate([np.stack(cols, 1), legendre_baseline(T, 3)], 1)
    if noise_seed is not None:
        rng = np.random.default_rng(noise_seed)
    a = rng.uniform(0.1, 0.8, V)
    b = rng.uniform(-0.5, 0.3, V)
    beta = rng.standard_normal((V, X.shape[1]))
    if snr_range is not None:
        amp = np.exp(rng.uniform(np.log(snr_range[0]), np.log(snr_range[1]), V))
        beta[:, :2] *= amp[:, None] / np.maximum(np.abs(beta[:, :2]).max(1, keepdims=True), 1e-6)
    noise = arma11_noise(rng, a, b, T)
    Y = (beta @ X.T + noise).astype(np.float32)
    ag, bg = config_grid()
    return Workload("plain_T200_p6", Y, X, tr, None, None, a, b, beta, ag, bg)


def _design_p12(rng, T: int, tr: float, censor_frac: float):
    """The p=12 design of gpu_config over all T acquired TRs and the retained
    TR indices (draws from ``rng`` in a fixed order)."""
    hrf = double_gamma_hrf(tr)
    cols = []
    for _ in range(4):
        on = np.sort(rng.uniform(0, T * tr - 10.0, 20))
        cols.append(boxcar_regressor(on, 10.0, T, tr, hrf))
    motion = np.zeros((T, 6))
    for j in range(6):
        m = 0.0
        for t in range(T):
            m = 0.9 * m + rng.standard_normal() * 0.1
            motion[t, j] = m
    motion -= motion.mean(0)
    lin = np.linspace(-1.0, 1.0, T)
    Xf = np.concatenate([np.stack(cols, 1), motion, np.ones((T, 1)), lin[:, None]], 1)
    n_cens = int(round(censor_frac * T))
    censored = np.sort(rng.choice(T, n_cens, replace=False)) if n_cens else np.zeros(0, int)
    keep = np.setdiff1d(np.arange(T), censored)
    return Xf, keep


def gpu_config(V: int = 200_000, T: int = 400, tr: float = 2.0, seed: int = 2, censor_frac: float = 0.05,
               noise_seed: int | None = None, chunk: int = 250_000, device=None) -> Workload:
    """p=12: 4 HRF-convolved boxcar task regressors (20 random 10 s blocks each),
    6 AR(1) random-walk motion-like columns, constant + linear; 5% of the TRs
    censored uniformly at random (the retained series keeps its true time
    index); ARMA(1,1) noise as in the plain config.  ``seed`` fixes the design
    and censoring; ``noise_seed`` (default: ``seed``) the voxels.  With
    ``device`` the voxels are drawn by torch on that device (Y stays there)."""
    rng = np.random.default_rng(seed)
    Xf, keep = _design_p12(rng, T, tr, censor_frac)
    Xk = Xf[keep]
    ag, bg = config_grid()
    if device is not None:
        import torch

        g = torch.Generator(device=device)
        g.manual_seed(seed if noise_seed is None else noise_seed)
        Y = torch.empty(V, keep.size, dtype=torch.float32, device=device)
        Xd = torch.from_numpy(Xf).to(device=device, dtype=torch.float32)
        keep_d = torch.from_numpy(keep).to(device)
        a_all = torch.empty(V, device=device)
        b_all = torch.empty(V, device=device)
        for c0 in range(0, V, chunk):
            c1 = min(V, c0 + chunk)
            n = c1 - c0
            a = torch.rand(n, generator=g, device=device) * 0.7 + 0.1
            b = torch.rand(n, generator=g, device=device) * 0.8 - 0.5
            beta = torch.randn(n, Xf.shape[1], generator=g, device=device)
            noise = arma11_noise_torch(g, a, b, T)
            Y[c0:c1] = (beta @ Xd.T + noise)[:, keep_d]
            a_all[c0:c1], b_all[c0:c1] = a, b
        return Workload(f"censored_T{T}_p12", Y, Xk, tr, [0], keep.astype(np.int64), a_all, b_all, None, ag, bg)
    if noise_seed is not None:
        rng = np.random.default_rng(noise_seed)
    a = rng.uniform(0.1, 0.8, V)
    b = rng.uniform(-0.5, 0.3, V)
    beta = rng.standard_normal((V, Xf.shape[1]))
    Y = np.empty((V, keep.size), dtype=np.float32)
    for c0 in range(0, V, chunk):
        c1 = min(V, c0 + chunk)
        noise = arma11_noise(rng, a[c0:c1], b[c0:c1], T)
        Y[c0:c1] = beta[c0:c1] @ Xk.T + noise[:, keep]
    return Workload(f"censored_T{T}_p12", Y, Xk, tr, [0], keep.astype(np.int64), a, b, beta, ag, bg)


def _gaussian_smooth3d(field, fwhm: float):
    """Separable Gaussian smoothing of a (nx, ny, nz) torch tensor (FWHM in voxels)."""
    import torch
    import torch.nn.functional as F

    sigma = fwhm / 2.3548200450309493
    r = int(math.ceil(3 * sigma))
    x = torch.arange(-r, r + 1, dtype=field.dtype, device=field.device)
    k = torch.exp(-0.5 * (x / sigma) ** 2)
    k = (k / k.sum()).view(1, 1, -1)
    f = field
    for dim in range(3):
        f = f.movedim(dim, -1)
        sh = f.shape
        f = F.conv1d(f.reshape(-1, 1, sh[-1]), k, padding=r).reshape(sh).movedim(-1, dim)
    return f


def learned_eval_config(shape=(100, 100, 50), T: int = 600, tr: float = 2.0, seed: int = 5, heavy_frac: float = 0.02,
                        fwhm: float = 4.0, censor_frac: float = 0.05, device=None) -> Workload:
    """Evaluation volume for the learned proposal: the p=12 design of
    gpu_config at T acquired TRs (5% censored), ARMA parameters that are
    spatially smooth over the voxel grid (two Gaussian random fields, FWHM in
    voxels, mapped through the normal CDF to a in [0.1, 0.8] and b in
    [-0.5, 0.3]), and a fraction of voxels with Student-t(3) innovations
    (scaled to unit variance) instead of Gaussian ones.  Voxels in C order of
    ``shape``; Y is drawn on ``device``."""
    import torch

    Xf, keep_np = _design_p12(np.random.default_rng(seed), T, tr, censor_frac)
    ag, bg = config_grid()
    V = int(np.prod(shape))
    g = torch.Generator(device=device)
    g.manual_seed(seed + 1000)
    u = []
    for _ in range(2):
        f = _gaussian_smooth3d(torch.randn(*shape, generator=g, device=device), fwhm)
        f = (f - f.mean()) / f.std()
        u.append((0.5 * (1 + torch.erf(f / math.sqrt(2.0)))).reshape(-1))
    a = 0.1 + 0.7 * u[0]
    b = -0.5 + 0.8 * u[1]
    heavy = torch.rand(V, generator=g, device=device) < heavy_frac
    Xf_full = torch.from_numpy(Xf).to(device=device, dtype=torch.float32)
    keep = torch.from_numpy(keep_np).to(device)
    Y = torch.empty(V, keep.numel(), dtype=torch.float32, device=device)
    chunk = 125_000
    for c0 in range(0, V, chunk):
        c1 = min(V, c0 + chunk)
        n = c1 - c0
        aa, bb, hh = a[c0:c1], b[c0:c1], heavy[c0:c1]
        sig2 = (1 - aa * aa) / (1 + 2 * aa * bb + bb * bb)
        sd = torch.sqrt(sig2)

        def innov(shape_):
            z = torch.randn(*shape_, generator=g, device=device)
            chi = (torch.randn(3, *shape_, generator=g, device=device) ** 2).sum(0)
            t3 = z / torch.sqrt(chi / 3.0) / math.sqrt(3.0)
            return torch.where(hh if len(shape_) == 1 else hh[None, :], t3, z)

        noise = torch.empty(n, T, dtype=torch.float32, device=device)
        e_prev = innov((n,)) * sd
        x = e_prev + torch.sqrt(torch.clamp(1.0 - sig2, min=0.0)) * innov((n,))
        noise[:, 0] = x
        zz = innov((T - 1, n)) * sd
        for t in range(1, T):
            e = zz[t - 1]
            x = aa * x + e + bb * e_prev
            noise[:, t] = x
            e_prev = e
        beta = torch.randn(n, Xf_full.shape[1], generator=g, device=device)
        Y[c0:c1] = (beta @ Xf_full.T + noise)[:, keep]
    return Workload(f"grf_T{T}_p12", Y, Xf[keep_np], tr, [0], keep_np.astype(np.int64), a, b, None, ag, bg)



def grouped_config(V: int = 200_000, T: int = 400, tr: float = 2.0, n_groups: int = 20, n_events: int = 64,
                   seed: int = 4, chunk: int = 100_000):
    """Uneven per-group designs: ``n_groups`` double-gamma HRFs (peak delay
    U(4, 7) s, undershoot at peak + 10 s with ratio U(0.1, 0.35)), one shared
    single-trial design of ``n_events`` 1 s events at random onsets (one
    regressor each, convolved with the group's HRF) + Legendre 0-3 (p = n_events
    + 4); Dirichlet(0.5) group sizes; ARMA(1,1) noise as in the plain config.
    Returns (Y (V, T) float32, designs_by_group {g: (T, p) float64},
    group_indices (V,) int64, a_grid, b_grid)."""
    rng = np.random.default_rng(seed)
    onsets = np.sort(rng.uniform(0.0, (T - 15) * tr, n_events))
    base = legendre_baseline(T, 3)
    designs = {}
    for g in range(n_groups):
        peak = rng.uniform(4.0, 7.0)
        hrf = double_gamma_hrf(tr, peak=peak, under=peak + 10.0, ratio=rng.uniform(0.1, 0.35))
        cols = [boxcar_regressor(np.array([o]), 1.0, T, tr, hrf) for o in onsets]
        designs[g] = np.concatenate([np.stack(cols, axis=1), base], axis=1)
    w = rng.dirichlet(np.full(n_groups, 0.5))
    groups = rng.choice(n_groups, size=V, p=w).astype(np.int64)
    a = rng.uniform(0.1, 0.8, V)
    b = rng.uniform(-0.5, 0.3, V)
    Y = np.empty((V, T), dtype=np.float32)
    p = n_events + 4
    for c0 in range(0, V, chunk):
        c1 = min(V, c0 + chunk)
        noise = arma11_noise(rng, a[c0:c1], b[c0:c1], T)
        beta = rng.standard_normal((c1 - c0, p))
        sig = np.empty((c1 - c0, T))
        for g in range(n_groups):
            m = groups[c0:c1] == g
            if m.any():
                sig[m] = beta[m] @ designs[g].T
        Y[c0:c1] = (sig + noise).astype(np.float32)
    ag, bg = config_grid()
    return Y, designs, groups, ag, bg


def im_config(V: int = 50_000, n_runs: int = 5, run_len: int = 300, tr: float = 2.0, n_events: int = 596,
              seed: int = 6, chunk: int = 25_000):
    """Single-trial (IM) design, the shape of the reference's IM REML
    benchmark (about 600 regressors, benchmark/stages/glm_im_reml.py:17-22):
    ``n_events`` 1 s events at random onsets over ``n_runs`` runs of
    ``run_len`` volumes, one double-gamma regressor each, plus Legendre 0-3
    over the series (p = n_events + 4); ARMA(1,1) noise per voxel as in the
    plain config, restarted at each run.  Returns (Y (V, T) float32, X (T, p)
    float64, run_starts, a_grid, b_grid)."""
    rng = np.random.default_rng(seed)
    T = n_runs * run_len
    hrf = double_gamma_hrf(tr)
    onsets = np.sort(rng.uniform(0.0, (T - 15) * tr, n_events))
    cols = [boxcar_regressor(np.array([o]), 1.0, T, tr, hrf) for o in onsets]
    X = np.concatenate([np.stack(cols, axis=1), legendre_baseline(T, 3)], axis=1)
    run_starts = [r * run_len for r in range(n_runs)]
    a = rng.uniform(0.1, 0.8, V)
    b = rng.uniform(-0.5, 0.3, V)
    Y = np.empty((V, T), dtype=np.float32)
    p = X.shape[1]
    for c0 in range(0, V, chunk):
        c1 = min(V, c0 + chunk)
        noise = np.concatenate([arma11_noise(rng, a[c0:c1], b[c0:c1], run_len) for _ in range(n_runs)], axis=1)
        beta = rng.standard_normal((c1 - c0, p)) * 0.5
        Y[c0:c1] = (beta @ X.T + 100.0 + noise).astype(np.float32)
    ag, bg = config_grid()
    return Y, X, run_starts, ag, bg
