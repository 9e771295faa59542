"""Learned (a, b) proposal + local REML search.

A small MLP maps the lag-1..10 autocorrelation of each voxel's OLS residual
(computed by the prep kernel) to (a_hat, b_hat, c), c = log-confidence.  It is
trained on the GPU by the project's own forward / backward / Adam kernels
(libffsreml.so, include/ffs_mlp.h) on labels produced by the full-grid GPU
search itself (the global argmin).  At inference the REML score is evaluated
only on the 5 x 5 lattice window around the snapped proposal; voxels whose
local optimum sits on a window side that the lattice continues past, or whose
confidence is below ``conf_min``, go on to the lambda band and then the full
grid.  Window scores
are bitwise equal to the full-grid surface, so an accepted local optimum is
exactly the global argmin whenever the latter lies inside the window.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .glm.arma import RemlGrid

N_IN, N_H, N_OUT = 10, 64, 3


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initialisation (zero biases) in the flat layout
    W1[10][64] b1[64] W2[64][64] b2[64] W3[64][3] b3[3]."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((N_IN, N_H)) * math.sqrt(2.0 / N_IN), np.zeros(N_H),
             rng.standard_normal((N_H, N_H)) * math.sqrt(2.0 / N_H), np.zeros(N_H),
             rng.standard_normal((N_H, N_OUT)) * math.sqrt(1.0 / N_H), np.zeros(N_OUT)]
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


@dataclass
class TrainLog:
    losses: list = field(default_factory=list)
    seconds: float = 0.0
    steps: int = 0


class ProposalMLP:
    """The proposal network, parameters resident on the device."""

    def __init__(self, device: torch.device | None = None, seed: int = 0, params: np.ndarray | None = None):
        self.device = _lib.require_device(device)
        lib = _lib.load()
        self.n_params = int(lib.ffs_mlp_param_count())
        p = init_params(seed) if params is None else np.asarray(params, dtype=np.float32)
        if p.shape != (self.n_params,):
            raise ValueError(f"expected {self.n_params} parameters, got {p.shape}")
        self.params = torch.from_numpy(p.copy()).to(self.device)
        self.conf_min = -math.inf

    def forward(self, feats: torch.Tensor, idx: torch.Tensor | None = None) -> torch.Tensor:
        B = feats.shape[0] if idx is None else idx.shape[0]
        out = torch.empty(B, N_OUT, dtype=torch.float32, device=self.device)
        rc = _lib.load().ffs_mlp_forward(_lib.ptr(self.params), _lib.ptr(feats), _lib.ptr(idx), B, _lib.ptr(out),
                                         _lib.stream_handle(self.device))
        _lib.check(rc, "mlp_forward")
        return out

    def loss_and_grad(self, feats, target, idx=None):
        lib = _lib.load()
        B = feats.shape[0] if idx is None else idx.shape[0]
        grad = torch.empty(self.n_params, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        ws = torch.empty(int(lib.ffs_mlp_workspace_floats(B)), dtype=torch.float32, device=self.device)
        rc = lib.ffs_mlp_grad(_lib.ptr(self.params), _lib.ptr(feats), _lib.ptr(target), _lib.ptr(idx), B,
                              _lib.ptr(grad), _lib.ptr(loss), _lib.ptr(ws), _lib.stream_handle(self.device))
        _lib.check(rc, "mlp_grad")
        return loss, grad

    def train(self, feats: torch.Tensor, target: torch.Tensor, epochs: int = 8, batch: int = 4096,
              lr: float = 2e-3, lr_final: float = 1e-4, seed: int = 0, log_every: int = 0,
              standardize: bool = True) -> TrainLog:
        """Adam on the device; mini-batches from a seeded host permutation.

        With ``standardize`` the network trains on z-scored features and the
        standardisation is folded exactly into the first layer afterwards
        (W1 <- W1 / sd, b1 <- b1 - (mean / sd) W1), so inference takes raw
        features."""
        if standardize:
            mu = feats.mean(dim=0)
            sd = feats.std(dim=0).clamp(min=1e-6)
            log = self.train((feats - mu) / sd, target, epochs, batch, lr, lr_final, seed, log_every, False)
            self._fold_standardization(mu.double().cpu().numpy(), sd.double().cpu().numpy())
            return log
        lib = _lib.load()
        dev = self.device
        n = feats.shape[0]
        rng = np.random.default_rng(seed)
        batch = max(1, min(batch, n))
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        ws = torch.empty(int(lib.ffs_mlp_workspace_floats(batch)), dtype=torch.float32, device=dev)
        steps_per_epoch = max(1, n // batch)
        total = epochs * steps_per_epoch
        log = TrainLog()
        st = _lib.stream_handle(dev)
        torch.cuda.synchronize(dev)  # the timing covers the training only, not work queued before it
        t0 = time.perf_counter()
        step = 0
        losses = torch.empty(steps_per_epoch, dtype=torch.float32, device=dev)
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n).astype(np.int32)).to(dev)
            # one epoch of steps launched from the library (no host loop per step)
            _lib.check(lib.ffs_mlp_train_steps(_lib.ptr(self.params), _lib.ptr(m), _lib.ptr(v), _lib.ptr(feats),
                                               _lib.ptr(target), _lib.ptr(perm), batch, steps_per_epoch, step, total,
                                               lr, lr_final, 0.9, 0.999, 1e-8, _lib.ptr(losses), _lib.ptr(ws), st),
                       "mlp_train_steps")
            if log_every:
                lh = losses.cpu().tolist()
                log.losses += [(step + i + 1, lh[i]) for i in range(steps_per_epoch) if (step + i + 1) % log_every == 0]
            step += steps_per_epoch
        torch.cuda.synchronize(dev)
        log.seconds = time.perf_counter() - t0
        log.steps = step
        return log

    def _fold_standardization(self, mu: np.ndarray, sd: np.ndarray) -> None:
        p = self.params.detach().cpu().double().numpy()
        W1 = p[:N_IN * N_H].reshape(N_IN, N_H) / sd[:, None]
        b1 = p[N_IN * N_H:N_IN * N_H + N_H] - mu @ W1
        p[:N_IN * N_H] = W1.reshape(-1)
        p[N_IN * N_H:N_IN * N_H + N_H] = b1
        self.params = torch.from_numpy(p.astype(np.float32)).to(self.device)

    def state(self) -> np.ndarray:
        return self.params.detach().cpu().numpy().copy()

    def save(self, path, meta: dict | None = None) -> None:
        """Write the network as plain arrays (``numpy.savez``): the float32
        parameters in the flat layout of ``init_params``, the confidence
        threshold and a JSON metadata string stored as bytes, so loading needs
        no pickle."""
        import json

        info = dict(meta or {})
        info.setdefault("layout", "W1[10][64] b1[64] W2[64][64] b2[64] W3[64][3] b3[3], raw lag-1..10 residual ACF in")
        blob = np.frombuffer(json.dumps(info, sort_keys=True).encode(), dtype=np.uint8)
        with open(path, "wb") as fh:
            np.savez(fh, params=self.state().astype(np.float32), conf_min=np.float64(self.conf_min), meta=blob)

    @classmethod
    def load(cls, path, device: torch.device | None = None) -> "ProposalMLP":
        """Read a network written by ``save`` (no pickle is loaded)."""
        import json

        with np.load(path, allow_pickle=False) as z:
            params = np.asarray(z["params"], dtype=np.float32)
            conf_min = float(z["conf_min"])
            meta = json.loads(bytes(np.asarray(z["meta"], dtype=np.uint8)).decode()) if "meta" in z else {}
        net = cls(device, params=params)
        net.conf_min = conf_min
        net.meta = meta
        return net


def features(grid: RemlGrid, Y: torch.Tensor) -> torch.Tensor:
    """Lag-1..10 autocorrelation of the OLS residual of every voxel (device)."""
    _, _, feats = grid.prep_features(Y)
    return feats


def synth_training_voxels(grid: RemlGrid, V: int, seed: int = 0, a_range=(0.1, 0.8), b_range=(-0.5, 0.3),
                          chunk: int = 250_000) -> torch.Tensor:
    """Training series for the proposal network on the grid's own timeline
    (device, (V, T) float32): stationary unit-variance ARMA(1,1) noise with a,
    b uniform in the given ranges, restarted at every run and sampled at the
    retained times only (censored samples are dropped), plus N(0, 1)
    coefficients on the grid's design."""
    from .glm.arma import run_bounds
    from .synth import arma11_noise_torch

    dev = grid.device
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    T = grid.T
    tau = None if grid.tau is None else np.asarray(torch.as_tensor(grid.tau).flatten().tolist(), dtype=np.int64)
    runs = []
    for s, e in run_bounds(T, grid.run_starts):
        t_run = np.arange(e - s) if tau is None else tau[s:e] - int(tau[s])
        runs.append((s, e, torch.from_numpy(t_run).to(dev), int(t_run[-1]) + 1))
    Xd = torch.from_numpy(grid.X64).to(dev, torch.float32)
    Y = torch.empty(V, T, dtype=torch.float32, device=dev)
    for c0 in range(0, V, chunk):
        c1 = min(V, c0 + chunk)
        n = c1 - c0
        a = torch.rand(n, generator=g, device=dev) * (a_range[1] - a_range[0]) + a_range[0]
        b = torch.rand(n, generator=g, device=dev) * (b_range[1] - b_range[0]) + b_range[0]
        beta = torch.randn(n, grid.p, generator=g, device=dev)
        Yc = beta @ Xd.T
        for s, e, t_run, length in runs:
            Yc[:, s:e] += arma11_noise_torch(g, a, b, length)[:, t_run]
        Y[c0:c1] = Yc
    return Y


def train_proposal(grid: RemlGrid, n_train: int = 1_000_000, epochs: int = 10, n_val: int = 100_000,
                   target: float = 0.9998, seed: int = 0, batch: int = 4096, lr: float = 3e-3,
                   lr_final: float = 1e-4, log_every: int = 200) -> ProposalMLP:
    """Train the proposal network for ``grid``'s design and timeline on
    synthetic voxels (``synth_training_voxels``) labelled by the full-grid GPU
    search (global argmin), then calibrate its confidence threshold on
    ``n_val`` fresh voxels for ``target`` agreement with the full grid (with
    the grid's current lambda bands).  ``train_log`` and ``meta`` on the
    returned network record what was done."""
    t0 = time.perf_counter()
    Ytr = synth_training_voxels(grid, n_train, seed=seed)
    labels, _ = full_grid_labels(grid, Ytr)
    feats = features(grid, Ytr)
    del Ytr
    mlp = ProposalMLP(grid.device, seed=seed)
    log = mlp.train(feats, labels, epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, seed=seed,
                    log_every=log_every)
    del feats, labels
    Yv = synth_training_voxels(grid, n_val, seed=seed + 7919)
    _, best_v = full_grid_labels(grid, Yv)
    calibrate_conf_min(mlp, grid, Yv, best_v, target=target)
    torch.cuda.synchronize(grid.device)
    mlp.train_log = log
    mlp.meta = {"T": grid.T, "p": grid.p, "K": grid.K, "n_train": n_train, "epochs": epochs, "n_val": n_val,
                "target": target, "seed": seed, "bands": list(grid.bands) if grid.bands else None,
                "train_s": time.perf_counter() - t0, "train_steps": log.steps}
    return mlp


def full_grid_labels(grid: RemlGrid, Y: torch.Tensor, batch: int = 500_000):
    """Global-argmin (a, b) of the full-grid GPU search, per voxel."""
    V = Y.shape[0]
    labels = torch.empty(V, 2, dtype=torch.float32, device=grid.device)
    best_all = torch.empty(V, dtype=torch.int32, device=grid.device)
    params = grid.params.to(grid.device, torch.float32)
    for b0 in range(0, V, batch):
        b1 = min(V, b0 + batch)
        Yt, _, _ = grid.prep(Y[b0:b1])
        S = grid.surface(Yt, b1 - b0)
        del Yt
        best, _ = grid.select(S, hierarchical=False)
        del S
        best_all[b0:b1] = best
        labels[b0:b1] = params[best.long()]
    return labels, best_all


@dataclass
class LocalResult:
    best: torch.Tensor
    score: torch.Tensor
    scale: torch.Tensor
    n_fallback: int
    props: torch.Tensor


class LocalSearcher:
    """Reusable device buffers for the learned-proposal search of one grid."""

    def __init__(self, grid: RemlGrid, mlp: ProposalMLP, V: int):
        self.grid, self.mlp, self.V = grid, mlp, V
        lib = _lib.load()
        nb = ctypes.c_int64()
        _lib.check(lib.ffs_reml_local_workspace(grid.handle, V, ctypes.byref(nb)), "local_workspace")
        dev = grid.device
        self.ws = torch.zeros(int(nb.value), dtype=torch.uint8, device=dev)
        self.best = torch.empty(V, dtype=torch.int32, device=dev)
        self.score = torch.empty(V, dtype=torch.float32, device=dev)
        self.need = torch.empty(V, dtype=torch.int32, device=dev)
        self.scale = torch.empty(V, dtype=torch.float32, device=dev)
        self.idx = torch.empty(V, dtype=torch.int32, device=dev)
        self.count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.feats = torch.empty(V, 10, dtype=torch.float32, device=dev)
        # residualisation coefficients of the features pass, reused by the
        # window / band / subset preparations (None where the plan keeps none)
        stride = ctypes.c_int32()
        _lib.check(lib.ffs_reml_plan_coef_stride(grid.handle, ctypes.byref(stride)), "plan_coef_stride")
        self.coef = (torch.empty(V, stride.value, dtype=torch.float64, device=dev) if stride.value > 0 else None)
        self.bws = None
        if grid.bands is not None:
            _lib.check(lib.ffs_reml_band_workspace(grid.handle, V, ctypes.byref(nb)), "band_workspace")
            self.bws = torch.zeros(int(nb.value), dtype=torch.uint8, device=dev)
        # full-grid voxels handled without a host round trip (more take a second pass)
        self.cap = max(1, min(V, max(4096, V // 16) if self.bws is not None else max(4096, V // 3)))
        _lib.check(lib.ffs_reml_subset_workspace(grid.handle, self.cap, ctypes.byref(nb)), "subset_workspace")
        self.sws = torch.empty(int(nb.value), dtype=torch.uint8, device=dev)

    def run(self, Y: torch.Tensor, timer: "StageTimer | None" = None, walk: int | None = None,
            stats: dict | None = None, band: bool = True, defer_count: bool = False) -> LocalResult:
        """Learned-proposal search of Y (V, T) float32 on the device.

        Voxels whose window optimum lies on a window edge take up to ``walk``
        further 5x5 windows centred on that optimum.  Those still on an edge,
        and those with low confidence, then take the band search when the grid
        has lambda bands (``RemlGrid.set_bands``) and ``band`` is set: every
        pair whose lambda lies near the proposal's.  What remains (band optimum
        on the band's edge, non-finite proposals) takes the full grid.
        ``walk`` defaults to 0 with bands (the band search handles window
        edges better than further windows) and 1 without.  ``timer``
        (optional) records a device event at the start of every stage;
        ``stats`` (optional, synchronises) receives the voxel count of each
        pass ("window": first pass, "walk": list, "band").  ``defer_count``:
        return without reading the full-grid voxel count on the host (no
        synchronisation; ``n_fallback`` is -1), so the caller can queue its
        next work first and then call ``settle``, which reads the count and
        searches any voxels past the device-side capacity."""
        grid, lib, dev = self.grid, _lib.load(), self.grid.device
        use_band = band and self.bws is not None
        if walk is None:
            walk = 0 if use_band else 1
        V = Y.shape[0]
        if V != self.V:
            raise ValueError(f"LocalSearcher was sized for {self.V} voxels, got {V}")
        st = _lib.stream_handle(dev)

        def mark(name):
            if timer is not None:
                timer.mark(name)

        mark("features")
        coef = _lib.ptr(self.coef) if self.coef is not None else None
        _lib.check(lib.ffs_reml_prep_features(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(self.scale), 1,
                                              _lib.ptr(self.feats), coef, st), "prep_features")
        mark("mlp")
        props = self.mlp.forward(self.feats)
        mark("local_prepare")
        ws, nb = _lib.ptr(self.ws), self.ws.numel()
        _lib.check(lib.ffs_reml_local_prepare(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(props),
                                              float(self.mlp.conf_min), _lib.ptr(self.need), _lib.ptr(self.scale),
                                              coef, ws, nb, st), "local_prepare")
        if stats is not None:
            stats["window"], stats["window_slots"] = window_load(grid, props, float(self.mlp.conf_min))
        mark("local_energy")
        _lib.check(lib.ffs_reml_local_energy(grid.handle, V, ws, nb, st), "local_energy")
        mark("local_project")
        _lib.check(lib.ffs_reml_local_project(grid.handle, V, _lib.ptr(self.best), _lib.ptr(self.score),
                                              _lib.ptr(self.need), ws, nb, st), "local_project")
        for _ in range(walk):
            if stats is not None:
                stats.setdefault("walk", []).append(int((self.need == 1).sum()))
            mark("walk_prepare")
            _lib.check(lib.ffs_reml_local_recenter(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(self.scale),
                                                   coef, ws, nb, st), "local_recenter")
            mark("walk_energy")
            _lib.check(lib.ffs_reml_local_energy(grid.handle, V, ws, nb, st), "local_energy")
            mark("walk_project")
            _lib.check(lib.ffs_reml_local_project(grid.handle, V, _lib.ptr(self.best), _lib.ptr(self.score),
                                                  _lib.ptr(self.need), ws, nb, st), "local_project")
        if use_band:
            if stats is not None:
                n_vox, n_slots = band_load(grid, props, self.need)
                stats["band"], stats["band_slots"] = n_vox, n_slots
            bws, bnb = _lib.ptr(self.bws), self.bws.numel()
            mark("band_prepare")
            _lib.check(lib.ffs_reml_band_prepare(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(props),
                                                 _lib.ptr(self.need), _lib.ptr(self.scale), coef, bws, bnb, st),
                       "band_prepare")
            mark("band_energy")
            _lib.check(lib.ffs_reml_band_energy(grid.handle, V, bws, bnb, st), "band_energy")
            mark("band_project")
            _lib.check(lib.ffs_reml_band_project(grid.handle, V, _lib.ptr(self.best), _lib.ptr(self.score),
                                                 _lib.ptr(self.need), bws, bnb, st), "band_project")
        mark("fallback")
        _lib.check(lib.ffs_reml_compact(_lib.ptr(self.need), V, _lib.ptr(self.idx), _lib.ptr(self.count), st),
                   "compact")
        # full grid for the first `cap` flagged voxels without reading the count
        # on the host; the count is read once everything is queued
        _lib.check(lib.ffs_reml_subset_search(grid.handle, _lib.ptr(Y), Y.stride(0), V, _lib.ptr(self.idx),
                                              _lib.ptr(self.count), self.cap, _lib.ptr(self.best),
                                              _lib.ptr(self.score), _lib.ptr(self.scale), coef, _lib.ptr(self.sws),
                                              self.sws.numel(), st), "subset_search")
        if defer_count:
            return LocalResult(self.best, self.score, self.scale, -1, props)
        n_fb = self.settle(Y, timer)
        return LocalResult(self.best, self.score, self.scale, n_fb, props)

    def settle(self, Y: torch.Tensor, timer: "StageTimer | None" = None) -> int:
        """Full-grid voxel count of the last ``run`` (reads it on the host);
        voxels past the capacity of the device-counted search are searched now
        (best / score updated in place).  Returns the count; a count above
        ``cap`` means ``best`` changed after a deferred ``run``."""
        grid, lib, dev = self.grid, _lib.load(), self.grid.device
        st = _lib.stream_handle(dev)
        n_fb = int(self.count.item())
        self.last_n_fallback = n_fb
        if n_fb > self.cap:
            if timer is not None:
                timer.mark("fallback_overflow")
            rows = self.idx[self.cap:n_fb]
            n_rest = n_fb - self.cap
            Yt, _, _ = grid.prep(Y, rows=rows, scale=self.scale)
            S = torch.empty(n_rest, grid.K, dtype=torch.float32, device=dev)
            _lib.check(lib.ffs_reml_energy(grid.handle, _lib.ptr(Yt), Yt.stride(0), n_rest, _lib.ptr(S),
                                           S.stride(0), st), "energy")
            _lib.check(lib.ffs_reml_project(grid.handle, _lib.ptr(Yt), Yt.stride(0), n_rest, _lib.ptr(S),
                                            S.stride(0), st), "project")
            del Yt
            b_fb, s_fb = grid.select(S, hierarchical=False)
            del S
            _lib.check(lib.ffs_reml_scatter_best(_lib.ptr(rows), n_rest, _lib.ptr(b_fb), _lib.ptr(s_fb),
                                                 _lib.ptr(self.best), _lib.ptr(self.score), st), "scatter_best")
        return n_fb


def window_load(grid: RemlGrid, props: torch.Tensor, conf_min: float) -> tuple[int, int]:
    """(voxels, total pair evaluations) of the first window pass for these
    proposals: the library's window choice (nearest lattice cell, centre kept
    two cells inside the lattice; confidence >= conf_min and finite proposal)
    restated, with the window sizes read from the plan's window table."""
    _, counts, _ = grid.window_info()
    n_a, n_b = grid.lut.shape
    half = 2
    a_ax = torch.tensor(grid.a_axis, dtype=torch.float32, device=props.device)
    b_ax = torch.tensor(grid.b_axis, dtype=torch.float32, device=props.device)
    ok = torch.isfinite(props).all(dim=1) & (props[:, 2] >= conf_min)
    ai = (props[:, 0:1] - a_ax[None]).abs().argmin(dim=1)
    bi = (props[:, 1:2] - b_ax[None]).abs().argmin(dim=1)
    if n_a >= 2 * half + 1:
        ai = ai.clamp(half, n_a - 1 - half)
    if n_b >= 2 * half + 1:
        bi = bi.clamp(half, n_b - 1 - half)
    n = counts.to(props.device).long()[ai * n_b + bi]
    return int(ok.sum()), int(n[ok].sum())


def band_load(grid: RemlGrid, props: torch.Tensor, need: torch.Tensor) -> tuple[int, int]:
    """(voxels, total pair evaluations) the band search takes for these
    proposals and flags: the band choice of the library (lambda of the
    proposal clamped to the lattice box, nearest band centre), restated."""
    info = grid.band_info()
    if info is None:
        return 0, 0
    _, cnt, _ = info
    hw, h = grid.bands
    a = props[:, 0].double().clamp(min(grid.a_axis), max(grid.a_axis))
    b = props[:, 1].double().clamp(min(grid.b_axis), max(grid.b_axis))
    ok = (need != 0) & torch.isfinite(props[:, 0]) & torch.isfinite(props[:, 1])
    den = 1.0 + 2.0 * a * b + b * b
    lam = torch.where(den > 0, (1.0 + a * b) * (a + b) / den.clamp(min=1e-300), torch.zeros_like(a)).clamp(-1.0, 1.0)
    h32 = float(np.float32(h))
    bins = torch.round((lam + 1.0) / h32).long().clamp(0, cnt.numel() - 1)
    n = cnt.to(props.device).long()[bins]
    ok = ok & (n > 0)
    return int(ok.sum()), int(n[ok].sum())


class StageTimer:
    """Device events at stage starts on the current stream; ``durations()``
    (after a synchronize) sums the time from each mark to the next by name."""

    def __init__(self, device):
        self.stream = torch.cuda.current_stream(device)
        self.events: list = []

    def mark(self, name: str):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(self.stream)
        self.events.append((name, ev))

    def durations(self) -> dict:
        out: dict = {}
        for (name, a), (_, b) in zip(self.events, self.events[1:]):
            out[name] = out.get(name, 0.0) + a.elapsed_time(b)
        return out

    def counts(self) -> dict:
        out: dict = {}
        for name, _ in self.events[:-1]:
            out[name] = out.get(name, 0) + 1
        return out


def calibrate_conf_min(mlp: ProposalMLP, grid: RemlGrid, Y_val: torch.Tensor, best_val: torch.Tensor,
                       target: float = 0.999, walk: int | None = None) -> float:
    """Smallest-fallback confidence threshold whose local-search agreement with
    the global argmin on a validation set is >= target.  Sets mlp.conf_min."""
    mlp.conf_min = -math.inf
    res = LocalSearcher(grid, mlp, Y_val.shape[0]).run(Y_val, walk=walk)
    c = res.props[:, 2].detach()
    ok = (res.best == best_val)
    # voxels that went to the full grid are already right; find the threshold
    # on c that also sends enough of the wrong ones there
    order = torch.argsort(c)
    wrong = (~ok).to(torch.float64)[order]
    n = c.numel()
    cum_wrong_below = torch.cumsum(wrong, 0)  # wrong voxels with c <= c[order[i]]
    total_wrong = float(wrong.sum())
    need = total_wrong - (1.0 - target) * n
    if need <= 0:
        mlp.conf_min = -math.inf
        return mlp.conf_min
    pos = int(torch.searchsorted(cum_wrong_below, torch.tensor(need, dtype=torch.float64, device=c.device)))
    pos = min(pos, n - 1)
    mlp.conf_min = float(c[order[pos]].item()) + 1e-6
    return mlp.conf_min
