"""Per-voxel ARMA(1,1) estimation for the FLOBS / binning callers
(design/flobs.py:2750-2940): the REML grid search on the engine, through the
same public boundary the reference uses (``precompute_reml_grid`` +
``search_voxels_precomputed_grid``)."""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib


def legendre_polynomials(n_timepoints: int, order: int, normalize: bool = False) -> np.ndarray:
    """design/builder.py:105-167: Legendre P_0..P_order at linspace(-1, 1, n)
    (AFNI -polort, unnormalised unless ``normalize``); (n, 0) for order < 0."""
    if order < 0:
        return np.zeros((n_timepoints, 0))
    t = np.linspace(-1.0, 1.0, n_timepoints)
    out = np.stack([np.polynomial.legendre.Legendre.basis(k)(t) for k in range(order + 1)], axis=1)
    if normalize:
        out = out / np.linalg.norm(out, axis=0, keepdims=True)
    return out


def _polort_per_run(polort, n_runs: int) -> list[int]:
    """design/polort.py:70-81."""
    if isinstance(polort, (int, np.integer)):
        return [int(polort)] * n_runs
    degrees = [int(p) for p in polort]
    if len(degrees) != n_runs:
        raise ValueError(f"per-run polort has {len(degrees)} entries for {n_runs} runs")
    return degrees


def estimate_arma11_per_voxel(per_run_data: list, per_run_task_designs: list, polort, *, device=None,
                              verbose: bool = True, batch: int = 500_000) -> np.ndarray:
    """design/flobs.py:2750-2940: design = [task designs stacked over runs |
    block-diagonal Legendre drift of each run's degree], run starts at the run
    boundaries, default (a, b) grids; exhaustive REML search per voxel
    (``search_voxels_precomputed_grid``).  Returns (n_voxels, 2) float32 (a, b)."""
    from ..glm.arma import get_default_arma_grids, precompute_reml_grid, search_voxels_precomputed_grid

    dev = _lib.require_device(device)
    n_runs = len(per_run_data)
    n_tp = [int(torch.as_tensor(d).shape[1]) for d in per_run_data]
    starts = [0]
    for n in n_tp[:-1]:
        starts.append(starts[-1] + n)
    total = int(sum(n_tp))
    task = torch.cat([torch.as_tensor(t, dtype=torch.float64) for t in per_run_task_designs], dim=0)
    degrees = _polort_per_run(polort, n_runs)
    if max(degrees) >= 0:
        Z = torch.zeros(total, sum(d + 1 for d in degrees if d >= 0), dtype=torch.float64)
        col = 0
        for r in range(n_runs):
            Zr = torch.from_numpy(legendre_polynomials(n_tp[r], degrees[r]))
            Z[starts[r]:starts[r] + n_tp[r], col:col + Zr.shape[1]] = Zr
            col += Zr.shape[1]
        X = torch.cat([task, Z], dim=1)
    else:
        X = task
    # float32 design, as the reference builds it
    X = X.to(torch.float32)
    Y = torch.cat([torch.as_tensor(d, dtype=torch.float32) for d in per_run_data], dim=1)
    a_grid, b_grid = get_default_arma_grids()
    if verbose:
        print(f"  per-voxel ARMA: REML grid {len(a_grid)} a x {len(b_grid)} b over {total} TR", flush=True)
    grid = precompute_reml_grid(X, total, a_grid, b_grid, device=dev, run_starts=starts)
    V = Y.shape[0]
    best = torch.empty(V, 2, dtype=torch.float32)
    for b0 in range(0, V, batch):
        b1 = min(V, b0 + batch)
        bp, _ = search_voxels_precomputed_grid(X, Y[b0:b1], grid, device=dev)
        best[b0:b1] = bp.detach().cpu().to(torch.float32)
    return best.numpy().astype(np.float32)
