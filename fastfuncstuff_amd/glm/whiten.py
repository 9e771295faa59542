"""ARMA(1,1) whiteners (glm/whiten.py) on the engine.

The reference stores, per (a, b), the dense per-run inverse Cholesky factor
L^-1 of the correlation matrix and applies it by GEMM.  Here L^-1 is the
innovations recurrence of the plan's FP64 tables (the same coefficients the
search and the GLS use, censoring gaps and run starts included), applied by the
``ffs_reml_whiten_cols`` kernel; a ``RunWhitener`` is a handle on one pair of a
shared plan.  ``dense()`` materialises L^-1 for tests, as the reference's does.
"""

from __future__ import annotations

from collections.abc import Sequence

import torch

from .. import _lib
from .arma import run_bounds  # noqa: F401  (glm/whiten.py:27-38, re-exported)


class RunWhitener:
    """glm/whiten.py:64-154: L^-1 of one (a, b) pair; ``cols(M)`` = L^-1 M for
    M (T, k) or (T,), ``rows(Y)`` = Y L^-T for Y (V, T), ``dense()`` = L^-1."""

    def __init__(self, grid, k: int, n: int, dtype: torch.dtype = torch.float32):
        self._grid = grid
        self._k = int(k)
        self.n = int(n)
        self._dtype = dtype
        self._pair = torch.tensor([self._k], dtype=torch.int32, device=grid.device)

    @property
    def device(self) -> torch.device:
        return self._grid.device

    @property
    def dtype(self) -> torch.dtype:
        return self._dtype

    @property
    def nbytes(self) -> int:
        """Device bytes of this pair's recurrence coefficients."""
        return 4 * 8 * self.n

    def to(self, device=None, dtype=None) -> "RunWhitener":
        """A handle with another output dtype (the tables stay on the plan's device)."""
        return RunWhitener(self._grid, self._k, self.n, dtype or self._dtype)

    def _series(self, S: torch.Tensor) -> torch.Tensor:
        """L^-1 applied to each row of S (m, T); fp64 on the device."""
        S = S.to(device=self.device, dtype=torch.float64).contiguous()
        if S.shape[1] != self.n:
            raise ValueError(f"series have {S.shape[1]} samples, the whitener {self.n}")
        out = torch.empty_like(S)
        lib = _lib.load()
        _lib.check(lib.ffs_reml_whiten_cols(self._grid.handle, _lib.ptr(self._pair), 1, S.shape[0], _lib.ptr(S),
                                            self.n, _lib.ptr(out), _lib.stream_handle(self.device)), "whiten_cols")
        return out

    def cols(self, M: torch.Tensor) -> torch.Tensor:
        squeeze = M.ndim == 1
        Mm = M.unsqueeze(1) if squeeze else M
        out = self._series(Mm.T).T.to(device=M.device, dtype=M.dtype)
        return out.squeeze(1) if squeeze else out

    def rows(self, Y: torch.Tensor) -> torch.Tensor:
        return self._series(Y).to(device=Y.device, dtype=Y.dtype)

    def dense(self) -> torch.Tensor:
        return self.cols(torch.eye(self.n, dtype=torch.float64, device=self.device)).to(self._dtype)


def stack_cols(whiteners: Sequence[RunWhitener], M: torch.Tensor) -> torch.Tensor:
    """glm/whiten.py:156-186: (K, T, k) = L_k^-1 M for every whitener."""
    return torch.stack([w.cols(M) for w in whiteners])


def arma11_whiteners(a_grid, b_grid, n: int, *, run_starts=None, tau=None, device=None,
                     dtype: torch.dtype = torch.float32):
    """glm/whiten.py:197-271: one whitener per valid (a, b) of the grid (the
    validity filter of the covariance builder, no zero insertion), their
    log det R (K,) and the pair list."""
    from .arma import RemlGrid, _grid_tensor

    dev = _lib.require_device(device)
    a_grid, b_grid = _grid_tensor(a_grid, "a_grid"), _grid_tensor(b_grid, "b_grid")
    try:
        grid = RemlGrid(torch.ones(n, 1, dtype=torch.float64), a_grid, b_grid, dev, run_starts=run_starts, tau=tau,
                        ensure_zero=False)
    except ValueError as e:
        if "No valid" in str(e):
            return [], torch.empty(0, dtype=dtype), []
        raise
    ws = [RunWhitener(grid, k, n, dtype) for k in range(grid.K)]
    return ws, torch.as_tensor(grid.logdet_R, dtype=dtype), list(grid.param_list)


def arma11_whitener(a: float, b: float, n: int, *, run_starts=None, tau=None, device=None,
                    dtype: torch.dtype = torch.float32) -> RunWhitener:
    """glm/whiten.py:274-300: the whitener of one (a, b) (a and b taken in
    ``dtype``); ValueError when the pair is not a valid ARMA(1,1) correlation."""
    ws, _, plist = arma11_whiteners(torch.tensor([a], dtype=dtype), torch.tensor([b], dtype=dtype), n,
                                    run_starts=run_starts, tau=tau, device=device, dtype=dtype)
    if not plist:
        raise ValueError(f"Invalid ARMA(1,1) parameters: a={a}, b={b}")
    return ws[0]
