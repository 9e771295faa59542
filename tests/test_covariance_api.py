"""Host-side covariance builders of the Python API (glm/arma.py:732-1103)
against the oracle's dense restatement: same pairs, same order, same R."""

import torch

from fastfuncstuff_amd.glm.arma import build_arma11_covariance, build_arma11_covariance_batch, valid_grid_pairs
from oracle import reml_oracle as O


def test_covariance_batch_matches_oracle_with_runs_and_censoring():
    g = torch.linspace(-0.8, 0.8, 9)
    tau = torch.tensor(list(range(12)) + list(range(14, 32)))
    R, P, plist = build_arma11_covariance_batch(g, g, 30, dtype=torch.float64, run_starts=[0, 12], tau=tau)
    assert R.shape[0] == len(plist) == valid_grid_pairs(g.double(), g.double()).shape[0]
    for k in (0, 7, len(plist) - 1):
        a, b = plist[k]
        torch.testing.assert_close(R[k], O.covariance(a, b, 30, run_starts=[0, 12], tau=tau), rtol=0, atol=0)
    assert torch.equal(P, valid_grid_pairs(g.double(), g.double()))


def test_single_covariance_and_invalid_pair():
    R = build_arma11_covariance(0.5, 0.2, 12, dtype=torch.float64)
    torch.testing.assert_close(R, O.covariance(0.5, 0.2, 12), rtol=0, atol=0)
    assert build_arma11_covariance(0.2, -0.6, 12) is None  # lambda < 0


def test_afni_faithful_mode_is_refused():
    """glm/arma.py:85-92: the AFNI-faithful mode (banded R) is not built, so
    switching it on raises, and a search with the flag set by hand raises too
    before any device work."""
    import numpy as np
    import pytest

    from fastfuncstuff_amd.glm import arma

    arma.set_afni_mode(False)
    with pytest.raises(NotImplementedError):
        arma.set_afni_mode(True)
    assert arma._REML_MODE.afni_faithful is False
    arma._REML_MODE.afni_faithful = True
    try:
        g = torch.linspace(-0.8, 0.8, 5)
        with pytest.raises(NotImplementedError):
            arma.RemlGrid(np.ones((20, 1)), g, g, device=None)
    finally:
        arma._REML_MODE.afni_faithful = False
