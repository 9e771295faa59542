"""GPU tests of the hierarchical descent on demand (ffs_reml_hier_search).

The descent of the reference's GPU default selector (glm/arma.py:2224-2322)
is run on pair sets that hold only the lattice points each level can visit.
Their scores equal the full-grid surface bitwise, so the bar is exact equality
of the selected pair index and of its float32 score with
``select(surface(prep(Y)), hierarchical=True)`` (itself checked against the
oracle's hierarchical selector in test_gpu_parity), and of the scales with
prep's.  Cases: each projection mode, the censored config-2 timeline, lattices
whose zero point is not on the top level's points, a non-square lattice, a
lattice small enough for a single pass (top level and the level below it), and
designs wider than 32 columns (p = 40 in fp32 and fp16x3, p = 68 in bf16x3: no
stored residualisation coefficients, every pass residualises from Y) and a
130-column design (pairs over 64-column segments) in fp32 and fp16x3.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from fastfuncstuff_amd import _lib

    _lib.load()
    return torch.device("cuda", 0)


def _grid_case(name):
    from fastfuncstuff_amd import synth

    if name == "config2":
        wl = synth.gpu_config(V=2500, T=400, seed=2, noise_seed=17)
        return wl.X, wl.Y, wl.a_grid, wl.b_grid, dict(run_starts=wl.run_starts, tau=wl.tau)
    wl = synth.plain_config(V=3001, T=200, seed=3, noise_seed=21)
    X = wl.X
    if name == "offcentre":  # 21 x 13 lattice, zero point off the top level's points
        ag = np.linspace(-0.55, 0.85, 21).astype(np.float32)
        bg = np.linspace(-0.6, 0.6, 13).astype(np.float32)
    elif name == "small":  # 5 x 6 lattice: top level 1, one pass
        ag = np.linspace(-0.6, 0.6, 5).astype(np.float32)
        bg = np.linspace(-0.5, 0.5, 6).astype(np.float32)
    elif name in ("wide", "wide68", "seg130"):  # p = 40 / 68 > 32; 130 > 96 (segmented pairs)
        rng = np.random.default_rng(0)
        extra = {"wide": 34, "wide68": 62, "seg130": 124}[name]
        X = np.concatenate([wl.X, rng.standard_normal((wl.X.shape[0], extra))], 1)
        ag, bg = wl.a_grid, wl.b_grid
    else:
        ag, bg = wl.a_grid, wl.b_grid
    return X, wl.Y, ag, bg, {}


@pytest.mark.parametrize("name,mode", [("plain", "fp16x3"), ("plain", "bf16x3"), ("plain", "fp32"),
                                       ("config2", "fp16x3"), ("offcentre", "fp16x3"), ("small", "fp16x3"),
                                       ("wide", "fp32"), ("wide", "fp16x3"), ("wide68", "bf16x3"),
                                       ("seg130", "fp32"), ("seg130", "fp16x3")])
def test_hier_search_equals_surface_descent(dev, name, mode):
    from fastfuncstuff_amd.glm.arma import RemlGrid

    print(f"[hier] {name} {mode}", flush=True)
    X, Y, ag, bg, kw = _grid_case(name)
    grid = RemlGrid(torch.from_numpy(np.ascontiguousarray(X)), torch.from_numpy(ag), torch.from_numpy(bg), dev, **kw)
    grid.set_projection(mode)
    Yd = torch.from_numpy(Y).to(dev)
    V = Y.shape[0]
    Yt, scale, _ = grid.prep(Yd, normalize=True)
    S = grid.surface(Yt, V)
    b_ref, s_ref = grid.select(S, hierarchical=True)
    best, score, sc = grid.hier_search(Yd)
    torch.cuda.synchronize()
    assert torch.equal(sc, scale)
    nd = int((best != b_ref).sum())
    assert nd == 0, f"{nd} of {V} voxels select another pair than the surface descent"
    assert torch.equal(score, s_ref)
    # a second call on the same plan (tables built once) gives the same result
    best2, score2, _ = grid.hier_search(Yd[: V // 2])
    assert torch.equal(best2, b_ref[: V // 2]) and torch.equal(score2, s_ref[: V // 2])


def test_fit_uses_descent_with_the_surface_results(dev):
    """fit_glm_arma11's default (hierarchical) selector runs the descent on
    demand; every output equals the fit that computes the full surface."""
    from fastfuncstuff_amd import synth
    from fastfuncstuff_amd.glm.arma import fit_glm_arma11

    print("[hier] fit_glm_arma11 descent vs surface", flush=True)
    wl = synth.gpu_config(V=1500, T=400, seed=2, noise_seed=23)
    kw = dict(a_grid=wl.a_grid, b_grid=wl.b_grid, run_starts=wl.run_starts, tau=wl.tau, device=dev)
    r_d = fit_glm_arma11(wl.Y, wl.X, 2.0, **kw)
    r_s = fit_glm_arma11(wl.Y, wl.X, 2.0, hierarchical_search="surface", **kw)
    assert torch.equal(r_d.pair_index, r_s.pair_index)
    for name in ("arma_params", "reml_likelihood", "betas", "tstats", "fstats", "sigma2", "r2"):
        assert torch.equal(getattr(r_d, name), getattr(r_s, name)), name
    with pytest.raises(ValueError):
        fit_glm_arma11(wl.Y[:8], wl.X, 2.0, hierarchical_search="grid", **kw)
