"""Output files (CPU): the NIfTI-1 writer / reader against the format's fixed
fields, AFNI attribute round trips, and the -Rvar / -Rlklhd / bucket layouts of
the reference (glm/arma.py:6181-6454, cli/reml.py:4990-5164,
glm/outputs.py:946-1088) on a hand-filled results object."""

import gzip
import struct

import numpy as np
import pytest
import torch

from fastfuncstuff_amd import outputs as O
from fastfuncstuff_amd.glm.arma import ARMA11Results


def _results(V=12, p=3, seed=0):
    rng = np.random.default_rng(seed)
    r = ARMA11Results()
    r.betas = torch.from_numpy(rng.standard_normal((V, p)).astype(np.float32))
    r.tstats = torch.from_numpy(rng.standard_normal((V, p)).astype(np.float32))
    r.fstats = torch.from_numpy(rng.random(V).astype(np.float32))
    ab = rng.uniform(-0.5, 0.8, (V, 2)).astype(np.float32)
    r.arma_params = torch.from_numpy(ab)
    a, b = ab[:, 0].astype(np.float64), ab[:, 1].astype(np.float64)
    r.arma_lambda = torch.from_numpy(((b + a) * (1 + a * b) / (1 + 2 * a * b + b * b + 1e-10)).astype(np.float32))
    r.sigma2 = torch.from_numpy(rng.uniform(0.5, 2, V).astype(np.float32))
    r.reml_likelihood = torch.from_numpy(rng.uniform(100, 200, V).astype(np.float32))
    r.ljung_box = torch.from_numpy(rng.uniform(0, 30, V).astype(np.float32))
    r.ljung_box_dof = 7
    r.dof = 93
    r.n_regressors_full = p
    return r


def test_nifti_header_and_round_trip(tmp_path):
    data = np.arange(2 * 3 * 4 * 5, dtype=np.float32).reshape(2, 3, 4, 5)
    aff = np.diag([2.0, 3.0, 4.0, 1.0])
    aff[:3, 3] = [-10, 5, 7]
    f = O.write_nifti(tmp_path / "x.nii", data, affine=aff, brick_labels=list("abcde"),
                      brick_stataux={4: O.stat_type_to_stataux("fict", (7,))})
    raw = f.read_bytes()
    assert struct.unpack_from("<i", raw, 0)[0] == 348 and raw[344:348] == b"n+1\x00"
    assert struct.unpack_from("<8h", raw, 40)[:5] == (4, 2, 3, 4, 5)
    assert struct.unpack_from("<hh", raw, 70) == (16, 32)  # float32
    vox_offset = int(struct.unpack_from("<f", raw, 108)[0])
    assert vox_offset % 16 == 0 and vox_offset >= 352 and raw[348] == 1
    esize, ecode = struct.unpack_from("<ii", raw, 352)
    assert ecode == 4 and esize % 16 == 0 and 352 + esize == vox_offset
    # data are stored x fastest
    first = np.frombuffer(raw, dtype=np.float32, count=3, offset=vox_offset)
    assert first.tolist() == [data[0, 0, 0, 0], data[1, 0, 0, 0], data[0, 1, 0, 0]]
    img = O.read_nifti(f)
    assert np.array_equal(img.data, data)
    assert np.allclose(img.affine, aff)
    assert img.labels == list("abcde")
    assert img.stataux == {4: (6, (7.0,))}
    g = O.write_nifti(tmp_path / "x.nii.gz", data)
    assert gzip.open(g).read()[:4] == struct.pack("<i", 348)
    assert np.array_equal(O.read_nifti(g).data, data) and O.read_nifti(g).labels == []


def test_rvar_round_trip_with_mask(tmp_path):
    r = _results(V=12)
    mask = np.zeros((3, 2, 4), dtype=bool)
    mask.reshape(-1)[[0, 2, 3, 5, 7, 8, 11, 12, 15, 18, 20, 23]] = True
    f = O.save_arma_rvar(r, tmp_path / "rvar.nii.gz", volume_shape=(3, 2, 4), voxel_mask=mask)
    img = O.read_nifti(f)
    assert img.data.shape == (3, 2, 4, 6)
    assert img.labels == ["a", "b", "lam", "StDev", "-LogLik", "LjungBox"]
    assert img.stataux == {5: (6, (7.0,))}
    got = img.data.reshape(-1, 6)[mask.reshape(-1)]
    want = np.stack([r.arma_params[:, 0].numpy(), r.arma_params[:, 1].numpy(), r.arma_lambda.numpy(),
                     np.sqrt(r.sigma2.numpy()), r.reml_likelihood.numpy(), r.ljung_box.numpy()], 1)
    assert np.array_equal(got, want)
    assert not img.data.reshape(-1, 6)[~mask.reshape(-1)].any()
    ab = O.load_arma_params(f, mask)
    assert np.array_equal(ab, r.arma_params.numpy())
    assert O.load_arma_params(f).shape == (24, 2)
    with pytest.raises(FileNotFoundError):
        O.load_arma_params(tmp_path / "missing.nii")


def test_rvar_zero_ljung_box_warns(tmp_path):
    r = _results(V=5)
    r.ljung_box, r.ljung_box_dof = None, None
    with pytest.warns(UserWarning, match="LjungBox"):
        f = O.save_arma_rvar(r, tmp_path / "r.nii")
    img = O.read_nifti(f)
    assert img.data.shape == (5, 1, 1, 6) and not img.data[..., 5].any() and img.stataux == {}


def test_cli_rvar_stack_order():
    r = _results(V=4)
    stack, labels, stataux = O.rvar_stack(r)
    assert labels == ["a", "b", "lambda", "StDev", "-LogLik", "LjungBox"]
    assert stataux == {5: (6, (7.0,))}
    assert np.array_equal(stack[:, 3], np.sqrt(r.sigma2.numpy()))


def test_likelihood_surface_labels(tmp_path):
    r = _results(V=6)
    params = [(0.0, 0.3), (-0.25, 0.05), (0.8, -0.8)]
    r.reml_surface_params = params
    r.reml_lklhd_surface = torch.arange(18, dtype=torch.float32).reshape(6, 3)
    img = O.read_nifti(O.save_reml_likelihood_surface(r, tmp_path / "lk.nii", volume_shape=(3, 2, 1)))
    assert img.labels == ["a=0.00_b=0.30", "a=-0.25_b=0.05", "a=0.80_b=-0.80"]
    assert np.array_equal(img.data.reshape(6, 3), r.reml_lklhd_surface.numpy())


def test_bucket_order(tmp_path):
    r = _results(V=6, p=3)
    subs, labels = O.bucket_subbricks(r, ["faces", "houses#1", "motion"])
    assert labels == ["Full_Fstat", "faces#0_Coef", "faces#0_Tstat", "houses#1_Coef", "houses#1_Tstat",
                      "motion#0_Coef", "motion#0_Tstat"]
    assert np.array_equal(subs[3].reshape(-1), r.betas[:, 1].numpy())
    _, default = O.bucket_subbricks(r)
    assert default[1] == "cond01#0_Coef"
    with pytest.raises(ValueError):
        O.bucket_subbricks(r, ["only_one"])
    img = O.read_nifti(O.save_bucket(r, tmp_path / "b.nii.gz", ["faces", "houses#1", "motion"]))
    assert img.labels == labels
    assert img.stataux[0] == (4, (3.0, 93.0)) and img.stataux[2] == (3, (93.0,))
