"""Output files of the REML fit: the -Rvar variance stack, the -Rlklhd
likelihood surface and the stats-bucket sub-brick order, written as NIfTI-1
(``.nii`` / ``.nii.gz``) with AFNI's attribute extension for sub-brick labels
and statistic tags.

Layouts follow the reference:
  * -Rvar stack: ``save_arma_rvar`` / ``load_arma_params`` (glm/arma.py:6181-6454,
    bricks a, b, lam, StDev, -LogLik, LjungBox with LjungBox tagged chi-squared
    with ``ljung_box_dof`` degrees of freedom) and the CLI's stack
    (cli/reml.py:4990-5100, labels a, b, lambda, StDev, -LogLik, LjungBox);
  * -Rlklhd: one sub-brick per grid pair labelled ``a=0.00_b=0.30``
    (cli/reml.py:5102-5164);
  * bucket: Full_Fstat, then ``<name>#0_Coef`` / ``<name>#0_Tstat`` per regressor
    (glm/outputs.py:946-1088).

The file format is written here from the NIfTI-1 header layout (348-byte
header, ``n+1`` magic, extensions after the 4-byte extender, data from
``vox_offset`` with x fastest) and AFNI's NIML attribute group (ecode 4):
``BRICK_LABS`` (labels joined by ``~``), ``BRICK_STATAUX`` (sub-brick index,
statistic code, parameter count, parameters) and ``BRICK_STATSYM``.  Host-side
I/O only; nothing here runs on the GPU.
"""

from __future__ import annotations

import gzip
import os
import re
import struct
from pathlib import Path

import numpy as np
import torch

_ECODE_AFNI = 4
# AFNI statistic codes of the sub-brick types used here (afni.h FUNC_*_TYPE).
STAT_CODES = {"fico": 2, "fitt": 3, "fift": 4, "fizt": 5, "fict": 6}
_NIFTI_TYPES = {np.dtype(np.float32): (16, 32), np.dtype(np.float64): (64, 64), np.dtype(np.int16): (4, 16),
                np.dtype(np.int32): (8, 32), np.dtype(np.uint8): (2, 8)}
_NIFTI_DTYPES = {code: dt for dt, (code, _) in _NIFTI_TYPES.items()}


def stat_type_to_stataux(stat_type: str, params) -> tuple[int, tuple[float, ...]]:
    """``("fict", (dof,))`` -> ``(6, (dof,))``: the (code, params) pair of a
    statistic sub-brick (io/afni.py:3409-3427)."""
    key = stat_type.lower().lstrip("-")
    if key not in STAT_CODES:
        raise ValueError(f"unknown AFNI stat type {stat_type!r}; expected one of {sorted(STAT_CODES)}")
    return STAT_CODES[key], tuple(float(p) for p in params)


def _statsym(code: int, params: tuple[float, ...]) -> str:
    if code == 2 and len(params) == 3:
        return f"Correl({int(params[0])},{int(params[1])},{int(params[2])})"
    if code == 3 and len(params) == 1:
        return f"Ttest({int(params[0])})"
    if code == 4 and len(params) == 2:
        return f"Ftest({int(params[0])},{int(params[1])})"
    if code == 5:
        return "Zscore()"
    if code == 6 and len(params) == 1:
        return f"Chisq({int(params[0])})"
    return "none"


def _afni_extension(n_sub: int, labels, stataux) -> bytes:
    """NIML attribute group of the AFNI extension (labels and statistic tags)."""
    idcode = "XYZ_" + os.urandom(11).hex()[:22]
    parts = ["<?xml version='1.0' ?>\n<AFNI_attributes\n", f'  self_idcode="{idcode}"\n',
             '  ni_form="ni_group" >\n']
    if labels is not None:
        labs = "~".join(str(x) for x in labels).replace('"', "'")
        parts.append('<AFNI_atr\n  ni_type="String"\n  ni_dimen="1"\n  atr_name="BRICK_LABS" >\n'
                     f' "{labs}"\n</AFNI_atr>\n')
    if stataux:
        vals: list[float] = []
        for idx in sorted(stataux):
            code, params = stataux[idx]
            vals += [float(idx), float(code), float(len(params))] + [float(p) for p in params]
        syms = ";".join(_statsym(*stataux[i]) if i in stataux else "none" for i in range(n_sub))
        body = " " + "\n ".join(f"{v:g}" for v in vals)
        parts.append(f'<AFNI_atr\n  ni_type="float"\n  ni_dimen="{len(vals)}"\n  atr_name="BRICK_STATAUX" >\n'
                     f"{body}\n</AFNI_atr>\n")
        parts.append('<AFNI_atr\n  ni_type="String"\n  ni_dimen="1"\n  atr_name="BRICK_STATSYM" >\n'
                     f' "{syms}"\n</AFNI_atr>\n')
    parts.append("</AFNI_attributes>\n\x00")
    return "".join(parts).encode("utf-8")


def write_nifti(path, data: np.ndarray, affine=None, brick_labels=None, brick_stataux=None,
                tr: float | None = None) -> Path:
    """Write ``data`` (up to 7-D, indexed [x, y, z, t, ...]) as NIfTI-1 with
    an sform from ``affine`` (identity when None).  ``brick_labels`` /
    ``brick_stataux`` ({sub-brick: (code, params)}) go into the AFNI
    extension.  ``.gz`` paths are gzip-compressed."""
    path = Path(path)
    data = np.asarray(data)
    if data.dtype not in _NIFTI_TYPES:
        data = data.astype(np.float32)
    if data.ndim < 1 or data.ndim > 7:
        raise ValueError(f"NIfTI data must have 1..7 dimensions, got {data.ndim}")
    code, bitpix = _NIFTI_TYPES[data.dtype]
    aff = np.eye(4) if affine is None else np.asarray(affine, dtype=np.float64)
    if aff.shape != (4, 4):
        raise ValueError("affine must be 4 x 4")
    n_sub = data.shape[3] if data.ndim >= 4 else 1
    ext = b""
    if brick_labels is not None or brick_stataux:
        if brick_labels is not None and len(brick_labels) != n_sub:
            raise ValueError(f"{len(brick_labels)} labels for {n_sub} sub-bricks")
        payload = _afni_extension(n_sub, brick_labels, brick_stataux)
        esize = (8 + len(payload) + 15) // 16 * 16
        ext = struct.pack("<ii", esize, _ECODE_AFNI) + payload + b"\x00" * (esize - 8 - len(payload))
    vox_offset = 352 + len(ext)
    hdr = bytearray(348)
    struct.pack_into("<i", hdr, 0, 348)
    hdr[38] = ord("r")
    dim = [data.ndim] + list(data.shape) + [1] * (7 - data.ndim)
    struct.pack_into("<8h", hdr, 40, *dim)
    struct.pack_into("<hhh", hdr, 70, code, bitpix, 0)
    vox = np.sqrt((aff[:3, :3] ** 2).sum(axis=0))
    pixdim = [1.0, *vox.tolist(), float(tr) if tr else 1.0, 1.0, 1.0, 1.0]
    struct.pack_into("<8f", hdr, 76, *pixdim)
    struct.pack_into("<fff", hdr, 108, float(vox_offset), 1.0, 0.0)
    hdr[123] = 2 | 8  # mm, s
    struct.pack_into("<hh", hdr, 252, 0, 1)  # qform_code, sform_code (scanner)
    struct.pack_into("<4f", hdr, 280, *aff[0].tolist())
    struct.pack_into("<4f", hdr, 296, *aff[1].tolist())
    struct.pack_into("<4f", hdr, 312, *aff[2].tolist())
    hdr[344:348] = b"n+1\x00"
    blob = bytes(hdr) + (b"\x01\x00\x00\x00" if ext else b"\x00\x00\x00\x00") + ext
    body = np.asarray(data).tobytes(order="F")
    path.parent.mkdir(parents=True, exist_ok=True)
    opener = gzip.open if path.suffix == ".gz" else open
    with opener(path, "wb") as f:
        f.write(blob)
        f.write(body)
    return path


class NiftiImage:
    """What ``read_nifti`` returns: data (x fastest axes order restored),
    affine (sform), sub-brick labels and statistic tags from the AFNI
    extension (empty when absent)."""

    def __init__(self, data, affine, labels, stataux):
        self.data, self.affine, self.labels, self.stataux = data, affine, labels, stataux


def read_nifti(path) -> NiftiImage:
    path = Path(path)
    opener = gzip.open if path.suffix == ".gz" else open
    with opener(path, "rb") as f:
        raw = f.read()
    if len(raw) < 352 or struct.unpack_from("<i", raw, 0)[0] != 348 or raw[344:347] not in (b"n+1", b"ni1"):
        raise ValueError(f"{path}: not a little-endian NIfTI-1 file")
    dim = struct.unpack_from("<8h", raw, 40)
    code = struct.unpack_from("<h", raw, 70)[0]
    if code not in _NIFTI_DTYPES:
        raise ValueError(f"{path}: unsupported NIfTI datatype {code}")
    vox_offset = int(struct.unpack_from("<f", raw, 108)[0])
    slope, inter = struct.unpack_from("<ff", raw, 112)
    shape = tuple(int(x) for x in dim[1:1 + dim[0]])
    dt = _NIFTI_DTYPES[code]
    n = int(np.prod(shape))
    data = np.frombuffer(raw, dtype=dt, count=n, offset=vox_offset).reshape(shape, order="F").copy()
    if slope not in (0.0, 1.0) or inter != 0.0:
        data = data.astype(np.float32) * (slope if slope else 1.0) + inter
    aff = np.eye(4)
    aff[0] = struct.unpack_from("<4f", raw, 280)
    aff[1] = struct.unpack_from("<4f", raw, 296)
    aff[2] = struct.unpack_from("<4f", raw, 312)
    labels, stataux = [], {}
    o = 352
    if raw[348] == 1:
        while o + 8 <= vox_offset:
            esize, ecode = struct.unpack_from("<ii", raw, o)
            if esize < 8:
                break
            if ecode == _ECODE_AFNI:
                xml = raw[o + 8:o + esize].decode("utf-8", errors="replace")
                m = re.search(r'atr_name="BRICK_LABS"\s*>\s*"([^"]*)"', xml)
                if m:
                    labels = m.group(1).split("~")
                m = re.search(r'atr_name="BRICK_STATAUX"\s*>([^<]*)<', xml)
                if m:
                    vals = [float(v) for v in m.group(1).split()]
                    i = 0
                    while i + 3 <= len(vals):
                        idx, c, npar = int(vals[i]), int(vals[i + 1]), int(vals[i + 2])
                        stataux[idx] = (c, tuple(vals[i + 3:i + 3 + npar]))
                        i += 3 + npar
            o += esize
    return NiftiImage(data, aff, labels, stataux)


def _np(x):
    if x is None:
        return None
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def _to_volume(values: np.ndarray, volume_shape, voxel_mask) -> np.ndarray:
    """(V,) or (V, n) per-voxel values -> volume (x, y, z[, n]); masked voxels
    are scattered into zeros, no shape keeps them as (V, 1, 1[, n])."""
    values = np.asarray(values, dtype=np.float32)
    tail = values.shape[1:]
    if volume_shape is None:
        return values.reshape(values.shape[0], 1, 1, *tail)
    volume_shape = tuple(int(s) for s in volume_shape)
    if voxel_mask is not None:
        mask = _np(voxel_mask).reshape(-1).astype(bool)
        full = np.zeros((int(np.prod(volume_shape)), *tail), dtype=np.float32)
        full[mask] = values
        return full.reshape(*volume_shape, *tail)
    return values.reshape(*volume_shape, *tail)


def rvar_stack(results, lam_label: str = "lambda"):
    """(V, n) per-voxel -Rvar bricks, their labels and statistic tags, in the
    CLI's order (cli/reml.py:5004-5043): a, b, lambda, StDev, -LogLik and, when
    present, LjungBox tagged chi-squared with ``ljung_box_dof`` dof."""
    cols, labels, stataux = [], [], None
    if results.arma_params is not None:
        ab = _np(results.arma_params)
        cols += [ab[:, 0], ab[:, 1]]
        labels += ["a", "b"]
    if results.arma_lambda is not None:
        cols.append(_np(results.arma_lambda))
        labels.append(lam_label)
    if results.sigma2 is not None:
        cols.append(np.sqrt(_np(results.sigma2)))
        labels.append("StDev")
    if results.reml_likelihood is not None:
        cols.append(_np(results.reml_likelihood))
        labels.append("-LogLik")
    if results.ljung_box is not None:
        cols.append(_np(results.ljung_box))
        labels.append("LjungBox")
        if results.ljung_box_dof:
            stataux = {len(labels) - 1: stat_type_to_stataux("fict", (float(results.ljung_box_dof),))}
    return np.stack([np.asarray(c, dtype=np.float32) for c in cols], axis=1), labels, stataux


def save_arma_rvar(results, output_path, volume_shape=None, voxel_mask=None, affine=None, max_lag=None,
                   tau=None) -> Path:
    """glm/arma.py:6181-6377: 4-D file with bricks a, b, lam, StDev, -LogLik,
    LjungBox (zeros, with a warning, when no Ljung-Box statistic was computed)."""
    import warnings

    if results.arma_params is None or results.arma_lambda is None or results.reml_likelihood is None:
        raise ValueError("results carry no ARMA parameters / likelihoods to write")
    if results.ljung_box is None:
        if results.residuals_whitened is not None:
            from .glm.arma import ljung_box_device

            rw = torch.as_tensor(results.residuals_whitened)
            lb = ljung_box_device(rw.cuda() if torch.cuda.is_available() else rw, max_lag=max_lag, tau=tau)
            ljung = _np(lb)
        else:
            warnings.warn("Rvar sub-brick 5 (LjungBox) written as zeros: no whiteness statistic was "
                          "computed. Pass want_ljung_box=True to fit_glm_arma11() to populate it.", stacklevel=2)
            ljung = np.zeros(len(_np(results.arma_params)), dtype=np.float32)
    else:
        ljung = _np(results.ljung_box)
    ab = _np(results.arma_params)
    cols = [ab[:, 0], ab[:, 1], _np(results.arma_lambda), np.sqrt(_np(results.sigma2)),
            _np(results.reml_likelihood), ljung]
    stack = np.stack([np.asarray(c, dtype=np.float32) for c in cols], axis=1)
    vol = _to_volume(stack, volume_shape, voxel_mask)
    if affine is None:
        affine = getattr(results, "affine", None)
    stataux = None
    if results.ljung_box_dof:
        stataux = {5: stat_type_to_stataux("fict", (float(results.ljung_box_dof),))}
    return write_nifti(output_path, vol, affine=affine,
                       brick_labels=["a", "b", "lam", "StDev", "-LogLik", "LjungBox"], brick_stataux=stataux)


def load_arma_params(filepath, voxel_mask=None) -> np.ndarray:
    """glm/arma.py:6380-6454: (n_voxels, 2) float32 (a, b) from bricks 0 and 1
    of a -Rvar file, restricted to ``voxel_mask`` when given."""
    filepath = Path(filepath)
    if not filepath.exists():
        raise FileNotFoundError(f"ARMA params file not found: {filepath}")
    data = read_nifti(filepath).data.astype(np.float32)
    if data.ndim != 4 or data.shape[-1] < 2:
        raise ValueError(f"Expected 4D NIfTI with at least 2 volumes, got shape {data.shape}")
    a = data[..., 0].reshape(-1)
    b = data[..., 1].reshape(-1)
    if voxel_mask is not None:
        m = _np(voxel_mask).reshape(-1).astype(bool)
        a, b = a[m], b[m]
    return np.stack([a, b], axis=1)


def save_reml_likelihood_surface(results, output_path, volume_shape=None, voxel_mask=None,
                                 affine=None) -> Path:
    """cli/reml.py:5102-5164: one sub-brick per grid pair (the surface of
    ``save_profile_likelihoods=True``), labelled ``a=0.00_b=0.30``."""
    if results.reml_lklhd_surface is None or results.reml_surface_params is None:
        raise ValueError("results carry no likelihood surface (fit with save_profile_likelihoods=True)")
    surf = _np(results.reml_lklhd_surface).astype(np.float32)
    labels = [f"a={float(a):.2f}_b={float(b):.2f}" for a, b in results.reml_surface_params]
    vol = _to_volume(surf, volume_shape, voxel_mask)
    return write_nifti(output_path, vol, affine=affine if affine is not None else results.affine,
                       brick_labels=labels)


def bucket_subbricks(results, condition_names=None, volume_shape=None, voxel_mask=None, dtype=np.float32):
    """glm/outputs.py:946-1021: the stats-bucket sub-bricks in AFNI order —
    Full_Fstat (when present), then per regressor ``<name>#0_Coef`` and, with
    t-statistics, ``<name>#0_Tstat`` (names already ending in ``#N`` get
    ``_Coef`` / ``_Tstat`` only).  Returns (sub-bricks, labels)."""
    betas = _np(results.betas)
    n_reg = betas.shape[1] if betas.ndim == 2 else 1
    if condition_names is None:
        condition_names = [f"cond{i + 1:02d}" for i in range(n_reg)]
    elif len(condition_names) != n_reg:
        raise ValueError(f"condition_names has length {len(condition_names)} but results have {n_reg} regressors")
    bvol = _to_volume(betas.reshape(betas.shape[0], -1), volume_shape, voxel_mask)
    tvol = None
    if results.tstats is not None:
        t = _np(results.tstats)
        tvol = _to_volume(t.reshape(t.shape[0], -1), volume_shape, voxel_mask)
    subbricks, labels = [], []
    if results.fstats is not None:
        subbricks.append(_to_volume(_np(results.fstats), volume_shape, voxel_mask).astype(dtype))
        labels.append("Full_Fstat")
    indexed = re.compile(r"#\d+$")
    for i, name in enumerate(condition_names):
        coef, tst = (f"{name}_Coef", f"{name}_Tstat") if indexed.search(name) else (f"{name}#0_Coef",
                                                                                  f"{name}#0_Tstat")
        subbricks.append(bvol[..., i].astype(dtype))
        labels.append(coef)
        if tvol is not None:
            subbricks.append(tvol[..., i].astype(dtype))
            labels.append(tst)
    return subbricks, labels


def save_bucket(results, output_path, condition_names=None, volume_shape=None, voxel_mask=None,
                affine=None) -> Path:
    """Stats bucket (bucket_subbricks order) with t sub-bricks tagged Student t
    (dof = results.dof) and Full_Fstat tagged F (p, dof)."""
    subbricks, labels = bucket_subbricks(results, condition_names, volume_shape, voxel_mask)
    stataux = {}
    for i, lab in enumerate(labels):
        if lab.endswith("_Tstat") and results.dof:
            stataux[i] = stat_type_to_stataux("fitt", (float(results.dof),))
        elif lab == "Full_Fstat" and results.dof:
            cols = results.fitted_column_indices
            n_f = len(cols) if cols is not None else (results.n_regressors_full or 1)
            stataux[i] = stat_type_to_stataux("fift", (float(n_f), float(results.dof)))
    return write_nifti(output_path, np.stack(subbricks, axis=-1), affine=affine if affine is not None
                       else results.affine, brick_labels=labels, brick_stataux=stataux or None)
