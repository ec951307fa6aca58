"""Golden fixtures (tests/golden/*.npz): the oracle reproduces them exactly on the CPU, the HIP path on the GPU."""

from __future__ import annotations

import hashlib
import pathlib

import numpy as np
import pytest
import torch

from oracle import oracle
from tests.gpu_util import DEV, assert_mesh_equal

GOLDEN = pathlib.Path(__file__).parent / "golden"
NAMES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    d = dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))
    shape = tuple(d["mask_shape"])
    d["masks"] = np.unpackbits(d["masks_bits"], axis=-1)[..., : shape[2]].astype(np.float32)
    d["convention"] = "opengl" if bool(d["opengl"]) else "opencv"
    return d


def blurred_masks(d):
    k, s = int(d["blur"][0]), float(d["blur"][1])
    if k == 0:
        return d["masks"]
    m = oracle.gaussian_blur_sparse(d["masks"].astype(np.float64), k, s, np.float32)
    assert hashlib.sha256(np.ascontiguousarray(m).tobytes()).hexdigest() == str(d["blurred_sha256"])
    return m


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_golden(name):
    d = load(name)
    m = blurred_masks(d)
    f = oracle.sparse_field(m, d["transforms"], int(d["level"]), d["corner"], float(d["length"]), bool(d["partial"]),
                            d["convention"])
    assert np.array_equal([len(c) for c in f.candidates], d["cand_sizes"])
    assert np.array_equal(np.concatenate(f.candidates), d["cand"])
    assert np.array_equal(f.ids, d["field_ids"])
    assert np.array_equal(f.values.view(np.uint32), d["field_values"].view(np.uint32))
    v, fc = oracle.visual_hull(None, None, int(d["level"]), d["corner"], float(d["length"]), bool(d["partial"]),
                               d["convention"], True, field=f)
    assert np.array_equal(fc, d["faces"]) and np.array_equal(v, d["verts"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_golden(name):
    import torchhull_amd as th

    d = load(name)
    masks = torch.from_numpy(d["masks"][..., None]).to(DEV)
    k, s = int(d["blur"][0]), float(d["blur"][1])
    if k:
        masks = th.gaussian_blur(masks, k, s, sparse=True, dtype=torch.float32)
        sha = hashlib.sha256(np.ascontiguousarray(masks[..., 0].cpu().numpy()).tobytes()).hexdigest()
        assert sha == str(d["blurred_sha256"])
    T = torch.from_numpy(d["transforms"]).to(DEV)
    v, f, cand = th.visual_hull_with_candidate_voxels(masks, T, int(d["level"]), d["corner"].tolist(),
                                                      float(d["length"]), bool(d["partial"]), d["convention"], True)
    assert np.array_equal(torch.cat(cand).cpu().numpy(), d["cand"])
    assert_mesh_equal(v, f, d["verts"], d["faces"], atol=1e-6 * float(d["length"]))
