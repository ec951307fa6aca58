"""Writes the golden fixtures of this directory from the synthetic generator and the CPU oracle.

The reference ships no golden vectors (its only asset, data/Armadillo.ply, is
absent and its dataset generator is unseeded), so these fixtures are
regression pins of the oracle: inputs (bit-packed binary masks, transforms,
parameters) with the oracle's outputs (blurred masks, candidate lists, sparse
field, mesh).  Run from the repository root: python tests/golden/make_golden.py
"""

from __future__ import annotations

import pathlib
import sys

import hashlib

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle  # noqa: E402
from torchhull_amd import synthetic  # noqa: E402

CASES = {
    "c0_l6": (0, {"level": 6}),
    "c2_small": (2, {"n_cameras": 6, "resolution": 128, "level": 5}),
}


def make(name: str) -> dict:
    cfg, over = CASES[name]
    sc = synthetic.config_scene(cfg, 0, spec_override=over)
    masks = sc.masks.numpy()[..., 0]
    m = masks
    blurred_sha = ""
    if sc.blur is not None:
        m = oracle.gaussian_blur_sparse(masks.astype(np.float64), sc.blur[0], sc.blur[1], np.float32)
        blurred_sha = hashlib.sha256(np.ascontiguousarray(m).tobytes()).hexdigest()
    T = sc.transforms.numpy()
    f = oracle.sparse_field(m, T, sc.level, sc.cube_corner_bfl, sc.cube_length, sc.masks_partial,
                            sc.transforms_convention)
    v, fc = oracle.visual_hull(None, None, sc.level, sc.cube_corner_bfl, sc.cube_length, sc.masks_partial,
                               sc.transforms_convention, True, field=f)
    return dict(
        masks_bits=np.packbits(masks.astype(np.uint8), axis=-1), mask_shape=np.array(masks.shape),
        transforms=T, level=np.array(sc.level), corner=np.array(sc.cube_corner_bfl, np.float32),
        length=np.array(sc.cube_length, np.float32), partial=np.array(sc.masks_partial),
        opengl=np.array(sc.transforms_convention == "opengl"),
        blur=np.array(sc.blur if sc.blur else (0, 0.0), np.float64), blurred_sha256=np.array(blurred_sha),
        cand_sizes=np.array([len(c) for c in f.candidates]), cand=np.concatenate(f.candidates),
        field_ids=f.ids, field_values=f.values, verts=v, faces=fc)


if __name__ == "__main__":
    for name in CASES:
        d = make(name)
        np.savez_compressed(pathlib.Path(__file__).parent / f"{name}.npz", **d)
        print(name, {k: v.shape for k, v in d.items()})
