"""Helpers shared by the GPU parity tests."""

from __future__ import annotations

import numpy as np
import torch

DEV = torch.device("cuda:0") if torch.cuda.is_available() else None


def is_closed_manifold(faces: np.ndarray) -> bool:
    """Every edge in exactly two faces (tests/test_visual_hull.py:22-29 of the reference)."""
    f = np.asarray(faces)
    if f.shape[0] == 0:
        return True
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e.sort(axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    return bool((counts == 2).all())


def assert_mesh_equal(v, f, ov, of, atol: float) -> None:
    """Faces bit-exact, vertices within atol (NaN positions must match)."""
    v = v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v
    f = f.detach().cpu().numpy() if isinstance(f, torch.Tensor) else f
    assert f.shape == of.shape, (f.shape, of.shape)
    assert v.shape == ov.shape, (v.shape, ov.shape)
    assert np.array_equal(f, of), "faces differ"
    nan = np.isnan(ov)
    assert np.array_equal(np.isnan(v), nan)
    d = np.abs(v[~nan] - ov[~nan])
    assert d.size == 0 or d.max() <= atol, f"max vertex deviation {d.max()}"
