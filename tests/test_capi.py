"""The C-ABI library loads, exports every entry point of include/torchhull_amd.h, and reports errors (no GPU needed)."""

from __future__ import annotations

import ctypes
import re

import pytest

from torchhull_amd import _lib


def declared_functions() -> list[str]:
    text = _lib.HEADER_PATH.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(thull_\w+)\s*\(", text, re.M)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("thull_visual_hull_run", "thull_visual_hull_fetch_mesh", "thull_gaussian_blur",
              "thull_marching_cubes_dense_run", "thull_marching_cubes_sparse_run", "thull_integral_image"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert L.thull_abi_version() == 1


def test_errors_are_reported_not_raised_across_the_abi():
    L = _lib.lib()
    # null arguments are rejected with a message, whatever the device situation
    assert L.thull_visual_hull_run(None, None, 0, 1, None, None) != 0
    assert b"null" in L.thull_last_error()
    assert L.thull_ctx_create(0, None) != 0
    assert L.thull_last_error() != b""


def test_api_rejects_cpu_tensors_like_the_reference():
    import torch

    import torchhull_amd as th

    m = torch.zeros((2, 8, 8, 1))
    T = torch.eye(4).repeat(2, 1, 1)
    with pytest.raises(RuntimeError, match="No backend implementation available"):
        th.visual_hull(m, T, 3, (-1, -1, -1), 2.0, False, "opengl")
    with pytest.raises(RuntimeError, match="No backend implementation available"):
        th.gaussian_blur(m, 3, 1.0)
    with pytest.raises(RuntimeError, match="No backend implementation available"):
        th.marching_cubes(torch.zeros((1, 4, 4, 4)), 0.0)


def test_serving_footprint_and_device_checks():
    import torch

    from torchhull_amd import serving

    r = serving.HullRequest(masks=torch.zeros(8, 256, 256, 1), transforms=torch.zeros(8, 4, 4), level=7,
                            cube_corner_bfl=(-1.0, -1.0, -1.0), cube_length=2.0, masks_partial=False,
                            transforms_convention="opengl", blur=(7, 1.5))
    fp = serving.footprint_bytes(r)
    # blur output + padded integral images alone
    assert fp > 8 * 256 * 256 * 4 * 2 + 8 * 257 * 257 * 4
    r10 = dataclasses_replace(r, level=10)
    assert serving.footprint_bytes(r10) > 32 * fp
    with pytest.raises(RuntimeError):
        serving.HullServer("cpu")


def dataclasses_replace(obj, **kw):
    import dataclasses

    return dataclasses.replace(obj, **kw)
