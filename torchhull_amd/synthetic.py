"""Seeded synthetic visual-hull scenes: unions of ellipsoids seen by pinhole cameras.

The camera conventions follow the reference's dataset generator
(data/generate_dataset.py): the OpenGL projection of `perspective` (:13-30),
the OpenCV intrinsics of `perspective_cv` (:33-50), the OpenCV view
V_cv = D V D with D = diag(1,-1,-1,1) (:110-115) and row-flipped OpenCV masks
(:240).  Masks are rendered by an exact per-pixel ray-ellipsoid test in
float64: pixel (y, x) of an OpenGL mask is tested at its centre,
NDC ((2x+1)/W - 1, (2y+1)/H - 1), with row 0 at NDC y = -1, the inverse of
the pixel mapping of image_utils.h:67-71.

Scene parameters are drawn from ``numpy.random.default_rng(seed)``; masks can
be rendered on any torch device (the bench renders on the GPU, tests on the
CPU).
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np
import torch


@dataclasses.dataclass
class Scene:
    """One visual-hull problem: masks, fused transforms and carving parameters."""

    masks: torch.Tensor  # [B,H,W,1] float32 in {0,1}
    transforms: torch.Tensor  # [B,4,4] float32 (projection @ view)
    level: int
    cube_corner_bfl: tuple[float, float, float]
    cube_length: float
    masks_partial: bool
    transforms_convention: str
    blur: tuple[int, float] | None  # (kernel_size, sigma) or None
    seed: int = 0

    @property
    def num_cameras(self) -> int:
        return int(self.masks.shape[0])


@dataclasses.dataclass
class SceneSpec:
    n_ellipsoids: int
    n_cameras: int
    resolution: int
    level: int
    masks_partial: bool
    convention: str
    blur: tuple[int, float] | None
    radius: float = 2.5
    fovy_deg: float = 40.0
    axes_range: tuple[float, float] = (0.15, 0.4)


# The five workloads (cube [-1,1]^3 throughout).
CONFIGS = {
    0: SceneSpec(3, 8, 256, 7, False, "opengl", None),
    1: SceneSpec(6, 64, 1024, 9, False, "opengl", (9, 2.0)),
    2: SceneSpec(12, 32, 2048, 10, True, "opencv", (15, 3.0), radius=1.6, fovy_deg=60.0, axes_range=(0.05, 0.5)),
}


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform rotation from a Shoemake quaternion (3 uniforms)."""
    u1, u2, u3 = rng.uniform(size=3)
    q = np.array([
        math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
        math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
        math.sqrt(u1) * math.sin(2 * math.pi * u3),
        math.sqrt(u1) * math.cos(2 * math.pi * u3),
    ])
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def perspective(fovy: float, aspect: float, near: float, far: float) -> np.ndarray:
    """OpenGL projection (data/generate_dataset.py:13-30)."""
    f = 1.0 / math.tan(fovy / 2.0)
    return np.array([
        [f / aspect, 0.0, 0.0, 0.0],
        [0.0, f, 0.0, 0.0],
        [0.0, 0.0, -(far + near) / (far - near), -(2.0 * far * near) / (far - near)],
        [0.0, 0.0, -1.0, 0.0],
    ])


def perspective_cv(fovy: float, aspect: float, height: int, width: int) -> np.ndarray:
    """OpenCV intrinsics as a 4x4 (data/generate_dataset.py:33-50)."""
    t = math.tan(fovy / 2.0)
    return np.array([
        [(width / 2.0) / (t * aspect), 0.0, width / 2.0, 0.0],
        [0.0, (height / 2.0) / t, height / 2.0, 0.0],
        [0.0, 0.0, 1.0, 0.0],
        [0.0, 0.0, 0.0, 1.0],
    ])


def look_at(eye: np.ndarray) -> np.ndarray:
    """OpenGL view matrix of a camera at `eye` looking at the origin, up = +y (fallback +z)."""
    fwd = -eye / np.linalg.norm(eye)
    up = np.array([0.0, 1.0, 0.0])
    if abs(float(np.dot(fwd, up))) > 0.999:
        up = np.array([0.0, 0.0, 1.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    true_up = np.cross(right, fwd)
    R = np.stack([right, true_up, -fwd])  # rows: camera axes in world coordinates
    V = np.eye(4)
    V[:3, :3] = R
    V[:3, 3] = -R @ eye
    return V


def fibonacci_cameras(n: int, radius: float) -> np.ndarray:
    k = np.arange(n, dtype=np.float64)
    y = 1.0 - 2.0 * (k + 0.5) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    phi = k * math.pi * (3.0 - math.sqrt(5.0))
    return radius * np.stack([r * np.cos(phi), y, r * np.sin(phi)], axis=1)


def render_masks_gl(views: np.ndarray, fovy: float, H: int, W: int, centers: np.ndarray, rotations: np.ndarray,
                    axes: np.ndarray, device: torch.device | str = "cpu") -> torch.Tensor:
    """Exact ray-ellipsoid masks [B,H,W,1] float32 in the OpenGL pixel convention."""
    device = torch.device(device)
    B = views.shape[0]
    t = math.tan(fovy / 2.0)
    aspect = W / H
    xs = (2.0 * torch.arange(W, dtype=torch.float64, device=device) + 1.0) / W - 1.0
    ys = (2.0 * torch.arange(H, dtype=torch.float64, device=device) + 1.0) / H - 1.0
    out = torch.empty((B, H, W, 1), dtype=torch.float32, device=device)
    for b in range(B):
        V = views[b]
        R = V[:3, :3]
        eye = -R.T @ V[:3, 3]
        # camera-space directions (x right, y up, looking down -z)
        dx = (xs * t * aspect).view(1, W).expand(H, W)
        dy = (ys * t).view(H, 1).expand(H, W)
        dirs_cam = torch.stack([dx, dy, -torch.ones_like(dx)], dim=-1)  # [H,W,3]
        Rt = torch.as_tensor(R.T, dtype=torch.float64, device=device)
        dirs = dirs_cam @ Rt.T  # world directions
        hit = torch.zeros((H, W), dtype=torch.bool, device=device)
        for c, Q, a in zip(centers, rotations, axes):
            # local frame: p' = diag(1/a) Q^T (p - c)
            M = torch.as_tensor(np.diag(1.0 / a) @ Q.T, dtype=torch.float64, device=device)
            o = M @ torch.as_tensor(eye - c, dtype=torch.float64, device=device)
            d = dirs @ M.T
            A = (d * d).sum(-1)
            Bq = 2.0 * (d @ o)
            C = float(o @ o) - 1.0
            disc = Bq * Bq - 4.0 * A * C
            ok = disc >= 0
            sq = torch.sqrt(torch.clamp(disc, min=0.0))
            t_far = (-Bq + sq) / (2.0 * A)
            hit |= ok & (t_far > 0)
        out[b, :, :, 0] = hit.to(torch.float32)
    return out


def make_scene(spec: SceneSpec, seed: int, device: torch.device | str = "cpu") -> Scene:
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-0.4, 0.4, size=(spec.n_ellipsoids, 3))
    axes = rng.uniform(spec.axes_range[0], spec.axes_range[1], size=(spec.n_ellipsoids, 3))
    rotations = [random_rotation(rng) for _ in range(spec.n_ellipsoids)]
    fovy = math.radians(spec.fovy_deg)
    H = W = spec.resolution
    eyes = fibonacci_cameras(spec.n_cameras, spec.radius)
    views = np.stack([look_at(e) for e in eyes])
    masks = render_masks_gl(views, fovy, H, W, centers, rotations, axes, device=device)
    if spec.convention == "opengl":
        P = perspective(fovy, W / H, 0.1, 1000.0)
        T = np.einsum("ij,bjk->bik", P, views)
    else:
        D = np.diag([1.0, -1.0, -1.0, 1.0])
        P = perspective_cv(fovy, W / H, H, W)
        T = np.einsum("ij,bjk->bik", P, np.einsum("ij,bjk,kl->bil", D, views, D))
        masks = masks.flip(1)
    transforms = torch.as_tensor(T.astype(np.float32), device=device)
    return Scene(masks=masks.contiguous(), transforms=transforms, level=spec.level, cube_corner_bfl=(-1.0, -1.0, -1.0),
                 cube_length=2.0, masks_partial=spec.masks_partial, transforms_convention=spec.convention,
                 blur=spec.blur, seed=seed)


def config_scene(config: int, scene_index: int = 0, device: torch.device | str = "cpu",
                 spec_override: dict | None = None) -> Scene:
    """Scene `scene_index` of workload `config` (0, 1 or 2), seed 1000*config + scene_index."""
    spec = dataclasses.replace(CONFIGS[config], **(spec_override or {}))
    return make_scene(spec, 1000 * config + scene_index, device=device)


def mixed_spec(rng: np.random.Generator, max_level: int = 9) -> SceneSpec:
    """One scene of the mixed multi-scene workload: 1..8 ellipsoids, 4..48 cameras,
    resolution in {256,512,1024}, level 6..max_level, partial masks with p=0.3, blur k7 sigma 1.5."""
    return SceneSpec(
        n_ellipsoids=int(rng.integers(1, 9)),
        n_cameras=int(rng.integers(4, 49)),
        resolution=int(rng.choice([256, 512, 1024])),
        level=int(rng.integers(6, max_level + 1)),
        masks_partial=bool(rng.uniform() < 0.3),
        convention="opengl",
        blur=(7, 1.5),
    )


def mixed_scenes(n: int = 64, seed: int = 3000, device: torch.device | str = "cpu", max_level: int = 9,
                 resolution_cap: int | None = None) -> list[Scene]:
    rng = np.random.default_rng(seed)
    scenes = []
    for i in range(n):
        spec = mixed_spec(rng, max_level=max_level)
        if resolution_cap is not None:
            spec = dataclasses.replace(spec, resolution=min(spec.resolution, resolution_cap))
        scenes.append(make_scene(spec, seed + 1 + i, device=device))
    return scenes
