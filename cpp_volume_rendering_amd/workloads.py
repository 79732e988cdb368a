"""Seeded synthetic workloads: volume, transfer function, camera and light.

Each config is a reproducible stand-in for a dataset the reference would read from
disk (datamanager.cpp:69-101).  Cameras follow the reference's camera model
(vis_utils/camera.cpp:17-34: fovy 45 degrees, aspect = width/height) on a fixed
orbit around the volume centre; lights give the light-camera axes the way
Camera::GetCameraVectors does (camera.cpp:336-341).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import native as N


def orbit_camera(grid_size, width, height, az_deg=30.0, el_deg=20.0, dist_diag=2.0, fovy=45.0) -> N.Camera:
    gx, gy, gz = grid_size
    diag = math.sqrt(gx * gx + gy * gy + gz * gz)
    el, az = math.radians(el_deg), math.radians(az_deg)
    d = dist_diag * diag
    eye = (d * math.cos(el) * math.sin(az), d * math.sin(el), d * math.cos(el) * math.cos(az))
    return N.Camera(tuple(eye), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), fovy, width / height)


def _norm(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def light_from_direction(direction, grid_size, dist_diag=2.0) -> N.Light:
    """A light 'camera' at centre + direction * 2 diag looking at the centre.
    forward = normalize(eye - centre) (towards the light), right = normalize(cross(up, forward)),
    up = normalize(cross(forward, right)), as Camera::GetCameraVectors."""
    gx, gy, gz = grid_size
    diag = math.sqrt(gx * gx + gy * gy + gz * gz)
    fwd = _norm(direction)
    up0 = np.array([0.0, 1.0, 0.0]) if abs(fwd[1]) < 0.99 else np.array([0.0, 0.0, 1.0])
    right = _norm(np.cross(up0, fwd))
    up = _norm(np.cross(fwd, right))
    pos = fwd * dist_diag * diag
    return N.Light(tuple(pos), tuple(fwd), tuple(up), tuple(right))


def seeded_tf(seed: int, n_points: int = 4, max_density: int = 255, alpha_max: float = 0.95):
    """Piecewise-linear TF with `n_points` RGBA control points: isovalues 0 < ... < max
    (interior ones sorted, seeded), colour U[0.1,1]^3, alpha 0 at isovalue 0 then rising
    seeded values <= alpha_max (extinction stays finite)."""
    rng = np.random.default_rng(seed)
    interior = np.sort(rng.choice(np.arange(20, max_density - 20), size=n_points - 2, replace=False))
    iso = np.concatenate([[0], interior, [max_density]]).astype(np.float64)
    rgb = rng.uniform(0.1, 1.0, size=(n_points, 3))
    alpha = np.sort(rng.uniform(0.02, alpha_max, size=n_points - 1))
    alpha = np.concatenate([[0.0], alpha])
    rgb_iso = np.column_stack([rgb, iso])
    a_iso = np.column_stack([alpha, iso])
    return rgb_iso, a_iso


@dataclass
class Workload:
    name: str
    volume: np.ndarray
    scale: tuple
    rgb_iso: np.ndarray
    alpha_iso: np.ndarray
    width: int
    height: int
    light_dir: tuple = (-0.45, 0.75, 0.5)
    max_density: int = 255
    extra: dict = field(default_factory=dict)

    @property
    def grid(self):
        nz, ny, nx = self.volume.shape
        return (nx * self.scale[0], ny * self.scale[1], nz * self.scale[2])

    def camera(self, az=30.0, el=20.0):
        return orbit_camera(self.grid, self.width, self.height, az, el)

    def light(self):
        return light_from_direction(self.light_dir, self.grid)


def plain_blobs(n=128, width=512, height=512, seed=1, n_blobs=8) -> Workload:
    """Plain rc1pass: n^3 uint8 sum of `n_blobs` Gaussian blobs, 4-point TF, step 0.5."""
    vol = N.synth_blobs((n, n, n), N.U8, n_blobs=n_blobs, seed=seed)
    rgb, a = seeded_tf(seed, 4)
    return Workload(f"rc1p_blobs{n}", vol, (1.0, 1.0, 1.0), rgb, a, width, height)


def banded_noise(n=256, width=1024, height=1024, seed=2) -> Workload:
    """Extinction-based shading: n^3 uint16 4-octave value noise cut into 3 bands
    (0.2, 0.5, 0.8 of the range); TF with control points at the band values."""
    vol = N.synth_noise((n, n, n), N.U16, octaves=4, seed=seed, bands=[0.2, 0.5, 0.8])
    rng = np.random.default_rng(seed)
    iso = np.array([0, 40, 51, 115, 128, 191, 204, 255], dtype=np.float64)
    rgb = rng.uniform(0.1, 1.0, size=(len(iso), 3))
    alpha = np.array([0.0, 0.0, 0.05, 0.05, 0.3, 0.3, 0.8, 0.8])
    return Workload(f"ebs_noise{n}", vol, (1.0, 1.0, 1.0), np.column_stack([rgb, iso]),
                    np.column_stack([alpha, iso]), width, height)


def cone_blobs(n=256, width=1024, height=1024, seed=3, n_blobs=64, noise_weight=0.0) -> Workload:
    """Cone-traced lighting: n^3 uint8 sum of 64 Gaussian blobs (+ optional value noise),
    4-point TF, one directional light."""
    vol = N.synth_blobs((n, n, n), N.U8, n_blobs=n_blobs, seed=seed, noise_weight=noise_weight, noise_octaves=4)
    rgb, a = seeded_tf(seed + 100, 4)
    return Workload(f"cone_blobs{n}", vol, (1.0, 1.0, 1.0), rgb, a, width, height)


def surrogate_blobs_noise(n=512, width=1024, height=1024, seed=4) -> Workload:
    """Learned surrogate: n^3 uint8 of 64 Gaussian blobs plus 4-octave value noise."""
    return cone_blobs(n, width, height, seed=seed, n_blobs=64, noise_weight=0.25)


BATCH_LIGHTS = ((-0.45, 0.75, 0.5), (0.6, 0.3, -0.7), (0.1, -0.8, 0.4), (-0.7, 0.2, -0.6))


def batch_scene(width=512, height=512, n_views=16, lights=BATCH_LIGHTS, scale_down=1):
    """Batched render: 4 volumes of uneven sizes and types, each with its own seeded TF
    (4-8 control points) — 128^3 u8 blobs, 192x256x160 u16 banded noise, 384^3 u8 blobs
    plus noise, 512^3 u16 blobs plus noise (seeds 5-8) — and per volume `n_views` orbit
    views x the light directions.  Returns (workloads, jobs) with jobs as
    (volume index, camera, light).  `scale_down` divides every volume size (tests)."""
    s = max(1, int(scale_down))
    specs = [
        ("blobs128_u8", lambda: N.synth_blobs((128 // s,) * 3, N.U8, n_blobs=8, seed=5), 4),
        ("noise192x256x160_u16",
         lambda: N.synth_noise((192 // s, 256 // s, 160 // s), N.U16, octaves=4, seed=6, bands=[0.25, 0.55, 0.8]), 6),
        ("blobs_noise384_u8", lambda: N.synth_blobs((384 // s,) * 3, N.U8, n_blobs=64, seed=7, noise_weight=0.25), 8),
        ("blobs_noise512_u16", lambda: N.synth_blobs((512 // s,) * 3, N.U16, n_blobs=64, seed=8, noise_weight=0.25), 5),
    ]
    wls, jobs = [], []
    for i, (name, make, npts) in enumerate(specs):
        vol = make()
        rgb, a = seeded_tf(105 + i, npts)
        wl = Workload(name, vol, (1.0, 1.0, 1.0), rgb, a, width, height)
        wls.append(wl)
        for v in range(n_views):
            az = 360.0 * v / n_views
            el = 20.0 if v % 2 == 0 else -15.0
            cam = orbit_camera(wl.grid, width, height, az, el)
            for ld in lights:
                jobs.append((i, cam, light_from_direction(ld, wl.grid)))
    return wls, jobs
