"""Scenes and cameras of the workloads the tracer is measured and checked on.

* ``f0``: the reference's default scene (Program.cs:24-71: two spheres r=0.1 at
  (0.3,0.3,0.3) and (0.7,0.7,0.7), BaseDepth 4, camera at (-2,0.5,0.5) with an
  identity facing), built to depth 8 instead of 11 for size;
* ``config1``: 640x480, 60 deg FoV camera on a radius-2 sphere (seeded pose) looking
  at the origin; unit-sphere SDF + 3-octave value noise (amplitude 0.05, seed 1),
  depth 8, world [-1.25, 1.25]^3 mapped onto the octree's unit cube;
* ``config2``: 1920x1080, depth-12 fBm heightfield (6 octaves, lacunarity 2,
  gain 0.5, seed 7) over a 4096^2 grid, 100 frames along a seeded orbit;
* ``config3``: the same heightfield law on a 16384^2 grid, streamed to depth 14 along a
  seeded 300-frame spline fly-through at 3840x2160 (seed 3);
* ``config4``: 64 cameras around the config-2 scene, resolutions log-uniform between
  256x256 and 1920x1080, poses on a hemisphere, apertures 1-4 px (seed 11);
* ``config5``: depth-10 sphere+noise scene (seed 3), 32 views of 512x512.

All randomness is SplitMix64 with the stated seeds.  Camera facing rows are
(forward, right, up): setupInitialTrace (wavefront.cl:563-595) maps image x to
-sin(u) along row 1, so row 1 points to the image's right.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Iterator, List, Sequence, Tuple

import numpy as np

from . import (GEOMETRY_FBM_HEIGHTFIELD, GEOMETRY_SPHERE_NOISE, GEOMETRY_SPHERES, Geometry, TraceInput,
               build_tree)
import ctypes as C

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)


def look_at(eye: Sequence[float], target: Sequence[float], world_up=(0.0, 0.0, 1.0)) -> List[float]:
    """Facing matrix rows (forward, right, up) for a camera at `eye` looking at `target`."""
    f = np.asarray(target, np.float64) - np.asarray(eye, np.float64)
    f /= np.linalg.norm(f)
    up = np.asarray(world_up, np.float64)
    r = np.cross(up, f)
    if np.linalg.norm(r) < 1e-9:
        r = np.cross(np.array([0.0, 1.0, 0.0]), f)
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    return [float(v) for v in (*f, *r, *u)]


def trace_input(origin, facing, width: int, height: int, fov_x: float, fov_y: float, base_depth: int,
                tick: int = 1, max_child_request_id: int = 6000) -> TraceInput:
    """TraceInput with the cone/material defaults of Program.cs:302-318."""
    t = TraceInput()
    for i in range(3):
        t.origin[i] = float(origin[i])
    for i in range(9):
        t.facing[i] = float(facing[i])
    t.fov[0] = fov_x
    t.fov[1] = fov_y
    t.dof[0] = 0.0
    t.dof[1] = 0.0
    t.screen_size[0] = width
    t.screen_size[1] = height
    t.max_opacity = 200
    t.base_depth = base_depth
    t.tick = tick
    t.max_child_request_id = max_child_request_id
    t.fov_multiplier = 0.2
    t.fov_constant = 0.2
    t.weighting_multiplier = -0.05
    t.weighting_constant = 0.5
    t.ambient_light_level = 100
    return t


@dataclass
class SceneSpec:
    name: str
    geometry: Geometry
    base_depth: int
    max_depth: int
    max_size: int
    keepalive: object = None

    def build(self):
        return build_tree(self.geometry, self.base_depth, self.max_depth, self.max_size)


def f0_scene(max_depth: int = 8) -> SceneSpec:
    shapes = (C.c_float * 8)(0.3, 0.3, 0.3, 0.1, 0.7, 0.7, 0.7, 0.1)
    g = Geometry()
    g.kind = GEOMETRY_SPHERES
    g.count = 2
    g.shapes = C.cast(shapes, C.POINTER(C.c_float))
    return SceneSpec("f0-two-spheres", g, 4, max_depth, 0xFFFFFFFF // 64, keepalive=shapes)


def f0_camera(width: int = 64, height: int = 48, tick: int = 1) -> TraceInput:
    """Program.cs:302-318 camera; FoV.x follows the aspect rule of StateManager.cs:105-111."""
    fov_x = width / height * math.pi / 4.0
    ident = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
    return trace_input((-2.0, 0.5, 0.5), ident, width, height, fov_x, math.pi / 4.0, 4, tick)


def sphere_noise_scene(seed: int, max_depth: int, base_depth: int = 4, amplitude: float = 0.05) -> SceneSpec:
    g = Geometry()
    g.kind = GEOMETRY_SPHERE_NOISE
    g.seed = seed
    g.octaves = 3
    g.lacunarity = 2.0
    g.gain = 0.5
    g.amplitude = amplitude
    g.frequency = 4.0
    g.half_extent = 1.25
    return SceneSpec(f"sphere-noise-s{seed}-d{max_depth}", g, base_depth, max_depth, 0xFFFFFFFF // 8)


def world_to_octree(p, s: float = 1.25):
    return [(float(v) + s) / (2.0 * s) for v in p]


def sphere_pose(rng: SplitMix64, radius_world: float = 2.0, s: float = 1.25):
    """Seeded camera position on a sphere of the given world radius, looking at the origin."""
    z = 2.0 * rng.uniform() - 1.0
    phi = 2.0 * math.pi * rng.uniform()
    rr = math.sqrt(max(0.0, 1.0 - z * z))
    eye_w = (radius_world * rr * math.cos(phi), radius_world * rr * math.sin(phi), radius_world * z)
    eye = world_to_octree(eye_w, s)
    target = world_to_octree((0.0, 0.0, 0.0), s)
    return eye, look_at(eye, target)


def config1_scene() -> SceneSpec:
    return sphere_noise_scene(seed=1, max_depth=8)


def config1_camera(width: int = 640, height: int = 480, tick: int = 1) -> TraceInput:
    rng = SplitMix64(1)
    eye, facing = sphere_pose(rng)
    fov_x = math.radians(60.0)
    return trace_input(eye, facing, width, height, fov_x, fov_x * height / width, 4, tick)


def config2_scene(max_depth: int = 12, resolution: int = 4096) -> SceneSpec:
    g = Geometry()
    g.kind = GEOMETRY_FBM_HEIGHTFIELD
    g.seed = 7
    g.octaves = 6
    g.lacunarity = 2.0
    g.gain = 0.5
    g.frequency = 4.0
    g.resolution = resolution
    g.z_min = 0.36
    g.z_max = 0.44
    return SceneSpec(f"fbm-heightfield-d{max_depth}", g, 4, max_depth, 0xFFFFFFFF // 8)


def config2_camera(frame: int, width: int = 1920, height: int = 1080, frames: int = 100,
                   tick: int = None) -> TraceInput:
    """Frame `frame` of the seeded orbit over the heightfield."""
    rng = SplitMix64(7)
    theta0 = 2.0 * math.pi * rng.uniform()
    radius = 0.30 + 0.05 * rng.uniform()
    height_z = 0.60 + 0.04 * rng.uniform()
    th = theta0 + 2.0 * math.pi * frame / frames
    eye = (0.5 + radius * math.cos(th), 0.5 + radius * math.sin(th), height_z)
    target = (0.5 + 0.12 * math.cos(th + 1.3), 0.5 + 0.12 * math.sin(th + 1.3), 0.40)
    fov_x = math.radians(60.0)
    t = (frame % 65533) + 1 if tick is None else tick
    return trace_input(eye, look_at(eye, target), width, height, fov_x, fov_x * height / width, 4, t)


def config3_geometry(resolution: int = 16384, seed: int = 7) -> Geometry:
    """The streamed terrain: the config-2 fBm heightfield law (6 octaves, lacunarity 2,
    gain 0.5) sampled on a 16384^2 grid, so that depth-14 cells have their own samples."""
    g = Geometry()
    g.kind = GEOMETRY_FBM_HEIGHTFIELD
    g.seed = seed
    g.octaves = 6
    g.lacunarity = 2.0
    g.gain = 0.5
    g.frequency = 4.0
    g.resolution = resolution
    g.z_min = 0.36
    g.z_max = 0.44
    return g


def _catmull_rom(p0, p1, p2, p3, t):
    t2, t3 = t * t, t * t * t
    return [0.5 * (2 * b + (-a + c) * t + (2 * a - 5 * b + 4 * c - d) * t2 + (-a + 3 * b - 3 * c + d) * t3)
            for a, b, c, d in zip(p0, p1, p2, p3)]


def config3_camera(frame: int, width: int = 3840, height: int = 2160, frames: int = 300, tick: int = None,
                   seed: int = 3, control_points: int = 8, max_child_request_id: int = 0,
                   base_depth: int = 4) -> TraceInput:
    """Frame `frame` of the seeded fly-through: a closed Catmull-Rom spline through
    `control_points` seeded points 0.05-0.09 above the terrain's top (z 0.49-0.53) over
    [0.15, 0.85]^2, travelled once in `frames` frames; the camera looks along the path,
    pitched 25 degrees down."""
    rng = SplitMix64(seed)
    pts = [(0.15 + 0.7 * rng.uniform(), 0.15 + 0.7 * rng.uniform(), 0.49 + 0.04 * rng.uniform())
           for _ in range(control_points)]

    def at(u):
        i = int(math.floor(u))
        return _catmull_rom(*[pts[(i + k - 1) % control_points] for k in range(4)], u - i)

    u = (frame % frames) / frames * control_points
    eye, ahead = at(u), at(u + 1e-3)
    dx, dy = ahead[0] - eye[0], ahead[1] - eye[1]
    norm = math.hypot(dx, dy) or 1.0
    pitch = math.radians(25.0)
    fwd = (math.cos(pitch) * dx / norm, math.cos(pitch) * dy / norm, -math.sin(pitch))
    target = (eye[0] + fwd[0], eye[1] + fwd[1], eye[2] + fwd[2])
    fov_x = math.radians(60.0)
    tk = (frame % 65533) + 1 if tick is None else tick
    cam = trace_input(eye, look_at(eye, target), width, height, fov_x, fov_x * height / width, base_depth, tk)
    cam.max_child_request_id = max_child_request_id or width * height
    return cam


def config4_cameras(count: int = 64, tick: int = 1, seed: int = 11) -> Tuple[List[TraceInput], List[float]]:
    """Batched cameras around the config-2 heightfield: resolution (256*7.5^t, 256*4.21875^t)
    with t uniform (pixel count log-uniform between 256^2 and 1920x1080), eye on a hemisphere
    of radius 0.3-0.4 above the scene centre, aperture an integer 1..4 pixels."""
    rng = SplitMix64(seed)
    cams, apertures = [], []
    centre = (0.5, 0.5, 0.40)
    for _ in range(count):
        t = rng.uniform()
        w = int(round(256.0 * (1920.0 / 256.0) ** t))
        h = int(round(256.0 * (1080.0 / 256.0) ** t))
        z = 0.35 + 0.6 * rng.uniform()
        phi = 2.0 * math.pi * rng.uniform()
        radius = 0.30 + 0.10 * rng.uniform()
        rr = math.sqrt(max(0.0, 1.0 - z * z))
        eye = (centre[0] + radius * rr * math.cos(phi), centre[1] + radius * rr * math.sin(phi),
               centre[2] + radius * z)
        target = (centre[0] + 0.1 * (rng.uniform() - 0.5), centre[1] + 0.1 * (rng.uniform() - 0.5), centre[2])
        fov_x = math.radians(60.0)
        cams.append(trace_input(eye, look_at(eye, target), w, h, fov_x, fov_x * h / w, 4, tick))
        apertures.append(float(1 + int(rng.uniform() * 4.0) % 4))
    return cams, apertures


def config5_scene() -> SceneSpec:
    return sphere_noise_scene(seed=3, max_depth=10)


def config5_cameras(views: int = 32, size: int = 512) -> List[TraceInput]:
    rng = SplitMix64(3)
    out = []
    fov = math.radians(60.0)
    for v in range(views):
        z = 0.15 + 0.85 * rng.uniform()  # upper hemisphere
        phi = 2.0 * math.pi * rng.uniform()
        rr = math.sqrt(max(0.0, 1.0 - z * z))
        eye = world_to_octree((2.0 * rr * math.cos(phi), 2.0 * rr * math.sin(phi), 2.0 * z))
        out.append(trace_input(eye, look_at(eye, world_to_octree((0, 0, 0))), size, size, fov, fov, 4, v + 1))
    return out
