"""Seeded scene descriptions shared by the tests, the smoke check and the benchmark.

A scene is a plain list of bodies and shapes, built with the same f32 arithmetic the
reference's scene builders use, so any stepper (the HIP library or the CPU oracle) can
be loaded from it identically.

References:
  large_pyramid / many_pyramids   src/bin/benchmarks/scenes.rs:105-198
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

F32 = np.float32

STATIC, KINEMATIC, DYNAMIC = 0, 1, 2
BOX, SPHERE, CAPSULE = 0, 1, 2


def f32(x) -> np.float32:
    return np.float32(x)


@dataclass
class BodySpec:
    body_type: int
    position: tuple  # f64 x3
    rotation: tuple = (0.0, 0.0, 0.0, 1.0)  # x, y, z, s (f32)
    linear_velocity: tuple = (0.0, 0.0, 0.0)
    angular_velocity: tuple = (0.0, 0.0, 0.0)
    enable_sleep: bool = True
    linear_damping: float = 0.0
    angular_damping: float = 0.0
    gravity_scale: float = 1.0


@dataclass
class ShapeSpec:
    body: int
    kind: int
    params: tuple  # box: (hx, hy, hz); sphere: (cx, cy, cz, r); capsule: (x1,y1,z1, x2,y2,z2, r)
    density: float = 1000.0
    friction: float = 0.6
    restitution: float = 0.0


@dataclass
class Scene:
    name: str
    bodies: list = field(default_factory=list)
    shapes: list = field(default_factory=list)
    enable_sleep: bool = True
    enable_continuous: bool = True
    dt: float = 1.0 / 60.0
    substeps: int = 4

    def add_body(self, spec: BodySpec) -> int:
        self.bodies.append(spec)
        return len(self.bodies) - 1

    def add_shape(self, spec: ShapeSpec) -> int:
        self.shapes.append(spec)
        return len(self.shapes) - 1

    @property
    def dynamic_count(self) -> int:
        return sum(1 for b in self.bodies if b.body_type == DYNAMIC)


def pyramid(base_count: int = 100, density: float = 100.0, half: float = 0.5, ground_half=(400.0, 1.0, 400.0)) -> Scene:
    """scenes.rs:105-139 (large_pyramid); base_count=20, density=1000 gives the 210-box scene."""
    sc = Scene(name=f"pyramid_{base_count}", enable_sleep=False)
    g = sc.add_body(BodySpec(STATIC, (0.0, -1.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, tuple(float(f32(v)) for v in ground_half)))
    h = f32(half)
    shift = h
    for i in range(base_count):
        y = (f32(2.0) * f32(i) + f32(1.0)) * shift
        for j in range(i, base_count):
            x = (f32(i) + f32(1.0)) * shift + f32(2.0) * f32(j - i) * shift - h * f32(base_count)
            b = sc.add_body(BodySpec(DYNAMIC, (float(x), float(y), 0.0)))
            sc.add_shape(ShapeSpec(b, BOX, (float(h), float(h), float(h)), density=density))
    return sc


def many_pyramids(base_count: int = 10, rows: int = 14, cols: int = 14) -> Scene:
    """scenes.rs:143-198 (many_pyramids)."""
    sc = Scene(name="many_pyramids", enable_sleep=True)
    extent = f32(0.5)
    ground_extent = extent * f32(cols) * (f32(base_count) + f32(1.0))
    g = sc.add_body(BodySpec(STATIC, (0.0, -1.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (float(ground_extent), 1.0, float(ground_extent))))
    base_width = f32(2.0) * extent * f32(base_count)
    base_z = -ground_extent + f32(2.0) * extent
    delta_z = f32(2.0) * (ground_extent - f32(2.0) * extent) / f32(rows - 1)
    z = base_z
    for _ in range(rows):
        for j in range(cols):
            center_x = -ground_extent + f32(j) * (base_width + f32(2.0) * extent) + f32(2.0) * extent
            for i in range(base_count):
                y = (f32(2.0) * f32(i) + f32(1.0)) * extent
                for k in range(i, base_count):
                    x = (f32(i) + f32(1.0)) * extent + f32(2.0) * f32(k - i) * extent + center_x - f32(0.5)
                    b = sc.add_body(BodySpec(DYNAMIC, (float(x), float(y), float(z)), enable_sleep=False))
                    sc.add_shape(ShapeSpec(b, BOX, (float(extent),) * 3, density=100.0))
        z = f32(z + delta_z)
    return sc
