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
    gravity: tuple = (0.0, -10.0, 0.0)
    restitution_iterations: int = 2
    restitution_threshold: float = 1.0
    restitution_propagation: bool = False

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


PI = f32(3.14159265359)  # base/constants.rs:6


def unwind_angle(radians: np.float32) -> np.float32:
    """IEEE remainder by 2*pi, computed exactly in f64 (math/functions.rs:134-150)."""
    x = float(radians)
    period = float(f32(2.0) * PI)
    n = float(round(x / period))  # ties to even
    r = x - n * period
    if r > 0.5 * period:
        n += 1.0
        r = x - n * period
    elif r < -0.5 * period:
        n -= 1.0
        r = x - n * period
    return f32(r)


def compute_cos_sin(radians: np.float32) -> tuple:
    """Rational cosine/sine approximation, normalised (math/functions.rs:101-130)."""
    x = unwind_angle(radians)
    pi2 = PI * PI
    if x < f32(-0.5) * PI:
        y = x + PI
        y2 = y * y
        c = -(pi2 - f32(4.0) * y2) / (pi2 + y2)
    elif x > f32(0.5) * PI:
        y = x - PI
        y2 = y * y
        c = -(pi2 - f32(4.0) * y2) / (pi2 + y2)
    else:
        y2 = x * x
        c = (pi2 - f32(4.0) * y2) / (pi2 + y2)
    if x < f32(0.0):
        y = x + PI
        s = f32(-16.0) * y * (PI - y) / (f32(5.0) * pi2 - f32(4.0) * y * (PI - y))
    else:
        s = f32(16.0) * x * (PI - x) / (f32(5.0) * pi2 - f32(4.0) * x * (PI - x))
    mag = np.sqrt(s * s + c * c)
    inv_mag = f32(1.0) / mag if mag > f32(0.0) else f32(0.0)
    return f32(c * inv_mag), f32(s * inv_mag)


def overflow_hub(ring_count: int = 5, per_ring: int = 5, enable_sleep: bool = True) -> Scene:
    """A tall dynamic hub ringed by small boxes, so one body has more contacts than the
    dynamic colours can hold and the overflow colour fills
    (src/tests/simulation/overflow_color_tests.rs:16-79)."""
    sc = Scene(name="overflow_hub", enable_sleep=enable_sleep)
    g = sc.add_body(BodySpec(STATIC, (0.0, -1.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (20.0, 1.0, 20.0)))
    hx, hy, hz = f32(0.5), f32(2.5), f32(0.5)
    hub = sc.add_body(BodySpec(DYNAMIC, (0.0, float(hy), 0.0)))
    sc.add_shape(ShapeSpec(hub, BOX, (float(hx), float(hy), float(hz)), density=50.0))
    nh = f32(0.2)
    ring_radius = hx + nh - f32(0.03)
    spacing = f32(0.5)
    base_y = nh + f32(0.05)
    for ring in range(ring_count):
        y = base_y + spacing * f32(ring)
        offset = PI / f32(per_ring) if ring & 1 else f32(0.0)
        for slot in range(per_ring):
            theta = offset + (f32(2.0) * PI * f32(slot)) / f32(per_ring)
            c, s = compute_cos_sin(theta)
            b = sc.add_body(BodySpec(DYNAMIC, (float(ring_radius * c), float(y), float(ring_radius * s))))
            sc.add_shape(ShapeSpec(b, BOX, (float(nh),) * 3))
    return sc


def graph_scene() -> Scene:
    """Three spheres and a box on a static slab (src/tests/graph_tests.rs:33-66)."""
    sc = Scene(name="graph_scene", enable_sleep=True)
    g = sc.add_body(BodySpec(STATIC, (0.0, 0.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (10.0, 0.5, 10.0), density=0.0))
    for p in ((0.0, 0.9, 0.0), (0.6, 0.9, 0.0), (0.0, 2.0, 0.0)):
        b = sc.add_body(BodySpec(DYNAMIC, tuple(float(f32(v)) for v in p)))
        sc.add_shape(ShapeSpec(b, SPHERE, (0.0, 0.0, 0.0, 0.5), density=1000.0))
    b = sc.add_body(BodySpec(DYNAMIC, (float(f32(-0.9)), float(f32(0.9)), 0.0)))
    sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5), density=1000.0))
    return sc


class XorShift32:
    """Seeded xorshift32 (Marsaglia) for the synthetic scenes."""

    def __init__(self, seed: int):
        self.x = (seed & 0xFFFFFFFF) or 0x9E3779B9

    def next_u32(self) -> int:
        x = self.x
        x ^= (x << 13) & 0xFFFFFFFF
        x ^= x >> 17
        x ^= (x << 5) & 0xFFFFFFFF
        self.x = x & 0xFFFFFFFF
        return self.x

    def uniform(self, lo: float, hi: float) -> np.float32:
        u = f32((self.next_u32() >> 8) * (1.0 / 16777216.0))
        return f32(lo) + (f32(hi) - f32(lo)) * u

    def below(self, n: int) -> int:
        return self.next_u32() % n


def _ground(sc: Scene, half_x: float, half_z: float) -> None:
    """Static ground box, top at y = 0, a metre wider than the footprint on each side."""
    g = sc.add_body(BodySpec(STATIC, (0.0, -0.5, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (half_x + 1.0, 0.5, half_z + 1.0)))


def _container(sc: Scene, half_x: float, half_z: float, height: float) -> None:
    """Static floor (top at y = 0) and four walls around a half_x x half_z footprint."""
    g = sc.add_body(BodySpec(STATIC, (0.0, -0.5, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (half_x + 1.0, 0.5, half_z + 1.0)))
    hy = height / 2.0 + 1.0
    for sx in (-1.0, 1.0):
        w = sc.add_body(BodySpec(STATIC, (sx * (half_x + 0.5), hy - 1.0, 0.0)))
        sc.add_shape(ShapeSpec(w, BOX, (0.5, hy, half_z + 1.0)))
    for sz in (-1.0, 1.0):
        w = sc.add_body(BodySpec(STATIC, (0.0, hy - 1.0, sz * (half_z + 0.5))))
        sc.add_shape(ShapeSpec(w, BOX, (half_x + 1.0, hy, 0.5)))


def pile(count: int = 10_000, seed: int = 0x5EED, size_x: int = 20, size_y: int = 60, size_z: int = 20,
         base_y: float = 1.0, container: bool = True) -> Scene:
    """Random pile: `count` dynamic bodies, a third each boxes, spheres and capsules.

    Sizes s ~ U[0.25, 0.75] m are full extents: box edge lengths (per axis), sphere
    diameter, capsule total length (radius s/4, segment s/2 along local y). Density
    1000, friction ~ U[0.4, 0.8]. Bodies take distinct cells of a 1 m lattice over a
    size_x x size_y x size_z region above the floor (seeded shuffle), jittered by at
    most (1 - s)/2 so no two start overlapping. The container is a static floor and
    four walls (or only the floor, `container=False`). Continuous collision and sleeping
    are off.
    """
    sc = Scene(name=f"pile_{count}", enable_sleep=False, enable_continuous=False)
    if container:
        _container(sc, size_x / 2.0, size_z / 2.0, float(size_y) + base_y + 2.0)
    else:
        _ground(sc, size_x / 2.0, size_z / 2.0)
    rng = XorShift32(seed)
    cells = size_x * size_y * size_z
    if count > cells:
        raise ValueError("more bodies than lattice cells")
    order = list(range(cells))
    for i in range(cells - 1, 0, -1):  # Fisher-Yates
        j = rng.below(i + 1)
        order[i], order[j] = order[j], order[i]
    n_box = count - 2 * (count // 3)
    for k in range(count):
        cell = order[k]
        cx, rest = cell % size_x, cell // size_x
        cy, cz = rest % size_y, rest // size_y
        kind = BOX if k < n_box else (SPHERE if k < n_box + count // 3 else CAPSULE)
        friction = float(rng.uniform(0.4, 0.8))
        if kind == BOX:
            s = [rng.uniform(0.25, 0.75) for _ in range(3)]
            smax = max(s)
        else:
            s = rng.uniform(0.25, 0.75)
            smax = s
        slack = (f32(1.0) - f32(smax)) * f32(0.5)
        jitter = [float(rng.uniform(-1.0, 1.0) * slack) for _ in range(3)]
        pos = (-size_x / 2.0 + cx + 0.5 + jitter[0], base_y + cy + 0.5 + jitter[1], -size_z / 2.0 + cz + 0.5 + jitter[2])
        pos = tuple(float(f32(p)) for p in pos)
        b = sc.add_body(BodySpec(DYNAMIC, pos, enable_sleep=False))
        if kind == BOX:
            sc.add_shape(ShapeSpec(b, BOX, tuple(float(f32(v) * f32(0.5)) for v in s), density=1000.0, friction=friction))
        elif kind == SPHERE:
            sc.add_shape(ShapeSpec(b, SPHERE, (0.0, 0.0, 0.0, float(f32(s) * f32(0.5))), density=1000.0, friction=friction))
        else:
            q = float(f32(s) * f32(0.25))
            sc.add_shape(ShapeSpec(b, CAPSULE, (0.0, -q, 0.0, 0.0, q, 0.0, q), density=1000.0, friction=friction))
    return sc


def stacks(count: int, seed: int, base_y: float = 0.5) -> Scene:
    """Towers of up to 8 mixed bodies (boxes, spheres, capsules in turn; sizes as in
    `pile`) on a 1.5 m grid, 1 m apart in height, on a ground box."""
    sc = Scene(name=f"stacks_{count}", enable_sleep=False, enable_continuous=False)
    towers = -(-count // 8)
    side = max(1, int(np.ceil(np.sqrt(towers))))
    half = 0.75 * side
    _ground(sc, half, half)
    rng = XorShift32(seed)
    for k in range(count):
        t, level = k // 8, k % 8
        x = -half + 0.75 + 1.5 * (t % side)
        z = -half + 0.75 + 1.5 * (t // side)
        kind = (BOX, SPHERE, CAPSULE)[k % 3]
        friction = float(rng.uniform(0.4, 0.8))
        pos = (float(f32(x)), float(f32(base_y + level + 0.5)), float(f32(z)))
        b = sc.add_body(BodySpec(DYNAMIC, pos, enable_sleep=False))
        if kind == BOX:
            h = tuple(float(rng.uniform(0.25, 0.75) * f32(0.5)) for _ in range(3))
            sc.add_shape(ShapeSpec(b, BOX, h, density=1000.0, friction=friction))
        elif kind == SPHERE:
            sc.add_shape(ShapeSpec(b, SPHERE, (0.0, 0.0, 0.0, float(rng.uniform(0.25, 0.75) * f32(0.5))), density=1000.0,
                                   friction=friction))
        else:
            q = float(rng.uniform(0.25, 0.75) * f32(0.25))
            sc.add_shape(ShapeSpec(b, CAPSULE, (0.0, -q, 0.0, 0.0, q, 0.0, q), density=1000.0, friction=friction))
    return sc


def batch_body_count(rng: XorShift32, lo: int = 8, hi: int = 256) -> int:
    """round(exp(U[ln lo, ln hi])): log-uniform body counts."""
    u = float(rng.uniform(0.0, 1.0))
    return int(round(float(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo))))))


def batched(count: int = 4096, seed: int = 0xB47C4, lo: int = 8, hi: int = 256) -> list:
    """Independent worlds with log-uniform body counts in [lo, hi]: even worlds are piles
    (the random-pile generator on a ground box, no walls), odd worlds are stacks."""
    out = []
    for k in range(count):
        rng = XorShift32((seed + 0x9E3779B9 * (k + 1)) & 0xFFFFFFFF)
        n = batch_body_count(rng, lo, hi)
        world_seed = rng.next_u32()
        if k % 2 == 0:
            side = max(2, int(np.ceil(np.sqrt(n / 3.0))))
            height = -(-n // (side * side)) + 1
            out.append(pile(n, seed=world_seed, size_x=side, size_y=height, size_z=side, base_y=0.5, container=False))
        else:
            out.append(stacks(n, seed=world_seed))
    return out



# Restitution scenes: src/tests/simulation/restitution_phase_tests.rs. The ground is a
# 80 x 2 x 80 box whose top face is y = 0; balls are spheres of radius 0.5.
IMPACT_SPEED = f32(5.0)  # restitution_phase_tests.rs:22
RESTITUTION_TIME_STEP = f32(1.0 / 60.0)


def _bounce_ground(sc: Scene, restitution: float) -> None:
    """make_ground (restitution_phase_tests.rs:35-44)."""
    g = sc.add_body(BodySpec(STATIC, (0.0, -1.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (40.0, 1.0, 40.0), friction=0.0, restitution=restitution))


def _ball(sc: Scene, x: float, y, vy: float, restitution: float) -> int:
    """make_ball (restitution_phase_tests.rs:46-58): f32 position widened to f64."""
    b = sc.add_body(BodySpec(DYNAMIC, (float(f32(x)), float(f32(y)), 0.0), linear_velocity=(0.0, float(f32(vy)), 0.0)))
    sc.add_shape(ShapeSpec(b, SPHERE, (0.0, 0.0, 0.0, 0.5), friction=0.0, restitution=restitution))
    return b


def ground_bounce(restitution: float, gap) -> Scene:
    """A ball driven at 5 m/s into the ground from `gap` above it, gravity off
    (measure_ground_bounce, restitution_phase_tests.rs:62-72)."""
    sc = Scene(name=f"ground_bounce_{restitution}", enable_sleep=False, gravity=(0.0, 0.0, 0.0))
    _bounce_ground(sc, restitution)
    _ball(sc, 0.0, f32(0.5) + f32(gap), -IMPACT_SPEED, restitution)
    return sc


def supported_bounce(restitution: float, support_count: int, propagate: bool) -> Scene:
    """A ball driven onto a column of resting balls on the ground, four restitution
    iterations (measure_supported_bounce, restitution_phase_tests.rs:77-96). The impactor is
    the last body."""
    sc = Scene(name=f"supported_{restitution}_{support_count}", enable_sleep=False, gravity=(0.0, 0.0, 0.0),
               restitution_iterations=4, restitution_propagation=propagate)
    _bounce_ground(sc, 0.0)
    for i in range(support_count):
        _ball(sc, 0.0, f32(0.5) + f32(i), 0.0, 0.0)
    y = f32(0.5) + f32(support_count) + f32(0.5) * IMPACT_SPEED * RESTITUTION_TIME_STEP
    _ball(sc, 0.0, y, -IMPACT_SPEED, restitution)
    return sc


def threshold_bounce(threshold: float, speed: float) -> Scene:
    """A ball resting on a bouncy ground at approach speed `speed` under a restitution
    threshold (threshold / threshold_above, restitution_phase_tests.rs:160-193)."""
    sc = Scene(name=f"threshold_{threshold}", enable_sleep=False, gravity=(0.0, 0.0, 0.0), restitution_threshold=threshold)
    _bounce_ground(sc, 0.9)
    _ball(sc, 0.0, 0.5, -speed, 0.9)
    return sc


def head_on(restitution: float, density_b: float) -> Scene:
    """Two free unit spheres closing head on at 2.5 m/s each, gravity off; the second has
    density `density_b` (measure_head_on, restitution_phase_tests.rs:202-238)."""
    sc = Scene(name=f"head_on_{restitution}_{density_b}", enable_sleep=False, gravity=(0.0, 0.0, 0.0))
    half = float(f32(0.5) * IMPACT_SPEED)
    for x, vx, rho in ((-1.0, half, 1.0), (1.0, -half, density_b)):
        b = sc.add_body(BodySpec(DYNAMIC, (x, 0.0, 0.0), linear_velocity=(vx, 0.0, 0.0)))
        sc.add_shape(ShapeSpec(b, SPHERE, (0.0, 0.0, 0.0, 0.5), density=rho, friction=0.0, restitution=restitution))
    return sc


def drop_ball(drop_height: float) -> Scene:
    """A perfectly elastic ball dropped from rest under gravity onto a dead ground
    (measure_apexes, restitution_phase_tests.rs:937-958). The ball is the last body."""
    sc = Scene(name=f"drop_{drop_height}", enable_sleep=False)
    _bounce_ground(sc, 0.0)
    _ball(sc, 0.0, drop_height, 0.0, 1.0)
    return sc


def resting_stack(box_count: int = 10) -> Scene:
    """Ten unit boxes stacked on a bouncy ground, all with restitution 0.9, gravity on
    (resting_stack, restitution_phase_tests.rs:1001-1049)."""
    sc = Scene(name="resting_stack", enable_sleep=False)
    _bounce_ground(sc, 0.9)
    for i in range(box_count):
        b = sc.add_body(BodySpec(DYNAMIC, (0.0, 0.5 + i, 0.0)))
        sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5), density=1.0, friction=0.6, restitution=0.9))
    return sc


def flat_landing(restitution: float, sub_steps: int = 4, spin: float = 0.0) -> Scene:
    """A flat 2 x 0.5 x 2 box driven at 5 m/s into frictionless ground with gravity off and
    four restitution iterations, optionally spinning about x (measure_flat_bounce and
    measure_spinning_bounce, restitution_phase_tests.rs:824-855, 889-918). The box is the
    last body."""
    sc = Scene(name=f"flat_landing_{restitution}_{spin}", enable_sleep=False, gravity=(0.0, 0.0, 0.0),
               restitution_iterations=4, substeps=sub_steps)
    _bounce_ground(sc, 0.0)
    y = f32(0.25) + f32(0.5) * IMPACT_SPEED * RESTITUTION_TIME_STEP
    b = sc.add_body(BodySpec(DYNAMIC, (0.0, float(y), 0.0), linear_velocity=(0.0, -float(IMPACT_SPEED), 0.0),
                             angular_velocity=(float(f32(spin)), 0.0, 0.0)))
    sc.add_shape(ShapeSpec(b, BOX, (1.0, 0.25, 1.0), density=1.0, friction=0.0, restitution=restitution))
    return sc
