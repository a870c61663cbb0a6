"""Seeded scene descriptions shared by the tests, the smoke check and the benchmark.

A scene is a plain list of bodies and shapes, built with the same f32 arithmetic the
reference's scene builders use, so any stepper (the HIP library or the CPU oracle) can
be loaded from it identically.

References:
  large_pyramid / many_pyramids   src/bin/benchmarks/scenes.rs:105-198
"""
from __future__ import annotations

import dataclasses
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
    is_bullet: bool = False
    safety_factor: float = 0.5
    sleep_threshold: float = 0.05
    motion_locks: int = 0  # MotionLocks bits: linear x, y, z = 1, 2, 4; angular x, y, z = 8, 16, 32
    allow_fast_rotation: bool = False
    enable_contact_recycling: bool = True


@dataclass
class ShapeSpec:
    body: int
    kind: int
    params: tuple  # box: (hx, hy, hz); sphere: (cx, cy, cz, r); capsule: (x1,y1,z1, x2,y2,z2, r)
    density: float = 1000.0
    friction: float = 0.6
    restitution: float = 0.0
    rolling_resistance: float = 0.0
    category_bits: int = 1
    mask_bits: int = 0xFFFF_FFFF_FFFF_FFFF
    group_index: int = 0
    is_sensor: bool = False
    enable_speculative_contact: bool = True
    enable_hit_events: bool = False


@dataclass
class JointSpec:
    """A revolute joint (RevoluteJointDef, types/joint_defs.rs:30-58, :154-167). Frames are
    (px, py, pz, qx, qy, qz, qs) in the body's frame; the hinge axis is frame A's z axis."""
    body_a: int
    body_b: int
    frame_a: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    frame_b: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    constraint_hertz: float = 60.0
    constraint_damping_ratio: float = 2.0
    collide_connected: bool = False
    enable_spring: bool = False
    hertz: float = 0.0
    damping_ratio: float = 0.0
    target_angle: float = 0.0
    enable_limit: bool = False
    lower_angle: float = 0.0
    upper_angle: float = 0.0
    enable_motor: bool = False
    max_motor_torque: float = 0.0
    motor_speed: float = 0.0


@dataclass
class Scene:
    name: str
    bodies: list = field(default_factory=list)
    shapes: list = field(default_factory=list)
    joints: list = field(default_factory=list)
    enable_sleep: bool = True
    enable_continuous: bool = True
    dt: float = 1.0 / 60.0
    substeps: int = 4
    gravity: tuple = (0.0, -10.0, 0.0)
    restitution_iterations: int = 2
    restitution_threshold: float = 1.0
    restitution_propagation: bool = False
    tags: dict = field(default_factory=dict)  # name -> body index, for scenes that name their bodies

    def add_body(self, spec: BodySpec) -> int:
        self.bodies.append(spec)
        return len(self.bodies) - 1

    def add_shape(self, spec: ShapeSpec) -> int:
        self.shapes.append(spec)
        return len(self.shapes) - 1

    def add_joint(self, spec: JointSpec) -> int:
        self.joints.append(spec)
        return len(self.joints) - 1

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
         base_y: float = 1.0, container: bool = True, wall_height: float | None = None) -> Scene:
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
        _container(sc, size_x / 2.0, size_z / 2.0, wall_height or float(size_y) + base_y + 2.0)
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


# Revolute-joint scenes. The reference's own joint tests (src/tests/revolute_joint_tests.rs,
# src/tests/simulation/joint_inspection_tests.rs) check replay and contact filtering; the
# other scenes below are built so a closed form fixes what the joint must do.
IDENTITY_FRAME = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)


def _frame(p, q=(0.0, 0.0, 0.0, 1.0)) -> tuple:
    return tuple(float(f32(v)) for v in tuple(p) + tuple(q))


def hinged_pendulum(limit=None, spring=None, motor=None, gravity=(0.0, -10.0, 0.0), spin=(0.0, 0.0, 0.0)) -> Scene:
    """A 0.4 m box (density 1000) hinged 0.5 m from its centre to a static body at
    (0, 3, 0), starting level (hinged_pendulum, revolute_joint_tests.rs:103-131). The hinge
    axis is z. `limit` = (lower, upper), `spring` = (hertz, damping ratio, target angle),
    `motor` = (speed, max torque); `spin` is the box's initial angular velocity."""
    sc = Scene(name="hinged_pendulum", enable_sleep=False, gravity=tuple(gravity))
    a = sc.add_body(BodySpec(STATIC, (0.0, 3.0, 0.0)))
    b = sc.add_body(BodySpec(DYNAMIC, (1.5, 3.0, 0.0), angular_velocity=tuple(float(f32(v)) for v in spin)))
    sc.add_shape(ShapeSpec(b, BOX, (0.2, 0.2, 0.2), density=1000.0))
    j = JointSpec(a, b, frame_a=_frame((1.0, 0.0, 0.0)), frame_b=_frame((-0.5, 0.0, 0.0)))
    if limit is not None:
        j.enable_limit, j.lower_angle, j.upper_angle = True, float(f32(limit[0])), float(f32(limit[1]))
    if spring is not None:
        j.enable_spring, j.hertz, j.damping_ratio, j.target_angle = True, *(float(f32(v)) for v in spring)
    if motor is not None:
        j.enable_motor, j.motor_speed, j.max_motor_torque = True, float(f32(motor[0])), float(f32(motor[1]))
    sc.add_joint(j)
    return sc


def touching_boxes(collide_connected: bool = False) -> Scene:
    """Two overlapping unit boxes, gravity off (create_joint_clears_contacts /
    create_joint_keeps_contacts, joint_inspection_tests.rs:460-567). The joint between
    them is added by the test after the first step."""
    sc = Scene(name="touching_boxes", gravity=(0.0, 0.0, 0.0))
    for x in (0.0, 0.5):
        b = sc.add_body(BodySpec(DYNAMIC, (x, 0.0, 0.0)))
        sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5), density=1.0))
    return sc


# A chain link: capsule along x, segment 0.4 m, radius 0.1 m; neighbours are hinged 0.5 m
# apart at the link ends, so adjacent capsules overlap by 0.1 m and only the joint's
# collide_connected = false keeps them from colliding.
LINK_HALF = 0.2
LINK_RADIUS = 0.1
LINK_PITCH = 0.5
# alternate hinges turn about z and about y: frame rotated 90 degrees about x
_QX90 = (float(f32(np.sin(np.pi / 4))), 0.0, 0.0, float(f32(np.cos(np.pi / 4))))


def add_chain(sc: Scene, start: tuple, links: int, anchor: int | None = None, alternate_axes: bool = True,
              density: float = 1000.0, friction: float = 0.6) -> list:
    """Append a chain of `links` capsule links along +x from `start` (the first link's
    centre), joined by revolute joints at the link ends (collide_connected false). With
    `anchor` the first link is also hinged to that body at the chain's start point."""
    ids = []
    x0, y0, z0 = start
    for i in range(links):
        b = sc.add_body(BodySpec(DYNAMIC, (float(f32(x0 + LINK_PITCH * i)), float(f32(y0)), float(f32(z0))), enable_sleep=False))
        sc.add_shape(ShapeSpec(b, CAPSULE, (-LINK_HALF, 0.0, 0.0, LINK_HALF, 0.0, 0.0, LINK_RADIUS), density=density,
                               friction=friction))
        ids.append(b)
    half = 0.5 * LINK_PITCH
    for i in range(links - 1):
        q = _QX90 if (alternate_axes and i % 2 == 1) else (0.0, 0.0, 0.0, 1.0)
        sc.add_joint(JointSpec(ids[i], ids[i + 1], frame_a=_frame((half, 0.0, 0.0), q), frame_b=_frame((-half, 0.0, 0.0), q)))
    if anchor is not None:
        ax = sc.bodies[anchor].position
        p = (float(f32(x0 - half - ax[0])), float(f32(y0 - ax[1])), float(f32(z0 - ax[2])))
        sc.add_joint(JointSpec(anchor, ids[0], frame_a=_frame(p), frame_b=_frame((-half, 0.0, 0.0))))
    return ids


def hanging_chain(links: int = 8) -> Scene:
    """A chain of capsule links hinged (about z) to a static body, starting level, falling
    under gravity onto nothing: it swings down and hangs."""
    sc = Scene(name=f"hanging_chain_{links}", enable_sleep=False)
    anchor = sc.add_body(BodySpec(STATIC, (0.0, 10.0, 0.0)))
    add_chain(sc, (0.5 * LINK_PITCH, 10.0, 0.0), links, anchor=anchor, alternate_axes=False)
    return sc


def chain_pile(count: int = 400, chains: int = 40, links: int = 16, seed: int = 0xC4A1, size_x: int = 12,
               size_y: int = 12, size_z: int = 12) -> Scene:
    """A mixed pile (`pile`'s generator) with chains of capsule links laid in rows above
    it, all falling into the container together."""
    per_row, rows = _chain_rows(chains, links, size_x)
    layers = -(-rows // max(1, size_z - 1))
    y0 = 1.0 + size_y + 1.0
    sc = pile(count, seed, size_x, size_y, size_z, wall_height=y0 + layers + 2.0)
    sc.name = f"chain_pile_{count}_{chains}x{links}"
    _add_chain_rows(sc, chains, links, size_x, size_z, y0)
    return sc


def _chain_rows(chains: int, links: int, size_x: int) -> tuple:
    """Chains per row along x, and rows."""
    per_row = max(1, int(size_x // (LINK_PITCH * links + 0.5)))
    return per_row, -(-chains // per_row)


def _add_chain_rows(sc: Scene, chains: int, links: int, size_x: int, size_z: int, y0: float) -> None:
    """Chains along x, `per_row` to a row, rows 1 m apart in z, layers 1 m apart in y."""
    per_row, rows = _chain_rows(chains, links, size_x)
    rows_per_layer = max(1, size_z - 1)
    length = LINK_PITCH * links
    k = 0
    for r in range(rows):
        layer, rz = divmod(r, rows_per_layer)
        for c in range(per_row):
            if k == chains:
                return
            x = -size_x / 2.0 + 0.5 + c * (length + 0.5) + 0.5 * LINK_PITCH
            z = -size_z / 2.0 + 1.0 + rz
            add_chain(sc, (x, y0 + layer, z), links)
            k += 1


def large_world(bodies: int = 100_000, chains: int = 2000, links: int = 16, seed: int = 0x5EED5) -> Scene:
    """100k mixed bodies (the pile generator over a 50 x 60 x 50 m region) plus 2,000
    chains of 16 capsule links (15 revolute joints each, collide_connected false) in layers
    above the pile; sleep and continuous collision off."""
    sc = chain_pile(bodies, chains, links, seed, 50, 60, 50)
    sc.name = f"large_world_{bodies}_{chains}x{links}"
    return sc


def joint_hub(satellites: int = 24) -> Scene:
    """A dynamic hub box with `satellites` small boxes hinged around it on a ring and two
    boxes dropped on top, over the ground: the hub's first 20 joints take colours 0..19 and
    the rest overflow (constraint_graph.rs:256-298), and so do the contacts of the boxes on
    top with the hub, so the overflow colour holds joints and contacts that share a body."""
    sc = Scene(name=f"joint_hub_{satellites}", enable_sleep=False)
    g = sc.add_body(BodySpec(STATIC, (0.0, -1.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (20.0, 1.0, 20.0)))
    hub = sc.add_body(BodySpec(DYNAMIC, (0.0, 0.5, 0.0)))
    sc.add_shape(ShapeSpec(hub, BOX, (0.5, 0.5, 0.5)))
    for x in (-0.25, 0.25):
        b = sc.add_body(BodySpec(DYNAMIC, (x, 1.3, 0.0)))
        sc.add_shape(ShapeSpec(b, BOX, (0.2, 0.2, 0.2)))
    for i in range(satellites):
        c, s = compute_cos_sin(f32(2.0 * np.pi * i / satellites))
        p = (float(f32(1.5) * c), 0.3, float(f32(1.5) * s))
        b = sc.add_body(BodySpec(DYNAMIC, p))
        sc.add_shape(ShapeSpec(b, BOX, (0.2, 0.2, 0.2)))
        # hinge point on the hub's rim, 0.2 m below its centre, as seen from each body
        rim = (float(f32(0.6) * c), -0.2, float(f32(0.6) * s))
        from_b = (float(f32(rim[0]) - f32(p[0])), 0.0, float(f32(rim[2]) - f32(p[2])))
        sc.add_joint(JointSpec(hub, b, frame_a=_frame(rim), frame_b=_frame(from_b)))
    return sc


def pushed_boxes(seed: int, boxes: int = 32) -> Scene:
    """A differentiable-rollout world: `boxes` cubes (half size U[0.2, 0.35] m, density 1000,
    friction U[0.3, 0.8]) resting on a ground box, 3 m apart on an 8-wide grid, pushed with a
    seeded horizontal velocity U[-2, 2] m/s per axis and a spin about the vertical U[-2, 2]
    rad/s. They slide to rest within the rollout without reaching each other or tipping, so
    every contact is a flat box-ground contact. Sleep and continuous collision are off."""
    rng = XorShift32(seed)
    cols = 8
    rows = -(-boxes // cols)
    sc = Scene(name=f"pushed_boxes_{seed}", enable_sleep=False, enable_continuous=False)
    _ground(sc, 1.5 * cols, 1.5 * rows)
    for i in range(boxes):
        h = rng.uniform(0.2, 0.35)
        x = -1.5 * cols + 1.5 + 3.0 * (i % cols)
        z = -1.5 * rows + 1.5 + 3.0 * (i // cols)
        v = (float(rng.uniform(-2.0, 2.0)), 0.0, float(rng.uniform(-2.0, 2.0)))
        w = (0.0, float(rng.uniform(-2.0, 2.0)), 0.0)
        friction = float(rng.uniform(0.3, 0.8))
        b = sc.add_body(BodySpec(DYNAMIC, (float(f32(x)), float(h), float(f32(z))), linear_velocity=v,
                                 angular_velocity=w, enable_sleep=False))
        sc.add_shape(ShapeSpec(b, BOX, (float(h), float(h), float(h)), density=1000.0, friction=friction))
    return sc


def pushed_box_targets(count: int, seed: int = 0x7A46) -> np.ndarray:
    """Seeded loss targets for the differentiable rollout: one centre-of-mass target per
    world, U[-1, 1] m in x and z about the origin, at height 0.25 m."""
    rng = XorShift32(seed)
    t = np.zeros((count, 3), np.float32)
    for k in range(count):
        t[k] = (rng.uniform(-1.0, 1.0), 0.25, rng.uniform(-1.0, 1.0))
    return t


def run_rollback(world, steps: int = 200, dt: float = 1.0 / 60.0, substeps: int = 4):
    """The reference's rollback benchmark driver (bin/benchmarks/feature_scenes.rs:702-734 and
    the harness loop bin/benchmarks.rs:334-360): before every step the state is saved (the
    last 8 kept); at every 30th step the oldest is restored and 7 steps replayed. With an
    exact save and restore the run ends on the plain run's state. `world` is any stepper
    with save_state, restore_state and step (the GPU world or the oracle)."""
    states = []
    for i in range(steps):
        states.append(world.save_state())
        if len(states) > 8:
            states.pop(0)
        if i % 30 == 29 and states:
            world.restore_state(states[0])
            for _ in range(7):
                world.step(dt, substeps)
            states.clear()
        world.step(dt, substeps)
    return world


# ---------------------------------------------------------------------------------
# Continuous collision: fast bodies and bullets against thin walls and floors


def ccd_wall(bullet: bool) -> Scene:
    """A 0.4 m box fired at 200 m/s at a 0.2 m static wall, one substep per step
    (ccd_tests.rs:30-55). The box is the last body."""
    sc = Scene(name=f"ccd_wall_{'bullet' if bullet else 'fast'}", enable_sleep=False, substeps=1)
    w = sc.add_body(BodySpec(STATIC, (0.0, 0.0, 0.0)))
    sc.add_shape(ShapeSpec(w, BOX, (0.1, 2.0, 2.0)))
    b = sc.add_body(BodySpec(DYNAMIC, (-5.0, 0.0, 0.0), linear_velocity=(200.0, 0.0, 0.0), is_bullet=bullet))
    sc.add_shape(ShapeSpec(b, BOX, (0.2, 0.2, 0.2), density=1000.0))
    return sc


def bullet_wall(offset: float = 0.0, enable_continuous: bool = True) -> Scene:
    """A 0.1 m sphere bullet at 200 m/s, gravity off, at a 0.1 m static wall 5 m away, with
    every position offset by `offset` (large_world_tests.rs:169-200;
    world_tuning_tests.rs:90-113 with continuous collision off). The bullet is the last body."""
    sc = Scene(name=f"bullet_wall_{offset:g}", enable_continuous=enable_continuous)
    w = sc.add_body(BodySpec(STATIC, (offset + 5.0, offset, offset)))
    sc.add_shape(ShapeSpec(w, BOX, (0.05, 5.0, 5.0)))
    b = sc.add_body(BodySpec(DYNAMIC, (offset, offset, offset), linear_velocity=(200.0, 0.0, 0.0), gravity_scale=0.0,
                             is_bullet=True))
    sc.add_shape(ShapeSpec(b, SPHERE, (0.0, 0.0, 0.0, 0.1), density=1.0))
    return sc


def overshoot(continuous: bool) -> Scene:
    """A perfectly elastic unit cube dropped flat from 10 m on a floor narrower than itself,
    four restitution iterations; with `continuous` its safety factor is 0.1 so it is swept
    onto the floor (measure_overshoot, restitution_phase_tests.rs:311-380). The cube is the
    last body."""
    sc = Scene(name=f"overshoot_{continuous}", enable_sleep=False, restitution_iterations=4)
    f = sc.add_body(BodySpec(STATIC, (0.0, -0.25, 0.0)))
    sc.add_shape(ShapeSpec(f, BOX, (0.375, 0.25, 0.375)))
    b = sc.add_body(BodySpec(DYNAMIC, (0.0, 10.0, 0.0), safety_factor=0.1 if continuous else 0.5))
    sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5), restitution=1.0))
    return sc


def fast_and_bullets(cubes: int = 640, shooters: int = 24) -> Scene:
    """A grid of cubes on a ground between two static walls, with small spheres fired at
    300 m/s across it, every third one a bullet (build_fast_and_sleepy,
    parallel_solver_tests.rs:278-326; sleeping is not modelled here, so bodies stay awake)."""
    sc = Scene(name="fast_and_bullets", enable_sleep=True)
    g = sc.add_body(BodySpec(STATIC, (0.0, -1.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (100.0, 1.0, 100.0)))
    for x in (-16.0, 16.0):
        w = sc.add_body(BodySpec(STATIC, (x, 4.0, 0.0)))
        sc.add_shape(ShapeSpec(w, BOX, (0.25, 4.0, 20.0)))
    for i in range(cubes):
        b = sc.add_body(BodySpec(DYNAMIC, (float((i % 16) - 7.5), 0.4 + (i // 256) * 0.8, float(((i // 16) % 16) - 7.5))))
        sc.add_shape(ShapeSpec(b, BOX, (0.4, 0.4, 0.4)))
    for i in range(shooters):
        v = (300.0 if i % 2 == 0 else -300.0, 0.0, float(i % 3) - 1.0)
        b = sc.add_body(BodySpec(DYNAMIC, (-12.0 if i % 2 == 0 else 12.0, 1.0 + (i // 2) * 0.5, float((i % 5) - 2)),
                                 linear_velocity=v, is_bullet=i % 3 == 0))
        sc.add_shape(ShapeSpec(b, SPHERE, (0.0, 0.0, 0.0, 0.1)))
    return sc


# ---------------------------------------------------------------------------------
# Islands and sleep


def box_sleeps() -> Scene:
    """A unit box resting on a static slab, sleep on (sleep_tests.rs:33-99). The box is the
    last body."""
    sc = Scene(name="box_sleeps", enable_sleep=True)
    g = sc.add_body(BodySpec(STATIC, (0.0, 0.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (10.0, 0.5, 10.0)))
    b = sc.add_body(BodySpec(DYNAMIC, (0.0, 0.6, 0.0)))
    sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5), density=1000.0))
    return sc


def island_split() -> Scene:
    """Two unit boxes stacked on a static slab, sleep off (island_split_tests.rs:46-76)."""
    sc = Scene(name="island_split", enable_sleep=False)
    g = sc.add_body(BodySpec(STATIC, (0.0, 0.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (10.0, 0.5, 10.0)))
    for i in range(2):
        b = sc.add_body(BodySpec(DYNAMIC, (0.0, 0.55 + i * 1.01, 0.0)))
        sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5), density=1000.0))
    return sc


def sleep_and_wake(drop_height: float = 8.0) -> Scene:
    """A box settles and falls asleep on the ground; a ball dropped from `drop_height` lands
    on it later and wakes it; both settle and sleep again. Bodies: ground, box, ball."""
    sc = Scene(name="sleep_and_wake", enable_sleep=True)
    g = sc.add_body(BodySpec(STATIC, (0.0, -0.5, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (10.0, 0.5, 10.0)))
    b = sc.add_body(BodySpec(DYNAMIC, (0.0, 0.5, 0.0)))
    sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5)))
    s = sc.add_body(BodySpec(DYNAMIC, (0.2, drop_height, 0.1)))
    sc.add_shape(ShapeSpec(s, SPHERE, (0.0, 0.0, 0.0, 0.25)))
    return sc


def roll_off() -> Scene:
    """A ball rolls off a resting box and away across the ground: the box-ball contact ends,
    the island holding both is split once the box is sleepy, and the box falls asleep while
    the ball still rolls. Bodies: ground, box, ball."""
    sc = Scene(name="roll_off", enable_sleep=True)
    g = sc.add_body(BodySpec(STATIC, (0.0, -0.5, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (20.0, 0.5, 20.0)))
    b = sc.add_body(BodySpec(DYNAMIC, (0.0, 0.5, 0.0)))
    sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5)))
    s = sc.add_body(BodySpec(DYNAMIC, (0.0, 1.25, 0.0), linear_velocity=(1.5, 0.0, 0.0)))
    sc.add_shape(ShapeSpec(s, SPHERE, (0.0, 0.0, 0.0, 0.25), friction=0.1))
    return sc


def jointed_sleepers(drop_height: float = 8.0) -> Scene:
    """Two boxes resting side by side on the ground, hinged together along their shared edge
    at the top (so one island holds both), fall asleep together; a ball dropped from
    `drop_height` onto the first later wakes both, and all settle and sleep again. Bodies:
    ground, box, box, ball; one joint."""
    sc = Scene(name="jointed_sleepers", enable_sleep=True)
    g = sc.add_body(BodySpec(STATIC, (0.0, -0.5, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (10.0, 0.5, 10.0)))
    a = sc.add_body(BodySpec(DYNAMIC, (-0.55, 0.5, 0.0)))
    sc.add_shape(ShapeSpec(a, BOX, (0.5, 0.5, 0.5)))
    b = sc.add_body(BodySpec(DYNAMIC, (0.55, 0.5, 0.0)))
    sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5)))
    sc.joints.append(JointSpec(a, b, frame_a=(0.55, 0.5, 0.0, 0.0, 0.0, 0.0, 1.0), frame_b=(-0.55, 0.5, 0.0, 0.0, 0.0, 0.0, 1.0)))
    s = sc.add_body(BodySpec(DYNAMIC, (-0.45, drop_height, 0.1)))
    sc.add_shape(ShapeSpec(s, SPHERE, (0.0, 0.0, 0.0, 0.25)))
    return sc


def sleepy_stacks(count: int, seed: int) -> Scene:
    """`stacks` with sleep on for the world and every body: towers topple, bodies come to rest
    in many islands that merge, split, fall asleep and are woken by falling neighbours."""
    sc = stacks(count, seed)
    sc.name = f"sleepy_stacks_{count}"
    sc.enable_sleep = True
    sc.bodies = [dataclasses.replace(b, enable_sleep=True) for b in sc.bodies]
    return sc


def pile_defaults(count: int = 10_000, seed: int = 0x5EED, **kw) -> Scene:
    """`pile` with the reference's defaults for the world and every body: sleep on and
    continuous collision on (falling bodies are swept; settled bodies fall asleep)."""
    sc = pile(count, seed, **kw)
    sc.name = f"pile_defaults_{count}"
    sc.enable_sleep = True
    sc.enable_continuous = True
    sc.bodies = [dataclasses.replace(b, enable_sleep=True) for b in sc.bodies]
    return sc


def kinematic_platform(boxes: int = 6) -> Scene:
    """A kinematic slab sliding and turning under a row of boxes, with a static ground below
    (kinematic bodies move at their set velocities and push dynamic bodies; no dynamic body
    pushes them back). Sleep off."""
    sc = Scene(name="kinematic_platform", enable_sleep=False)
    g = sc.add_body(BodySpec(STATIC, (0.0, -2.0, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (20.0, 0.5, 20.0)))
    k = sc.add_body(BodySpec(KINEMATIC, (0.0, 0.0, 0.0), linear_velocity=(0.5, 0.0, 0.0), angular_velocity=(0.0, 0.3, 0.0)))
    sc.add_shape(ShapeSpec(k, BOX, (4.0, 0.25, 2.0)))
    for i in range(boxes):
        b = sc.add_body(BodySpec(DYNAMIC, (-3.0 + 1.2 * i, 0.76, 0.0)))
        sc.add_shape(ShapeSpec(b, BOX, (0.5, 0.5, 0.5), friction=0.8))
    return sc


LOCK_LINEAR_Z, LOCK_ANGULAR_X, LOCK_ANGULAR_Y = 0x4, 0x8, 0x10
LOCK_ROTATION = 0x8 | 0x10 | 0x20


def def_features(enable_sleep: bool = True, enable_continuous: bool = True) -> Scene:
    """Every BodyDef and ShapeDef field the C ABI carries, in one scene on a static ground
    (body_def.rs:58-104, shape_def.rs:90-135). Named bodies (Scene.tags):
      roll_sphere / slide_sphere   spheres launched at 3 m/s, rolling resistance 0.3 and 0
      roll_capsule                 a capsule rolling across x, rolling resistance 0.2
      planar_*                     tilted boxes locked to the x-y plane (linear z, angular x, y)
      fixed_rotation               a box with every rotation locked, launched spinning
      group_a / group_b            boxes of filter group -1 dropped onto each other (never touch)
      mask_a / mask_b              category 2 boxes whose masks exclude category 2
      sensor_drop                  a ball falling through a static sensor box
      bullet_ball                  a ball at 60 m/s (continuous) through a thin static sensor wall
      fast_spin / clamped_spin     spin 300 rad/s above the clamp (0.25 pi per substep), with and
                                   without allow_fast_rotation
      no_recycle                   contact recycling off
      sleepy                       sleep threshold 0.5 m/s
      no_speculative / hit_events  shape flags
    """
    sc = Scene(name="def_features", enable_sleep=enable_sleep, enable_continuous=enable_continuous)
    g = sc.add_body(BodySpec(STATIC, (0.0, -0.5, 0.0)))
    sc.add_shape(ShapeSpec(g, BOX, (30.0, 0.5, 30.0)))

    def body(tag, pos, **kw):
        shape = kw.pop("shape")
        b = sc.add_body(BodySpec(DYNAMIC, pos, **kw))
        kind, params, skw = shape
        sc.add_shape(ShapeSpec(b, kind, params, **skw))
        sc.tags[tag] = b
        return b

    body("roll_sphere", (0.0, 0.3, -6.0), linear_velocity=(3.0, 0.0, 0.0),
         shape=(SPHERE, (0.0, 0.0, 0.0, 0.3), {"rolling_resistance": 0.3}))
    body("slide_sphere", (0.0, 0.3, -4.0), linear_velocity=(3.0, 0.0, 0.0), shape=(SPHERE, (0.0, 0.0, 0.0, 0.3), {}))
    body("roll_capsule", (0.0, 0.2, -2.0), linear_velocity=(2.0, 0.0, 0.0),
         shape=(CAPSULE, (0.0, 0.0, -0.4, 0.0, 0.0, 0.4, 0.2), {"rolling_resistance": 0.2}))
    tilt = (0.0, 0.0, float(np.sin(0.15)), float(np.cos(0.15)))
    for k in range(3):
        body(f"planar_{k}", (-6.0 + 0.3 * k, 0.6 + 1.1 * k, 0.0), rotation=tilt, motion_locks=LOCK_LINEAR_Z | LOCK_ANGULAR_X | LOCK_ANGULAR_Y,
             angular_velocity=(1.0, 2.0, 0.5), shape=(BOX, (0.5, 0.5, 0.5), {}))
    body("fixed_rotation", (-3.0, 0.5, 3.0), linear_velocity=(2.0, 0.0, 0.0), angular_velocity=(0.0, 3.0, 0.0),
         motion_locks=LOCK_ROTATION, shape=(BOX, (0.5, 0.5, 0.5), {}))
    body("group_a", (3.0, 0.5, 3.0), shape=(BOX, (0.5, 0.5, 0.5), {"group_index": -1}))
    body("group_b", (3.2, 1.6, 3.1), shape=(BOX, (0.5, 0.5, 0.5), {"group_index": -1}))
    body("mask_a", (6.0, 0.5, 3.0), shape=(BOX, (0.5, 0.5, 0.5), {"category_bits": 2, "mask_bits": 0xFFFF_FFFF_FFFF_FFFD}))
    body("mask_b", (6.1, 1.6, 3.0), shape=(BOX, (0.5, 0.5, 0.5), {"category_bits": 2, "mask_bits": 0xFFFF_FFFF_FFFF_FFFD}))
    s = sc.add_body(BodySpec(STATIC, (9.0, 2.0, 3.0)))
    sc.add_shape(ShapeSpec(s, BOX, (1.0, 0.5, 1.0), is_sensor=True))
    body("sensor_drop", (9.0, 4.0, 3.0), shape=(SPHERE, (0.0, 0.0, 0.0, 0.25), {}))
    wall = sc.add_body(BodySpec(STATIC, (14.0, 1.0, -6.0)))
    sc.add_shape(ShapeSpec(wall, BOX, (0.05, 1.0, 1.0), is_sensor=True))
    body("bullet_ball", (8.0, 1.0, -6.0), linear_velocity=(60.0, 0.0, 0.0), gravity_scale=0.0,
         shape=(SPHERE, (0.0, 0.0, 0.0, 0.1), {}))
    body("fast_spin", (0.0, 0.6, 6.0), angular_velocity=(0.0, 300.0, 0.0), allow_fast_rotation=True,
         shape=(BOX, (0.5, 0.5, 0.5), {}))
    body("clamped_spin", (3.0, 0.6, 6.0), angular_velocity=(0.0, 300.0, 0.0), shape=(BOX, (0.5, 0.5, 0.5), {}))
    body("no_recycle", (3.0, 0.5, -3.0), linear_velocity=(0.5, 0.0, 0.0), enable_contact_recycling=False,
         shape=(BOX, (0.5, 0.5, 0.5), {}))
    body("sleepy", (6.0, 0.5, -3.0), sleep_threshold=0.5, shape=(BOX, (0.5, 0.5, 0.5), {}))
    body("no_speculative", (-6.0, 0.5, -3.0), linear_velocity=(0.0, -1.0, 0.0),
         shape=(BOX, (0.5, 0.5, 0.5), {"enable_speculative_contact": False}))
    body("hit_events", (-3.0, 0.5, -3.0), shape=(SPHERE, (0.0, 0.0, 0.0, 0.5), {"enable_hit_events": True}))
    return sc
