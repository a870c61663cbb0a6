"""Collision queries on the device.

time_of_impact (src/collision/time_of_impact.rs:406-598): a query is two point proxies (1 to
8 support points and a radius) and their sweeps; a sweep is (local_center, c1, c2, q1, q2)
with quaternions as (x, y, z, s).

collide / Narrow.collide: the narrow phase's collide routines (convex_manifold.rs:1419-1437,
convex_manifold_shapes.rs:21-797) on shape pairs in the register table's order (hull, then
capsule, then sphere). A shape is ("box", [hx, hy, hz, px, py, pz, qx, qy, qz, qs]) (the box
placed by p, q in its shape frame), ("sphere", [cx, cy, cz, r]) or ("capsule", [x1, y1, z1, x2,
y2, z2, r]); a transform is ((px, py, pz), (qx, qy, qz, qs)), B in A's frame.

Queries run one lane each through the same device code the step uses.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi

TOI_STATES = ("unknown", "failed", "overlapped", "hit", "separated")


def _sweep(values) -> np.ndarray:
    s = np.concatenate([np.asarray(x, np.float32).ravel() for x in values])
    if s.size != 17:
        raise ValueError("a sweep is local_center, c1, c2 (3 each) and q1, q2 (4 each)")
    return s


def time_of_impact(queries, device: int = 0) -> list:
    """`queries`: iterable of (points_a, radius_a, sweep_a, points_b, radius_b, sweep_b,
    max_fraction). Returns one dict per query: state, point, normal, fraction, distance and
    the distance, push-back and root iteration counts."""
    queries = list(queries)
    n = len(queries)
    inputs = (capi.ToiInput * max(n, 1))()
    for k, (pa, ra, sa, pb, rb, sb, max_fraction) in enumerate(queries):
        pa = np.asarray(pa, np.float32).reshape(-1, 3)
        pb = np.asarray(pb, np.float32).reshape(-1, 3)
        if not (1 <= len(pa) <= 8 and 1 <= len(pb) <= 8):
            raise ValueError("a proxy has 1 to 8 points")
        q = inputs[k]
        for i, p in enumerate(pa):
            q.points_a[i][:] = [float(x) for x in p]
        for i, p in enumerate(pb):
            q.points_b[i][:] = [float(x) for x in p]
        q.count_a, q.count_b = len(pa), len(pb)
        q.radius_a, q.radius_b = float(ra), float(rb)
        q.sweep_a[:] = [float(x) for x in _sweep(sa)]
        q.sweep_b[:] = [float(x) for x in _sweep(sb)]
        q.max_fraction = float(max_fraction)
    outputs = (capi.ToiOutput * max(n, 1))()
    lib = capi.load_library()
    capi.check(lib.nsl_time_of_impact(int(device), inputs, outputs, n))
    result = []
    for k in range(n):
        o = outputs[k]
        result.append({"state": TOI_STATES[o.state], "point": np.array(o.point, np.float32),
                       "normal": np.array(o.normal, np.float32), "fraction": np.float32(o.fraction),
                       "distance": np.float32(o.distance), "distance_iterations": o.distance_iterations,
                       "push_back_iterations": o.push_back_iterations, "root_iterations": o.root_iterations})
    return result


_SHAPE_TYPES = {"box": capi.SHAPE_HULL, "sphere": capi.SHAPE_SPHERE, "capsule": capi.SHAPE_CAPSULE}
_IDENTITY7 = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)


class CollideCache:
    """The cache a contact carries between collide calls: SATCache (box-box: separation, type,
    index_a, index_b, hit) or SimplexCache (box-sphere, box-capsule: metric, count, index_a[4],
    index_b[4]). `f` and `i` hold (value) and (kind, index_a[4], index_b[4]) like the oracle's."""

    def __init__(self):
        self.f = np.zeros(1, np.float32)
        self.i = np.zeros(9, np.int32)
        self.hit = 0

    @property
    def sat(self) -> dict:
        return {"separation": np.float32(self.f[0]), "type": int(self.i[0]), "index_a": int(self.i[1]),
                "index_b": int(self.i[5]), "hit": int(self.hit)}

    @property
    def simplex(self) -> dict:
        return {"metric": np.float32(self.f[0]), "count": int(self.i[0]), "index_a": self.i[1:5].tolist(),
                "index_b": self.i[5:9].tolist()}


def _fill_geometry(g, shape) -> None:
    kind, values = shape
    values = np.asarray(values, np.float32)
    g.type = _SHAPE_TYPES[kind]
    if kind == "box":
        g.params[:] = [float(values[0]), float(values[1]), float(values[2]), 0.0, 0.0, 0.0, 0.0]
        g.frame[:] = [float(x) for x in values[3:10]]
    else:
        g.params[:] = [float(x) for x in values] + [0.0] * (7 - len(values))
        g.frame[:] = _IDENTITY7


def collide(queries, device: int = 0) -> list:
    """`queries`: iterable of (shape_a, shape_b, transform, cache or None). Runs every call in
    one launch and returns per query a dict with normal, points [(point, separation, feature
    id)] and the updated cache (the one passed in, updated in place, or a new one)."""
    queries = list(queries)
    n = len(queries)
    arr = (capi.CollideQuery * max(n, 1))()
    caches = []
    for k, (a, b, xf, cache) in enumerate(queries):
        q = arr[k]
        _fill_geometry(q.a, a)
        _fill_geometry(q.b, b)
        q.transform_b_to_a[:] = [float(x) for x in (*xf[0], *xf[1])]
        cache = cache if cache is not None else CollideCache()
        caches.append(cache)
        q.cache.value = float(cache.f[0])
        q.cache.kind = int(cache.i[0])
        q.cache.index_a[:] = [int(x) for x in cache.i[1:5]]
        q.cache.index_b[:] = [int(x) for x in cache.i[5:9]]
        q.cache.hit = int(cache.hit)
    out = (capi.LocalManifold * max(n, 1))()
    capi.check(capi.load_library().nsl_collide(int(device), arr, out, n))
    result = []
    for k in range(n):
        m, c = out[k], caches[k]
        c.f[0] = m.cache.value
        c.i[0] = m.cache.kind
        c.i[1:5] = list(m.cache.index_a)
        c.i[5:9] = list(m.cache.index_b)
        c.hit = m.cache.hit
        pts = [(np.array(m.point[j], np.float32), np.float32(m.separation[j]), int(m.feature_id[j]))
               for j in range(m.point_count)]
        result.append({"normal": np.array(m.normal, np.float32), "points": pts, "cache": c})
    return result


class Narrow:
    """One collide call at a time on the device (the interface of the oracle's narrow-phase
    hooks: collide(a, b, xf, cache) -> normal, points [(point, separation, pair, feature id)])."""

    def __init__(self, device: int = 0):
        self.device = device

    def cache(self) -> CollideCache:
        return CollideCache()

    def collide(self, a, b, xf=((0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0)), cache=None) -> dict:
        r = collide([(a, b, xf, cache)], self.device)[0]
        r["points"] = [(p, s, _pair_of(fid), fid) for p, s, fid in r["points"]]
        return r


def _pair_of(feature_id: int) -> tuple:
    """(owner1, index1, owner2, index2) of a feature id (manifold.rs:58-94)."""
    return ((feature_id >> 24) & 0xFF, (feature_id >> 16) & 0xFF, (feature_id >> 8) & 0xFF, feature_id & 0xFF)
