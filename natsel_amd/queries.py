"""Collision queries on the device: time of impact (src/collision/time_of_impact.rs:406-598).

A query is two point proxies (1 to 8 support points and a radius) and their sweeps; a sweep
is (local_center, c1, c2, q1, q2) with quaternions as (x, y, z, s). Queries run one lane
each through the same device code the continuous-collision pass of a step uses.
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
