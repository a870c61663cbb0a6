"""nsl_collide_noalloc (staging in the library's fixed device array) against nsl_collide
(staging allocated per call), through the C ABI: two calls in a row, the second smaller,
each byte for byte equal to nsl_collide's output for the same queries; a call above the
array's 8192-query limit is refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _queries(n, seed):
    from natsel_amd import capi

    rng = np.random.default_rng(seed)
    kinds = [(capi.SHAPE_HULL, capi.SHAPE_HULL), (capi.SHAPE_HULL, capi.SHAPE_SPHERE), (capi.SHAPE_HULL, capi.SHAPE_CAPSULE),
             (capi.SHAPE_CAPSULE, capi.SHAPE_SPHERE), (capi.SHAPE_CAPSULE, capi.SHAPE_CAPSULE), (capi.SHAPE_SPHERE, capi.SHAPE_SPHERE)]
    arr = (capi.CollideQuery * n)()
    for k in range(n):
        q = arr[k]
        q.a.type, q.b.type = kinds[k % len(kinds)]
        for g in (q.a, q.b):
            if g.type == capi.SHAPE_HULL:
                g.params[:] = [float(x) for x in rng.uniform(0.2, 0.6, 3)] + [0.0] * 4
            elif g.type == capi.SHAPE_SPHERE:
                g.params[:] = [float(x) for x in rng.uniform(-0.1, 0.1, 3)] + [float(rng.uniform(0.1, 0.5))] + [0.0] * 3
            else:
                g.params[:] = [float(x) for x in rng.uniform(-0.4, 0.4, 6)] + [float(rng.uniform(0.1, 0.3))]
            g.frame[:] = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0]
        quat = rng.normal(size=4)
        quat /= np.linalg.norm(quat)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        q.transform_b_to_a[:] = [float(x) for x in d * rng.uniform(0.0, 1.2)] + [float(x) for x in quat]
    return arr


def test_noalloc_equals_collide_on_two_calls(gpu_available):
    from natsel_amd import capi

    lib = capi.load_library()
    for n, seed in ((3000, 1), (500, 2)):  # the second call is smaller than the first
        q = _queries(n, seed)
        want = (capi.LocalManifold * n)()
        got = (capi.LocalManifold * n)()
        capi.check(lib.nsl_collide(0, q, want, n))
        capi.check(lib.nsl_collide_noalloc(0, q, got, n))
        assert ctypes.string_at(got, ctypes.sizeof(got)) == ctypes.string_at(want, ctypes.sizeof(want)), f"n={n}"
        assert sum(got[k].point_count for k in range(n)) > n // 2


def test_noalloc_refuses_more_than_its_array(gpu_available):
    from natsel_amd import capi

    lib = capi.load_library()
    n = 8193
    q = (capi.CollideQuery * n)()
    for k in range(n):
        q[k].a.type = q[k].b.type = capi.SHAPE_SPHERE
    out = (capi.LocalManifold * n)()
    assert lib.nsl_collide_noalloc(0, q, out, n) == -1
    assert b"8192" in lib.nsl_last_error()
