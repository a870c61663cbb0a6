"""Batches of independent worlds over the HIP library (include/natsel_gpu.h, nsl_batch_*).

Each world of a batch is built with the reference's World calls (create_body,
create_hull_shape(make_box_hull), create_sphere_shape, create_capsule_shape) and steps
exactly as that World would alone (World::step, src/simulation/solver_integration.rs:807),
so its world_state_hash (src/simulation/body_queries.rs:432) matches. All worlds step
together on the device with no host work between steps.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi
from .scenes import BOX, SPHERE, Scene
from .world import (BodyDef, BoxHull, Capsule, ShapeDef, Sphere, WorldOptions, _material, body_def_of, c_body_def, c_options,
                    make_box_hull, shape_def_of)


class BatchWorld:
    """World `index` of a batch, with the reference World's creation calls."""

    def __init__(self, batch: "Batch", index: int):
        self._b = batch
        self.index = index

    def create_body(self, d: BodyDef) -> int:
        c = c_body_def(d)
        out = ctypes.c_int32()
        capi.check(self._b._lib.nsl_batch_create_body(self._b._h, self.index, ctypes.byref(c), ctypes.byref(out)))
        return out.value

    def create_hull_shape(self, body: int, d: ShapeDef, hull: BoxHull) -> int:
        out = ctypes.c_int32()
        h = (ctypes.c_float * 3)(hull.hx, hull.hy, hull.hz)
        capi.check(self._b._lib.nsl_batch_create_box(self._b._h, self.index, body, h, ctypes.byref(_material(d)), ctypes.byref(out)))
        return out.value

    def create_sphere_shape(self, body: int, d: ShapeDef, s: Sphere) -> int:
        out = ctypes.c_int32()
        c = (ctypes.c_float * 3)(*s.center)
        capi.check(self._b._lib.nsl_batch_create_sphere(self._b._h, self.index, body, c, s.radius, ctypes.byref(_material(d)),
                                                        ctypes.byref(out)))
        return out.value

    def create_capsule_shape(self, body: int, d: ShapeDef, c: Capsule) -> int:
        out = ctypes.c_int32()
        c1 = (ctypes.c_float * 3)(*c.center1)
        c2 = (ctypes.c_float * 3)(*c.center2)
        capi.check(self._b._lib.nsl_batch_create_capsule(self._b._h, self.index, body, c1, c2, c.radius,
                                                         ctypes.byref(_material(d)), ctypes.byref(out)))
        return out.value

    def body_arrays(self):
        return self._b.body_arrays(self.index)


class Batch:
    def __init__(self, device: int = 0, options: WorldOptions | None = None):
        self._lib = capi.load_library()
        self._opts = options or WorldOptions()
        handle = ctypes.c_void_p()
        capi.check(self._lib.nsl_batch_create(device, ctypes.byref(c_options(self._opts)), ctypes.byref(handle)))
        self._h = handle
        self.worlds: list[BatchWorld] = []

    def close(self) -> None:
        if self._h:
            capi.check(self._lib.nsl_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_world(self) -> BatchWorld:
        out = ctypes.c_int32()
        capi.check(self._lib.nsl_batch_add_world(self._h, ctypes.byref(out)))
        w = BatchWorld(self, out.value)
        self.worlds.append(w)
        return w

    def step(self, time_step: float, sub_step_count: int, steps: int = 1) -> None:
        capi.check(self._lib.nsl_batch_step(self._h, time_step, sub_step_count, steps))

    def state_hashes(self) -> np.ndarray:
        out = np.zeros(len(self.worlds), np.uint64)
        capi.check(self._lib.nsl_batch_state_hashes(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return out

    def body_arrays(self, world: int):
        n = ctypes.c_int32()
        capi.check(self._lib.nsl_batch_body_count(self._h, world, ctypes.byref(n)))
        n = n.value
        pos = np.zeros((n, 3), np.float64)
        rot = np.zeros((n, 4), np.float32)
        v = np.zeros((n, 3), np.float32)
        w = np.zeros((n, 3), np.float32)
        f = lambda a, t: a.ctypes.data_as(ctypes.POINTER(t))  # noqa: E731
        capi.check(self._lib.nsl_batch_get_bodies(self._h, world, f(pos, ctypes.c_double), f(rot, ctypes.c_float),
                                                  f(v, ctypes.c_float), f(w, ctypes.c_float)))
        return pos, rot, v, w

    def contact_colors(self, world: int):
        """Per contact slot of world `world`: colour (-1 when not in the graph), local index, liveness."""
        n = ctypes.c_int32()
        capi.check(self._lib.nsl_batch_contact_capacity(self._h, world, ctypes.byref(n)))
        color, local, alive = (np.zeros(n.value, np.int32) for _ in range(3))
        f = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        capi.check(self._lib.nsl_batch_get_contact_colors(self._h, world, f(color), f(local), f(alive)))
        return color, local, alive

    def contact_shapes(self, world: int):
        """Per contact slot of world `world`: its shape ids (primary first; -1 when free)."""
        n = ctypes.c_int32()
        capi.check(self._lib.nsl_batch_contact_capacity(self._h, world, ctypes.byref(n)))
        a, b = np.zeros(n.value, np.int32), np.zeros(n.value, np.int32)
        f = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        capi.check(self._lib.nsl_batch_get_contact_shapes(self._h, world, f(a), f(b)))
        return a, b

    def shape_bounds(self, world: int, n_shapes: int):
        """(fat AABBs [n, 6] as lower xyz, upper xyz; moved flags [n]) of world `world`."""
        fat = np.zeros((n_shapes, 6), np.float32)
        moved = np.zeros(n_shapes, np.uint8)
        capi.check(self._lib.nsl_batch_get_shape_bounds(self._h, world, fat.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                        moved.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return fat, moved

    # --- differentiable rollout ---------------------------------------------------------
    def record(self, time_step: float, sub_step_count: int, steps: int, max_constraints: int = 64) -> None:
        """Step `steps` steps like step() while recording the tape for adjoint()."""
        capi.check(self._lib.nsl_batch_record(self._h, time_step, sub_step_count, steps, max_constraints))

    def adjoint(self, targets: np.ndarray) -> dict:
        """Loss |com - target|^2 per world after the recorded steps and its gradients per
        dynamic body (rows in the batch's body order: each world's dynamic bodies in creation
        order, world after world) w.r.t. the first step's velocity, angular velocity and the
        body's friction."""
        targets = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        n = int(self.stats()["bodies"])
        loss = np.zeros(len(targets), np.float32)
        gv, gw = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32)
        gf = np.zeros(n, np.float32)
        status = np.zeros(len(targets), np.int32)
        f = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
        capi.check(self._lib.nsl_batch_adjoint(self._h, f(targets), f(loss), f(gv), f(gw), f(gf),
                                               status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return {"loss": loss, "grad_v": gv, "grad_w": gw, "grad_friction": gf, "status": status}

    def model_losses(self, worlds, params, deltas, targets: np.ndarray, frozen: bool = False) -> np.ndarray:
        """The adjoint's model in double precision, one perturbed parameter per case
        (param = body * 7 + {0..2: v, 3..5: w, 6: friction}); `frozen` takes every branch
        from the last adjoint()'s record."""
        worlds = np.ascontiguousarray(worlds, np.int32)
        params = np.ascontiguousarray(params, np.int32)
        deltas = np.ascontiguousarray(deltas, np.float64)
        targets = np.ascontiguousarray(targets, np.float32).reshape(-1, 3)
        out = np.zeros(len(worlds), np.float64)
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))  # noqa: E731
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
        capi.check(self._lib.nsl_batch_model_losses(self._h, ip(worlds), ip(params), dp(deltas), len(worlds),
                                                    targets.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(frozen), dp(out)))
        return out

    def rollout_times(self) -> tuple:
        """Device milliseconds of the last record() and adjoint() calls."""
        r, a = ctypes.c_double(), ctypes.c_double()
        capi.check(self._lib.nsl_batch_rollout_times(self._h, ctypes.byref(r), ctypes.byref(a)))
        return r.value, a.value

    @staticmethod
    def _tuning_dict(t) -> dict:
        return {"collide_blocks": t.collide_blocks, "use_graph": bool(t.use_graph), "candidates": t.candidates,
                "step_ms": t.step_ms, "previous_ms": t.previous_ms}

    def tuning(self) -> dict:
        """The batched step's launch shape (results do not depend on it)."""
        t = capi.BatchTuning()
        capi.check(self._lib.nsl_batch_get_tuning(self._h, ctypes.byref(t)))
        return self._tuning_dict(t)

    def set_tuning(self, collide_blocks: int = 0, use_graph: bool = False) -> None:
        """collide_blocks: workgroups of the collide launch (0: as many as are resident);
        use_graph: replay a captured HIP graph of 8 steps."""
        t = capi.BatchTuning(collide_blocks=int(collide_blocks), use_graph=int(bool(use_graph)))
        capi.check(self._lib.nsl_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, time_step: float, sub_step_count: int, steps: int = 8) -> dict:
        """Time each candidate launch shape over `steps` steps from a snapshot of the batch,
        check that all reach the same world hashes, restore the snapshot and keep the
        fastest shape (nsl_batch_autotune)."""
        t = capi.BatchTuning()
        capi.check(self._lib.nsl_batch_autotune(self._h, time_step, sub_step_count, steps, ctypes.byref(t)))
        return self._tuning_dict(t)

    def stats(self) -> dict:
        s = capi.BatchStats()
        capi.check(self._lib.nsl_batch_get_stats(self._h, ctypes.byref(s)))
        return {n: getattr(s, n) for n, _ in capi.BatchStats._fields_}


def load_batch(scene_list: list[Scene], device: int = 0) -> Batch:
    """One batch world per scene. The worlds of a batch share their options, so the scenes
    must agree on sleep, continuous collision, time step and substeps."""
    first = scene_list[0]
    for s in scene_list:
        key = lambda x: (x.enable_sleep, x.enable_continuous, x.dt, x.substeps, tuple(x.gravity), x.restitution_iterations,  # noqa: E731
                         x.restitution_threshold, x.restitution_propagation)
        if key(s) != key(first):
            raise ValueError(f"scene {s.name} differs from {first.name} in its world options (sleep, continuous collision, "
                             "dt, substeps, gravity or restitution settings)")
    b = Batch(device, WorldOptions(enable_sleep=first.enable_sleep, enable_continuous=first.enable_continuous,
                                   gravity=tuple(first.gravity), restitution_iterations=first.restitution_iterations,
                                   restitution_threshold=first.restitution_threshold,
                                   enable_restitution_propagation=first.restitution_propagation))
    for scene in scene_list:
        w = b.add_world()
        ids = [w.create_body(body_def_of(s)) for s in scene.bodies]
        for s in scene.shapes:
            d = shape_def_of(s)
            if s.kind == BOX:
                w.create_hull_shape(ids[s.body], d, make_box_hull(*s.params))
            elif s.kind == SPHERE:
                w.create_sphere_shape(ids[s.body], d, Sphere(tuple(s.params[:3]), s.params[3]))
            else:
                w.create_capsule_shape(ids[s.body], d, Capsule(tuple(s.params[:3]), tuple(s.params[3:6]), s.params[6]))
    return b
