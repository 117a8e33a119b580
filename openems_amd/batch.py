"""Batched engine: many independent simulations advanced together.

Each scene gets its own Operator, EngineHIP and probes (exactly as
:class:`~openems_amd.simulation.Simulation` sets them up); the batch then
advances all of them with one persistent launch per half step over a work
table of (instance, tile) items (include/ems_hip.h, ``ems_batch_*``).  Tile
shape, z-sweep depth and grid size come from the library's autotuner, which
times the candidates on the batch itself and restores the state afterwards.
Every instance is bit-identical to running its engine alone.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import NativeError, check, lib, up
from .simulation import Simulation


class BatchEngine:
    def __init__(self, scenes, device=0, layout="compressed", nthreads=0):
        """scenes: list of scene dicts; layout: coefficient layout of every
        instance ("compressed" falls back to "plain" for the whole batch when
        one instance cannot be compressed)."""
        self.sims = []
        self.b = None
        try:
            self._build(scenes, device, layout, nthreads)
        except NativeError:
            if layout != "compressed":
                raise
            self.close()
            self._build(scenes, device, "plain", nthreads)
        self.layout = layout

    def _build(self, scenes, device, layout, nthreads):
        self.sims = []
        for s in scenes:
            self.sims.append(Simulation(s, device=device, nthreads=nthreads, layout=layout, fused=False))
        L = lib()
        hs = (C.c_void_p * len(self.sims))(*[sim.eng.h.value for sim in self.sims])
        b = C.c_void_p()
        check(L.ems_batch_create(C.byref(b), hs, len(self.sims)))
        self.b = b

    @property
    def count(self):
        return len(self.sims)

    def autotune(self, steps=6):
        """Pick tile shape, z depths and grid by timing on the batch; returns ms per step."""
        ms = C.c_double()
        check(lib().ems_batch_autotune(self.b, int(steps), C.byref(ms)))
        return ms.value

    def set_config(self, tile_x=0, zdepth=None, grid=0):
        zd = None
        if zdepth is not None:
            zd = np.ascontiguousarray(zdepth, np.uint32)
            zd_p = zd.ctypes.data_as(up)
        else:
            zd_p = None
        check(lib().ems_batch_set_config(self.b, int(tile_x), zd_p, int(grid)))

    def set_grouping(self, group_bytes):
        """Working-set budget of a group of instances stepped together (0: one group)."""
        check(lib().ems_batch_set_grouping(self.b, int(group_bytes)))

    def config(self):
        tx, grid, nitems = C.c_uint(), C.c_uint(), C.c_uint()
        zd = np.zeros(self.count, np.uint32)
        check(lib().ems_batch_get_config(self.b, C.byref(tx), zd.ctypes.data_as(up), C.byref(grid),
                                         C.byref(nitems)))
        gb, ng = C.c_ulonglong(), C.c_uint()
        check(lib().ems_batch_get_grouping(self.b, C.byref(gb), C.byref(ng)))
        return {"tile_x": tx.value, "zdepth": zd.tolist(), "grid": grid.value, "work_items": nitems.value,
                "group_bytes": gb.value, "groups": ng.value}

    def set_graph(self, steps):
        """Replay each group's launches of `steps` steps as a captured HIP
        graph (0: direct launches)."""
        check(lib().ems_batch_set_graph(self.b, int(steps)))

    def graph(self):
        v = C.c_uint()
        check(lib().ems_batch_get_graph(self.b, C.byref(v)))
        return v.value

    def set_streams(self, n):
        """Run the groups on n streams (1..4) so that several advance at once."""
        check(lib().ems_batch_set_streams(self.b, int(n)))

    def streams(self):
        v = C.c_uint()
        check(lib().ems_batch_get_streams(self.b, C.byref(v)))
        return v.value

    def iterate(self, nsteps):
        check(lib().ems_batch_iterate(self.b, int(nsteps)))

    def last_ms(self):
        ms = C.c_double()
        check(lib().ems_batch_last_ms(self.b, C.byref(ms)))
        return ms.value

    def results(self):
        return [sim.results() for sim in self.sims]

    def close(self):
        if self.b is not None:
            lib().ems_batch_destroy(self.b)
            self.b = None
        for sim in self.sims:
            sim.close()
        self.sims = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
