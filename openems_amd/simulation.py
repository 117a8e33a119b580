"""Simulation driver: the timestep loop of openEMS::RunFDTD on the HIP engine.

``Simulation(scene).run()`` builds the operator (Operator), creates the
engine with its extensions (EngineHIP), registers the probes and dumps, and
advances NrTS steps with the end criterion disabled, sampling at numTS = 0
and after every step that a probe is due (openems.cpp:1388-1494).
"""
from __future__ import annotations

import time

from .engine import EngineHIP
from .operator import Operator
from .processing import FieldDumpFD, Probe


class Simulation:
    def __init__(self, scene, device=0, nthreads=0, layout="auto", fused=None):
        self.scene = scene
        self.nrts = int(scene.get("nrts", 1000))
        t0 = time.perf_counter()
        self.op = Operator(scene, nthreads=nthreads)
        self.setup_s = time.perf_counter() - t0
        self.eng = EngineHIP(self.op, device=device, layout=layout, fused=fused)
        os_ = int(scene.get("oversampling", 4))
        self.probes = [Probe(self.op, p, self.nrts, os_) for p in scene.get("probes", [])]
        self.dumps = [FieldDumpFD(self.op, d, self.nrts) for d in scene.get("dumps", [])]
        for p in self.probes:
            p.register(self.eng)
        for d in self.dumps:
            d.register(self.eng)

    def run(self, nsteps=None, chunk=0):
        n = self.nrts if nsteps is None else int(nsteps)
        done = 0
        while done < n:
            k = n - done if chunk <= 0 else min(chunk, n - done)
            self.eng.IterateTS(k)
            done += k
        return self

    def results(self):
        out = {"probes": {}, "dumps": {}, "dT": self.op.dT, "numTS": self.eng.numTS}
        for p in self.probes:
            td, fd = p.collect(self.eng)
            out["probes"][p.name] = {"td": td, "fd": fd, "freqs": p.freqs}
        for d in self.dumps:
            out["dumps"][d.name] = {"fd": d.collect(self.eng), "freqs": d.freqs}
        return out

    def close(self):
        self.eng.close()
        self.op.close()
