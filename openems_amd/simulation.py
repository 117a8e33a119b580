"""Simulation driver: the timestep loop of openEMS::RunFDTD on the HIP engine.

``Simulation(scene).run()`` builds the operator (Operator), creates the
engine with its extensions (EngineHIP), registers the probes and dumps, and
advances NrTS steps, sampling at numTS = 0 and after every step that a probe
is due (openems.cpp:1388-1494).  For a periodic excitation it sets up the
steady-state detection of openems.cpp:1205-1234; with
scene["end_criteria"] > 0 the run stops once its last difference falls to the
criterion (checked every period, where the reference checks every few
seconds of wall time).  Without a periodic excitation the criterion is the
reference's energy decay (field energy over its largest sample), checked at a
fixed step interval.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

from .engine import EngineHIP
from .operator import Operator
from .processing import FieldDumpFD, FieldDumpTD, Probe, make_probe


class Simulation:
    def __init__(self, scene, device=0, nthreads=0, layout="auto", fused=None, graph=None):
        self.scene = scene
        self.nrts = int(scene.get("nrts", 1000))
        t0 = time.perf_counter()
        self.op = Operator(scene, nthreads=nthreads)
        self.setup_s = time.perf_counter() - t0
        self.ss_period = 0
        ss = None
        if self.op.period > 0 and scene.get("steady_state", True):
            ss = steady_state_probes(self.op)
            self.ss_period = ss[0]
        self.eng = EngineHIP(self.op, device=device, layout=layout, fused=fused, steady_state=ss)
        if graph is not None:
            self.eng.set_graph(graph)
        os_ = int(scene.get("oversampling", 4))
        self.probes = [make_probe(self.op, p, self.nrts, os_) for p in scene.get("probes", [])]
        self.dumps = [FieldDumpTD(self.op, d, self.nrts, os_) if d.get("domain") == "td"
                      else FieldDumpFD(self.op, d, self.nrts) for d in scene.get("dumps", [])]
        for p in self.probes:
            p.register(self.eng)
        for d in self.dumps:
            d.register(self.eng)

    def SetAbort(self, val=True):
        """openEMS::SetAbort (openems.h:103): a running or later run stops
        gracefully at its next abort check."""
        self._abort = bool(val)

    def check_abort(self):
        """openEMS::CheckAbortCond (openems.cpp:1364-1385): the abort flag, or a
        file named ABORT in the working directory."""
        if getattr(self, "_abort", False):
            return True
        if os.path.exists("ABORT"):
            print('Simulation.check_abort: found file "ABORT", aborting simulation gracefully...', file=sys.stderr)
            return True
        return False

    def run(self, nsteps=None, chunk=0):
        """Advance nsteps (default NrTS) in chunks, checking the end criterion
        and the abort conditions (CheckAbortCond: SetAbort, an ABORT file,
        SIGINT) between chunks of at most scene["abort_check"] steps (default
        1000) as RunFDTD's loop does (openems.cpp:1427-1480).  self.aborted
        tells whether the run stopped early on an abort."""
        n = self.nrts if nsteps is None else int(nsteps)
        end = float(self.scene.get("end_criteria", 0.0))
        self.aborted = False
        if end > 0 and not self.ss_period:
            return self._run_energy(n, end)
        if end > 0 and self.ss_period:
            chunk = self.ss_period if chunk <= 0 else min(chunk, self.ss_period)
        every = max(1, int(self.scene.get("abort_check", 1000)))
        chunk = every if chunk <= 0 else min(chunk, every)
        done = 0
        try:
            while done < n:
                if self.check_abort():
                    self.aborted = True
                    break
                k = min(chunk, n - done)
                self.eng.IterateTS(k)
                done += k
                if end > 0 and self.ss_period and self.eng.steady_state(1)[0] <= end:
                    break
        except KeyboardInterrupt:
            print("Simulation.run: received SIGINT, aborting simulation gracefully...", file=sys.stderr)
            self.aborted = True
        return self

    def _run_energy(self, n, end):
        """The energy end criterion of openEMS::RunFDTD (openems.cpp:1424-1480):
        the field energy (CalcFastEnergy) is sampled at the step of the
        excitation's largest amplitude (Excitation::GetMaxExcitationTimestep,
        FDTD/excitation.cpp:135-148) and then every scene["energy_check"]
        steps (default 1000; the reference samples every 4 s of wall time);
        the run ends once energy / its largest sample falls to end or below."""
        check = max(1, int(self.scene.get("energy_check", 1000)))
        sig_v, _ = self.op.signal()
        amp = np.abs(np.asarray(sig_v, np.float32))
        max_step = int(np.argmax(amp)) if amp.size and amp.max() > 0 else 0
        start = self.eng.numTS
        max_e, change = 0.0, 1.0
        self.stopped_at = None
        while self.eng.numTS - start < n and change > end:
            if self.check_abort():
                self.aborted = True
                break
            ts = self.eng.numTS
            nxt = min(start + n, (ts // check + 1) * check)
            if ts < max_step:
                nxt = min(nxt, max_step)
            self.eng.IterateTS(nxt - ts)
            ts = self.eng.numTS
            if ts == max_step or ts % check == 0:
                e = self.eng.energy()
                max_e = max(max_e, e)
                if ts % check == 0 and max_e:
                    change = e / max_e
        if change <= end:
            self.stopped_at = self.eng.numTS
        return self

    def results(self):
        out = {"probes": {}, "dumps": {}, "dT": self.op.dT, "numTS": self.eng.numTS}
        for p in self.probes:
            td, fd = p.collect(self.eng)
            out["probes"][p.name] = {"td": td, "fd": fd, "freqs": p.freqs}
        for d in self.dumps:
            if isinstance(d, FieldDumpTD):
                ts, t, fields = d.collect(self.eng)
                out["dumps"][d.name] = {"ts": ts, "time": t, "td": fields}
            else:
                out["dumps"][d.name] = {"fd": d.collect(self.eng), "freqs": d.freqs}
        return out

    def write_probes(self, directory):
        """Probe text files in the reference's layout (one per probe, plus
        <name>_FD where the probe has frequencies), as openEMS leaves them in
        its simulation folder; returns the paths written."""
        paths = []
        for p in self.probes:
            td, fd = p.collect(self.eng)
            paths += p.write_files(td, fd, directory)
        return paths

    def write_dumps(self, directory):
        """Field dumps as the reference writes them: FileType 1 as <name>.h5
        (openems_amd.hdf5_out), FileType 0 as VTK RectilinearGrid files
        (openems_amd.vtk_out); returns the paths written."""
        import os
        from . import hdf5_out, vtk_out
        paths = []
        for d in self.dumps:
            if not getattr(d, "enabled", True):
                continue
            ftype = int(d.spec.get("file_type", 1))
            if ftype not in (0, 1):
                raise ValueError(f"dump {d.name!r}: file type {ftype} is not supported")
            base = os.path.join(directory, d.name)
            if isinstance(d, FieldDumpTD):
                ts, t, fields = d.collect(self.eng)
                if fields is None:
                    continue
                if ftype == 1:
                    paths.append(hdf5_out.write_td_dump(base + ".h5", self.op, d, ts, t, fields))
                else:
                    paths += vtk_out.write_td_dump(base, self.op, d, ts, fields)
            else:
                fd = d.collect(self.eng)
                if fd is None:
                    continue
                if ftype == 1:
                    paths.append(hdf5_out.write_fd_dump(base + ".h5", self.op, d, fd, d.freqs))
                else:
                    paths += vtk_out.write_fd_dump(base, self.op, d, fd, d.freqs)
        return paths

    def close(self):
        self.eng.close()
        self.op.close()


def steady_state_probes(op):
    """Steady-state detection of openems.cpp:1205-1234: period
    m_TS_period = round(T / dT) (Operator_Ext_SteadyState::BuildExtension,
    operator_ext_steadystate.cpp:50-56) and E probes in all three directions
    at the mesh centre, then, for each axis n, at the centre with pos[n]
    multiplied by 1/4 and by 3/4 -- integer divisions, so pos[n] = 0 both
    times (21 probes)."""
    period = int(round(op.period / op.dT))
    dirs, pos = [], []
    centre = [op.N[k] // 2 for k in range(3)]
    for d in range(3):
        dirs.append(d)
        pos.append(list(centre))
    for n in range(3):
        p = list(centre)
        p[n] *= 1 // 4
        for d in range(3):
            dirs.append(d)
            pos.append(list(p))
        p[n] *= 3 // 4
        for d in range(3):
            dirs.append(d)
            pos.append(list(p))
    return period, dirs, pos
