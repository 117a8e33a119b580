"""Processing: probes and frequency-domain dumps evaluated by the engine.

Host-side mirror of the reference's Processing classes.  The snapping and the
order of the terms of every integral are fixed here, exactly as the
reference's probe classes visit them, and handed to the device, which
evaluates them at each sampling step (ems_add_probe); the time-domain rows and
the running DFT are then assembled here in the reference's operation order.

* ProcessVoltage            Common/processvoltage.cpp:37-47 ->
                            Engine_Interface_FDTD::CalcVoltageIntegral (FDTD/engine_interface_fdtd.cpp:206-232)
* ProcessCurrent            Common/processcurrent.cpp:42-171
* ProcessFieldProbe         Common/processfieldprobe.cpp:77-92, GetRawField (FDTD/engine_interface_fdtd.cpp:263-300)
* ProcessIntegral::Process  Common/processintegral.cpp:134-171 (TD rows, DFT)
* Processing scheduling     Common/processing.cpp:82-170
* ProcessFieldsFD           Common/processfields_fd.cpp:72-107 (E field, node interpolation)
"""
from __future__ import annotations

import cmath
import ctypes
import ctypes.util
import math

import numpy as np

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.cosf.restype = ctypes.c_float
_libm.cosf.argtypes = [ctypes.c_float]
_libm.sinf.restype = ctypes.c_float
_libm.sinf.argtypes = [ctypes.c_float]

PROBE_NAMES = {0: "voltage", 1: "current", 2: "E-field", 3: "H-field"}
# GetProcessingName / GetIntegralName of ProcessVoltage, ProcessCurrent and
# ProcessFieldProbe (Common/processvoltage.h:30, processvoltage.cpp:30-35,
# processcurrent.h:29, processcurrent.cpp:35-40, processfieldprobe.cpp:32-64)
PROCESSING_NAMES = {0: "voltage integration", 1: "current integration", 2: "electric field probe",
                    3: "magnetic field probe"}
INTEGRAL_NAMES = {0: ["voltage"], 1: ["current"], 2: ["Ex/(V/m)", "Ey/(V/m)", "Ez/(V/m)"],
                  3: ["Hx/(A/m)", "Hy/(A/m)", "Hz/(A/m)"]}


def _g(v, prec=6):
    """A double as a C++ ostream prints it at the given precision (%g form)."""
    return "%.*g" % (prec, v)


def process_window(spec, dT):
    """Processing::SetProcessStartStopTime (Common/processing.cpp:211-226):
    steps [floor(start/dT), ceil(stop/dT)] from spec["start_time"] /
    spec["stop_time"] (seconds, 0 = open); an empty window means the whole run."""
    t0, t1 = float(spec.get("start_time", 0.0)), float(spec.get("stop_time", 0.0))
    a, b = 0, 0xFFFFFFFF
    if t0 > 0:
        a = int(math.floor(t0 / dT))
    if t1 > 0:
        b = int(math.ceil(t1 / dT))
    if b <= a:
        a, b = 0, 0xFFFFFFFF
    return a, b


def nyquist_num(f, dT):
    """CalcNyquistNum (tools/useful.cpp:30-36)."""
    if f == 0:
        return 0xFFFFFFFF
    return int(math.floor(1 / f / 2 / dT))


class Probe:
    """One ProcessIntegral-family probe (type 0 voltage, 1 current, 2 E point, 3 H point)."""

    def __init__(self, op, spec, nrts, oversampling=4):
        self.op = op
        self.spec = spec
        self.name = spec.get("name", "")
        self.ptype = int(spec["type"])
        self.weight = float(spec.get("weight", 1.0))
        self.dual = self.ptype in (1, 3)
        nyq = op.nyquist
        self.interval = max(1, nyq // int(oversampling))
        self.fd_interval = 0
        self.freqs = []
        for f in spec.get("freqs", []):          # Processing::AddFrequency
            nts = nyquist_num(f, op.dT)
            if nts == 0:
                continue
            if self.fd_interval == 0:
                self.fd_interval = nyq
            if self.fd_interval > nts:
                self.fd_interval = nts
            self.freqs.append(float(f))
        self.start_ts, self.stop_ts = process_window(spec, op.dT)
        self.enabled = True
        self._terms()
        self.steps = [ts for ts in range(nrts + 1) if self.due(ts)]
        self.pid = None

    def due(self, ts):
        if not self.enabled or ts < self.start_ts or ts > self.stop_ts:
            return False
        return ts % self.interval == 0 or (self.fd_interval != 0 and ts % self.fd_interval == 0)

    def _terms(self):
        op = self.op
        s = self.spec
        self.comp, self.pos, self.sign, self.div = [], [], [], None
        if self.ptype == 0:
            start, _ = op.snap_to_mesh(s["start"], False)
            stop, _ = op.snap_to_mesh(s["stop"], False)
            self.start, self.stop = start, stop
            if sum(1 for n in range(3) if start[n] != stop[n]) != 1:
                # not a line: the probe stays on and records zeros
                # (CalcVoltageIntegral, FDTD/engine_interface_fdtd.cpp:208-212)
                return
            for n in range(3):
                if start[n] < stop[n]:
                    p = list(start)
                    while p[n] < stop[n]:
                        self._add(n, p, 1)
                        p[n] += 1
                elif start[n] > stop[n]:
                    p = list(stop)
                    while p[n] < start[n]:
                        self._add(n, p, -1)
                        p[n] += 1
            return
        # ProcessCurrent snaps with SnapMethod 1 (processcurrent.cpp:28), field probes with 0
        dim, st, sp, sin, ein = op.snap_box(s["start"], s["stop"], self.dual, method=1 if self.ptype == 1 else 0)
        if dim < 0:
            self.enabled = False
            return
        self.start, self.stop, self.sin, self.ein = st, sp, sin, ein
        if self.ptype == 1:
            self._current_terms(s, dim)
            return
        self.div = []
        for n in range(3):
            self.comp.append(n)
            self.pos.append(list(st))
            self.sign.append(1)
            self.div.append(op.edge_length(n, st[n], self.ptype == 3))

    def _add(self, n, p, sg):
        self.comp.append(n)
        self.pos.append(list(p))
        self.sign.append(sg)

    def _current_terms(self, s, dim):
        """ProcessCurrent::DefineStartStopCoord + CalcIntegral term order."""
        op = self.op
        st, sp = self.start, self.stop
        nd = int(s.get("norm_dir", -1))
        ds, de = s["start"], s["stop"]
        if nd < 0 or nd > 2:
            if dim != 2:
                self.enabled = False
                return
            for n in range(3):
                if sp[n] == st[n]:
                    nd = n
        else:
            for n in range(3):
                if n == nd:
                    continue
                if ds[n] == de[n]:
                    if op.disc_line(n, st[n], True) > ds[n] and st[n] > 0:
                        st[n] -= 1
                    if op.disc_line(n, st[n], True) < ds[n] and sp[n] < op.N[n] - 1:
                        sp[n] += 1
        if st[nd] != sp[nd]:
            self.enabled = False
            return
        si, ei = self.sin, self.ein
        seq = []

        def seg(n_comp, fixed, axis, lo, hi, sg):
            for i in range(lo + 1, hi + 1):
                p = list(fixed)
                p[axis] = i
                seq.append((n_comp, p, sg))
        if nd == 0:
            if ei[0] and si[2]:
                seg(1, [sp[0], 0, st[2]], 1, st[1], sp[1], 1)
            if ei[0] and ei[1]:
                seg(2, [sp[0], sp[1], 0], 2, st[2], sp[2], 1)
            if si[0] and ei[2]:
                seg(1, [st[0], 0, sp[2]], 1, st[1], sp[1], -1)
            if si[0] and si[1]:
                seg(2, [st[0], st[1], 0], 2, st[2], sp[2], -1)
        elif nd == 1:
            if si[0] and si[1]:
                seg(2, [st[0], st[1], 0], 2, st[2], sp[2], 1)
            if ei[1] and ei[2]:
                seg(0, [0, sp[1], sp[2]], 0, st[0], sp[0], 1)
            if ei[0] and ei[1]:
                seg(2, [sp[0], sp[1], 0], 2, st[2], sp[2], -1)
            if si[1] and si[2]:
                seg(0, [0, st[1], st[2]], 0, st[0], sp[0], -1)
        else:
            if si[1] and si[2]:
                seg(0, [0, st[1], st[2]], 0, st[0], sp[0], 1)
            if ei[0] and si[2]:
                seg(1, [sp[0], 0, st[2]], 1, st[1], sp[1], 1)
            if ei[1] and ei[2]:
                seg(0, [0, sp[1], sp[2]], 0, st[0], sp[0], -1)
            if si[0] and ei[2]:
                seg(1, [st[0], 0, sp[2]], 1, st[1], sp[1], -1)
        for n, p, sg in seq:
            self._add(n, p, sg)
        self.norm_dir = nd

    # ---- probe files ----------------------------------------------------------------
    def _coords_header(self):
        op = self.op
        gd = float(op.scene.get("delta_unit", 1.0))
        out = []
        for label, p in (("start", self.start), ("stop", self.stop)):
            xyz = ",".join(_g(op.disc_line(n, p[n]) * gd) for n in range(3))
            idx = ",".join(str(int(v)) for v in p)
            out.append(f"% {label}-coordinates: ({xyz}) m -> [{idx}]\n")
        return out

    def write_files(self, td, fd, directory, version="openems_amd", when=None):
        """Write the probe's time-domain file <name> and, with frequencies, its
        frequency-domain file <name>_FD in the reference's text layout:
        header of ProcessIntegral::InitProcess (Common/processintegral.cpp:41-72),
        rows of ProcessIntegral::Process (:142-151; time and values at 12
        significant digits) and ProcessIntegral::Dump_FD_Data (:92-132; 6
        digits).  A disabled probe writes nothing.  Returns the paths written."""
        import os
        import time as _time
        if not self.enabled:
            return []
        stamp = _time.ctime(when)
        names = INTEGRAL_NAMES[self.ptype]
        proc = PROCESSING_NAMES[self.ptype]
        path = os.path.join(directory, self.name)
        lines = [f"% time-domain {proc} by openEMS {version} @{stamp}\n"] + self._coords_header()
        lines.append("% t/s" + "".join("\t" + n for n in names) + "\n")
        for row in np.asarray(td):
            lines.append(_g(float(row[0]), 12) + "".join("\t" + _g(float(v), 12) for v in row[1:]) + "\n")
        with open(path, "w") as f:
            f.writelines(lines)
        paths = [path]
        if self.freqs:
            fd = np.asarray(fd)
            lines = [f"% frequency-domain {proc} by openEMS {version} @{stamp}\n"] + self._coords_header()
            lines.append("% f/Hz" + "".join("\t" + n + "\t" for n in names) + "\n")
            lines.append("%" + "\treal\timag" * len(names) + "\n")
            for j, f0 in enumerate(self.freqs):
                lines.append(_g(f0) + "".join("\t" + _g(fd[i, j].real) + "\t" + _g(fd[i, j].imag)
                                              for i in range(len(names))) + "\n")
            with open(path + "_FD", "w") as f:
                f.writelines(lines)
            paths.append(path + "_FD")
        return paths

    # ---- engine registration and results -----------------------------------------
    def register(self, eng):
        if not self.enabled or not self.comp:
            return
        pos = np.array(self.pos, np.uint32).T.copy()     # [3][nelem]
        self.pid = eng.add_probe(self.ptype, self.comp, pos, self.sign, self.div, self.interval,
                                 self.fd_interval, self.start_ts, self.stop_ts, len(self.steps))

    def collect(self, eng):
        """Returns (td, fd): td rows (t, values*weight...), fd [nint][nfreq] complex."""
        nint = 3 if self.ptype in (2, 3) else 1
        if self.pid is None:
            ts = np.array(self.steps, np.uint32)
            vals = np.zeros((len(ts), 3))
        else:
            ts, vals = eng.probe_results(self.pid)
        return self._series(ts, vals, nint)

    def _series(self, ts, vals, nint):
        """TD rows and FD sums of sampled integrals (ProcessIntegral::Process,
        Common/processintegral.cpp:134-177)."""
        dT = self.op.dT
        td = []
        fd = np.zeros((nint, len(self.freqs)), np.complex128)
        for k in range(len(ts)):
            t_int = int(ts[k])
            t = (t_int + (0.5 if self.dual else 0.0)) * dT   # GetTime(dualTime)
            v = [float(vals[k, i]) for i in range(nint)]
            if t_int % self.interval == 0:
                td.append([t] + [x * self.weight for x in v])
            if self.fd_interval and t_int % self.fd_interval == 0:
                for j, f in enumerate(self.freqs):
                    th = ((-2.0 * math.pi) * f) * t
                    e = cmath.exp(complex(-0.0, th))
                    for i, x in enumerate(v):
                        a = x * self.weight
                        re = ((a * e.real) * 2.0) * dT * float(self.fd_interval)
                        im = ((a * e.imag) * 2.0) * dT * float(self.fd_interval)
                        fd[i, j] += complex(re, im)
        return np.array(td) if td else np.zeros((0, 1 + nint)), fd


def evaluate_function(text, names, var):
    """An expression of the function-parser syntax at every point of var
    [..., nvar], through the operator library's parser."""
    import ctypes as C
    from ._native import check_op, lib, ptr
    v = np.ascontiguousarray(var, np.float64)
    shape = v.shape[:-1]
    flat = v.reshape(-1, v.shape[-1])
    out = np.zeros(flat.shape[0])
    check_op(lib().emsop_eval_function(text.encode(), names.encode(), flat.shape[1], ptr(flat, C.c_double),
                                       flat.shape[0], ptr(out, C.c_double)))
    return out.reshape(shape)


def make_probe(op, spec, nrts, oversampling=4):
    """The processing object of a probe spec (openEMS::SetupProcessing, openems.cpp:520-575)."""
    if int(spec["type"]) not in PROBE_NAMES:
        raise ValueError(f"probe type {spec['type']} is not supported (mode matching, types 10 and 11, "
                         "is outside the engine's scope)")
    return Probe(op, spec, nrts, oversampling)


class FieldDumpFD:
    """ProcessFieldsFD (Common/processfields_fd.cpp:72-107) of the E field
    (dump type 10, spec "field": "E") or the H field (type 11, "H"), with no
    (spec "interp": 0), node (1, the default) or cell (2) interpolation.  As
    set up by openEMS::SetupProcessing (openems.cpp:612-681): an H dump samples
    at the dual time and, without interpolation, on the dual mesh; cell
    interpolation always uses the dual mesh; the box snaps with SnapMethod 1
    (processfields.cpp:40, operator.cpp:323-337)."""

    def __init__(self, op, spec, nrts):
        self.op = op
        self.spec = spec
        self.name = spec.get("name", "")
        fld = spec.get("field", "E")
        self.field = {"E": 0, "H": 1, 0: 0, 1: 1, 10: 0, 11: 1}[fld]
        self.interp = int(spec.get("interp", 1))
        if self.interp not in (0, 1, 2):
            raise ValueError("dump interpolation must be 0 (none), 1 (node) or 2 (cell)")
        self.dual_time = self.field == 1
        dual = self.interp == 2 or (self.field == 1 and self.interp == 0)
        self.dual_mesh = dual
        dim, st, sp, _, _ = op.snap_box(spec["start"], spec["stop"], dual, method=1)
        # a box outside the mesh disables the dump (Processing::DefineStartStopCoord, processing.cpp:180-188)
        self.enabled = dim >= 0
        self.start, self.stop = st, sp
        nyq = op.nyquist
        self.fd_interval = 0
        self.freqs = []
        for f in spec.get("freqs", []) if self.enabled else []:
            nts = nyquist_num(f, op.dT)
            if nts == 0:
                continue
            if self.fd_interval == 0:
                self.fd_interval = nyq
            if self.fd_interval > nts:
                self.fd_interval = nts
            self.freqs.append(float(f))
        # samples at the multiples of fd_interval inside the window (Processing::CheckTimestep)
        self.start_ts, self.stop_ts = process_window(spec, op.dT)
        self.first, self.nsamples = 0, 0
        if self.fd_interval:
            iv = self.fd_interval
            self.first = (self.start_ts + iv - 1) // iv * iv
            last = min(nrts, self.stop_ts)
            self.nsamples = (last - self.first) // iv + 1 if last >= self.first else 0
        self.did = None

    def factors(self):
        """exp(-2 pi i f T) * 2 * dT * FD_Interval in complex<float>, per sample and frequency."""
        dT = self.op.dT
        out = np.zeros((self.nsamples, len(self.freqs), 2), np.float32)
        scale = np.float32(dT * self.fd_interval)
        for k in range(self.nsamples):
            # GetTime(dualTime) = (numTS + 0.5 dualTime) dT (engine_interface_fdtd.h:52)
            T = (self.first + k * self.fd_interval + (0.5 if self.dual_time else 0.0)) * dT
            for j, f in enumerate(self.freqs):
                th = np.float32(((-2.0 * math.pi) * f) * T)
                c = np.float32(_libm.cosf(th))
                s = np.float32(_libm.sinf(th))
                out[k, j, 0] = np.float32(c * np.float32(2)) * scale
                out[k, j, 1] = np.float32(s * np.float32(2)) * scale
        return out

    def register(self, eng):
        if not self.fd_interval or not self.nsamples:
            return
        self.did, self.shape = eng.add_fd_dump(self.start, self.stop, self.fd_interval, self.factors(),
                                               field=self.field, interp=self.interp, first_ts=self.first)

    def collect(self, eng):
        if self.did is None:
            return None
        raw = eng.fd_dump_results(self.did, self.shape)
        return raw[..., 0].astype(np.float32) + 1j * raw[..., 1].astype(np.float32)


class FieldDumpTD:
    """ProcessFieldsTD (Common/processfields_td.cpp:50-91) of the E field
    (dump type 0, spec "field": "E") or the H field (type 1, "H") with no,
    node or cell interpolation (spec "interp", default 1), set up as
    FieldDumpFD: the field of the box at every step that
    Processing::CheckTimestep accepts (processing.cpp:82-106) -- inside the
    window spec "start_time" .. "stop_time" (seconds, SetProcessStartStopTime,
    processing.cpp:211-226) and a multiple of ProcessInterval =
    Nyquist / OverSampling (openems.cpp:631; 0 takes no samples).  The samples
    stay on the device until collected."""

    def __init__(self, op, spec, nrts, oversampling=4):
        fd = FieldDumpFD(op, dict(spec, freqs=[]), nrts)
        self.op = op
        self.spec = spec
        self.name = spec.get("name", "")
        self.field, self.interp = fd.field, fd.interp
        self.dual_time, self.dual_mesh = fd.dual_time, fd.dual_mesh
        self.enabled = fd.enabled
        self.start, self.stop = fd.start, fd.stop
        self.interval = op.nyquist // int(oversampling)
        self.start_ts, self.stop_ts = process_window(spec, op.dT)
        self.steps = [ts for ts in range(nrts + 1) if self.due(ts)]
        self.did = None

    def due(self, ts):
        return (self.enabled and self.interval > 0 and self.start_ts <= ts <= self.stop_ts
                and ts % self.interval == 0)

    def register(self, eng):
        if not self.steps:
            return
        self.did, self.box = eng.add_td_dump(self.start, self.stop, self.interval, field=self.field,
                                             interp=self.interp, start_ts=self.start_ts, stop_ts=self.stop_ts,
                                             capacity=len(self.steps))

    def collect(self, eng):
        """(steps, times, fields [count][3][nx][ny][nz] float32); time at the
        dual time for H, GetTime(dualTime) (engine_interface_fdtd.h:52)."""
        if self.did is None:
            return np.zeros(0, np.uint32), np.zeros(0), None
        ts, out = eng.td_dump_results(self.did, self.box)
        t = (ts.astype(np.float64) + (0.5 if self.dual_time else 0.0)) * self.op.dT
        return ts, t, out
