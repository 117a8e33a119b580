"""Engine_HIP: the MI355X engine behind include/ems_hip.h.

Mirrors the reference's Engine API (FDTD/engine.h:41-101): ``IterateTS``,
``GetNumberOfTimesteps`` and bulk field access, with the engine extensions
(excitation, UPML, Mur) created from an :class:`~openems_amd.operator.Operator`
in the reference's creation order.  There is no CPU path: without a GPU or
without the native library construction raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import ProbeDesc, check, dp, fp, lib, ptr

RLC_KEYS = ("i2v", "ilv", "vvd", "vv2", "vj1", "vj2", "ib0", "b1", "b2")

# device coefficient layouts (include/ems_hip.h)
LAYOUT_AUTO, LAYOUT_PLAIN, LAYOUT_COMPRESSED = 0, 1, 2
LAYOUTS = {"auto": LAYOUT_AUTO, "plain": LAYOUT_PLAIN, "compressed": LAYOUT_COMPRESSED}


def compress_operator(N, vv, vi, ii, iv, max_unique=65536):
    """Per-cell compression of an N-I-J-K operator as the engine applies it
    (host only): returns (idx [Nz][Ny][Nx] uint16, tabV [U][8], tabI [U][8])."""
    arrs = [np.ascontiguousarray(a, np.float32) for a in (vv, vi, ii, iv)]
    tot = int(np.prod(N))
    idx = np.empty(tot, np.uint16)
    tv = np.empty((max_unique, 8), np.float32)
    ti = np.empty((max_unique, 8), np.float32)
    U = check(lib().ems_compress_operator((C.c_uint * 3)(*N), *[ptr(a, C.c_float) for a in arrs],
                                          ptr(idx, C.c_ushort), ptr(tv, C.c_float), ptr(ti, C.c_float),
                                          int(max_unique)))
    return idx.reshape(N[2], N[1], N[0]), tv[:U].copy(), ti[:U].copy()


def _table_ptrs(tables):
    """[3] arrays of n_unique x 4 floats -> (float*)[3]"""
    arrs = [np.ascontiguousarray(t, np.float32) for t in tables]
    return arrs, (fp * 3)(*[ptr(a, C.c_float) for a in arrs])


def expand_sse_compressed(N, op_index, vv4, vi4, ii4, iv4):
    """N-I-J-K arrays from the reference's SSE-compressed operator (host only)."""
    opi = np.ascontiguousarray(op_index, np.uint32)
    U = len(vv4[0])
    keep, out = [], []
    ptrs = []
    for t in (vv4, vi4, ii4, iv4):
        a, p = _table_ptrs(t)
        keep.append(a)
        ptrs.append(p)
    for _ in range(4):
        out.append(np.empty((3,) + tuple(N), np.float32))
    check(lib().ems_expand_sse_compressed((C.c_uint * 3)(*N), ptr(opi, C.c_uint), U, *ptrs,
                                          *[ptr(o, C.c_float) for o in out]))
    return out


STEP_FORMS = {"halves": 0, "lds": 1, "columns": 2, "quads": 3}


class EngineHIP:
    def __init__(self, op, device=0, layout="auto", sse_compressed=None, fused=None, steady_state=None,
                 lorentz=None, rlc=None, quad=None):
        """op: Operator; layout: "auto" | "plain" | "compressed";
        sse_compressed: optional (op_index, vv4, vi4, ii4, iv4), the reference's
        SSE-compressed form of the same operator, used instead of op's arrays;
        steady_state: optional (period_steps, dirs, positions) of a
        steady-state detection (Engine_Ext_SteadyState);
        lorentz: dispersion orders (Engine_Ext_LorentzMaterial tables), each a
        dict pos [3][K], v_int, v_ext, v_lor (or None), i_int, i_ext, i_lor
        [3][K], volt, curr (bool); None takes the operator's own
        (Operator.lorentz_tables);
        rlc: lumped RLC table (Engine_Ext_LumpedRLC), a dict pos [3][K], dir
        [K] and the nine coefficient arrays i2v, ilv, vvd, vv2, vj1, vj2, ib0,
        b1, b2 [K]; None takes the operator's own (Operator.rlc_table);
        fused: step form, None (automatic), False or "halves" (two half
        steps), True or "lds" (single sweep staging planes in LDS, k_step) or
        "columns" (single sweep of independent columns, k_step2); a required
        single-sweep form fails when an active extension has none;
        quad: half steps with four x cells per lane (k_volt4 / k_curr4), None
        (automatic: used when the x lines are a multiple of 4, the step runs
        two half steps and the mesh has no UPML), False (one cell per lane)
        or True (required)."""
        L = lib()
        if L.ems_device_count() <= 0:
            raise RuntimeError("EngineHIP needs a HIP device; none is visible")
        self.op = op
        self.N = op.N
        h = C.c_void_p()
        n = (C.c_uint * 3)(*op.N)
        lay = LAYOUTS[layout] if isinstance(layout, str) else int(layout)
        if sse_compressed is not None:
            opi, *tabs = sse_compressed
            opi = np.ascontiguousarray(opi, np.uint32)
            keep, ptrs = [opi], []
            for t in tabs:
                a, p = _table_ptrs(t)
                keep.append(a)
                ptrs.append(p)
            check(L.ems_create_sse_compressed(C.byref(h), n, ptr(opi, C.c_uint), len(tabs[0][0]), *ptrs,
                                              int(device), lay))
        else:
            check(L.ems_create_layout(C.byref(h), n, op.coeff_ptr("vv"), op.coeff_ptr("vi"), op.coeff_ptr("ii"),
                                      op.coeff_ptr("iv"), int(device), lay))
        self.h = h
        self._keep = []
        if fused is not None:
            check(L.ems_set_fused(h, STEP_FORMS[fused] if isinstance(fused, str) else (1 if fused else 0)))
        if quad is not None:
            check(L.ems_set_quad(h, 1 if quad else 0))
        if steady_state is not None:
            self.set_steady_state(*steady_state)
        # the operator's own Lorentz / lumped RLC tables unless tables are given
        if lorentz is None and hasattr(op, "lorentz_tables"):
            lorentz = [o for o in op.lorentz_tables() if o["pos"].shape[1] > 0 and (o["volt"] or o["curr"])]
        if rlc is None and hasattr(op, "rlc_table"):
            rlc = op.rlc_table()
        self._lorentz = lorentz or []
        self._rlc = rlc if rlc else None
        try:
            self._extensions()
        except Exception:
            self.close()
            raise

    def _extensions(self):
        L = lib()
        op = self.op
        sv, sc = op.signal()
        ve, ce = op.excitation(False), op.excitation(True)
        arrs = [sv, sc, ve["index"], ve["dir"], ve["amp"], ve["delay"], ce["index"], ce["dir"], ce["amp"], ce["delay"]]
        arrs = [np.ascontiguousarray(a) for a in arrs]
        self._keep += arrs
        sv, sc, vi, vd, va, vdl, ci, cd, ca, cdl = arrs
        check(L.ems_set_excitation(self.h, ptr(sv, C.c_float), ptr(sc, C.c_float), len(sv), op.period_steps,
                                   len(vd), ptr(vi, C.c_uint), ptr(vd, C.c_ushort), ptr(va, C.c_float),
                                   ptr(vdl, C.c_uint), len(cd), ptr(ci, C.c_uint), ptr(cd, C.c_ushort),
                                   ptr(ca, C.c_float), ptr(cdl, C.c_uint)))
        for f in op.mur_faces():
            cP = np.ascontiguousarray(f["coeff_nyP"], np.float32)
            cPP = np.ascontiguousarray(f["coeff_nyPP"], np.float32)
            check(L.ems_add_mur(self.h, f["ny"], f["line"], f["line_shift"], ptr(cP, C.c_float), ptr(cPP, C.c_float),
                                f["start_ts"]))
        for s in op.upml_slabs():
            st = (C.c_uint * 3)(*s["start"])
            nm = (C.c_uint * 3)(*s["num"])
            check(L.ems_add_upml(self.h, st, nm, *s["_ptrs"]))
        for o in self._lorentz:
            pos = np.ascontiguousarray(o["pos"], np.uint32)
            arrs = {k: (None if o.get(k) is None else np.ascontiguousarray(o[k], np.float32))
                    for k in ("v_int", "v_ext", "v_lor", "i_int", "i_ext", "i_lor")}
            self._keep += [pos] + [a for a in arrs.values() if a is not None]
            fpn = lambda a: None if a is None else ptr(a, C.c_float)
            check(L.ems_add_lorentz(self.h, pos.shape[1], ptr(pos, C.c_uint), fpn(arrs["v_int"]), fpn(arrs["v_ext"]),
                                    fpn(arrs["v_lor"]), fpn(arrs["i_int"]), fpn(arrs["i_ext"]), fpn(arrs["i_lor"]),
                                    int(bool(o.get("volt", True))), int(bool(o.get("curr", True)))))
        if self._rlc is not None:
            r = self._rlc
            pos = np.ascontiguousarray(r["pos"], np.uint32)
            d = np.ascontiguousarray(r["dir"], np.int32)
            cs = [np.ascontiguousarray(r[k], np.float32) for k in RLC_KEYS]
            self._keep += [pos, d] + cs
            cp = (fp * 9)(*[ptr(a, C.c_float) for a in cs])
            check(L.ems_add_lumped_rlc(self.h, pos.shape[1], ptr(pos, C.c_uint), ptr(d, C.c_int), cp))
        check(L.ems_init(self.h))

    # -- Engine API ------------------------------------------------------------
    def IterateTS(self, n):
        check(lib().ems_iterate(self.h, int(n)))
        return True

    iterate = IterateTS

    def GetNumberOfTimesteps(self):
        return lib().ems_get_numts(self.h)

    @property
    def numTS(self):
        return self.GetNumberOfTimesteps()

    def get_volt(self):
        out = np.empty((3,) + tuple(self.N), np.float32)
        check(lib().ems_get_volt(self.h, ptr(out, C.c_float)))
        return out

    def get_curr(self):
        out = np.empty((3,) + tuple(self.N), np.float32)
        check(lib().ems_get_curr(self.h, ptr(out, C.c_float)))
        return out

    def set_volt(self, a):
        a = np.ascontiguousarray(a, np.float32)
        check(lib().ems_set_volt(self.h, ptr(a, C.c_float)))

    def set_curr(self, a):
        a = np.ascontiguousarray(a, np.float32)
        check(lib().ems_set_curr(self.h, ptr(a, C.c_float)))

    def read_field(self, curr, start, stop, comp=-1):
        """Voltages (curr False) or currents of the box start..stop (inclusive
        mesh indices): [3][nx][ny][nz] for comp -1, else [nx][ny][nz]."""
        st = (C.c_uint * 3)(*[int(v) for v in start])
        sp = (C.c_uint * 3)(*[int(v) for v in stop])
        n = tuple(int(b) - int(a) + 1 for a, b in zip(start, stop))
        out = np.empty(((3,) if comp < 0 else ()) + n, np.float32)
        check(lib().ems_read_field(self.h, int(bool(curr)), int(comp), st, sp, ptr(out, C.c_float)))
        return out

    def write_field(self, curr, start, stop, values, comp=-1):
        st = (C.c_uint * 3)(*[int(v) for v in start])
        sp = (C.c_uint * 3)(*[int(v) for v in stop])
        a = np.ascontiguousarray(values, np.float32)
        check(lib().ems_write_field(self.h, int(bool(curr)), int(comp), st, sp, ptr(a, C.c_float)))

    def field_at(self, field, pos, interp=0):
        """E (field 0) or H (field 1) at mesh node pos with interpolation 0
        (none), 1 (node) or 2 (cell): Engine_Interface_FDTD::GetEField / GetHField."""
        el = [np.ascontiguousarray(self.op.edge_lengths(n, dual=bool(field)), np.float64) for n in range(3)]
        arr = (dp * 3)(*[ptr(e, C.c_double) for e in el])
        p = (C.c_uint * 3)(*[int(v) for v in pos])
        out = np.zeros(3, np.float64)
        check(lib().ems_get_field_at(self.h, int(field), int(interp), p, arr, ptr(out, C.c_double)))
        return out

    def energy(self):
        e = C.c_double()
        check(lib().ems_energy(self.h, C.byref(e)))
        return e.value

    def sync(self):
        check(lib().ems_sync(self.h))

    def set_async(self, flag):
        check(lib().ems_set_async(self.h, int(flag)))

    def set_profiling(self, flag):
        check(lib().ems_set_profiling(self.h, int(flag)))

    def set_graph(self, chunk_steps):
        """Replay chunks of chunk_steps timesteps as captured HIP graphs (0: direct launches)."""
        check(lib().ems_set_graph(self.h, int(chunk_steps)))

    def graph(self):
        """(chunk steps, graphs captured, chunks replayed)"""
        a, b, c = C.c_uint(), C.c_uint(), C.c_uint()
        check(lib().ems_get_graph(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def layout(self):
        """(layout, n_unique, pml_layout, n_unique_pml) as names and table sizes"""
        a, b, c, d = C.c_int(), C.c_uint(), C.c_int(), C.c_uint()
        check(lib().ems_get_layout(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        names = {0: None, 1: "plain", 2: "compressed"}
        return names[a.value], b.value, names[c.value], d.value

    def set_steady_state(self, period_steps, dirs, positions, history=4096):
        """Engine_Ext_SteadyState: period in timesteps, E probes (dir, (x, y, z))."""
        d = np.ascontiguousarray(dirs, np.int32)
        pos = np.ascontiguousarray(np.asarray(positions, np.uint32).T)
        self._keep += [d, pos]
        check(lib().ems_set_steady_state(self.h, int(period_steps), len(d), ptr(d, C.c_int), ptr(pos, C.c_uint),
                                         int(history)))

    def steady_state(self, cap=4096):
        """(GetLastDiff, [(timestep, diff) of every evaluation])"""
        last, n = C.c_double(), C.c_uint()
        ts = np.zeros(cap, np.uint32)
        df = np.zeros(cap)
        check(lib().ems_steady_state(self.h, C.byref(last), C.byref(n), ptr(ts, C.c_uint), ptr(df, C.c_double), cap))
        m = min(n.value, cap)
        return last.value, list(zip(ts[:m].tolist(), df[:m].tolist()))

    def fused(self):
        """(active step form: 0 two half steps, 1 LDS sweep, 2 column sweep;
        z planes per workgroup)"""
        f, d = C.c_int(), C.c_uint()
        check(lib().ems_get_fused(self.h, C.byref(f), C.byref(d)))
        return f.value, d.value

    def quad(self):
        """(1 when the half steps run four x cells per lane, else 0; z planes per workgroup)"""
        q, d = C.c_int(), C.c_uint()
        check(lib().ems_get_quad(self.h, C.byref(q), C.byref(d)))
        return q.value, d.value

    def last_timing(self):
        t, v, c = C.c_double(), C.c_double(), C.c_double()
        check(lib().ems_last_timing(self.h, C.byref(t), C.byref(v), C.byref(c)))
        return t.value, v.value, c.value

    # -- probes and dumps ---------------------------------------------------------
    def add_probe(self, ptype, comp, pos, sign, divisor, interval, fd_interval, start_ts, stop_ts, capacity):
        comp = np.ascontiguousarray(comp, np.uint32)
        pos = np.ascontiguousarray(pos, np.uint32)
        sign = np.ascontiguousarray(sign if sign is not None else np.ones(len(comp)), np.int32)
        div = np.ascontiguousarray(divisor if divisor is not None else np.zeros(len(comp)), np.float64)
        self._keep += [comp, pos, sign, div]
        d = ProbeDesc(int(ptype), len(comp), ptr(comp, C.c_uint), ptr(pos, C.c_uint), ptr(sign, C.c_int),
                      ptr(div, C.c_double), int(interval), int(fd_interval), int(start_ts), int(stop_ts),
                      int(capacity))
        pid = C.c_int()
        check(lib().ems_add_probe(self.h, C.byref(d), C.byref(pid)))
        return pid.value

    def probe_results(self, pid):
        L = lib()
        cnt = C.c_uint()
        check(L.ems_probe_results(self.h, pid, C.byref(cnt), None, None))
        n = cnt.value
        ts = np.zeros(n, np.uint32)
        vals = np.zeros((n, 3), np.float64)
        if n:
            check(L.ems_probe_results(self.h, pid, C.byref(cnt), ptr(ts, C.c_uint), ptr(vals, C.c_double)))
        return ts, vals

    def add_fd_dump(self, start, stop, fd_interval, factors, field=0, interp=1, first_ts=0):
        """Frequency-domain dump of the E (field 0) or H (field 1) field over
        the mesh-index box start..stop, interpolation 0 none, 1 node, 2 cell."""
        L = lib()
        factors = np.ascontiguousarray(factors, np.float32)   # [nsamples][nfreq][2]
        ns, nf = factors.shape[0], factors.shape[1]
        els = [np.ascontiguousarray(self.op.edge_lengths(n, field == 1), np.float64) for n in range(3)]
        self._keep += [factors] + els
        elp = (dp * 3)(*[ptr(e, C.c_double) for e in els])
        st = (C.c_uint * 3)(*start)
        sp = (C.c_uint * 3)(*stop)
        did = C.c_int()
        check(L.ems_add_fd_dump(self.h, int(field), int(interp), st, sp, int(fd_interval), int(first_ts), nf, ns,
                                ptr(factors, C.c_float), elp, C.byref(did)))
        shape = (nf, 3) + tuple(int(stop[n]) - int(start[n]) + 1 for n in range(3))
        return did.value, shape

    def add_td_dump(self, start, stop, interval, field=0, interp=1, start_ts=0, stop_ts=0xFFFFFFFF, capacity=1):
        """Time-domain dump of the E (field 0) or H (field 1) field over the
        mesh-index box start..stop after every step that is a multiple of
        interval inside [start_ts, stop_ts]; keeps up to capacity samples."""
        L = lib()
        els = [np.ascontiguousarray(self.op.edge_lengths(n, field == 1), np.float64) for n in range(3)]
        self._keep += els
        elp = (dp * 3)(*[ptr(e, C.c_double) for e in els])
        st = (C.c_uint * 3)(*start)
        sp = (C.c_uint * 3)(*stop)
        did = C.c_int()
        check(L.ems_add_td_dump(self.h, int(field), int(interp), st, sp, int(interval), int(start_ts), int(stop_ts),
                                int(capacity), elp, C.byref(did)))
        return did.value, tuple(int(stop[n]) - int(start[n]) + 1 for n in range(3))

    def td_dump_results(self, did, box):
        """(steps [count], fields [count][3][nx][ny][nz] float32) sampled so far."""
        L = lib()
        cnt = C.c_uint()
        check(L.ems_td_dump_results(self.h, did, C.byref(cnt), None, None))
        n = cnt.value
        ts = np.zeros(n, np.uint32)
        out = np.zeros((n, 3) + tuple(box), np.float32)
        if n:
            check(L.ems_td_dump_results(self.h, did, C.byref(cnt), ptr(ts, C.c_uint), ptr(out, C.c_float)))
        return ts, out

    def fd_dump_results(self, did, shape):
        out = np.zeros(shape + (2,), np.float32)
        check(lib().ems_fd_dump_results(self.h, did, ptr(out, C.c_float)))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().ems_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
