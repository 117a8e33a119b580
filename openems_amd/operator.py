"""Operator: host build of the EC-FDTD operator (include/ems_operator.h).

Python face of the native operator builder, mirroring the reference's
``Operator`` (FDTD/operator.h): timestep, update coefficients vv/vi/ii/iv,
excitation signal and cell lists, UPML slabs and Mur faces, mesh queries.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import check_op, dp, fp, ip, lib, ptr, up, usp

BC_CODES = {"PEC": 0, "PMC": 1, "MUR": 2}
SIGNALS = {"gauss": 0, "sinus": 1, "dirac": 2, "step": 3}


def _v3(v):
    if np.ndim(v) == 0:
        return np.full(3, float(v))
    return np.asarray(v, dtype=np.float64).reshape(3)


def parse_bc(bc):
    """openEMS.SetBoundaryCond strings/ints -> (types, pml sizes) (python/openEMS/openEMS.pyx:282-309)."""
    types, sizes = [], []
    for b in bc:
        if isinstance(b, str):
            if b.startswith("PML_"):
                types.append(3)
                sizes.append(int(b[4:]))
                continue
            types.append(BC_CODES[b])
        else:
            types.append(int(b))
        sizes.append(8)
    return types, sizes


RLC_COEFS = ("i2v", "ilv", "vvd", "vv2", "vj1", "vj2", "ib0", "b1", "b2")
DISPERSION_KEYS = ("eps_plasma", "eps_relax", "eps_pole", "mue_plasma", "mue_relax", "mue_pole")


def dispersion_params(orders):
    """[norder][6][3] float64 of a material's "lorentz" entry: a list with one
    dict per dispersion order holding eps_plasma, eps_relax, eps_pole,
    mue_plasma, mue_relax, mue_pole (plasma and pole frequencies in Hz,
    relaxation times in s; scalars or per-direction triples, 0 = unset)."""
    out = np.zeros((len(orders), 6, 3), np.float64)
    for o, d in enumerate(orders):
        for k, key in enumerate(DISPERSION_KEYS):
            out[o, k] = _v3(d.get(key, 0.0))
    return np.ascontiguousarray(out)


def debye_params(orders):
    """[norder][2][3] float64 of a material's "debye" entry: one dict per
    order with eps_delta and eps_relax (s), scalars or per-direction triples."""
    out = np.zeros((len(orders), 2, 3), np.float64)
    for o, d in enumerate(orders):
        out[o, 0] = _v3(d.get("eps_delta", 0.0))
        out[o, 1] = _v3(d.get("eps_relax", 0.0))
    return np.ascontiguousarray(out)


class Operator:
    """Builds the operator of a scene dict (see openems_amd.scenes)."""

    def __init__(self, scene, nthreads=0):
        L = lib()
        self.scene = scene
        h = C.c_void_p()
        check_op(L.emsop_create(C.byref(h)))
        self.h = h
        try:
            self._setup(scene, nthreads)
            check_op(L.emsop_build(self.h, int(scene.get("nrts", 1000))))
        except Exception:
            self.close()
            raise
        n = (C.c_uint * 3)()
        L.emsop_num_lines(self.h, n)
        self.N = (n[0], n[1], n[2])

    def _setup(self, s, nthreads):
        L = lib()
        lines = [np.ascontiguousarray(l, np.float64) for l in s["lines"]]
        self._lines = lines
        check_op(L.emsop_set_grid(self.h, ptr(lines[0], C.c_double), len(lines[0]), ptr(lines[1], C.c_double),
                                  len(lines[1]), ptr(lines[2], C.c_double), len(lines[2]),
                                  float(s.get("delta_unit", 1.0))))
        bg = s.get("background", {})
        check_op(L.emsop_set_background(self.h, float(bg.get("eps", 1.0)), float(bg.get("mue", 1.0)),
                                        float(bg.get("kappa", 0.0)), float(bg.get("sigma", 0.0))))
        for m in s.get("materials", []):
            if "grid" in m:
                # material grid: per-cell eps_r / kappa over whole mesh cells
                g = m["grid"]
                c0 = np.asarray(g["cell0"], np.uint32)
                nc = np.asarray(g["ncells"], np.uint32)
                eps = np.ascontiguousarray(np.asarray(g["eps"], np.float64).reshape(-1))
                kap = g.get("kappa")
                kap = None if kap is None else np.ascontiguousarray(np.asarray(kap, np.float64).reshape(-1))
                self._keep_grid = getattr(self, "_keep_grid", []) + [c0, nc, eps, kap]
                check_op(L.emsop_add_material_grid(self.h, ptr(c0, C.c_uint), ptr(nc, C.c_uint), ptr(eps, C.c_double),
                                                   None if kap is None else ptr(kap, C.c_double),
                                                   float(m.get("mue", 1.0)), float(m.get("sigma", 0.0)),
                                                   int(m.get("priority", 0))))
                continue
            arrs = [_v3(m.get(k, d)) for k, d in (("eps", 1.0), ("mue", 1.0), ("kappa", 0.0), ("sigma", 0.0))]
            st, sp = _v3(m["start"]), _v3(m["stop"])
            check_op(L.emsop_add_material(self.h, *[ptr(a, C.c_double) for a in arrs], int(m.get("priority", 0)),
                                          ptr(st, C.c_double), ptr(sp, C.c_double)))
            if m.get("lorentz"):
                prm = dispersion_params(m["lorentz"])
                check_op(L.emsop_set_dispersion(self.h, prm.shape[0], ptr(prm, C.c_double)))
            elif m.get("debye"):
                prm = debye_params(m["debye"])
                check_op(L.emsop_set_debye(self.h, prm.shape[0], ptr(prm, C.c_double)))
        for m in s.get("metals", []):
            if "curve" in m:
                pts = np.ascontiguousarray(np.asarray(m["curve"], np.float64).reshape(-1, 3))
                check_op(L.emsop_add_metal_curve(self.h, pts.shape[0], ptr(pts, C.c_double)))
                continue
            st, sp = _v3(m["start"]), _v3(m["stop"])
            if "radius" in m:
                check_op(L.emsop_add_metal_cylinder(self.h, int(m.get("priority", 0)), ptr(st, C.c_double),
                                                    ptr(sp, C.c_double), float(m["radius"]),
                                                    float(m.get("shell_width", 0.0))))
                continue
            check_op(L.emsop_add_metal(self.h, int(m.get("priority", 0)), ptr(st, C.c_double), ptr(sp, C.c_double)))
        for le in s.get("lumped", []):
            st, sp = _v3(le["start"]), _v3(le["stop"])
            check_op(L.emsop_add_lumped(self.h, int(le["dir"]), float(le.get("R", -1.0)), float(le.get("C", 0.0)),
                                        float(le.get("L", 0.0)), int(bool(le.get("caps", True))),
                                        int(le.get("type", "parallel") == "parallel"), int(le.get("priority", 0)),
                                        ptr(st, C.c_double), ptr(sp, C.c_double)))
        for e in s.get("excitations", []):
            if "curve" in e:
                pts = np.ascontiguousarray(np.asarray(e["curve"], np.float64).reshape(-1, 3))
                val = _v3(e.get("val", 0.0))
                check_op(L.emsop_add_excitation_curve(self.h, int(e.get("type", 0)), ptr(val, C.c_double),
                                                      float(e.get("delay", 0.0)), pts.shape[0], ptr(pts, C.c_double)))
            else:
                st, sp, val = _v3(e["start"]), _v3(e["stop"]), _v3(e.get("val", 0.0))
                if "radius" in e:
                    check_op(L.emsop_add_excitation_cylinder(
                        self.h, int(e.get("type", 0)), ptr(val, C.c_double), float(e.get("delay", 0.0)),
                        int(e.get("priority", 0)), ptr(st, C.c_double), ptr(sp, C.c_double), float(e["radius"]),
                        float(e.get("shell_width", 0.0))))
                else:
                    check_op(L.emsop_add_excitation(self.h, int(e.get("type", 0)), ptr(val, C.c_double),
                                                    float(e.get("delay", 0.0)), int(e.get("priority", 0)),
                                                    ptr(st, C.c_double), ptr(sp, C.c_double)))
            if e.get("weight") is not None:
                w = (C.c_char_p * 3)(*[None if f is None else str(f).encode() for f in e["weight"]])
                check_op(L.emsop_set_excitation_weight(self.h, w))
        bc = np.asarray(s.get("bc", [0] * 6), np.int32)
        pml = np.asarray(s.get("pml_size", [8] * 6), np.uint32)
        vph = np.asarray(s.get("mur_vph", [0.0] * 6), np.float64)
        check_op(L.emsop_set_boundary(self.h, ptr(bc, C.c_int), ptr(pml, C.c_uint), ptr(vph, C.c_double)))
        sig = s["signal"]
        if sig["type"] == "custom":
            check_op(L.emsop_set_custom_signal(self.h, str(sig["function"]).encode(), float(sig.get("f0", 0.0))))
        else:
            check_op(L.emsop_set_signal(self.h, SIGNALS[sig["type"]], float(sig.get("f0", 0.0)),
                                        float(sig.get("fc", 0.0))))
        check_op(L.emsop_set_timestep(self.h, int(s.get("ts_method", 3)), float(s.get("forced_dt", 0.0)),
                                      float(s.get("ts_factor", 1.0))))
        check_op(L.emsop_set_cell_constant_material(self.h, int(bool(s.get("cell_constant_material", False)))))
        check_op(L.emsop_set_threads(self.h, int(nthreads)))

    # ---- dispersive materials and lumped RLC elements -----------------------------
    def lorentz_tables(self):
        """Operator_Ext_LorentzMaterial tables, one dict per dispersion order in
        the form EngineHIP(lorentz=...) takes (empty without dispersive materials)."""
        L = lib()
        out = []
        for o in range(L.emsop_lorentz_orders(self.h)):
            cnt = C.c_uint()
            on = (C.c_int * 4)()
            pos = (C.POINTER(C.c_uint) * 3)()
            tabs = (C.POINTER(C.c_float) * 18)()
            check_op(L.emsop_lorentz_get(self.h, o, C.byref(cnt), on, pos, tabs))
            K = cnt.value

            def arr(k):
                return np.array([np.ctypeslib.as_array(tabs[3 * k + n], (K,)) if K else np.zeros(0, np.float32)
                                 for n in range(3)], np.float32)
            d = {"pos": np.array([np.ctypeslib.as_array(pos[n], (K,)) if K else np.zeros(0, np.uint32)
                                  for n in range(3)], np.uint32),
                 "volt": bool(on[0]), "curr": bool(on[1])}
            d["v_int"], d["v_ext"] = (arr(0), arr(1)) if on[0] else (np.ones((3, K), np.float32), np.zeros((3, K), np.float32))
            d["i_int"], d["i_ext"] = (arr(2), arr(3)) if on[1] else (np.ones((3, K), np.float32), np.zeros((3, K), np.float32))
            d["v_lor"] = arr(4) if on[2] else None
            d["i_lor"] = arr(5) if on[3] else None
            out.append(d)
        return out

    def rlc_table(self):
        """Operator_Ext_LumpedRLC table in the form EngineHIP(rlc=...) takes, or None."""
        L = lib()
        K = L.emsop_rlc_count(self.h)
        if K == 0:
            return None
        pos = (C.POINTER(C.c_uint) * 3)()
        dr = C.POINTER(C.c_int)()
        coef = (C.POINTER(C.c_float) * 9)()
        check_op(L.emsop_rlc_get(self.h, pos, C.byref(dr), coef))
        d = {"pos": np.array([np.ctypeslib.as_array(pos[n], (K,)) for n in range(3)], np.uint32),
             "dir": np.ctypeslib.as_array(dr, (K,)).astype(np.int32)}
        for k, name in enumerate(RLC_COEFS):
            d[name] = np.ctypeslib.as_array(coef[k], (K,)).astype(np.float32)
        return d

    # ---- scalars ------------------------------------------------------------
    @property
    def dT(self):
        return lib().emsop_timestep(self.h)

    @property
    def nyquist(self):
        return lib().emsop_nyquist(self.h)

    @property
    def period(self):
        return lib().emsop_signal_period(self.h)

    @property
    def period_steps(self):
        return lib().emsop_signal_period_steps(self.h)

    @property
    def signal_length(self):
        return lib().emsop_signal_length(self.h)

    @property
    def num_cells(self):
        return int(self.N[0]) * int(self.N[1]) * int(self.N[2])

    # ---- arrays (copies) ------------------------------------------------------
    def _arr(self, p, count, dtype=np.float32):
        if not p:
            raise RuntimeError("array not available")
        return np.ctypeslib.as_array(p, shape=(count,)).astype(dtype, copy=True)

    def signal(self):
        L = lib()
        n = self.signal_length
        return self._arr(L.emsop_signal_volt(self.h), n), self._arr(L.emsop_signal_curr(self.h), n)

    def coeff(self, which):
        """which in 'vv','vi','ii','iv' -> float32 [3][Nx][Ny][Nz]."""
        k = ("vv", "vi", "ii", "iv").index(which)
        return self._arr(lib().emsop_coeff(self.h, k), 3 * self.num_cells).reshape((3,) + self.N)

    def coeff_ptr(self, which):
        k = ("vv", "vi", "ii", "iv").index(which)
        return lib().emsop_coeff(self.h, k)

    def ec(self, which):
        """which in 'C','G','L','R' -> float32 [3][Nx][Ny][Nz]."""
        k = ("C", "G", "L", "R").index(which)
        return self._arr(lib().emsop_ec(self.h, k), 3 * self.num_cells).reshape((3,) + self.N)

    def effmat_at(self, ny, pos):
        out = (C.c_double * 4)()
        p = (C.c_uint * 3)(*pos)
        check_op(lib().emsop_effmat_at(self.h, ny, p, out))
        return tuple(out)

    def excitation(self, curr=False):
        L = lib()
        K = L.emsop_exc_count(self.h, int(curr))
        idx = np.zeros((3, K), np.uint32)
        d = np.zeros(K, np.uint16)
        a = np.zeros(K, np.float32)
        dl = np.zeros(K, np.uint32)
        if K:
            check_op(L.emsop_exc_get(self.h, int(curr), ptr(idx, C.c_uint), ptr(d, C.c_ushort), ptr(a, C.c_float),
                                     ptr(dl, C.c_uint)))
        return {"index": idx, "dir": d, "amp": a, "delay": dl}

    def upml_slabs(self):
        L = lib()
        out = []
        for k in range(L.emsop_upml_count(self.h)):
            st = (C.c_uint * 3)()
            nm = (C.c_uint * 3)()
            arr = (fp * 6)()
            check_op(L.emsop_upml_get(self.h, k, st, nm, arr))
            num = tuple(nm)
            cnt = 3 * num[0] * num[1] * num[2]
            names = ("vv", "vvfo", "vvfn", "ii", "iifo", "iifn")
            d = {"start": tuple(st), "num": num, "_ptrs": [arr[i] for i in range(6)]}
            for i, nme in enumerate(names):
                d[nme] = self._arr(arr[i], cnt).reshape((3,) + num)
            out.append(d)
        return out

    def mur_faces(self):
        L = lib()
        out = []
        for k in range(L.emsop_mur_count(self.h)):
            ny = C.c_int()
            line, shift, sts = C.c_uint(), C.c_uint(), C.c_uint()
            cP, cPP = fp(), fp()
            check_op(L.emsop_mur_get(self.h, k, C.byref(ny), C.byref(line), C.byref(shift), C.byref(sts),
                                     C.byref(cP), C.byref(cPP)))
            n = ny.value
            na, nb = self.N[(n + 1) % 3], self.N[(n + 2) % 3]
            out.append({"ny": n, "line": line.value, "line_shift": shift.value, "start_ts": sts.value,
                        "coeff_nyP": self._arr(cP, na * nb).reshape(na, nb),
                        "coeff_nyPP": self._arr(cPP, na * nb).reshape(na, nb)})
        return out

    # ---- mesh queries ---------------------------------------------------------
    def disc_line(self, n, pos, dual=False):
        return lib().emsop_disc_line(self.h, n, int(pos), int(dual))

    def edge_length(self, n, pos, dual=False):
        return lib().emsop_edge_length(self.h, n, int(pos), int(dual))

    def edge_lengths(self, n, dual=False):
        return np.array([self.edge_length(n, i, dual) for i in range(self.N[n])])

    def snap_to_mesh(self, coord, dual=False):
        c = np.asarray(coord, np.float64)
        out = (C.c_uint * 3)()
        ins = (C.c_int * 3)()
        lib().emsop_snap_to_mesh(self.h, ptr(c, C.c_double), int(dual), out, ins)
        return list(out), [bool(v) for v in ins]

    def snap_box(self, start, stop, dual=False, method=0):
        """SnapBox2Mesh (FDTD/operator.cpp:297-358): (dim, start index, stop
        index, start inside, stop inside).  Method 0 snaps each corner to its
        nearest line; method 1 then widens a non-flat axis to the next line
        outside where the snapped line falls inside the box."""
        a = np.asarray(start, np.float64)
        b = np.asarray(stop, np.float64)
        us, ue = (C.c_uint * 3)(), (C.c_uint * 3)()
        si, ei = (C.c_int * 3)(), (C.c_int * 3)()
        dim = lib().emsop_snap_box(self.h, ptr(a, C.c_double), ptr(b, C.c_double), int(dual), us, ue, si, ei)
        us, ue = list(us), list(ue)
        if method == 1 and dim >= 0:
            dim = 0
            for n in range(3):
                lo, hi = min(a[n], b[n]), max(a[n], b[n])
                if ue[n] > us[n]:
                    if self.disc_line(n, us[n], dual) > lo and us[n] > 0:
                        us[n] -= 1
                    if self.disc_line(n, ue[n], dual) < hi and ue[n] < self.N[n] - 1:
                        ue[n] += 1
                if ue[n] > us[n]:
                    dim += 1
        return dim, us, ue, [bool(v) for v in si], [bool(v) for v in ei]

    def find_path(self, start, stop):
        """Operator::FindPath (FDTD/operator.cpp:390-496): list of (dir, (x, y, z))
        mesh edges from the node nearest start to the node nearest stop."""
        a = np.asarray(start, np.float64)
        b = np.asarray(stop, np.float64)
        cnt = C.c_uint()
        # a step moves along an axis only in the sign of the line's direction,
        # so a path has at most sum(N) steps plus 3 closing edges
        cap = int(sum(self.N)) + 3
        pos = np.zeros((cap, 3), np.uint32)
        dirs = np.zeros(cap, np.int32)
        check_op(lib().emsop_find_path(self.h, ptr(a, C.c_double), ptr(b, C.c_double), cap, ptr(pos, C.c_uint),
                                       ptr(dirs, C.c_int), C.byref(cnt)))
        return [(int(dirs[t]), tuple(int(v) for v in pos[t])) for t in range(cnt.value)]

    def close(self):
        if getattr(self, "h", None):
            lib().emsop_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
