"""openEMS's Python interface (python/openEMS/openEMS.pyx, ports.py and the
CSXCAD objects they use) for the HIP engine, without CSXCAD.

A script written for the reference's Python binding runs with its imports
swapped::

    from openems_amd.pyopenems import openEMS, ContinuousStructure

    FDTD = openEMS(NrTS=10000, EndCriteria=1e-4)
    FDTD.SetGaussExcite(0, 20e9)
    FDTD.SetBoundaryCond(['PML_8'] * 6)
    CSX = ContinuousStructure()
    FDTD.SetCSX(CSX)
    mesh = CSX.GetGrid()
    mesh.SetDeltaUnit(1e-3)
    mesh.AddLine('x', np.linspace(-20, 20, 81)) ...
    CSX.AddMaterial('sub', epsilon=3.66).AddBox([-10, -10, 0], [10, 10, 1.5], priority=10)
    port = FDTD.AddLumpedPort(1, 50, [0, 0, 0], [0, 0, 1.5], 'z', excite=1)
    FDTD.Run(sim_path)
    port.CalcPort(sim_path, freqs)

The supported subset is what the operator builder takes: rectilinear
Cartesian meshes, Box and Curve primitives (and Cylinder / CylindricalShell
for metals and excitations), Material / Lorentz / Debye / Metal / Excitation
(with weight functions) / Probe / Dump / LumpedElement properties, lumped and microstrip ports
and Gaussian, sinusoidal, Dirac, step or custom excitation.  ``Run`` builds
the operator and the engine, runs up to NrTS steps or the end criterion, and
writes the probe files and HDF5 dumps into sim_path, as openEMS does.
Properties of different kinds with equal priorities over the same cell are
ranked materials, metals, lumped elements, excitations (CSXCAD ranks them by
creation order).
"""
from __future__ import annotations

import os
import shutil

import numpy as np

_DIRS = {"x": 0, "y": 1, "z": 2, 0: 0, 1: 1, 2: 2}


def _dir(d):
    if isinstance(d, str):
        d = d.lower()
        if d.startswith("r"):
            raise ValueError("cylindrical meshes are not supported")
    return _DIRS[d]


def DFT_time2freq(t, val, freq, signal_type="pulse"):
    """Single-sided DFT of a sampled time signal at the given frequencies,
    scaled by the sample spacing ('pulse') or 1/len ('periodic'), the
    transform the reference's port processing applies to probe files
    (python/openEMS/utilities.py:21-35)."""
    t = np.asarray(t, np.float64)
    val = np.asarray(val, np.float64)
    freq = np.atleast_1d(np.asarray(freq, np.float64))
    if len(t) != len(val) or freq.size == 0:
        raise ValueError("time and value arrays must match and freq must not be empty")
    out = np.array([np.sum(val * np.exp(-1j * 2 * np.pi * f * t)) for f in freq])
    if signal_type == "pulse":
        out = out * (t[1] - t[0])
    elif signal_type == "periodic":
        out = out / len(t)
    else:
        raise ValueError(f"unknown signal type {signal_type!r}")
    return 2 * out


# ----------------------------------------------------------------------------
# structure
# ----------------------------------------------------------------------------
class RectilinearGrid:
    def __init__(self):
        self.lines = [np.zeros(0), np.zeros(0), np.zeros(0)]
        self.delta_unit = 1.0

    def SetDeltaUnit(self, unit):
        self.delta_unit = float(unit)

    def GetDeltaUnit(self):
        return self.delta_unit

    def SetLines(self, d, lines):
        self.lines[_dir(d)] = np.unique(np.asarray(lines, np.float64).ravel())

    def AddLine(self, d, lines):
        n = _dir(d)
        self.lines[n] = np.unique(np.concatenate([self.lines[n], np.atleast_1d(np.asarray(lines, np.float64))]))

    def GetLines(self, d, do_sort=True):
        return self.lines[_dir(d)].copy()

    def GetQtyLines(self, d):
        return len(self.lines[_dir(d)])

    def ClearLines(self, d):
        self.lines[_dir(d)] = np.zeros(0)


class Property:
    """A CSXCAD property: a kind, its parameters, and its primitives."""

    def __init__(self, kind, name, **params):
        self.kind = kind
        self.name = name
        self.params = params
        self.prims = []          # ("box", priority, start, stop), ("curve", priority, points) or
                                 # ("cylinder", priority, start, stop, radius, shell width)

    def GetName(self):
        return self.name

    def AddBox(self, start, stop, priority=0, **kw):
        self.prims.append(("box", int(priority), [float(v) for v in start], [float(v) for v in stop]))
        return self

    def AddCurve(self, points, priority=0, **kw):
        """points: (3, N) as CSXCAD takes them (also for N = 3), or (N, 3)."""
        p = np.asarray(points, np.float64)
        if p.ndim != 2 or 3 not in p.shape:
            raise ValueError("curve points must be (3, N) or (N, 3)")
        if p.shape[0] == 3:          # CSXCAD's (3, N) layout
            p = p.T
        self.prims.append(("curve", int(priority), [tuple(r) for r in p.tolist()]))
        return self

    def AddCylinder(self, start, stop, radius, priority=0, **kw):
        """CSPrimCylinder: axis start -> stop, radius (metal and excitation properties)."""
        self.prims.append(("cylinder", int(priority), [float(v) for v in start], [float(v) for v in stop],
                           float(radius), 0.0))
        return self

    def AddCylindricalShell(self, start, stop, radius, shell_width, priority=0, **kw):
        """CSPrimCylindricalShell: mid radius and wall width (metal and excitation properties)."""
        self.prims.append(("cylinder", int(priority), [float(v) for v in start], [float(v) for v in stop],
                           float(radius), float(shell_width)))
        return self

    def SetWeightFunction(self, func):
        """Excitation weights, one expression per direction in x, y, z, rho, r, a."""
        if len(func) != 3:
            raise ValueError("a weight function per direction (3) is needed")
        self.params["weight"] = [str(f) for f in func]

    def SetMaterialProperty(self, **kw):
        self.params.update(kw)

    def SetEnabled(self, val):
        self.params["enabled"] = bool(val)


class ContinuousStructure:
    def __init__(self, **kw):
        self.grid = RectilinearGrid()
        self.props = []
        self.background = {"eps": 1.0, "mue": 1.0, "kappa": 0.0, "sigma": 0.0}

    def GetGrid(self):
        return self.grid

    def _add(self, kind, name, **params):
        p = Property(kind, name, **params)
        self.props.append(p)
        return p

    def AddMaterial(self, name, epsilon=1.0, mue=1.0, kappa=0.0, sigma=0.0, **kw):
        return self._add("material", name, epsilon=epsilon, mue=mue, kappa=kappa, sigma=sigma, **kw)

    def AddLorentzMaterial(self, name, epsilon=1.0, mue=1.0, kappa=0.0, sigma=0.0, **kw):
        """Dispersion terms as SetMaterialProperty keywords: eps_plasma_frequency,
        eps_relaxation_time, eps_lorentz_pole_frequency and the mue_ ones,
        scalars or per-order lists."""
        return self._add("lorentz", name, epsilon=epsilon, mue=mue, kappa=kappa, sigma=sigma, **kw)

    def AddDebyeMaterial(self, name, epsilon=1.0, mue=1.0, kappa=0.0, sigma=0.0, **kw):
        """eps_delta and eps_relaxation_time as keywords, scalars or per-order lists."""
        return self._add("debye", name, epsilon=epsilon, mue=mue, kappa=kappa, sigma=sigma, **kw)

    def AddMetal(self, name, **kw):
        return self._add("metal", name)

    def AddExcitation(self, name, exc_type, exc_val, delay=0.0, **kw):
        return self._add("excitation", name, exc_type=int(exc_type), exc_val=[float(v) for v in exc_val],
                         delay=float(delay))

    def AddProbe(self, name, p_type, weight=1.0, norm_dir=-1, frequency=None, mode_function=None, **kw):
        return self._add("probe", name, p_type=int(p_type), weight=float(weight), norm_dir=int(norm_dir),
                         frequency=list(np.atleast_1d(frequency)) if frequency is not None else [],
                         mode_function=None if mode_function is None else [str(f) for f in mode_function])

    def AddDump(self, name, dump_type=0, dump_mode=1, file_type=1, frequency=None, **kw):
        return self._add("dump", name, dump_type=int(dump_type), dump_mode=int(dump_mode),
                         file_type=int(file_type),
                         frequency=list(np.atleast_1d(frequency)) if frequency is not None else [])

    def AddLumpedElement(self, name, ny, caps=True, R=None, C=None, L=None, LEtype=0, **kw):
        return self._add("lumped", name, ny=_dir(ny), caps=bool(caps), R=-1.0 if R is None else float(R),
                         C=0.0 if C is None else float(C), L=0.0 if L is None else float(L), LEtype=int(LEtype))

    def SetBackgroundMaterial(self, epsilon=1.0, mue=1.0, kappa=0.0, sigma=0.0):
        self.background = {"eps": epsilon, "mue": mue, "kappa": kappa, "sigma": sigma}

    # ---- scene dict for the operator builder
    def _orders(self, p, keys):
        vals = [np.atleast_1d(p.params.get(k, 0.0)) for k, _ in keys]
        n = max(len(v) for v in vals)
        return [{dst: float(v[o]) if o < len(v) else 0.0 for (k, dst), v in zip(keys, vals)} for o in range(n)]

    def scene_parts(self):
        mats, metals, lumped, excs, probes, dumps = [], [], [], [], [], []
        for p in self.props:
            P = p.params
            for prim in p.prims:
                kind, prio = prim[0], prim[1]
                if kind in ("curve", "cylinder") and p.kind not in ("metal", "excitation"):
                    raise ValueError(f"{kind} primitives of {p.kind} {p.name!r} are not supported")
                cyl = {"radius": prim[4], "shell_width": prim[5]} if kind == "cylinder" else {}
                if p.kind in ("material", "lorentz", "debye"):
                    m = {"eps": P["epsilon"], "mue": P["mue"], "kappa": P["kappa"], "sigma": P["sigma"],
                         "priority": prio, "start": prim[2], "stop": prim[3]}
                    if p.kind == "lorentz":
                        m["lorentz"] = self._orders(p, (("eps_plasma_frequency", "eps_plasma"),
                                                        ("eps_relaxation_time", "eps_relax"),
                                                        ("eps_lorentz_pole_frequency", "eps_pole"),
                                                        ("mue_plasma_frequency", "mue_plasma"),
                                                        ("mue_relaxation_time", "mue_relax"),
                                                        ("mue_lorentz_pole_frequency", "mue_pole")))
                    elif p.kind == "debye":
                        m["debye"] = self._orders(p, (("eps_delta", "eps_delta"),
                                                      ("eps_relaxation_time", "eps_relax")))
                    mats.append(m)
                elif p.kind == "metal":
                    metals.append({"curve": prim[2]} if kind == "curve" else
                                  dict(priority=prio, start=prim[2], stop=prim[3], **cyl))
                elif p.kind == "excitation":
                    if not P.get("enabled", True):
                        continue
                    e = {"type": P["exc_type"], "val": tuple(P["exc_val"]), "delay": P["delay"]}
                    if P.get("weight") is not None:
                        e["weight"] = list(P["weight"])
                    excs.append(dict(e, curve=prim[2]) if kind == "curve" else
                                dict(e, priority=prio, start=prim[2], stop=prim[3], **cyl))
                elif p.kind == "lumped":
                    lumped.append({"dir": P["ny"], "R": P["R"], "C": P["C"], "L": P["L"], "caps": P["caps"],
                                   "type": "series" if P["LEtype"] == 1 else "parallel", "priority": prio,
                                   "start": prim[2], "stop": prim[3]})
                elif p.kind == "probe":
                    probes.append({"type": P["p_type"], "name": p.name, "start": prim[2], "stop": prim[3],
                                   "weight": P["weight"], "freqs": P["frequency"], "norm_dir": P["norm_dir"]})
                    if P.get("mode_function") is not None:
                        probes[-1]["mode_function"] = [str(f) for f in P["mode_function"]]
                elif p.kind == "dump":
                    dt = P["dump_type"]
                    if dt not in (0, 1, 10, 11):
                        raise ValueError(f"dump type {dt} of {p.name!r} is not supported")
                    d = {"name": p.name, "start": prim[2], "stop": prim[3], "field": "E" if dt in (0, 10) else "H",
                         "interp": P["dump_mode"], "file_type": P["file_type"]}
                    if dt < 10:
                        d["domain"] = "td"
                    else:
                        d["freqs"] = P["frequency"]
                    dumps.append(d)
        return {"materials": mats, "metals": metals, "lumped": lumped, "excitations": excs,
                "probes": probes, "dumps": dumps}


# ----------------------------------------------------------------------------
# ports
# ----------------------------------------------------------------------------
# python/openEMS/physical_constants.py
C0 = 299792458.0
MUE0 = 4e-7 * np.pi
EPS0 = 1.0 / (MUE0 * C0 ** 2)
Z0 = np.sqrt(MUE0 / EPS0)


class UI_data:
    """Voltage / current probe files of a port read back (np.loadtxt with '%'
    comments, as the reference's port processing reads them)."""

    def __init__(self, fns, path, freq, signal_type="pulse"):
        self.fns = [fns] if isinstance(fns, str) else list(fns)
        self.freq = np.atleast_1d(freq)
        self.ui_time, self.ui_val, self.ui_f_val = [], [], []
        for fn in self.fns:
            d = np.loadtxt(os.path.join(path, fn), comments="%")
            self.ui_time.append(d[:, 0])
            self.ui_val.append(d[:, 1])
            self.ui_f_val.append(DFT_time2freq(d[:, 0], d[:, 1], self.freq, signal_type))


class LumpedPort:
    """A lumped port (ports.LumpedPort): a resistor (or a metal short for
    R = 0) over the port box, an optional soft excitation against the port
    direction, a voltage probe along it (weight -1) and a current probe
    across its middle; CalcPort splits the port voltage and current into
    incident and reflected waves against Z_ref = R."""

    def __init__(self, CSX, port_nr, R, start, stop, exc_dir, excite=0, **kw):
        self.CSX = CSX
        self.number = port_nr
        self.R = R
        self.start = np.array(start, np.float64)
        self.stop = np.array(stop, np.float64)
        self.excite = excite
        self.priority = kw.get("priority", 0)
        self.delay = kw.get("delay", 0.0)
        prefix = kw.get("PortNamePrefix", "")
        lbl = prefix + "port_{}" + "_{}".format(port_nr)
        self.Z_ref = None
        self.port_props = []
        ny = _dir(exc_dir)
        self.exc_ny = ny
        if self.start[ny] == self.stop[ny]:
            raise ValueError("LumpedPort: start and stop may not be identical in excitation direction")
        self.direction = np.sign(self.stop[ny] - self.start[ny])
        if R > 0:
            CSX.AddLumpedElement(lbl.format("resist"), ny=ny, caps=True, R=R).AddBox(
                self.start, self.stop, priority=self.priority)
        elif R == 0:
            CSX.AddMetal(lbl.format("resist")).AddBox(self.start, self.stop, priority=self.priority)
        if excite != 0:
            v = np.zeros(3)
            v[ny] = -1 * self.direction * excite
            self.port_props.append(CSX.AddExcitation(lbl.format("excite"), exc_type=0, exc_val=v,
                                                     delay=self.delay))
            self.port_props[-1].AddBox(
                self.start, self.stop, priority=self.priority)
        self.U_filenames = [lbl.format("ut")]
        us, ue = 0.5 * (self.start + self.stop), 0.5 * (self.start + self.stop)
        us[ny], ue[ny] = self.start[ny], self.stop[ny]
        CSX.AddProbe(self.U_filenames[0], p_type=0, weight=-1).AddBox(us, ue)
        self.I_filenames = [lbl.format("it")]
        i_s, i_e = self.start.copy(), self.stop.copy()
        i_s[ny] = i_e[ny] = 0.5 * (self.start[ny] + self.stop[ny])
        CSX.AddProbe(self.I_filenames[0], p_type=1, weight=self.direction, norm_dir=ny).AddBox(i_s, i_e)

    def SetEnabled(self, val):
        """Switches the port's excitation on or off (ports.Port.SetEnabled,
        python/openEMS/ports.py:86-96); enabling a port without one raises."""
        found = False
        for prop in self.port_props:
            if prop.kind == "excitation":
                prop.SetEnabled(val)
                found = True
        if not found and val:
            raise ValueError("Unable to enable port! No excitation found!")

    def ReadUIData(self, sim_path, freq, signal_type="pulse"):
        self.u_data = UI_data(self.U_filenames, sim_path, freq, signal_type)
        self.uf_tot = sum(self.u_data.ui_f_val)
        self.ut_tot = sum(self.u_data.ui_val)
        self.i_data = UI_data(self.I_filenames, sim_path, freq, signal_type)
        self.if_tot = sum(self.i_data.ui_f_val)
        self.it_tot = sum(self.i_data.ui_val)

    def CalcPort(self, sim_path, freq, ref_impedance=None, ref_plane_shift=None, signal_type="pulse"):
        """Incident / reflected split (ports.Port.CalcPort, python/openEMS/ports.py:108-150):
        against ref_impedance, else the port's own Z_ref (R for a lumped port,
        the measured line impedance for a microstrip port); ref_plane_shift
        moves the reference plane along a line port by beta * shift."""
        self.ReadUIData(sim_path, freq, signal_type)
        if ref_impedance is not None:
            self.Z_ref = ref_impedance
        if self.Z_ref is None:
            self.Z_ref = getattr(self, "R", None)
        if self.Z_ref is None:
            raise ValueError("Port Z_ref should not be None!")
        Z = self.Z_ref
        if ref_plane_shift is not None:
            if not hasattr(self, "beta"):
                raise ValueError("Port has no beta attribute!")
            shift = ref_plane_shift
            if getattr(self, "measplane_shift", 0):
                shift -= self.measplane_shift
            shift *= self.CSX.GetGrid().GetDeltaUnit()
            phase = np.real(self.beta) * shift
            uf = self.uf_tot * np.cos(-phase) + 1j * self.if_tot * Z * np.sin(-phase)
            itf = self.if_tot * np.cos(-phase) + 1j * self.uf_tot / Z * np.sin(-phase)
            self.uf_tot, self.if_tot = uf, itf
        self.uf_inc = 0.5 * (self.uf_tot + self.if_tot * Z)
        self.if_inc = 0.5 * (self.if_tot + self.uf_tot / Z)
        self.uf_ref = self.uf_tot - self.uf_inc
        self.if_ref = self.if_inc - self.if_tot
        if isinstance(Z, (int, float, np.floating)):
            self.ut_inc = 0.5 * (self.ut_tot + self.it_tot * Z)
            self.it_inc = 0.5 * (self.it_tot + self.ut_tot / Z)
            self.ut_ref = self.ut_tot - self.ut_inc
            self.it_ref = self.it_inc - self.it_tot
        self.P_inc = 0.5 * np.real(self.uf_inc * np.conj(self.if_inc))
        self.P_ref = 0.5 * np.real(self.uf_ref * np.conj(self.if_ref))
        self.P_acc = 0.5 * np.real(self.uf_tot * np.conj(self.if_tot))


class WaveguidePort(LumpedPort):
    """A waveguide port (ports.WaveguidePort, python/openEMS/ports.py:341-400):
    a soft E excitation weighted by the mode's E pattern in the plane
    start[exc] (when excited), and mode-matching probes (types 10 and 11) of
    the E and H patterns in the plane stop[exc]; CalcPort takes the analytic
    wave impedance k Z0 / beta (beta = sqrt(k^2 - kc^2)) as Z_ref."""

    def __init__(self, CSX, port_nr, start, stop, exc_dir, E_WG_func, H_WG_func, kc, excite=0, **kw):
        self.CSX = CSX
        self.number = port_nr
        self.start = np.array(start, np.float64)
        self.stop = np.array(stop, np.float64)
        self.excite = excite
        self.priority = kw.get("priority", 0)
        self.delay = kw.get("delay", 0.0)
        self.Z_ref = None
        self.port_props = []
        lbl = kw.get("PortNamePrefix", "") + "port_{}" + "_{}".format(port_nr)
        ny = _dir(exc_dir)
        self.exc_ny, self.ny_P, self.ny_PP = ny, (ny + 1) % 3, (ny + 2) % 3
        self.direction = np.sign(self.stop[ny] - self.start[ny])
        self.ref_index = 1
        if excite != 0 and self.stop[ny] == self.start[ny]:
            raise ValueError("Port length in excitation direction may not be zero if port is excited!")
        self.kc = kc
        self.E_func, self.H_func = E_WG_func, H_WG_func
        if excite != 0:
            a, b = self.start.copy(), self.stop.copy()
            b[ny] = a[ny]
            v = np.ones(3)
            v[ny] = 0
            exc = CSX.AddExcitation(lbl.format("excite"), exc_type=0, exc_val=v, delay=self.delay)
            self.port_props.append(exc)
            exc.SetWeightFunction([str(x) for x in self.E_func])
            exc.AddBox(a, b, priority=self.priority)
        a, b = self.start.copy(), self.stop.copy()
        a[ny] = b[ny]
        self.measplane_shift = abs(self.stop[ny] - self.start[ny])
        self.U_filenames = [lbl.format("ut")]
        CSX.AddProbe(self.U_filenames[0], p_type=10, mode_function=self.E_func).AddBox(a, b)
        self.I_filenames = [lbl.format("it")]
        CSX.AddProbe(self.I_filenames[0], p_type=11, weight=self.direction, mode_function=self.H_func).AddBox(a, b)

    def CalcPort(self, sim_path, freq, ref_impedance=None, ref_plane_shift=None, signal_type="pulse"):
        k = 2.0 * np.pi * np.asarray(freq, np.float64) / C0 * self.ref_index
        self.beta = np.sqrt(k ** 2 - self.kc ** 2 + 0j)
        self.ZL = k * Z0 / self.beta
        if ref_impedance is None:
            self.Z_ref = self.ZL
        super().CalcPort(sim_path, freq, ref_impedance, ref_plane_shift, signal_type)


class RectWGPort(WaveguidePort):
    """A rectangular waveguide port (ports.RectWGPort, ports.py:402-460): the
    TE_mn patterns of an a x b guide (sizes in meters) written as mode
    functions of the two transverse coordinates relative to start, in
    drawing units; TE modes only."""

    def __init__(self, CSX, port_nr, start, stop, exc_dir, a, b, mode_name, excite=0, **kw):
        ny = _dir(exc_dir)
        nP, nPP = (ny + 1) % 3, (ny + 2) % 3
        if len(mode_name) != 4:
            raise ValueError("Invalid mode definition")
        if not mode_name.startswith("TE"):
            raise ValueError("Currently only TE-modes are supported! Mode found: {}".format(mode_name))
        self.WG_size, self.WG_mode = [a, b], mode_name
        M, N = float(mode_name[2]), float(mode_name[3])
        self.M, self.N = M, N
        unit = CSX.GetGrid().GetDeltaUnit()
        xyz = "xyz"
        sP = "({}-{})".format(xyz[nP], float(start[nP]))
        sPP = "({}-{})".format(xyz[nPP], float(start[nPP]))
        kc = np.sqrt((M * np.pi / a) ** 2 + (N * np.pi / b) ** 2)
        a, b = a / unit, b / unit
        E, H = ["0", "0", "0"], ["0", "0", "0"]
        if N > 0:
            E[nP] = "{}*cos({}*{})*sin({}*{})".format(N / b, M * np.pi / a, sP, N * np.pi / b, sPP)
            H[nPP] = "{}*cos({}*{})*sin({}*{})".format(N / b, M * np.pi / a, sP, N * np.pi / b, sPP)
        if M > 0:
            E[nPP] = "{}*sin({}*{})*cos({}*{})".format(-1 * M / a, M * np.pi / a, sP, N * np.pi / b, sPP)
            H[nP] = "{}*sin({}*{})*cos({}*{})".format(M / a, M * np.pi / a, sP, N * np.pi / b, sPP)
        super().__init__(CSX, port_nr, start, stop, exc_dir, E, H, kc, excite, **kw)


class MSLPort(LumpedPort):
    """A microstrip port (ports.MSLPort, python/openEMS/ports.py:190-340): the
    strip as a metal box in the plane start[exc] of the port box, three
    voltage probes across the substrate at the mesh lines around the
    measurement plane (half the port length by default, MeasPlaneShift), two
    current probes around the strip between them, an optional soft feed at
    FeedShift and a feed resistance Feed_R at the port start.  ReadUIData
    takes the middle voltage and the mean current and derives the
    propagation constant beta and the line impedance Z_ref from the
    differences (the method of the reference's port)."""

    def __init__(self, CSX, port_nr, metal_prop, start, stop, prop_dir, exc_dir, excite=0, **kw):
        self.CSX = CSX
        self.number = port_nr
        self.start = np.array(start, np.float64)
        self.stop = np.array(stop, np.float64)
        self.excite = excite
        self.priority = kw.get("priority", 0)
        self.delay = kw.get("delay", 0.0)
        self.Z_ref = None
        self.port_props = []
        lbl = kw.get("PortNamePrefix", "") + "port_{}" + "_{}".format(port_nr)
        ex, pr = _dir(exc_dir), _dir(prop_dir)
        self.exc_ny, self.prop_ny = ex, pr
        if not np.all(self.start != self.stop):
            raise ValueError("Start coordinate must not be equal to stop coordinate")
        if ex == pr:
            raise ValueError("Excitation direction must not be equal to propagation direction")
        self.direction = np.sign(self.stop[pr] - self.start[pr])
        self.upside_down = np.sign(self.stop[ex] - self.start[ex])
        self.feed_shift = kw.get("FeedShift", 0)
        self.measplane_shift = kw.get("MeasPlaneShift", 0.5 * abs(self.start[pr] - self.stop[pr]))
        self.measplane_pos = self.start[pr] + self.measplane_shift * self.direction
        self.feed_R = kw.get("Feed_R", np.inf)
        strip_stop = self.stop.copy()
        strip_stop[ex] = self.start[ex]
        metal_prop.AddBox(self.start, strip_stop, priority=self.priority)
        lines = CSX.GetGrid().GetLines(pr)
        if len(lines) <= 5:
            raise ValueError("At least 5 lines in propagation direction required!")
        k = int(np.clip(np.argmin(np.abs(lines - self.measplane_pos)), 1, len(lines) - 2))
        self.measplane_shift = abs(self.start[pr] - lines[k])
        idx = np.array([k - 1, k, k + 1])
        if self.direction < 0:
            idx = idx[::-1]
        u_pos = lines[idx]
        self.U_delta = np.diff(u_pos)
        self.U_filenames = []
        mid = 0.5 * (self.start + self.stop)
        for n, sfx in enumerate("ABC"):
            a, b = mid.copy(), mid.copy()
            a[pr] = b[pr] = u_pos[n]
            a[ex], b[ex] = self.start[ex], self.stop[ex]
            self.U_filenames.append(lbl.format("ut") + sfx)
            CSX.AddProbe(self.U_filenames[-1], p_type=0).AddBox(a, b)
        i_pos = u_pos[:2] + np.diff(u_pos) / 2.0
        self.I_delta = np.diff(i_pos)
        self.I_filenames = []
        a, b = self.start.copy(), self.stop.copy()
        b[ex] = self.start[ex]
        for n, sfx in enumerate("AB"):
            a[pr] = b[pr] = i_pos[n]
            self.I_filenames.append(lbl.format("it") + sfx)
            CSX.AddProbe(self.I_filenames[-1], p_type=1, weight=self.direction, norm_dir=pr).AddBox(a.copy(), b.copy())
        if excite != 0:
            kf = int(np.argmin(np.abs(lines - (self.start[pr] + self.feed_shift * self.direction))))
            a, b = self.start.copy(), self.stop.copy()
            a[pr] = b[pr] = lines[kf]
            v = np.zeros(3)
            v[ex] = -1 * self.upside_down * excite
            self.port_props.append(CSX.AddExcitation(lbl.format("excite"), exc_type=0, exc_val=v,
                                                     delay=self.delay))
            self.port_props[-1].AddBox(
                a, b, priority=self.priority)
        if self.feed_R >= 0 and not np.isinf(self.feed_R):
            a, b = self.start.copy(), self.stop.copy()
            b[pr] = a[pr]
            if self.feed_R == 0:
                metal_prop.AddBox(a, b)
            else:
                CSX.AddLumpedElement(lbl.format("resist"), ny=ex, caps=True, R=self.feed_R).AddBox(a, b)

    def ReadUIData(self, sim_path, freq, signal_type="pulse"):
        self.u_data = UI_data(self.U_filenames, sim_path, freq, signal_type)
        self.uf_tot = self.u_data.ui_f_val[1]
        self.ut_tot = self.u_data.ui_val[1]
        self.i_data = UI_data(self.I_filenames, sim_path, freq, signal_type)
        self.if_tot = 0.5 * (self.i_data.ui_f_val[0] + self.i_data.ui_f_val[1])
        self.it_tot = 0.5 * (self.i_data.ui_val[0] + self.i_data.ui_val[1])
        unit = self.CSX.GetGrid().GetDeltaUnit()
        Et = self.u_data.ui_f_val[1]
        dEt = (self.u_data.ui_f_val[2] - self.u_data.ui_f_val[0]) / (np.sum(np.abs(self.U_delta)) * unit)
        Ht = self.if_tot
        dHt = (self.i_data.ui_f_val[1] - self.i_data.ui_f_val[0]) / (np.abs(self.I_delta[0]) * unit)
        beta = np.sqrt(-dEt * dHt / (Ht * Et))
        beta[np.real(beta) < 0] *= -1
        self.beta = beta
        self.Z_ref = np.sqrt(Et * dEt / (Ht * dHt))


# ----------------------------------------------------------------------------
# simulation
# ----------------------------------------------------------------------------
class openEMS:
    """The reference's openEMS Python class (python/openEMS/openEMS.pyx:73-579)
    on the HIP engine."""

    def __init__(self, NrTS=1e9, EndCriteria=1e-5, OverSampling=None, TimeStepFactor=None, TimeStepMethod=None,
                 CellConstantMaterial=None, CoordSystem=0, **kw):
        if CoordSystem not in (0, None):
            raise ValueError("cylindrical coordinates are not supported")
        self.nrts = int(NrTS)
        self.end_criteria = float(EndCriteria)
        self.oversampling = OverSampling
        self.ts_factor = TimeStepFactor
        self.ts_method = TimeStepMethod
        self.time_step = None
        self.cell_constant = CellConstantMaterial
        self.signal = None
        self.bc = [0] * 6
        self.pml = [8] * 6
        self.CSX = None

    # ---- settings (openEMS.pyx:123-309)
    def SetNumberOfTimeSteps(self, val):
        self.nrts = int(val)

    def SetEndCriteria(self, val):
        self.end_criteria = float(val)

    def SetOverSampling(self, val):
        self.oversampling = int(val)

    def SetCellConstantMaterial(self, val):
        self.cell_constant = bool(val)

    def SetTimeStepMethod(self, val):
        self.ts_method = int(val)

    def SetTimeStep(self, val):
        self.time_step = float(val)

    def SetTimeStepFactor(self, val):
        self.ts_factor = float(val)

    def SetGaussExcite(self, f0, fc):
        self.signal = {"type": "gauss", "f0": float(f0), "fc": float(fc)}

    def SetSinusExcite(self, f0):
        self.signal = {"type": "sinus", "f0": float(f0)}

    def SetDiracExcite(self, f_max):
        self.signal = {"type": "dirac", "f_max": float(f_max)}

    def SetStepExcite(self, f_max):
        self.signal = {"type": "step", "f_max": float(f_max)}

    def SetCustomExcite(self, _str, f0, fmax):
        self.signal = {"type": "custom", "function": str(_str), "f0": float(f0), "f_max": float(fmax)}

    def SetBoundaryCond(self, BC):
        if len(BC) != 6:
            raise ValueError("Invalid boundary condition size!")
        for n, b in enumerate(BC):
            if isinstance(b, (int, np.integer)):
                self.bc[n] = int(b)
            elif b in ("PEC", "PMC", "MUR"):
                self.bc[n] = ("PEC", "PMC", "MUR").index(b)
            elif isinstance(b, str) and b.startswith("PML_"):
                self.bc[n] = 3
                self.pml[n] = int(b[4:])
            else:
                raise ValueError("Unknown boundary condition")

    def SetCSX(self, CSX):
        self.CSX = CSX

    def GetCSX(self):
        return self.CSX

    def AddRectWaveGuidePort(self, port_nr, start, stop, p_dir, a, b, mode_name, excite=0, **kw):
        if self.CSX is None:
            raise ValueError("AddRectWaveGuidePort: CSX is not set!")
        return RectWGPort(self.CSX, port_nr, start, stop, p_dir, a, b, mode_name, excite, **kw)

    def AddMSLPort(self, port_nr, metal_prop, start, stop, prop_dir, exc_dir, excite=0, **kw):
        if self.CSX is None:
            raise ValueError("AddMSLPort: CSX is not set!")
        port = MSLPort(self.CSX, port_nr, metal_prop, start, stop, prop_dir, exc_dir, excite, **kw)
        e2g = kw.get("edges2grid")
        if e2g is not None:
            grid = self.CSX.GetGrid()
            for n in ([0, 1, 2] if e2g == "all" else [_dir(c) for c in e2g]):
                grid.AddLine(n, start[n])
                if start[n] != stop[n]:
                    grid.AddLine(n, stop[n])
        return port

    def AddLumpedPort(self, port_nr, R, start, stop, p_dir, excite=0, **kw):
        if self.CSX is None:
            raise ValueError("AddLumpedPort: CSX is not set!")
        port = LumpedPort(self.CSX, port_nr, R, start, stop, p_dir, excite, **kw)
        e2g = kw.get("edges2grid")
        if e2g is not None:
            grid = self.CSX.GetGrid()
            for n in ([0, 1, 2] if e2g == "all" else [_dir(c) for c in e2g]):
                grid.AddLine(n, start[n])
                if start[n] != stop[n]:
                    grid.AddLine(n, stop[n])
        return port

    # ---- scene, XML and run
    def scene(self):
        """The simulation as a scene dict (openems_amd.scenes layout)."""
        if self.CSX is None:
            raise ValueError("CSX is not set")
        if self.signal is None:
            raise ValueError("no excitation signal set")
        g = self.CSX.GetGrid()
        sc = {"name": "pyopenems", "lines": [l.copy() for l in g.lines], "delta_unit": g.delta_unit,
              "background": dict(self.CSX.background), "bc": list(self.bc), "pml_size": list(self.pml),
              "signal": dict(self.signal), "nrts": self.nrts, "end_criteria": self.end_criteria}
        if self.oversampling and self.oversampling > 1:
            sc["oversampling"] = int(self.oversampling)
        if self.ts_method:
            sc["ts_method"] = int(self.ts_method)
        if self.time_step:
            sc["forced_dt"] = self.time_step
        if self.ts_factor:
            sc["ts_factor"] = float(self.ts_factor)
        if self.cell_constant:
            sc["cell_constant_material"] = True
        sc.update(self.CSX.scene_parts())
        return sc

    def Write2XML(self, file):
        from .xml_setup import write_xml
        write_xml(self.scene(), file)

    def Run(self, sim_path, cleanup=False, setup_only=False, **kw):
        """Build, run up to NrTS steps or the end criterion, and write the
        probe files and HDF5 dumps into sim_path.  Returns 0, or the
        Simulation when setup_only is set (not run)."""
        from .simulation import Simulation
        if cleanup and os.path.exists(sim_path):
            shutil.rmtree(sim_path, ignore_errors=True)
        os.makedirs(sim_path, exist_ok=True)
        sim = Simulation(self.scene(), **{k: v for k, v in kw.items() if k in ("device", "layout", "fused", "graph")})
        if setup_only:
            return sim
        try:
            sim.run()
            self.results = sim.results()
            sim.write_probes(sim_path)
            sim.write_dumps(sim_path)
        finally:
            sim.close()
        return 0
