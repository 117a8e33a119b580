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

    def AddProbe(self, name, p_type, weight=1.0, norm_dir=-1, frequency=None, **kw):
        """Probe types 0-3 (voltage, current, E and H point); the mode-matching
        types 10 / 11 are outside the engine's scope and refused at Run."""
        return self._add("probe", name, p_type=int(p_type), weight=float(weight), norm_dir=int(norm_dir),
                         frequency=list(np.atleast_1d(frequency)) if frequency is not None else [])

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
# SI constants of the reference's python/openEMS/physical_constants.py
C0 = 299792458.0
MUE0 = 4e-7 * np.pi
EPS0 = 1.0 / (MUE0 * C0 ** 2)
Z0 = np.sqrt(MUE0 / EPS0)


class UI_data:
    """The probe files of a port read back: per file the sample times, the
    values and their single-sided spectrum at `freq` (DFT_time2freq).  The
    files are the probe text files of the simulation ('%' comment lines, time
    in the first column, value in the second)."""

    def __init__(self, fns, path, freq, signal_type="pulse"):
        self.fns = [fns] if isinstance(fns, str) else list(fns)
        self.freq = np.atleast_1d(freq)
        tables = [np.loadtxt(os.path.join(path, fn), comments="%") for fn in self.fns]
        self.ui_time = [tab[:, 0] for tab in tables]
        self.ui_val = [tab[:, 1] for tab in tables]
        self.ui_f_val = [DFT_time2freq(tab[:, 0], tab[:, 1], self.freq, signal_type) for tab in tables]


def line_waves(U, I, Z):
    """Forward and backward waves at one plane of a line of characteristic
    impedance Z, from the total voltage U and current I there.

    On the line U = U_f + U_b and Z I = U_f - U_b (the backward wave's current
    flows against the reference direction).  Adding and subtracting,
        U_f = (U + Z I) / 2,   U_b = (U - Z I) / 2,
    and each wave carries the current of its own voltage over Z:
        I_f = U_f / Z,   I_b = U_b / Z   (so I = I_f - I_b).
    Works on spectra (complex Z) and on time series (real Z)."""
    zi = Z * I
    u_f = 0.5 * (U + zi)
    u_b = 0.5 * (U - zi)
    return u_f, u_b, u_f / Z, u_b / Z


def move_plane(U, I, Z, theta):
    """Total voltage and current a distance d further along a lossless line,
    theta = beta * d: the forward wave is delayed by exp(-j theta), the
    backward wave advanced by exp(+j theta); recombined,
        U' = U cos(theta) - j Z I sin(theta),  I' = I cos(theta) - j (U / Z) sin(theta)."""
    u_f, u_b, _, _ = line_waves(U, I, Z)
    fwd = np.exp(-1j * theta)
    u_f, u_b = u_f * fwd, u_b / fwd
    return u_f + u_b, (u_f - u_b) / Z


def line_from_three_planes(U_a, U_b, U_c, span_u, I_1, I_2, span_i):
    """Propagation constant beta and characteristic impedance of a uniform
    line from the spectra of three voltage probes (planes a, b, c in the
    propagation direction, a to c `span_u` metres) and two current probes
    (between a-b and b-c, `span_i` metres apart).

    The telegrapher's equations at plane b,
        dU/dz = -Z' I,   dI/dz = -Y' U,
    with the central differences dU/dz = (U_c - U_a) / span_u,
    dI/dz = (I_2 - I_1) / span_i and the plane's values U = U_b,
    I = (I_1 + I_2) / 2, give the per-length impedance Z' and admittance Y'.
    A forward wave exp(-gamma z) has gamma^2 = Z' Y' and impedance
    sqrt(Z' / Y'); with gamma = j beta on a line of small loss:
        beta = sqrt(-gamma^2) = sqrt(-(dU/dz)(dI/dz) / (U I)),
        Z_L  = sqrt((dU/dz) U / ((dI/dz) I)),
    beta on the branch with a non-negative real part (the forward wave)."""
    I_b = 0.5 * (I_1 + I_2)
    dU = (U_c - U_a) / span_u
    dI = (I_2 - I_1) / span_i
    beta = np.sqrt(-(dU * dI) / (U_b * I_b))
    beta = np.where(np.real(beta) < 0, -beta, beta)
    return beta, np.sqrt((dU * U_b) / (dI * I_b)), U_b, I_b


def _half_power(u, i):
    """Time-averaged power of phasors u, i (peak amplitudes): Re(u i*) / 2."""
    return 0.5 * np.real(u * np.conj(i))


class _Port:
    """What the reference's ports share: the excitation switch and the
    post-processing of the port's probe files into total, incident and
    reflected voltages, currents and powers (python/openEMS/ports.py:48-153
    is the interface; the splitting follows line_waves / move_plane)."""

    Z_ref = None
    measplane_shift = 0.0

    def _setup_common(self, CSX, port_nr, start, stop, excite, kw):
        self.CSX = CSX
        self.number = port_nr
        self.start = np.array(start, np.float64)
        self.stop = np.array(stop, np.float64)
        self.excite = excite
        self.priority = kw.get("priority", 0)
        self.delay = kw.get("delay", 0.0)
        self.port_props = []
        self._label = kw.get("PortNamePrefix", "") + "port_{}_" + str(port_nr)

    def _name(self, what):
        return self._label.format(what)

    def _add_excitation(self, vec, start, stop):
        exc = self.CSX.AddExcitation(self._name("excite"), exc_type=0, exc_val=vec, delay=self.delay)
        exc.AddBox(start, stop, priority=self.priority)
        self.port_props.append(exc)
        return exc

    def SetEnabled(self, val):
        """Switches the port's excitation on or off; a port built without an
        excitation cannot be switched on (ValueError)."""
        excitations = [p for p in self.port_props if p.kind == "excitation"]
        if val and not excitations:
            raise ValueError(f"port {self.number} has no excitation to enable")
        for p in excitations:
            p.SetEnabled(val)

    def ReadUIData(self, sim_path, freq, signal_type="pulse"):
        """Total voltage and current of the port: the sums over its probe files."""
        self.u_data = UI_data(self.U_filenames, sim_path, freq, signal_type)
        self.i_data = UI_data(self.I_filenames, sim_path, freq, signal_type)
        self.ut_tot, self.uf_tot = sum(self.u_data.ui_val), sum(self.u_data.ui_f_val)
        self.it_tot, self.if_tot = sum(self.i_data.ui_val), sum(self.i_data.ui_f_val)

    def _port_impedance(self):
        return self.Z_ref

    def CalcPort(self, sim_path, freq, ref_impedance=None, ref_plane_shift=None, signal_type="pulse"):
        """Reads the probe files and splits the port's voltage and current
        into incident (forward) and reflected (backward) waves against
        ref_impedance, or the port's own impedance (R for a lumped port, the
        measured line impedance of a microstrip port).  ref_plane_shift (in
        drawing units, from the port start) moves the reference plane along
        a line port before the split.  Sets uf_/if_/ut_/it_ tot, inc and ref
        and the powers P_inc, P_ref, P_acc."""
        self.ReadUIData(sim_path, freq, signal_type)
        if ref_impedance is not None:
            self.Z_ref = ref_impedance
        else:
            self.Z_ref = self._port_impedance()
        if self.Z_ref is None:
            raise ValueError(f"port {self.number}: no reference impedance")
        Z = self.Z_ref
        if ref_plane_shift is not None:
            if getattr(self, "beta", None) is None:
                raise ValueError(f"port {self.number}: no propagation constant to move the reference plane")
            d = (ref_plane_shift - self.measplane_shift) * self.CSX.GetGrid().GetDeltaUnit()
            self.uf_tot, self.if_tot = move_plane(self.uf_tot, self.if_tot, Z, np.real(self.beta) * d)
        self.uf_inc, self.uf_ref, self.if_inc, self.if_ref = line_waves(self.uf_tot, self.if_tot, Z)
        if np.isrealobj(Z) and np.ndim(Z) == 0:
            self.ut_inc, self.ut_ref, self.it_inc, self.it_ref = line_waves(self.ut_tot, self.it_tot, Z)
        self.P_inc = _half_power(self.uf_inc, self.if_inc)
        self.P_ref = _half_power(self.uf_ref, self.if_ref)
        self.P_acc = _half_power(self.uf_tot, self.if_tot)


class LumpedPort(_Port):
    """A lumped port (the interface of ports.LumpedPort): a resistor R with
    end caps over the port box (a metal short for R = 0), optionally a soft
    E excitation of -direction * excite along the port axis, a voltage probe
    along the axis through the box centre (weight -1) and a current probe
    over the box's mid plane (weight = direction, normal = the axis).  Its
    reference impedance is R."""

    def __init__(self, CSX, port_nr, R, start, stop, exc_dir, excite=0, **kw):
        self._setup_common(CSX, port_nr, start, stop, excite, kw)
        self.R = R
        ax = _dir(exc_dir)
        self.exc_ny = ax
        length = self.stop[ax] - self.start[ax]
        if length == 0:
            raise ValueError(f"lumped port {port_nr}: the box has no extent along the port axis")
        self.direction = np.sign(length)
        if R > 0:
            CSX.AddLumpedElement(self._name("resist"), ny=ax, caps=True, R=R).AddBox(
                self.start, self.stop, priority=self.priority)
        elif R == 0:
            CSX.AddMetal(self._name("resist")).AddBox(self.start, self.stop, priority=self.priority)
        if excite != 0:
            vec = np.zeros(3)
            vec[ax] = -self.direction * excite
            self._add_excitation(vec, self.start, self.stop)
        centre = 0.5 * (self.start + self.stop)
        u0, u1 = centre.copy(), centre.copy()
        u0[ax], u1[ax] = self.start[ax], self.stop[ax]
        self.U_filenames = [self._name("ut")]
        CSX.AddProbe(self.U_filenames[0], p_type=0, weight=-1).AddBox(u0, u1)
        i0, i1 = self.start.copy(), self.stop.copy()
        i0[ax] = i1[ax] = centre[ax]
        self.I_filenames = [self._name("it")]
        CSX.AddProbe(self.I_filenames[0], p_type=1, weight=self.direction, norm_dir=ax).AddBox(i0, i1)

    def _port_impedance(self):
        return self.R


class MSLPort(_Port):
    """A microstrip port (the interface of ports.MSLPort): the strip as a
    metal box in the plane start[exc] of the port box; three voltage probes
    across the substrate on the mesh lines around the measurement plane
    (half the port length from the start unless MeasPlaneShift is given) and
    two current probes around the strip half way between them, ordered in
    the propagation direction; an optional soft feed at FeedShift and a
    feed resistance Feed_R at the port start.  ReadUIData takes the middle
    voltage and the mean current as the port's totals and derives beta and
    the line impedance from the telegrapher's equations
    (line_from_three_planes); the line impedance is the reference impedance."""

    def __init__(self, CSX, port_nr, metal_prop, start, stop, prop_dir, exc_dir, excite=0, **kw):
        self._setup_common(CSX, port_nr, start, stop, excite, kw)
        ex, pr = _dir(exc_dir), _dir(prop_dir)
        self.exc_ny, self.prop_ny = ex, pr
        if np.any(self.start == self.stop):
            raise ValueError(f"microstrip port {port_nr}: the box must have an extent along every axis")
        if ex == pr:
            raise ValueError(f"microstrip port {port_nr}: excitation and propagation directions coincide")
        self.direction = np.sign(self.stop[pr] - self.start[pr])
        self.upside_down = np.sign(self.stop[ex] - self.start[ex])
        self.feed_shift = kw.get("FeedShift", 0)
        shift = kw.get("MeasPlaneShift", 0.5 * abs(self.stop[pr] - self.start[pr]))
        self.measplane_pos = self.start[pr] + shift * self.direction
        self.feed_R = kw.get("Feed_R", np.inf)
        strip_end = self.stop.copy()
        strip_end[ex] = self.start[ex]
        metal_prop.AddBox(self.start, strip_end, priority=self.priority)
        lines = CSX.GetGrid().GetLines(pr)
        if len(lines) <= 5:
            raise ValueError(f"microstrip port {port_nr}: more than 5 mesh lines along the line are needed")
        # the mesh line nearest the measurement plane and its two neighbours,
        # in the propagation direction
        k = int(np.clip(np.argmin(np.abs(lines - self.measplane_pos)), 1, len(lines) - 2))
        self.measplane_shift = abs(self.start[pr] - lines[k])
        u_at = lines[[k - 1, k, k + 1]][:: int(self.direction) or 1]
        self.U_delta = np.diff(u_at)
        i_at = 0.5 * (u_at[:-1] + u_at[1:])
        self.I_delta = np.diff(i_at)
        centre = 0.5 * (self.start + self.stop)
        self.U_filenames = [self._name("ut") + tag for tag in "ABC"]
        for fn, pos in zip(self.U_filenames, u_at):
            u0, u1 = centre.copy(), centre.copy()
            u0[pr] = u1[pr] = pos
            u0[ex], u1[ex] = self.start[ex], self.stop[ex]
            CSX.AddProbe(fn, p_type=0).AddBox(u0, u1)
        self.I_filenames = [self._name("it") + tag for tag in "AB"]
        for fn, pos in zip(self.I_filenames, i_at):
            i0, i1 = self.start.copy(), self.stop.copy()
            i1[ex] = self.start[ex]
            i0[pr] = i1[pr] = pos
            CSX.AddProbe(fn, p_type=1, weight=self.direction, norm_dir=pr).AddBox(i0, i1)
        if excite != 0:
            kf = int(np.argmin(np.abs(lines - (self.start[pr] + self.feed_shift * self.direction))))
            f0, f1 = self.start.copy(), self.stop.copy()
            f0[pr] = f1[pr] = lines[kf]
            vec = np.zeros(3)
            vec[ex] = -self.upside_down * excite
            self._add_excitation(vec, f0, f1)
        if self.feed_R >= 0 and np.isfinite(self.feed_R):
            r0, r1 = self.start.copy(), self.stop.copy()
            r1[pr] = r0[pr]
            if self.feed_R == 0:
                metal_prop.AddBox(r0, r1)
            else:
                CSX.AddLumpedElement(self._name("resist"), ny=ex, caps=True, R=self.feed_R).AddBox(r0, r1)

    def ReadUIData(self, sim_path, freq, signal_type="pulse"):
        self.u_data = UI_data(self.U_filenames, sim_path, freq, signal_type)
        self.i_data = UI_data(self.I_filenames, sim_path, freq, signal_type)
        unit = self.CSX.GetGrid().GetDeltaUnit()
        Uf, If = self.u_data.ui_f_val, self.i_data.ui_f_val
        self.beta, self.Z_line, self.uf_tot, self.if_tot = line_from_three_planes(
            Uf[0], Uf[1], Uf[2], np.sum(np.abs(self.U_delta)) * unit, If[0], If[1], abs(self.I_delta[0]) * unit)
        self.ut_tot = self.u_data.ui_val[1]
        self.it_tot = 0.5 * (self.i_data.ui_val[0] + self.i_data.ui_val[1])

    def _port_impedance(self):
        return self.Z_line


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
        self._sim = sim
        try:
            sim.SetAbort(getattr(self, "_abort", False))
            sim.run()
            self.results = sim.results()
            sim.write_probes(sim_path)
            sim.write_dumps(sim_path)
        finally:
            self._sim = None
            sim.close()
        return 0

    def SetAbort(self, val):
        """openEMS.SetAbort (python/openEMS/openEMS.pyx:566-567): a run in
        progress (from another thread) or the next run stops gracefully."""
        self._abort = bool(val)
        sim = getattr(self, "_sim", None)
        if sim is not None:
            sim.SetAbort(val)
