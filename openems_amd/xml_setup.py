"""openEMS XML setups (the file ``openEMS <file.xml>`` reads) for the HIP engine.

``read_xml`` turns an openEMS XML file into a scene dict that
:class:`~openems_amd.simulation.Simulation` runs; ``write_xml`` writes a scene
dict back in the same layout; ``run_xml`` is the whole ``openEMS file.xml``
run: read, build, iterate, and leave the probe files in a folder.

The ``<FDTD>`` block follows openEMS::Parse_XML_FDTDSetup
(openems.cpp:837-975) attribute by attribute: NumberOfTimesteps, endCriteria
(0 or absent means 1e-6), OverSampling (used when > 1), CellConstantMaterial,
TimeStepMethod, TimeStep, TimeStepFactor, f_max, ``<BoundaryCond>`` (integer
codes or PEC / PMC / MUR / PML_n, Mur phase velocities) and ``<Excitation>``
(Type 0 Gaussian f0/fc, 1 sinusoid f0, 2 Dirac and 3 step with f_max, 10
custom Function of t with f0).

The ``<ContinuousStructure>`` block is CSXCAD's format, read here for the
subset the engine takes without CSXCAD: ``RectilinearGrid`` (DeltaUnit,
X/Y/ZLines), ``BackgroundMaterial``, and the properties Material, Metal,
Excitation, ProbeBox, DumpBox (time- and frequency-domain E / H dumps),
LumpedElement, LorentzMaterial and
DebyeMaterial with Box primitives (P1, P2, Priority) and, for Metal and
Excitation, Curve (Vertex points), Cylinder (P1, P2, Radius) and
CylindricalShell (P1, P2, Radius, ShellWidth) primitives; an Excitation's
Weight element (X, Y, Z expressions); ProbeBox types 0-3 (mode matching,
types 10 / 11, is outside the engine's scope and refused).  Properties are
entered in file
order, primitives in property order, as CSXCAD hands them to openEMS
(openems.cpp:483-694 for probes and dumps).  Anything else (other primitive
types, cylindrical meshes) is refused with a
ValueError naming it.
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as ET

import numpy as np

_BOUNDS = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
_SIGNAL_TYPES = {0: "gauss", 1: "sinus", 2: "dirac", 3: "step", 10: "custom"}
_SIGNAL_CODES = {v: k for k, v in _SIGNAL_TYPES.items()}


def _floats(text):
    text = (text or "").replace(";", ",").strip()
    return [float(v) for v in text.split(",") if v.strip()] if text else []


def _vec(el, name, default):
    """An attribute holding one value (isotropic) or x,y,z."""
    v = el.get(name) if el is not None else None
    if v is None:
        return default
    f = _floats(v)
    if len(f) == 1:
        return f[0]
    if len(f) == 3:
        return tuple(f)
    raise ValueError(f"attribute {name}={v!r}: expected 1 or 3 values")


def _int(el, name, default):
    v = el.get(name)
    if v is None:
        return default
    return int(float(v))


def _point(el):
    return tuple(float(el.get(k, 0.0)) for k in ("X", "Y", "Z"))


def _primitives(prop):
    """(kind, priority, data) for each primitive of a property, in order."""
    prims = prop.find("Primitives")
    out = []
    for p in (list(prims) if prims is not None else []):
        prio = _int(p, "Priority", 0)
        if p.tag == "Box":
            p1, p2 = p.find("P1"), p.find("P2")
            if p1 is None or p2 is None:
                raise ValueError(f"Box of {prop.get('Name')!r} needs P1 and P2")
            out.append(("box", prio, (_point(p1), _point(p2))))
        elif p.tag == "Curve":
            pts = [_point(v) for v in p.findall("Vertex")]
            out.append(("curve", prio, pts))
        elif p.tag in ("Cylinder", "CylindricalShell"):
            p1, p2 = p.find("P1"), p.find("P2")
            if p1 is None or p2 is None or p.get("Radius") is None:
                raise ValueError(f"{p.tag} of {prop.get('Name')!r} needs P1, P2 and Radius")
            width = float(p.get("ShellWidth", 0.0)) if p.tag == "CylindricalShell" else 0.0
            out.append(("cylinder", prio, (_point(p1), _point(p2), float(p.get("Radius")), width)))
        else:
            raise ValueError(f"primitive {p.tag!r} of {prop.get('Name')!r} is not supported")
    return out


def _fd_samples(prop):
    el = prop.find("FD_Samples")
    if el is not None:
        return _floats(el.text)
    return _floats(prop.get("FD_Samples", ""))


def _bc(bc_el):
    """openEMS::Parse_XML_FDTDSetup boundary block (openems.cpp:877-929)."""
    types, sizes, vph = [0] * 6, [8] * 6, [0.0] * 6
    for n, name in enumerate(_BOUNDS):
        v = bc_el.get(name) if bc_el is not None else None
        if v is None:
            continue
        try:
            types[n] = int(v)
            continue
        except ValueError:
            pass
        if v == "PEC":
            types[n] = 0
        elif v == "PMC":
            types[n] = 1
        elif v == "MUR":
            types[n] = 2
        elif v.startswith("PML_"):
            types[n] = 3
            sizes[n] = int(v[4:] or 0)
        else:
            types[n] = 0      # unknown: the reference warns and keeps PEC
    if bc_el is not None and bc_el.get("MUR_PhaseVelocity") is not None:
        vph = [float(bc_el.get("MUR_PhaseVelocity"))] * 6
    for n, name in enumerate(_BOUNDS):
        v = bc_el.get("MUR_PhaseVelocity_" + name) if bc_el is not None else None
        if v is not None:
            vph[n] = float(v)
    return types, sizes, vph


_LORENTZ_ATTRS = (("eps_plasma", "EpsilonPlasmaFrequency"), ("eps_relax", "EpsilonRelaxTime"),
                  ("eps_pole", "EpsilonLorPoleFrequency"), ("mue_plasma", "MuePlasmaFrequency"),
                  ("mue_relax", "MueRelaxTime"), ("mue_pole", "MueLorPoleFrequency"))
_DEBYE_ATTRS = (("eps_delta", "EpsilonDelta"), ("eps_relax", "EpsilonRelaxTime"))


def _orders(prop, el, attrs):
    """Per-order terms of a dispersive property (attribute Order, default 1):
    attribute names carry the order suffix _1, _2, ... (the first order also
    without a suffix); a missing term is 0 (unset)."""
    out = []
    for o in range(_int(prop, "Order", 1)):
        sufs = ["_%d" % (o + 1)] + ([""] if o == 0 else [])
        d = {}
        for key, attr in attrs:
            d[key] = 0.0
            for sfx in sufs:
                v = _vec(el, attr + sfx, None)
                if v is not None:
                    d[key] = v
                    break
        out.append(d)
    return out


def read_xml(source):
    """Scene dict of an openEMS XML setup (a path or the XML text)."""
    if isinstance(source, str) and source.lstrip().startswith("<"):
        root = ET.fromstring(source)
    else:
        root = ET.parse(source).getroot()
    if root.tag != "openEMS":
        raise ValueError("not an openEMS setup: the root element must be <openEMS>")
    fdtd = root.find("FDTD")
    if fdtd is None:
        raise ValueError("no <FDTD> block")
    csx = root.find("ContinuousStructure")
    if csx is None:
        raise ValueError("no <ContinuousStructure> block")
    scene = {"name": "xml"}
    # ---- FDTD block (openems.cpp:837-975)
    nts = float(fdtd.get("NumberOfTimesteps", 0))
    scene["nrts"] = 0 if nts < 0 else int(nts)
    if _int(fdtd, "CylinderCoords", 0) == 1:
        raise ValueError("cylindrical coordinates are not supported")
    end = float(fdtd.get("endCriteria", 0))
    scene["end_criteria"] = 1e-6 if end == 0 else end
    os_ = _int(fdtd, "OverSampling", 0)
    scene["oversampling"] = os_ if os_ > 1 else 4
    if fdtd.get("CellConstantMaterial") is not None:
        scene["cell_constant_material"] = _int(fdtd, "CellConstantMaterial", 0) == 1
    types, sizes, vph = _bc(fdtd.find("BoundaryCond"))
    scene["bc"], scene["pml_size"], scene["mur_vph"] = types, sizes, vph
    exc = fdtd.find("Excitation")
    if exc is None:
        raise ValueError("no <Excitation> in the FDTD block")
    etype = _int(exc, "Type", -1)
    if etype not in _SIGNAL_TYPES:
        raise ValueError(f"excitation type {etype} is not supported")
    f_max = float(fdtd.get("f_max", 0.0))
    sig = {"type": _SIGNAL_TYPES[etype], "f0": float(exc.get("f0", 0.0)), "fc": float(exc.get("fc", 0.0))}
    if etype in (2, 3, 10):
        sig["f_max"] = f_max
    if etype == 10:
        sig["function"] = exc.get("Function", "")
    scene["signal"] = sig
    if fdtd.get("TimeStepMethod") is not None:
        scene["ts_method"] = _int(fdtd, "TimeStepMethod", 3)
    if fdtd.get("TimeStep") is not None:
        scene["forced_dt"] = float(fdtd.get("TimeStep"))
    if fdtd.get("TimeStepFactor") is not None:
        scene["ts_factor"] = float(fdtd.get("TimeStepFactor"))
    # ---- structure
    grid = csx.find("RectilinearGrid")
    if grid is None:
        raise ValueError("no <RectilinearGrid>")
    if _int(grid, "CoordSystem", 0) != 0 or _int(csx, "CoordSystem", 0) != 0:
        raise ValueError("only Cartesian meshes are supported")
    lines = []
    for tag in ("XLines", "YLines", "ZLines"):
        el = grid.find(tag)
        vals = np.unique(np.asarray(_floats(el.text if el is not None else ""), np.float64))
        if vals.size < 3:
            raise ValueError(f"{tag}: at least 3 mesh lines are required")
        lines.append(vals)
    scene["lines"] = lines
    scene["delta_unit"] = float(grid.get("DeltaUnit", 1.0))
    bg = csx.find("BackgroundMaterial")
    if bg is not None:
        scene["background"] = {"eps": float(bg.get("Epsilon", 1.0)), "mue": float(bg.get("Mue", 1.0)),
                               "kappa": float(bg.get("Kappa", 0.0)), "sigma": float(bg.get("Sigma", 0.0))}
    mats, metals, lumped, excs, probes, dumps, skipped = [], [], [], [], [], [], []
    props = csx.find("Properties")
    for prop in (list(props) if props is not None else []):
        name = prop.get("Name", "")
        prims = _primitives(prop)
        boxes = [(pr, d) for k, pr, d in prims if k == "box"]
        curves = [d for k, pr, d in prims if k == "curve"]
        if prop.tag not in ("Metal", "Excitation"):
            for k, pr, d in prims:
                if k != "box":
                    raise ValueError(f"{k} primitives of {prop.tag} {name!r} are not supported")
        if prop.tag in ("Material", "LorentzMaterial", "DebyeMaterial"):
            el = prop.find("Property")
            base = {"eps": _vec(el, "Epsilon", 1.0), "mue": _vec(el, "Mue", 1.0),
                    "kappa": _vec(el, "Kappa", 0.0), "sigma": _vec(el, "Sigma", 0.0)}
            if prop.tag == "LorentzMaterial":
                base["lorentz"] = _orders(prop, el, _LORENTZ_ATTRS)
            elif prop.tag == "DebyeMaterial":
                base["debye"] = _orders(prop, el, _DEBYE_ATTRS)
            for prio, (p1, p2) in boxes:
                mats.append(dict(base, priority=prio, start=p1, stop=p2))
        elif prop.tag == "Metal":
            for kind, prio, d in prims:
                m = {"curve": d} if kind == "curve" else {"priority": prio, "start": d[0], "stop": d[1]}
                if kind == "cylinder":
                    m.update(radius=d[2], shell_width=d[3])
                metals.append(m)
        elif prop.tag == "Excitation":
            e = {"type": _int(prop, "Type", 0), "val": _vec(prop, "Excite", 0.0),
                 "delay": float(prop.get("Delay", 0.0))}
            if isinstance(e["val"], float):
                e["val"] = (e["val"],) * 3
            w = prop.find("Weight")
            if w is not None:
                e["weight"] = [w.get(k) for k in ("X", "Y", "Z")]
            for kind, prio, d in prims:
                x = dict(e, curve=d) if kind == "curve" else dict(e, priority=prio, start=d[0], stop=d[1])
                if kind == "cylinder":
                    x.update(radius=d[2], shell_width=d[3])
                excs.append(x)
        elif prop.tag == "LumpedElement":
            le = {"dir": _int(prop, "Direction", -1), "R": float(prop.get("R", -1.0) or -1.0),
                  "C": float(prop.get("C", 0.0) or 0.0), "L": float(prop.get("L", 0.0) or 0.0),
                  "caps": _int(prop, "Caps", 1) == 1,
                  "type": "series" if _int(prop, "LEtype", 0) == 1 else "parallel"}
            for prio, (p1, p2) in boxes:
                lumped.append(dict(le, priority=prio, start=p1, stop=p2))
        elif prop.tag == "ProbeBox":
            ptype = _int(prop, "Type", 0)
            if ptype not in (0, 1, 2, 3):
                raise ValueError(f"probe type {ptype} of {name!r} is not supported")
            for nb, (prio, (p1, p2)) in enumerate(boxes):
                probes.append({"type": ptype, "name": name if len(boxes) == 1 else f"{name}_{nb}",
                               "start": list(p1), "stop": list(p2), "weight": float(prop.get("Weight", 1.0)),
                               "freqs": _fd_samples(prop), "norm_dir": _int(prop, "NormDir", -1),
                               "start_time": float(prop.get("StartTime", 0.0)),
                               "stop_time": float(prop.get("StopTime", 0.0))})
        elif prop.tag == "DumpBox":
            dtype = _int(prop, "DumpType", 0)
            mode = _int(prop, "DumpMode", 0)
            if dtype not in (0, 1, 10, 11):
                skipped.append((name, f"dump type {dtype}"))
                continue
            for nb, (prio, (p1, p2)) in enumerate(boxes):
                d = {"name": name if len(boxes) == 1 else f"{name}_{nb}", "start": list(p1), "stop": list(p2),
                     "field": "E" if dtype in (0, 10) else "H", "interp": mode,
                     "file_type": _int(prop, "FileType", 0)}
                d.update(start_time=float(prop.get("StartTime", 0.0)), stop_time=float(prop.get("StopTime", 0.0)))
                if dtype < 10:
                    d["domain"] = "td"
                else:
                    d["freqs"] = _fd_samples(prop)
                dumps.append(d)
        else:
            raise ValueError(f"property {prop.tag} {name!r} is not supported")
    scene.update(materials=mats, metals=metals, lumped=lumped, excitations=excs, probes=probes, dumps=dumps)
    scene["skipped"] = skipped
    return scene


# ----------------------------------------------------------------------------
# writer
# ----------------------------------------------------------------------------
def _fmt(v):
    if isinstance(v, (tuple, list, np.ndarray)):
        return ",".join(repr(float(x)) for x in v)
    return repr(float(v))


def _box(parent, prio, p1, p2):
    b = ET.SubElement(parent, "Box", Priority=str(int(prio)))
    ET.SubElement(b, "P1", X=repr(float(p1[0])), Y=repr(float(p1[1])), Z=repr(float(p1[2])))
    ET.SubElement(b, "P2", X=repr(float(p2[0])), Y=repr(float(p2[1])), Z=repr(float(p2[2])))


def _prim(parent, d):
    """A box, or a cylinder / cylindrical shell when d has a radius."""
    if "radius" not in d:
        _box(parent, d.get("priority", 0), d["start"], d["stop"])
        return
    w = float(d.get("shell_width", 0.0))
    attrs = {"Priority": str(int(d.get("priority", 0))), "Radius": repr(float(d["radius"]))}
    if w > 0:
        attrs["ShellWidth"] = repr(w)
    c = ET.SubElement(parent, "CylindricalShell" if w > 0 else "Cylinder", **attrs)
    for tag, p in (("P1", d["start"]), ("P2", d["stop"])):
        ET.SubElement(c, tag, X=repr(float(p[0])), Y=repr(float(p[1])), Z=repr(float(p[2])))


def _curve(parent, pts):
    c = ET.SubElement(parent, "Curve", Priority="0")
    for p in pts:
        ET.SubElement(c, "Vertex", X=repr(float(p[0])), Y=repr(float(p[1])), Z=repr(float(p[2])))


def write_xml(scene, path=None):
    """The scene as an openEMS XML setup (returned as text; also written to
    path when given).  Material grids are not representable and are refused."""
    root = ET.Element("openEMS")
    sig = scene["signal"]
    fd = ET.SubElement(root, "FDTD", NumberOfTimesteps=str(int(scene.get("nrts", 1000))),
                       endCriteria=repr(float(scene.get("end_criteria", 1e-6)) or 1e-6),
                       OverSampling=str(int(scene.get("oversampling", 4))))
    if sig["type"] in ("dirac", "step", "custom"):
        fd.set("f_max", repr(float(sig.get("f_max", 0.0))))
    if "ts_method" in scene:
        fd.set("TimeStepMethod", str(int(scene["ts_method"])))
    if scene.get("forced_dt", 0) > 0:
        fd.set("TimeStep", repr(float(scene["forced_dt"])))
    if "ts_factor" in scene:
        fd.set("TimeStepFactor", repr(float(scene["ts_factor"])))
    if scene.get("cell_constant_material"):
        fd.set("CellConstantMaterial", "1")
    ex = ET.SubElement(fd, "Excitation", Type=str(_SIGNAL_CODES[sig["type"]]), f0=repr(float(sig.get("f0", 0.0))),
                       fc=repr(float(sig.get("fc", 0.0))))
    if sig["type"] == "custom":
        ex.set("Function", sig["function"])
    bc = ET.SubElement(fd, "BoundaryCond")
    pml = scene.get("pml_size", [8] * 6)
    for n, name in enumerate(_BOUNDS):
        t = int(scene.get("bc", [0] * 6)[n])
        bc.set(name, f"PML_{int(pml[n])}" if t == 3 else str(t))
    for n, name in enumerate(_BOUNDS):
        v = scene.get("mur_vph", [0.0] * 6)[n]
        if v:
            bc.set("MUR_PhaseVelocity_" + name, repr(float(v)))
    csx = ET.SubElement(root, "ContinuousStructure", CoordSystem="0")
    bg = scene.get("background", {})
    ET.SubElement(csx, "BackgroundMaterial", Epsilon=_fmt(bg.get("eps", 1.0)), Mue=_fmt(bg.get("mue", 1.0)),
                  Kappa=_fmt(bg.get("kappa", 0.0)), Sigma=_fmt(bg.get("sigma", 0.0)))
    props = ET.SubElement(csx, "Properties")
    pid = 0

    def prop(tag, name, **attrs):
        nonlocal pid
        el = ET.SubElement(props, tag, ID=str(pid), Name=name, **attrs)
        pid += 1
        return el
    for k, m in enumerate(scene.get("materials", [])):
        if "grid" in m:
            raise ValueError("material grids have no XML form")
        tag = "LorentzMaterial" if m.get("lorentz") else ("DebyeMaterial" if m.get("debye") else "Material")
        attrs = {}
        orders = m.get("lorentz") or m.get("debye") or []
        if orders:
            attrs["Order"] = str(len(orders))
        el = prop(tag, f"mat{k}", **attrs)
        _box(ET.SubElement(el, "Primitives"), m.get("priority", 0), m["start"], m["stop"])
        pe = ET.SubElement(el, "Property", Epsilon=_fmt(m.get("eps", 1.0)), Mue=_fmt(m.get("mue", 1.0)),
                           Kappa=_fmt(m.get("kappa", 0.0)), Sigma=_fmt(m.get("sigma", 0.0)))
        for o, d in enumerate(m.get("lorentz") or []):
            for key, attr in (("eps_plasma", "EpsilonPlasmaFrequency"), ("eps_relax", "EpsilonRelaxTime"),
                              ("eps_pole", "EpsilonLorPoleFrequency"), ("mue_plasma", "MuePlasmaFrequency"),
                              ("mue_relax", "MueRelaxTime"), ("mue_pole", "MueLorPoleFrequency")):
                pe.set(f"{attr}_{o + 1}", _fmt(d.get(key, 0.0)))
        for o, d in enumerate(m.get("debye") or []):
            pe.set(f"EpsilonDelta_{o + 1}", _fmt(d.get("eps_delta", 0.0)))
            pe.set(f"EpsilonRelaxTime_{o + 1}", _fmt(d.get("eps_relax", 0.0)))
    for k, m in enumerate(scene.get("metals", [])):
        el = prop("Metal", f"metal{k}")
        pr = ET.SubElement(el, "Primitives")
        if "curve" in m:
            _curve(pr, m["curve"])
        else:
            _prim(pr, m)
    for k, le in enumerate(scene.get("lumped", [])):
        el = prop("LumpedElement", f"lumped{k}", Direction=str(int(le["dir"])), Caps=str(int(bool(le.get("caps", True)))),
                  R=_fmt(le.get("R", -1.0)), C=_fmt(le.get("C", 0.0)), L=_fmt(le.get("L", 0.0)),
                  LEtype="1" if le.get("type", "parallel") == "series" else "0")
        _box(ET.SubElement(el, "Primitives"), le.get("priority", 0), le["start"], le["stop"])
    for k, e in enumerate(scene.get("excitations", [])):
        el = prop("Excitation", f"excite{k}", Type=str(int(e.get("type", 0))), Excite=_fmt(e.get("val", (0, 0, 0))),
                  Delay=_fmt(e.get("delay", 0.0)))
        if e.get("weight") is not None:
            ET.SubElement(el, "Weight", **{k: ("1" if f is None else str(f)) for k, f in zip("XYZ", e["weight"])})
        pr = ET.SubElement(el, "Primitives")
        if "curve" in e:
            _curve(pr, e["curve"])
        else:
            _prim(pr, e)
    for k, p in enumerate(scene.get("probes", [])):
        el = prop("ProbeBox", p.get("name") or f"probe{k}", Type=str(int(p["type"])),
                  Weight=_fmt(p.get("weight", 1.0)), NormDir=str(int(p.get("norm_dir", -1))))
        for key, attr in (("start_time", "StartTime"), ("stop_time", "StopTime")):
            if p.get(key):
                el.set(attr, repr(float(p[key])))
        _box(ET.SubElement(el, "Primitives"), 0, p["start"], p["stop"])
        if p.get("freqs"):
            ET.SubElement(el, "FD_Samples").text = ",".join(repr(float(f)) for f in p["freqs"])
    for k, d in enumerate(scene.get("dumps", [])):
        fld = {"E": 0, "H": 1, 0: 0, 1: 1}[d.get("field", "E")] + (0 if d.get("domain") == "td" else 10)
        el = prop("DumpBox", d.get("name") or f"dump{k}", DumpType=str(fld), DumpMode=str(int(d.get("interp", 1))),
                  FileType=str(int(d.get("file_type", 0))))
        for key, attr in (("start_time", "StartTime"), ("stop_time", "StopTime")):
            if d.get(key):
                el.set(attr, repr(float(d[key])))
        _box(ET.SubElement(el, "Primitives"), 0, d["start"], d["stop"])
        if d.get("freqs"):
            ET.SubElement(el, "FD_Samples").text = ",".join(repr(float(f)) for f in d["freqs"])
    grid = ET.SubElement(csx, "RectilinearGrid", DeltaUnit=repr(float(scene.get("delta_unit", 1.0))),
                         CoordSystem="0")
    for tag, l in zip(("XLines", "YLines", "ZLines"), scene["lines"]):
        ET.SubElement(grid, tag).text = ",".join(repr(float(v)) for v in l)
    text = ET.tostring(root, encoding="unicode")
    if path:
        with open(path, "w") as f:
            f.write('<?xml version="1.0" encoding="UTF-8" standalone="yes" ?>\n' + text + "\n")
    return text


def run_xml(path, out_dir=None, **sim_kwargs):
    """``openEMS <path>`` on the HIP engine: read the setup, build the
    operator, run up to NumberOfTimesteps (or the end criterion), write the
    probe files and the HDF5 field dumps (FileType 1) into out_dir (default:
    the setup's folder) and return
    (Simulation results, list of files written)."""
    from .simulation import Simulation
    scene = read_xml(path)
    out_dir = out_dir or os.path.dirname(os.path.abspath(path))
    sim = Simulation(scene, **sim_kwargs)
    try:
        sim.run()
        res = sim.results()
        files = sim.write_probes(out_dir) + sim.write_dumps(out_dir)
    finally:
        sim.close()
    return res, files
