"""The reference's Python interface mirrored by openems_amd.pyopenems (CPU):
a script's CSX objects give the same operator as the equivalent scene dict,
a lumped port lays out the reference's properties, and the port DFT."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd.operator import Operator
from openems_amd.pyopenems import DFT_time2freq, ContinuousStructure, openEMS


def _wire_via_api(n=30):
    sc = scenes.wire_antenna(n, 100)
    FDTD = openEMS(NrTS=100, EndCriteria=0)
    FDTD.SetGaussExcite(0, 20e9)
    FDTD.SetBoundaryCond(["PML_6"] * 6)
    CSX = ContinuousStructure()
    FDTD.SetCSX(CSX)
    mesh = CSX.GetGrid()
    mesh.SetDeltaUnit(1e-3)
    for d, l in zip("xyz", sc["lines"]):
        mesh.AddLine(d, l)
    m = sc["materials"][0]
    CSX.AddMaterial("diel", epsilon=m["eps"], kappa=m["kappa"]).AddBox(m["start"], m["stop"], priority=m["priority"])
    wire = CSX.AddMetal("wire")
    for mt in sc["metals"]:
        wire.AddCurve(np.array(mt["curve"]).T)
    for k, e in enumerate(sc["excitations"]):
        CSX.AddExcitation(f"exc{k}", exc_type=e["type"], exc_val=e["val"], delay=e.get("delay", 0.0)).AddCurve(
            np.array(e["curve"]))
    for p in sc["probes"]:
        CSX.AddProbe(p["name"], p_type=p["type"], weight=p["weight"]).AddBox(p["start"], p["stop"])
    return FDTD, sc


def test_script_objects_build_the_scene_operator(native):
    FDTD, sc = _wire_via_api()
    api = FDTD.scene()
    sc["end_criteria"] = 0.0
    a, b = Operator(sc, nthreads=2), Operator(api, nthreads=2)
    assert a.dT == b.dT
    for name in ("vv", "vi", "ii", "iv"):
        assert np.array_equal(a.coeff(name), b.coeff(name)), name
    ea, eb = a.excitation(False), b.excitation(False)
    for k in ("index", "dir", "amp", "delay"):
        assert np.array_equal(ea[k], eb[k])
    assert [p["name"] for p in api["probes"]] == ["feed_ut", "et"]


def test_lumped_port_properties():
    """ports.LumpedPort: resistor with caps over the box, soft excitation
    against the port direction, V probe along it (weight -1), I probe across
    its middle (weight = direction, normal = port axis)."""
    FDTD = openEMS(NrTS=10)
    CSX = ContinuousStructure()
    FDTD.SetCSX(CSX)
    port = FDTD.AddLumpedPort(1, 50, [1, 2, 3], [1, 2, 1], "z", excite=1, priority=5)
    kinds = [(p.kind, p.name) for p in CSX.props]
    assert kinds == [("lumped", "port_resist_1"), ("excitation", "port_excite_1"), ("probe", "port_ut_1"),
                     ("probe", "port_it_1")]
    parts = CSX.scene_parts()
    assert parts["lumped"][0]["R"] == 50 and parts["lumped"][0]["caps"] and parts["lumped"][0]["priority"] == 5
    assert parts["excitations"][0]["val"] == (0.0, 0.0, 1.0)          # -direction * excite, direction = -1
    ut, it = parts["probes"]
    assert ut["weight"] == -1 and ut["start"] == [1, 2, 3] and ut["stop"] == [1, 2, 1]
    assert it["weight"] == -1 and it["norm_dir"] == 2 and it["start"][2] == it["stop"][2] == 2.0
    short = openEMS()
    c2 = ContinuousStructure()
    short.SetCSX(c2)
    short.AddLumpedPort(2, 0, [0, 0, 0], [0, 0, 1], 2)
    assert c2.props[0].kind == "metal"
    with pytest.raises(ValueError):
        short.AddLumpedPort(3, 50, [0, 0, 0], [0, 1, 0], "z")


def test_port_dft_closed_form():
    """DFT of a sampled cosine at its own frequency over whole periods:
    'pulse' scaling gives amplitude x duration, 'periodic' the amplitude."""
    f = 2e9
    t = np.arange(400) * (1 / f / 40)          # 10 periods, 40 samples each
    v = 3.0 * np.cos(2 * np.pi * f * t)
    assert abs(DFT_time2freq(t, v, [f], "periodic")[0] - 3.0) < 1e-9
    assert abs(DFT_time2freq(t, v, f, "pulse")[0] - 3.0 * 400 * (t[1] - t[0])) < 1e-18
    with pytest.raises(ValueError):
        DFT_time2freq(t, v, f, "other")
