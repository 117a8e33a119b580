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


def coax_setup(res=5.0, length=1000.0, nrts=5000):
    """The structure of the reference's TESTSUITE/combinedtests/Coax.m as a
    script of this interface: a coaxial line along z (inner PEC cylinder of
    radius 100, outer PEC shell from 230 to 240, drawing unit 1 mm), PEC on
    five faces and PML_8 at z max, a soft E excitation weighted x / rho^2,
    y / rho^2 on a shell between the conductors in the first half cell, a
    voltage probe across the gap and a current probe around the inner
    conductor at z = length / 2.  Returns (openEMS object, Z0 of the line)."""
    eps0, mue0 = 8.85418781762e-12, 1.256637062e-6
    ri, rai, raa = 100.0, 230.0, 240.0
    FDTD = openEMS(NrTS=nrts, EndCriteria=1e-6)
    FDTD.SetGaussExcite(0, 1e9)
    FDTD.SetBoundaryCond(["PEC", "PEC", "PEC", "PEC", "PEC", "PML_8"])
    CSX = ContinuousStructure()
    FDTD.SetCSX(CSX)
    mesh = CSX.GetGrid()
    mesh.SetDeltaUnit(1e-3)
    x = -2.5 * res - raa + res * np.arange(int(round((2 * raa + 5 * res) / res)) + 1)
    mesh.SetLines("x", x)
    mesh.SetLines("y", x.copy())
    mesh.SetLines("z", np.linspace(0, length, int(round(length / res)) + 1))
    pec = CSX.AddMetal("PEC")
    pec.AddCylinder([0, 0, 0], [0, 0, length], ri, priority=1)
    pec.AddCylindricalShell([0, 0, 0], [0, 0, length], 0.5 * (raa + rai), raa - rai, priority=0)
    exc = CSX.AddExcitation("excite", 0, [1, 1, 0])
    exc.SetWeightFunction(["(x)/(x*x+y*y)", "y/pow(rho,2)", "0"])
    exc.AddCylindricalShell([0, 0, 0], [0, 0, res / 2], 0.5 * (ri + rai), rai - ri)
    CSX.AddProbe("ut1", 0).AddBox([ri, 0, length / 2], [rai, 0, length / 2])
    mid = ri + 3 * res
    CSX.AddProbe("it1", 1).AddBox([-mid, -mid, length / 2], [mid, mid, length / 2])
    Z0 = np.sqrt(mue0 / eps0) * np.log(rai / ri) / (2 * np.pi)
    return FDTD, Z0


def test_cylinders_and_weights_match_oracle(native):
    """Cylinder and cylindrical-shell metals and a weighted excitation on a
    shell (a coarse coax of the reference's Coax test): coefficients and
    excitation lists bit-identical to the oracle, and the excitation
    amplitudes follow the weight x / rho^2, y / rho^2 times the edge length."""
    from tests.test_operator_parity import _compare
    FDTD, _ = coax_setup(res=10.0, length=200.0, nrts=100)
    scene = FDTD.scene()
    op, ref = _compare(scene)
    vv = op.coeff("vv")
    L = [np.asarray(l, np.float64) for l in scene["lines"]]
    # E_z edges inside the inner conductor (rho <= 100) are PEC, those in the gap are not
    xi = int(np.argmin(np.abs(L[0] - 2.5)))
    assert vv[2, xi, xi, 5] == 0
    xg = int(np.argmin(np.abs(L[0] - 162.5)))
    assert vv[2, xg, xi, 5] != 0
    e = op.excitation(False)
    assert e["dir"].size > 100 and set(np.unique(e["dir"])) == {0, 1}
    idx, d, amp = e["index"], e["dir"], e["amp"]
    pos = np.stack(np.unravel_index(idx, tuple(op.N)), 0) if idx.ndim == 1 else idx
    dual = [(l[:-1] + l[1:]) / 2 for l in L]
    for k in range(0, d.size, 7):
        p = pos[:, k]
        n = int(d[k])
        P = [dual[a][p[a]] if a == n else L[a][p[a]] for a in range(3)]
        rho2 = P[0] ** 2 + P[1] ** 2
        w = P[n] / rho2
        assert np.isclose(amp[k], w * (L[n][p[n] + 1] - L[n][p[n]]) * 1e-3, rtol=1e-6)


def test_cylinder_arguments_are_checked(native):
    with pytest.raises(Exception):
        FDTD, _ = coax_setup(res=10.0, length=200.0, nrts=100)
        FDTD.CSX.AddMetal("bad").AddCylinder([0, 0, 0], [0, 0, 0], 5.0)
        Operator(FDTD.scene(), nthreads=1)
    FDTD, _ = coax_setup(res=10.0, length=200.0, nrts=100)
    FDTD.CSX.props[-3].SetWeightFunction(["x+", "0", "0"])
    with pytest.raises(Exception, match="weight"):
        Operator(FDTD.scene(), nthreads=1)


def _msl_setup(nrts=3000):
    """A 50 Ohm-class microstrip (eps_r 3.66, 0.254 mm substrate, 0.6 mm strip,
    drawing unit 1 um) between two microstrip ports of 1 mm, port 1 fed,
    both ends terminated with Feed_R = 50, 1.5 mm of substrate and an 8-cell
    UPML beyond each end."""
    from openems_amd.pyopenems import MSLPort  # noqa: F401
    FDTD = openEMS(NrTS=nrts, EndCriteria=1e-4)
    FDTD.SetGaussExcite(0, 20e9)
    FDTD.SetBoundaryCond(["PML_8"] * 4 + ["PEC", "PML_8"])
    CSX = ContinuousStructure()
    FDTD.SetCSX(CSX)
    mesh = CSX.GetGrid()
    mesh.SetDeltaUnit(1e-6)
    mesh.SetLines("x", np.arange(-7500, 7501, 100.0))
    mesh.SetLines("y", np.concatenate([np.arange(-2000, -400, 100.0), np.arange(-400, 401, 50.0),
                                       np.arange(500, 2001, 100.0)]))
    mesh.SetLines("z", np.concatenate([np.linspace(0, 254, 5), np.arange(354, 2001, 100.0)]))
    CSX.AddMaterial("RO4350B", epsilon=3.66).AddBox([-7500, -2000, 0], [7500, 2000, 254])
    pec = CSX.AddMetal("PEC")
    p1 = FDTD.AddMSLPort(1, pec, [-6000, -300, 254], [-5000, 300, 0], "x", "z", excite=-1, Feed_R=50, priority=10)
    p2 = FDTD.AddMSLPort(2, pec, [6000, -300, 254], [5000, 300, 0], "x", "z", MeasPlaneShift=500, Feed_R=50,
                         priority=10)
    pec.AddBox([-5000, -300, 254], [5000, 300, 254], priority=10)
    return FDTD, p1, p2


def test_msl_port_structure():
    """ports.MSLPort: three voltage probes across the substrate on the mesh
    lines around the measurement plane, two current probes around the strip
    half way between them, the feed on the port's first line, the strip in
    the plane start[exc], and Feed_R as a lumped element at the port start."""
    FDTD, p1, p2 = _msl_setup()
    parts = FDTD.CSX.scene_parts()
    probes = {p["name"]: p for p in parts["probes"]}
    assert p1.U_filenames == ["port_ut_1A", "port_ut_1B", "port_ut_1C"]
    assert [probes[n]["start"][0] for n in p1.U_filenames] == [-5600.0, -5500.0, -5400.0]
    assert all(probes[n]["start"][2] == 254 and probes[n]["stop"][2] == 0 for n in p1.U_filenames)
    assert [probes[n]["start"][0] for n in p1.I_filenames] == [-5550.0, -5450.0]
    assert all(probes[n]["weight"] == 1 and probes[n]["norm_dir"] == 0 for n in p1.I_filenames)
    # port 2 runs towards -x: probes in descending x, current weight -1
    assert [probes[n]["start"][0] for n in p2.U_filenames] == [5600.0, 5500.0, 5400.0]
    assert all(probes[n]["weight"] == -1 for n in p2.I_filenames)
    assert p1.measplane_shift == 500 and p2.measplane_shift == 500
    exc = [e for e in parts["excitations"]]
    assert len(exc) == 1 and exc[0]["start"][0] == exc[0]["stop"][0] == -6000.0
    assert exc[0]["val"] == (0.0, 0.0, -1.0)          # -upside_down * excite, upside_down = -1, excite = -1
    assert len(parts["lumped"]) == 2 and all(le["R"] == 50 and le["dir"] == 2 for le in parts["lumped"])


def test_msl_port_beta_and_impedance_of_a_travelling_wave(tmp_path):
    """ReadUIData's finite-difference beta and Z_ref on probe files of an
    ideal travelling wave u(x, t) = g(t - x / v), i = u / Z0, pinned against
    the closed forms: beta -> 2 pi f / v and Z_ref -> Z0 as the probe spacing
    shrinks against the wavelength; a plane shift of the reference plane by
    the measurement-plane offset leaves the incident wave at the port start."""
    from openems_amd.pyopenems import MSLPort
    FDTD, p1, _ = _msl_setup()
    v, Z0 = 1.6e8, 48.0
    dt = 5e-14
    t = np.arange(6000) * dt
    g = lambda tt: np.exp(-((tt - 6e-11) / 1.5e-11) ** 2)
    unit = 1e-6
    probes = {p["name"]: p for p in FDTD.CSX.scene_parts()["probes"]}
    for n in p1.U_filenames:
        x = (probes[n]["start"][0] + 6000) * unit
        np.savetxt(str(tmp_path / n), np.c_[t, g(t - x / v)], header="t u", comments="% ")
    for n in p1.I_filenames:
        x = (probes[n]["start"][0] + 6000) * unit
        np.savetxt(str(tmp_path / n), np.c_[t + dt / 2, g(t + dt / 2 - x / v) / Z0], header="t i", comments="% ")
    f = np.array([2e9, 5e9, 10e9])
    p1.CalcPort(str(tmp_path), f)
    assert np.allclose(np.real(p1.beta), 2 * np.pi * f / v, rtol=2e-2)
    assert np.allclose(np.abs(p1.Z_ref), Z0, rtol=2e-2)
    assert np.all(np.abs(p1.uf_ref) < 0.03 * np.abs(p1.uf_inc))
    assert isinstance(p1, MSLPort)


def test_port_set_enabled():
    """Port.SetEnabled switches a port's excitation: a disabled fed port adds
    no excitation to the scene; enabling a port without one raises."""
    FDTD, p1, p2 = _msl_setup()
    assert len(FDTD.scene()["excitations"]) == 1
    p1.SetEnabled(False)
    assert FDTD.scene()["excitations"] == []
    p1.SetEnabled(True)
    assert len(FDTD.scene()["excitations"]) == 1
    with pytest.raises(ValueError, match="no excitation"):
        p2.SetEnabled(True)
    p2.SetEnabled(False)


def test_line_waves_and_plane_move_closed_forms():
    """line_waves on a line of impedance Z: a pure forward wave (U = Z I)
    has no backward part, a short (U = 0) and an open end (I = 0) reflect
    fully with -1 and +1, and the parts recombine to U and I.  move_plane by
    theta delays the forward wave by exp(-j theta) and advances the backward
    one; moving by theta and back by -theta is the identity; a quarter wave
    turns a short into an open (the impedance inverter Z^2 / Z_L)."""
    from openems_amd.pyopenems import line_waves, move_plane
    Z = 50.0
    uf, ub, i_f, i_b = line_waves(2.0 + 1.0j, (2.0 + 1.0j) / Z, Z)
    assert abs(ub) < 1e-15 and abs(uf - (2 + 1j)) < 1e-15 and abs(i_f - uf / Z) < 1e-15
    uf, ub, _, _ = line_waves(0.0, 0.1, Z)
    assert abs(ub / uf + 1) < 1e-15
    uf, ub, _, _ = line_waves(3.0, 0.0, Z)
    assert abs(ub / uf - 1) < 1e-15
    rng = np.random.default_rng(1)
    U, I = rng.normal(size=8) + 1j * rng.normal(size=8), rng.normal(size=8) + 1j * rng.normal(size=8)
    uf, ub, i_f, i_b = line_waves(U, I, Z)
    assert np.allclose(uf + ub, U) and np.allclose(i_f - i_b, I)
    th = 0.7
    U2, I2 = move_plane(U, I, Z, th)
    uf2, ub2, _, _ = line_waves(U2, I2, Z)
    assert np.allclose(uf2, uf * np.exp(-1j * th)) and np.allclose(ub2, ub * np.exp(1j * th))
    U3, I3 = move_plane(U2, I2, Z, -th)
    assert np.allclose(U3, U) and np.allclose(I3, I)
    Us, Is = move_plane(np.zeros(1, complex), np.ones(1, complex), Z, np.pi / 2)    # short, a quarter wave on
    assert abs(Is[0]) < 1e-12 and abs(Us[0]) > 1                                      # -> open


def test_line_from_three_planes_on_a_lossy_line():
    """line_from_three_planes on exact phasors of a forward wave on a lossy
    line U(z) = exp(-gamma z), I(z) = U(z) / Z_L: beta tends to Im(gamma) and
    Z_L to its value as the probe spacing shrinks (second-order central
    differences), with a backward wave the estimate is no longer Z_L."""
    from openems_amd.pyopenems import line_from_three_planes
    gamma = 2.0 + 300.0j            # 1/m
    ZL = 48.0 - 2.0j
    for h, tol in ((1e-3, 2e-2), (2e-4, 1e-3)):
        U = lambda z: np.exp(-gamma * z)
        za, zb, zc = -h, 0.0, h
        z1, z2 = -h / 2, h / 2
        beta, Z, Ub, Ib = line_from_three_planes(U(za), U(zb), U(zc), 2 * h, U(z1) / ZL, U(z2) / ZL, h)
        assert abs(beta - gamma.imag + 1j * gamma.real) / abs(gamma) < tol   # beta = -j gamma
        assert abs(Z - ZL) / abs(ZL) < tol


def test_abort_conditions(tmp_path, monkeypatch):
    """openEMS::CheckAbortCond restated in Simulation.check_abort: the abort
    flag and a file named ABORT in the working directory (CPU: the check
    alone, no engine)."""
    from openems_amd.simulation import Simulation
    sim = Simulation.__new__(Simulation)
    monkeypatch.chdir(tmp_path)
    assert not sim.check_abort()
    sim.SetAbort(True)
    assert sim.check_abort()
    sim.SetAbort(False)
    assert not sim.check_abort()
    (tmp_path / "ABORT").write_text("")
    assert sim.check_abort()
