"""Pins the oracle (the CPU restatement of the reference) against what the
mathematics fixes and against the reference's own cross-engine invariant
(TESTSUITE/enginetests/cavity.m: basic, sse, sse-compressed and
multithreaded engines agree on the same setup).  CPU only."""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle.operator_ref import C0, EPS0, MUE0, Z0, RefOperator, nyquist_num
from oracle.oracle import RefEngine, RefProbe, run_reference
from openems_amd import scenes


def _uniform_scene(n=16, d=1.0, bc=(0,) * 6, **kw):
    l = np.arange(n) * d
    s = {"lines": [l, l.copy(), l.copy()], "delta_unit": 1e-3, "bc": list(bc),
         "signal": {"type": "gauss", "f0": 0.0, "fc": 10e9}, "nrts": 100,
         "excitations": [{"type": 0, "val": (0, 0, 1), "start": (n // 2 * d,) * 2 + (n // 2 * d,),
                          "stop": (n // 2 * d, n // 2 * d, (n // 2 + 1) * d)}]}
    s.update(kw)
    return s


@pytest.mark.parametrize("d", [1.0, 0.5, 2.0])
def test_timestep_uniform_vacuum_closed_form(d):
    op = RefOperator(_uniform_scene(12, d)).build()
    delta = d * 1e-3
    exact = delta * math.sqrt(EPS0 * MUE0 / 3.0)
    # the reference forms the CFL sum from float32 EC values: relative error ~1e-8
    assert abs(op.dT - exact) / exact < 1e-7


def test_timestep_var1_uniform_closed_form():
    op = RefOperator(_uniform_scene(12, 1.0, ts_method=1)).build()
    exact = 1e-3 * math.sqrt(EPS0 * MUE0 / 4.0)      # 2/sqrt(16/(LC)) with L=mu0*d, C=eps0*d
    assert abs(op.dT - exact) / exact < 1e-6


def test_vacuum_coefficients_closed_form():
    op = RefOperator(_uniform_scene(12)).build()
    d = 1e-3
    inner = (slice(None), slice(2, 9), slice(2, 9), slice(2, 9))
    assert np.all(op.vv[inner] == 1.0) and np.all(op.ii[inner] == 1.0)
    np.testing.assert_allclose(op.vi[inner], op.dT / (EPS0 * d), rtol=2e-7)
    np.testing.assert_allclose(op.iv[inner], op.dT / (MUE0 * d), rtol=2e-7)


def test_pec_faces_and_last_current_lines_zeroed():
    op = RefOperator(_uniform_scene(10)).build()
    N = op.N
    for n in range(3):
        for c in range(3):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[n] = 0
            hi[n] = N[n] - 1
            if c != n:
                assert not op.vv[(c,) + tuple(lo)].any() and not op.vi[(c,) + tuple(lo)].any()
            assert not op.vv[(c,) + tuple(hi)].any()
            assert not op.ii[(c,) + tuple(hi)].any() and not op.iv[(c,) + tuple(hi)].any()


def test_gaussian_signal_properties():
    op = RefOperator(_uniform_scene(12, nrts=1000)).build()
    fc = 10e9
    v = op.sig_volt
    assert op.exc_length == math.ceil(18.0 / (2 * math.pi * fc) / op.dT)
    # the pulse is cut to the number of timesteps when that is shorter (excitation.cpp:154-159)
    assert RefOperator(_uniform_scene(12, nrts=100)).build().exc_length == 100
    # f0 = 0: s(0) = exp(-9); peak at t0 = 9/(2 pi fc); symmetric about t0
    assert abs(v[0] - math.exp(-9)) <= 2e-7 * math.exp(-9) * 8
    t0 = 9.0 / (2 * math.pi * fc)
    assert abs(int(np.argmax(v)) - t0 / op.dT) <= 1.0
    ts = np.arange(len(v)) * op.dT
    mirror = np.interp(2 * t0 - ts, ts, v)
    ok = (2 * t0 - ts) >= 0
    np.testing.assert_allclose(v[ok], mirror[ok], atol=2e-3)
    assert op.nyquist == nyquist_num(fc, op.dT)
    # current signal is the same pulse half a step later
    np.testing.assert_allclose(op.sig_curr[:-1], np.interp(ts[:-1] + 0.5 * op.dT, ts, v), atol=5e-3)


def test_sinusoid_timestep_snapped_to_period():
    f0 = 15e9
    op = RefOperator(_uniform_scene(12, signal={"type": "sinus", "f0": f0})).build()
    steps = (1 / f0) / op.dT
    assert abs(steps - round(steps)) < 1e-9
    assert op.exc_length == round(2.0 / f0 / op.dT)
    assert op.sig_volt[0] == 0
    n = np.arange(1, op.exc_length)
    np.testing.assert_allclose(op.sig_volt[1:], np.sin(2 * np.pi * f0 * n * op.dT), atol=1e-6)


def test_upml_grading_integrates_to_design_reflection():
    """The default grading kappa(D) of operator_ext_upml.cpp:30 is designed so that
    the round-trip attenuation exp(-2 Z0 * integral(kappa)) equals 1e-6."""
    W, size = 8e-3, 8
    dl = W / size
    val, _ = integrate.quad(lambda D: RefOperator.grading(D, dl, W, Z0, size), 0.0, W)
    assert abs(2 * Z0 * val - math.log(1e6)) < 1e-9 * math.log(1e6)
    # grading grows geometrically by 2.5 per cell
    assert abs(RefOperator.grading(2 * dl, dl, W, Z0, size) / RefOperator.grading(dl, dl, W, Z0, size) - 2.5) < 1e-12


def test_mur_coefficient_vacuum_closed_form():
    s = _uniform_scene(12, bc=[2] * 6)
    op = RefOperator(s).build()
    d = 1e-3
    c0t = C0 * op.dT
    want = np.float32((c0t - d) / (c0t + d))
    for f in op.mur:
        assert np.all(f.coeff_nyP == want) and np.all(f.coeff_nyPP == want)


def test_quarter_cell_averaging_at_a_dielectric_interface():
    """Half-space eps_r=4 for z >= z_k: tangential edges on the interface see the
    area average (1+4)/2, normal edges crossing it see 4 (AverageMatQuarterCell)."""
    n, k = 12, 6
    s = _uniform_scene(n, materials=[{"eps": 4.0, "start": (-1, -1, k), "stop": (n, n, n + 1)}])
    op = RefOperator(s)
    op.calc_ec()
    e_x = op.effmat[0][0][4, 4, k] / EPS0
    e_z = op.effmat[2][0][4, 4, k] / EPS0
    e_z_below = op.effmat[2][0][4, 4, k - 1] / EPS0
    assert abs(e_x - 2.5) < 1e-12
    assert abs(e_z - 4.0) < 1e-12
    assert abs(e_z_below - 1.0) < 1e-12


def _cavity_like_scene():
    """TESTSUITE/enginetests/cavity.m-like: 27x11x33 lines, BC
    MUR, PML_8, PMC, PEC, PEC, PEC, an eps_r=3.66 brick and a soft source."""
    lx = np.arange(27) * 1.0
    ly = np.arange(11) * 1.0
    lz = np.arange(33) * 1.0
    return {"lines": [lx, ly, lz], "delta_unit": 1e-3, "bc": [2, 3, 1, 0, 0, 0], "pml_size": [8] * 6,
            "materials": [{"eps": 3.66, "start": (3, 2, 5), "stop": (10, 8, 20), "priority": 1}],
            "excitations": [{"type": 0, "val": (0, 1, 1), "start": (6, 5, 12), "stop": (6, 6, 13)}],
            "signal": {"type": "gauss", "f0": 10e9, "fc": 8e9}, "nrts": 300}


@pytest.mark.parametrize("kind,threads", [("sse", 1), ("sse_compressed", 1), ("multithreaded", 2),
                                          ("multithreaded", 3)])
def test_cross_engine_identity(oracle_lib, kind, threads):
    op = RefOperator(_cavity_like_scene()).build()
    ref = RefEngine(op, "basic")
    ref.iterate(300)
    v0, c0 = ref.fields()
    e = RefEngine(op, kind, threads)
    e.iterate(300)
    v1, c1 = e.fields()
    assert np.max(np.abs(v0)) > 0
    # bit-identical apart from flush-to-zero of subnormals in the SSE engines
    tiny = 1e-37
    assert np.all((v0 == v1) | ((np.abs(v0) < tiny) & (np.abs(v1) < tiny)))
    assert np.all((c0 == c1) | ((np.abs(c0) < tiny) & (np.abs(c1) < tiny)))


def test_compressed_index_first_seen_order(oracle_lib):
    op = RefOperator(_cavity_like_scene()).build()
    e = RefEngine(op, "sse_compressed")
    idx = e.opindex().ravel()
    nu = e.num_unique()
    assert idx[0] == 0 and idx.max() == nu - 1
    seen = -1
    for v in idx:                       # a new index is always the next unused one
        assert v <= seen + 1
        seen = max(seen, int(v))


def test_pec_cavity_resonance(oracle_lib):
    """Lowest TM110 resonance of a 20x16x12 mm PEC box, c/2*sqrt(1/a^2+1/b^2),
    within 1 % (numerical dispersion) from an E_z probe spectrum."""
    a, b, c = 20, 16, 12
    s = {"lines": [np.arange(a + 1) * 1.0, np.arange(b + 1) * 1.0, np.arange(c + 1) * 1.0],
         "delta_unit": 1e-3, "bc": [0] * 6, "signal": {"type": "gauss", "f0": 12e9, "fc": 8e9},
         "nrts": 6000,
         "excitations": [{"type": 0, "val": (0, 0, 1), "start": (3, 4, 5), "stop": (3, 4, 6)}]}
    op = RefOperator(s).build()
    pr = RefProbe(op, 2, (13, 9, 6), (13, 9, 6), oversampling=4)
    run_reference(op, 6000, probes=[pr])
    td = pr.td_array()
    sig = td[:, 3] - td[:, 3].mean()
    dt = td[1, 0] - td[0, 0]
    spec = np.abs(np.fft.rfft(sig * np.hanning(len(sig)), 1 << 17))
    f = np.fft.rfftfreq(1 << 17, dt)
    band = (f > 8e9) & (f < 16e9)
    fpk = f[band][np.argmax(spec[band])]
    f110 = C0 / 2 * math.sqrt((1 / (a * 1e-3)) ** 2 + (1 / (b * 1e-3)) ** 2)
    assert abs(fpk - f110) / f110 < 0.01


def test_voltage_probe_is_line_sum(oracle_lib):
    s = scenes.pec_cavity(24, 60)
    op = RefOperator(s).build()
    pr = RefProbe(op, 0, (15, 15, 0), (15, 15, 23))
    eng = run_reference(op, 60, probes=[pr])
    v, _ = eng.raw_views()
    assert pr.td[-1][1][0] == pytest.approx(float(np.sum(v[2, 15, 15, :23].astype(np.float64))), rel=1e-12)
    pr2 = RefProbe(op, 0, (15, 15, 23), (15, 15, 0))          # reversed direction flips the sign
    pr2.process(60, v, None)
    assert pr2.td[-1][1][0] == -pr.td[-1][1][0]


def test_steady_state_probe_set_and_convergence():
    """openems.cpp:1205-1234: three centre probes, then per axis the centre
    with pos[n] * (1/4) and * (3/4) in integer arithmetic (so pos[n] = 0,
    twice): 21 probes.  With no excitation the probes never hold power and
    the detection reports 1; with a sinusoid in an absorbing box the last
    difference falls well below 1 after the transient."""
    import numpy as np
    from openems_amd.operator import Operator
    from openems_amd.simulation import steady_state_probes
    from openems_amd import scenes
    from oracle.operator_ref import RefOperator
    from oracle.oracle import RefEngine
    s = scenes.mur_layers(20, 18, 16, 2000)
    op = Operator(s)
    per, dirs, pos = steady_state_probes(op)
    assert per == int(round(op.period / op.dT)) and per > 0
    assert len(dirs) == 21 and dirs[:3] == [0, 1, 2]
    c = [n // 2 for n in op.N]
    assert pos[:3] == [c, c, c]
    for n in range(3):
        q = list(c)
        q[n] = 0
        assert pos[3 + 6 * n: 9 + 6 * n] == [q] * 6
    # no excitation: never valid -> 1
    s0 = dict(s)
    s0["excitations"] = [dict(s["excitations"][0], val=(0.0, 0.0, 0.0))]
    ref0 = RefEngine(RefOperator(s0).build(), "basic", steady_state=(per, dirs, pos))
    ref0.iterate(6 * per)
    h0 = ref0.steady_history()
    assert [t for t, _ in h0] == [k * per for k in range(2, 6)]
    assert all(d == 1.0 for _, d in h0)
    ref0.close()
    ref = RefEngine(RefOperator(s).build(), "basic", steady_state=(per, dirs, pos))
    ref.iterate(2000)
    h = ref.steady_history()
    assert len(h) >= 10
    ds = [d for _, d in h]
    assert ds[-1] < 0.1 and ds[-1] < 0.5 * max(ds)
    assert all(b <= a for a, b in zip(ds[len(ds) // 2:], ds[len(ds) // 2 + 1:]))   # settling
    ref.close()
    op.close()


def _drude_scene(tau=2e-11, fp=15e9, pole=0.0, mue=False):
    """Uniform 20^3 PEC box with one dispersive box."""
    l = np.arange(20, dtype=np.float64)
    d = {"eps_plasma": fp, "eps_relax": tau, "eps_pole": pole}
    if mue:
        d.update({"mue_plasma": fp, "mue_relax": tau, "mue_pole": pole})
    return {"lines": [l, l.copy(), l.copy()], "delta_unit": 1e-3, "bc": [0] * 6,
            "materials": [{"eps": 2.0, "mue": 1.5, "priority": 1, "start": (5, 5, 5), "stop": (12, 11, 13),
                           "lorentz": [d]}],
            "excitations": [{"type": 0, "val": (0, 0, 1.0), "start": (3, 3, 3), "stop": (3, 3, 4)}],
            "signal": {"type": "gauss", "f0": 0.0, "fc": 10e9}, "nrts": 100}


def test_drude_coefficients_closed_form():
    """Drude order (plasma frequency wp, relaxation time tau) on an edge of
    capacitance C: L_D = 1/(wp^2 C), R_D = L_D/tau, so
    v_int = (2 - dT/tau)/(2 + dT/tau), independent of C, and
    v_ext = dT wp^2 C / (1 + dT/(2 tau)) * vi; the Lorentz pole wl adds
    v_lor = dT wl^2 / (wp^2 C vi).  Same on the current side with L, iv."""
    tau, fp, pole = 2e-11, 15e9, 9e9
    ref = RefOperator(_drude_scene(tau, fp, pole, mue=True)).build()
    (o,) = ref.lorentz
    dT = ref.dT
    C, G, L, R = ref.EC
    wp, wl = 2 * np.pi * fp, 2 * np.pi * pole
    x, y, z = o["pos"].astype(np.int64)
    assert o["volt"] and o["curr"] and o["volt_lor"] and o["curr_lor"]
    for n in range(3):
        c = C[n][x, y, z].astype(np.float64)
        vi = ref.vi[n][x, y, z].astype(np.float64)
        act = (o["v_ext"][n] != 0)
        assert act.sum() > 50
        assert np.allclose(o["v_int"][n][act], (2 - dT / tau) / (2 + dT / tau), rtol=2e-7, atol=0)
        want = dT * wp ** 2 * c / (1 + dT / (2 * tau)) * vi
        assert np.allclose(o["v_ext"][n][act], want[act], rtol=1e-6, atol=0)
        assert np.allclose(o["v_lor"][n][act], (dT * wl ** 2 / (wp ** 2 * c * vi))[act], rtol=1e-6, atol=0)
        assert np.all(o["v_int"][n][~act] == 1) and np.all(o["v_lor"][n][~act] == 0)
        lc = L[n][x, y, z].astype(np.float64)
        iv = ref.iv[n][x, y, z].astype(np.float64)
        ac = (o["i_ext"][n] != 0)
        assert ac.sum() > 50
        assert np.allclose(o["i_int"][n][ac], (2 - dT / tau) / (2 + dT / tau), rtol=2e-7, atol=0)
        assert np.allclose(o["i_ext"][n][ac], (dT * wp ** 2 * lc / (1 + dT / (2 * tau)) * iv)[ac], rtol=1e-6, atol=0)
        assert np.allclose(o["i_lor"][n][ac], (dT * wl ** 2 / (wp ** 2 * lc * iv))[ac], rtol=1e-6, atol=0)


def test_dispersive_cells_are_those_inside_the_box():
    """Brute force over the mesh: a cell is listed exactly when one of its
    voltage edge centres (primary) or current edge centres (dual) lies in the
    closed dispersive box with a non-zero coefficient; Drude only (no pole)."""
    sc = _drude_scene()
    ref = RefOperator(sc).build()
    (o,) = ref.lorentz
    assert not o["volt_lor"] and not o["curr"]
    lo, hi = np.array(sc["materials"][0]["start"]), np.array(sc["materials"][0]["stop"])
    N = ref.N
    want = []
    for x in range(N[0]):
        for y in range(N[1]):
            for z in range(N[2]):
                on = False
                for n in range(3):
                    p = [x, y, z]
                    if p[n] >= N[n] - 1 or ref.vi[n][x, y, z] == 0:
                        continue
                    pt = np.array(p, np.float64)
                    pt[n] += 0.5
                    on |= bool(np.all(pt >= lo) & np.all(pt <= hi))
                if on:
                    want.append((x, y, z))
    assert [tuple(int(v) for v in c) for c in o["pos"].T] == want


def _rlc_scene(kind, n0, R=20.0, L=2e-9, Cap=1e-12):
    l = np.arange(16, dtype=np.float64)
    return {"lines": [l, l.copy(), l.copy()], "delta_unit": 1e-3, "bc": [0] * 6,
            "lumped": [{"dir": 2, "R": R, "C": Cap, "L": L, "caps": False, "type": kind,
                        "start": (7, 7, 6), "stop": (7, 7, 6 + n0)}],
            "excitations": [{"type": 0, "val": (0, 0, 1.0), "start": (3, 3, 3), "stop": (3, 3, 4)}],
            "signal": {"type": "gauss", "f0": 0.0, "fc": 10e9}, "nrts": 100}


def test_series_rlc_spread_over_edges():
    """A series RLC over n0 edges in series: each edge carries L/n0, R/n0 and
    C*n0, so L*C and R*C per edge stay those of the element.  Hence ib0
    doubles and b1, b2 halve from one edge to two, and for one edge
    ib0 = 2 dT C / (4 L C + 2 dT R C + dT^2)."""
    a = RefOperator(_rlc_scene("series", 1)).build().rlc
    b = RefOperator(_rlc_scene("series", 2)).build().rlc
    assert a["dir"].size == 1 and b["dir"].size == 2
    assert np.allclose(b["ib0"], 2 * a["ib0"][0], rtol=1e-6)
    assert np.allclose(b["b1"], a["b1"][0] / 2, rtol=1e-6) and np.allclose(b["b2"], a["b2"][0] / 2, rtol=1e-6)
    ref = RefOperator(_rlc_scene("series", 1)).build()
    dT, R, L, Cap = ref.dT, 20.0, 2e-9, 1e-12
    assert a["ib0"][0] == pytest.approx(2 * dT * Cap / (4 * L * Cap + 2 * dT * R * Cap + dT * dT), rel=1e-6)
    assert a["b1"][0] == pytest.approx((dT * dT - 4 * L * Cap) / (dT * Cap), rel=1e-6)
    assert np.all(a["i2v"] == 0) and np.all(a["ilv"] == 0)
    # the edge's own conductance is cleared and its coefficients recomputed
    assert ref.EC[1][2][7, 7, 6] == 0


def test_parallel_rlc_conductance_and_inductance():
    """Parallel RLC over one edge: C and G = 1/R replace the edge's own, so
    vi = dT/C / (1 + dT/(2 R C)) and i2v equals it; ilv = dT/L."""
    ref = RefOperator(_rlc_scene("parallel", 1)).build()
    t = ref.rlc
    dT = ref.dT
    assert t["dir"].size == 1
    want = (dT / 1e-12) / (1 + dT / (2 * 20.0 * 1e-12))
    assert t["i2v"][0] == pytest.approx(want, rel=1e-6)
    assert ref.vi[2][7, 7, 6] == pytest.approx(want, rel=1e-6)
    assert t["ilv"][0] == pytest.approx(dT / 2e-9, rel=1e-6)
    assert t["vvd"][0] == 1 and t["ib0"][0] == 0
    # a parallel element without L is the main operator's R||C, not an RLC entry
    assert RefOperator(_rlc_scene("parallel", 1, L=0.0)).build().rlc is None


def test_lorentz_pole_flags_are_per_side():
    """m_volt_Lor_ADE_On and m_curr_Lor_ADE_On are separate
    (engine_ext_lorentzmaterial.cpp:86-173): an order with a current-side
    pole and no voltage-side pole applies the pole to the currents only.
    A zero voltage pole table equals none; dropping the current pole
    changes the fields."""
    sc = _drude_scene(mue=True)
    sc["materials"][0]["lorentz"][0].update({"eps_pole": 0.0, "mue_pole": 9e9})
    ref = RefOperator(sc).build()
    (o,) = ref.lorentz
    assert o["curr_lor"] and not o["volt_lor"]

    def run(v_lor, i_lor):
        d = dict(o)
        d["v_lor"], d["i_lor"] = v_lor, i_lor
        e = RefEngine(ref, "basic", lorentz=[d])
        e.iterate(60)
        v, c = (np.array(a) for a in e.raw_views())
        e.close()
        return v, c

    base = run(None, o["i_lor"])
    same = run(np.zeros_like(o["v_lor"]), o["i_lor"])
    nopole = run(None, None)
    assert np.array_equal(base[0], same[0]) and np.array_equal(base[1], same[1])
    assert not np.array_equal(base[1], nopole[1])


def test_debye_coefficients_closed_form():
    """Debye order (delta eps, tau > 2 dT) on a uniform mesh of cell d:
    C_L = eps0 delta_eps d^2 / d, R_D = tau / C_L, so v_int = (2 dT - tau)/tau,
    v_ext = 2 C_L vi / tau and v_lor = dT/(C_L vi); a second order with
    tau < 2 dT stays off.  Voltages only."""
    l = np.arange(16, dtype=np.float64)
    d = 1e-3
    de, tau = 3.0, 2e-11
    sc = {"lines": [l, l.copy(), l.copy()], "delta_unit": d, "bc": [0] * 6,
          "materials": [{"eps": 2.0, "priority": 1, "start": (4, 4, 4), "stop": (10, 11, 9),
                         "debye": [{"eps_delta": de, "eps_relax": tau}, {"eps_delta": 1.0, "eps_relax": 1e-13}]}],
          "excitations": [{"type": 0, "val": (0, 0, 1.0), "start": (2, 2, 2), "stop": (2, 2, 3)}],
          "signal": {"type": "gauss", "f0": 0.0, "fc": 10e9}, "nrts": 100}
    ref = RefOperator(sc).build()
    o0, o1 = ref.lorentz
    dT = ref.dT
    assert o0["volt"] and o0["volt_lor"] and not o0["curr"]
    assert o1["pos"].shape[1] == 0
    CL = EPS0 * de * d * d / d
    x, y, z = o0["pos"].astype(np.int64)
    for n in range(3):
        vi = ref.vi[n][x, y, z].astype(np.float64)
        act = o0["v_ext"][n] != 0
        assert act.sum() > 100
        assert np.allclose(o0["v_int"][n][act], (2 * dT - tau) / tau, rtol=1e-6, atol=0)
        assert np.allclose(o0["v_ext"][n][act], (2 * CL * vi / tau)[act], rtol=1e-6, atol=0)
        assert np.allclose(o0["v_lor"][n][act], (dT / (CL * vi))[act], rtol=1e-6, atol=0)
