"""A script in the reference's Python interface run on the engine
(openems_amd.pyopenems): probe files written into sim_path, the lumped
port's incident / reflected split computed from them, equal to the same
scene run directly."""
import numpy as np
import pytest

from openems_amd.pyopenems import ContinuousStructure, openEMS
from openems_amd.simulation import Simulation

pytestmark = pytest.mark.gpu


def test_lumped_port_script(gpu, tmp_path):
    FDTD = openEMS(NrTS=1500, EndCriteria=1e-9)
    FDTD.SetGaussExcite(8e9, 6e9)
    FDTD.SetBoundaryCond(["PML_6"] * 6)
    CSX = ContinuousStructure()
    FDTD.SetCSX(CSX)
    mesh = CSX.GetGrid()
    mesh.SetDeltaUnit(1e-3)
    for d in "xyz":
        mesh.SetLines(d, np.arange(-7.0, 7.01, 0.5))
    CSX.AddMaterial("substrate", epsilon=3.66, kappa=0.002).AddBox([-5, -5, -1.5], [5, 5, 0], priority=1)
    CSX.AddMetal("patch").AddBox([-3, -2.5, 0], [3, 2.5, 0], priority=10)
    CSX.AddMetal("ground").AddBox([-5, -5, -1.5], [5, 5, -1.5], priority=10)
    port = FDTD.AddLumpedPort(1, 50, [-1.5, 0, -1.5], [-1.5, 0, 0], "z", excite=1, priority=5)
    CSX.AddDump("Ez", dump_type=10, dump_mode=1, file_type=1, frequency=[8e9]).AddBox([-6, -6, -0.5], [6, 6, -0.5])
    print("case lumped port script, 1500 steps", flush=True)
    assert FDTD.Run(str(tmp_path)) == 0
    assert (tmp_path / "port_ut_1").exists() and (tmp_path / "port_it_1").exists() and (tmp_path / "Ez.h5").exists()
    f = np.linspace(4e9, 12e9, 9)
    port.CalcPort(str(tmp_path), f)
    assert np.all(port.P_inc > 0) and np.all(np.isfinite(port.uf_ref))
    Zin = port.uf_tot / port.if_tot
    assert np.all(np.isfinite(Zin)) and np.all(np.real(Zin) > 0)
    s11 = port.uf_ref / port.uf_inc
    assert np.all(np.abs(s11) <= 1.0 + 1e-6)
    # the same scene run directly gives the same probe rows
    sim = Simulation(FDTD.scene())
    sim.run()
    td = sim.results()["probes"]["port_ut_1"]["td"]
    back = np.loadtxt(str(tmp_path / "port_ut_1"), comments="%")
    np.testing.assert_allclose(back, td, rtol=1e-11, atol=1e-300)
    sim.close()


def test_field_probes_equal_field_dumps(gpu, tmp_path):
    """The reference's probe / dump consistency test (TESTSUITE/probes/
    fieldprobes.m): an infinitesimal hard-source dipole with Mur on every
    face; E and H time-domain HDF5 dumps (no interpolation) on the six faces
    of a box and E / H point probes at the face centres.  At every sample the
    dump at the probe point equals the probe (relative difference <= 1e-7,
    times within 1e-13 s), and the fields are not trivially small."""
    from openems_amd import hdf5_out
    c0 = 299792458.0
    unit = 1e-6
    f_max = 1e9
    dl = c0 / f_max / unit / 50
    FDTD = openEMS(NrTS=1200, EndCriteria=1e-6, OverSampling=10)
    FDTD.SetGaussExcite(0, f_max)
    FDTD.SetBoundaryCond([2] * 6)
    CSX = ContinuousStructure()
    FDTD.SetCSX(CSX)
    mesh = CSX.GetGrid()
    mesh.SetDeltaUnit(unit)
    for d in "xyz":
        mesh.SetLines(d, np.arange(-40, 41) * (dl / 2))     # the dipole's line x = y = 0 exactly
    CSX.AddExcitation("infDipole", 1, [0, 0, 1]).AddBox([0, 0, -dl / 2], [0, 0, dl / 2], priority=1)
    s1 = np.array([-4.5] * 3) * dl / 2
    s2 = np.array([4.5] * 3) * dl / 2
    faces = {"xn": (s1, [s1[0], s2[1], s2[2]]), "xp": ([s2[0], s1[1], s1[2]], s2),
             "yn": (s1, [s2[0], s1[1], s2[2]]), "yp": ([s1[0], s2[1], s1[2]], s2),
             "zn": (s1, [s2[0], s2[1], s1[2]]), "zp": ([s1[0], s1[1], s2[2]], s2)}
    for k, (a, b) in faces.items():
        CSX.AddDump(f"Et_{k}", dump_type=0, dump_mode=0, file_type=1).AddBox(a, b)
        CSX.AddDump(f"Ht_{k}", dump_type=1, dump_mode=0, file_type=1).AddBox(a, b)
    coords = [[s1[0], 0, 0], [s2[0], 0, 0], [0, s1[1], 0], [0, s2[1], 0], [0, 0, s1[2]], [0, 0, s2[2]]]
    for n, c in enumerate(coords):
        CSX.AddProbe(f"et{n + 1}", 2).AddBox(c, c)
        CSX.AddProbe(f"ht{n + 1}", 3).AddBox(c, c)
    print("case fieldprobes: dipole, Mur, 12 TD dumps, 12 point probes, 1200 steps", flush=True)
    assert FDTD.Run(str(tmp_path)) == 0
    for n, k in enumerate(faces):
        for fld, lim in (("E", 5e-3), ("H", 1e-7)):
            path = str(tmp_path / f"{fld}t_{k}.h5")
            lines = [hdf5_out.read_dataset(path, f"/Mesh/{a}")[0] for a in "xyz"]
            idx = [int(np.argmin(np.abs(lines[a] - coords[n][a] * unit))) for a in range(3)]
            probe = np.loadtxt(str(tmp_path / f"{fld.lower()}t{n + 1}"), comments="%")
            names = sorted(nm for nm in _td_names(path))
            assert len(names) == probe.shape[0]
            vals = np.zeros((len(names), 3))
            times = np.zeros(len(names))
            for t, nm in enumerate(names):
                data, attrs = hdf5_out.read_dataset(path, "/FieldData/TD/" + nm)
                vals[t] = data[:, idx[2], idx[1], idx[0]]
                times[t] = attrs["time"][0]
            assert np.all(np.abs(times - probe[:, 0]) <= 1e-13)
            with np.errstate(divide="ignore", invalid="ignore"):
                rel = np.abs((vals - probe[:, 1:4]) / vals)
            assert not np.any(rel > 1e-7), (fld, k, np.nanmax(rel))
            assert max(np.abs(vals[:, 0]).max(), np.abs(vals[:, 1]).max(), np.abs(vals[:, 2]).max()) >= lim


def _td_names(path):
    from openems_amd import hdf5_out
    return hdf5_out.list_group(path, "/FieldData/TD")


def test_engine_cross_check_cavity(gpu, tmp_path):
    """The reference's engine cross-check (TESTSUITE/enginetests/cavity.m):
    a 27 x 11 x 33 cavity with boundaries MUR, PML_8, PMC, PEC, PEC, PEC, an
    eps_r = 3.66 brick, a soft [1 1 1] curve excitation, E and H point probes
    and whole-mesh E / H time-domain HDF5 dumps without interpolation, 1000
    steps.  Every dumped sample is bit-identical to the oracle engine's (the
    reference requires identity between its engines)."""
    from openems_amd import hdf5_out
    from oracle.operator_ref import RefOperator
    from oracle.oracle import RefFieldDumpTD, run_reference
    a, b, d = 5e-2, 2e-2, 6e-2
    FDTD = openEMS(NrTS=1000, EndCriteria=0)
    FDTD.SetGaussExcite(4.5e9, 4.5e9)
    FDTD.SetBoundaryCond(["MUR", "PML_8", "PMC", "PEC", "PEC", "PEC"])
    CSX = ContinuousStructure()
    FDTD.SetCSX(CSX)
    mesh = CSX.GetGrid()
    mesh.SetDeltaUnit(1)
    x, y, z = np.linspace(0, a, 27), np.linspace(0, b, 11), np.linspace(0, d, 33)
    for k, l in zip("xyz", (x, y, z)):
        mesh.SetLines(k, l)
    CSX.AddExcitation("excite1", 0, [1, 1, 1]).AddCurve([[x[17], x[18]], [y[6], y[7]], [z[21], z[22]]])
    pt = [x[8], y[2], z[10]]
    CSX.AddProbe("E_probe", 2).AddBox(pt, pt)
    CSX.AddProbe("H_probe", 3).AddBox(pt, pt)
    CSX.AddMaterial("RO4350B", epsilon=3.66).AddBox([x[2], y[2], z[2]], [x[4], y[3], z[5]], priority=100)
    lo, hi = [x[0], y[0], z[0]], [x[-1], y[-1], z[-1]]
    CSX.AddDump("Et", dump_type=0, dump_mode=0, file_type=1).AddBox(lo, hi)
    CSX.AddDump("Ht", dump_type=1, dump_mode=0, file_type=1).AddBox(lo, hi)
    print("case enginetests cavity 27x11x33, 1000 steps", flush=True)
    scene = FDTD.scene()
    assert FDTD.Run(str(tmp_path)) == 0
    ref = RefOperator(scene).build()
    assert ref.vexc.dir.size > 0
    dumps = [RefFieldDumpTD(ref, k, lo, hi, interp=0) for k in (0, 1)]
    run_reference(ref, 1000, dumps=dumps)
    for name, rd in zip(("Et", "Ht"), dumps):
        path = str(tmp_path / f"{name}.h5")
        names = hdf5_out.list_group(path, "/FieldData/TD")
        assert [int(nm) for nm in names] == [sm[0] for sm in rd.samples]
        assert len(names) > 100
        for nm, (ts, t, want) in zip(names, rd.samples):
            got, attrs = hdf5_out.read_dataset(path, "/FieldData/TD/" + nm)
            assert np.array_equal(got, np.transpose(want, (0, 3, 2, 1))), (name, nm)
        assert np.abs(rd.samples[-1][2]).max() > 0


def _fft_time2freq(t, val):
    """Single-sided spectrum of a sampled signal as the reference's test
    helper computes it (matlab/FFT_time2freq.m): zero-padded FFT to the next
    power of two, times the sample spacing, doubled, phase referred to t[0]."""
    dt = t[1] - t[0]
    nfft = 1 << int(np.ceil(np.log2(len(val))))
    v = np.fft.fft(val, nfft) * dt
    f = 1 / (2 * dt) * np.linspace(0, 1, nfft // 2 + 1)
    v = 2 * v[:nfft // 2 + 1]
    return f, v * np.exp(-1j * 2 * np.pi * f * t[0])


def _check_frequency(f, val, f_upper, f_lower, rel_amplitude, kind):
    """TESTSUITE/helperscripts/check_frequency.m: the largest value between
    each pair of limits (nearest samples) is at least (inside) or at most
    (outside) rel_amplitude times the overall largest value."""
    m = val.max()
    for lo, hi in zip(f_lower, f_upper):
        i1, i2 = int(np.argmin(np.abs(f - lo))), int(np.argmin(np.abs(f - hi)))
        w = val[i1:i2 + 1].max()
        if kind == "inside" and w < m * rel_amplitude:
            return False
        if kind == "outside" and w > m * rel_amplitude:
            return False
    return True


def test_cavity_resonances_against_closed_form(gpu, tmp_path):
    """The reference's physics test TESTSUITE/combinedtests/cavity.m on the
    engine: a 5 x 2 x 6 cm PEC cavity excited by a [1 1 1] curve source,
    20000 steps; the spectra of three voltage probes peak at the analytic
    TE101/102/201/202 resonances within +-0.13 % (at 60 % of the maximum) and
    TM110/111 within -0.25 % / +0 % (27 %), and stay below 17 % of the maximum
    between them (2 % guard bands)."""
    c0 = 299792458.0
    a, b, d = 5e-2, 2e-2, 6e-2
    f_start, f_stop = 1e9, 10e9
    FDTD = openEMS(NrTS=20000, EndCriteria=1e-6)
    FDTD.SetGaussExcite((f_stop - f_start) / 2, (f_stop - f_start) / 2)
    FDTD.SetBoundaryCond([0] * 6)
    CSX = ContinuousStructure()
    FDTD.SetCSX(CSX)
    mesh = CSX.GetGrid()
    mesh.SetDeltaUnit(1)
    x, y, z = np.linspace(0, a, 26), np.linspace(0, b, 11), np.linspace(0, d, 32)
    for k, l in zip("xyz", (x, y, z)):
        mesh.SetLines(k, l)
    CSX.AddExcitation("excite1", 0, [1, 1, 1]).AddCurve([[x[16], x[17]], [y[6], y[7]], [z[20], z[21]]])
    CSX.AddProbe("ut1x", 0).AddBox([x[5], y[4], z[5]], [x[6], y[4], z[5]])
    CSX.AddProbe("ut1y", 0).AddBox([x[5], y[4], z[5]], [x[5], y[5], z[5]])
    CSX.AddProbe("ut1z", 0).AddBox([x[12], y[4], z[5]], [x[12], y[4], z[6]])
    print("case combinedtests cavity, 20000 steps", flush=True)
    assert FDTD.Run(str(tmp_path)) == 0
    spectra = []
    for name in ("ut1x", "ut1y", "ut1z"):
        td = np.loadtxt(str(tmp_path / name), comments="%")
        t, v = td[:, 0], td[:, 1]
        i0 = int(np.argmin(np.abs(t - 7e-10)))
        f, u = _fft_time2freq(t[i0:], v[i0:])
        spectra.append(u)
    i1, i2 = int(np.argmin(np.abs(f - f_start))), int(np.argmin(np.abs(f - f_stop)))
    f = f[i1:i2 + 1]
    uy, uz = np.abs(spectra[1][i1:i2 + 1]), np.abs(spectra[2][i1:i2 + 1])
    k = lambda m, n, l: np.sqrt((m * np.pi / a) ** 2 + (n * np.pi / b) ** 2 + (l * np.pi / d) ** 2)
    f_TE = np.array([c0 / (2 * np.pi) * k(*mnl) for mnl in ((1, 0, 1), (1, 0, 2), (2, 0, 1), (2, 0, 2))])
    f_TM = np.array([c0 / (2 * np.pi) * k(*mnl) for mnl in ((1, 1, 0), (1, 1, 1))])

    def outer(fs):
        tmp = np.concatenate([[f_start], fs, [f_stop]])
        return tmp[:-1] * (1 + 0.02), tmp[1:] * (1 - 0.02)
    o1, o2 = outer(f_TE)
    o1m, o2m = outer(f_TM)
    assert _check_frequency(f, uy, f_TE * (1 + 1.3e-3), f_TE * (1 - 1.3e-3), 0.6, "inside")
    assert _check_frequency(f, uz, f_TM * (1 + 0), f_TM * (1 - 2.5e-3), 0.27, "inside")
    assert _check_frequency(f, uy, o2, o1, 0.17, "outside")
    assert _check_frequency(f, uz, o2m, o1m, 0.17, "outside")


def test_coax_impedance_against_closed_form(gpu, tmp_path):
    """The reference's physics test TESTSUITE/combinedtests/Coax.m on the
    engine: a 1 m coaxial line (inner radius 100 mm, outer 230 mm, 5 mm mesh,
    PML_8 at the far end) excited by a soft E source weighted x / rho^2,
    y / rho^2 in the plane z = 0; |u / i| of the voltage and current probes
    at the middle stays within -1 % / +3 % of the closed form
    Z0 = sqrt(mu0 / eps0) ln(230 / 100) / (2 pi) from 0 to 1 GHz."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_pyopenems import coax_setup
    FDTD, Z0 = coax_setup()
    print("case combinedtests Coax, 102 x 102 x 201 lines, up to 5000 steps", flush=True)
    assert FDTD.Run(str(tmp_path)) == 0
    tu = np.loadtxt(str(tmp_path / "ut1"), comments="%")
    ti = np.loadtxt(str(tmp_path / "it1"), comments="%")
    f, u = _fft_time2freq(tu[:, 0], tu[:, 1])
    _, i = _fft_time2freq(ti[:, 0], ti[:, 1])
    k1, k2 = int(np.argmin(np.abs(f - 0.0))), int(np.argmin(np.abs(f - 1e9)))
    Z = np.abs(u[k1:k2 + 1] / i[k1:k2 + 1])
    print("Z0 %.4f, |u/i| from %.4f to %.4f over %d frequencies" % (Z0, Z.min(), Z.max(), Z.size), flush=True)
    assert Z.size > 10
    assert np.all(Z <= Z0 * 1.03) and np.all(Z >= Z0 * 0.99)


def test_msl_ports_through_line(gpu, tmp_path):
    """Two microstrip ports (ports.MSLPort) on a 10 mm line of a 50 Ohm-class
    strip: the measured line impedance lies within 35..60 Ohm (Hammerstad's
    closed form gives 47 Ohm for w/h = 2.36, eps_r 3.66), beta between the
    free-space and the substrate wave numbers, and the through line matched
    (|S11| < 0.3, |S21| > 0.8) from 2 to 15 GHz."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_pyopenems import _msl_setup
    FDTD, p1, p2 = _msl_setup()
    print("case microstrip ports, 151 x 48 x 22 lines, up to 3000 steps", flush=True)
    assert FDTD.Run(str(tmp_path)) == 0
    f = np.linspace(2e9, 15e9, 27)
    p1.CalcPort(str(tmp_path), f)
    p2.CalcPort(str(tmp_path), f, ref_impedance=p1.Z_ref)
    k0 = 2 * np.pi * f / 299792458.0
    s11 = p1.uf_ref / p1.uf_inc
    s21 = p2.uf_ref / p1.uf_inc
    print("Z_ref %.1f..%.1f Ohm, beta/k0 %.3f..%.3f, max|S11| %.3f, min|S21| %.3f" % (
        np.real(p1.Z_ref).min(), np.real(p1.Z_ref).max(), (np.real(p1.beta) / k0).min(),
        (np.real(p1.beta) / k0).max(), np.abs(s11).max(), np.abs(s21).min()), flush=True)
    assert np.all((np.real(p1.Z_ref) > 35) & (np.real(p1.Z_ref) < 60))
    assert np.all((np.real(p1.beta) > k0) & (np.real(p1.beta) < k0 * np.sqrt(3.66)))
    assert np.all(np.abs(s11) < 0.3) and np.all(np.abs(s21) > 0.8)
