"""The product's host operator builder (C++, include/ems_operator.h) against the
oracle's numpy restatement: every float32 coefficient, EC value, timestep,
excitation list, UPML slab and Mur coefficient must be bit-identical.  CPU only."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd.operator import Operator
from oracle.operator_ref import RefOperator


def _compare(scene):
    op = Operator(scene, nthreads=4)
    ref = RefOperator(scene).build()
    assert op.N == ref.N
    assert op.dT == ref.dT
    assert op.nyquist == ref.nyquist
    for name, arr in zip("CGLR", ref.EC):
        assert np.array_equal(op.ec(name), arr), f"EC {name}"
    for name in ("vv", "vi", "ii", "iv"):
        got, want = op.coeff(name), getattr(ref, name)
        assert np.array_equal(got, want), f"{name}: {np.count_nonzero(got != want)} mismatches"
    sv, sc = op.signal()
    assert np.array_equal(sv, ref.sig_volt) and np.array_equal(sc, ref.sig_curr)
    for curr, lst in ((False, ref.vexc), (True, ref.cexc)):
        e = op.excitation(curr)
        assert np.array_equal(e["index"], lst.index)
        assert np.array_equal(e["dir"], lst.dir)
        assert np.array_equal(e["amp"], lst.amp)
        assert np.array_equal(e["delay"], lst.delay)
    slabs = op.upml_slabs()
    assert len(slabs) == len(ref.upml)
    for a, b in zip(slabs, ref.upml):
        assert tuple(a["start"]) == tuple(b.start) and tuple(a["num"]) == tuple(b.num)
        for k in ("vv", "vvfo", "vvfn", "ii", "iifo", "iifn"):
            assert np.array_equal(a[k], getattr(b, k)), f"UPML {k}"
    faces = op.mur_faces()
    assert len(faces) == len(ref.mur)
    for a, b in zip(faces, ref.mur):
        assert (a["ny"], a["line"], a["line_shift"], a["start_ts"]) == (b.ny, b.line, b.line_shift, b.start_ts)
        assert np.array_equal(a["coeff_nyP"], b.coeff_nyP) and np.array_equal(a["coeff_nyPP"], b.coeff_nyPP)
    return op, ref


def test_pec_cavity(native):
    _compare(scenes.pec_cavity(24, 100))


def test_upml_bricks_lumped_port(native):
    op, ref = _compare(scenes.upml_bricks(40, 300, nbricks=8))
    assert len(op.upml_slabs()) == 6
    assert op.excitation()["dir"].size == 16          # 2x2 edge columns x 4 cells of the port box


def test_mur_layers(native):
    op, _ = _compare(scenes.mur_layers(40, 36, 32, 100))
    assert len(op.mur_faces()) == 6


def test_mixed_boundaries_and_materials(native):
    """TESTSUITE/enginetests/cavity.m boundary mix (MUR, PML_8, PMC, PEC, PEC, PEC),
    an anisotropic lossy brick, a magnetic-loss brick, a metal plate and a hard source."""
    l = [np.arange(27) * 1.0, np.arange(11) * 1.5, np.arange(33) * 0.75]
    s = {"lines": l, "delta_unit": 1e-3, "bc": [2, 3, 1, 0, 0, 0], "pml_size": [8] * 6,
         "materials": [{"eps": (3.66, 2.0, 4.0), "kappa": (0.01, 0.0, 0.2), "start": (3, 1.5, 3.75),
                        "stop": (10, 12, 15), "priority": 2},
                       {"eps": 2.0, "mue": 1.5, "sigma": 30.0, "start": (12, 3, 6), "stop": (16, 9, 12), "priority": 1}],
         "metals": [{"start": (4, 4.5, 18), "stop": (8, 9, 18), "priority": 5}],
         "excitations": [{"type": 1, "val": (0, 1, 1), "start": (6, 7.5, 9), "stop": (6, 9, 9.75)},
                         {"type": 2, "val": (1, 0, 0), "delay": 3e-12, "start": (14, 3.7, 14.6), "stop": (15, 5.3, 15.4)}],
         "signal": {"type": "gauss", "f0": 10e9, "fc": 8e9}, "nrts": 400}
    op, ref = _compare(s)
    assert op.excitation(True)["dir"].size > 0


@pytest.mark.parametrize("extra", [{"ts_method": 1}, {"cell_constant_material": True},
                                   {"ts_factor": 0.8}, {"signal": {"type": "sinus", "f0": 12e9}},
                                   {"signal": {"type": "step"}}])
def test_operator_options(native, extra):
    s = scenes.upml_bricks(32, 200, nbricks=5)
    s.update(extra)
    _compare(s)


def test_graded_full_size_256(native):
    """The full 256^3 UPML workload's operator (bit-identical)."""
    _compare(scenes.upml_bricks(256, 10000, dump=False))


def test_snapping_matches_oracle(native):
    s = scenes.upml_bricks(40, 100, nbricks=3)
    op = Operator(s)
    ref = RefOperator(s)
    rng = np.random.default_rng(7)
    lo = [l[0] - 1 for l in s["lines"]]
    hi = [l[-1] + 1 for l in s["lines"]]
    for _ in range(200):
        c = [rng.uniform(lo[n], hi[n]) for n in range(3)]
        for dual in (False, True):
            assert op.snap_to_mesh(c, dual) == ref.mesh.snap(c, dual)


def test_material_grid_matches_one_cell_boxes(native):
    """A material grid (per-cell eps_r / kappa) builds the same operator as the
    oracle's expansion into one-cell boxes of equal priority, and its
    quarter-cell weights reproduce C and dC/d eps_r."""
    s, _, _ = scenes.adjoint_design(30, 28, 26, pml=6, design=(6, 5, 3), nrts=100)
    g = s["materials"][0]["grid"]
    g["kappa"] = np.random.default_rng(3).uniform(0, 0.05, size=g["eps"].shape)
    op, ref = _compare(s)
    import ctypes as C
    from openems_amd._native import lib
    cells = (C.c_uint * 12)()
    w = (C.c_double * 4)()
    c0, nc = g["cell0"], g["ncells"]
    eps = g["eps"]
    Cc = op.ec("C")
    for n, pos in [(0, (c0[0] + 2, c0[1] + 2, c0[2] + 1)), (1, (c0[0] + 1, c0[1] + 3, c0[2] + 2)),
                   (2, (c0[0] + 3, c0[1] + 1, c0[2] + 1))]:
        assert lib().emsop_eps_weights(op.h, n, (C.c_uint * 3)(*pos), cells, w) == 0
        total = 0.0
        for q in range(4):
            c = [cells[3 * q + a] for a in range(3)]
            r = [c[a] - c0[a] for a in range(3)]
            assert all(0 <= r[a] < nc[a] for a in range(3))
            total += w[q] * eps[r[2], r[1], r[0]]
        assert total == pytest.approx(float(Cc[n][pos]), rel=1e-6)
    # derivative: raising one cell's eps_r by 1 changes C by its weight
    n, pos = 0, (c0[0] + 2, c0[1] + 2, c0[2] + 1)
    lib().emsop_eps_weights(op.h, n, (C.c_uint * 3)(*pos), cells, w)
    c = [cells[a] for a in range(3)]
    s2 = scenes.adjoint_design(30, 28, 26, pml=6, design=(6, 5, 3), nrts=100)[0]
    s2["materials"][0]["grid"]["eps"] = eps.copy()
    s2["materials"][0]["grid"]["eps"][c[2] - c0[2], c[1] - c0[1], c[0] - c0[0]] += 1.0
    s2["materials"][0]["grid"]["kappa"] = g["kappa"]
    op2 = Operator(s2)
    dCn = float(op2.ec("C")[n][pos]) - float(Cc[n][pos])
    assert dCn == pytest.approx(w[0], rel=1e-5)
    op2.close()
    op.close()


def test_lorentz_and_lumped_rlc_tables(native):
    """Operator_Ext_LorentzMaterial and Operator_Ext_LumpedRLC: the product's
    tables and the operator they rewrite equal the oracle's, bit for bit
    (Drude box reaching into the UPML, two-order Lorentz box with a mue pole,
    series RLC, parallel RLC on the edges' own capacitance, LC with R = 0)."""
    scene = scenes.dispersive_rlc()
    op, ref = _compare(scene)
    lt = op.lorentz_tables()
    assert len(lt) == len(ref.lorentz) == 2
    for a, b in zip(lt, ref.lorentz):
        assert np.array_equal(a["pos"], b["pos"])
        assert (a["volt"], a["curr"], a["v_lor"] is not None, a["i_lor"] is not None) == \
            (b["volt"], b["curr"], b["volt_lor"], b["curr_lor"])
        for k in ("v_int", "v_ext", "i_int", "i_ext", "v_lor", "i_lor"):
            if a[k] is not None:
                assert np.array_equal(a[k], b[k]), k
    assert lt[0]["pos"].shape[1] > 0 and lt[1]["curr"]
    rt = op.rlc_table()
    assert np.array_equal(rt["pos"], ref.rlc["pos"]) and np.array_equal(rt["dir"], ref.rlc["dir"])
    for k in ("i2v", "ilv", "vvd", "vv2", "vj1", "vj2", "ib0", "b1", "b2"):
        assert np.array_equal(rt[k], ref.rlc[k]), k
    # series: 4 z edges x 2 x 2 columns; parallel: 4 z edges x 2 x 1; LC: 2 y edges
    assert rt["dir"].size == 16 + 8 + 2
    assert rt["dir"].tolist() == [2] * 24 + [1] * 2


def test_wire_paths_match_oracle(native):
    """Operator::FindPath: the product's mesh-edge paths equal the oracle's
    for random segments (inside, partly outside and wholly outside a graded
    mesh), edge for edge."""
    scene = scenes.wire_antenna(30, 10)
    op = Operator(scene, nthreads=2)
    ref = RefOperator(scene)
    L = scene["lines"]
    rng = np.random.default_rng(11)
    lengths = []
    for _ in range(400):
        a = [rng.uniform(l[0] - 3, l[-1] + 3) for l in L]
        b = [rng.uniform(l[0] - 3, l[-1] + 3) for l in L]
        got = op.find_path(a, b)
        assert got == ref.find_path(a, b), (a, b)
        lengths.append(len(got))
    assert max(lengths) > 20 and min(lengths) == 0


def test_wire_curves_metal_and_excitation(native):
    """Metal polylines (CalcPEC_Curves) and soft / hard excitation curves
    (Operator_Ext_Excitation): coefficients and excitation lists bit-identical."""
    scene = scenes.wire_antenna(36, 100)
    op, ref = _compare(scene)
    e = op.excitation(False)
    assert e["dir"].size == 14
    # the metal wire's edges carry no voltage update
    arm = ref.curve_edges(ref.metal_curves[0])
    assert len(arm) > 10
    vv, vi = op.coeff("vv"), op.coeff("vi")
    for n, pos in arm:
        assert vv[(n,) + pos] == 0 and vi[(n,) + pos] == 0


@pytest.mark.parametrize("field,interp", [("E", 0), ("E", 1), ("E", 2), ("H", 0), ("H", 1), ("H", 2)])
def test_field_dump_boxes_match_oracle(native, field, interp):
    """ProcessFields box snapping (SnapMethod 1, dual mesh for H without
    interpolation and for cell interpolation): the product's dump box equals
    the oracle's for random boxes on a graded mesh."""
    from openems_amd.processing import FieldDumpFD
    from oracle.oracle import RefFieldDumpFD
    scene = scenes.wire_antenna(30, 50)
    op = Operator(scene, nthreads=2)
    ref = RefOperator(scene).build()
    L = scene["lines"]
    rng = np.random.default_rng(5)
    for _ in range(40):
        a = [rng.uniform(l[0] - 1, l[-1] + 1) for l in L]
        b = [rng.uniform(l[0] - 1, l[-1] + 1) for l in L]
        if rng.random() < 0.3:
            b[2] = a[2]
        spec = {"start": a, "stop": b, "freqs": [5e9], "field": field, "interp": interp}
        d = FieldDumpFD(op, spec, 50)
        r = RefFieldDumpFD(ref, 0 if field == "E" else 1, a, b, [5e9], interp=interp)
        assert d.enabled == r.enabled
        if not r.enabled:
            assert d.fd_interval == 0
            continue
        assert list(d.start) == [int(p[0]) for p in r.pos]
        assert list(d.stop) == [int(p[-1]) for p in r.pos]
        assert d.dual_mesh == r.dual_mesh


@pytest.mark.parametrize("ptype", [0, 1, 2, 3])
def test_probe_boxes_match_oracle(native, ptype):
    """Probe snapping: voltage lines (SnapToMesh), current surfaces (dual
    mesh, SnapMethod 1, normal-direction expansion) and field points: the
    product's probe box and enable state equal the oracle's for random
    off-line boxes."""
    from openems_amd.processing import Probe
    from oracle.oracle import RefProbe
    scene = scenes.wire_antenna(30, 50)
    op = Operator(scene, nthreads=2)
    ref = RefOperator(scene).build()
    L = scene["lines"]
    rng = np.random.default_rng(17 + ptype)
    seen = 0
    for _ in range(60):
        a = [rng.uniform(l[0], l[-1]) for l in L]
        b = [rng.uniform(l[0], l[-1]) for l in L]
        nd = -1
        if ptype == 1:
            nd = int(rng.integers(-1, 3))
            k = nd if nd >= 0 else int(rng.integers(0, 3))
            b[k] = a[k]
        elif ptype == 0:
            k = int(rng.integers(0, 3))
            b = list(a)
            b[k] = rng.uniform(L[k][0], L[k][-1])
        spec = {"type": ptype, "start": a, "stop": b, "weight": 1.0, "freqs": [], "norm_dir": nd}
        p = Probe(op, spec, 50)
        r = RefProbe(ref, ptype, a, b, 1.0, [], nd)
        assert p.enabled == r.enabled
        if r.enabled:
            seen += 1
            assert list(p.start) == list(r.start) and list(p.stop) == list(r.stop)
    assert seen > 10


def test_current_probe_box_widens_with_snap_method_1(native):
    """An off-line current box widens to the next dual line outside on each
    non-flat axis (SnapMethod 1 of ProcessCurrent)."""
    from oracle.oracle import RefProbe
    l = np.arange(16.0)
    sc = {"lines": [l, l.copy(), l.copy()], "delta_unit": 1e-3,
          "signal": {"type": "gauss", "f0": 0.0, "fc": 10e9}, "nrts": 10}
    ref = RefOperator(sc).build()
    # dual lines at i + 0.5; x 3.7..9.2 snaps to 3 (3.5) .. 8 (8.5), both inside -> widened to 3 .. 9
    r = RefProbe(ref, 1, (3.7, 4.6, 7.5), (9.2, 10.4, 7.5), 1.0, [], 2)
    assert r.enabled and r.start[:2] == [3, 4] and r.stop[:2] == [9, 10]


@pytest.mark.parametrize("function", [
    "sin(2*pi*3e9*t)*exp(-((t-1e-9)/3e-10)^2)",
    "if(t<2e-10, t/2e-10, 1) * cos(2*pi*5e9*t)",
    "-t^2*1e18 + 3*t*1e9 - sqrt(abs(sin(t*1e10)))",
    "max(0, min(1, t*4e9)) * (t>1e-10 & t<9e-10)",
    "2^-1 + atan2(t*1e9, 1) - floor(t*7e9)/ceil(t*3e9+0.5) + int(t*2.5e9) + log10(1+t*1e9)",
])
def test_custom_excitation_signal(native, function):
    """Custom excitation (CalcCustomExcitation): the product's parser and the
    oracle's (Python's own expression parser) give the same float32 samples."""
    sc = scenes.pec_cavity(16, 300)
    sc["signal"] = {"type": "custom", "function": function, "f0": 5e9}
    op, ref = _compare(sc)
    assert op.signal()[0].size == 300


@pytest.mark.parametrize("function,message", [("sin(2*pi*t", "expected at position"),
                                               ("t+", "unexpected end"), ("2*t-", "unexpected end"),
                                               ("t<", "unexpected end"), ("if(t,1,)", "unexpected")])
def test_custom_excitation_parse_error(native, function, message):
    """Malformed expressions, including a missing last operand, fail the build
    with the position of the error."""
    from openems_amd._native import NativeError
    sc = scenes.pec_cavity(16, 30)
    sc["signal"] = {"type": "custom", "function": function, "f0": 5e9}
    with pytest.raises(NativeError, match=message):
        Operator(sc)


def test_processing_windows_match_oracle(native):
    """Processing::SetProcessStartStopTime: probes sample only inside
    [floor(start/dT), ceil(stop/dT)], a frequency-domain dump's first sample
    is the first multiple of its interval in the window; an inverted window
    means the whole run.  Product and oracle agree step for step."""
    from openems_amd.processing import FieldDumpFD, Probe
    from oracle.oracle import RefFieldDumpFD, RefProbe
    scene = scenes.upml_bricks(40, 400, nbricks=4)
    op = Operator(scene, nthreads=2)
    ref = RefOperator(scene).build()
    dT = op.dT
    p0 = scene["probes"][0]
    for t0, t1 in ((37.2 * dT, 211.5 * dT), (0.0, 100.1 * dT), (150 * dT, 0.0), (300 * dT, 100 * dT)):
        spec = dict(p0, start_time=t0, stop_time=t1, freqs=[5e9])
        p = Probe(op, spec, 400)
        r = RefProbe(ref, p0["type"], p0["start"], p0["stop"], p0["weight"], [5e9], p0["norm_dir"],
                     start_time=t0, stop_time=t1)
        assert p.steps == [ts for ts in range(401) if r.check(ts)]
        d = scene["dumps"][0]
        fd = FieldDumpFD(op, dict(d, start_time=t0, stop_time=t1), 400)
        rd = RefFieldDumpFD(ref, 0, d["start"], d["stop"], d["freqs"], start_time=t0, stop_time=t1)
        due = [ts for ts in range(401) if rd.due(ts)]
        assert [fd.first + k * fd.fd_interval for k in range(fd.nsamples)] == due
    # closed form of the window: floor / ceil, inverted -> whole run
    r = RefProbe(ref, 2, p0["start"], p0["start"], 1.0, [], -1, start_time=10.5 * dT, stop_time=20.2 * dT)
    assert (r.start_ts, r.stop_ts) == (10, 21)


def test_tilted_cylinders_and_weighted_boxes(native):
    """Cylinders and shells on oblique axes (metal, and hard / soft excitation
    with weights in x, y, z, rho, r, a), priorities that overlap a dielectric
    box and a metal box, on a graded mesh: coefficients and excitation lists
    bit-identical to the oracle."""
    sc = scenes.wire_antenna(28, 20)
    sc["metals"] = [
        {"priority": 2, "start": (-6.0, -4.0, -3.0), "stop": (5.0, 7.0, 9.0), "radius": 2.2},
        {"priority": 1, "start": (4.0, -8.0, 0.5), "stop": (-3.0, 6.0, -2.0), "radius": 3.1, "shell_width": 1.3},
        {"priority": 0, "start": (-2.0, -2.0, -2.0), "stop": (1.0, 3.0, 2.0)},
    ]
    sc["excitations"] = [
        {"type": 0, "val": (1.0, -0.5, 0.25), "priority": 3, "start": (-7.0, 5.0, -6.0), "stop": (6.0, -5.0, 4.0),
         "radius": 1.7, "shell_width": 0.9, "weight": ["x*y + 0.5", "cos(a)*rho", None]},
        {"type": 1, "val": (0.0, 0.0, 2.0), "priority": 1, "start": (2.0, 2.0, -8.0), "stop": (2.5, 3.0, 8.0),
         "radius": 1.2, "weight": [None, None, "exp(-r/4)*(z>0)"]},
        {"type": 2, "val": (0.3, 0.3, 0.3), "priority": 0, "start": (-8.0, -8.0, 3.0), "stop": (8.0, 8.0, 5.0),
         "weight": ["1", "y/3", "sqrt(x*x+1)"]},
        {"type": 0, "val": (0.0, 0.0, 1.0), "weight": ["0", "0", "2+sin(z)"],
         "curve": [(0.3, 0.4, -3.0), (0.3, 0.4, 3.0)]},
    ]
    op, ref = _compare(sc)
    assert op.excitation(False)["dir"].size > 20 and op.excitation(True)["dir"].size > 20
    bare = dict(sc, metals=[], excitations=[])
    zeros = lambda o: np.count_nonzero(o.coeff("vv") == 0)
    assert zeros(op) > zeros(Operator(bare, nthreads=2)) + 100


def test_eval_function_checks_its_arguments(native):
    """emsop_eval_function: values at several points, NaN where the
    expression has no value, and errors for a malformed expression, an
    unknown variable and a variable list that does not match nvar."""
    from openems_amd._native import NativeError
    from openems_amd.processing import evaluate_function
    v = np.array([[0.0, 2.0], [1.0, -1.0], [2.0, 0.5]])
    got = evaluate_function("x^2 + y*pi - min(x, y) + if(x>1, 10, 0)", "x, y", v)
    want = [x ** 2 + y * np.pi - min(x, y) + (10 if x > 1 else 0) for x, y in v]
    assert np.array_equal(got, want)
    assert np.isnan(evaluate_function("sqrt(x)", "x", np.array([[-1.0]]))[0])
    for text, names, msg in (("x+", "x", "unexpected end"), ("q*2", "x", "unknown name"),
                             ("x", "x,y,z", "variable list")):
        with pytest.raises(NativeError, match=msg):
            evaluate_function(text, names, np.zeros((1, 1)))
