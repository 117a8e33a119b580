"""Quad half steps (k_volt4 / k_curr4: four x cells per lane, 16-byte
accesses) against the one-cell-per-lane kernels and the oracle's restated
reference engine.  Bar: bit-identical fields.

The scenes cover every branch of the quad kernels: the x-1 / x+1 neighbour
quads at the first and last quad of a row, UPML slabs whose x boundaries fall
inside a quad (6-cell UPML: quads holding cells of two slabs, or of a slab and
the interior), the hook-only UPML cells of the current update on the last
line of each direction (random initial fields), Mur faces, Lorentz / RLC
hooks around the quad half steps, both coefficient layouts and both UPML
layouts (plain and compressed), and a mesh whose x lines are not a multiple
of 4 (quads unavailable: automatic mode keeps one cell per lane, a required
quad mode fails loudly)."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd._native import NativeError
from openems_amd.engine import EngineHIP
from openems_amd.operator import Operator
from oracle.operator_ref import RefOperator
from oracle.oracle import RefEngine

pytestmark = pytest.mark.gpu


def _oracle_ade(ref):
    """The oracle operator's Lorentz orders and RLC table in the engine's form
    (None / empty when the scene has none)."""
    lor = []
    for o in getattr(ref, "lorentz", None) or []:
        if o["pos"].shape[1] == 0 or not (o["volt"] or o["curr"]):
            continue
        d = dict(o)
        d["v_lor"] = o["v_lor"] if o["volt_lor"] else None
        d["i_lor"] = o["i_lor"] if o["curr_lor"] else None
        lor.append(d)
    return lor, getattr(ref, "rlc", None)


def _run(op, nrts, layout, quad, init=None):
    eng = EngineHIP(op, layout=layout, fused=False, quad=quad)
    try:
        q, _ = eng.quad()
        assert q == (1 if quad else 0)
        if init is not None:
            eng.set_volt(init[0])
            eng.set_curr(init[1])
        eng.IterateTS(nrts)
        return eng.get_volt(), eng.get_curr(), eng.layout()
    finally:
        eng.close()


CASES = [
    ("pec_cavity_32", lambda: scenes.pec_cavity(32, 120)),
    ("upml_bricks_48_pml8", lambda: scenes.upml_bricks(48, 150, nbricks=8, dump=False)),
    ("upml_bricks_52_pml6", lambda: scenes.upml_bricks(52, 150, nbricks=8, pml=6, dump=False)),
    ("mur_layers_40", lambda: scenes.mur_layers(40, 36, 32, 200)),
    ("dispersive_rlc_40", lambda: scenes.dispersive_rlc(40, 200)),
]


@pytest.mark.parametrize("layout", ["plain", "compressed"])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_quad_matches_oracle(gpu, name, make, layout):
    scene = make()
    nrts = scene["nrts"]
    print("case", name, "layout", layout, "steps", nrts, flush=True)
    op = Operator(scene)
    v, c, lay = _run(op, nrts, layout, True)
    assert lay[0] == layout
    if op.upml_slabs():
        assert lay[2] == layout
    ref_op = RefOperator(scene).build()
    lor, rlc = _oracle_ade(ref_op)
    ref = RefEngine(ref_op, "basic", lorentz=lor, rlc=rlc)
    ref.iterate(nrts)
    rv, rc = ref.raw_views()
    assert np.max(np.abs(rv)) > 0
    assert np.array_equal(v, rv), f"voltage mismatches: {np.count_nonzero(v != rv)}"
    assert np.array_equal(c, rc), f"current mismatches: {np.count_nonzero(c != rc)}"
    ref.close()


@pytest.mark.parametrize("layout", ["plain", "compressed"])
@pytest.mark.parametrize("pml", [6, 8])
def test_quad_random_initial_fields(gpu, pml, layout):
    """Random fields everywhere, including the last line of every direction
    (hook-only UPML cells, cells the main update skips): quads and one cell
    per lane agree bit for bit, and both match the oracle."""
    scene = scenes.upml_bricks(44, 30, nbricks=6, pml=pml, dump=False)
    print("case upml_bricks_44 random fields, pml", pml, "layout", layout, flush=True)
    op = Operator(scene)
    ref = RefEngine(RefOperator(scene).build(), "basic")
    rv, rc = ref.raw_views()
    rng = np.random.default_rng(17 + pml)
    v0 = rng.standard_normal(rv.shape).astype(np.float32)
    c0 = rng.standard_normal(rc.shape).astype(np.float32)
    rv[...] = v0
    rc[...] = c0
    vq, cq, _ = _run(op, 30, layout, True, (v0, c0))
    v1, c1, _ = _run(op, 30, layout, False, (v0, c0))
    ref.iterate(30)
    assert np.array_equal(vq, v1) and np.array_equal(cq, c1)
    assert np.array_equal(vq, rv), f"voltage mismatches {np.count_nonzero(vq != rv)}"
    assert np.array_equal(cq, rc), f"current mismatches {np.count_nonzero(cq != rc)}"
    ref.close()


def test_quad_unavailable_for_odd_rows(gpu):
    """x lines not a multiple of 4: automatic mode runs one cell per lane, a
    required quad mode is refused."""
    scene = scenes.pec_cavity(30, 10)
    op = Operator(scene)
    eng = EngineHIP(op, fused=False)
    assert eng.quad()[0] == 0
    eng.close()
    with pytest.raises(NativeError, match="quad"):
        EngineHIP(op, fused=False, quad=True)


def test_quad_default_policy(gpu):
    """Automatic mode: quads on meshes without UPML, one cell per lane on
    UPML meshes and where the x lines are not a multiple of 4."""
    for make, want in ((lambda: scenes.mur_layers(40, 36, 32, 10), 1),
                       (lambda: scenes.pec_cavity(32, 10), 1),
                       (lambda: scenes.upml_bricks(48, 10, nbricks=4, dump=False), 0),
                       (lambda: scenes.upml_bricks(46, 10, nbricks=4, dump=False), 0)):
        op = Operator(make())
        eng = EngineHIP(op)
        assert eng.quad()[0] == want
        eng.close()


def test_quad_on_headline_mesh(gpu):
    """The 256^3 UPML workload with quads required: 40 steps bit-identical
    to one cell per lane."""
    scene = scenes.upml_bricks(256, 40, dump=False)
    print("case upml_bricks_256 quad vs one cell per lane, 40 steps", flush=True)
    op = Operator(scene)
    vq, cq, _ = _run(op, 40, "auto", True)
    v1, c1, _ = _run(op, 40, "auto", False)
    assert np.max(np.abs(v1)) > 0
    assert np.array_equal(vq, v1) and np.array_equal(cq, c1)


# ---- quad column sweep (single-sweep step form 3, k_step4)

SWEEP_CASES = [
    ("pec_cavity_32", lambda: scenes.pec_cavity(32, 120)),
    ("pec_cavity_128_tile_edge_source", lambda: scenes.pec_cavity(128, 60)),
    ("wire_antenna_36_pec", lambda: dict(scenes.wire_antenna(36, 120), bc=[0] * 6)),
    # Mur on every face, a sheet source across the whole mesh (on the face and
    # neighbour lines of the x and y faces too), 300 steps (past the Mur
    # start step), and Mur next to PEC and PMC faces
    ("mur_layers_40", lambda: scenes.mur_layers(40, 36, 32, 300)),
    ("mur_pec_pmc_mix", lambda: dict(scenes.mur_layers(36, 28, 44, 250), bc=[2, 0, 1, 2, 2, 0])),
    # 66 quads per row and 22 rows: two 64 x 4 tiles along x (the second with
    # two quads) and a last row of tiles with two rows
    ("mur_layers_264x22x18", lambda: scenes.mur_layers(264, 22, 18, 200)),
]


@pytest.mark.parametrize("tile", ["0", "1"])
@pytest.mark.parametrize("layout", ["plain", "compressed"])
@pytest.mark.parametrize("name,make", SWEEP_CASES, ids=[c[0] for c in SWEEP_CASES])
def test_quad_sweep_matches_oracle(gpu, name, make, layout, tile, monkeypatch):
    """Fields after the steps of the quad column sweep equal the oracle's bit
    for bit (sources at the mesh centre, on quad and tile boundaries, and on
    curves), one column per lane (k_step4) and in 64 x 4 quad tiles
    (k_step4t; rows of fewer than 64 quads leave lanes idle)."""
    monkeypatch.setenv("EMS_Q4TILE", tile)
    scene = make()
    nrts = scene["nrts"]
    print("case quad sweep", name, "layout", layout, "tile", tile, "steps", nrts, flush=True)
    op = Operator(scene)
    eng = EngineHIP(op, layout=layout, fused="quads")
    assert eng.fused()[0] == 3
    eng.IterateTS(nrts)
    v, c = eng.get_volt(), eng.get_curr()
    eng.close()
    ref = RefEngine(RefOperator(scene).build(), "basic")
    ref.iterate(nrts)
    rv, rc = ref.raw_views()
    assert np.max(np.abs(rv)) > 0
    assert np.array_equal(v, rv), f"voltage mismatches: {np.count_nonzero(v != rv)}"
    assert np.array_equal(c, rc), f"current mismatches: {np.count_nonzero(c != rc)}"
    ref.close()


@pytest.mark.parametrize("tile", ["0", "1"])
def test_quad_sweep_random_initial_fields(gpu, tile, monkeypatch):
    """Random fields everywhere: the last line of each direction keeps its
    currents, every other cell matches the oracle bit for bit."""
    monkeypatch.setenv("EMS_Q4TILE", tile)
    scene = scenes.pec_cavity(40, 30)
    op = Operator(scene)
    ref = RefEngine(RefOperator(scene).build(), "basic")
    rv, rc = ref.raw_views()
    rng = np.random.default_rng(23)
    v0 = rng.standard_normal(rv.shape).astype(np.float32)
    c0 = rng.standard_normal(rc.shape).astype(np.float32)
    rv[...] = v0
    rc[...] = c0
    eng = EngineHIP(op, fused="quads")
    eng.set_volt(v0)
    eng.set_curr(c0)
    eng.IterateTS(31)          # odd: the result sits in the second buffers
    v, c = eng.get_volt(), eng.get_curr()
    eng.close()
    ref.iterate(31)
    assert np.array_equal(c[:, -1, :, :], c0[:, -1, :, :])
    assert np.array_equal(v, rv) and np.array_equal(c, rc)
    ref.close()


def test_quad_sweep_refused_where_unavailable(gpu):
    for make, why in ((lambda: scenes.upml_bricks(40, 10, nbricks=4, dump=False), "UPML"),
                      (lambda: scenes.pec_cavity(30, 10), "multiple of 4"),
                      (lambda: scenes.dispersive_rlc(40, 10), "fused form")):
        with pytest.raises(NativeError, match=why):
            EngineHIP(Operator(make()), fused="quads")


@pytest.mark.parametrize("pml", [6, 8])
def test_quad_split_boxes_match_single_box(gpu, pml, monkeypatch):
    """A UPML mesh split into the box without UPML cells (run without UPML
    code) and the boxes around it gives the same fields as one box over the
    whole mesh and as one cell per lane, from random initial fields (6-cell
    UPML: the box's x faces fall inside quads, whose other cells go to the
    outer boxes; the outer boxes at one plane and at 3 planes per
    workgroup)."""
    scene = scenes.upml_bricks(52, 25, nbricks=6, pml=pml, dump=False)
    op = Operator(scene)
    rng = np.random.default_rng(31 + pml)
    shape = (3,) + tuple(op.N)
    init = (rng.standard_normal(shape).astype(np.float32), rng.standard_normal(shape).astype(np.float32))
    out = []
    for split, zout, quad in (("1", "1", True), ("1", "3", True), ("0", "1", True), ("1", "1", False)):
        monkeypatch.setenv("EMS_QSPLIT", split)
        monkeypatch.setenv("EMS_QZOUT", zout)
        out.append(_run(op, 25, "compressed", quad, init))
    for v, c, _ in out[1:]:
        assert np.array_equal(out[0][0], v) and np.array_equal(out[0][1], c)


@pytest.mark.parametrize("chunk", [8, 7])
def test_tiled_sweep_graph_chunks_match_direct_launches(gpu, chunk):
    """The automatic tiled quad sweep with Mur faces (x-face hooks inside the
    sweep), E-point and voltage probes: timestep chunks replayed as HIP graphs
    give the same fields and probe series as direct launches, over several
    iterate calls (odd and even step counts, so both buffer parities are
    captured)."""
    from openems_amd.simulation import Simulation
    scene = scenes.mur_layers(264, 22, 18, 420)
    scene["steady_state"] = False
    print("case mur_layers_264x22x18 sweep graphs, chunk", chunk, flush=True)
    ref = Simulation(scene, graph=0)
    sim = Simulation(scene, graph=chunk)
    assert sim.eng.fused()[0] == 3 and ref.eng.fused()[0] == 3
    for k in (37, 100, 83, 200):
        ref.eng.IterateTS(k)
        sim.eng.IterateTS(k)
        assert np.array_equal(sim.eng.get_volt(), ref.eng.get_volt())
        assert np.array_equal(sim.eng.get_curr(), ref.eng.get_curr())
    assert sim.eng.graph()[2] > 0
    a, b = sim.results(), ref.results()
    for name, pr in b["probes"].items():
        assert np.array_equal(a["probes"][name]["td"], pr["td"])
    assert max(np.abs(pr["td"][:, 1:]).max() for pr in b["probes"].values()) > 0
    sim.close()
    ref.close()
