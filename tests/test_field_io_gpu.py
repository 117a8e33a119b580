"""Box-level field access and point fields through the C ABI
(ems_read_field / ems_write_field, the reference's per-element
Engine::GetVolt / SetVolt over a box; ems_get_field_at, the reference's
Engine_Interface_FDTD::GetEField / GetHField with its three interpolation
types), against the whole-field reads and the device field dumps."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd.engine import EngineHIP
from openems_amd.operator import Operator

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fused", [False, "columns"])
def test_read_write_box(gpu, fused):
    scene = scenes.upml_bricks(40, 30, nbricks=6, dump=False)
    op = Operator(scene)
    eng = EngineHIP(op, fused=fused)
    try:
        eng.IterateTS(30)
        V, I = eng.get_volt(), eng.get_curr()
        st, sp = (3, 5, 7), (20, 11, 33)
        box = tuple(slice(a, b + 1) for a, b in zip(st, sp))
        assert np.array_equal(eng.read_field(False, st, sp), V[(slice(None),) + box])
        assert np.array_equal(eng.read_field(True, st, sp, comp=2), I[(2,) + box])
        # one cell, one component
        assert eng.read_field(False, (9, 9, 9), (9, 9, 9), comp=1)[0, 0, 0] == V[1, 9, 9, 9]
        # write a box, read the whole field back; then step on and compare
        # with an engine given the same fields whole
        rng = np.random.default_rng(4)
        vals = rng.standard_normal((3,) + tuple(b - a + 1 for a, b in zip(st, sp))).astype(np.float32)
        eng.write_field(False, st, sp, vals)
        V2 = V.copy()
        V2[(slice(None),) + box] = vals
        assert np.array_equal(eng.get_volt(), V2)
        eng.write_field(True, (1, 2, 3), (1, 2, 3), np.array([[[5.0]]], np.float32), comp=0)
        I2 = I.copy()
        I2[0, 1, 2, 3] = 5.0
        assert np.array_equal(eng.get_curr(), I2)
        eng.IterateTS(7)
        ref = EngineHIP(op, fused=fused)
        ref.IterateTS(30)
        ref.set_volt(V2)
        ref.set_curr(I2)
        ref.IterateTS(7)
        assert np.array_equal(eng.get_volt(), ref.get_volt()) and np.array_equal(eng.get_curr(), ref.get_curr())
        ref.close()
    finally:
        eng.close()


@pytest.mark.parametrize("interp", [0, 1, 2])
@pytest.mark.parametrize("field", [0, 1])
def test_field_at_matches_td_dump(gpu, field, interp):
    """Point fields equal the time-domain dump of the same box, field and
    interpolation at the same step (both in the dump's float32), including
    nodes on the mesh's first and last lines."""
    scene = scenes.upml_bricks(32, 40, nbricks=4, dump=False)
    scene["dumps"] = [{"field": "EH"[field], "domain": "td", "interp": interp,
                       "start": [scene["lines"][a][0] for a in range(3)],
                       "stop": [scene["lines"][a][-1] for a in range(3)], "name": "all"}]
    from openems_amd.simulation import Simulation
    sim = Simulation(scene)
    last = sim.dumps[0].steps[-1]          # the fields as they stand at the last sample
    assert last > 0
    sim.run(last)
    res = sim.results()["dumps"]["all"]
    ts, td = res["ts"], res["td"]
    k = len(ts) - 1
    assert ts[k] == sim.eng.numTS == last
    st, sp = sim.dumps[0].start, sim.dumps[0].stop
    checked = 0
    for pos in ((0, 0, 0), (31, 31, 31), (5, 17, 3), (16, 16, 16), (0, 9, 31), (31, 1, 2), (1, 30, 1)):
        if not all(st[a] <= pos[a] <= sp[a] for a in range(3)):
            continue
        got = sim.eng.field_at(field, pos, interp).astype(np.float32)
        want = td[k][:, pos[0] - st[0], pos[1] - st[1], pos[2] - st[2]]
        assert np.array_equal(got, want), (pos, got, want)
        checked += 1
    assert checked >= 4 and np.abs(td[k]).max() > 0
    sim.close()
