"""Batched engine (include/ems_hip.h ems_batch_*) through the C ABI: every
instance of a batch of uneven grids is bit-identical to the oracle engine
and to its own engine run alone, for the autotuned and for hand-set tile
shapes, z depths and grid sizes."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd.batch import BatchEngine
from openems_amd.simulation import Simulation
from oracle.operator_ref import RefOperator
from oracle.oracle import RefEngine

pytestmark = pytest.mark.gpu


def _small_batch(count, nrts, seed=20260204):
    return scenes.batched_sweep(count=count, nrts=nrts, seed=seed, lo=26, hi=44)


def _alone(scene, nsteps):
    sim = Simulation(scene, fused=False)
    sim.run(nsteps)
    out = (sim.eng.get_volt(), sim.eng.get_curr(), sim.results()["probes"])
    sim.close()
    return out


def _check_same(batch, ref, nsteps):
    for sim, s in zip(batch.sims, ref):
        v, c, probes = s
        assert sim.eng.numTS == nsteps
        assert np.array_equal(sim.eng.get_volt(), v) and np.array_equal(sim.eng.get_curr(), c)
        got = sim.results()["probes"]
        for name, p in probes.items():
            assert np.array_equal(got[name]["td"], p["td"])
            assert got[name]["td"].shape[0] > 0


def test_batch_instances_match_oracle(gpu):
    scs = _small_batch(5, 120)
    print("case batch of 5 uneven instances vs oracle, 120 steps", flush=True)
    b = BatchEngine(scs)
    cfg = b.config()
    assert cfg["work_items"] > 0
    b.iterate(120)
    assert b.last_ms() > 0
    for sim, s in zip(b.sims, scs):
        ref = RefEngine(RefOperator(s).build(), "basic")
        ref.iterate(120)
        rv, rc = ref.raw_views()
        assert np.max(np.abs(rv)) > 0
        assert np.array_equal(sim.eng.get_volt(), rv), s["name"]
        assert np.array_equal(sim.eng.get_curr(), rc), s["name"]
        ref.close()
    b.close()


@pytest.mark.parametrize("config", ["autotune", "tiles32_depth3_grid7", "tiles64_own_depths_full_grid",
                                    "groups_of_one_or_two", "graph_chunks_of_32", "graph_chunks_of_7",
                                    "three_streams", "two_streams_graphs"])
def test_batch_equals_engines_alone(gpu, config):
    scs = _small_batch(4, 90, seed=77)
    print("case batch of 4 vs engines alone, 90 steps,", config, flush=True)
    ref = [_alone(s, 90) for s in scs]
    b = BatchEngine(scs)
    if config == "autotune":
        ms = b.autotune(steps=3)
        assert ms > 0
    elif config == "tiles32_depth3_grid7":
        b.set_config(tile_x=32, zdepth=[3] * b.count, grid=7)
    elif config == "groups_of_one_or_two":
        b.set_grouping(3 << 20)          # a few MiB: several groups, stepped one after another
        assert b.config()["groups"] > 1
    elif config == "graph_chunks_of_32":
        assert b.graph() == 0            # direct launches by default
        b.set_graph(32)
    elif config == "three_streams":
        b.set_grouping(3 << 20)
        b.set_streams(3)
        assert b.streams() == 3 and b.config()["groups"] > 2
    elif config == "two_streams_graphs":
        b.set_grouping(3 << 20)
        b.set_streams(2)
        b.set_graph(16)
    elif config == "graph_chunks_of_7":
        b.set_grouping(3 << 20)
        b.set_graph(7)                   # 40 = 5 x 7 + 5, 50 = 7 x 7 + 1 steps per call
    else:
        b.set_config(tile_x=64, zdepth=None, grid=0)
    print("config", b.config(), "graph", b.graph(), flush=True)
    b.iterate(40)
    b.iterate(50)
    _check_same(b, ref, 90)
    b.close()


def test_batch_refuses_mixed_layouts(gpu):
    scs = _small_batch(2, 10)
    from openems_amd._native import NativeError, lib
    import ctypes as C
    sims = [Simulation(scs[0], layout="plain", fused=False), Simulation(scs[1], layout="compressed", fused=False)]
    hs = (C.c_void_p * 2)(*[s.eng.h.value for s in sims])
    b = C.c_void_p()
    assert lib().ems_batch_create(C.byref(b), hs, 2) < 0
    assert b"layout" in lib().ems_last_error()
    for s in sims:
        s.close()
