"""The memory-scale workload at full size (1024 x 1024 x 768 lines, about
805 M cells, 12 seeded dielectric layers, Mur on every face, sinusoidal E_x
sheet source, 16 E-point probes and a z-line voltage probe): the compressed
layout (uint16 index, 1.6 GB) and the plain layout (38.6 GB of coefficients)
run bit-identically -- same energy sums and probe samples -- and the
energy is finite and non-zero.  Parity with the oracle on the same scene at
40 x 36 x 32 lines is covered by test_engine_gpu.py."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd.engine import EngineHIP
from openems_amd.operator import Operator
from openems_amd.processing import Probe

pytestmark = pytest.mark.gpu

STEPS = 40


def _run(op, scene, layout):
    eng = EngineHIP(op, layout=layout)
    assert eng.layout()[0] == layout
    probes = [Probe(op, p, scene["nrts"], scene["oversampling"]) for p in scene["probes"]]
    for p in probes:
        p.register(eng)
    eng.IterateTS(STEPS)
    e = eng.energy()
    tds = [p.collect(eng)[0] for p in probes]
    u = eng.layout()[1]
    eng.close()
    return e, tds, u


def test_memory_scale_layouts_agree(gpu):
    scene = scenes.mur_layers()
    print("case mur_layers 1024x1024x768: operator build", flush=True)
    op = Operator(scene)
    assert op.num_cells == 1024 * 1024 * 768
    print("case mur_layers 1024x1024x768: compressed layout", STEPS, "steps", flush=True)
    ec, tc, u = _run(op, scene, "compressed")
    assert 0 < u < 1000
    print("case mur_layers 1024x1024x768: plain layout", STEPS, "steps", flush=True)
    ep, tp, _ = _run(op, scene, "plain")
    op.close()
    assert np.isfinite(ec) and ec > 0
    assert ec == ep
    # the probes sit 284 cells from the source plane: after 40 steps they
    # still read zero, identically in both layouts
    for a, b in zip(tc, tp):
        assert a.shape == b.shape and np.array_equal(a, b)
