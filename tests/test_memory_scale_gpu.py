"""The memory-scale workload at full size (1024 x 1024 x 768 lines, about
805 M cells, 12 seeded dielectric layers, Mur on every face, sinusoidal E_x
sheet source, 16 E-point probes and a z-line voltage probe):

* the automatic step form (the tiled quad column sweep with the Mur hooks
  around it) and two half steps -- separate kernels -- give the same fields,
  probe series and energy over 1200 steps, long enough for the wave from the
  source plane to reach the probes 284 cells away;
* the compressed layout (uint16 index, 1.6 GB) and the plain layout (38.6 GB
  of coefficients) run bit-identically for 40 steps.

Parity with the oracle on the same scene at 40 x 36 x 32 and 264 x 22 x 18
lines, for both step forms, is covered by test_engine_gpu.py and
test_quad_gpu.py."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd.engine import EngineHIP
from openems_amd.operator import Operator
from openems_amd.processing import Probe

pytestmark = pytest.mark.gpu

STEPS = 40


def _run(op, scene, layout, steps=STEPS, fused=None, planes=()):
    eng = EngineHIP(op, layout=layout, fused=fused)
    assert eng.layout()[0] == layout
    form = eng.fused()[0]
    probes = [Probe(op, p, scene["nrts"], scene["oversampling"]) for p in scene["probes"]]
    for p in probes:
        p.register(eng)
    eng.IterateTS(steps)
    e = eng.energy()
    tds = [p.collect(eng)[0] for p in probes]
    u = eng.layout()[1]
    N = op.N
    # whole x-y planes of both fields
    cuts = [(eng.read_field(False, (0, 0, z), (N[0] - 1, N[1] - 1, z)),
             eng.read_field(True, (0, 0, z), (N[0] - 1, N[1] - 1, z))) for z in planes]
    eng.close()
    return e, tds, u, form, cuts


def test_memory_scale_sweep_matches_half_steps(gpu):
    steps = 1200
    scene = scenes.mur_layers(nrts=steps)
    print("case mur_layers 1024x1024x768 sweep vs half steps: operator build", flush=True)
    op = Operator(scene)
    planes = (0, 1, 99, 100, 101, 383, 384, 766, 767)
    print("case mur_layers 1024x1024x768: automatic form", steps, "steps", flush=True)
    ea, ta, _, form, ca = _run(op, scene, "compressed", steps, None, planes)
    assert form == 3
    print("case mur_layers 1024x1024x768: two half steps", steps, "steps", flush=True)
    eh, th, _, formh, ch = _run(op, scene, "compressed", steps, "halves", planes)
    op.close()
    assert formh == 0
    assert np.isfinite(ea) and ea > 0 and ea == eh
    # the E probes in the middle plane have seen the wave
    assert max(np.abs(t[:, 1:]).max() for t in ta[:16]) > 0
    for a, b in zip(ta, th):
        assert a.shape == b.shape and np.array_equal(a, b)
    for (va, ia), (vh, ih) in zip(ca, ch):
        assert np.array_equal(va, vh) and np.array_equal(ia, ih)
    assert np.abs(ca[4][0]).max() > 0


def test_memory_scale_layouts_agree(gpu):
    scene = scenes.mur_layers()
    print("case mur_layers 1024x1024x768: operator build", flush=True)
    op = Operator(scene)
    assert op.num_cells == 1024 * 1024 * 768
    print("case mur_layers 1024x1024x768: compressed layout", STEPS, "steps", flush=True)
    ec, tc, u, _, _ = _run(op, scene, "compressed")
    assert 0 < u < 1000
    print("case mur_layers 1024x1024x768: plain layout", STEPS, "steps", flush=True)
    ep, tp, _, _, _ = _run(op, scene, "plain")
    op.close()
    assert np.isfinite(ec) and ec > 0
    assert ec == ep
    # the probes sit 284 cells from the source plane: after 40 steps they
    # still read zero, identically in both layouts
    for a, b in zip(tc, tp):
        assert a.shape == b.shape and np.array_equal(a, b)
