"""An openEMS XML setup run end to end on the engine (openems_amd.xml_setup.
run_xml, the ``openEMS file.xml`` path): probe files in the reference layout
equal to the same scene run from its dict; the energy end criterion stops a
decaying run early."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd.simulation import Simulation
from openems_amd.xml_setup import read_xml, run_xml, write_xml

pytestmark = pytest.mark.gpu


def test_xml_run_matches_the_scene_and_writes_probe_files(gpu, tmp_path):
    from openems_amd import hdf5_out
    scene = scenes.wire_antenna(36, 400)
    scene["probes"][0]["freqs"] = [8e9, 12e9]
    scene["end_criteria"] = 1e-9            # far below what 400 steps reach
    L = scene["lines"]
    box = ((L[0][10], L[1][10], L[2][18]), (L[0][26], L[1][26], L[2][18]))
    scene["dumps"] = [{"name": "Et", "start": box[0], "stop": box[1], "field": "E", "interp": 1, "domain": "td",
                       "file_type": 1},
                      {"name": "Hf", "start": box[0], "stop": box[1], "field": "H", "interp": 2, "freqs": [9e9],
                       "file_type": 1}]
    path = tmp_path / "wire.xml"
    write_xml(scene, str(path))
    print("case wire_antenna_36 from XML, 400 steps", flush=True)
    res, files = run_xml(str(path), out_dir=str(tmp_path))
    assert sorted(f.split("/")[-1] for f in files) == ["Et.h5", "Hf.h5", "et", "feed_ut", "feed_ut_FD"]
    td = res["dumps"]["Et"]
    last, attrs = hdf5_out.read_dataset(str(tmp_path / "Et.h5"), "/FieldData/TD/%08d" % td["ts"][-1])
    assert np.array_equal(last, np.transpose(td["td"][-1], (0, 3, 2, 1))) and np.abs(last).max() > 0
    re, _ = hdf5_out.read_dataset(str(tmp_path / "Hf.h5"), "/FieldData/FD/f0_real")
    assert np.array_equal(re, np.transpose(res["dumps"]["Hf"]["fd"][0].real, (0, 3, 2, 1)))
    sim = Simulation(read_xml(str(path)))
    sim.run()
    want = sim.results()
    assert want["numTS"] == res["numTS"] == 400
    for name, p in want["probes"].items():
        assert np.array_equal(p["td"], res["probes"][name]["td"])
        back = np.loadtxt(str(tmp_path / name), comments="%")
        np.testing.assert_allclose(back, p["td"].reshape(back.shape), rtol=1e-11, atol=1e-300)
    assert np.abs(want["probes"]["feed_ut"]["td"][:, 1]).max() > 0
    sim.close()


def test_energy_end_criterion_stops_a_decaying_run(gpu):
    """The run ends at the first check (every 100 steps) where the field
    energy has fallen to 1e-3 of the largest sample, the samples being the
    checks and the step of the excitation's peak (openems.cpp:1424-1480).
    The scene has a 50 Ohm port: after the pulse the soft source keeps
    adding the signal's first sample (exp(-9) for a Gaussian,
    engine_ext_excitation.cpp:50-54), which the port's resistor dissipates."""
    scene = scenes.upml_bricks(40, 6000, nbricks=8, dump=False)
    scene["end_criteria"] = 1e-3
    scene["energy_check"] = 100
    print("case upml_bricks_40 energy end criterion 1e-3", flush=True)
    sim = Simulation(scene)
    sim.run()
    stop = sim.stopped_at
    assert stop is not None and 100 <= stop < 6000 and stop % 100 == 0
    sig_v, _ = sim.op.signal()
    max_step = int(np.argmax(np.abs(sig_v)))
    assert 0 < max_step < 100
    ref = Simulation(scene)
    ref.eng.IterateTS(max_step)
    peak = ref.eng.energy()
    ref.eng.IterateTS(100 - max_step)
    first = None
    while ref.eng.numTS <= stop:
        e = ref.eng.energy()
        peak = max(peak, e)
        if e / peak <= 1e-3:
            first = ref.eng.numTS
            break
        ref.eng.IterateTS(100)
    assert first == stop
    sim.close()
    ref.close()
