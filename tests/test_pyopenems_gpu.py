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
