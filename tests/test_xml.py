"""openEMS XML setups (openems_amd.xml_setup): the FDTD block as
openEMS::Parse_XML_FDTDSetup reads it (openems.cpp:837-975), a structure in
CSXCAD's layout, and write -> read round trips that rebuild bit-identical
operators.  CPU only."""
import numpy as np
import pytest

from openems_amd import scenes
from openems_amd.operator import Operator
from openems_amd.xml_setup import read_xml, write_xml

SETUP = """<?xml version="1.0" encoding="UTF-8" standalone="yes" ?>
<openEMS>
  <FDTD NumberOfTimesteps="2e3" endCriteria="0" OverSampling="1" f_max="7e9" TimeStepFactor="0.9">
    <Excitation Type="0" f0="2e9" fc="5e9"/>
    <BoundaryCond xmin="PML_8" xmax="MUR" ymin="0" ymax="1" zmin="PEC" zmax="bogus" MUR_PhaseVelocity="2.5e8"
                  MUR_PhaseVelocity_xmax="2e8"/>
  </FDTD>
  <ContinuousStructure CoordSystem="0">
    <BackgroundMaterial Epsilon="1" Mue="1" Kappa="0" Sigma="0"/>
    <Properties>
      <Material ID="0" Name="substrate" Isotropy="1">
        <FillColor R="1" G="2" B="3" a="4"/>
        <Primitives>
          <Box Priority="10"><P1 X="0" Y="0" Z="0"/><P2 X="20" Y="20" Z="1.5"/></Box>
        </Primitives>
        <Property Epsilon="3.66" Mue="1" Kappa="0.001" Sigma="0" Density="0"/>
      </Material>
      <Metal ID="1" Name="wire">
        <Primitives>
          <Curve Priority="5"><Vertex X="2" Y="2" Z="2"/><Vertex X="2" Y="2" Z="8"/><Vertex X="6" Y="4" Z="8"/></Curve>
          <Box Priority="4"><P1 X="10" Y="10" Z="4"/><P2 X="12" Y="11" Z="6"/></Box>
        </Primitives>
      </Metal>
      <LumpedElement ID="2" Name="port_resist_1" Direction="2" Caps="1" R="50" C="0" L="0">
        <Primitives><Box Priority="5"><P1 X="8" Y="8" Z="3"/><P2 X="8" Y="8" Z="5"/></Box></Primitives>
      </LumpedElement>
      <Excitation ID="3" Name="port_excite_1" Number="0" Type="0" Excite="-0,-0,-1" Delay="1e-10">
        <Primitives><Box Priority="5"><P1 X="8" Y="8" Z="3"/><P2 X="8" Y="8" Z="5"/></Box></Primitives>
        <Weight X="1" Y="1" Z="1"/>
      </Excitation>
      <ProbeBox ID="4" Name="port_ut1" Number="0" NormDir="-1" Type="0" Weight="-1">
        <Primitives><Box Priority="5"><P1 X="8" Y="8" Z="3"/><P2 X="8" Y="8" Z="5"/></Box></Primitives>
      </ProbeBox>
      <ProbeBox ID="5" Name="port_it1" Number="0" NormDir="2" Type="1" Weight="1">
        <Primitives><Box Priority="5"><P1 X="8" Y="8" Z="4"/><P2 X="8" Y="8" Z="4"/></Box></Primitives>
        <FD_Samples>1e9,2e9</FD_Samples>
      </ProbeBox>
      <DumpBox ID="6" Name="Ef" DumpType="10" DumpMode="2" FileType="1">
        <Primitives><Box Priority="0"><P1 X="0" Y="0" Z="4"/><P2 X="20" Y="20" Z="4"/></Box></Primitives>
        <FD_Samples>3e9</FD_Samples>
      </DumpBox>
      <DumpBox ID="7" Name="Et" DumpType="0" DumpMode="0" FileType="0">
        <Primitives><Box Priority="0"><P1 X="0" Y="0" Z="4"/><P2 X="20" Y="20" Z="4"/></Box></Primitives>
      </DumpBox>
    </Properties>
    <RectilinearGrid DeltaUnit="0.001" CoordSystem="0">
      <XLines>0,1,2,3,4,5,6,7,8,9,10,11,12,13,14,15,16,17,18,19,20</XLines>
      <YLines>20,19,18,17,16,15,14,13,12,11,10,9,8,7,6,5,4,3,2,1,0</YLines>
      <ZLines>0,0.5,1,1.5,2,3,4,5,6,7,8,9,10</ZLines>
    </RectilinearGrid>
  </ContinuousStructure>
</openEMS>
"""


def test_fdtd_block_as_the_reference_reads_it():
    sc = read_xml(SETUP)
    assert sc["nrts"] == 2000                              # NumberOfTimesteps read as a double
    assert sc["end_criteria"] == 1e-6                      # 0 -> the 1e-6 default
    assert sc["oversampling"] == 4                         # values <= 1 keep the default 4
    assert sc["bc"] == [3, 2, 0, 1, 0, 0]                  # PML_8, MUR, 0, 1, PEC, unknown -> PEC
    assert sc["pml_size"][0] == 8
    assert sc["mur_vph"] == [2.5e8, 2e8, 2.5e8, 2.5e8, 2.5e8, 2.5e8]
    assert sc["signal"] == {"type": "gauss", "f0": 2e9, "fc": 5e9}
    assert sc["ts_factor"] == 0.9 and "forced_dt" not in sc


def test_structure_subset():
    sc = read_xml(SETUP)
    assert sc["delta_unit"] == 1e-3
    assert np.array_equal(sc["lines"][1], np.arange(21.0))      # sorted
    assert sc["materials"] == [{"eps": 3.66, "mue": 1.0, "kappa": 0.001, "sigma": 0.0, "priority": 10,
                                "start": (0.0, 0.0, 0.0), "stop": (20.0, 20.0, 1.5)}]
    assert sc["metals"][0] == {"curve": [(2.0, 2.0, 2.0), (2.0, 2.0, 8.0), (6.0, 4.0, 8.0)]}
    assert sc["metals"][1]["priority"] == 4
    assert sc["lumped"][0]["dir"] == 2 and sc["lumped"][0]["R"] == 50 and sc["lumped"][0]["type"] == "parallel"
    e = sc["excitations"][0]
    assert e["val"] == (-0.0, -0.0, -1.0) and e["delay"] == 1e-10 and e["type"] == 0
    ut, it = sc["probes"]
    assert (ut["name"], ut["type"], ut["weight"], ut["freqs"]) == ("port_ut1", 0, -1.0, [])
    assert (it["name"], it["type"], it["norm_dir"], it["freqs"]) == ("port_it1", 1, 2, [1e9, 2e9])
    assert sc["dumps"] == [{"name": "Ef", "start": [0.0, 0.0, 4.0], "stop": [20.0, 20.0, 4.0], "freqs": [3e9],
                            "field": "E", "interp": 2, "file_type": 1, "start_time": 0.0, "stop_time": 0.0},
                           {"name": "Et", "start": [0.0, 0.0, 4.0], "stop": [20.0, 20.0, 4.0], "field": "E",
                            "interp": 0, "file_type": 0, "domain": "td", "start_time": 0.0, "stop_time": 0.0}]
    assert sc["skipped"] == []
    op = Operator(sc, nthreads=2)
    assert op.N == (21, 21, 13)
    assert op.excitation()["dir"].size == 2


def _coax_scene():
    """Cylinder and cylindrical-shell metals and a weighted excitation shell."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from test_pyopenems import coax_setup
    FDTD, _ = coax_setup(res=10.0, length=200.0, nrts=100)
    scene = FDTD.scene()
    assert any("radius" in m for m in scene["metals"]) and scene["excitations"][0]["weight"]
    return scene


def test_unsupported_parts_are_refused():
    with pytest.raises(ValueError, match="Cylinder|cylindrical"):
        read_xml(SETUP.replace('<FDTD NumberOfTimesteps="2e3"', '<FDTD CylinderCoords="1" NumberOfTimesteps="2e3"'))
    with pytest.raises(ValueError, match="excitation type 7"):
        read_xml(SETUP.replace('<Excitation Type="0" f0="2e9" fc="5e9"/>', '<Excitation Type="7" f0="2e9"/>'))
    custom = read_xml(SETUP.replace('<Excitation Type="0" f0="2e9" fc="5e9"/>',
                                    '<Excitation Type="10" f0="2e9" Function="sin(2*pi*2e9*t)"/>'))
    assert custom["signal"] == {"type": "custom", "f0": 2e9, "fc": 0.0, "f_max": 7e9, "function": "sin(2*pi*2e9*t)"}
    with pytest.raises(ValueError, match="Sphere"):
        read_xml(SETUP.replace('<Box Priority="4">', '<Sphere Priority="4"/><Box Priority="4">'))


@pytest.mark.parametrize("make", [
    lambda: scenes.upml_bricks(40, 100, nbricks=8),
    lambda: scenes.dispersive_rlc(32, 100),
    lambda: scenes.wire_antenna(30, 100),
    lambda: scenes.mur_layers(32, 30, 28, 100),
    lambda: _coax_scene(),
])
def test_round_trip_rebuilds_the_same_operator(native, make, tmp_path):
    """write_xml -> read_xml reproduces every coefficient, timestep and
    excitation list of the scene it came from."""
    scene = make()
    path = tmp_path / "setup.xml"
    write_xml(scene, str(path))
    back = read_xml(str(path))
    a, b = Operator(scene, nthreads=2), Operator(back, nthreads=2)
    assert a.dT == b.dT and a.N == b.N
    for name in ("vv", "vi", "ii", "iv"):
        assert np.array_equal(a.coeff(name), b.coeff(name)), name
    for curr in (False, True):
        ea, eb = a.excitation(curr), b.excitation(curr)
        for k in ("index", "dir", "amp", "delay"):
            assert np.array_equal(ea[k], eb[k])
    assert [p["name"] for p in back["probes"]] == [p["name"] for p in scene.get("probes", [])]


def test_mode_matching_probes_are_refused(tmp_path):
    """Probe types 10 / 11 (ProcessModeMatch) are outside the engine's scope:
    the reader refuses them by name instead of writing an empty probe."""
    path = tmp_path / "mm.xml"
    path.write_text(SETUP.replace('NormDir="-1" Type="0"', 'NormDir="-1" Type="10" ModeFunctionX="0"'))
    with pytest.raises(ValueError, match="probe type 10"):
        read_xml(str(path))
