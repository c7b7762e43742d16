"""Scene files in the reference's scene.xml schema (fluids3_amd/scene_file.py)."""
import numpy as np
import pytest

from fluids3_amd import derive_params
from fluids3_amd.scene_file import SceneFileError, gravity, load_scene, scene_ids

TEXT = """
<Fluid>
  <DT>        0.002      </DT>
  <SimScale>  0.005      </SimScale>
  <Viscosity> 0.3        </Viscosity>
  <RestDensity> 800.0    </RestDensity>
  <IntStiff>  1.5        </IntStiff>
  <BoundDamp> 100.0      </BoundDamp>
  <PlaneGravDir> 0,-9.8,0 </PlaneGravDir>
  <PointGravAmt> 0.0     </PointGravAmt>
</Fluid>
<Scene id="1">
  <Name>  Test Pool </Name>
  <Num>   4096      </Num>
  <VolMin> -50, 0, -25 </VolMin>
  <VolMax>  50, 50, 25 </VolMax>
  <InitMin> -45, 2, -20 </InitMin>
  <InitMax> -10, 40, 20 </InitMax>
  <WaveForceFreq> 8.0 </WaveForceFreq>
  <GroundSlope> 0.04 </GroundSlope>
  <Viscosity> 0.4 </Viscosity>
</Scene>
<Scene id="2">
  <Num> 100 </Num>
  <VolMin> 0,0,0 </VolMin> <VolMax> 40,40,40 </VolMax>
  <InitMin> 1,1,1 </InitMin> <InitMax> 10,10,10 </InitMax>
  <WaveForceMax> 20.0 </WaveForceMax>
</Scene>
"""


def test_scene_overrides_fluid_section():
    assert scene_ids(TEXT) == ["1", "2"]
    name, s, unused = load_scene(TEXT, 1)
    assert name == "Test Pool"
    assert s.num_points == 4096
    assert s.dt == pytest.approx(0.002) and s.sim_scale == pytest.approx(0.005)
    assert s.visc == pytest.approx(0.4)  # the scene's value wins
    assert s.rest_density == pytest.approx(800.0) and s.int_stiff == pytest.approx(1.5)
    assert s.bound_damp == pytest.approx(100.0) and s.bound_slope == pytest.approx(0.04)
    assert s.bound_wall_freq == pytest.approx(8.0)
    assert s.bound_min.tolist() == [-50, 0, -25] and s.init_max.tolist() == [-10, 40, 20]
    g, gm = gravity(s)
    assert g == pytest.approx([0.0, -9.8, 0.0]) and gm == pytest.approx(9.8)
    assert unused == {"PointGravAmt": "0.0"}
    P = derive_params(s)  # a valid scene for the step
    assert P.pvisc == pytest.approx(0.4) and P.prest_dens == pytest.approx(800.0)
    assert np.allclose(P.gravity.tolist(), [0.0, -9.8, 0.0])


def test_unused_forces_are_reported_and_errors_raised():
    _, s, unused = load_scene(TEXT, 2)
    assert s.num_points == 100 and unused["WaveForceMax"] == "20.0"
    with pytest.raises(SceneFileError):
        load_scene(TEXT, 3)
    with pytest.raises(SceneFileError):
        load_scene(TEXT.replace("<GroundSlope>", "<Bogus>").replace("</GroundSlope>", "</Bogus>"), 1)
    with pytest.raises(SceneFileError):
        load_scene(TEXT.replace("-50, 0, -25", "-50, 0"), 1)
