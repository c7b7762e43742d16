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


# ------------------------------------------- the committed schema fixture

import json
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture_text():
    with open(os.path.join(GOLDEN, "scene_schema.xml")) as fh:
        return fh.read()


def test_fixture_matches_hand_written_values():
    """tests/golden/scene_schema.xml against values read off the file by hand:
    the XML declaration and comments are skipped, an entity is decoded, the
    scene's <Viscosity> overrides the Fluid section's, and keys of forces the
    step lacks come back as unused."""
    text = fixture_text()
    with open(os.path.join(GOLDEN, "scene_schema_expected.json")) as fh:
        want = json.load(fh)
    assert scene_ids(text) == want["ids"]
    for sid in want["ids"]:
        name, s, unused = load_scene(text, sid)
        exp = want[sid]
        assert name == exp["name"]
        assert unused == exp["unused"]
        for k, v in exp.items():
            if k in ("name", "unused"):
                continue
            got = getattr(s, k)
            if isinstance(v, list):
                assert got.tolist() == pytest.approx(v), (sid, k)
            else:
                assert got == pytest.approx(v, rel=1e-6), (sid, k)


def test_truncated_and_malformed_files_are_refused_not_crashed():
    """Every prefix of the fixture, and a set of broken variants, either parse
    or raise SceneFileError (the reader is bounds-checked C++)."""
    text = fixture_text()
    for cut in range(0, len(text), 7):  # every 7th prefix
        try:
            load_scene(text[:cut], 7)
        except SceneFileError:
            pass
    bad = [
        text.replace("</Fluid>", ""),                         # unclosed section
        text.replace("<DT>", "<DT a=1>"),                     # unquoted attribute
        text.replace("0.0025", "0.0025x"),                    # not a number
        text.replace("&amp;", "&bogus;"),                     # unknown entity
        text.replace("<Num>           12000", "<Num> -5"),    # negative count
        text.replace("</Name>", "</Nmae>"),                   # mismatched end tag
        text.replace("<VolMin>        -40, 0, -20", "<VolMin> 1,2,3,4"),  # four components
        "<Scene id=\"7\"><Num>10</Num><Inner><X>1</X></Inner></Scene>",  # nested element
    ]
    for t in bad:
        with pytest.raises(SceneFileError):
            load_scene(t, 7)
    with pytest.raises(SceneFileError):
        load_scene(text, "seven")
