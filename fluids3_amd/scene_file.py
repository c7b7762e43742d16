"""Scene/parameter files in the reference's scene.xml schema -> fx_scene.

The schema (README.md "Scene parameters"; the reference's example file
fluids5.0/assets/scene.xml) is one <Fluid> element followed by <Scene id="k">
elements, each holding parameters as text elements; "all parameters are
permitted in either Fluid or Scene sections" and a scene's values override
the Fluid section's (FluidSystem::ParseXML called for "Fluid" then for the
scene, fluids3.2/fluids/fluid_system.cpp:2268-2309).  Vectors are written
"x, y, z".  The file has several top-level elements, so it is parsed inside a
synthetic root.

Keys the SPH step of this library uses map onto fx_scene fields; keys for
forces the step does not have (point gravity, the X+ wave wall) or for the
viewer are reported back in `unused` with their values, so a caller can tell
when a scene relies on them.  Fields a file does not set keep the dam-break
defaults of scenes.dam_break.
"""
import math
import xml.etree.ElementTree as ET

from . import scenes
from .abi import f3

SCALAR_KEYS = {
    "DT": "dt",
    "SimScale": "sim_scale",
    "Viscosity": "visc",
    "RestDensity": "rest_density",
    "Mass": "mass",
    "Radius": "pradius",
    "IntStiff": "int_stiff",
    "BoundStiff": "bound_stiff",
    "BoundDamp": "bound_damp",
    "AccelLimit": "accel_limit",
    "VelLimit": "vel_limit",
    "GroundSlope": "bound_slope",
    "WaveForceMin": "bound_wall_force",  # amplitude of the X- wave wall
    "WaveForceFreq": "bound_wall_freq",
    "GridDensity": "grid_density",
    "Spacing": "spacing",
    "SmoothRadius": "smooth_radius",
}
VECTOR_KEYS = {"VolMin": "bound_min", "VolMax": "bound_max", "InitMin": "init_min", "InitMax": "init_max"}
# keys of the schema that the step does not use
UNUSED_KEYS = ("PointGravAmt", "PointGravPos", "WaveForceMax", "SearchDist", "DrawMode", "DrawGrid", "DrawText")


class SceneFileError(ValueError):
    pass


def _vec(text, key):
    parts = [p for p in text.replace(",", " ").split() if p]
    if len(parts) != 3:
        raise SceneFileError(f"{key}: expected three numbers, got {text!r}")
    return tuple(float(p) for p in parts)


def _sections(text):
    try:
        root = ET.fromstring("<SceneFile>" + text + "</SceneFile>")
    except ET.ParseError as e:
        raise SceneFileError(f"not a scene file: {e}") from None
    fluid = root.find("Fluid")
    sc = {el.get("id"): el for el in root.findall("Scene")}
    return fluid, sc


def scene_ids(text):
    """The ids of the <Scene> sections, in file order."""
    _, sc = _sections(text)
    return list(sc)


def load_scene(text, scene_id, seed=scenes.SEED):
    """(name, fx_scene, unused) of scene `scene_id` (the <Fluid> section then
    the scene's own values).  `text` is the file's contents."""
    fluid, sc = _sections(text)
    key = str(scene_id)
    if key not in sc:
        raise SceneFileError(f"no <Scene id=\"{key}\"> (ids: {', '.join(sc)})")
    s = scenes.dam_break(0, (0, 0, 0), (1, 1, 1), (0, 0, 0), (1, 1, 1), seed=seed)
    name = ""
    unused = {}
    for section in (fluid, sc[key]):
        if section is None:
            continue
        for el in section:
            tag, val = el.tag, (el.text or "").strip()
            if tag == "Name":
                name = val
            elif tag == "Num":
                s.num_points = int(float(val))
            elif tag in SCALAR_KEYS:
                setattr(s, SCALAR_KEYS[tag], float(val))
            elif tag in VECTOR_KEYS:
                setattr(s, VECTOR_KEYS[tag], f3(_vec(val, tag)))
            elif tag == "PlaneGravDir":
                g = _vec(val, tag)
                s.grav_dir = f3(g)
                s.grav_amt = 1.0
            elif tag in UNUSED_KEYS:
                unused[tag] = val
            else:
                raise SceneFileError(f"unknown key <{tag}>")
    if s.num_points <= 0:
        raise SceneFileError(f"scene {key}: <Num> missing or not positive")
    return name, s, unused


def load_scene_file(path, scene_id, seed=scenes.SEED):
    with open(path, "r") as fh:
        return load_scene(fh.read(), scene_id, seed=seed)


def gravity(scene):
    """Gravity vector of a scene (grav_dir * grav_amt) and its magnitude."""
    g = [scene.grav_dir.x * scene.grav_amt, scene.grav_dir.y * scene.grav_amt, scene.grav_dir.z * scene.grav_amt]
    return g, math.sqrt(sum(x * x for x in g))
