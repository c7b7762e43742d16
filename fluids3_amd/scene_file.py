"""Scene/parameter files in the reference's scene.xml schema -> fx_scene.

A thin mirror of the C ABI's reader (fx_scene_from_xml / fx_scene_xml_ids in
libfluidsx.so, fluids3_amd/csrc/scene_xml.cpp), so that a C++ host and Python
read a file the same way.  The schema (the reference's README.md "Scene parameters"; the
reference's example file fluids5.0/assets/scene.xml) is one <Fluid> element
followed by <Scene id="k"> elements, each holding parameters as text
elements; a scene's values override the Fluid section's (FluidSystem::ParseXML
called for "Fluid" then for the scene, fluids3.2/fluids/fluid_system.cpp:
2268-2309).  Vectors are written "x, y, z".

Keys the SPH step uses map onto fx_scene fields; keys for forces the step does
not have (point gravity, the X+ wave wall) or for the viewer are reported back
in `unused` with their values, so a caller can tell when a scene relies on
them.  Fields a file does not set keep the dam-break defaults of
scenes.dam_break.
"""
import ctypes as C
import math

from . import scenes
from .particles import lib


class SceneFileError(ValueError):
    pass


def _text(text):
    if isinstance(text, str):
        text = text.encode()
    if b"\0" in text:
        raise SceneFileError("not a scene file: NUL byte")
    return text


def scene_ids(text):
    """The ids of the <Scene> sections, in file order (as strings)."""
    t = _text(text)
    ids = (C.c_int32 * 256)()
    count = C.c_int32()
    st = lib().fx_scene_xml_ids(t, ids, 256, C.byref(count))
    if st != 0:
        raise SceneFileError(lib().fx_last_error().decode(errors="replace"))
    return [str(ids[k]) for k in range(count.value)]


def load_scene(text, scene_id, seed=scenes.SEED):
    """(name, fx_scene, unused) of scene `scene_id` (the <Fluid> section then
    the scene's own values).  `text` is the file's contents."""
    t = _text(text)
    try:
        sid = int(str(scene_id))
    except ValueError:
        raise SceneFileError(f"scene id {scene_id!r} is not an integer") from None
    s = scenes.dam_break(0, (0, 0, 0), (1, 1, 1), (0, 0, 0), (1, 1, 1), seed=seed)
    name = C.create_string_buffer(256)
    unused = C.create_string_buffer(4096)
    st = lib().fx_scene_from_xml(t, sid, C.byref(s), name, 256, unused, 4096)
    if st != 0:
        raise SceneFileError(lib().fx_last_error().decode(errors="replace"))
    extra = {}
    for line in unused.value.decode().splitlines():
        k, _, v = line.partition("=")
        extra[k] = v
    return name.value.decode(), s, extra


def load_scene_file(path, scene_id, seed=scenes.SEED):
    with open(path, "r") as fh:
        return load_scene(fh.read(), scene_id, seed=seed)


def gravity(scene):
    """Gravity vector of a scene (grav_dir * grav_amt) and its magnitude."""
    g = [scene.grav_dir.x * scene.grav_amt, scene.grav_dir.y * scene.grav_amt, scene.grav_dir.z * scene.grav_amt]
    return g, math.sqrt(sum(x * x for x in g))
