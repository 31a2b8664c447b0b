"""Scene files for the device path.

The reference reads render settings from JSON configuration files (configs/*.json, with //
comments) and leaves their "scene" key unimplemented (apps/main.cpp:62-65).  This module
reads those configuration files (render settings -> render jobs) and a JSON scene format
of this project's own that fills the device tables: meshes from Wavefront OBJ files (imported
as the reference's ObjImporter does, rbr_obj_parse), spheres, quads, boxes, planes,
materials, lights, an optional medium, background and camera.

Scene file:
  {
    "camera":     {"position": [x, y, z], "lookAt": [x, y, z], "up": [0, 1, 0], "fov": 40,
                   "width": 256, "height": 256},
    "materials":  {"white": {"baseColor": [.73, .73, .73], "metallic": 0, "roughness": 1,
                             "ior": 1.5, "transmission": 0, "emission": [0, 0, 0]}, ...},
    "objects":    [{"type": "mesh", "file": "bunny.obj", "material": "white",
                    "scale": 1.0, "translate": [0, 0, 0], "smooth": true},
                   {"type": "sphere", "center": [x, y, z], "radius": r, "material": ...},
                   {"type": "quad", "origin": [..], "u": [..], "v": [..], "material": ...},
                   {"type": "box", "size": [..], "rotateY": deg, "translate": [..], "material": ...},
                   {"type": "plane", "normal": [..], "d": 0, "material": ...}],
    "lights":     [{"type": "area", "position": [..], "u": [..], "v": [..], "color": [..], "intensity": 1},
                   {"type": "point", "position": [..], "color": [..], "intensity": ..},
                   {"type": "directional", "direction": [..], "color": [..], "intensity": ..}],
    "medium":     {"type": "homogeneous" | "exponential", "sigmaA": [..], "sigmaS": [..],
                   "falloff": 0, "baseHeight": 0, "g": 0},
    "background": [0, 0, 0]
  }
Material keys follow Material's parameters (material.h); "diffuse": [r, g, b] and
"emissive": [r, g, b] are the reference's Material::diffuse / Material::emissive
shorthands.  Relative OBJ paths are taken from the scene file's folder; a mesh whose file
has vertex normals for every corner shades with them (the reference's importer keeps them
then) unless "smooth" is false; scale and translate do not change them.  An area light of a
visible emitter also needs the emitting quad as an object, as in the reference's scenes.
"""
from __future__ import annotations

import json
import math
import os
import re

import numpy as np

import rbr_amd as R
import scenes as S

F = np.float32


def _strip_comments(text: str) -> str:
    """Removes // and /* */ comments outside strings (the reference's config files use them)."""
    out, i, n, in_str = [], 0, len(text), False
    while i < n:
        c = text[i]
        if in_str:
            out.append(c)
            if c == "\\" and i + 1 < n:
                out.append(text[i + 1])
                i += 2
                continue
            if c == '"':
                in_str = False
            i += 1
        elif c == '"':
            in_str = True
            out.append(c)
            i += 1
        elif text.startswith("//", i):
            j = text.find("\n", i)
            i = n if j < 0 else j
        elif text.startswith("/*", i):
            j = text.find("*/", i + 2)
            i = n if j < 0 else j + 2
        else:
            out.append(c)
            i += 1
    return re.sub(r",(\s*[}\]])", r"\1", "".join(out))   # trailing commas


def read_json(path: str) -> dict:
    with open(path) as f:
        return json.loads(_strip_comments(f.read()))


def _v3(x, default=(0.0, 0.0, 0.0)):
    return tuple(float(a) for a in (default if x is None else x))


def _material(m: dict) -> dict:
    if "diffuse" in m:
        return S.diffuse(_v3(m["diffuse"]))
    if "emissive" in m:
        return S.emissive(_v3(m["emissive"]))
    return S.material(_v3(m.get("baseColor"), (0.8, 0.8, 0.8)), m.get("metallic", 0.0), m.get("roughness", 1.0),
                      m.get("ior", 1.5), m.get("transmission", 0.0), _v3(m.get("emission")))


def scene_from_dict(d: dict, base_dir: str = ".", width: int | None = None, height: int | None = None,
                    camera: dict | None = None) -> S.SceneDesc:
    """A SceneDesc from a parsed scene file; width/height/camera override the file's camera
    (a render configuration's image and camera sections)."""
    b = S._Builder()
    names = {}
    for name, m in d.get("materials", {}).items():
        names[name] = b.mat(_material(m))

    def mat_of(o):
        m = o.get("material")
        if isinstance(m, dict):
            return b.mat(_material(m))
        if m not in names:
            raise ValueError(f"unknown material {m!r}")
        return names[m]

    for o in d.get("objects", []):
        t = o.get("type")
        if t == "mesh":
            path = o["file"] if os.path.isabs(o["file"]) else os.path.join(base_dir, o["file"])
            tris, normals = R.load_obj(path, with_normals=True)
            tris = tris.reshape(-1, 3, 3) * F(o.get("scale", 1.0)) + np.asarray(_v3(o.get("translate")), F)
            b.tri_mesh(tris.reshape(-1, 9).astype(F), mat_of(o), normals if o.get("smooth", True) else None)
        elif t == "sphere":
            b.sphere(_v3(o["center"]), float(o["radius"]), mat_of(o))
        elif t == "quad":
            b.quad(_v3(o["origin"]), _v3(o["u"]), _v3(o["v"]), mat_of(o))
        elif t == "box":
            b.box(_v3(o["size"]), float(o.get("rotateY", 0.0)), _v3(o.get("translate")), mat_of(o))
        elif t == "plane":
            n = S.normalized(_v3(o["normal"]))
            b.planes.append(np.array([n[0], n[1], n[2], float(o.get("d", 0.0))], F))
            b.pm.append(mat_of(o))
        else:
            raise ValueError(f"unknown object type {t!r}")
    for li in d.get("lights", []):
        t = li.get("type")
        col, inten = _v3(li.get("color"), (1.0, 1.0, 1.0)), float(li.get("intensity", 1.0))
        if t == "area":
            b.lights.append(S.area_light(_v3(li["position"]), _v3(li["u"]), _v3(li["v"]), col, inten))
        elif t == "point":
            b.lights.append(S.point_light(_v3(li["position"]), col, inten))
        elif t == "directional":
            b.lights.append(S.directional_light(_v3(li["direction"]), col, inten))
        else:
            raise ValueError(f"unknown light type {t!r}")
    med = None
    if d.get("medium"):
        m = d["medium"]
        med = dict(type=1 if m.get("type", "homogeneous") == "homogeneous" else 2,
                   sigma_a=_v3(m.get("sigmaA")), sigma_s=_v3(m.get("sigmaS")), falloff=float(m.get("falloff", 0.0)),
                   base_height=float(m.get("baseHeight", 0.0)), g=float(m.get("g", 0.0)))
    cam = dict(d.get("camera", {}))
    cam.update(camera or {})
    w = int(width or cam.get("width", 256))
    h = int(height or cam.get("height", 256))
    c = S.look_at(_v3(cam.get("position"), (0.0, 0.0, 5.0)), _v3(cam.get("lookAt")), _v3(cam.get("up"), (0, 1, 0)),
                  w, h, float(cam.get("fov", 45.0)))
    return b.build(c, medium=med, background=_v3(d.get("background")), name=d.get("name", "scene"))


def load_scene(path: str, **kw) -> S.SceneDesc:
    return scene_from_dict(read_json(path), os.path.dirname(os.path.abspath(path)), **kw)


# Render configuration (the reference's configs/*.json) -> render settings.  Integrator types
# with a device path (config.cpp:45-52): "pathtracing" / "pt", "volumepath" / "vpt",
# "whitted", "normal", "depth", "ao"; cameras: "perspective" (pinhole; an aperture > 0 is
# refused, there is no thin lens on the device).  "ambient" is the scene's ambient light, which
# the Whitted integrator uses (the reference's default 0.1).
_INTEGRATORS = {"pathtracing": R.PATH, "pt": R.PATH, "volumepath": R.VOLPATH, "vpt": R.VOLPATH,
                "whitted": R.WHITTED, "normal": R.NORMAL, "depth": R.DEPTH, "ao": R.AO}


def load_config(path: str) -> dict:
    c = read_json(path)
    integ = c.get("integrator", {})
    typ = integ.get("type", "pathtracing")
    if typ not in _INTEGRATORS:
        raise ValueError(f"integrator {typ!r} has no device path ({', '.join(sorted(_INTEGRATORS))})")
    cam = c.get("camera", {})
    if cam.get("type", "perspective") != "perspective" or float(cam.get("aperture", 0.0)) > 0.0:
        raise ValueError("only the pinhole perspective camera has a device path")
    img = c.get("image", {})
    out = {
        "integrator": _INTEGRATORS[typ],
        "depth": int(integ.get("maxDepth", 10)),
        "spp": int(integ.get("samplesPerPixel", 128)),
        "nee": bool(integ.get("nee", True)),
        "mis": bool(integ.get("mis", True)),
        "width": int(img.get("width", 800)),
        "height": int(img.get("height", 600)),
        "output": img.get("output", "results/output.png"),
        "tone_map": bool(img.get("toneMapping", True)),
        "gamma": float(img.get("gamma", 2.2)),
        "exposure": float(img.get("exposure", 1.0)),
        "camera": {"position": cam.get("position", [0.0, 2.0, 8.0]), "lookAt": cam.get("lookAt", [0.0, 1.0, 0.0]),
                   "up": cam.get("up", [0.0, 1.0, 0.0]), "fov": cam.get("fov", 60.0)},
        "scene": c.get("scene", ""),
        "background": c.get("background"),
        "ambient": tuple(float(a) for a in c.get("ambient", (0.1, 0.1, 0.1))),
    }
    return out


def render_config(ctx: R.Context, config_path: str, scene_path: str | None = None, seed: int = 1,
                  write: bool = True) -> np.ndarray:
    """Renders a reference configuration on the GPU with a scene file (the config's "scene"
    key, resolved from the config's folder, or scene_path); writes the config's 8-bit output
    (gamma / tone mapping as Film::toRGB8*) and a .pfm beside it.  Returns the film."""
    cfg = load_config(config_path)
    sp = scene_path or cfg["scene"]
    if not sp:
        raise ValueError("no scene file: the reference's built-in demo scenes have no device description")
    if not os.path.isabs(sp):
        sp = os.path.join(os.path.dirname(os.path.abspath(config_path)), sp)
    d = read_json(sp)
    if cfg["background"] is not None:
        d["background"] = cfg["background"]
    desc = scene_from_dict(d, os.path.dirname(os.path.abspath(sp)), cfg["width"], cfg["height"], cfg["camera"])
    sc = R.scene_from_desc(ctx, desc)
    film = R.render(ctx, [dict(scene=sc, spp=cfg["spp"], depth=cfg["depth"], seed=seed, integrator=cfg["integrator"],
                               nee=cfg["nee"], mis=cfg["mis"], ambient=cfg["ambient"])])[0]
    sc.close()
    if write:
        out = cfg["output"]
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        R.write_png(out if out.lower().endswith(".png") else out + ".png",
                    R.film_to_rgb8(film, cfg["gamma"], cfg["tone_map"], cfg["exposure"]))
        R.write_pfm(os.path.splitext(out)[0] + ".pfm", film)
    return film
