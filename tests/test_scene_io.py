"""Scene ingestion (host code, CPU): OBJ import equals the reference's own ObjImporter (compiled
from its sources) triangle for triangle and bit for bit, including its index rules and its
failures; the JSON scene format rebuilds the Cornell workload exactly; the reference-style
configuration files (// comments) map to render settings."""
import os

import numpy as np
import pytest

import oracle_py as O
import rbr_amd as R
import scene_io
import scenes as S

OBJ_CASES = {
    "tri": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",
    "quad_fan_and_slashes": (b"# a quad and a pentagon\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0\n"
                             b"vt 0 0\nvt 1 0\nvn 0 0 1\nf 1/1/1 2/2/1 3//1 4\nf 1 2 3 5 4\n"),
    "negative_indices": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf -3 -2 -1\nv 2 2 2\nf -1 -3 -2\n",
    "crlf_spaces_tabs": b"v\t0 0 0\r\nv  1e0 0 0 \r\n  v 0 1.0e+00 0\r\n\r\nf 1 2 3\r\n",
    "odd_numbers": b"v .5 -0 1.\nv 1e-3 3.4028235e38 -1.17549435e-38\nv 0.1 0.2 0.3\nf 1 2 3\n",
    "short_vertex_skipped": b"v 0 0 0\nv 1 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",
    "bad_texture_index_ignored": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nf 1/9 2/1 3/-5\n",
    "unknown_records": b"o thing\ng group\nusemtl red\nmtllib x.mtl\ns 1\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\nl 1 2\n",
    "degenerate_faces": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\nf 1\nf 1 2 3\n",
}
OBJ_FAILURES = {
    "missing_vertex": b"v 0 0 0\nv 1 0 0\nf 1 2 3\n",
    "zero_index": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",
    "face_before_vertices": b"f 1 2 3\nv 0 0 0\nv 1 0 0\nv 0 1 0\n",
    "empty": b"# nothing\n",
    "no_faces": b"v 0 0 0\nv 1 0 0\nv 0 1 0\n",
}


@pytest.mark.skipif(not O.ref_available(), reason="reference library not built")
@pytest.mark.parametrize("name", sorted(OBJ_CASES))
def test_obj_matches_reference_importer(name):
    text = OBJ_CASES[name]
    ref = O.ref_obj_parse(text)
    mine = R.parse_obj(text)
    assert ref is not None
    assert mine.shape == ref.shape
    assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))


@pytest.mark.skipif(not O.ref_available(), reason="reference library not built")
@pytest.mark.parametrize("name", sorted(OBJ_FAILURES))
def test_obj_failures_match_reference(name):
    assert O.ref_obj_parse(OBJ_FAILURES[name]) is None
    with pytest.raises(ValueError):
        R.parse_obj(OBJ_FAILURES[name])


def _mesh_obj(tris: np.ndarray) -> bytes:
    """An indexed OBJ of a triangle soup (vertices printed with 9 significant digits)."""
    v = tris.reshape(-1, 3)
    uniq, inv = np.unique(v, axis=0, return_inverse=True)
    lines = [f"v {x:.9g} {y:.9g} {z:.9g}" for x, y, z in uniq]
    f = inv.reshape(-1, 3) + 1
    lines += [f"f {a} {b} {c}" for a, b, c in f]
    return ("\n".join(lines) + "\n").encode()


@pytest.mark.skipif(not O.ref_available(), reason="reference library not built")
def test_obj_mesh_round_trip():
    tris = R.icosphere(12, 100.0, 0.1, 0xC1)
    text = _mesh_obj(tris)
    mine = R.parse_obj(text)
    assert np.array_equal(mine.view(np.uint32), O.ref_obj_parse(text).view(np.uint32))
    assert np.array_equal(mine, tris)   # 9 significant digits round-trip FP32 exactly


CORNELL_JSON = {
    "materials": {"white": {"diffuse": [0.73, 0.73, 0.73]}, "red": {"diffuse": [0.65, 0.05, 0.05]},
                  "green": {"diffuse": [0.12, 0.45, 0.15]}, "light": {"emissive": [17, 12, 4]}},
    "objects": [
        {"type": "quad", "origin": [0, 0, 0], "u": [555, 0, 0], "v": [0, 0, 555], "material": "white"},
        {"type": "quad", "origin": [0, 555, 0], "u": [555, 0, 0], "v": [0, 0, 555], "material": "white"},
        {"type": "quad", "origin": [0, 0, 555], "u": [555, 0, 0], "v": [0, 555, 0], "material": "white"},
        {"type": "quad", "origin": [555, 0, 0], "u": [0, 555, 0], "v": [0, 0, 555], "material": "red"},
        {"type": "quad", "origin": [0, 0, 0], "u": [0, 555, 0], "v": [0, 0, 555], "material": "green"},
        {"type": "quad", "origin": [213, 554, 227], "u": [130, 0, 0], "v": [0, 0, 105], "material": "light"},
        {"type": "box", "size": [165, 330, 165], "rotateY": 15, "translate": [265, 0, 295], "material": "white"},
        {"type": "box", "size": [165, 165, 165], "rotateY": -18, "translate": [130, 0, 65], "material": "white"},
    ],
    "lights": [{"type": "area", "position": [278, 554, 279.5], "u": [130, 0, 0], "v": [0, 0, 105],
                "color": [17, 12, 4], "intensity": 1}],
    "camera": {"position": [278, 278, -800], "lookAt": [278, 278, 0], "up": [0, 1, 0], "fov": 40,
               "width": 256, "height": 256},
}


def test_json_scene_rebuilds_cornell():
    a = scene_io.scene_from_dict(CORNELL_JSON)
    b = S.cornell_c0(256, 256)
    assert np.array_equal(a.quads, b.quads) and len(a.tris) == 0 and len(a.spheres) == 0
    # materials by content: the same per-quad material parameters
    ma = [a.materials[i] for i in a.quad_mat]
    mb = [b.materials[i] for i in b.quad_mat]
    assert ma == mb
    assert a.lights == b.lights
    for f in ("position", "forward", "right", "up"):
        assert np.array_equal(getattr(a.camera, f), getattr(b.camera, f))
    assert (a.camera.fx, a.camera.width, a.camera.height) == (b.camera.fx, b.camera.width, b.camera.height)


def test_json_scene_mesh_sphere_plane_medium(tmp_path):
    tris = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0], [0, 0, 1, 1, 0, 1, 0, 1, 1]], np.float32)
    (tmp_path / "m.obj").write_bytes(_mesh_obj(tris))
    d = {"materials": {"m": {"baseColor": [0.5, 0.4, 0.3], "metallic": 1, "roughness": 0.2}},
         "objects": [{"type": "mesh", "file": "m.obj", "material": "m", "scale": 2.0, "translate": [1, 2, 3]},
                     {"type": "sphere", "center": [0, 5, 0], "radius": 2, "material": {"diffuse": [1, 1, 1]}},
                     {"type": "plane", "normal": [0, 2, 0], "d": -1, "material": "m"}],
         "lights": [{"type": "point", "position": [0, 10, 0], "color": [1, 1, 1], "intensity": 100},
                    {"type": "directional", "direction": [0, -1, 0], "color": [1, 1, 1]}],
         "medium": {"type": "exponential", "sigmaA": [0.1, 0.1, 0.1], "sigmaS": [0.4, 0.4, 0.4], "falloff": 2, "g": 0.3},
         "background": [0.1, 0.2, 0.3], "camera": {"width": 64, "height": 48}}
    (tmp_path / "s.json").write_text(__import__("json").dumps(d))
    desc = scene_io.load_scene(str(tmp_path / "s.json"))
    assert np.array_equal(desc.tris, (tris.reshape(-1, 3, 3) * np.float32(2) + np.float32([1, 2, 3])).reshape(-1, 9))
    assert desc.spheres.tolist() == [[0, 5, 0, 2]] and desc.planes.tolist() == [[0, 1, 0, -1]]
    assert desc.materials[desc.tri_mat[0]]["metallic"] == 1.0 and desc.medium["type"] == 2
    assert desc.background == (0.1, 0.2, 0.3)
    assert (desc.camera.width, desc.camera.height) == (64, 48) and len(desc.lights) == 2
    with pytest.raises(ValueError):
        scene_io.scene_from_dict({"objects": [{"type": "sphere", "center": [0, 0, 0], "radius": 1, "material": "x"}]})


CONFIG_TEXT = """{
    // render settings, in the reference's configuration format
    "integrator": {
        "type": "pathtracing",   // pathtracing or volumepath
        "maxDepth": 12,
        "samplesPerPixel": 32,
        "nee": true,
        "mis": false,
    },
    /* camera */
    "camera": {"type": "perspective", "position": [278, 278, -800], "lookAt": [278, 278, 0], "up": [0, 1, 0],
               "fov": 40.0, "aperture": 0.0},
    "image": {"width": 64, "height": 48, "output": "out/cornell.png", "toneMapping": false, "gamma": 2.2},
    "scene": "cornell.json",
    "note": "a // inside a string stays"
}
"""


def test_config_mapping(tmp_path):
    (tmp_path / "c.json").write_text(CONFIG_TEXT)
    raw = scene_io.read_json(str(tmp_path / "c.json"))
    assert raw["note"] == "a // inside a string stays"
    cfg = scene_io.load_config(str(tmp_path / "c.json"))
    assert (cfg["integrator"], cfg["depth"], cfg["spp"], cfg["nee"], cfg["mis"]) == (R.PATH, 12, 32, True, False)
    assert (cfg["width"], cfg["height"], cfg["tone_map"], cfg["scene"]) == (64, 48, False, "cornell.json")
    bad = CONFIG_TEXT.replace('"aperture": 0.0', '"aperture": 0.1')
    (tmp_path / "d.json").write_text(bad)
    with pytest.raises(ValueError):
        scene_io.load_config(str(tmp_path / "d.json"))
    (tmp_path / "e.json").write_text(CONFIG_TEXT.replace('"pathtracing",', '"simple",'))
    with pytest.raises(ValueError):
        scene_io.load_config(str(tmp_path / "e.json"))
    (tmp_path / "f.json").write_text(CONFIG_TEXT.replace('"pathtracing",', '"whitted",').replace(
        '"scene": "cornell.json",', '"scene": "cornell.json", "ambient": [0.0, 0.0, 0.0],'))
    w = scene_io.load_config(str(tmp_path / "f.json"))
    assert w["integrator"] == R.WHITTED and w["ambient"] == (0.0, 0.0, 0.0) and cfg["ambient"] == (0.1, 0.1, 0.1)


@pytest.mark.gpu
def test_render_config_on_gpu(tmp_path):
    """A reference-style configuration with a scene file renders on the GPU to the same film
    as the equivalent scene description, and writes its PNG (Film::toRGB8) and PFM."""
    import json
    (tmp_path / "cornell.json").write_text(json.dumps(CORNELL_JSON))
    (tmp_path / "c.json").write_text(CONFIG_TEXT.replace("out/cornell.png", str(tmp_path / "out" / "cornell.png")))
    with R.Context(0) as ctx:
        film = scene_io.render_config(ctx, str(tmp_path / "c.json"), seed=5)
        cam = {"position": [278, 278, -800], "lookAt": [278, 278, 0], "up": [0, 1, 0], "fov": 40.0}
        desc = scene_io.scene_from_dict(CORNELL_JSON, width=64, height=48, camera=cam)
        sc = R.scene_from_desc(ctx, desc)
        ref = R.render(ctx, [dict(scene=sc, spp=32, depth=12, seed=5, nee=True, mis=False)])[0]
    assert film.shape == (48, 64, 3) and np.array_equal(film, ref)
    assert np.array_equal(R.read_pfm(str(tmp_path / "out" / "cornell.pfm")), film)
    assert (tmp_path / "out" / "cornell.png").stat().st_size > 64 * 48 * 3


OBJ_NORMAL_CASES = {
    "all_corners": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 2\nvn 0 0.6 0.8\nvn 1 1 1\nf 1//1 2//2 3//3\n",
    "quad_fan_shared": (b"v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\nvn 0.1 0 1\n"
                        b"f 1/1/1 2/1/2 3/1/1 4/1/2\n"),
    "negative_indices": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvn 1 0 0\nf -3//-2 -2//-1 -1//-2\n",
    "one_corner_without": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3\n",
    "zero_normal": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvn 0 0 0\nf 1//1 2//2 3//1\n",
    "out_of_range": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//5 3//1\n",
    "no_normals": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n",
    "degenerate_face_without": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2//1 3//1\nf 1 2\n",
    "tiny_normal": b"v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1e-30\nf 1//1 2//1 3//1\n",
}


@pytest.mark.skipif(not O.ref_available(), reason="reference library not built")
@pytest.mark.parametrize("name", sorted(OBJ_NORMAL_CASES))
def test_obj_normals_match_reference_importer(name):
    """Vertex normals kept (or dropped) exactly as the reference's importer does, per corner."""
    text = OBJ_NORMAL_CASES[name]
    ref = O.ref_obj_parse_normals(text)
    mine = R.parse_obj_normals(text)
    assert (ref is None) == (mine is None), (ref, mine)
    if ref is not None:
        assert np.array_equal(mine.view(np.uint32), ref.view(np.uint32))
    assert np.array_equal(R.parse_obj(text), O.ref_obj_parse(text))


@pytest.mark.skipif(not O.ref_available(), reason="reference library not built")
def test_shading_normal_matches_reference_triangle():
    """rbr_triangle_shading_normal (the function the renderer evaluates for triangles with
    vertex normals) against the reference's Triangle::intersect on random triangles, corner
    normals (unnormalized, some facing away) and rays: the same bits after the reference's
    setFaceNormal (faced towards the ray)."""
    rng = np.random.default_rng(17)
    checked = 0
    for _ in range(3000):
        v = rng.normal(size=9).astype(np.float32) * np.float32(3)
        n = rng.normal(size=9).astype(np.float32)
        bary = rng.dirichlet([1, 1, 1])
        target = (bary[0] * v[0:3] + bary[1] * v[3:6] + bary[2] * v[6:9]).astype(np.float32)
        o = (target + rng.normal(size=3) * 4).astype(np.float32)
        d = (target - o).astype(np.float32)
        hit = O.ref_triangle_normal(v, n, o, d)
        if hit is None:
            continue
        rn, front, dn = hit
        mine = R.triangle_shading_normal(v, n, o, dn)
        # the reference's frontFace test: dot(ray.direction, outward) < 0 in float
        f32 = np.float32
        dotv = f32(f32(dn[0] * mine[0]) + f32(dn[1] * mine[1])) + f32(dn[2] * mine[2])
        faced = mine if dotv < 0 else -mine
        assert bool(dotv < 0) == front
        assert np.array_equal(faced.view(np.uint32), rn.view(np.uint32)), (v, n, o, dn, faced, rn)
        checked += 1
    assert checked > 2000


def test_json_scene_smooth_mesh(tmp_path):
    text = OBJ_NORMAL_CASES["all_corners"]
    (tmp_path / "m.obj").write_bytes(text)
    d = {"materials": {"m": {"diffuse": [0.5, 0.5, 0.5]}},
         "objects": [{"type": "mesh", "file": "m.obj", "material": "m", "scale": 2.0},
                     {"type": "mesh", "file": "m.obj", "material": "m", "smooth": False},
                     {"type": "sphere", "center": [0, 5, 0], "radius": 1, "material": "m"}],
         "camera": {"width": 8, "height": 8}}
    desc = scene_io.scene_from_dict(d, str(tmp_path))
    assert desc.tri_normals is not None and desc.tri_normals.shape == (2, 9)
    assert np.array_equal(desc.tri_normals[0], R.parse_obj_normals(text)[0]) and not desc.tri_normals[1].any()
    flat = scene_io.scene_from_dict({**d, "objects": d["objects"][1:]}, str(tmp_path))
    assert flat.tri_normals is None
