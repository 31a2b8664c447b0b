"""Pins the oracle's restated integrators (oracle/orc_render.cpp) against the reference's
own Scene / PathTracingIntegrator / VolumetricPathIntegrator compiled from its sources
(oracle/_ref): with the same sampler stream (mt19937_64 reseeded per row) the images must
agree bit for bit, up to pixels whose paths cross an exact BVH tie."""
import numpy as np
import pytest

import oracle_py as O
import scenes as S

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref/libref.so not built")


def compare(desc, spp, depth, integrator=0, nee=True, mis=True):
    o = O.OracleRender(desc).render(spp, depth, integrator, nee, mis)
    r, _ = O.RefRender(desc).render(spp, depth, integrator, nee, mis)
    exact = (o == r).all(-1).mean()
    return o, r, exact


@pytest.mark.parametrize("nee,mis", [(True, True), (True, False), (False, False)])
def test_cornell_path_bit_exact(nee, mis):
    o, r, exact = compare(S.cornell_c0(48, 48), 8, 8, 0, nee, mis)
    assert exact == 1.0
    assert o.mean() > 0.01


@pytest.mark.parametrize("smooth", [True, False])
def test_smooth_mesh_path_bit_exact(smooth):
    """Triangles with vertex normals: the oracle's interpolated shading normal and the
    reference's Triangle::intersect (triangle.cpp:40-48) give the same image bit for bit,
    except where a path meets a shared edge of two triangles (both report the same t; the
    reference keeps the one it tested last, the oracle the lowest id) or the ball's contact
    point with the floor: about 1 pixel in 1,000, smooth or flat alike."""
    o, r, exact = compare(S.smooth_ball(40, 30, smooth), 8, 4)
    assert exact >= 0.998, exact
    assert abs(float(o.mean()) - float(r.mean())) < 3e-3 * float(r.mean())
    assert o.mean() > 0.01


def test_materials_path_bit_exact():
    # Spheres rest on the floor: a path through the exact contact point is a tie between
    # two primitives, which the reference resolves by test order and the oracle by lowest
    # id; after that the two paths diverge, so a pixel or two may differ.
    o, r, exact = compare(S.materials_scene(40, 30), 8, 8)
    assert exact >= 0.998, exact
    assert abs(float(o.mean()) - float(r.mean())) < 1e-3 * float(r.mean())


@pytest.mark.parametrize("job", [0, 9])
def test_fog_volpath_bit_exact(job):
    jobs = S.fog_jobs_c3(scale=0.05)
    desc = jobs[job]["scene"]
    o, r, exact = compare(desc, 4, 6, integrator=1)
    assert exact >= 0.999, exact
    assert o.mean() > 0.0


def test_moments_agree_with_render_mean():
    d = S.cornell_c0(24, 24)
    orc = O.OracleRender(d)
    s1, s2 = orc.moments(256, 8, seed=3)
    img = orc.render(256, 8)
    m = (s1 / 256).reshape(img.shape)
    var = (s2 / 256 - (s1 / 256) ** 2).reshape(img.shape)
    z = (img - m) / np.sqrt(2 * var / 256 + 1e-12)
    assert abs(z.mean()) < 0.2 and (np.abs(z) > 5).mean() < 0.01
