"""Edge-rounded shading normals (RT/intersect.rs:196-233) in the oracle (CPU).

Pinned by the properties the reference states and tests for
shading_normal_near_edge: radius 0 returns the facet normal bit for bit
(intersect.rs:423-460 traces a whole grid both ways); away from every
other plane by more than the radius the facet normal is kept; at a shared
edge the normal is the bisector of the two facet normals, the same from
either side (no seam); the blend eases monotonically in between; and a
0.02 radius measurably changes the face-up image (> 0.5 % in Y,
intersect.rs:462-520).  The device follows the oracle path for path in
tests/test_film_gpu.py.
"""
import ctypes as C

import numpy as np

import oracle_lib
from indicatrix_amd import tracer

F3 = C.c_float * 3


def planes():
    return tracer.standard_round_brilliant()


def shade(pl, hit, facet, radius):
    n = pl[facet].normal
    out = F3()
    oracle_lib.lib().ixo_shading_normal_near_edge(pl, len(pl), F3(*hit), facet, F3(*n), radius, out)
    return np.array(out[:], np.float32)


def cube():
    import ctypes
    from indicatrix_amd import abi
    rows = [((1, 0, 0), -0.5), ((-1, 0, 0), -0.5), ((0, 1, 0), -0.5), ((0, -1, 0), -0.5),
            ((0, 0, 1), -0.5), ((0, 0, -1), -0.5)]
    arr = (abi.FacetPlane * 6)()
    for i, (n, d) in enumerate(rows):
        arr[i].normal[:] = n
        arr[i].d = d
    return arr


def test_radius_zero_and_far_from_edges_keep_the_facet_normal():
    pl = planes()
    table = np.array(pl[0].normal, np.float32)
    for r in [0.0, -1.0]:
        assert shade(pl, [0.3, 0.32, 0.1], 0, r).tobytes() == table.tobytes()
    # the table centre is 0.13 from the star facets: a 0.02 radius leaves it flat
    assert shade(pl, [0.0, 0.32, 0.0], 0, 0.02).tobytes() == table.tobytes()


def test_edge_normal_is_the_bisector_from_either_side():
    pl = cube()
    # a point on the edge shared by +x (0) and +y (2)
    edge = [0.5, 0.5, 0.1]
    a = shade(pl, edge, 0, 0.1)
    b = shade(pl, edge, 2, 0.1)
    bis = np.array([1.0, 1.0, 0.0], np.float32) / np.float32(np.sqrt(2.0))
    assert np.abs(a - bis).max() < 1e-6 and np.abs(b - bis).max() < 1e-6
    assert np.abs(a - b).max() < 1e-6


def test_blend_eases_monotonically_toward_the_edge():
    pl = cube()
    prev = 0.0
    for x in np.linspace(0.3, 0.5, 21):
        n = shade(pl, [0.5, float(x), 0.0], 0, 0.1)   # on +x, approaching the +y edge
        assert abs(np.linalg.norm(n.astype(np.float64)) - 1.0) < 1e-6
        assert n[1] >= prev - 1e-7
        prev = n[1]
    assert shade(pl, [0.5, 0.39, 0.0], 0, 0.1)[1] == 0.0       # 0.11 from the edge: flat
    assert abs(prev - np.float32(np.sqrt(0.5))) < 1e-6       # at the edge: the bisector


def _frame(radius, grid=16, spp=128):
    cam = tracer.camera_new(0.35, 0.28, 5.0, 18.0, grid, grid, spp)
    m = tracer.GemMaterial.diamond().to_abi()
    m.edge_rounding_radius = radius
    env = tracer.LightingPreset.studio("RingLights", 1.0, 0.85, 0.95)
    sc = tracer.FrameScene(cam, grid, grid, planes(), m, 12, env)
    return oracle_lib.accumulate(sc, 0, spp)


def test_rounding_measurably_changes_the_face_up_image():
    sharp = _frame(0.0)
    zero = _frame(-0.0)
    assert sharp.tobytes() == zero.tobytes()
    rounded = _frame(0.02)
    rel = abs(rounded[:, 1].sum() - sharp[:, 1].sum()) / sharp[:, 1].sum()
    assert rel > 0.005, rel
