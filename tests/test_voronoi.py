"""The Voronoi tessellation the engine library builds (storm_voronoi_build),
pinned on the CPU: a regular lattice of generators gives exactly the
Cartesian cells; on random generators every cell is the set of points
closest to its generator (brute force), the cells fill the box, each face
between two cells appears in both with opposite normals and equal areas, the
box faces tile the box's sides, and face polygons are counter-clockwise
about their outward normals with their centroid in their plane.
"""
from __future__ import annotations

import numpy as np
import pytest

from storm_amd import problems as P


def _plane_ok(m):
    for f in range(m.nfaces):
        v = m.vertices[m.face_vertex_start[f]:m.face_vertex_start[f + 1]]
        n = m.face_normal[f]
        assert abs(np.linalg.norm(n) - 1.0) < 1e-12
        # counter-clockwise seen from outside, vertices and centroid on one plane
        area = 0.5 * sum(np.dot(np.cross(v[i] - v[0], v[i + 1] - v[0]), n) for i in range(1, len(v) - 1))
        assert area == pytest.approx(m.face_area[f], rel=1e-9, abs=1e-15)
        assert np.all(np.abs((v - v[0]) @ n) < 1e-10)
        assert abs(np.dot(m.face_centroid[f] - v[0], n)) < 1e-10


def test_lattice_generators_give_the_cartesian_cells():
    n = 4
    g = (np.arange(n) + 0.5) / n
    pts = np.array([[x, y, z] for x in g for y in g for z in g])
    m = P.voronoi_mesh(pts)
    assert m.ncells == n ** 3
    assert np.all(np.diff(m.face_start) == 6)
    assert np.allclose(m.volume, 1.0 / n ** 3, rtol=1e-12)
    assert np.allclose(m.face_area, 1.0 / n ** 2, rtol=1e-12)
    # every normal is an axis, and the cell's generator sits at its centre
    assert np.allclose(np.sort(np.abs(m.face_normal), axis=1), [0, 0, 1])
    cof = m.cell_of_face()
    d = np.abs(np.sum((m.face_centroid - m.mesh_point[cof]) * m.face_normal, axis=1))
    assert np.allclose(d, 0.5 / n, rtol=1e-12)
    _plane_ok(m)


@pytest.mark.parametrize("n", [50, 700])
def test_random_tessellation_invariants(n):
    rng = np.random.default_rng(n)
    lo, hi = np.array([0.0, -0.5, 1.0]), np.array([2.0, 0.5, 1.5])
    pts = lo + (hi - lo) * rng.random((n, 3))
    m = P.voronoi_mesh(pts, lo, hi)
    assert m.ncells == n
    assert m.volume.sum() == pytest.approx(np.prod(hi - lo), rel=1e-12)
    # cells numbered in Z order of their generators: same points, a permutation
    assert np.allclose(np.sort(m.mesh_point[:, 0]), np.sort(pts[:, 0]))
    cof = m.cell_of_face()
    # reciprocity: (a -> b) and (b -> a) with opposite normals and equal areas
    pair = {}
    for f in range(m.nfaces):
        if m.face_neighbor[f] >= 0:
            pair[(cof[f], m.face_neighbor[f])] = f
    for (a, b), f in pair.items():
        g = pair[(b, a)]
        assert np.allclose(m.face_normal[f], -m.face_normal[g], atol=1e-12)
        assert m.face_area[f] == pytest.approx(m.face_area[g], rel=1e-9, abs=1e-14)
    # the box faces tile each side
    side_area = {0: (hi[1] - lo[1]) * (hi[2] - lo[2]), 2: (hi[0] - lo[0]) * (hi[2] - lo[2]),
                 4: (hi[0] - lo[0]) * (hi[1] - lo[1])}
    for s in range(6):
        a = m.face_area[m.face_neighbor == -1 - s].sum()
        assert a == pytest.approx(side_area[s - s % 2], rel=1e-12)
    # each bisector face lies halfway between the two generators
    for (a, b), f in pair.items():
        mid = 0.5 * (m.mesh_point[a] + m.mesh_point[b])
        assert abs(np.dot(m.face_centroid[f] - mid, m.face_normal[f])) < 1e-10
    _plane_ok(m)
    # brute force: a point lies in the cell of its nearest generator
    x = lo + (hi - lo) * rng.random((2000, 3))
    near = np.argmin(((x[:, None, :] - m.mesh_point[None, :, :]) ** 2).sum(-1), axis=1)
    for p, c in zip(x, near):
        f0, f1 = m.face_start[c], m.face_start[c + 1]
        s = np.sum((p - m.face_centroid[f0:f1]) * m.face_normal[f0:f1], axis=1)
        assert np.all(s <= 1e-12)


def test_bad_generators_are_rejected():
    with pytest.raises(RuntimeError, match="outside"):
        P.voronoi_mesh(np.array([[0.5, 0.5, 1.5]]))
    with pytest.raises(RuntimeError, match="repeated"):
        P.voronoi_mesh(np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0.2, 0.3, 0.4]]))
