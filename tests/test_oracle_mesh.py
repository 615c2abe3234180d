"""Pins of the oracle's mesh generators and geometry precompute.

Known-answer values measured on a build of the real reference (ΣA-4π on
quad_sphere(4), h0 on icosphere_hex(5), area_coeff on quad_sphere(4), face
counts) plus invariants and closed forms that an indexing, sign or operand
mistake would break.
"""
import math

import numpy as np
import pytest

import oracle_lib as O

GAM = 2 - 1 / (4 / 3)


def _edges(m):
    fo, fv = m.foff, m.fv
    a = fv
    nxt = np.empty_like(fv)
    for f in range(m.nf):
        s, e = fo[f], fo[f + 1]
        nxt[s:e] = np.roll(fv[s:e], -1)
    return a, nxt


@pytest.mark.parametrize("n,F", [(0, 12), (1, 42), (2, 162), (3, 642), (4, 2562), (5, 10242)])
def test_icosphere_hex_counts_and_euler(n, F):
    m = O.Mesh("hex", n)
    assert m.nf == F == 10 * 4 ** n + 2
    assert m.nv == 20 * 4 ** n
    assert m.nh == 6 * F - 12
    assert m.nv - m.nh // 2 + m.nf == 2
    val = np.diff(m.foff)
    assert np.count_nonzero(val == 5) == 12 and np.all((val == 5) | (val == 6))


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4])
def test_quad_sphere_counts(n):
    m = O.Mesh("quad", n)
    assert m.nf == 6 * 4 ** n and np.all(np.diff(m.foff) == 4)
    assert m.nv - m.nh // 2 + m.nf == 2


@pytest.mark.parametrize("nx,ny", [(16, 8), (64, 32), (5, 3)])
def test_uv_sphere_counts(nx, ny):
    m = O.Mesh("uv", nx, ny)
    assert m.nf == nx * ny
    val = np.diff(m.foff)
    assert np.count_nonzero(val == 3) == 2 * nx and np.count_nonzero(val == 4) == nx * (ny - 2)
    assert m.nv == nx * (ny - 1) + 2
    # pole on +y (sph_gen.h:220)
    assert np.array_equal(m.xyz[0], [0.0, 1.0, 0.0]) and np.array_equal(m.xyz[-1], [0.0, -1.0, 0.0])


@pytest.mark.parametrize("kind,args", [("hex", (3,)), ("quad", (3,)), ("uv", (16, 8)), ("ico", (2,))])
def test_vertices_on_unit_sphere_and_faces_outward(kind, args):
    m = O.Mesh(kind, *args)
    assert np.max(np.abs(np.linalg.norm(m.xyz, axis=1) - 1)) < 4e-16
    g = O.Geometry(m)
    # counter-clockwise faces seen from outside: polygon normal along the centre
    assert np.all(np.sum(g.normal * g.fc, axis=1) > 0)


def test_icosahedron_level0_closed_form():
    """icosphere(0) vertices are the golden-ratio icosahedron of sph_gen.h:162-178,
    projected, in insertion order."""
    m = O.Mesh("ico", 0)
    phi = (1 + math.sqrt(5)) / 2
    a, b = 1.0, 1 / phi
    raw = np.array([[0, b, -a], [b, a, 0], [-b, a, 0], [0, b, a], [0, -b, a], [-a, 0, b],
                    [0, -b, -a], [a, 0, -b], [a, 0, b], [-a, 0, -b], [b, -a, 0], [-b, -a, 0]])
    assert np.allclose(m.xyz, raw / np.linalg.norm(raw, axis=1)[:, None], atol=1e-15)
    # triangle vertex sets of sph_gen.h:183-202 (a face's circulator may start at any vertex)
    tri = [(3, 2, 1), (2, 3, 4), (6, 5, 4), (5, 9, 4), (8, 7, 1), (7, 10, 1), (12, 11, 5), (11, 12, 7),
           (10, 6, 3), (6, 10, 12), (9, 8, 2), (8, 9, 11), (3, 6, 4), (9, 2, 4), (10, 3, 1),
           (2, 8, 1), (12, 10, 7), (8, 11, 7), (6, 12, 5), (11, 9, 5)]
    for f, t in enumerate(tri):
        got = list(m.fv[m.foff[f]:m.foff[f + 1]])
        want = [x - 1 for x in t]
        k = got.index(want[0])
        assert got[k:] + got[:k] == want


def test_reference_known_answers():
    """Values a build of the real reference produced."""
    q4 = O.Mesh("quad", 4)
    g = O.Geometry(q4)
    assert abs(g.area.sum() - 4 * np.pi) == pytest.approx(1.7e-12, rel=0.05)
    h5 = O.Geometry(O.Mesh("hex", 5))
    assert round(h5.h0, 5) == 0.01476
    s = O.Solver(q4, np.ones((q4.nf, 5)), 5, O.HLLCPLUS, GAM, 0.0, O.make_params(acc_width=0.1))
    assert round(s.area_coeff(), 5) == 0.47729


def _unit(x):
    return x / np.linalg.norm(x, axis=-1)[..., None]


@pytest.mark.parametrize("kind,args", [("hex", (4,)), ("quad", (4,)), ("uv", (32, 16))])
def test_geometry_invariants(kind, args):
    m = O.Mesh(kind, *args)
    g = O.Geometry(m)
    a, b = _edges(m)
    va, vb = m.xyz[a], m.xyz[b]
    owner = np.repeat(np.arange(m.nf), np.diff(m.foff))
    # spherical areas tile the sphere and match the solid-angle formula of
    # Van Oosterom & Strackee for the (centre, v_j, v_j+1) fan
    c = _unit(g.fc)[owner]
    num = np.abs(np.einsum("ij,ij->i", c, np.cross(va, vb)))
    den = 1 + np.einsum("ij,ij->i", c, va) + np.einsum("ij,ij->i", va, vb) + np.einsum("ij,ij->i", vb, c)
    tri = 2 * np.arctan2(num, den)
    area = np.bincount(owner, weights=tri)
    assert np.max(np.abs(area - g.area) / g.area) < 1e-6
    assert abs(g.area.sum() - 4 * np.pi) < 1e-9
    # neighbour across each edge holds both edge vertices; twin is an involution
    assert np.all(g.twin[g.twin] == np.arange(m.nh))
    assert np.array_equal(owner[g.twin], g.nb_edge)
    assert np.array_equal(a[g.twin], b) and np.array_equal(b[g.twin], a)
    # edge normals: unit, tangent to the sphere at the edge midpoint, normal to
    # the edge, pointing away from the owner, antisymmetric bit for bit
    mid = _unit(va + vb)
    assert np.allclose(np.linalg.norm(g.en, axis=1), 1, atol=1e-15)
    assert np.max(np.abs(np.einsum("ij,ij->i", g.en, mid))) < 1e-14
    assert np.max(np.abs(np.einsum("ij,ij->i", g.en, va - vb))) < 1e-14
    assert np.all(np.einsum("ij,ij->i", g.en, (va + vb) / 2 - g.fc[owner]) > 0)
    assert np.array_equal(g.en[g.twin], -g.en)
    # BM_dist = great-circle distance centre -> edge midpoint
    bm = 2 * np.arcsin(np.linalg.norm(_unit(g.fc[owner]) - mid, axis=1) / 2)
    assert np.max(np.abs(bm - g.bm_dist)) < 1e-14
    # stencil weights and orientation: H+ faces downstream of the edge, H- upstream
    assert np.max(np.abs(g.beta_plus.sum(1) - 1)) < 1e-12
    assert np.max(np.abs(g.beta_minus.sum(1) - 1)) < 1e-12
    BM = (va + vb) / 2 - g.fc[owner]
    assert np.all(np.einsum("ij,ij->i", g.fc[g.ff_plus[:, 0]] - g.fc[owner], BM) > 0)
    assert np.all(np.einsum("ij,ij->i", g.fc[g.ff_minus[:, 0]] - g.fc[owner], BM) < 0)
    assert np.all(g.h_plus > 0) and np.all(g.h_minus > 0)
    # h0 is the minimum great-circle edge length
    arc = 2 * np.arcsin(np.linalg.norm(va - vb, axis=1) / 2)
    assert g.h0 == pytest.approx(arc.min(), rel=1e-15)
