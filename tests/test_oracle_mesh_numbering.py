"""Pins of the oracle's mesh numbering by derivation on tiny inputs.

The oracle's generators (oracle/oracle_mesh.c) restate pmp's half-edge
machinery step by step.  These tests do not: they state, as closed-form
rules, what that machinery yields after one subdivision or dual step, and
build the expected meshes from the level-0 vertex and face lists with those
rules alone.  The rules follow from the reference's code read as follows
(src/pmp/surface_mesh.cpp, src/pmp/algorithms/subdivision.cpp,
src/geometry/sph_gen.h):

R1 (add_face :244-423) a new face's first half-edge is the one entering its
   first vertex, so its vertex circulator starts at the first vertex given;
   an edge gets the next edge index the first time a face lists it, scanning
   the faces in order and each face's sides (v_i, v_i+1) in order; its
   half-edge 0 runs v_i -> v_i+1 inside that face.
R2 (insert_vertex :598-649) splitting edge e adds vertex nv + e (edges are
   split in index order) and leaves, in each face of e, the face's half-edge
   on the first half of e (in that face's orientation).  The new vertex's
   half-edge is its half-edge towards the end of half-edge 0; the end vertex
   of half-edge 0 gets the outgoing half-edge of the other face's side.
R3 (insert_edge :651-685) cuts off, as a new face numbered after all faces,
   the part that follows the first argument's next half-edge, starting the
   new face's circulator at the second argument's end vertex; the old face
   keeps the first argument as its half-edge.
R4 (faces around a vertex, surface_mesh.h ccw_rotated_halfedge) start at the
   face of the vertex's half-edge and go to the face holding the directed
   edge from the vertex to its predecessor in the current face.
R5 (dual, sph_gen.h:72-96) one vertex per face (vertex mean), then one face
   per vertex listing those face vertices in R4 order.

A dropped or swapped index in the oracle's half-edge restatement changes at
least one face list or vertex position below.
"""
import math

import numpy as np
import pytest

import oracle_lib as O

ICOSAHEDRON_TRIANGLES = [(3, 2, 1), (2, 3, 4), (6, 5, 4), (5, 9, 4), (8, 7, 1), (7, 10, 1), (12, 11, 5),
                         (11, 12, 7), (10, 6, 3), (6, 10, 12), (9, 8, 2), (8, 9, 11), (3, 6, 4), (9, 2, 4),
                         (10, 3, 1), (2, 8, 1), (12, 10, 7), (8, 11, 7), (6, 12, 5), (11, 9, 5)]
HEXAHEDRON_QUADS = [(3, 2, 1, 0), (2, 6, 5, 1), (5, 6, 7, 4), (0, 4, 7, 3), (3, 7, 6, 2), (1, 5, 4, 0)]


def _unit(p):
    p = np.asarray(p, dtype=float)
    return p / np.linalg.norm(p, axis=-1)[..., None]


def _icosahedron():
    phi = (1 + math.sqrt(5)) / 2
    a, b = 1.0, 1 / phi
    raw = np.array([[0, b, -a], [b, a, 0], [-b, a, 0], [0, b, a], [0, -b, a], [-a, 0, b],
                    [0, -b, -a], [a, 0, -b], [a, 0, b], [-a, 0, -b], [b, -a, 0], [-b, -a, 0]], dtype=float)
    return _unit(raw), [tuple(v - 1 for v in t) for t in ICOSAHEDRON_TRIANGLES]


def _hexahedron():
    a = 1 / math.sqrt(3)
    xyz = np.array([[-a, -a, -a], [a, -a, -a], [a, a, -a], [-a, a, -a],
                    [-a, -a, a], [a, -a, a], [a, a, a], [-a, a, a]])
    return xyz, list(HEXAHEDRON_QUADS)


def _edge_table(faces):
    """R1: edges in order of first listing, each as (start, end, face of its
    half-edge 0, face on the other side)."""
    index, edges = {}, []
    for f, fv in enumerate(faces):
        n = len(fv)
        for i in range(n):
            a, b = fv[i], fv[(i + 1) % n]
            key = frozenset((a, b))
            if key not in index:
                index[key] = len(edges)
                edges.append([a, b, f, None])
            else:
                edges[index[key]][3] = f
    assert all(e[3] is not None for e in edges)
    return edges


def _index(edges):
    return {frozenset((a, b)): k for k, (a, b, _, _) in enumerate(edges)}


def _neighbours(faces, nv):
    nb = [set() for _ in range(nv)]
    for fv in faces:
        n = len(fv)
        for i in range(n):
            nb[fv[i]].add(fv[(i + 1) % n])
            nb[fv[(i + 1) % n]].add(fv[i])
    return nb


def _rotate_to_last_split(fv, index):
    """The face's vertices starting at x, where (x, y) is its side of highest
    edge index (R2: the face's half-edge is then x -> m_xy)."""
    n = len(fv)
    sides = [index[frozenset((fv[i], fv[(i + 1) % n]))] for i in range(n)]
    k = int(np.argmax(sides))
    return [fv[(k + i) % n] for i in range(n)]


def _split_edges(edges, nv, corner):
    """R2 on every edge: the two halves keep / get edge indices e and ne + e
    (half-edge 0 of start -> m_e and m_e -> end, in the sub-faces at those
    corners of the creating face), and each vertex's half-edge face."""
    ne = len(edges)
    out = [None] * (2 * ne)
    start_face = {}
    for e, (a, b, fm, fo) in enumerate(edges):
        mv = nv + e
        out[e] = [a, mv, corner[(fm, a)], corner[(fo, a)]]
        out[ne + e] = [mv, b, corner[(fm, b)], corner[(fo, b)]]
        start_face[mv] = corner[(fm, b)]  # m_e -> end(e), in the creating face
        start_face[b] = corner[(fo, b)]   # end(e) -> m_e, in the other face (the last e wins)
    assert len(start_face) == nv + ne
    return out, start_face


def derive_loop(xyz, faces, edges=None):
    """One loop_subdivision (subdivision.cpp:178-300) plus projection, by
    R1-R3.  edges: the mesh's edge table (R1 for a mesh built face by face);
    returns positions, faces, the new edge table and each vertex's half-edge
    face (for R4)."""
    edges = _edge_table(faces) if edges is None else edges
    nv, nf, ne = len(xyz), len(faces), len(edges)
    index = _index(edges)
    nb = _neighbours(faces, nv)
    # positions from the old ones: Loop's vertex and edge rules
    new = np.zeros((nv + ne, 3))
    for v in range(nv):
        k = len(nb[v])
        p = xyz[sorted(nb[v])].sum(0) / k
        beta = 0.625 - (0.375 + 0.25 * math.cos(2 * math.pi / k)) ** 2
        new[v] = xyz[v] * (1 - beta) + beta * p

    def opposite_corner(f, a, b):
        return [v for v in faces[f] if v not in (a, b)][0]

    for e, (a, b, fm, fo) in enumerate(edges):
        c, d = opposite_corner(fm, a, b), opposite_corner(fo, a, b)
        new[nv + e] = (3 * (xyz[a] + xyz[b]) + xyz[c] + xyz[d]) * 0.125
    new = _unit(new)

    def m(a, b):
        return nv + index[frozenset((a, b))]

    out = [None] * (4 * nf)
    corner = {}  # (old face, corner vertex) -> sub-face
    inner = []
    for f, fv in enumerate(faces):
        x, y, z = _rotate_to_last_split(fv, index)
        out[nf + 3 * f] = (m(y, z), m(x, y), y)
        out[nf + 3 * f + 1] = (m(z, x), m(y, z), z)
        out[nf + 3 * f + 2] = (m(x, y), m(z, x), x)
        out[f] = (m(z, x), m(x, y), m(y, z))
        corner[(f, y)], corner[(f, z)], corner[(f, x)] = nf + 3 * f, nf + 3 * f + 1, nf + 3 * f + 2
        # R3: the cut edges, half-edge 0 inside the centre face
        inner += [[m(x, y), m(y, z), f, nf + 3 * f], [m(y, z), m(z, x), f, nf + 3 * f + 1],
                  [m(z, x), m(x, y), f, nf + 3 * f + 2]]
    halves, start_face = _split_edges(edges, nv, corner)
    return new, out, halves + inner, start_face


def derive_catmull_clark(xyz, faces, edges=None):
    """One catmull_clark_subdivision (subdivision.cpp:9-176) plus projection
    of a quad mesh, by R1-R3."""
    edges = _edge_table(faces) if edges is None else edges
    nv, nf, ne = len(xyz), len(faces), len(edges)
    index = _index(edges)
    nb = _neighbours(faces, nv)
    fp = np.array([xyz[list(fv)].mean(0) for fv in faces])
    new = np.zeros((nv + ne + nf, 3))
    for e, (a, b, fm, fo) in enumerate(edges):
        new[nv + e] = (xyz[a] + xyz[b] + fp[fm] + fp[fo]) * 0.25
    for v in range(nv):
        k = len(nb[v])
        adj = [f for f, fv in enumerate(faces) if v in fv]
        assert len(adj) == k
        p = (xyz[sorted(nb[v])].sum(0) + fp[adj].sum(0)) / (k * k)
        new[v] = p + ((k - 2.0) / k) * xyz[v]
    new[nv + ne:] = fp
    new = _unit(new)

    def m(a, b):
        return nv + index[frozenset((a, b))]

    out = [None] * (4 * nf)
    corner = {}
    inner = []
    for f, fv in enumerate(faces):
        x, y, z, w = _rotate_to_last_split(fv, index)
        c = nv + ne + f
        out[nf + 3 * f] = (c, m(x, y), y, m(y, z))
        out[nf + 3 * f + 1] = (m(z, w), c, m(y, z), z)
        out[nf + 3 * f + 2] = (m(w, x), c, m(z, w), w)
        out[f] = (c, m(w, x), x, m(x, y))
        corner[(f, x)], corner[(f, y)], corner[(f, z)], corner[(f, w)] = f, nf + 3 * f, nf + 3 * f + 1, nf + 3 * f + 2
        # R3 then R2 on the first cut: m_xy -> c, c -> m_yz, c -> m_zw, c -> m_wx
        inner += [[m(x, y), c, f, nf + 3 * f], [c, m(y, z), nf + 3 * f + 1, nf + 3 * f],
                  [c, m(z, w), nf + 3 * f + 2, nf + 3 * f + 1], [c, m(w, x), f, nf + 3 * f + 2]]
    halves, _ = _split_edges(edges, nv, corner)
    return new, out, halves + inner


def derive_dual(xyz, faces, start_face):
    """R4 + R5: the dual's vertices (face means, projected) and faces."""
    centres = _unit(np.array([xyz[list(fv)].mean(0) for fv in faces]))
    holder = {}
    for f, fv in enumerate(faces):
        n = len(fv)
        for i in range(n):
            holder[(fv[i], fv[(i + 1) % n])] = f
    out = []
    for v in range(len(xyz)):
        ring = [start_face[v]]
        while True:
            fv = faces[ring[-1]]
            pred = fv[fv.index(v) - 1]
            nxt = holder[(v, pred)]
            if nxt == ring[0]:
                break
            ring.append(nxt)
        out.append(tuple(ring))
    return centres, out


def _faces(m):
    return [tuple(int(x) for x in m.fv[m.foff[f]:m.foff[f + 1]]) for f in range(m.nf)]


def test_hexahedron_closed_form():
    """quad_sphere(0) is sph_gen.h:31-57's cube: corner coordinates +-1/sqrt3
    and the six quads, each circulated from its first listed vertex (R1)."""
    m = O.Mesh("quad", 0)
    xyz, faces = _hexahedron()
    assert np.array_equal(m.xyz, xyz)
    assert _faces(m) == faces


@pytest.mark.parametrize("ns,nt", [(5, 3), (16, 8), (7, 4)])
def test_uv_sphere_closed_form(ns, nt):
    """uv_sphere (sph_gen.h:215-265): pole (0,1,0), then ring i of n_slices
    vertices at polar angle pi(i+1)/n_stacks, then (0,-1,0); faces: per slice
    the top triangle then the bottom one, then the quads stack by stack."""
    m = O.Mesh("uv", ns, nt)
    pts = [(0.0, 1.0, 0.0)]
    for i in range(nt - 1):
        phi = math.pi * (i + 1) / nt
        for j in range(ns):
            th = 2.0 * math.pi * j / ns
            pts.append((math.sin(phi) * math.cos(th), math.cos(phi), math.sin(phi) * math.sin(th)))
    pts.append((0.0, -1.0, 0.0))
    assert np.max(np.abs(m.xyz - np.array(pts))) <= 2e-16
    bottom = len(pts) - 1
    want = []
    for i in range(ns):
        want.append((0, (i + 1) % ns + 1, i + 1))
        b0, b1 = i + ns * (nt - 2) + 1, (i + 1) % ns + ns * (nt - 2) + 1
        want.append((bottom, b0, b1))
    for j in range(nt - 2):
        j0, j1 = j * ns + 1, (j + 1) * ns + 1
        for i in range(ns):
            want.append((j0 + i, j0 + (i + 1) % ns, j1 + (i + 1) % ns, j1 + i))
    assert _faces(m) == want


def test_icosphere_level1_numbering():
    xyz0, faces0 = _icosahedron()
    xyz, faces, _, _ = derive_loop(xyz0, faces0)
    m = O.Mesh("ico", 1)
    assert (m.nv, m.nf) == (42, 80)
    assert _faces(m) == faces
    assert np.max(np.abs(m.xyz - xyz)) < 1e-15


def test_icosphere_level2_numbering():
    """The rules applied twice: level 1's faces already come out of a split,
    so this also pins edge numbering on a subdivided mesh."""
    xyz0, faces0 = _icosahedron()
    xyz1, faces1, edges1, _ = derive_loop(xyz0, faces0)
    xyz2, faces2, _, _ = derive_loop(xyz1, faces1, edges1)
    m = O.Mesh("ico", 2)
    assert _faces(m) == faces2
    assert np.max(np.abs(m.xyz - xyz2)) < 1e-15


@pytest.mark.parametrize("level", [1, 2, 3])
def test_icosphere_hex_numbering(level):
    """icosphere_hex(n) = dual(icosphere(n)), projected: face v of the dual
    lists the faces around vertex v of icosphere(n), starting at the face
    that holds v's half-edge."""
    xyz, faces = _icosahedron()
    edges = None
    for _ in range(level):
        xyz, faces, edges, start_face = derive_loop(xyz, faces, edges)
    dx, dfaces = derive_dual(xyz, faces, start_face)
    m = O.Mesh("hex", level)
    assert _faces(m) == dfaces
    assert np.max(np.abs(m.xyz - dx)) < 1e-15


def test_quad_sphere_level1_numbering():
    xyz0, faces0 = _hexahedron()
    xyz, faces, _ = derive_catmull_clark(xyz0, faces0)
    m = O.Mesh("quad", 1)
    assert (m.nv, m.nf) == (26, 24)
    assert _faces(m) == faces
    assert np.max(np.abs(m.xyz - xyz)) < 1e-15


def test_quad_sphere_level2_numbering():
    xyz0, faces0 = _hexahedron()
    xyz1, faces1, edges1 = derive_catmull_clark(xyz0, faces0)
    xyz2, faces2, _ = derive_catmull_clark(xyz1, faces1, edges1)
    m = O.Mesh("quad", 2)
    assert _faces(m) == faces2
    assert np.max(np.abs(m.xyz - xyz2)) < 1e-15


def test_derivation_detects_a_planted_mistake():
    """The derivation is sensitive to the numbering: swapping two edge
    indices of the icosahedron (as a mis-ordered edge scan would) changes the
    derived level-1 faces, and they then differ from the oracle's."""
    xyz0, faces0 = _icosahedron()
    _, good, _, _ = derive_loop(xyz0, faces0)
    edges = _edge_table(faces0)
    edges[3], edges[4] = edges[4], edges[3]  # a mis-ordered edge scan
    _, bad, _, _ = derive_loop(xyz0, faces0, edges)
    assert bad != good
