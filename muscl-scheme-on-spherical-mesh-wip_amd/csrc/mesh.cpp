// Spherical mesh generators with pmp-identical numbering.  Built with
// -ffp-contract=off: every coordinate is the same double the reference's
// GCC build computes.
#include "mesh.hpp"

#include <cmath>
#include <stdexcept>
#include <utility>

namespace splash {

namespace {
// pmp norm(): sqrt(((0 + x*x) + y*y) + z*z)   (mat_vec.h:512-525)
inline double pnorm(double x, double y, double z) {
    double s = 0.0;
    s += x * x;
    s += y * y;
    s += z * z;
    return std::sqrt(s);
}
}  // namespace

bool HalfedgeMesh::boundary_v(int v) const {
    int h = v_he_[v];
    return !(h != kInvalid && h_face_[h] != kInvalid);
}

int HalfedgeMesh::add_vertex(double x, double y, double z) {
    v_he_.push_back(kInvalid);
    pos_.push_back({x, y, z});
    return int(v_he_.size()) - 1;
}

// surface_mesh.h:1943-1963: h0 (even) points to `end`, h1 to `start`.
int HalfedgeMesh::new_edge(int start, int end) {
    int h0 = int(h_to_.size());
    h_to_.push_back(end);
    h_to_.push_back(start);
    for (auto* a : {&h_face_, &h_next_, &h_prev_}) {
        a->push_back(kInvalid);
        a->push_back(kInvalid);
    }
    return h0;
}

int HalfedgeMesh::new_face() {
    f_he_.push_back(kInvalid);
    return int(f_he_.size()) - 1;
}

// surface_mesh.cpp:172-190
int HalfedgeMesh::find_halfedge(int start, int end) const {
    int h = v_he_[start];
    const int first = h;
    if (h == kInvalid) return kInvalid;
    do {
        if (h_to_[h] == end) return h;
        h = cw(h);
    } while (h != first);
    return kInvalid;
}

// surface_mesh.cpp:198-214
void HalfedgeMesh::adjust_outgoing(int v) {
    int h = v_he_[v];
    const int first = h;
    if (h == kInvalid) return;
    do {
        if (boundary_h(h)) {
            v_he_[v] = h;
            return;
        }
        h = cw(h);
    } while (h != first);
}

// surface_mesh.cpp:244-423
int HalfedgeMesh::add_face(const int* vs, int n) {
    std::vector<int>& he = af_he_;
    std::vector<int>& fresh = af_fresh_;
    std::vector<int>& adjust = af_adjust_;
    std::vector<std::pair<int, int>>& links = af_links_;
    he.assign(size_t(n), 0);
    fresh.assign(size_t(n), 0);
    adjust.assign(size_t(n), 0);
    links.clear();
    links.reserve(6 * size_t(n));
    for (int i = 0, ii = 1; i < n; ++i, ii = (ii + 1) % n) {
        if (!boundary_v(vs[i])) throw std::runtime_error("add_face: complex vertex");
        he[i] = find_halfedge(vs[i], vs[ii]);
        fresh[i] = he[i] == kInvalid;
        if (!fresh[i] && !boundary_h(he[i])) throw std::runtime_error("add_face: complex edge");
    }
    // re-link patches where two consecutive existing halfedges are not adjacent
    for (int i = 0, ii = 1; i < n; ++i, ii = (ii + 1) % n) {
        if (fresh[i] || fresh[ii]) continue;
        int inner_prev = he[i], inner_next = he[ii];
        if (h_next_[inner_prev] == inner_next) continue;
        int boundary_prev = opp(inner_next);
        do {
            boundary_prev = opp(h_next_[boundary_prev]);
        } while (!boundary_h(boundary_prev) || boundary_prev == inner_prev);
        int boundary_next = h_next_[boundary_prev];
        if (boundary_next == inner_next) throw std::runtime_error("add_face: patch re-linking failed");
        links.emplace_back(boundary_prev, h_next_[inner_prev]);
        links.emplace_back(h_prev_[inner_next], boundary_next);
        links.emplace_back(inner_prev, inner_next);
    }
    for (int i = 0, ii = 1; i < n; ++i, ii = (ii + 1) % n)
        if (fresh[i]) he[i] = new_edge(vs[i], vs[ii]);

    int f = new_face();
    f_he_[f] = he[n - 1];

    for (int i = 0, ii = 1; i < n; ++i, ii = (ii + 1) % n) {
        int v = vs[ii];
        int inner_prev = he[i], inner_next = he[ii];
        int id = (fresh[i] ? 1 : 0) | (fresh[ii] ? 2 : 0);
        if (id) {
            int outer_prev = opp(inner_next);
            int outer_next = opp(inner_prev);
            if (id == 1) {
                links.emplace_back(h_prev_[inner_next], outer_next);
                v_he_[v] = outer_next;
            } else if (id == 2) {
                int boundary_next = h_next_[inner_prev];
                links.emplace_back(outer_prev, boundary_next);
                v_he_[v] = boundary_next;
            } else if (v_he_[v] == kInvalid) {
                v_he_[v] = outer_next;
                links.emplace_back(outer_prev, outer_next);
            } else {
                int boundary_next = v_he_[v];
                links.emplace_back(h_prev_[boundary_next], outer_next);
                links.emplace_back(outer_prev, boundary_next);
            }
            links.emplace_back(inner_prev, inner_next);
        } else {
            adjust[ii] = (v_he_[v] == inner_next);
        }
        h_face_[he[i]] = f;
    }
    for (auto& l : links) link(l.first, l.second);
    for (int i = 0; i < n; ++i)
        if (adjust[i]) adjust_outgoing(vs[i]);
    return f;
}

// surface_mesh.cpp:598-649
int HalfedgeMesh::insert_vertex(int h0, int v) {
    int h2 = h_next_[h0];
    int o0 = opp(h0);
    int o2 = h_prev_[o0];
    int v2 = h_to_[h0];
    int fh = h_face_[h0];
    int fo = h_face_[o0];
    int h1 = new_edge(v, v2);
    int o1 = opp(h1);
    link(h1, h2);
    link(h0, h1);
    h_to_[h0] = v;
    h_to_[h1] = v2;
    h_face_[h1] = fh;
    link(o1, o0);
    link(o2, o1);
    h_to_[o1] = v;
    h_face_[o1] = fo;
    v_he_[v2] = o1;
    adjust_outgoing(v2);
    v_he_[v] = h1;
    adjust_outgoing(v);
    if (fh != kInvalid) f_he_[fh] = h0;
    if (fo != kInvalid) f_he_[fo] = o1;
    return o1;
}

void HalfedgeMesh::insert_vertex_on_edge(int e, const double* p) {
    int v = add_vertex(p[0], p[1], p[2]);
    insert_vertex(e << 1, v);
}

// surface_mesh.cpp:651-685
int HalfedgeMesh::insert_edge(int h0, int h1) {
    int v0 = h_to_[h0], v1 = h_to_[h1];
    int h2 = h_next_[h0], h3 = h_next_[h1];
    int h4 = new_edge(v0, v1);
    int h5 = opp(h4);
    int f0 = h_face_[h0];
    int f1 = new_face();
    f_he_[f0] = h0;
    f_he_[f1] = h1;
    link(h0, h4);
    link(h4, h3);
    h_face_[h4] = f0;
    link(h1, h5);
    link(h5, h2);
    int h = h2;
    do {
        h_face_[h] = f1;
        h = h_next_[h];
    } while (h != h2);
    return h4;
}

// sph_gen.h:140-148: p = (1./|p|) * p
void HalfedgeMesh::project_to_unit_sphere() {
    for (auto& p : pos_) {
        double s = 1. / pnorm(p.x, p.y, p.z);
        p.x *= s;
        p.y *= s;
        p.z *= s;
    }
}

HalfedgeMesh::Vec HalfedgeMesh::centroid(int f) const {
    Vec c{0, 0, 0};
    double n = 0;
    int h = f_he_[f];
    const int first = h;
    do {
        const Vec& p = pos_[h_to_[h]];
        c.x += p.x;
        c.y += p.y;
        c.z += p.z;
        n += 1;
        h = h_next_[h];
    } while (h != first);
    c.x /= n;
    c.y /= n;
    c.z /= n;
    return c;
}

void HalfedgeMesh::reserve(int nv, int ne, int nf) {
    pos_.reserve(size_t(nv));
    v_he_.reserve(size_t(nv));
    for (std::vector<int>* a : {&h_to_, &h_face_, &h_next_, &h_prev_}) a->reserve(2 * size_t(ne));
    f_he_.reserve(size_t(nf));
}

// subdivision.cpp:180-342 (closed mesh without feature edges)
void HalfedgeMesh::loop_subdivision() {
    const int nv = n_vertices(), ne = n_halfedges() / 2, nf = n_faces();
    reserve(nv + ne, 2 * ne + 3 * nf, 4 * nf);  // each edge split, three new edges per triangle
    std::vector<Vec> vpoint(nv), epoint(ne);
    for (int v = 0; v < nv; ++v) {
        Vec p{0, 0, 0};
        double k = 0;
        int h = v_he_[v];
        const int first = h;
        do {
            const Vec& q = pos_[h_to_[h]];
            p.x += q.x;
            p.y += q.y;
            p.z += q.z;
            k += 1;
            h = ccw(h);
        } while (h != first);
        p.x /= k;
        p.y /= k;
        p.z /= k;
        const double beta = (0.625 - std::pow(0.375 + 0.25 * std::cos(2.0 * M_PI / k), 2.0));
        const double keep = 1.0 - beta;
        const Vec& x = pos_[v];
        vpoint[v] = {x.x * keep + p.x * beta, x.y * keep + p.y * beta, x.z * keep + p.z * beta};
    }
    for (int e = 0; e < ne; ++e) {
        const int h0 = 2 * e, h1 = 2 * e + 1;
        const Vec& a = pos_[h_to_[h0]];
        const Vec& b = pos_[h_to_[h1]];
        const Vec& c = pos_[h_to_[h_next_[h0]]];
        const Vec& d = pos_[h_to_[h_next_[h1]]];
        auto comp = [](double pa, double pb, double pc, double pd) {
            double p = pa;
            p += pb;
            p *= 3.0;
            p += pc;
            p += pd;
            p *= 0.125;
            return p;
        };
        epoint[e] = {comp(a.x, b.x, c.x, d.x), comp(a.y, b.y, c.y, d.y), comp(a.z, b.z, c.z, d.z)};
    }
    pos_ = vpoint;
    for (int e = 0; e < ne; ++e) {
        const double p[3] = {epoint[e].x, epoint[e].y, epoint[e].z};
        insert_vertex_on_edge(e, p);
    }
    for (int f = 0; f < nf; ++f) {
        int h = f_he_[f];
        insert_edge(h, h_next_[h_next_[h]]);
        h = h_next_[h];
        insert_edge(h, h_next_[h_next_[h]]);
        h = h_next_[h];
        insert_edge(h, h_next_[h_next_[h]]);
    }
}

// subdivision.cpp:9-178 (closed mesh without feature edges)
void HalfedgeMesh::catmull_clark_subdivision() {
    const int nv = n_vertices(), ne = n_halfedges() / 2, nf = n_faces();
    reserve(nv + ne + nf, 4 * ne, 2 * ne);  // closed mesh: the faces' corners number 2 ne
    std::vector<Vec> vpoint(nv), epoint(ne), fpoint(nf);
    for (int f = 0; f < nf; ++f) fpoint[f] = centroid(f);
    for (int e = 0; e < ne; ++e) {
        const Vec& a = pos_[h_to_[2 * e]];
        const Vec& b = pos_[h_to_[2 * e + 1]];
        const Vec& c = fpoint[h_face_[2 * e]];
        const Vec& d = fpoint[h_face_[2 * e + 1]];
        auto comp = [](double pa, double pb, double pc, double pd) {
            double p = 0;
            p += pa;
            p += pb;
            p += pc;
            p += pd;
            p *= 0.25;
            return p;
        };
        epoint[e] = {comp(a.x, b.x, c.x, d.x), comp(a.y, b.y, c.y, d.y), comp(a.z, b.z, c.z, d.z)};
    }
    for (int v = 0; v < nv; ++v) {
        const int first = v_he_[v];
        double k = 0;
        int h = first;
        do {
            k += 1;
            h = ccw(h);
        } while (h != first);
        Vec p{0, 0, 0};
        h = first;
        do {
            const Vec& q = pos_[h_to_[h]];
            p.x += q.x;
            p.y += q.y;
            p.z += q.z;
            h = ccw(h);
        } while (h != first);
        h = first;
        do {
            const Vec& q = fpoint[h_face_[h]];
            p.x += q.x;
            p.y += q.y;
            p.z += q.z;
            h = ccw(h);
        } while (h != first);
        const double kk = k * k;
        p.x /= kk;
        p.y /= kk;
        p.z /= kk;
        const double w = (k - 2.0) / k;
        const Vec& x = pos_[v];
        vpoint[v] = {p.x + x.x * w, p.y + x.y * w, p.z + x.z * w};
    }
    pos_ = vpoint;
    for (int e = 0; e < ne; ++e) {
        const double p[3] = {epoint[e].x, epoint[e].y, epoint[e].z};
        insert_vertex_on_edge(e, p);
    }
    for (int f = 0; f < nf; ++f) {
        const int h0 = f_he_[f];
        insert_edge(h0, h_next_[h_next_[h0]]);
        const int h1 = h_next_[h0];
        const double p[3] = {fpoint[f].x, fpoint[f].y, fpoint[f].z};
        insert_vertex_on_edge(h1 >> 1, p);
        int h = h_next_[h_next_[h_next_[h1]]];
        while (h != h0) {
            insert_edge(h1, h);
            h = h_next_[h_next_[h_next_[h1]]];
        }
    }
}

// sph_gen.h:72-96: one vertex per face centroid, one face per old vertex in
// the order of pmp's counter-clockwise face circulator, then assign.
void HalfedgeMesh::dual() {
    HalfedgeMesh out;
    out.reserve(n_faces(), n_halfedges() / 2, n_vertices());
    std::vector<int> fvertex(n_faces());
    for (int f = 0; f < n_faces(); ++f) {
        Vec c = centroid(f);
        fvertex[f] = out.add_vertex(c.x, c.y, c.z);
    }
    std::vector<int> ring;
    for (int v = 0; v < n_vertices(); ++v) {
        ring.clear();
        int h = v_he_[v];
        if (h != kInvalid && boundary_h(h)) {
            do h = ccw(h);
            while (boundary_h(h));
        }
        const int first = h;
        do {
            ring.push_back(fvertex[h_face_[h]]);
            do h = ccw(h);
            while (boundary_h(h));
        } while (h != first);
        out.add_face(ring.data(), int(ring.size()));
    }
    *this = std::move(out);
}

// MUSCL_geometry.hpp:66-88: vertices in index order, face vertex lists from
// to_vertex(halfedge(f)) along next_halfedge.
PolyMesh HalfedgeMesh::flatten() const {
    PolyMesh m;
    m.xyz.reserve(3 * pos_.size());
    for (const Vec& p : pos_) {
        m.xyz.push_back(p.x);
        m.xyz.push_back(p.y);
        m.xyz.push_back(p.z);
    }
    m.foff.reserve(f_he_.size() + 1);
    m.fv.reserve(h_to_.size() / 2);
    for (int f = 0; f < n_faces(); ++f) {
        m.foff.push_back(int32_t(m.fv.size()));
        int h = f_he_[f];
        const int first = h;
        do {
            m.fv.push_back(h_to_[h]);
            h = h_next_[h];
        } while (h != first);
    }
    m.foff.push_back(int32_t(m.fv.size()));
    return m;
}

namespace {

// sph_gen.h:158-205
HalfedgeMesh icosahedron() {
    HalfedgeMesh m;
    const double phi = (1. + std::sqrt(5.)) * 0.5;
    const double a = 1.0;
    const double b = 1.0 / phi;
    const double verts[12][3] = {{0, b, -a}, {b, a, 0},  {-b, a, 0}, {0, b, a},
                                 {0, -b, a}, {-a, 0, b}, {0, -b, -a}, {a, 0, -b},
                                 {a, 0, b},  {-a, 0, -b}, {b, -a, 0}, {-b, -a, 0}};
    for (const auto& v : verts) m.add_vertex(v[0], v[1], v[2]);
    m.project_to_unit_sphere();
    // triangles of sph_gen.h:183-202 with the 1-based vertex names shifted to ids
    const int tri[20][3] = {{2, 1, 0},  {1, 2, 3},  {5, 4, 3},  {4, 8, 3},  {7, 6, 0},
                            {6, 9, 0},  {11, 10, 4}, {10, 11, 6}, {9, 5, 2},  {5, 9, 11},
                            {8, 7, 1},  {7, 8, 10}, {2, 5, 3},  {8, 1, 3},  {9, 2, 0},
                            {1, 7, 0},  {11, 9, 6}, {7, 10, 6}, {5, 11, 4}, {10, 8, 4}};
    for (const auto& t : tri) m.add_face(t, 3);
    return m;
}

// sph_gen.h:31-57
HalfedgeMesh hexahedron() {
    HalfedgeMesh m;
    const double a = 1. / std::sqrt(3.);
    const double verts[8][3] = {{-a, -a, -a}, {a, -a, -a}, {a, a, -a}, {-a, a, -a},
                                {-a, -a, a},  {a, -a, a},  {a, a, a},  {-a, a, a}};
    for (const auto& v : verts) m.add_vertex(v[0], v[1], v[2]);
    const int quads[6][4] = {{3, 2, 1, 0}, {2, 6, 5, 1}, {5, 6, 7, 4},
                             {0, 4, 7, 3}, {3, 7, 6, 2}, {1, 5, 4, 0}};
    for (const auto& q : quads) m.add_face(q, 4);
    return m;
}

void check_level(int n) {
    if (n < 0 || n > 12) throw std::invalid_argument("subdivision level must be in [0, 12]");
}

}  // namespace

// sph_gen.h:215-265 (pole on +y)
PolyMesh uv_sphere(int n_slices, int n_stacks) {
    if (n_slices < 3 || n_stacks < 2) throw std::invalid_argument("uv_sphere needs n_slices>=3, n_stacks>=2");
    HalfedgeMesh m;
    const int v0 = m.add_vertex(0, 1, 0);
    for (int i = 0; i < n_stacks - 1; i++) {
        const double phi = M_PI * double(i + 1) / double(n_stacks);
        for (int j = 0; j < n_slices; j++) {
            const double theta = 2.0 * M_PI * double(j) / double(n_slices);
            m.add_vertex(std::sin(phi) * std::cos(theta), std::cos(phi), std::sin(phi) * std::sin(theta));
        }
    }
    const int v1 = m.add_vertex(0, -1, 0);
    for (int i = 0; i < n_slices; ++i) {
        const int top[3] = {v0, (i + 1) % n_slices + 1, i + 1};
        m.add_face(top, 3);
        const int base = n_slices * (n_stacks - 2) + 1;
        const int bottom[3] = {v1, i + base, (i + 1) % n_slices + base};
        m.add_face(bottom, 3);
    }
    for (int j = 0; j < n_stacks - 2; j++) {
        const int j0 = j * n_slices + 1, j1 = (j + 1) * n_slices + 1;
        for (int i = 0; i < n_slices; i++) {
            const int q[4] = {j0 + i, j0 + (i + 1) % n_slices, j1 + (i + 1) % n_slices, j1 + i};
            m.add_face(q, 4);
        }
    }
    return m.flatten();
}

// sph_gen.h:267-276
PolyMesh icosphere(int n_subdivisions) {
    check_level(n_subdivisions);
    HalfedgeMesh m = icosahedron();
    for (int i = 0; i < n_subdivisions; i++) {
        m.loop_subdivision();
        m.project_to_unit_sphere();
    }
    return m.flatten();
}

// sph_gen.h:278-284
PolyMesh icosphere_hex(int n_subdivisions) {
    check_level(n_subdivisions);
    HalfedgeMesh m = icosahedron();
    for (int i = 0; i < n_subdivisions; i++) {
        m.loop_subdivision();
        m.project_to_unit_sphere();
    }
    m.dual();
    m.project_to_unit_sphere();
    return m.flatten();
}

// sph_gen.h:286-295
PolyMesh quad_sphere(int n_subdivisions) {
    check_level(n_subdivisions);
    HalfedgeMesh m = hexahedron();
    for (int i = 0; i < n_subdivisions; i++) {
        m.catmull_clark_subdivision();
        m.project_to_unit_sphere();
    }
    return m.flatten();
}

}  // namespace splash
