// Host-side spherical mesh construction for the SPLASH MUSCL step.
//
// HalfedgeMesh is a compact half-edge container whose operations reproduce
// the element numbering of the vendored pmp library that the reference
// builds its meshes with (src/pmp/surface_mesh.cpp add_face :244-423,
// insert_vertex :598-649, insert_edge :651-685), so face ids, the vertex
// order inside each face and every coordinate bit come out identical to the
// reference's generators (src/geometry/sph_gen.h).
#pragma once

#include <cstdint>
#include <vector>

namespace splash {

// Flattened polygon mesh: face f has vertices fv[foff[f] .. foff[f+1]-1].
struct PolyMesh {
    std::vector<double> xyz;      // [nv][3]
    std::vector<int32_t> foff;    // [nf+1]
    std::vector<int32_t> fv;      // [nh]
    int32_t n_vertices() const { return int32_t(xyz.size() / 3); }
    int32_t n_faces() const { return foff.empty() ? 0 : int32_t(foff.size() - 1); }
    int32_t n_halfedges() const { return foff.empty() ? 0 : foff.back(); }
};

class HalfedgeMesh {
public:
    static constexpr int kInvalid = -1;

    int add_vertex(double x, double y, double z);
    int add_face(const int* vs, int n);
    int insert_vertex(int h0, int v);
    int insert_edge(int h0, int h1);
    void insert_vertex_on_edge(int e, const double* p);

    void project_to_unit_sphere();
    void loop_subdivision();
    void catmull_clark_subdivision();
    void dual();

    PolyMesh flatten() const;

    int n_vertices() const { return int(v_he_.size()); }
    int n_faces() const { return int(f_he_.size()); }
    int n_halfedges() const { return int(h_to_.size()); }

private:
    // capacity for nv vertices, ne edges and nf faces, so that growing to a
    // known size does not copy the arrays repeatedly
    void reserve(int nv, int ne, int nf);
    struct Vec {
        double x, y, z;
    };
    static int opp(int h) { return (h & 1) ? h - 1 : h + 1; }
    int cw(int h) const { return h_next_[opp(h)]; }
    int ccw(int h) const { return opp(h_prev_[h]); }
    bool boundary_h(int h) const { return h_face_[h] == kInvalid; }
    bool boundary_v(int v) const;
    void link(int h, int nh) {
        h_next_[h] = nh;
        h_prev_[nh] = h;
    }
    int new_edge(int start, int end);
    int new_face();
    int find_halfedge(int start, int end) const;
    void adjust_outgoing(int v);
    Vec centroid(int f) const;

    std::vector<int> h_to_, h_face_, h_next_, h_prev_;
    std::vector<int> v_he_;
    std::vector<Vec> pos_;
    std::vector<int> f_he_;
    // add_face's per-call work lists, kept between calls so that building a
    // mesh of millions of faces does not allocate per face
    std::vector<int> af_he_, af_fresh_, af_adjust_;
    std::vector<std::pair<int, int>> af_links_;
};

// Generators (sph_gen.h:215-295); all return pmp-identical numbering.
PolyMesh uv_sphere(int n_slices, int n_stacks);
PolyMesh icosphere(int n_subdivisions);
PolyMesh icosphere_hex(int n_subdivisions);
PolyMesh quad_sphere(int n_subdivisions);

}  // namespace splash
