// The drop-in constructors taking an existing pmp::SurfaceMesh.  pmp itself
// cannot be compiled in this image (its mat_vec.h needs Eigen), so a minimal
// class with the members of pmp's interface that from_pmp() uses (vertices(),
// faces(), position(Vertex), vertices(Face), handle idx()) stands in for it,
// filled from a generated mesh.
//   pmp_adapter flatten   from_pmp() reproduces the generator's arrays (no GPU)
//   pmp_adapter solve     MUSCL_HLLCplus / MUSCL_HLLE built from the pmp-style
//                         mesh step exactly as from the flattened one
// Exits 0 on success.
#include <cmath>
#include <iostream>
#include <string>
#include <vector>

#include "splash/MUSCL.hpp"

namespace mockpmp {
struct Vertex {
    int i;
    int idx() const { return i; }
};
struct Face {
    int i;
    int idx() const { return i; }
};
struct Point {
    double x[3];
    double operator[](int k) const { return x[k]; }
};
class SurfaceMesh {
public:
    explicit SurfaceMesh(const splash::SurfaceMesh& m) {
        const auto& xyz = m.vertices();
        for (size_t v = 0; v < m.n_vertices(); ++v) pts_.push_back(Point{{xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]}});
        const auto& fo = m.face_offsets();
        const auto& fv = m.face_vertices();
        for (size_t f = 0; f < m.n_faces(); ++f) {
            std::vector<Vertex> ring;
            for (int32_t k = fo[f]; k < fo[f + 1]; ++k) ring.push_back(Vertex{fv[size_t(k)]});
            faces_.push_back(ring);
        }
    }
    std::vector<Vertex> vertices() const {
        std::vector<Vertex> out;
        for (size_t v = 0; v < pts_.size(); ++v) out.push_back(Vertex{int(v)});
        return out;
    }
    std::vector<Face> faces() const {
        std::vector<Face> out;
        for (size_t f = 0; f < faces_.size(); ++f) out.push_back(Face{int(f)});
        return out;
    }
    const Point& position(Vertex v) const { return pts_[size_t(v.idx())]; }
    const std::vector<Vertex>& vertices(Face f) const { return faces_[size_t(f.idx())]; }
    size_t n_faces() const { return faces_.size(); }

private:
    std::vector<Point> pts_;
    std::vector<std::vector<Vertex>> faces_;
};
}  // namespace mockpmp

static_assert(splash::is_pmp_mesh<mockpmp::SurfaceMesh>::value, "pmp interface detected");
static_assert(!splash::is_pmp_mesh<splash::SurfaceMesh>::value, "the flat mesh is not taken for a pmp mesh");

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "flatten";
    bool ok = true;
    for (const splash::SurfaceMesh& m : {splash::icosphere_hex(3), splash::quad_sphere(3), splash::uv_sphere(16, 8)}) {
        const mockpmp::SurfaceMesh pm(m);
        const splash::SurfaceMesh back = splash::from_pmp(pm);
        const bool same = back.vertices() == m.vertices() && back.face_offsets() == m.face_offsets() &&
                          back.face_vertices() == m.face_vertices() && back.fingerprint() == m.fingerprint();
        std::cout << "flatten F=" << m.n_faces() << (same ? " equal" : " DIFFERENT") << "\n";
        ok = ok && same;
        if (mode != "solve") continue;
        const double gam = 2 - 1 / (4. / 3);
        std::vector<std::vector<double>> U(m.n_faces(), std::vector<double>(5));
        for (size_t f = 0; f < m.n_faces(); ++f)
            U[f] = {1.0 + 0.01 * double(f % 7), 0.01 * double(f % 3), -0.02, 0.005 * double(f % 5), 0.3};
        splash::MUSCL_HLLCplus a(pm, U, 5, gam, 0.0, false, 8), b(m, U, 5, gam, 0.0, false, 8);
        a.set_verbose(false);
        b.set_verbose(false);
        a.do_steps(1e-4, 10);
        b.do_steps(1e-4, 10);
        std::vector<std::vector<double>> U4(m.n_faces(), std::vector<double>(4));
        for (size_t f = 0; f < m.n_faces(); ++f) U4[f] = {1.0 + 0.01 * double(f % 5), 0, 0, 0};
        splash::MUSCL_HLLE c(pm, U4, 4, gam), d(m, U4, 4, gam);
        c.set_verbose(false);
        d.set_verbose(false);
        c.do_steps(1e-3, 10);
        d.do_steps(1e-3, 10);
        const bool equal = a.get_state() == b.get_state() && a.time() == b.time() && c.get_state() == d.get_state();
        std::cout << "solve F=" << m.n_faces() << (equal ? " equal" : " DIFFERENT") << "\n";
        ok = ok && equal;
    }
    return ok ? 0 : 1;
}
