// Host-side flattening of a spherical polygon mesh into the structure-of-
// arrays stencil tables the HIP kernels stream.  The values restate
// MUSCL_base_geometry (src/geometry/MUSCL_geometry.hpp:43-575, :981-1030)
// bit for bit; derived per-half-edge constants (unit edge midpoint, its
// recomputed norm, n x R, arc/area, chord) are evaluated here once with the
// exact expression the reference evaluates on every call.
#pragma once

#include <cstdint>
#include <vector>

#include "mesh.hpp"

namespace splash {

struct Geometry {
    int32_t nf = 0, nh = 0, nv = 0;
    std::vector<int32_t> foff;            // [nf+1]

    // per face
    std::vector<double> fc;               // [nf][3] vertex-mean centre (|fc| < 1)
    std::vector<double> normal;           // [nf][3]
    std::vector<double> area;             // [nf]
    std::vector<double> fcn;              // [nf][3] fc/|fc|

    // per half-edge h = foff[f] + j
    std::vector<int32_t> he_face;         // owning face
    std::vector<int32_t> nb_edge;         // face across the edge
    std::vector<int32_t> twin;            // half-edge of nb_edge pointing back
    std::vector<int32_t> ff_plus;         // [nh][2]
    std::vector<int32_t> ff_minus;        // [nh][2]
    std::vector<double> beta_plus;        // [nh][2]
    std::vector<double> beta_minus;       // [nh][2]
    std::vector<double> h_plus, h_minus, bm_dist;
    std::vector<double> en;               // [nh][3] outward edge normal
    std::vector<double> ec;               // [nh][3] unit edge midpoint ((a+b)/2 normalised)
    std::vector<double> ecn;              // [nh]    |ec| as recomputed by vector3d::norm
    std::vector<double> chord;            // [nh]    |v_j - v_j1|
    std::vector<double> ev;               // [nh][3] v_j - v_j1 (write_t_curl)
    std::vector<double> arc;              // [nh]    great-circle length of the edge

    double h0 = 0;                        // MUSCL_base.hpp:62-91
    double round_diff = 0;                // MUSCL_base.hpp:242
    int32_t max_valence = 0;
};

// Throws std::runtime_error when the mesh is not a closed manifold whose
// every edge has exactly one neighbour (the reference's assumption).
// threads <= 0: OpenMP's default team (every host thread unless
// OMP_NUM_THREADS says otherwise); the tables do not depend on it
Geometry build_geometry(const PolyMesh& mesh, int threads = 0);

}  // namespace splash
