// Host-side set-up of one run: geometry, the internal GPU layout, the
// per-half-edge and per-face tables in the layout's order, the physical
// parameters of the kernels, and (mixed precision) the packed FP32 tables.
// One resident run (Engine) and an ensemble of runs (Ensemble) build their
// device data from it.
#pragma once

#include <cstdint>
#include <vector>

#include "engine_types.hpp"
#include "geometry.hpp"
#include "layout.hpp"
#include "mesh.hpp"

namespace splash {

struct RunParams {
    int32_t accretion_on = 0;
    int32_t friction_on = 1;
    double acc_rate = 3.25e-7;
    double e_acc = 0.02;
    double omega_acc_abs = 0.4;
    double tilt_angle = 0.4;
    double acc_width = 0.1;
    double alpha = 1e-6;
};

struct RunHost {
    Geometry geo;
    Layout L;
    int dim = 0, solver = 0;
    double area_coeff = 0, total_mass = 0;
    // internal half-edge order
    std::vector<int32_t> he_fv, sten;             // owner | valence << 27; stencil faces
    std::vector<double> beta, hsc, ec, en;         // [nh][4] each
    // internal face order
    std::vector<double> area, fcn4, acc;           // [nf], [nf][4] (centre, round_diff / area), [nf][8]
    std::vector<uint8_t> light_mask;               // [nf] (dim 5): bit k set when light-curve observer k sees the face
    PhysParams phys{};
    RunParams params;
};

// Checks the arguments (throws std::invalid_argument), builds the geometry
// with `threads` OpenMP threads and the layout with the given patch shape;
// with fallback_faces > 0 a layout whose face sets overflow at max_faces is
// built at fallback_faces instead.
RunHost build_run_host(const PolyMesh& mesh, const double* U_in, int dim, int solver, double gam, double omega_ns,
                       const RunParams& params, int32_t slots, int32_t max_faces, int threads,
                       int32_t fallback_faces = 0);

// Rebuilds the layout and every table in its order for another patch shape
// (the geometry and physical parameters are kept).
void relayout(RunHost& r, int32_t slots, int32_t max_faces);

// The four light-curve observers of get_light_curves: obs[4 k .. 4 k + 2]
// their positions, obs[4 k + 3] the norm; the readout forms a visible
// face's weight as |r . o| / |o| * A (r the unit face centre, A the area).
void light_observers(double obs[16]);

// FP32 tables of the mixed kernel, per extended slot row (block b's own then
// halo half-edges start at foff[f0] + halo_off[b]) and per edge of a block
// (a block holds both slots of each of its edges, so n_slots/2 edges from
// row (foff[f0] + halo_off[b]) / 2), the edge oriented for the slot that
// evaluates its flux:
//   slot_a [n_ext][2]: x = plus0 | plus1 << 10 | minus0 << 20 | negate-en << 30,
//                      y = minus1 | owner << 10 | edge << 20 | valence << 28
//                      (face-set locals)
//   slot_b [n_ext][4]: H+/H-, BM_dist/H+, beta+0, beta-0 (beta.1 = 1 - beta.0
//                      to 3e-13 on every generated mesh; H+/BM_dist is the
//                      reciprocal of the second entry)
//   edge32 [n_ext/2][8]: unit edge midpoint, arc length, normal, chord
//   edge_rec [n_ext/2]: evaluating slot | twin slot << 9 | lower mesh index << 18
//   rate [nf]: accretion rate profile, rounded to FP32 (the FP32 source is its only user)
//   inv_area [nf]: 1 / area (the limiter's Courant number is arc * dt / area)
//   fc32 [nf][4]: unit face centre rounded to FP32, the rate above (stage 1,
//                 whose result only feeds the FP32 input of stage 2)
struct MixedTables {
    std::vector<uint32_t> slot_a, edge_rec;
    std::vector<float> slot_b, edge32, inv_area, fc32;
    std::vector<float> rate;
    int32_t fset_max = 1;
    int32_t slot_max = 4;  // largest own + halo slot count of a patch, a multiple of 4
};
MixedTables pack_mixed_tables(const RunHost& r);

}  // namespace splash
