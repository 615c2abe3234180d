// Plain-old-data types shared by the host engine and the HIP kernels.
#pragma once

#include <cstdint>

namespace splash {

enum SolverKind : int32_t { kHLLE = 0, kHLLE_P = 1, kHLLC = 2, kHLLCPLUS = 3 };
enum Precision : int32_t { kFP64 = 0, kMixed = 1 };

// Sticky error bits (the reference's stop_check causes).
enum StopCode : int32_t {
    kStopNaNFlux = 1,        // MUSCL_base.hpp:264-268
    kStopNaNWaveSpeed = 2,   // adiabatic.hpp:671-676
    kStopWaveBranch = 4,     // HLLC.hpp:114-119, HLLCplus.hpp:169-174, HLLE.hpp:51-55
    kStopNegativeDt = 8,     // adiabatic.hpp:573-577 (g_eff < 0 somewhere)
    kStopBadDim = 16,        // adiabatic.hpp:26-30, isothermal.hpp:18-22
};

// Device-resident per-run scalars.  The time step never leaves the GPU: the
// FP64 path's last kernel of a step leaves per-block partials of the
// CFL/friction limits (GeomPtrs::dt_partial) and a small kernel turns them
// into dt; the mixed path folds them into a few accumulator words
// (GeomPtrs::dt_acc) that its next stage-1 kernel reduces itself.
struct Scalars {
    double dt;                    // clamped dt of the current step
    double t;                     // simulation time
    double dt0;                   // requested dt
    unsigned long long steps;
    int32_t stop;                 // StopCode bits
    int32_t acc_buf;              // mixed path: the dt_acc buffer this step's stage 2 fills
    unsigned long long first_err; // (code << 40) | half-edge or face index, atomicMin
};

// Mixed path: the dt limits of a run are folded into kDtSlots accumulator
// slots (patch b into slot b % kDtSlots), each an order-preserving key pair
// {0.1 h0 / max(|v|, c_s), min friction dt} kept by atomic min; a run has
// two buffers of slots, one filled by a step's stage 2 while the other is
// read and then cleared by the next stage 1.  Each (run, buffer, slot) key
// pair has 4 KiB to itself, so no two entries share a cache line, memory
// channel or neighbourhood: measured on a 14,000-patch stage kernel, the 16
// slots packed into 256 bytes made it 40% slower, and the two buffers of a
// slot 256 bytes apart (slots 4 KiB apart) still cost 2% of a step, while
// 4 KiB entries cost nothing measurable.  128 KiB per run.
constexpr int kDtSlots = 16;
constexpr int kDtEntryWords = 512;  // 4 KiB
constexpr int kDtRunWords = 2 * kDtSlots * kDtEntryWords;

struct PhysParams {
    int32_t dim, solver, accretion_on, friction_on;
    double gam, omega_ns;
    double h0, round_diff;
    double e_acc, alpha;
    double dmdt_coef;   // -(fall_eff*acc_rate*4*pi)/total_mass (adiabatic.hpp:463)
    double om_acc[3];   // accretion angular velocity (0, sin(tilt), cos(tilt)) * omega_acc_abs
    // constants of the mixed-precision kernel's FP32 source terms
    double gam_h;       // gam / (gam - 1)
    double gam3d_rho;   // gam3d / (2 gam3d - 1), gam3d = 1 / (2 - gam)
    double cool_k;      // 12/5 k_m (3/4 c_sigma)^(1/4) of the cooling coefficient C
    // FP32 copies of what the mixed kernel's FP32 arithmetic uses, so its
    // waves neither divide nor convert the uniform FP64 parameters
    struct {
        float gam, inv_gm1, kq, gm1;  // gam, 1/(gam-1), (gam+1)/(2 gam), gam-1
        float gam_h, gam3d_rho, cool_k, e_acc, alpha, omega_ns, dmdt_coef;
        float om_acc[3];
    } f;
};

// Device tables in the internal face order (faces ordered by a coordinate
// bisection or a cube-map Morton curve of their centres; half-edges
// face-major in that order).  Every
// per-half-edge value is the one of the reference's [face][edge] table for
// the same half-edge.
struct GeomPtrs {
    int32_t nf, nh;
    const int32_t* foff;      // [nf+1]
    const int32_t* he_fv;     // [nh] owner face | valence << 27
    const int32_t* sten;      // [nh][4]: plus0, plus1, minus0, minus1 (internal face ids)
    const double* beta;       // [nh][4]: bp0, bp1, bm0, bm1
    const double* hsc;        // [nh][4]: H+, H-, BM_dist, arc/area of the owner
    const double* ec;         // [nh][4]: unit edge midpoint xyz, |ec| as recomputed
    const double* en;         // [nh][4]: edge normal xyz, chord length
    const int32_t* twin_slot; // [nh][2]: twin half-edge, its LDS slot in the owner's block
    const double* area;       // [nf]
    const double* fcn;        // [nf][4]: unit face centre xyz, round_diff / area
    const double* acc;        // [nf][8]: accretion rate, omega_acc x r (3), r x (omega_acc x r) (3), unused
    const int32_t* blk_face;  // [n_blocks+1] face range of each stage block
    const int32_t* halo_off;  // [n_blocks+1] range of each block's halo in halo_he
    const int32_t* halo_he;   // outside twins of each block's boundary half-edges
    // mixed-precision kernel only (null in FP64 runs).  Per extended slot
    // row (block b's own then halo half-edges start at foff[f0] +
    // halo_off[b]):
    //   slot_a [n_ext][2]: packed record x, y
    //     x = plus0 | plus1 << 10 | minus0 << 20 | negate-en << 30
    //     y = minus1 | owner << 10 | edge << 20 | valence << 28
    //     (face fields: face-set locals)
    //   slot_b [n_ext][4]: H+/H-, BM_dist/H+, beta+0, beta-0
    // and per edge of a block (n_slots/2 edges from row (foff[f0] +
    // halo_off[b]) / 2, oriented for the slot that evaluates the flux):
    //   edge32 [n_ext/2][8]: unit edge midpoint xyz, arc length, normal xyz, chord
    //   edge_rec [n_ext/2]: evaluating slot | twin slot << 9 | low << 18
    const uint32_t* slot_a;
    const float* slot_b;
    const float* edge32;
    const uint32_t* edge_rec;
    const int32_t* fset_off;  // [n_blocks+1]
    const int32_t* fset_face; // internal face ids, each block's own faces first
    const float* acc_rate;    // [nf] accretion rate profile (acc[.][0]) in FP32, for the FP32 source
    const float* inv_area;    // [nf] 1 / area (FP32)
    const float* fc32;        // [nf][4] unit face centre in FP32, accretion rate (the FP32 stage-1 epilogue)
    int32_t fset_max;         // largest face set (dynamic LDS of the mixed kernel)
    int32_t slot_max;         // largest patch slot count, a multiple of 4 (dynamic LDS)
    int32_t n_blocks;         // multiple of 8 (trailing blocks may be empty)
    double* dt_partial;       // [n_blocks][2] by logical block: max(|v|, c_s), min friction dt (FP64 path)
    double* dt_partial2;      // [n_runs][16][2]: the dt kernel's per-workgroup partials (FP64 path)
    unsigned* dt_ticket;      // [n_runs]: the dt kernel's arrival counters (0 between launches)
    unsigned long long* dt_acc;  // mixed path: [n_runs][2][kDtSlots][kDtEntryWords], a key pair per entry
    // ensembles: several runs' patches in one grid (null / 1 for one run)
    const int32_t* blk_run;   // [n_blocks] run of each logical block
    const int32_t* run_blk;   // [n_runs+1] block range of each run
    int32_t n_runs;
};

}  // namespace splash
