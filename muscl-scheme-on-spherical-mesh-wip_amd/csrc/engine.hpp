// One SPLASH MUSCL run resident on one GPU: geometry tables, state and the
// per-step launch sequence (optionally replayed from a captured HIP graph).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine_types.hpp"
#include "geometry.hpp"
#include "launch.hpp"
#include "mesh.hpp"
#include "run_setup.hpp"

namespace splash {

class HipError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class Engine {
public:
    // U_in is [nf][dim] in mesh face order (slot 4 = energy for dim 5).
    // Mixed precision: patch_faces = faces per patch of the stage kernel,
    // 1..64 (0 = 64), block_threads = threads per stage workgroup, 128, 192
    // or 256 (0 = 256), and patch_faces = -1 takes the fastest of a few
    // measured (threads, faces) candidates instead
    Engine(const PolyMesh& mesh, const double* U_in, int dim, int solver, double gam, double omega_ns,
           const RunParams& params, int precision, int geometry_threads = 0, int patch_faces = 0,
           int block_threads = 0);
    // from a prepared run, whose layout sets the patch shape
    Engine(const RunHost& r, const double* U_in, int precision);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void step(double dt0);              // enqueue one do_step
    // enqueue n do_steps: the first call launches them one by one and then
    // captures kGraphSteps steps into a HIP graph; later calls replay that
    // graph for each group of kGraphSteps and launch the rest one by one
    void steps(double dt0, int n);
    static constexpr int kGraphSteps = 8;
    // false: steps() never captures a graph (every step is launched eagerly)
    void set_use_graph(bool on) { use_graph_ = on; }
    void synchronize();
    // Runs n eager steps with HIP events around each launch; ms receives the
    // average duration of [finalize_dt, stage 1, stage 2, whole step] (the
    // mixed path has no dt kernel: its stage 1 forms dt, and ms[0] = 0).
    static constexpr int kKernelsPerStep = 3;
    void profile(double dt0, int n, double* ms);

    double time();                      // synchronises
    double last_dt();                   // synchronises
    uint64_t step_count();
    int stop_bits();
    uint64_t first_error();
    void get_state(double* U_out);      // [nf][dim], slot 4 = energy
    // resume: replaces the state (mesh face order), the time and the sticky
    // stop bits; the next step's dt comes from the new state
    void set_state(const double* U_in, double t, int stop_bits);
    void diagnostics(double out[5]);    // sum of (Sigma, L, E) * area
    void light_curves(double out[4]);   // get_light_curves (dim 5)
    // the same nine sums into device memory d_out[9] on stream s (null: the
    // run's stream), enqueued without waiting and without allocating
    void readout_noalloc(double* d_out, hipStream_t s);
    // write_t_curl's per-face values [nf] (mesh order) of the current state;
    // needs an Engine built from a mesh
    void curl(double* out);

    int32_t n_faces() const { return nf_; }
    int32_t n_halfedges() const { return nh_; }
    int dim() const { return dim_; }
    int precision() const { return precision_; }
    double area_coeff() const { return area_coeff_; }
    double total_mass() const { return total_mass_; }
    int32_t patch_faces() const { return patch_faces_; }
    int32_t block_threads() const { return block_threads_; }
    // the host geometry tables (empty for an Engine built from a RunHost)
    const Geometry& geometry() const { return geo_; }
    hipStream_t stream() const { return stream_; }
    const StepArgs& args() const { return args_; }

private:
    void enqueue_step(double dt0);
    void enqueue_prepare();
    // a: the launch arguments (default: args_, whose dt kernel takes dt0 by
    // value)
    void enqueue_kernel(int which, double dt0, const KernelEvents* ev = nullptr, const StepArgs* a = nullptr);
    void capture_graph();
    void init(const RunHost& r, const double* U_in);
    void upload(const RunHost& r);
    Scalars read_scalars();
    // the diagnostics sums, light curves and scalars of the current state,
    // read once per state: the driver's per-step queries share one readout
    void ensure_readout();

    Geometry geo_;
    int dim_, solver_, precision_;
    double gam_, omega_ns_;
    RunParams params_;
    double area_coeff_ = 0, total_mass_ = 0;
    int32_t patch_faces_ = 0;
    int32_t block_threads_ = 0;
    int32_t nf_ = 0, nh_ = 0;

    // internal layout: faces in bisection or curve order, half-edges face-major
    std::vector<int32_t> perm_;              // internal -> mesh face
    std::vector<int32_t> he_old_;            // internal -> mesh half-edge

    hipStream_t stream_ = nullptr;
    int device_ = 0;                    // the GPU the run lives on
    std::vector<void*> allocations_;
    StepArgs args_{};
    uint8_t* d_light_mask_ = nullptr;   // [nf] light-curve observer visibility bits (dim 5)
    double* d_light_obs_ = nullptr;     // [16] observer positions and norms
    double* d_readout_part_ = nullptr;  // [readout blocks][9] per-block sums
    // mapped pinned host memory that the readout kernel writes: the
    // diagnostics sums, the light curves and the scalars; readout_valid_
    // holds until the next launch or upload that changes the state
    struct Readout {
        double sums[9];
        Scalars sc;
    };
    Readout* h_readout_ = nullptr;
    Readout* d_readout_view_ = nullptr;  // its device address
    bool readout_valid_ = false;
    bool stepped_ = false;              // a step ran since the state was set
    double* d_curl_ = nullptr;          // lazily: ev [nh][3], out [nf], W_tmp [nf][dim]

    // kGraphSteps steps captured after the first steps() call; their dt
    // kernel reads dt0 from dt0_dev_, last set to graph_dt0_
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    double* dt0_dev_ = nullptr;
    double graph_dt0_ = 0;
    bool graph_dt0_set_ = false;
    bool use_graph_ = true;
};

// A patch shape of the mixed kernel: threads per workgroup (two LDS slots
// each) and the face cap of a patch.
struct PatchShape {
    int32_t threads, faces;
};

// The mixed-precision patch shape with the shortest measured step: each
// candidate re-lays out r, runs a trial Engine on its own copy of the state
// (time_trial: at least 20 ms of warm-up steps, then 32 timed ones at dt0)
// and r is left laid out for the winner, which is returned.
PatchShape autotune_patch_shape(RunHost& r, const double* U_in, double dt0);

// faces per patch the mixed kernel allows (two face waves of 64), and the
// candidates of the tuner
constexpr int32_t kPatchFacesMax = 128;
constexpr PatchShape kPatchShapeCandidates[] = {{256, 112}, {256, 96}, {256, 64}, {256, 48},
                                                {256, 32}, {192, 48}, {192, 32}, {128, 32}};

// The patch face cap when none is set: 112 (two face waves) for 256-thread
// workgroups on meshes of mean valence at most 4.5, whose 512 slots then
// hold about 110 quads; 64 (one face wave) otherwise, since hexagons fill
// the slots at about 70 faces and a second wave would update a handful.
inline int32_t default_patch_faces(const PolyMesh& m, int32_t threads) {
    const bool low_valence = 2 * int64_t(m.n_halfedges()) <= 9 * int64_t(m.n_faces());
    return threads == 256 && low_valence ? 112 : 64;
}
constexpr int32_t kPatchFacesOneWave = 64;

// Seconds for `timed` steps of a tuner trial, after untimed groups of 16
// steps (each followed by a wait) for at least warm_ms: after an idle gap
// the GPU needs about 20 ms of load to reach its working clock, so a fixed
// short warm-up would time the first candidates slow.  step(n) enqueues n
// steps, sync() waits for them.
template <class Step, class Sync>
double time_trial(Step&& step, Sync&& sync, int timed, double warm_ms = 20.0) {
    using clock = std::chrono::steady_clock;
    const auto w0 = clock::now();
    do {
        step(16);
        sync();
    } while (std::chrono::duration<double, std::milli>(clock::now() - w0).count() < warm_ms);
    const auto t0 = clock::now();
    step(timed);
    sync();
    return std::chrono::duration<double>(clock::now() - t0).count();
}

}  // namespace splash
