// Several independent SPLASH runs (uneven meshes, their own parameters and
// dt0) advanced together on one GPU.  Their patches are concatenated into
// one grid, so each RK stage is a single launch of the FP32-flux /
// FP64-state kernel for every run; the dt of each run is reduced over its
// own patches by one workgroup per run.  Each run evolves exactly as a
// single mixed-precision Engine on the same inputs would.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "engine.hpp"
#include "run_setup.hpp"

namespace splash {

class Ensemble {
public:
    struct Run {
        const PolyMesh* mesh;
        const double* U_in;  // [nf][dim] in the mesh's face order
        double gam, omega_ns;
        RunParams params;
    };
    // patch_faces as for Engine: 1..64, 0 = 64, -1 = the fastest of a few
    // candidates measured on trial ensembles
    Ensemble(const std::vector<Run>& runs, int dim, int solver, int geometry_threads = 0, int patch_faces = 0,
             int block_threads = 0);
    ~Ensemble();
    Ensemble(const Ensemble&) = delete;
    Ensemble& operator=(const Ensemble&) = delete;

    // dt0: [n_runs]; n steps: the first call launches them one by one and
    // then captures kGraphSteps steps into a HIP graph; later calls replay it
    // for each group of kGraphSteps and launch the rest one by one
    void steps(const double* dt0, int n);
    static constexpr int kGraphSteps = 8;
    // false: steps() never captures a graph
    void set_use_graph(bool on) { use_graph_ = on; }
    void synchronize();
    static constexpr int kKernelsPerStep = 3;
    // n back-to-back steps, each kernel timed with dispatch-bound HIP events:
    // ms = mean [dt finalize (0: stage 1 forms dt), stage 1, stage 2, whole step]
    void profile(const double* dt0, int n, double* ms);

    int n_runs() const { return int(faces_.size()); }
    int32_t n_faces(int r) const { return faces_.at(size_t(r)); }
    int64_t total_faces() const { return args_.geom.nf; }
    int64_t total_halfedges() const { return args_.geom.nh; }
    int32_t n_blocks() const { return args_.geom.n_blocks; }
    int32_t patch_faces() const { return patch_faces_; }
    int32_t block_threads() const { return block_threads_; }
    double time(int r);
    double last_dt(int r);
    int stop_bits(int r);
    void get_state(int r, double* U_out);
    // resume run r as Engine::set_state does: its state (mesh face order),
    // time and sticky stop bits; the other runs continue unchanged
    void set_state(int r, const double* U_in, double t, int stop_bits);
    // run r's do_step diagnostics sums and (dim 5) light curves, as
    // Engine::diagnostics / light_curves: one device readout covers every
    // run and serves until the next step
    void diagnostics(int r, double out[5]);
    void light_curves(int r, double out[4]);

private:
    Ensemble(const std::vector<RunHost>& hosts, const std::vector<const double*>& U_in, int dim, int solver);
    void init(const std::vector<RunHost>& hosts, const std::vector<const double*>& U_in);
    void upload_dt0(const double* dt0);
    void enqueue_step();
    void capture_graph();
    void enqueue_kernel(int which, const KernelEvents* ev = nullptr);
    Scalars read_scalars(int r);
    void ensure_readout();
    void* dev(size_t bytes, const void* host);

    int dim_, solver_;
    int32_t patch_faces_ = 0;
    int32_t block_threads_ = 0;
    std::vector<int32_t> faces_, face_off_;       // per run
    std::vector<std::vector<int32_t>> perm_;      // per run: internal -> mesh face
    std::vector<double> dt0_host_;

    hipStream_t stream_ = nullptr;
    std::vector<void*> allocations_;
    StepArgs args_{};
    double* dt0_dev_ = nullptr;
    // readout: per-workgroup face ranges and runs, each run's first
    // workgroup, per-workgroup sums, per-run results; the host copies
    // (sums [n_runs][9], then the scalars) stay valid until the next launch
    const int2* d_seg_ = nullptr;
    const int* d_seg_run_ = nullptr;
    const int* d_run_part_ = nullptr;
    int n_seg_ = 0;
    uint8_t* d_light_mask_ = nullptr;
    double* d_light_obs_ = nullptr;
    double* d_readout_part_ = nullptr;
    // mapped pinned host memory the readout kernel writes ([n_runs][9] sums,
    // [n_runs] scalars), and its device addresses
    double* h_sums_ = nullptr;
    Scalars* h_scalars_ = nullptr;
    double* d_sums_view_ = nullptr;
    Scalars* d_scalars_view_ = nullptr;
    bool readout_valid_ = false;
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;
    bool use_graph_ = true;
};

}  // namespace splash
