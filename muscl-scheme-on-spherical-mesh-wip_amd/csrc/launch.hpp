// Launch interface between the host engine and the two kernel translation
// units (FP64 and mixed precision).
#pragma once

#include <hip/hip_runtime.h>

#include "engine_types.hpp"

namespace splash {

struct StepArgs {
    GeomPtrs geom;
    PhysParams phys;                // one run's parameters (FP64 stage kernel argument)
    const PhysParams* phys_dev;     // [n_runs] on the device
    const double* dt0_dev;          // [n_runs] per-run dt0 (ensembles), or null: the dt0 argument
    Scalars* sc;                    // [n_runs]
    double* U;  // [nf][dim] E-form state
    double* W;   // [nf][dim] p-form input of stage 1 (and output of stage 2)
    double* W2;  // [nf][dim] p-form input of stage 2
    float* W32;  // mixed runs: FP32 copy of the stage input, [nf] float4 (Sigma, L) then (dim 5) [nf] pressure
    float* W32b; // mixed runs: the other stage's FP32 input
    int block_threads;  // mixed runs: threads per stage-kernel workgroup (128, 192 or 256)
};

// Optional start/stop events of one launch: bound to the kernel's own
// dispatch (hipExtLaunchKernel), so they time the kernel alone.
struct KernelEvents {
    hipEvent_t start, stop;
};

namespace fp64 {
void enqueue_prepare(const StepArgs& a, hipStream_t s);
void enqueue_finalize_dt(const StepArgs& a, double dt0, hipStream_t s, const KernelEvents* ev = nullptr);
void enqueue_stage(const StepArgs& a, int stage, hipStream_t s, const KernelEvents* ev = nullptr);
// The diagnostics sums and (dim 5) the light curves of the current state
// into out[9] = {Sigma A, L A (3), E A, flux_0, flux_45, flux_90, flux_180},
// and a copy of the scalars into sc_out[1] (both device views of mapped host
// memory): two launches, with per-block sums in
// partial[readout_blocks(nf)][9].
// light_mask [nf] / light_obs [16]: the light-curve observers' visibility
// bits and positions (dim 5; unused for dim 4).
int readout_faces();  // faces per readout workgroup
int readout_blocks(int nf);
void enqueue_readout(const StepArgs& a, const uint8_t* light_mask, const double* light_obs, double* partial,
                     double* out, Scalars* sc_out, hipStream_t s);
// The same for every run of an ensemble: n_part workgroups, seg[b] the face
// range and seg_run[b] the run of workgroup b, run_part[r] the first
// workgroup of run r (run_part[n_runs] = n_part); out[n_runs][9],
// sc_out[n_runs].
void enqueue_readout_runs(const StepArgs& a, const uint8_t* light_mask, const double* light_obs, const int2* seg,
                          const int* seg_run, const int* run_part, int n_part, double* partial, double* out,
                          Scalars* sc_out, hipStream_t s);
// write_t_curl per face [nf] from the last stage-2 input (W2, or W32b widened
// into W_tmp [nf][dim]); stepped = false gives the value before any step.
// ev = v_j - v_j1 per half-edge [nh][3].
void enqueue_curl(const StepArgs& a, bool stepped, const double* ev, double* W_tmp, double* out, hipStream_t s);
int block_size();
int finalize_blocks();  // dt-kernel workgroups per run
// enqueue_readout with its per-workgroup partials (and the scalars copy) in
// the library's fixed per-GPU array instead of a caller buffer: out is a
// device array of 9 doubles.  Allocates nothing; returns
// hipErrorInvalidValue for a mesh of more than readout_noalloc_max_faces()
// faces.  Calls that overlap on one GPU share the array and race.
int readout_noalloc_max_faces();
hipError_t enqueue_readout_noalloc(const StepArgs& a, const uint8_t* light_mask, const double* light_obs, int device,
                                   double* out, hipStream_t s);
// *p = v, ordered on stream s (a one-thread kernel: no memset node in a graph)
void enqueue_set_scalar(double* p, double v, hipStream_t s);
}  // namespace fp64

namespace mixed {
// the stage inputs of the current state and its dt limits (which the next
// stage 1 turns into dt)
void enqueue_prepare(const StepArgs& a, hipStream_t s);
// one stage of a step: there is no separate dt kernel; stage 1 forms dt
// from dt0 (a.dt0_dev[run] when set, else the argument) and the limits the
// previous stage 2 or enqueue_prepare left
void enqueue_stage(const StepArgs& a, int stage, double dt0, hipStream_t s, const KernelEvents* ev = nullptr);
int slots(int threads);  // LDS half-edge slots per block of `threads` threads
int max_faces();         // faces per block
int default_threads();   // threads per block unless set or tuned
bool valid_threads(int threads);
}  // namespace mixed

}  // namespace splash
