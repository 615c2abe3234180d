// FP64 instantiation of the stage kernels.  This translation unit is built
// with -ffp-contract=off: its arithmetic follows the reference's operation
// order so the plain FP64 path can be compared with the CPU restatement.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/splash.h"
#include "engine_types.hpp"
#include "launch.hpp"

namespace splash {
namespace fp64 {
#include "stage_kernels.inc"

namespace {

template <int DIM, int SOLVER>
void stage(const StepArgs& a, int stage, hipStream_t s, const KernelEvents* ev) {
    if (stage == 1)
        launch(k_stage<double, DIM, SOLVER, 1>, a.geom.n_blocks, kBlock, 0, s, ev, a.geom, a.phys, a.sc, a.U, a.W,
               a.W2);
    else
        launch(k_stage<double, DIM, SOLVER, 2>, a.geom.n_blocks, kBlock, 0, s, ev, a.geom, a.phys, a.sc, a.U, a.W2,
               a.W);
}

}  // namespace

void enqueue_stage(const StepArgs& a, int st, hipStream_t s, const KernelEvents* ev) {
    switch (a.phys.solver) {
    case kHLLE: stage<4, kHLLE>(a, st, s, ev); break;
    case kHLLE_P: stage<5, kHLLE_P>(a, st, s, ev); break;
    case kHLLC: stage<5, kHLLC>(a, st, s, ev); break;
    default: stage<5, kHLLCPLUS>(a, st, s, ev); break;
    }
}

void enqueue_prepare(const StepArgs& a, hipStream_t s) {
    if (a.phys.dim == 5) k_prepare<5><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys_dev, a.U, a.W, a.W32);
    else k_prepare<4><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys_dev, a.U, a.W, a.W32);
}

void enqueue_finalize_dt(const StepArgs& a, double dt0, hipStream_t s, const KernelEvents* ev) {
    const int runs = a.geom.n_runs;
    if (a.phys.dim == 5)
        launch(k_finalize_dt<5>, runs * kFinalizeBlocks, kFinalizeThreads, 0, s, ev, a.geom, a.phys_dev, a.sc, dt0,
               a.dt0_dev);
    else launch(k_finalize_dt<4>, runs * kFinalizeBlocks, kFinalizeThreads, 0, s, ev, a.geom, a.phys_dev, a.sc, dt0,
                a.dt0_dev);
}


int readout_faces() { return kReadoutFaces; }
int readout_blocks(int nf) { return (nf + kReadoutFaces - 1) / kReadoutFaces; }

void enqueue_readout(const StepArgs& a, const uint8_t* light_mask, const double* light_obs, double* partial,
                     double* out, Scalars* sc_out, hipStream_t s) {
    enqueue_readout_runs(a, light_mask, light_obs, nullptr, nullptr, nullptr, readout_blocks(a.geom.nf), partial, out,
                         sc_out, s);
}

void enqueue_readout_runs(const StepArgs& a, const uint8_t* light_mask, const double* light_obs, const int2* seg,
                          const int* seg_run, const int* run_part, int n_part, double* partial, double* out,
                          Scalars* sc_out, hipStream_t s) {
    if (a.phys.dim == 5)
        k_readout_partial<5><<<n_part, kBlock, 0, s>>>(a.geom, a.phys_dev, a.U, light_mask, light_obs, seg, seg_run,
                                                       partial);
    else k_readout_partial<4><<<n_part, kBlock, 0, s>>>(a.geom, a.phys_dev, a.U, light_mask, light_obs, seg, seg_run,
                                                         partial);
    k_readout_final<<<run_part ? a.geom.n_runs : 1, kBlock, 0, s>>>(partial, run_part, n_part, out, a.sc, sc_out);
}

void enqueue_curl(const StepArgs& a, bool stepped, const double* ev, double* W_tmp, double* out, hipStream_t s) {
    const int gf = (a.geom.nf + kBlock - 1) / kBlock;
    const double* W = nullptr;
    if (stepped) {
        if (a.W2) {
            W = a.W2;
        } else {  // FP32 stage-2 input, widened
            if (a.phys.dim == 5) k_widen32<5><<<gf, kBlock, 0, s>>>(a.geom.nf, a.W32b, W_tmp);
            else k_widen32<4><<<gf, kBlock, 0, s>>>(a.geom.nf, a.W32b, W_tmp);
            W = W_tmp;
        }
    }
    if (a.phys.dim == 5) k_curl<5><<<gf, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.U, W, ev, out);
    else k_curl<4><<<gf, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.U, W, ev, out);
}

__global__ void k_set_scalar(double* p, double v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

void enqueue_set_scalar(double* p, double v, hipStream_t s) { k_set_scalar<<<1, 64, 0, s>>>(p, v); }

// ---- the allocation-free readout ----
// The per-workgroup partial sums of k_readout_partial, followed by the copy
// of the scalars that k_readout_final writes, in one fixed array per GPU
// (each GPU that loads the library holds its copy until the process ends):
// kScratchBlocks * kReadout + 8 doubles, 288 KiB.
constexpr int kScratchBlocks = 4096;
constexpr size_t kScratchDoubles = size_t(kScratchBlocks) * kReadout + 8;
static_assert(sizeof(Scalars) <= 8 * sizeof(double), "the scalars copy fits its slot");
static_assert(kScratchBlocks * kReadoutFaces == SPLASH_READOUT_NOALLOC_MAX_FACES, "the limit include/splash.h states");
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;
}  // namespace

// The device address of `device`'s copy of scratch, looked up once per
// device with that device current (the host-side name of a __device__ array
// is no device address), then kept.
static hipError_t scratch_address(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch[device] = static_cast<double*>(p);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

int readout_noalloc_max_faces() { return kScratchBlocks * kReadoutFaces; }

hipError_t enqueue_readout_noalloc(const StepArgs& a, const uint8_t* light_mask, const double* light_obs, int device,
                                   double* out, hipStream_t s) {
    if (readout_blocks(a.geom.nf) > kScratchBlocks) return hipErrorInvalidValue;
    double* part = nullptr;
    const hipError_t e = scratch_address(device, &part);
    if (e != hipSuccess) return e;
    Scalars* sc_copy = reinterpret_cast<Scalars*>(part + size_t(kScratchBlocks) * kReadout);
    enqueue_readout(a, light_mask, light_obs, part, out, sc_copy, s);
    return hipGetLastError();
}

int block_size() { return kBlock; }
int finalize_blocks() { return kFinalizeBlocks; }

}  // namespace fp64
}  // namespace splash
