// FP64 instantiation of the stage kernels.  This translation unit is built
// with -ffp-contract=off: its arithmetic follows the reference's operation
// order so the plain FP64 path can be compared with the CPU restatement.
#include <hip/hip_runtime.h>

#include "engine_types.hpp"
#include "launch.hpp"

namespace splash {
namespace fp64 {
#include "stage_kernels.inc"

namespace {

template <int DIM, int SOLVER>
void stages(const StepArgs& a, hipStream_t s) {
    const int gh = (a.geom.nh + kBlock - 1) / kBlock;
    const double* W = a.W;
    double* Wm = a.W;
    double* UP = static_cast<double*>(a.UP);
    k_reconstruct<double, DIM><<<gh, kBlock, 0, s>>>(a.geom, a.phys, a.sc, W, UP);
    k_flux_update<double, DIM, SOLVER, 1><<<a.n_flux_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, UP, a.U, Wm);
    k_reconstruct<double, DIM><<<gh, kBlock, 0, s>>>(a.geom, a.phys, a.sc, W, UP);
    k_flux_update<double, DIM, SOLVER, 2><<<a.n_flux_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, UP, a.U, Wm);
}

}  // namespace

void enqueue_prepare(const StepArgs& a, hipStream_t s) {
    const int gf = (a.geom.nf + kBlock - 1) / kBlock;
    if (a.phys.dim == 5) k_prepare<5><<<gf, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.U, a.W);
    else k_prepare<4><<<gf, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.U, a.W);
}

void enqueue_finalize_dt(const StepArgs& a, double dt0, hipStream_t s) {
    if (a.phys.dim == 5) k_finalize_dt<5><<<1, 64, 0, s>>>(a.phys, a.sc, dt0);
    else k_finalize_dt<4><<<1, 64, 0, s>>>(a.phys, a.sc, dt0);
}

void enqueue_stages(const StepArgs& a, hipStream_t s) {
    switch (a.phys.solver) {
    case kHLLE: stages<4, kHLLE>(a, s); break;
    case kHLLE_P: stages<5, kHLLE_P>(a, s); break;
    case kHLLC: stages<5, kHLLC>(a, s); break;
    default: stages<5, kHLLCPLUS>(a, s); break;
    }
}

void enqueue_diag(const StepArgs& a, double* partial, hipStream_t s) {
    const int gf = (a.geom.nf + kBlock - 1) / kBlock;
    if (a.phys.dim == 5) k_diag_partial<5><<<gf, kBlock, 0, s>>>(a.geom, a.U, partial);
    else k_diag_partial<4><<<gf, kBlock, 0, s>>>(a.geom, a.U, partial);
}

int block_size() { return kBlock; }

}  // namespace fp64
}  // namespace splash
