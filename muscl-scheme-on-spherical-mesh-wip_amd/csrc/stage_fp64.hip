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
void flux_update(const StepArgs& a, int stage, hipStream_t s) {
    double* UP = static_cast<double*>(a.UP);
    if (stage == 1)
        k_flux_update<double, DIM, SOLVER, 1><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, UP, a.U, a.W);
    else
        k_flux_update<double, DIM, SOLVER, 2><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, UP, a.U, a.W);
}

}  // namespace

void enqueue_reconstruct(const StepArgs& a, hipStream_t s) {
    const int gh = (a.geom.nh + kBlock - 1) / kBlock;
    double* UP = static_cast<double*>(a.UP);
    if (a.phys.dim == 5) k_reconstruct<double, 5><<<gh, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.W, UP);
    else k_reconstruct<double, 4><<<gh, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.W, UP);
}

void enqueue_flux_update(const StepArgs& a, int stage, hipStream_t s) {
    switch (a.phys.solver) {
    case kHLLE: flux_update<4, kHLLE>(a, stage, s); break;
    case kHLLE_P: flux_update<5, kHLLE_P>(a, stage, s); break;
    case kHLLC: flux_update<5, kHLLC>(a, stage, s); break;
    default: flux_update<5, kHLLCPLUS>(a, stage, s); break;
    }
}

void enqueue_prepare(const StepArgs& a, hipStream_t s) {
    if (a.phys.dim == 5) k_prepare<5><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.U, a.W);
    else k_prepare<4><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.U, a.W);
}

void enqueue_finalize_dt(const StepArgs& a, double dt0, hipStream_t s) {
    if (a.phys.dim == 5) k_finalize_dt<5><<<1, kBlock, 0, s>>>(a.geom, a.phys, a.sc, dt0);
    else k_finalize_dt<4><<<1, kBlock, 0, s>>>(a.geom, a.phys, a.sc, dt0);
}


void enqueue_diag(const StepArgs& a, double* partial, hipStream_t s) {
    const int gf = (a.geom.nf + kBlock - 1) / kBlock;
    if (a.phys.dim == 5) k_diag_partial<5><<<gf, kBlock, 0, s>>>(a.geom, a.U, partial);
    else k_diag_partial<4><<<gf, kBlock, 0, s>>>(a.geom, a.U, partial);
}

int block_size() { return kBlock; }

}  // namespace fp64
}  // namespace splash
