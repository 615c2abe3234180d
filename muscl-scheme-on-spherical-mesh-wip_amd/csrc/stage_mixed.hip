// Mixed-precision instantiation of the stage kernels: reconstruction slopes,
// limiter and Riemann fluxes in FP32; state, RK combine, sources, dt
// reductions and geometry in FP64.
#include <hip/hip_runtime.h>

#include "engine_types.hpp"
#include "launch.hpp"

namespace splash {
namespace mixed {
#include "stage_kernels.inc"

namespace {

template <int DIM, int SOLVER>
void flux_update(const StepArgs& a, int stage, hipStream_t s) {
    float* UP = static_cast<float*>(a.UP);
    if (stage == 1)
        k_flux_update<float, DIM, SOLVER, 1><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, UP, a.U, a.W);
    else
        k_flux_update<float, DIM, SOLVER, 2><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, UP, a.U, a.W);
}

}  // namespace

void enqueue_reconstruct(const StepArgs& a, hipStream_t s) {
    const int gh = (a.geom.nh + kBlock - 1) / kBlock;
    float* UP = static_cast<float*>(a.UP);
    if (a.phys.dim == 5) k_reconstruct<float, 5><<<gh, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.W, UP);
    else k_reconstruct<float, 4><<<gh, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.W, UP);
}

void enqueue_flux_update(const StepArgs& a, int stage, hipStream_t s) {
    switch (a.phys.solver) {
    case kHLLE: flux_update<4, kHLLE>(a, stage, s); break;
    case kHLLE_P: flux_update<5, kHLLE_P>(a, stage, s); break;
    case kHLLC: flux_update<5, kHLLC>(a, stage, s); break;
    default: flux_update<5, kHLLCPLUS>(a, stage, s); break;
    }
}


}  // namespace mixed
}  // namespace splash
