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
void stages(const StepArgs& a, hipStream_t s) {
    const int gh = (a.geom.nh + kBlock - 1) / kBlock;
    float* UP = static_cast<float*>(a.UP);
    k_reconstruct<float, DIM><<<gh, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.W, UP);
    k_flux_update<float, DIM, SOLVER, 1><<<a.n_flux_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, UP, a.U, a.W);
    k_reconstruct<float, DIM><<<gh, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.W, UP);
    k_flux_update<float, DIM, SOLVER, 2><<<a.n_flux_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, UP, a.U, a.W);
}

}  // namespace

void enqueue_stages(const StepArgs& a, hipStream_t s) {
    switch (a.phys.solver) {
    case kHLLE: stages<4, kHLLE>(a, s); break;
    case kHLLE_P: stages<5, kHLLE_P>(a, s); break;
    case kHLLC: stages<5, kHLLC>(a, s); break;
    default: stages<5, kHLLCPLUS>(a, s); break;
    }
}

}  // namespace mixed
}  // namespace splash
