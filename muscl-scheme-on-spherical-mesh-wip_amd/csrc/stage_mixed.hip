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
void stage(const StepArgs& a, int stage, hipStream_t s) {
    if (stage == 1)
        k_stage<float, DIM, SOLVER, 1><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.U, a.W, a.W2);
    else
        k_stage<float, DIM, SOLVER, 2><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.U, a.W2, a.W);
}

}  // namespace

void enqueue_stage(const StepArgs& a, int st, hipStream_t s) {
    switch (a.phys.solver) {
    case kHLLE: stage<4, kHLLE>(a, st, s); break;
    case kHLLE_P: stage<5, kHLLE_P>(a, st, s); break;
    case kHLLC: stage<5, kHLLC>(a, st, s); break;
    default: stage<5, kHLLCPLUS>(a, st, s); break;
    }
}


}  // namespace mixed
}  // namespace splash
