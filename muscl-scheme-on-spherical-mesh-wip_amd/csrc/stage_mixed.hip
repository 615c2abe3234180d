// Mixed-precision instantiation of the stage kernels: reconstruction slopes,
// limiter and Riemann fluxes in FP32; state, RK combine, sources, dt
// reductions and geometry in FP64.
#include <hip/hip_runtime.h>

#include "engine_types.hpp"
#include "launch.hpp"

namespace splash {
namespace mixed {

// FP32 math of the mixed path on the native v_sqrt / v_log / v_exp / v_rcp
// instructions (about 1 ulp each, far inside the path's 1e-5 error budget)
__device__ __forceinline__ float m_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float m_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float m_pow(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));  // x > 0 at every call site
}
__device__ __forceinline__ float m_tanh(float x) {
    // tanh|x| = 1 - 2 / (e^{2|x|} + 1), sign restored; saturates to 1 for large |x|
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * fabsf(x));  // 2/ln 2
    return copysignf(1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f), x);
}

#include "stage_kernels.inc"

// ----------------------------------------------------------------------
// The mixed-precision stage kernel.  It follows k_stage (same patches, LDS
// slots and face loop) but reads FP32 copies: the stage input as 8 floats
// per face (two float4 loads per stencil gather) and one float4 record per
// half-edge table, so a half-edge costs 5 vector loads instead of 10 and
// 25 scalar gathers.  Face areas, the state, the source terms, the dt
// reductions and the flux accumulation (compensated, Neumaier) stay FP64.

template <int DIM>
__device__ __forceinline__ void load_w32(const float* __restrict__ W32, int i, float* w) {
    if constexpr (DIM == 5) {
        const float4 a = reinterpret_cast<const float4*>(W32)[2 * size_t(i)];
        const float4 b = reinterpret_cast<const float4*>(W32)[2 * size_t(i) + 1];
        w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w; w[4] = b.x;
    } else {
        const float4 a = reinterpret_cast<const float4*>(W32)[size_t(i)];
        w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
    }
}

template <int DIM>
__device__ __forceinline__ void store_w32(float* __restrict__ W32, int i, const double* w) {
    if constexpr (DIM == 5) {
        reinterpret_cast<float4*>(W32)[2 * size_t(i)] = make_float4(float(w[0]), float(w[1]), float(w[2]), float(w[3]));
        reinterpret_cast<float4*>(W32)[2 * size_t(i) + 1] = make_float4(float(w[4]), 0.f, 0.f, 0.f);
    } else {
        reinterpret_cast<float4*>(W32)[size_t(i)] = make_float4(float(w[0]), float(w[1]), float(w[2]), float(w[3]));
    }
}

// find_U_edges for one half-edge in FP32 (MUSCL_base.hpp:375-433)
template <int DIM>
__device__ __forceinline__ void reconstruct32(const GeomPtrs& g, const PhysParams& P, float dt,
                                              const float* __restrict__ W32, int h, float* out) {
    const int fv = g.he_fv[h];
    const int i = fv & 0x07ffffff;
    const int valence = fv >> 27;
    const int4 st = reinterpret_cast<const int4*>(g.sten)[h];
    const float4 bt = reinterpret_cast<const float4*>(g.beta32)[h];
    const float4 hs = reinterpret_cast<const float4*>(g.hsc32)[h];  // 1/H+, 1/H-, BM, arc/area
    const float4 e4 = reinterpret_cast<const float4*>(g.ec32)[h];
    const float4 n4 = reinterpret_cast<const float4*>(g.en32)[h];
    float wi[DIM], a[DIM], b[DIM], c[DIM], d[DIM];
    load_w32<DIM>(W32, i, wi);
    load_w32<DIM>(W32, st.x, a);
    load_w32<DIM>(W32, st.y, b);
    load_w32<DIM>(W32, st.z, c);
    load_w32<DIM>(W32, st.w, d);
    float pp[DIM], r[DIM], lim[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        pp[k] = (bt.x * a[k] + bt.y * b[k] - wi[k]) * hs.x;
        const float pm = (wi[k] - (bt.z * c[k] + bt.w * d[k])) * hs.y;
        r[k] = m_div(pm, pp[k]);
    }
    const V3<float> ec = v3(e4.x, e4.y, e4.z);
    const V3<float> en = v3(n4.x, n4.y, n4.z);
    const float eta_m = __builtin_amdgcn_rcpf(hs.z * hs.y), eta_p = __builtin_amdgcn_rcpf(hs.z * hs.x);  // H-+/BM
    if constexpr (DIM == 5) {
        const V3<float> vel = cross(ec, v3(wi[1], wi[2], wi[3])) / (-wi[0] * e4.w);
        const float cs = m_sqrt(m_div(float(P.gam) * wi[4], wi[0]));
        const float nu = (cs + dot(vel, en)) * dt * hs.w;
        limiter_adiabatic<float, DIM>(r, nu, eta_m, eta_p, valence, lim);
    } else {
        const V3<float> vel = cross(ec, v3(wi[1], wi[2], wi[3])) / (-wi[0]);
        const float nu = (1.0f + dot(vel, en)) * dt * hs.w;
        limiter_isothermal<float, DIM>(r, nu, eta_m, valence, lim);
    }
#pragma unroll
    for (int k = 0; k < DIM; ++k) out[k] = wi[k] + pp[k] * lim[k] * hs.z;
    if constexpr (DIM == 5) {
        // E_fc (MUSCL_base.hpp:541-562)
        const V3<float> vel = cross(ec, v3(out[1], out[2], out[3])) / (-out[0]);
        out[4] = out[4] * float(1 / (P.gam - 1)) + out[0] * dot(vel, vel) * 0.5f;
    }
}

template <int DIM, int SOLVER, int STAGE>
__global__ __launch_bounds__(kBlock) void k_stage32(GeomPtrs g, PhysParams P, Scalars* __restrict__ sc,
                                                    double* __restrict__ U, const float* __restrict__ W32,
                                                    float* __restrict__ W32_out, const double* __restrict__ W,
                                                    double* __restrict__ W_out) {
    __shared__ float sUP[kBlock * DIM];
    __shared__ float sF[kBlock * DIM];
    const int b = xcd_block();
    const int f0 = g.blk_face[b], f1 = g.blk_face[b + 1];
    const int hbase = g.foff[f0];
    const int n_own = g.foff[f1] - hbase;
    const int hoff = g.halo_off[b];
    const int n_slots = n_own + (g.halo_off[b + 1] - hoff);
    const int t = threadIdx.x;
    const double dt = sc->dt;

    if (t < n_slots) {
        const int h = t < n_own ? hbase + t : g.halo_he[hoff + t - n_own];
        reconstruct32<DIM>(g, P, float(dt), W32, h, &sUP[t * DIM]);
    }
    __syncthreads();

    // one flux per edge: the block evaluates its task half-edges from the
    // edge's lower-index side and writes +F / -F into both own slots, so the
    // two faces of an edge exchange exactly opposite fluxes (mass, momentum
    // and energy conserved to round-off)
    const int toff = g.task_off[b];
    if (t < g.task_off[b + 1] - toff) {
        const int h = g.task_he[toff + t];
        const int sh = h - hbase;
        const int2 tw = reinterpret_cast<const int2*>(g.twin_slot)[h];
        const bool low = h < tw.x;
        const int he = low ? h : tw.x;
        const float4 e4 = reinterpret_cast<const float4*>(g.ec32)[he];
        const float4 n4 = reinterpret_cast<const float4*>(g.en32)[he];
        Edge<float> e;
        e.ec = v3(e4.x, e4.y, e4.z);
        e.ecn = e4.w;
        e.en = v3(n4.x, n4.y, n4.z);
        float uL[DIM], uR[DIM], F[DIM];
        const int sl = low ? sh : tw.y, sr = low ? tw.y : sh;
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            uL[k] = sUP[sl * DIM + k];
            uR[k] = sUP[sr * DIM + k];
        }
        const int err = flux_star<float, DIM, SOLVER>(float(P.gam), uL, uR, e, F);
        if (err) raise_stop(sc, err, h);
        if (low) {
#pragma unroll
            for (int k = 0; k < DIM; ++k) sF[sh * DIM + k] = F[k];
            if (tw.y < n_own)
#pragma unroll
                for (int k = 0; k < DIM; ++k) sF[tw.y * DIM + k] = -F[k];
        } else {
#pragma unroll
            for (int k = 0; k < DIM; ++k) sF[sh * DIM + k] = -F[k];
        }
    }
    __syncthreads();

    const int i = f0 + t;
    double vmax = 1e-8, dtx = 1e20;
    if (i < f1) {
        const double dt_here = (STAGE == 1) ? dt / 2 : dt;
        const V3<double> fcn = v3(g.fcn[4 * size_t(i)], g.fcn[4 * size_t(i) + 1], g.fcn[4 * size_t(i) + 2]);
        double res[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) res[k] = 0;
        if constexpr (DIM == 5) {
            if (P.accretion_on) {
                double w[5], src[5];
#pragma unroll
                for (int k = 0; k < 5; ++k) w[k] = W[size_t(i) * DIM + k];
                source_adiabatic(P, w, fcn, g.acc + 8 * size_t(i), src);
#pragma unroll
                for (int k = 0; k < 5; ++k) res[k] = dt_here * src[k];
            }
        }
        const double rd_dt_A = P.round_diff * dt_here / g.area[i];
        const int hb = g.foff[i], he = g.foff[i + 1];
        double sum[DIM], comp[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) sum[k] = comp[k] = 0;
        bool nan_flux = false;
        for (int hh = hb; hh < he; ++hh) {
            const double coef = rd_dt_A * double(reinterpret_cast<const float4*>(g.en32)[hh].w);
            const float* Fh = &sF[(hh - hbase) * DIM];
#pragma unroll
            for (int k = 0; k < DIM; ++k) {
                nan_flux |= __builtin_isnan(Fh[k]);
                const double x = coef * double(Fh[k]);
                const double s = sum[k] + x;
                comp[k] += (fabs(sum[k]) >= fabs(x)) ? (sum[k] - s) + x : (x - s) + sum[k];
                sum[k] = s;
            }
        }
        if (nan_flux) raise_stop(sc, splash::kStopNaNFlux, i);
        double un[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) un[k] = (res[k] - (sum[k] + comp[k])) + U[size_t(i) * DIM + k];
        if (un[0] < 1e-8) un[0] = 1e-8;
        double wn[DIM];
        face_terms<DIM>(P, un, fcn, wn, vmax, dtx);
        store_w32<DIM>(W32_out, i, wn);
        if (P.accretion_on) {
#pragma unroll
            for (int k = 0; k < DIM; ++k) W_out[size_t(i) * DIM + k] = wn[k];
        }
        if constexpr (STAGE == 2) {
#pragma unroll
            for (int k = 0; k < DIM; ++k) U[size_t(i) * DIM + k] = un[k];
        }
    }
    if constexpr (STAGE == 2) {
        reduce_dt_limits(g.dt_partial, vmax, dtx);
        if (blockIdx.x == 0 && t == 0) {
            sc->t += sc->dt;
            sc->steps += 1;
        }
    }
}

namespace {

template <int DIM, int SOLVER>
void stage(const StepArgs& a, int stage, hipStream_t s) {
    if (stage == 1)
        k_stage32<DIM, SOLVER, 1><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.U, a.W32, a.W32b, a.W,
                                                                     a.W2);
    else
        k_stage32<DIM, SOLVER, 2><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys, a.sc, a.U, a.W32b, a.W32, a.W2,
                                                                     a.W);
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
