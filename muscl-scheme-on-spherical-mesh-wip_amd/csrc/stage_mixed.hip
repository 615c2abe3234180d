// Mixed-precision instantiation of the stage kernels: reconstruction slopes,
// limiter and Riemann fluxes in FP32; state, RK combine, sources, dt
// reductions and geometry in FP64.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <stdexcept>

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

#include "stage_kernels.inc"

// ----------------------------------------------------------------------
// The mixed-precision stage kernel.  It follows k_stage (patches of faces,
// LDS slots for their half-edges and halo, one face loop) but works on FP32
// copies: the stage input as one float4 (Sigma, L) row per face plus (dim 5)
// one pressure per face, staged per block in LDS with 1 / area; per slot an
// 8-byte packed record (stencil face-set indices, owner, edge, valence) and a
// float4 {H+/H-, BM/H+, beta+0, beta-0}; per edge a float4 {ec, arc}, a
// float4 {en, chord} and a 4-byte slot record.  The state, the RK combine,
// the dt limits and reductions and the flux accumulation (Kahan-compensated
// where it updates the state) stay FP64; the accretion / cooling / friction
// sources run in FP32 on the FP32 stage input.

// Patches: at most kFaces faces (waves 0 and 1 update them, 64 each) whose
// own half-edges plus halo fit 2 * NT LDS slots, two slots per thread of an
// NT-thread workgroup.  NT is one of the block shapes below, chosen per run
// (or ensemble) by the caller or the tuner.
constexpr int kFaces = 128;
constexpr int kThreadChoices[] = {128, 192, 256};
static_assert(2 * 256 <= 512, "slot indices are packed in 9 bits");

// Element `idx` of a device table through a 32-bit byte offset (each table
// of the mixed path is below 4 GiB, checked at upload), so an access is a
// scalar base plus one 32-bit vector offset instead of 64-bit address math.
template <class T>
__device__ __forceinline__ T* row(T* base, uint32_t idx) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * uint32_t(sizeof(T)));
}
template <class T>
__device__ __forceinline__ const T* row(const T* base, uint32_t idx) {
    return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * uint32_t(sizeof(T)));
}

// The dt limits of a patch's two face waves, combined in LDS without a
// barrier: each wave folds its max(|v|, c_s) and min friction dt into the
// block's LDS words with workgroup-scope atomics, and the second wave to
// arrive folds the block's limits into its run's accumulator slot with
// device-scope atomic minima.  The CFL bound enters as the quotient
// 0.1 h0 / max(|v|, c_s) itself: a correctly rounded quotient is monotone in
// the divisor, so the minimum over patches equals the quotient of the
// maximum bit for bit.  vmax > 0 orders like its bit pattern; dtx (any sign,
// never NaN: the comparisons skip NaN candidates) and the quotient are mapped
// to unsigned keys of the same order.
struct DtCombine {
    unsigned long long v, x;
    unsigned arrived;
};
__device__ __forceinline__ unsigned long long order_key(double d) {
    const long long b = __double_as_longlong(d);
    return b >= 0 ? (unsigned long long)(b) | 0x8000000000000000ull : ~(unsigned long long)(b);
}
__device__ __forceinline__ double from_order_key(unsigned long long k) {
    return __longlong_as_double((long long)((k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k));
}
__device__ __forceinline__ void dt_combine_init(DtCombine* c) {
    c->v = (unsigned long long)(__double_as_longlong(1e-8));
    c->x = order_key(1e20);
    c->arrived = 0;
}
static_assert(kDtSlots == 16, "the first 16 lanes (one DPP row) reduce the dt slots");
// slot s of buffer buf of a run's dt accumulators: {CFL key, friction key}
__device__ __forceinline__ unsigned long long* dt_slot(const GeomPtrs& g, int run, int buf, int s) {
    return g.dt_acc + ((size_t(run) * 2 + size_t(buf)) * kDtSlots + size_t(s)) * kDtEntryWords;
}
// The minimum of a 64-bit key over each row of 16 lanes, through DPP lane
// moves (no LDS round trip): swaps within quads, then row rotations by 4
// and 8.  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_move(unsigned long long v) {
    const int lo = __builtin_amdgcn_mov_dpp(int(unsigned(v)), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(int(unsigned(v >> 32)), CTRL, 0xF, 0xF, false);
    return (unsigned long long)(unsigned(hi)) << 32 | unsigned(lo);
}
__device__ __forceinline__ unsigned long long row_min(unsigned long long v) {
    unsigned long long o = dpp_move<0xB1>(v);  // quad_perm [1,0,3,2]: lane ^ 1
    v = o < v ? o : v;
    o = dpp_move<0x4E>(v);  // quad_perm [2,3,0,1]: lane ^ 2
    v = o < v ? o : v;
    o = dpp_move<0x124>(v);  // row_ror:4
    v = o < v ? o : v;
    o = dpp_move<0x128>(v);  // row_ror:8
    return o < v ? o : v;
}
// lane 0 of each of the two face waves, once per block; acc = the block's
// accumulator slot
__device__ __forceinline__ void dt_combine(DtCombine* c, unsigned long long* acc, double h0, double vmax,
                                           double dtx) {
    __hip_atomic_fetch_max(&c->v, (unsigned long long)(__double_as_longlong(vmax)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_min(&c->x, order_key(dtx), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (__hip_atomic_fetch_add(&c->arrived, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 1u) {
        const double v =
            __longlong_as_double((long long)(__hip_atomic_load(&c->v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)));
        const unsigned long long kx = __hip_atomic_load(&c->x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_min(acc, order_key(h0 * 0.1 / v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_min(acc + 1, kx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int DIM>
__device__ __forceinline__ void store_w32(float* __restrict__ W32, int nf, int i, const double* w) {
    *row(reinterpret_cast<float4*>(W32), uint32_t(i)) = make_float4(float(w[0]), float(w[1]), float(w[2]), float(w[3]));
    if constexpr (DIM == 5) *row(W32, uint32_t(4 * nf + i)) = float(w[4]);
}

// h(r) of adiabatic.hpp:685-828 in FP32: 1 - tanh(a) for 0 < r < 1 with
// a = 4 r^2 (1-r)^0.1, tanh(a)^6 for r >= 1 with a = 10 (r-1)^3, else 0.
// Both arguments are >= 0, so tanh(a) = 1 - 2q with q = 1 / (e^{2a} + 1),
// and 2a/ln 2 is formed directly as the exponent of one exp2 (the constant
// 8/ln 2 of the first enters as log2(8/ln 2) under the power).
__device__ __forceinline__ float h_switch32(float r) {
    const bool mid = r < 1.0f && r > 0.0f;
    const bool high = r >= 1.0f;
    const float x = r - 1.0f;
    const float y = mid ? (r * r) * __builtin_amdgcn_exp2f(fmaf(0.1f, __builtin_amdgcn_logf(-x), 3.5287663729448977f))
                        : (x * x) * (x * 28.853900817779268f);
    const float q = __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(y) + 1.0f);
    const float t = fmaf(-2.0f, q, 1.0f), t2 = t * t;
    return mid ? 2.0f * q : (high ? t2 * t2 * t2 : 0.0f);
}

// The limiters of adiabatic.hpp:685-828 and isothermal.hpp:269-320 on
// r = rho * D (rho = H+/H-, D = (w - u-) / (u+ - w)), with e = H+/BM =
// eta_plus: eta_minus * r = e * D, and eta_minus = e / rho.  The reference
// sets the limiter of a NaN r to 0.  Here D = d- * rcp(d+) is NaN only when
// d+ = d- = 0, or when the state already holds NaN or inf: with a NaN D the
// hardware min / max give a finite limiter and h(NaN) = 0, and the slope
// d+ = 0 makes U+ = w exactly as a zero limiter does (a non-finite state
// stays non-finite either way), so the override needs no instructions.
template <int DIM>
__device__ __forceinline__ void limiter32(const float* D, float rho, float e, float nu, int valence, float* lim) {
    if constexpr (DIM == 5) {
        const float ed = e * __builtin_amdgcn_rcpf(float(2 * valence) * nu);  // eta_minus r / den = ed D
        const float c3 = (1.0f + nu) * (1.0f / 3.0f);
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            const float rk = rho * D[k];
            const float sb = lmax(0.0f, lmax(lmin(1.0f, D[k] * ed), lmin(rk, e)));
            const float to = lmax(0.0f, lmin(sb, fmaf(c3, rk - 1.0f, 1.0f)));
            lim[k] = fmaf(h_switch32(rk), sb - to, to);  // (1-h) third_order + h superbee
        }
    } else {
        const float ed = e * __builtin_amdgcn_rcpf(float(valence) * nu);  // 2 eta r / den = ed D
        const float eta = e * __builtin_amdgcn_rcpf(rho);
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            const float rk = rho * D[k];
            lim[k] = lmax(0.0f, lmax(lmin(1.0f, D[k] * ed), lmin(rk, eta)));
        }
    }
}

// find_U_edges for one half-edge in FP32 (MUSCL_base.hpp:375-433) from the
// block's LDS face states (sW: Sigma, L; sP: pressure; sC: sound speed; sIA:
// 1 / area) and the step's dt32: li = owner, plus0, plus1, minus0, minus1 local indices; hs =
// {H+/H-, BM/H+, beta+0, beta-0} (each pair's second weight is 1 - the
// first); arc = the edge's arc length; ec = the unit edge midpoint; en = the
// half-edge's outward normal.
// The slopes stay raw differences d+ = u+ - w, d- = w - u-: the reference's
// r = (d-/H-) / (d+/H+) is (H+/H-) d-/d+ and its U+ = w + (d+/H+) lim BM is
// w + d+ (BM/H+) lim.
template <int DIM>
__device__ __forceinline__ void reconstruct_lds(const float4* sW, const float* sP, const float* sC,
                                                const float* sIA, float dt32, const int* li, int valence, float4 hs,
                                                float arc, V3<float> ec, V3<float> en, float* out) {
    const float bp = hs.z, bm = hs.w;
    float w[5][DIM];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const float4 x = sW[li[q]];
        w[q][0] = x.x; w[q][1] = x.y; w[q][2] = x.z; w[q][3] = x.w;
        if constexpr (DIM == 5) w[q][4] = sP[li[q]];
    }
    const float* wi = w[0];
    float dp[DIM], D[DIM], lim[DIM];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
        dp[k] = fmaf(bp, w[1][k] - w[2][k], w[2][k]) - wi[k];
        const float dm = wi[k] - fmaf(bm, w[3][k] - w[4][k], w[4][k]);
        D[k] = dm * __builtin_amdgcn_rcpf(dp[k]);
    }
    // v . en with v = (ec x L) / (-Sigma |ec|) (dim 5; |ec| is 1 in FP32) or
    // (ec x L) / -Sigma
    const float cl = dot(cross(ec, v3(wi[1], wi[2], wi[3])), en);
    const float e = __builtin_amdgcn_rcpf(hs.y);  // H+/BM
    const float cfl = arc * (dt32 * sIA[li[0]]);  // dt arc / area
    if constexpr (DIM == 5) {
        const float nu = (sC[li[0]] - cl * __builtin_amdgcn_rcpf(wi[0])) * cfl;
        limiter32<DIM>(D, hs.x, e, nu, valence, lim);
    } else {
        const float nu = (1.0f - cl * __builtin_amdgcn_rcpf(wi[0])) * cfl;
        limiter32<DIM>(D, hs.x, e, nu, valence, lim);
    }
    // dim 5: slot 4 stays the limited pressure; flux32 forms E_fc from it
#pragma unroll
    for (int k = 0; k < DIM; ++k) out[k] = fmaf(dp[k] * hs.y, lim[k], wi[k]);
}

// flux_star in FP32 (isothermal.hpp:141-260, HLLE_p.hpp:34-55, HLLC.hpp:
// 18-122, HLLCplus.hpp:17-177): the formulas, branches and stop codes of the
// FP64 flux_star, with |ec| = 1 (the FP64 unit midpoint's norm rounds to
// exactly 1.0f, checked when the tables are packed, so the divisions by it
// drop out), one reciprocal of each side's Sigma, Sigma |v|^2 without the
// square root, and HLLC+'s velocity-jump factors formed before the star
// flux so the two side velocities are dead by then.  ec, en: unit edge
// midpoint and the normal of the left state's face; kq = (gam+1)/(2 gam),
// inv_gm1 = 1/(gam-1).  Adiabatic states arrive with slot 4 = the limited
// pressure p+: the reference converts it to E at the edge midpoint (E_fc,
// MUSCL_base.hpp:541-562) and pressure() converts back with the same
// velocity, so p = max(1e-16, p+) and E = p+/(gam-1) + Sigma |v|^2 / 2 are
// formed here from p+ directly.
template <int DIM, int SOLVER>
__device__ __forceinline__ int flux32(float gam, float kq, float inv_gm1, const float* uL, const float* uR,
                                      V3<float> ec, V3<float> en, float* F) {
    const float iL = __builtin_amdgcn_rcpf(uL[0]), iR = __builtin_amdgcn_rcpf(uR[0]);
    const V3<float> vL = cross(ec, v3(uL[1], uL[2], uL[3])) * (-iL);  // (ec x L) / -Sigma
    const V3<float> vR = cross(ec, v3(uR[1], uR[2], uR[3])) * (-iR);
    const float VL = dot(en, vL), VR = dot(en, vR);
    if constexpr (SOLVER == splash::kHLLE) {
        // sound speed 1, PI = Sigma
        const float SL = smin(VL, VR) - 1.0f, SR = smax(VL, VR) + 1.0f;
        const V3<float> nxR = cross(en, ec);
        const float FL[4] = {uL[0] * VL, uL[1] * VL - nxR.x * uL[0], uL[2] * VL - nxR.y * uL[0],
                             uL[3] * VL - nxR.z * uL[0]};
        const float FR[4] = {uR[0] * VR, uR[1] * VR - nxR.x * uR[0], uR[2] * VR - nxR.y * uR[0],
                             uR[3] * VR - nxR.z * uR[0]};
        return hlle_combine<float, 4>(SL, SR, uL, uR, FL, FR, F) ? splash::kStopWaveBranch : 0;
    } else {
        int err = 0;
        const float pL = smax(1e-16f, uL[4]), pR = smax(1e-16f, uR[4]);
        const float eL[5] = {uL[0], uL[1], uL[2], uL[3], fmaf(uL[4], inv_gm1, uL[0] * dot(vL, vL) * 0.5f)};
        const float eR[5] = {uR[0], uR[1], uR[2], uR[3], fmaf(uR[4], inv_gm1, uR[0] * dot(vR, vR) * 0.5f)};
        const float aL = m_sqrt(gam * pL * iL), aR = m_sqrt(gam * pR * iR);
        // PVRS wave speeds (adiabatic.hpp:594-683)
        const float p_star = smax(0.0f, 0.5f * (pL + pR) - 0.125f * (VR - VL) * (uL[0] + uR[0]) * (aL + aR));
        float qL = 1.0f, qR = 1.0f;
        if (!(p_star <= pR)) {
            qR = m_sqrt(1.0f + kq * (p_star * __builtin_amdgcn_rcpf(pR) - 1.0f));
            if (m_isnan(qR) || m_isinf(qR)) qR = 1.0f;
        }
        if (!(p_star <= pL)) {
            qL = m_sqrt(1.0f + kq * (p_star * __builtin_amdgcn_rcpf(pL) - 1.0f));
            if (m_isnan(qL) || m_isinf(qL)) qL = 1.0f;
        }
        const float SL = VL - aL * qL, SR = VR + aR * qR;
        if (m_isnan(SL) || m_isnan(SR)) err |= splash::kStopNaNWaveSpeed;
        if constexpr (SOLVER == splash::kHLLE_P) {
            const V3<float> nxR = cross(en, ec);
            float FL[5], FR[5];
            flux_adiabatic(eL, vL, pL, nxR, Edge<float>{ec, 1.0f, en}, FL);
            flux_adiabatic(eR, vR, pR, nxR, Edge<float>{ec, 1.0f, en}, FR);
            if (hlle_combine<float, 5>(SL, SR, eL, eR, FL, FR, F)) err |= splash::kStopWaveBranch;
            return err;
        } else {
            const float phi_L = uL[0] * (SL - VL), phi_R = uR[0] * (SR - VR);
            float S_star = (pR - pL + phi_L * VL - phi_R * VR) * __builtin_amdgcn_rcpf(phi_L - phi_R);
            if (m_isnan(S_star) || m_isinf(S_star)) S_star = (VL + VR) * 0.5f;
            const float P_LR = (pL + pR + phi_L * (S_star - VL) + phi_R * (S_star - VR)) * 0.5f;
            // HLLCplus.hpp:123-165 factors: f* - 1, 1 - (p ratio)^M, the
            // velocity jump, V_R - V_L
            float fm1 = 0.0f, g = 0.0f;
            V3<float> dvel = v3(0.0f, 0.0f, 0.0f);
            if constexpr (SOLVER == splash::kHLLCPLUS) {
                const float M = smin(1.0f, smax(m_sqrt(dot(vL, vL)) * __builtin_amdgcn_rcpf(aL),
                                                m_sqrt(dot(vR, vR)) * __builtin_amdgcn_rcpf(aR)));
                const bool lambd = ((VL - aL) > 0.0f && (VR - aR) < 0.0f) || ((VL + aL) > 0.0f && (VR + aR) < 0.0f) ||
                                   ((VR - aR) > 0.0f && (VL - aL) < 0.0f) || ((VR + aR) > 0.0f && (VL + aL) < 0.0f);
                const float omm = 1.0f - M * M;
                const float f_star = lambd ? 1.0f : M * m_sqrt(4.0f + omm * omm) * __builtin_amdgcn_rcpf(1.0f + M * M);
                fm1 = f_star - 1.0f;
                const float hh = smin(pL * __builtin_amdgcn_rcpf(pR), pR * __builtin_amdgcn_rcpf(pL));
                g = 1.0f - m_pow(hh, M);
                dvel = vR - vL;
            }
            // branch of HLLC.hpp:98-119
            int side;  // 0: F_L, 1: F_L*, 2: F_R*, 3: F_R
            if (SL >= 0.0f) side = 0;
            else if (SL < 0.0f && S_star >= 0.0f) side = 1;
            else if (S_star < 0.0f && SR >= 0.0f) side = 2;
            else if (SR < 0.0f) side = 3;
            else {
                side = 3;
                err |= splash::kStopWaveBranch;
            }
            const bool left = side <= 1;
            float u[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) u[k] = left ? eL[k] : eR[k];
            const V3<float> nxR = cross(en, ec);
            float FK[5];
            flux_adiabatic(u, left ? vL : vR, left ? pL : pR, nxR, Edge<float>{ec, 1.0f, en}, FK);
            if (side == 0 || side == 3) {
#pragma unroll
                for (int k = 0; k < 5; ++k) F[k] = FK[k];
                return err;
            }
            const float S_K = left ? SL : SR;
            const float iK = __builtin_amdgcn_rcpf(S_K - S_star);
            const float Dv[5] = {0.0f, -nxR.x, -nxR.y, -nxR.z, S_star};
#pragma unroll
            for (int k = 0; k < 5; ++k) F[k] = (S_star * (S_K * u[k] - FK[k]) + S_K * P_LR * Dv[k]) * iK;
            if constexpr (SOLVER == splash::kHLLCPLUS) {
                const float dV = VR - VL;
                const float coef = S_K * iK * g;
                const float fd = fm1 * dV;
                const V3<float> vff = v3(en.x * fd + (dvel.x - en.x * dV) * coef, en.y * fd + (dvel.y - en.y * dV) * coef,
                                         en.z * fd + (dvel.z - en.z * dV) * coef);
                const V3<float> am = cross(ec, vff);
                const float pp = phi_L * phi_R * __builtin_amdgcn_rcpf(phi_R - phi_L);
                F[1] += am.x * pp;
                F[2] += am.y * pp;
                F[3] += am.z * pp;
                F[4] += fm1 * dV * S_star * pp;
            }
            return err;
        }
    }
}

// The source terms of adiabatic.hpp:396-541 (as source_adiabatic) in FP32
// on the face's FP32 stage input (Sigma, L, p): accretion with omega_acc x r
// and r x (omega_acc x r) formed from the face centre, radiative cooling and
// friction.  Cooling uses C = cool_k Sigma/p (g_eff Sigma)^(1/4), the
// reference's 12/5 k_m Sigma/p (3/4 c_sigma g_eff Sigma)^(1/4) with the
// constant factor taken out so nothing leaves FP32 range, and works with
// 1 - beta directly (the reference clamps beta to [1e-8, 1 - 1e-8] and maps
// NaN to 1e-8; in FP32 1 - 1e-8 rounds to 1).  Each term is scaled by dt
// (<= 1e-4) and added to the FP64 state, so FP32 rounding enters at ~1e-11
// of the state per step.
__device__ __forceinline__ void source32(const PhysParams& P, float4 x, float p, V3<float> r, float rate,
                                         float* res) {
    const float sig = x.x;
    const float inv0 = __builtin_amdgcn_rcpf(sig), invp = __builtin_amdgcn_rcpf(p);
    const V3<float> L = v3(x.y, x.z, x.w);
    const V3<float> vel = cross(r, L) * (-inv0);
    const V3<float> omxr = cross(v3(P.f.om_acc[0], P.f.om_acc[1], P.f.om_acc[2]), r);
    const V3<float> rxo = cross(r, omxr);
    res[0] = rate;
    res[1] = rate * rxo.x;
    res[2] = rate * rxo.y;
    res[3] = rate * rxo.z;
    const V3<float> dvv = omxr - vel;
    res[4] = rate * (P.f.e_acc + dot(dvv, dvv) * 0.5f);
    const float vn2 = dot(vel, vel);
    const float h_spec = P.f.gam_h * p * inv0;  // gam/(gam-1) p/Sigma
    if (!P.friction_on) {
        const V3<float> rxv = cross(r, vel);
        const float dmdt = P.f.dmdt_coef * sig;
        res[0] += dmdt;
        res[1] += dmdt * rxv.x;
        res[2] += dmdt * rxv.y;
        res[3] += dmdt * rxv.z;
        res[4] += dmdt * (h_spec + vn2 * 0.5f);
    }
    const float g_eff = float(kGM) - vn2;
    const float C = P.f.cool_k * sig * invp * m_sqrt(m_sqrt(g_eff * sig));
    constexpr float beta_switch = 0.736194670678821f;
    constexpr float C_switch = -1 - 1 / (beta_switch - 1);
    float omb;  // 1 - beta
    if (C <= C_switch) {
        omb = __builtin_amdgcn_rcpf(1 + C);
    } else {
        const float q = 2 * __builtin_amdgcn_rcpf(C), q2 = q * q;
        omb = q2 * q2;
    }
    if (!(omb <= 1 - 1e-8f)) omb = 1 - 1e-8f;  // beta below the floor, or NaN
    if (omb < 1e-8f) omb = 1e-8f;              // beta above the ceiling
    res[4] -= g_eff * float(1 / 3.4e6) * omb;  // kappa
    if (P.friction_on) {
        const float rho = P.f.gam3d_rho * g_eff * sig * sig * invp;
        const V3<float> vel0 = cross(v3(0.0f, 0.0f, P.f.omega_ns), r);
        const V3<float> dv0 = vel - vel0;
        const float a = P.f.alpha * rho * m_sqrt(dot(dv0, dv0));
        const V3<float> wr = dv0 * (-a);
        const V3<float> l_fr = cross(r, wr);
        const float a0 = a * inv0;
        res[0] -= a;
        res[1] += l_fr.x - a0 * L.x;
        res[2] += l_fr.y - a0 * L.y;
        res[3] += l_fr.z - a0 * L.z;
        res[4] += dot(wr, vel0) - a * (vn2 * 0.5f + h_spec);
    }
}

// pressure_fc, find_v_max and extra_dt_constr terms of one E-form state
// (as face_terms) with one FP64 reciprocal of the density instead of one
// division per use, and the friction dt 0.01 Sigma / (alpha rho |v - v0|)
// with rho = gam3d/(2 gam3d - 1) g_eff Sigma^2 / p as one quotient: the
// p-form feeds the FP32 stage input, and the dt limits stay FP64 within a
// few ulp of the exact quotients.
template <int DIM>
__device__ __forceinline__ void face_terms_mixed(const PhysParams& P, const double* u, V3<double> fcn, double* w,
                                                 double& vmax, double& dtx) {
    const double inv = 1.0 / u[0];
    const V3<double> rl = cross(fcn, v3(u[1], u[2], u[3]));
    const double vv2 = dot(rl, rl) * (inv * inv);
#pragma unroll
    for (int k = 0; k < DIM; ++k) w[k] = u[k];
    double cs2;  // c_s^2
    if constexpr (DIM == 5) {
        const double p = smax(1e-16, (u[4] - u[0] * vv2 * 0.5) * (P.gam - 1));
        w[4] = p;
        cs2 = P.gam * p * inv;
        const double g_eff = kGM - vv2;
        const V3<double> vel = rl * (-inv);
        const V3<double> vel0 = cross(v3(0.0, 0.0, P.omega_ns), fcn);
        const double cand = 0.01 * p / (P.alpha * P.gam3d_rho * g_eff * u[0] * norm(vel - vel0));
        if (dtx > cand) dtx = cand;
    } else {
        cs2 = 1.0;
    }
    // max(|v|, c_s) as one square root (sqrt is monotonic; fmax skips a NaN
    // operand, as the two comparisons of find_v_max do)
    const double vc = sqrt(__builtin_fmax(vv2, cs2));
    if (vc > vmax) vmax = vc;
}

// One slot's rows: packed record (x, y), {H+/H-, BM/H+, beta+0, beta-0} and
// its edge's {ec, arc}, {en, chord}.
struct SlotRows {
    uint2 ra;
    float4 hs, e4, n4;
};

__device__ __forceinline__ SlotRows load_slot(const GeomPtrs& g, uint32_t ext, uint32_t edge_base) {
    SlotRows r;
    r.ra = *row(reinterpret_cast<const uint2*>(g.slot_a), ext);
    r.hs = *row(reinterpret_cast<const float4*>(g.slot_b), ext);
    const uint32_t edge = edge_base + ((r.ra.y >> 20) & 255u);
    r.e4 = *row(reinterpret_cast<const float4*>(g.edge32), 2 * edge);
    r.n4 = *row(reinterpret_cast<const float4*>(g.edge32), 2 * edge + 1);
    return r;
}

// Record bits (face fields are face-set locals): x = plus0 | plus1 << 10 |
// minus0 << 20 | negate-en << 30; y = minus1 | owner << 10 | edge << 20 |
// valence << 28.
template <int DIM>
__device__ __forceinline__ void reconstruct_slot(const float4* sW, const float* sP, const float* sC,
                                                 const float* sIA, float dt32, const SlotRows& r, float* out) {
    const uint2 ra = r.ra;
    const int li[5] = {int((ra.y >> 10) & 1023u), int(ra.x & 1023u), int((ra.x >> 10) & 1023u),
                       int((ra.x >> 20) & 1023u), int(ra.y & 1023u)};
    const float sg = (ra.x >> 30) & 1u ? -1.0f : 1.0f;
    reconstruct_lds<DIM>(sW, sP, sC, sIA, dt32, li, int(ra.y >> 28), r.hs, r.e4.w, v3(r.e4.x, r.e4.y, r.e4.z),
                         v3(sg * r.n4.x, sg * r.n4.y, sg * r.n4.z), out);
}

// dt0_arg / dt0v: the requested dt (dt0v[run] when dt0v is set), read by
// stage 1, which forms the step's dt.
template <int DIM, int SOLVER, int STAGE, int NT>
__global__ __launch_bounds__(NT, 5) void k_stage32(GeomPtrs g, const PhysParams* __restrict__ Pv,
                                                       Scalars* __restrict__ scv, double* __restrict__ U,
                                                       const float* __restrict__ W32,
                                                       float* __restrict__ W32_out, double dt0_arg,
                                                       const double* __restrict__ dt0v) {
    // dynamic LDS: the reconstructed half-edge states [slot_max][DIM] (after
    // the flux phase has read them, the same rows hold each own slot's flux
    // times chord length), then the face set's FP32 states (float4 Sigma, L
    // rows; dim 5: pressures, sound speeds) and 1 / area, each sized by the
    // layout
    extern __shared__ float4 lds_dyn[];
    float* const sUP = reinterpret_cast<float*>(lds_dyn);
    float* const sF = sUP;
    float4* const sW = lds_dyn + (g.slot_max * DIM) / 4;     // slot_max is a multiple of 4
    float* const sP = reinterpret_cast<float*>(sW + g.fset_max);  // dim 5: pressure
    float* const sC = sP + g.fset_max;                              // dim 5: sound speed
    float* const sIA = DIM == 5 ? sC + g.fset_max : sP;             // 1 / area
    const int b = xcd_block();
    const int run = g.blk_run ? g.blk_run[b] : 0;  // ensembles: this patch's run
    const PhysParams P = Pv[run];
    Scalars* __restrict__ sc = scv + run;
    const int f0 = g.blk_face[b], f1 = g.blk_face[b + 1];
    const int hbase = g.foff[f0];
    const int n_own = g.foff[f1] - hbase;
    const int hoff = g.halo_off[b];
    const int n_slots = n_own + (g.halo_off[b + 1] - hoff);
    const int t = threadIdx.x;
    const uint32_t ext0 = uint32_t(hbase + hoff), edge_base = ext0 >> 1;
    const bool first_patch = b == (g.run_blk ? g.run_blk[run] : 0);

    const int i = f0 + t;
    const bool face = i < f1;  // t < kFaces: waves 0 and 1

    // The step's dt.  Stage 2 reads it from the scalars.  Stage 1 forms it
    // (MUSCL_base.hpp:100-115, adiabatic.hpp:543-580) from the run's
    // accumulator buffer that the previous stage 2, or the prepare kernel,
    // filled: wave 0 of every block reduces the kDtSlots slots (the same
    // keys give every block the same dt) after its phase-0 gathers, while
    // the other waves finish theirs, and the run's first patch records dt,
    // raises the negative-dt stop, clears the other buffer for this step's
    // stage 2 and names it in the scalars.
    __shared__ DtCombine s_dt;
    __shared__ double s_dt_step;
    double dt = 0.0;
    int acc_rd = 0, acc_wr = 0;
    ulonglong2 acc_k = make_ulonglong2(~0ull, ~0ull);
    if constexpr (STAGE == 1) {
        acc_rd = int(sc->steps & 1ull);
        if (t < kDtSlots) acc_k = *reinterpret_cast<const ulonglong2*>(dt_slot(g, run, acc_rd, t));
    } else {
        dt = sc->dt;
        acc_wr = sc->acc_buf & 1;
        if (t == 0) dt_combine_init(&s_dt);  // before the first barrier
    }

    // phase 0: stage the FP32 states of the block's face set in LDS (own
    // faces first, so face f0 + t sits at local index t); the first slot's
    // rows are requested first so the loads overlap
    SlotRows r0;
    if (t < n_slots) r0 = load_slot(g, ext0 + uint32_t(t), edge_base);
    {
        // a face set holds at most 1024 faces (kPer per thread): every index
        // is requested before the first gather, so each thread waits for two
        // round trips, not two per face
        const int fs0 = g.fset_off[b], n_fs = g.fset_off[b + 1] - fs0;
        const float4* W4 = reinterpret_cast<const float4*>(W32);
        constexpr int kPer = (1024 + NT - 1) / NT;
        uint32_t f[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int q = t + j * NT;
            f[j] = q < n_fs ? uint32_t(*row(g.fset_face, uint32_t(fs0 + q))) : 0u;
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int q = t + j * NT;
            if (q < n_fs) {
                const float4 w = *row(W4, f[j]);
                sW[q] = w;
                sIA[q] = *row(g.inv_area, f[j]);
                if constexpr (DIM == 5) {
                    const float pq = *row(W32, uint32_t(4 * g.nf) + f[j]);
                    sP[q] = pq;
                    sC[q] = m_sqrt(m_div(P.f.gam * pq, w.x));
                }
            }
        }
    }
    if constexpr (STAGE == 1) {
        if (t < 64) {  // wave 0, every lane active (row_min)
            const unsigned long long kc = row_min(acc_k.x), kx = row_min(acc_k.y);
            if (t == 0) {
                const double dt0 = dt0v ? dt0v[run] : dt0_arg;
                double d = dt0;
                const double dt_cfl = from_order_key(kc);  // 0.1 h0 / max(|v|, c_s)
                if (d > dt_cfl) d = dt_cfl;
                if constexpr (DIM == 5) {
                    double extra = from_order_key(kx);
                    if (extra < 0) {
                        if (first_patch) atomicOr(&sc->stop, splash::kStopNegativeDt);
                        extra = 1e20;
                    }
                    if (d > extra) d = extra;
                }
                s_dt_step = d;
                if (first_patch) {
                    sc->dt0 = dt0;
                    sc->dt = d;
                    sc->acc_buf = acc_rd ^ 1;
                }
            }
            if (first_patch && t < kDtSlots)
                *reinterpret_cast<ulonglong2*>(dt_slot(g, run, acc_rd ^ 1, t)) = make_ulonglong2(~0ull, ~0ull);
        }
    }
    __syncthreads();
    if constexpr (STAGE == 1) dt = s_dt_step;
    const float dt32 = float(dt);

    // phase 1: reconstruct own and halo half-edges into LDS, slots t and
    // t + NT; thread k then requests the rows of edge k for phase 2
    const int n_edges = n_slots >> 1;
    {
        const int t1 = t + NT;
        SlotRows r1;
        if (t1 < n_slots) r1 = load_slot(g, ext0 + uint32_t(t1), edge_base);
        if (t < n_slots) reconstruct_slot<DIM>(sW, sP, sC, sIA, dt32, r0, &sUP[t * DIM]);
        if (t1 < n_slots) reconstruct_slot<DIM>(sW, sP, sC, sIA, dt32, r1, &sUP[t1 * DIM]);
    }
    // face data of phase 3, requested here so the flux phase hides the
    // loads without holding them through the reconstruction.  Stage 1's
    // result only feeds the FP32 input of stage 2, so it takes the face
    // centre and the accretion rate from one FP32 row; stage 2 updates the
    // FP64 state and reads the FP64 centre and round_diff / area.
    float rate = 0.0f;
    V3<double> fcn = v3(0.0, 0.0, 0.0);
    double rd_area = 0.0;
    int hb = 0, he = 0;
    if (face) {
        hb = *row(g.foff, uint32_t(i));
        he = *row(g.foff, uint32_t(i + 1));
        if constexpr (STAGE == 1) {
            const float4 c4 = *row(reinterpret_cast<const float4*>(g.fc32), uint32_t(i));
            fcn = v3(double(c4.x), double(c4.y), double(c4.z));
            rate = c4.w;
        } else {
            const double4 c4 = *row(reinterpret_cast<const double4*>(g.fcn), uint32_t(i));
            fcn = v3(c4.x, c4.y, c4.z);
            rd_area = c4.w;
            if (DIM == 5 && P.accretion_on) rate = *row(g.acc_rate, uint32_t(i));
        }
    }
    uint32_t rec = 0;
    float4 e4, n4;
    if (t < n_edges) {
        const uint32_t edge = edge_base + uint32_t(t);
        e4 = *row(reinterpret_cast<const float4*>(g.edge32), 2 * edge);
        n4 = *row(reinterpret_cast<const float4*>(g.edge32), 2 * edge + 1);
        rec = *row(g.edge_rec, edge);
    }
    __syncthreads();

    // phase 2: one flux per edge, thread k for edge k, evaluated from the
    // edge's half-edge of lower mesh index (the twin side has the same
    // midpoint and the negated normal) and written as +F*chord / -F*chord, so
    // the two faces of an edge exchange exactly opposite fluxes, whatever the
    // patch shape
    const int sw = int(rec & 511u);          // the evaluating (own) slot
    const int tws = int((rec >> 9) & 511u);  // its twin's slot
    const bool low = (rec >> 18) & 1u;
    float F[DIM];
    if (t < n_edges) {
        const V3<float> ec = v3(e4.x, e4.y, e4.z);
        const V3<float> en = low ? v3(n4.x, n4.y, n4.z) : v3(-n4.x, -n4.y, -n4.z);
        float uL[DIM], uR[DIM];
        const int sl = low ? sw : tws, sr = low ? tws : sw;
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            uL[k] = sUP[sl * DIM + k];
            uR[k] = sUP[sr * DIM + k];
        }
        const int err = flux32<DIM, SOLVER>(P.f.gam, P.f.kq, P.f.inv_gm1, uL, uR, ec, en, F);
        if (err) raise_stop(sc, err, hbase + sw);
        const float chord = low ? n4.w : -n4.w;
#pragma unroll
        for (int k = 0; k < DIM; ++k) F[k] *= chord;
    }
    __syncthreads();  // every edge's states are read: sF may overwrite them
    if (t < n_edges) {
#pragma unroll
        for (int k = 0; k < DIM; ++k) sF[sw * DIM + k] = F[k];
        if (low && tws < n_own)
#pragma unroll
            for (int k = 0; k < DIM; ++k) sF[tws * DIM + k] = -F[k];
    }
    __syncthreads();

    // phase 3: res2d (FP64 sum of the edge terms, compensated in stage 2),
    // source, RK combine and density floor for one face per thread
    double vmax = 1e-8, dtx = 1e20;
    if (face) {
        double u0[DIM];  // the FP64 state (stage 2, and stage 1's faces formed from it)
        double coef;     // round_diff / area * dt_here
        const double dt_here = (STAGE == 1) ? dt / 2 : dt;
        if constexpr (STAGE == 1) {
            coef = 0.5 * P.round_diff * double(dt32 * sIA[t]);  // dt / area in FP32
        } else {
            // U0 is first needed after the source and the edge sum, which
            // hide this load
            const double* ui = row(U, uint32_t(i * DIM));
#pragma unroll
            for (int k = 0; k < DIM; ++k) u0[k] = ui[k];
            coef = rd_area * dt_here;
        }
        float src[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) src[k] = 0.0f;
        if constexpr (DIM == 5) {
            // on the face's stage input (p-form), from LDS
            if (P.accretion_on) source32(P, sW[t], sP[t], v3(float(fcn.x), float(fcn.y), float(fcn.z)), rate, src);
        }
        // FP64 sum of the FP32 edge terms.  The terms are FP32 numbers, so
        // the partial sums are exact unless a face's terms span more than
        // 2^26 in magnitude.  Stage 2, which updates the FP64 state, adds
        // them Kahan-compensated (c carries the negated low part each
        // addition loses into the next term); stage 1's sums are rounded to
        // FP32, far coarser than anything the compensation could recover,
        // so it adds them plainly (c stays 0).
        double sum[DIM], c[DIM];
#pragma unroll
        for (int k = 0; k < DIM; ++k) sum[k] = c[k] = 0;
        bool nan_flux = false;
        for (int hh = hb; hh < he; ++hh) {
            const float* Fh = &sF[(hh - hbase) * DIM];
#pragma unroll
            for (int k = 0; k < DIM; ++k) {
                nan_flux |= __builtin_isnan(Fh[k]);
                if constexpr (STAGE == 1) {
                    sum[k] += double(Fh[k]);
                } else {
                    const double y = double(Fh[k]) - c[k];
                    const double s = sum[k] + y;
                    c[k] = (s - sum[k]) - y;
                    sum[k] = s;
                }
            }
        }
        if (nan_flux) raise_stop(sc, splash::kStopNaNFlux, i);
        // stage 1's result only feeds the FP32 input of stage 2: the
        // increments leave the FP64 sums rounded to FP32 and are
        // added to the face's own FP32 stage input (LDS entry t, the p-form
        // of U0); the pressure moves by (gam-1) (dE - dK), K = |r x L|^2 /
        // (2 Sigma), so no E-form state is formed.  That difference cancels
        // where the kinetic energy dwarfs the pressure (near vacuum, a
        // floored density): a face whose K exceeds 8 times its pressure
        // before or after (or with any non-finite value) is formed from the
        // FP64 state instead, as stage 2 forms its faces, so the FP32
        // rounding stays within a few ulp of the pressure.
        bool fast = STAGE == 1;
        if constexpr (STAGE == 1) {
            float inc[DIM];
#pragma unroll
            for (int k = 0; k < DIM; ++k) inc[k] = float(dt_here * double(src[k]) - coef * (sum[k] - c[k]));
            const float4 x = sW[t];
            float s1 = x.x + inc[0];
            if (s1 < 1e-8f) s1 = 1e-8f;
            const float4 w1 = make_float4(s1, x.y + inc[1], x.z + inc[2], x.w + inc[3]);
            float p1 = 0.0f;
            if constexpr (DIM == 5) {
                const V3<float> cf = v3(float(fcn.x), float(fcn.y), float(fcn.z));  // the FP32 centre, exact
                const V3<float> r0 = cross(cf, v3(x.y, x.z, x.w)), r1 = cross(cf, v3(w1.y, w1.z, w1.w));
                const float k0 = 0.5f * dot(r0, r0) * __builtin_amdgcn_rcpf(x.x);
                const float k1 = 0.5f * dot(r1, r1) * __builtin_amdgcn_rcpf(w1.x);
                p1 = smax(1e-16f, fmaf(P.f.gm1, inc[4] - (k1 - k0), sP[t]));
                fast = smax(k0, k1) <= 8.0f * smin(sP[t], p1);  // false for NaN
            }
            if (fast) {
                *row(reinterpret_cast<float4*>(W32_out), uint32_t(i)) = w1;
                if constexpr (DIM == 5) *row(W32_out, uint32_t(4 * g.nf + i)) = p1;
            } else {
                const double* ui = row(U, uint32_t(i * DIM));  // U0 (stage 2 has not written U1 yet)
#pragma unroll
                for (int k = 0; k < DIM; ++k) u0[k] = ui[k];
                const double4 c4 = *row(reinterpret_cast<const double4*>(g.fcn), uint32_t(i));  // the FP64 centre
                fcn = v3(c4.x, c4.y, c4.z);
            }
        }
        if (!fast) {
            double un[DIM];
#pragma unroll
            for (int k = 0; k < DIM; ++k) un[k] = (dt_here * double(src[k]) - coef * (sum[k] - c[k])) + u0[k];
            if (un[0] < 1e-8) un[0] = 1e-8;
            double wn[DIM];
            face_terms_mixed<DIM>(P, un, fcn, wn, vmax, dtx);
            store_w32<DIM>(W32_out, g.nf, i, wn);
            if constexpr (STAGE == 2) {
#pragma unroll
                for (int k = 0; k < DIM; ++k) row(U, uint32_t(i * DIM))[k] = un[k];
            }
        }
    }
    if constexpr (STAGE == 2) {
        // this block's partial of max(|v|, c_s) and min friction dt: its
        // faces are in waves 0 and 1 (kFaces = 128; wave 1 holds neutral
        // values when it has none), each reduces its own, and the two are
        // combined in LDS (dt_combine), so neither waits for the other
        static_assert(kFaces == 128, "waves 0 and 1 hold a patch's faces");
        if (t < 128) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_xor(vmax, o);
                const double ox = __shfl_xor(dtx, o);
                if (ov > vmax) vmax = ov;
                if (dtx > ox) dtx = ox;
            }
            if ((t & 63) == 0) dt_combine(&s_dt, dt_slot(g, run, acc_wr, b % kDtSlots), P.h0, vmax, dtx);
            if (t == 0 && first_patch) {  // the run's first patch advances its clock
                sc->t += sc->dt;
                sc->steps += 1;
            }
        }
    }
}

namespace {

template <int DIM, int SOLVER, int NT>
void stage_nt(const StepArgs& a, int stage, double dt0, hipStream_t s, const KernelEvents* ev) {
    const size_t lds = size_t(a.geom.fset_max) * (DIM == 5 ? 28 : 20) + sizeof(float) * DIM * size_t(a.geom.slot_max);
    if (a.geom.fset_max > 1024 || a.geom.slot_max > 2 * NT || a.geom.slot_max % 4 || lds > 64 * 1024)
        throw std::runtime_error("stage block LDS exceeds the limit");
    if (stage == 1)
        launch(k_stage32<DIM, SOLVER, 1, NT>, a.geom.n_blocks, NT, lds, s, ev, a.geom, a.phys_dev, a.sc, a.U, a.W32,
               a.W32b, dt0, a.dt0_dev);
    else
        launch(k_stage32<DIM, SOLVER, 2, NT>, a.geom.n_blocks, NT, lds, s, ev, a.geom, a.phys_dev, a.sc, a.U, a.W32b,
               a.W32, dt0, a.dt0_dev);
}

template <int DIM, int SOLVER>
void stage(const StepArgs& a, int stage, double dt0, hipStream_t s, const KernelEvents* ev) {
    switch (a.block_threads) {
    case 128: stage_nt<DIM, SOLVER, 128>(a, stage, dt0, s, ev); break;
    case 192: stage_nt<DIM, SOLVER, 192>(a, stage, dt0, s, ev); break;
    case 256: stage_nt<DIM, SOLVER, 256>(a, stage, dt0, s, ev); break;
    default: throw std::invalid_argument("unsupported stage-kernel block size");
    }
}

}  // namespace

// Every run's two dt accumulator buffers set to the neutral key (a kernel,
// not a memset, so no memset node can enter a captured graph).
__global__ __launch_bounds__(256) void k_clear_dt_acc(unsigned long long* __restrict__ acc, int n) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) acc[k] = ~0ull;
}

// E-form state -> FP32 stage input and this state's dt limits (into the
// accumulator buffer the next stage 1 reads), with the arithmetic of
// k_stage32's epilogue (face_terms_mixed), so a run started or resumed from
// a state continues exactly as one that reached it by stepping.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_prepare32(GeomPtrs g, const PhysParams* __restrict__ Pv,
                                                      const Scalars* __restrict__ scv, const double* __restrict__ U,
                                                      float* __restrict__ W32) {
    const int b = xcd_block();
    const int run = g.blk_run ? g.blk_run[b] : 0;
    const PhysParams P = Pv[run];
    const int acc_buf = int(scv[run].steps & 1ull);
    const int t = threadIdx.x;
    __shared__ DtCombine s_dt;
    if (t == 0) dt_combine_init(&s_dt);
    __syncthreads();
    if (t >= kFaces) return;  // a patch's faces fit waves 0 and 1
    const int i = g.blk_face[b] + t;
    double vmax = 1e-8, dtx = 1e20;
    if (i < g.blk_face[b + 1]) {
        double u[DIM], w[DIM];
        const double* ui = row(U, uint32_t(i * DIM));
#pragma unroll
        for (int k = 0; k < DIM; ++k) u[k] = ui[k];
        const double4 c4 = *row(reinterpret_cast<const double4*>(g.fcn), uint32_t(i));
        face_terms_mixed<DIM>(P, u, v3(c4.x, c4.y, c4.z), w, vmax, dtx);
        store_w32<DIM>(W32, g.nf, i, w);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(vmax, o);
        const double ox = __shfl_xor(dtx, o);
        if (ov > vmax) vmax = ov;
        if (dtx > ox) dtx = ox;
    }
    if ((t & 63) == 0) dt_combine(&s_dt, dt_slot(g, run, acc_buf, b % kDtSlots), P.h0, vmax, dtx);
}

void enqueue_prepare(const StepArgs& a, hipStream_t s) {
    const int n_acc = a.geom.n_runs * kDtRunWords;
    k_clear_dt_acc<<<(n_acc + 255) / 256, 256, 0, s>>>(a.geom.dt_acc, n_acc);
    if (a.phys.dim == 5) k_prepare32<5><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys_dev, a.sc, a.U, a.W32);
    else k_prepare32<4><<<a.geom.n_blocks, kBlock, 0, s>>>(a.geom, a.phys_dev, a.sc, a.U, a.W32);
}

int slots(int threads) { return 2 * threads; }
int max_faces() { return kFaces; }
int default_threads() { return 256; }
bool valid_threads(int threads) {
    for (int c : kThreadChoices)
        if (c == threads) return true;
    return false;
}

void enqueue_stage(const StepArgs& a, int st, double dt0, hipStream_t s, const KernelEvents* ev) {
    switch (a.phys.solver) {
    case kHLLE: stage<4, kHLLE>(a, st, dt0, s, ev); break;
    case kHLLE_P: stage<5, kHLLE_P>(a, st, dt0, s, ev); break;
    case kHLLC: stage<5, kHLLC>(a, st, dt0, s, ev); break;
    default: stage<5, kHLLCPLUS>(a, st, dt0, s, ev); break;
    }
}


}  // namespace mixed
}  // namespace splash
