// Host side of one resident SPLASH run: uploads the stencil tables, keeps
// the state on the GPU and enqueues the per-step kernel sequence.
#include "engine.hpp"
#include "layout.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace splash {

namespace {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

inline double norm3(double x, double y, double z) {
    double s = 0.0;
    s = s + x * x;
    s = s + y * y;
    s = s + z * z;
    return std::sqrt(s);
}

}  // namespace

Engine::Engine(const PolyMesh& mesh, const double* U_in, int dim, int solver, double gam, double omega_ns,
               const RunParams& params, int precision, int geometry_threads)
    : dim_(dim), solver_(solver), precision_(precision), gam_(gam), omega_ns_(omega_ns), params_(params) {
    if (solver < kHLLE || solver > kHLLCPLUS) throw std::invalid_argument("unknown solver");
    if ((solver == kHLLE && dim != 4) || (solver != kHLLE && dim != 5))
        throw std::invalid_argument("dim must be 4 for MUSCL_HLLE and 5 for the adiabatic solvers");
    if (precision != kFP64 && precision != kMixed) throw std::invalid_argument("unknown precision mode");
    if (!U_in) throw std::invalid_argument("U_in is null");
    geo_ = build_geometry(mesh, geometry_threads);

    const int32_t nf = geo_.nf;
    if (dim == 5) {
        // adiabatic.hpp:70-90 (the index of the :75 loop starts at face 0)
        double area_coeff = 0;
        const double mu = M_PI / 2;
        const double sigma = params.acc_width / 2.355;
        for (int32_t f = 0; f < nf; f++) {
            const double* c = &geo_.fc[3 * size_t(f)];
            const double theta = std::acos(c[2] / norm3(c[0], c[1], c[2]));
            area_coeff += geo_.area[f] / std::sqrt(2 * M_PI * sigma * sigma) *
                          std::exp(-(theta - mu) * (theta - mu) / (2 * sigma * sigma));
        }
        area_coeff /= 4 * M_PI;
        area_coeff_ = area_coeff;
        double tm = 0;
        for (int32_t f = 0; f < nf; f++) tm += U_in[size_t(f) * dim] * geo_.area[f];
        total_mass_ = tm;
    }

    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    try {
        upload();
        std::vector<double> U(size_t(nf) * dim);
        for (int32_t k = 0; k < nf; ++k)
            for (int c = 0; c < dim; ++c) U[size_t(k) * dim + c] = U_in[size_t(perm_[k]) * dim + c];
        check(hipMemcpy(args_.U, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice), "upload state");
        fp64::enqueue_prepare(args_, stream_);
        check(hipGetLastError(), "prepare launch");
        check(hipStreamSynchronize(stream_), "prepare");
    } catch (...) {
        for (void* p : allocations_) (void)hipFree(p);
        (void)hipStreamDestroy(stream_);
        throw;
    }
}

Engine::~Engine() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : allocations_) (void)hipFree(p);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::upload() {
    const Geometry& g = geo_;
    const int32_t nf = g.nf, nh = g.nh;
    if (g.max_valence >= 32) throw std::runtime_error("face valence >= 32 is not supported");
    {
        Layout L = precision_ == kMixed ? build_layout(g, mixed::slots(), mixed::max_faces())
                                        : build_layout(g, fp64::block_size(), fp64::block_size());
        validate_layout(g, L);
        perm_ = std::move(L.perm);
        iperm_ = std::move(L.iperm);
        foff_ = std::move(L.foff);
        he_old_ = std::move(L.he_old);
        he_new_ = std::move(L.he_new);
        blk_face_ = std::move(L.blk_face);
        halo_off_ = std::move(L.halo_off);
        halo_he_ = std::move(L.halo_he);
        twin_slot_ = std::move(L.twin_slot);
        fset_off_ = std::move(L.fset_off);
        fset_face_ = std::move(L.fset_face);
        ext_loc_ = std::move(L.ext_loc);
    }
    auto dev = [&](size_t bytes, const void* host) -> void* {
        void* p = nullptr;
        check(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc");
        allocations_.push_back(p);
        if (host && bytes) check(hipMemcpy(p, host, bytes, hipMemcpyHostToDevice), "upload");
        return p;
    };

    std::vector<int32_t> he_fv(nh), sten(4 * size_t(nh));
    std::vector<double> beta(4 * size_t(nh)), hsc(4 * size_t(nh)), ec(4 * size_t(nh)), en(4 * size_t(nh));
    for (int32_t hn = 0; hn < nh; ++hn) {
        const int32_t h = he_old_[hn];
        const int32_t f = g.he_face[h];
        he_fv[hn] = iperm_[f] | ((g.foff[f + 1] - g.foff[f]) << 27);
        sten[4 * size_t(hn) + 0] = iperm_[g.ff_plus[2 * size_t(h)]];
        sten[4 * size_t(hn) + 1] = iperm_[g.ff_plus[2 * size_t(h) + 1]];
        sten[4 * size_t(hn) + 2] = iperm_[g.ff_minus[2 * size_t(h)]];
        sten[4 * size_t(hn) + 3] = iperm_[g.ff_minus[2 * size_t(h) + 1]];
        beta[4 * size_t(hn) + 0] = g.beta_plus[2 * size_t(h)];
        beta[4 * size_t(hn) + 1] = g.beta_plus[2 * size_t(h) + 1];
        beta[4 * size_t(hn) + 2] = g.beta_minus[2 * size_t(h)];
        beta[4 * size_t(hn) + 3] = g.beta_minus[2 * size_t(h) + 1];
        hsc[4 * size_t(hn) + 0] = g.h_plus[h];
        hsc[4 * size_t(hn) + 1] = g.h_minus[h];
        hsc[4 * size_t(hn) + 2] = g.bm_dist[h];
        // adiabatic.hpp:757-758: distance(v_j, v_j1) / surface_area[n_face]
        hsc[4 * size_t(hn) + 3] = g.arc[h] / g.area[f];
        for (int c = 0; c < 3; ++c) {
            ec[4 * size_t(hn) + c] = g.ec[3 * size_t(h) + c];
            en[4 * size_t(hn) + c] = g.en[3 * size_t(h) + c];
        }
        ec[4 * size_t(hn) + 3] = g.ecn[h];
        en[4 * size_t(hn) + 3] = g.chord[h];
    }
    std::vector<double> area(nf), fcn(4 * size_t(nf), 0.0);
    for (int32_t k = 0; k < nf; ++k) {
        const int32_t f = perm_[k];
        area[k] = g.area[f];
        for (int c = 0; c < 3; ++c) fcn[4 * size_t(k) + c] = g.fcn[3 * size_t(f) + c];
    }

    // accretion profile per face (adiabatic.hpp:425-449), evaluated once with
    // the expressions the reference evaluates on every source() call
    std::vector<double> acc(8 * size_t(nf), 0.0);
    if (dim_ == 5 && params_.accretion_on) {
        const double mu = M_PI / 2;
        const double sigma = params_.acc_width / 2.355;
        const double tilt = params_.tilt_angle;
        for (int32_t k = 0; k < nf; ++k) {
            const double* r = &fcn[4 * size_t(k)];
            const double theta_acc = std::acos(r[1] * std::sin(tilt) + r[2] * std::cos(tilt));
            const double rate = 1 / area_coeff_ * params_.acc_rate / std::sqrt(2 * M_PI * sigma * sigma) *
                                std::exp(-(theta_acc - mu) * (theta_acc - mu) / (2 * sigma * sigma));
            const double ox = 0, oy = std::sin(tilt) * params_.omega_acc_abs, oz = std::cos(tilt) * params_.omega_acc_abs;
            // omxr = omega_acc x r ; rxv = r x omxr
            const double mx = oy * r[2] - oz * r[1], my = oz * r[0] - ox * r[2], mz = ox * r[1] - oy * r[0];
            double* a = &acc[8 * size_t(k)];
            a[0] = rate;
            a[1] = mx;
            a[2] = my;
            a[3] = mz;
            a[4] = r[1] * mz - r[2] * my;
            a[5] = r[2] * mx - r[0] * mz;
            a[6] = r[0] * my - r[1] * mx;
        }
    }

    GeomPtrs& d = args_.geom;
    d.nf = nf;
    d.nh = nh;
    d.foff = static_cast<const int32_t*>(dev(sizeof(int32_t) * (nf + 1), foff_.data()));
    d.he_fv = static_cast<const int32_t*>(dev(sizeof(int32_t) * nh, he_fv.data()));
    d.sten = static_cast<const int32_t*>(dev(sizeof(int32_t) * sten.size(), sten.data()));
    d.beta = static_cast<const double*>(dev(sizeof(double) * beta.size(), beta.data()));
    d.hsc = static_cast<const double*>(dev(sizeof(double) * hsc.size(), hsc.data()));
    d.ec = static_cast<const double*>(dev(sizeof(double) * ec.size(), ec.data()));
    d.en = static_cast<const double*>(dev(sizeof(double) * en.size(), en.data()));
    d.twin_slot = static_cast<const int32_t*>(dev(sizeof(int32_t) * twin_slot_.size(), twin_slot_.data()));
    d.area = static_cast<const double*>(dev(sizeof(double) * nf, area.data()));
    d.fcn = static_cast<const double*>(dev(sizeof(double) * fcn.size(), fcn.data()));
    d.acc = static_cast<const double*>(dev(sizeof(double) * acc.size(), acc.data()));
    d.blk_face = static_cast<const int32_t*>(dev(sizeof(int32_t) * blk_face_.size(), blk_face_.data()));
    d.halo_off = static_cast<const int32_t*>(dev(sizeof(int32_t) * halo_off_.size(), halo_off_.data()));
    d.halo_he = static_cast<const int32_t*>(dev(sizeof(int32_t) * halo_he_.size(), halo_he_.data()));
    if (precision_ == kMixed) {
        // FP32 tables of the mixed kernel, per extended slot row (block b's
        // own then halo half-edges start at foff[f0] + halo_off[b]):
        //   slot_a: packed record (x, y) and beta+0, beta-0 (beta.1 = 1 - beta.0
        //           to 3e-13 on every generated mesh, invisible in FP32)
        //   slot_b: 1/H+, 1/H-, BM_dist, arc/area
        // and per edge of a block (a block holds both slots of each of its
        // edges, so it has n_slots/2 edges, starting at row (foff[f0] +
        // halo_off[b]) / 2), oriented for the slot that evaluates its flux:
        //   edge:   {ec xyz, |ec|}, {en xyz, chord}
        // Record bits: x = plus0 | plus1 << 10 | minus0 << 20 | negate-en << 30;
        // y = minus1 | owner << 10 | edge << 20 | valence << 28 (face-set
        // locals).  Per edge, edge_rec = evaluating slot | twin slot << 9 |
        // (evaluating slot has the lower half-edge index) << 18.
        const size_t n_ext = ext_loc_.size() / 4;
        if (n_ext % 2) throw std::runtime_error("odd slot count");
        // the kernel addresses every table with 32-bit byte offsets
        if (16 * n_ext >= (size_t(1) << 32) || 8 * size_t(nf) * size_t(dim_) >= (size_t(1) << 32))
            throw std::runtime_error("mesh too large for the mixed-precision path's 32-bit table offsets");
        std::vector<uint32_t> sa(4 * n_ext, 0u);
        std::vector<float> sb(4 * n_ext), ed(4 * n_ext);  // ed: n_ext/2 edges of 8 floats
        std::vector<uint32_t> erec(n_ext / 2, 0u);
        const int32_t nb = int32_t(blk_face_.size()) - 1;
        std::vector<int32_t> eidx, own_of_halo;
        for (int32_t b = 0; b < nb; ++b) {
            const int32_t h0 = foff_[blk_face_[b]], h1 = foff_[blk_face_[b + 1]];
            const int32_t n_own = h1 - h0;
            const int32_t n_slots = n_own + halo_off_[b + 1] - halo_off_[b];
            const size_t e0 = size_t(h0) + size_t(halo_off_[b]);
            if (n_slots % 2 || e0 % 2) throw std::runtime_error("block slots do not pair into edges");
            eidx.assign(size_t(n_slots), -1);
            own_of_halo.assign(size_t(n_slots), -1);
            auto tw_of = [&](int32_t s) { return twin_slot_[2 * size_t(h0 + s) + 1]; };
            auto low_of = [&](int32_t s) { return h0 + s < twin_slot_[2 * size_t(h0 + s)]; };
            int32_t n_edges = 0;
            for (int32_t s = 0; s < n_own; ++s)  // representatives: the slots that evaluate a flux
                if (low_of(s) || tw_of(s) >= n_own) eidx[size_t(s)] = n_edges++;
            if (2 * n_edges != n_slots) throw std::runtime_error("edge count is not half the slot count");
            for (int32_t s = 0; s < n_own; ++s) {
                const int32_t t = tw_of(s);
                if (t >= n_own) own_of_halo[size_t(t)] = s;
                else if (eidx[size_t(s)] < 0) eidx[size_t(s)] = eidx[size_t(t)];
            }
            for (int32_t s = n_own; s < n_slots; ++s) {
                if (own_of_halo[size_t(s)] < 0) throw std::runtime_error("halo slot without an own twin");
                eidx[size_t(s)] = eidx[size_t(own_of_halo[size_t(s)])];
            }
            for (int32_t s = 0; s < n_slots; ++s) {
                const size_t e = e0 + size_t(s);
                const int32_t* L = &ext_loc_[4 * e];
                const size_t hn = size_t(L[3]);
                const uint32_t owner = uint32_t(L[0] & 0xffff), valence = uint32_t(L[0] >> 16);
                const uint32_t q[4] = {uint32_t(L[1] & 0xffff), uint32_t(L[1] >> 16), uint32_t(L[2] & 0xffff),
                                       uint32_t(L[2] >> 16)};
                if (owner > 1023 || q[0] > 1023 || q[1] > 1023 || q[2] > 1023 || q[3] > 1023 || valence > 15)
                    throw std::runtime_error("slot record does not fit its 10-bit local indices / 4-bit valence");
                const bool own = s < n_own;
                const bool rep = own && (low_of(s) || tw_of(s) >= n_own);
                const uint32_t ei = uint32_t(eidx[size_t(s)]);
                if (ei > 255 || n_slots > 512) throw std::runtime_error("slot record does not fit its fields");
                sa[4 * e] = q[0] | q[1] << 10 | q[2] << 20 | (rep ? 0u : 1u) << 30;
                sa[4 * e + 1] = q[3] | owner << 10 | ei << 20 | valence << 28;
                if (rep)
                    erec[e0 / 2 + ei] = uint32_t(s) | uint32_t(tw_of(s)) << 9 | (low_of(s) ? 1u : 0u) << 18;
                const float bp = float(beta[4 * hn + 0]), bm = float(beta[4 * hn + 2]);
                std::memcpy(&sa[4 * e + 2], &bp, 4);
                std::memcpy(&sa[4 * e + 3], &bm, 4);
                sb[4 * e + 0] = float(1.0 / hsc[4 * hn + 0]);
                sb[4 * e + 1] = float(1.0 / hsc[4 * hn + 1]);
                sb[4 * e + 2] = float(hsc[4 * hn + 2]);
                sb[4 * e + 3] = float(hsc[4 * hn + 3]);
                if (rep) {
                    float* r = &ed[8 * (e0 / 2 + ei)];
                    for (int c = 0; c < 4; ++c) {
                        r[c] = float(ec[4 * hn + c]);
                        r[4 + c] = float(en[4 * hn + c]);
                    }
                }
            }
        }
        std::vector<double> rate(nf);
        for (int32_t k = 0; k < nf; ++k) rate[k] = acc[8 * size_t(k)];
        d.acc_rate = static_cast<const double*>(dev(sizeof(double) * rate.size(), rate.data()));
        std::vector<double> rda(nf);
        for (int32_t k = 0; k < nf; ++k) rda[k] = g.round_diff / area[k];
        d.rd_area = static_cast<const double*>(dev(sizeof(double) * rda.size(), rda.data()));
        d.slot_a = static_cast<const uint32_t*>(dev(sizeof(uint32_t) * sa.size(), sa.data()));
        d.slot_b = static_cast<const float*>(dev(sizeof(float) * sb.size(), sb.data()));
        d.edge32 = static_cast<const float*>(dev(sizeof(float) * ed.size(), ed.data()));
        d.edge_rec = static_cast<const uint32_t*>(dev(sizeof(uint32_t) * erec.size(), erec.data()));
        d.fset_off = static_cast<const int32_t*>(dev(sizeof(int32_t) * fset_off_.size(), fset_off_.data()));
        d.fset_face = static_cast<const int32_t*>(dev(sizeof(int32_t) * fset_face_.size(), fset_face_.data()));
        int32_t mx = 1;
        for (size_t q = 0; q + 1 < fset_off_.size(); ++q) mx = std::max(mx, fset_off_[q + 1] - fset_off_[q]);
        d.fset_max = mx;
        const size_t row = dim_ == 5 ? 5 : 4;  // float4 (Sigma, L) rows, then the pressures
        args_.W32 = static_cast<float*>(dev(sizeof(float) * row * nf, nullptr));
        args_.W32b = static_cast<float*>(dev(sizeof(float) * row * nf, nullptr));
    } else {
        d.slot_a = d.edge_rec = nullptr;
        d.acc_rate = d.rd_area = nullptr;
        d.fset_off = d.fset_face = nullptr;
        d.slot_b = d.edge32 = nullptr;
        d.fset_max = 0;
        args_.W32 = args_.W32b = nullptr;
    }
    d.n_blocks = int32_t(blk_face_.size()) - 1;
    d.dt_partial = static_cast<double*>(dev(sizeof(double) * 2 * size_t(d.n_blocks), nullptr));

    PhysParams& P = args_.phys;
    P.dim = dim_;
    P.solver = solver_;
    P.accretion_on = params_.accretion_on;
    P.friction_on = params_.friction_on;
    P.gam = gam_;
    P.omega_ns = omega_ns_;
    P.h0 = g.h0;
    P.round_diff = g.round_diff;
    P.e_acc = params_.e_acc;
    P.alpha = params_.alpha;
    const double fall_eff = 1;
    P.dmdt_coef = dim_ == 5 ? -(fall_eff * params_.acc_rate * 4 * M_PI) / total_mass_ : 0.0;
    P.om_acc[0] = 0;
    P.om_acc[1] = std::sin(params_.tilt_angle) * params_.omega_acc_abs;
    P.om_acc[2] = std::cos(params_.tilt_angle) * params_.omega_acc_abs;
    {
        const double gam3d = 1 / (2 - gam_);
        P.gam_h = gam_ / (gam_ - 1);
        P.gam3d_rho = gam3d / (2 * gam3d - 1);
        P.cool_k = 12. / 5 * 1.6e-13 * std::pow(3. / 4 * 4.85e36, 0.25);  // adiabatic.hpp cooling constants
    }

    args_.U = static_cast<double*>(dev(sizeof(double) * size_t(nf) * dim_, nullptr));
    args_.W = static_cast<double*>(dev(sizeof(double) * size_t(nf) * dim_, nullptr));
    args_.W2 = static_cast<double*>(dev(sizeof(double) * size_t(nf) * dim_, nullptr));

    Scalars s{};
    s.first_err = ~0ull;
    args_.sc = static_cast<Scalars*>(dev(sizeof(Scalars), &s));

    const int B = fp64::block_size();
    n_diag_blocks_ = (nf + B - 1) / B;
    d_partial_ = static_cast<double*>(dev(sizeof(double) * 5 * n_diag_blocks_, nullptr));
}

void Engine::enqueue_kernel(int which, double dt0, const KernelEvents* ev) {
    if (which == 0) fp64::enqueue_finalize_dt(args_, dt0, stream_, ev);
    else if (precision_ == kFP64) fp64::enqueue_stage(args_, which, stream_, ev);
    else mixed::enqueue_stage(args_, which, stream_, ev);
}

void Engine::enqueue_step(double dt0) {
    for (int k = 0; k < kKernelsPerStep; ++k) enqueue_kernel(k, dt0);
}

void Engine::profile(double dt0, int n, double* ms) {
    // n steps launched back to back on this run's stream, each kernel with
    // start/stop HIP events bound to its own dispatch, one wait at the end
    constexpr int K = kKernelsPerStep;
    const size_t nk = size_t(n > 0 ? n : 0) * K;
    std::vector<hipEvent_t> ev(2 * nk, nullptr);
    double acc[K + 1] = {0};
    try {
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
        for (size_t q = 0; q < nk; ++q) {
            const KernelEvents ke{ev[2 * q], ev[2 * q + 1]};
            enqueue_kernel(int(q % K), dt0, &ke);
        }
        check(hipGetLastError(), "profile launch");
        if (nk) check(hipEventSynchronize(ev.back()), "hipEventSynchronize");
        for (size_t q = 0; q < nk; ++q) {
            float t = 0;
            check(hipEventElapsedTime(&t, ev[2 * q], ev[2 * q + 1]), "hipEventElapsedTime");
            acc[q % K] += t;
            if (q % K == K - 1) {
                check(hipEventElapsedTime(&t, ev[2 * (q - K + 1)], ev[2 * q + 1]), "hipEventElapsedTime");
                acc[K] += t;
            }
        }
    } catch (...) {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k <= K; ++k) ms[k] = n > 0 ? acc[k] / n : 0.0;
}

void Engine::step(double dt0) {
    enqueue_step(dt0);
    check(hipGetLastError(), "step launch");
}

void Engine::steps(double dt0, int n) {
    if (n <= 0) return;
    constexpr int kGraphSteps = 8;
    if (n >= kGraphSteps) {
        if (!graph_exec_ || graph_dt0_ != dt0 || graph_steps_ != kGraphSteps) {
            if (graph_exec_) check(hipGraphExecDestroy(graph_exec_), "hipGraphExecDestroy");
            if (graph_) check(hipGraphDestroy(graph_), "hipGraphDestroy");
            graph_exec_ = nullptr;
            graph_ = nullptr;
            check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            for (int k = 0; k < kGraphSteps; ++k) enqueue_step(dt0);
            check(hipStreamEndCapture(stream_, &graph_), "hipStreamEndCapture");
            check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
            graph_dt0_ = dt0;
            graph_steps_ = kGraphSteps;
        }
        while (n >= kGraphSteps) {
            check(hipGraphLaunch(graph_exec_, stream_), "hipGraphLaunch");
            n -= kGraphSteps;
        }
    }
    for (int k = 0; k < n; ++k) enqueue_step(dt0);
    check(hipGetLastError(), "steps launch");
}

void Engine::synchronize() { check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

Scalars Engine::read_scalars() {
    Scalars s;
    check(hipMemcpyAsync(&s, args_.sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream_), "read scalars");
    synchronize();
    return s;
}

double Engine::time() { return read_scalars().t; }
double Engine::last_dt() { return read_scalars().dt; }
uint64_t Engine::step_count() { return read_scalars().steps; }
int Engine::stop_bits() { return read_scalars().stop; }
uint64_t Engine::first_error() {
    // kernels record internal indices: map back to the mesh's numbering
    const uint64_t e = read_scalars().first_err;
    if (e == ~0ull) return e;
    const uint64_t code = e >> 40, where = e & 0xffffffffffull;
    const uint64_t mapped = code == kStopNaNFlux ? uint64_t(perm_[where]) : uint64_t(he_old_[where]);
    return (code << 40) | mapped;
}

void Engine::get_state(double* U_out) {
    std::vector<double> U(size_t(geo_.nf) * dim_);
    check(hipMemcpyAsync(U.data(), args_.U, sizeof(double) * U.size(), hipMemcpyDeviceToHost, stream_), "read state");
    synchronize();
    for (int32_t k = 0; k < geo_.nf; ++k)
        for (int c = 0; c < dim_; ++c) U_out[size_t(perm_[k]) * dim_ + c] = U[size_t(k) * dim_ + c];
}

void Engine::diagnostics(double out[5]) {
    fp64::enqueue_diag(args_, d_partial_, stream_);
    check(hipGetLastError(), "diag launch");
    std::vector<double> part(5 * size_t(n_diag_blocks_));
    check(hipMemcpyAsync(part.data(), d_partial_, sizeof(double) * part.size(), hipMemcpyDeviceToHost, stream_),
          "read diagnostics");
    synchronize();
    for (int k = 0; k < 5; ++k) out[k] = 0;
    for (int b = 0; b < n_diag_blocks_; ++b)
        for (int k = 0; k < 5; ++k) out[k] += part[5 * size_t(b) + k];
}

}  // namespace splash
