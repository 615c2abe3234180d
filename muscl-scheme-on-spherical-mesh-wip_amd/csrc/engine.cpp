// Host side of one resident SPLASH run: uploads the stencil tables, keeps
// the state on the GPU and enqueues the per-step kernel sequence.
#include "engine.hpp"
#include "layout.hpp"
#include "run_setup.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace splash {

namespace {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

Engine::Engine(const PolyMesh& mesh, const double* U_in, int dim, int solver, double gam, double omega_ns,
               const RunParams& params, int precision, int geometry_threads, int patch_faces, int block_threads)
    : dim_(dim), solver_(solver), precision_(precision), gam_(gam), omega_ns_(omega_ns), params_(params) {
    if (precision != kFP64 && precision != kMixed) throw std::invalid_argument("unknown precision mode");
    if (patch_faces < -1 || patch_faces > kPatchFacesMax) throw std::invalid_argument("patch_faces out of range");
    if (block_threads != 0 && !mixed::valid_threads(block_threads))
        throw std::invalid_argument("block_threads must be 0, 128, 192 or 256");
    const int32_t threads = block_threads ? block_threads : mixed::default_threads();
    const int32_t cap = patch_faces > 0 ? patch_faces : default_patch_faces(mesh, threads);
    // a defaulted cap falls back to one face wave where its face sets overflow
    RunHost r = precision == kMixed
                    ? build_run_host(mesh, U_in, dim, solver, gam, omega_ns, params, mixed::slots(threads), cap,
                                     geometry_threads, patch_faces > 0 ? 0 : kPatchFacesOneWave)
                    : build_run_host(mesh, U_in, dim, solver, gam, omega_ns, params, fp64::block_size(),
                                     fp64::block_size(), geometry_threads);
    if (precision == kMixed && patch_faces == -1) autotune_patch_shape(r, U_in, 1e-4);
    init(r, U_in);
    geo_ = std::move(r.geo);
}

Engine::Engine(const RunHost& r, const double* U_in, int precision)
    : dim_(r.dim), solver_(r.solver), precision_(precision), gam_(r.phys.gam), omega_ns_(r.phys.omega_ns),
      params_(r.params) {
    if (precision != kFP64 && precision != kMixed) throw std::invalid_argument("unknown precision mode");
    init(r, U_in);
}

void Engine::init(const RunHost& r, const double* U_in) {
    area_coeff_ = r.area_coeff;
    total_mass_ = r.total_mass;
    patch_faces_ = r.L.max_faces;
    if (precision_ == kMixed && patch_faces_ > mixed::max_faces())
        throw std::logic_error("patch face cap exceeds the faces the stage kernel updates per patch");
    block_threads_ = precision_ == kMixed ? r.L.slots / 2 : fp64::block_size();
    nf_ = r.geo.nf;
    nh_ = r.geo.nh;
    const int32_t nf = r.geo.nf;
    check(hipGetDevice(&device_), "hipGetDevice");
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    try {
        upload(r);
        args_.block_threads = block_threads_;
        std::vector<double> U(size_t(nf) * dim_);
        for (int32_t k = 0; k < nf; ++k)
            for (int c = 0; c < dim_; ++c) U[size_t(k) * dim_ + c] = U_in[size_t(perm_[k]) * dim_ + c];
        check(hipMemcpy(args_.U, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice), "upload state");
        enqueue_prepare();
        check(hipStreamSynchronize(stream_), "prepare");
    } catch (...) {
        (void)hipStreamSynchronize(stream_);
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_exec_ = nullptr;
        graph_ = nullptr;
        for (void* p : allocations_) (void)hipFree(p);
        allocations_.clear();
        if (h_readout_) (void)hipHostFree(h_readout_);
        h_readout_ = nullptr;
        (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
        throw;
    }
}

PatchShape autotune_patch_shape(RunHost& r, const double* U_in, double dt0) {
    PatchShape best{mixed::default_threads(), kPatchFacesOneWave};
    double best_t = 0;
    for (const PatchShape& c : kPatchShapeCandidates) {
        try {
            relayout(r, mixed::slots(c.threads), c.faces);
        } catch (const FaceSetOverflow&) {
            continue;  // this shape's face sets do not fit the LDS face slots
        }
        Engine e(r, U_in, kMixed);
        const double t = time_trial([&](int n) { e.steps(dt0, n); }, [&] { e.synchronize(); }, 32);
        if (best_t == 0 || t < best_t) {
            best_t = t;
            best = c;
        }
    }
    if (r.L.max_faces != best.faces || r.L.slots != mixed::slots(best.threads))
        relayout(r, mixed::slots(best.threads), best.faces);
    return best;
}

Engine::~Engine() {
    // graph launches still in flight use the executable graph and every
    // buffer: wait for them before anything is released
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (void* p : allocations_) (void)hipFree(p);
    if (h_readout_) (void)hipHostFree(h_readout_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::upload(const RunHost& r) {
    const Geometry& g = r.geo;
    const int32_t nf = g.nf, nh = g.nh;
    const Layout& L = r.L;
    auto dev = [&](size_t bytes, const void* host) -> void* {
        void* p = nullptr;
        check(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc");
        allocations_.push_back(p);
        if (host && bytes) check(hipMemcpy(p, host, bytes, hipMemcpyHostToDevice), "upload");
        return p;
    };

    GeomPtrs& d = args_.geom;
    d.nf = nf;
    d.nh = nh;
    d.foff = static_cast<const int32_t*>(dev(sizeof(int32_t) * (nf + 1), L.foff.data()));
    d.he_fv = static_cast<const int32_t*>(dev(sizeof(int32_t) * nh, r.he_fv.data()));
    d.sten = static_cast<const int32_t*>(dev(sizeof(int32_t) * r.sten.size(), r.sten.data()));
    d.beta = static_cast<const double*>(dev(sizeof(double) * r.beta.size(), r.beta.data()));
    d.hsc = static_cast<const double*>(dev(sizeof(double) * r.hsc.size(), r.hsc.data()));
    d.ec = static_cast<const double*>(dev(sizeof(double) * r.ec.size(), r.ec.data()));
    d.en = static_cast<const double*>(dev(sizeof(double) * r.en.size(), r.en.data()));
    d.twin_slot = static_cast<const int32_t*>(dev(sizeof(int32_t) * L.twin_slot.size(), L.twin_slot.data()));
    d.area = static_cast<const double*>(dev(sizeof(double) * nf, r.area.data()));
    d.fcn = static_cast<const double*>(dev(sizeof(double) * r.fcn4.size(), r.fcn4.data()));
    d.acc = static_cast<const double*>(dev(sizeof(double) * r.acc.size(), r.acc.data()));
    d.blk_face = static_cast<const int32_t*>(dev(sizeof(int32_t) * L.blk_face.size(), L.blk_face.data()));
    d.halo_off = static_cast<const int32_t*>(dev(sizeof(int32_t) * L.halo_off.size(), L.halo_off.data()));
    d.halo_he = static_cast<const int32_t*>(dev(sizeof(int32_t) * L.halo_he.size(), L.halo_he.data()));
    d.slot_a = d.edge_rec = nullptr;
    d.slot_b = d.edge32 = nullptr;
    d.acc_rate = nullptr;
    d.inv_area = nullptr;
    d.fc32 = nullptr;
    d.fset_off = d.fset_face = nullptr;
    d.fset_max = 0;
    d.slot_max = 0;
    args_.W32 = args_.W32b = nullptr;
    if (precision_ == kMixed) {
        const size_t n_ext = L.ext_loc.size() / 4;
        // the kernel addresses every table with 32-bit byte offsets
        if (16 * n_ext >= (size_t(1) << 32) || 8 * size_t(nf) * size_t(dim_) >= (size_t(1) << 32))
            throw std::runtime_error("mesh too large for the mixed-precision path's 32-bit table offsets");
        const MixedTables m = pack_mixed_tables(r);
        d.slot_a = static_cast<const uint32_t*>(dev(sizeof(uint32_t) * m.slot_a.size(), m.slot_a.data()));
        d.slot_b = static_cast<const float*>(dev(sizeof(float) * m.slot_b.size(), m.slot_b.data()));
        d.edge32 = static_cast<const float*>(dev(sizeof(float) * m.edge32.size(), m.edge32.data()));
        d.edge_rec = static_cast<const uint32_t*>(dev(sizeof(uint32_t) * m.edge_rec.size(), m.edge_rec.data()));
        d.acc_rate = static_cast<const float*>(dev(sizeof(float) * m.rate.size(), m.rate.data()));
        d.inv_area = static_cast<const float*>(dev(sizeof(float) * m.inv_area.size(), m.inv_area.data()));
        d.fc32 = static_cast<const float*>(dev(sizeof(float) * m.fc32.size(), m.fc32.data()));
        d.fset_off = static_cast<const int32_t*>(dev(sizeof(int32_t) * L.fset_off.size(), L.fset_off.data()));
        d.fset_face = static_cast<const int32_t*>(dev(sizeof(int32_t) * L.fset_face.size(), L.fset_face.data()));
        d.fset_max = m.fset_max;
        d.slot_max = m.slot_max;
        const size_t row = dim_ == 5 ? 5 : 4;  // float4 (Sigma, L) rows, then the pressures
        args_.W32 = static_cast<float*>(dev(sizeof(float) * row * nf, nullptr));
        args_.W32b = static_cast<float*>(dev(sizeof(float) * row * nf, nullptr));
    }
    d.n_blocks = L.n_blocks();
    d.dt_partial = d.dt_partial2 = nullptr;
    d.dt_ticket = nullptr;
    d.dt_acc = nullptr;
    if (precision_ == kFP64) {
        // per-block partials and the dt kernel's own
        d.dt_partial = static_cast<double*>(dev(sizeof(double) * 2 * size_t(d.n_blocks), nullptr));
        const std::vector<double> p2(2 * size_t(fp64::finalize_blocks()), 0.0);
        const unsigned zero = 0;
        d.dt_partial2 = static_cast<double*>(dev(sizeof(double) * p2.size(), p2.data()));
        d.dt_ticket = static_cast<unsigned*>(dev(sizeof(unsigned), &zero));
    } else {
        // the two accumulator buffers (cleared again before each prepare)
        const std::vector<unsigned long long> keys(size_t(kDtRunWords), ~0ull);
        d.dt_acc = static_cast<unsigned long long*>(dev(sizeof(unsigned long long) * keys.size(), keys.data()));
    }
    d.blk_run = d.run_blk = nullptr;
    d.n_runs = 1;

    args_.phys = r.phys;
    args_.phys_dev = static_cast<const PhysParams*>(dev(sizeof(PhysParams), &r.phys));
    args_.dt0_dev = nullptr;  // eager steps take dt0 as a kernel argument
    {
        const double zero = 0;
        dt0_dev_ = static_cast<double*>(dev(sizeof(double), &zero));
    }
    args_.U = static_cast<double*>(dev(sizeof(double) * size_t(nf) * dim_, nullptr));
    args_.W = args_.W2 = nullptr;  // FP64 p-form stage inputs (the mixed path keeps FP32 ones)
    if (precision_ == kFP64) {
        args_.W = static_cast<double*>(dev(sizeof(double) * size_t(nf) * dim_, nullptr));
        args_.W2 = static_cast<double*>(dev(sizeof(double) * size_t(nf) * dim_, nullptr));
    }

    Scalars s{};
    s.first_err = ~0ull;
    args_.sc = static_cast<Scalars*>(dev(sizeof(Scalars), &s));

    d_readout_part_ = static_cast<double*>(dev(sizeof(double) * 9 * size_t(fp64::readout_blocks(nf)), nullptr));
    if (dim_ == 5) {
        double obs[16];
        light_observers(obs);
        d_light_mask_ = static_cast<uint8_t*>(dev(r.light_mask.size(), r.light_mask.data()));
        d_light_obs_ = static_cast<double*>(dev(sizeof(obs), obs));
    }
    check(hipHostMalloc(reinterpret_cast<void**>(&h_readout_), sizeof(Readout),
                        hipHostMallocMapped | hipHostMallocCoherent),
          "hipHostMalloc");
    check(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_readout_view_), h_readout_, 0),
          "hipHostGetDevicePointer");

    perm_ = L.perm;
    he_old_ = L.he_old;
}

void Engine::enqueue_prepare() {
    readout_valid_ = false;
    if (precision_ == kMixed) mixed::enqueue_prepare(args_, stream_);
    else fp64::enqueue_prepare(args_, stream_);
    check(hipGetLastError(), "prepare launch");
}

void Engine::enqueue_kernel(int which, double dt0, const KernelEvents* ev, const StepArgs* a) {
    const StepArgs& args = a ? *a : args_;
    if (precision_ == kFP64) {
        if (which == 0) fp64::enqueue_finalize_dt(args, dt0, stream_, ev);
        else fp64::enqueue_stage(args, which, stream_, ev);
    } else if (which > 0) {
        mixed::enqueue_stage(args, which, dt0, stream_, ev);  // stage 1 forms dt: no dt kernel
    }
}

void Engine::enqueue_step(double dt0) {
    for (int k = 0; k < kKernelsPerStep; ++k) enqueue_kernel(k, dt0);
}

void Engine::profile(double dt0, int n, double* ms) {
    readout_valid_ = false;
    // n steps launched back to back on this run's stream, each kernel with
    // start/stop HIP events bound to its own dispatch, one wait at the end
    constexpr int K = kKernelsPerStep;
    const int k0 = precision_ == kMixed ? 1 : 0;  // the mixed path launches no dt kernel
    const int kn = K - k0;
    const size_t nk = size_t(n > 0 ? n : 0) * size_t(kn);
    std::vector<hipEvent_t> ev(2 * nk, nullptr);
    double acc[K + 1] = {0};
    try {
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
        for (size_t q = 0; q < nk; ++q) {
            const KernelEvents ke{ev[2 * q], ev[2 * q + 1]};
            enqueue_kernel(k0 + int(q % kn), dt0, &ke);
        }
        check(hipGetLastError(), "profile launch");
        if (nk) check(hipEventSynchronize(ev.back()), "hipEventSynchronize");
        for (size_t q = 0; q < nk; ++q) {
            float t = 0;
            check(hipEventElapsedTime(&t, ev[2 * q], ev[2 * q + 1]), "hipEventElapsedTime");
            acc[k0 + q % kn] += t;
            if (q % kn == size_t(kn - 1)) {
                check(hipEventElapsedTime(&t, ev[2 * (q - kn + 1)], ev[2 * q + 1]), "hipEventElapsedTime");
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
    stepped_ = true;
    readout_valid_ = false;
    enqueue_step(dt0);
    check(hipGetLastError(), "step launch");
}

void Engine::capture_graph() {
    // the graph's dt kernel reads dt0 from dt0_dev_, so this one graph serves
    // every dt0; a failed capture is ended before the error propagates
    StepArgs ga = args_;
    ga.dt0_dev = dt0_dev_;
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        for (int k = 0; k < kGraphSteps; ++k)
            for (int w = 0; w < kKernelsPerStep; ++w) enqueue_kernel(w, 0.0, nullptr, &ga);
        check(hipGetLastError(), "capture launch");
    } catch (...) {
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(stream_, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
    }
    check(hipStreamEndCapture(stream_, &graph_), "hipStreamEndCapture");
    check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    // upload now (stream-ordered) so the first replay does not pay for it
    check(hipGraphUpload(graph_exec_, stream_), "hipGraphUpload");
}

void Engine::steps(double dt0, int n) {
    if (n <= 0) return;
    stepped_ = true;
    readout_valid_ = false;
    if (!graph_exec_) {
        // the first call launches its steps one by one and captures the
        // graph afterwards (nothing is launched while capturing), so every
        // kernel has run before it is captured and later calls replay it
        for (int k = 0; k < n; ++k) enqueue_step(dt0);
        check(hipGetLastError(), "steps launch");
        if (use_graph_) capture_graph();
        return;
    }
    if (n >= kGraphSteps) {
        if (dt0 != graph_dt0_ || !graph_dt0_set_) {
            // stream-ordered: replays already queued keep reading the old value
            fp64::enqueue_set_scalar(dt0_dev_, dt0, stream_);
            graph_dt0_ = dt0;
            graph_dt0_set_ = true;
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
    if (readout_valid_) return h_readout_->sc;
    check(hipMemcpyAsync(&h_readout_->sc, args_.sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream_),
          "read scalars");
    synchronize();
    return h_readout_->sc;
}

void Engine::ensure_readout() {
    if (readout_valid_) return;
    fp64::enqueue_readout(args_, d_light_mask_, d_light_obs_, d_readout_part_, d_readout_view_->sums,
                          &d_readout_view_->sc, stream_);
    check(hipGetLastError(), "readout launch");
    synchronize();
    readout_valid_ = true;
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
    std::vector<double> U(size_t(nf_) * dim_);
    check(hipMemcpyAsync(U.data(), args_.U, sizeof(double) * U.size(), hipMemcpyDeviceToHost, stream_), "read state");
    synchronize();
    for (int32_t k = 0; k < nf_; ++k)
        for (int c = 0; c < dim_; ++c) U_out[size_t(perm_[k]) * dim_ + c] = U[size_t(k) * dim_ + c];
}

void Engine::set_state(const double* U_in, double t, int stop_bits) {
    if (!U_in) throw std::invalid_argument("U_in is null");
    synchronize();  // steps in flight still read the old state
    readout_valid_ = false;
    std::vector<double> U(size_t(nf_) * dim_);
    for (int32_t k = 0; k < nf_; ++k)
        for (int c = 0; c < dim_; ++c) U[size_t(k) * dim_ + c] = U_in[size_t(perm_[k]) * dim_ + c];
    check(hipMemcpy(args_.U, U.data(), U.size() * sizeof(double), hipMemcpyHostToDevice), "upload state");
    Scalars s = read_scalars();
    s.t = t;
    s.stop = stop_bits;
    s.first_err = ~0ull;
    check(hipMemcpy(args_.sc, &s, sizeof(Scalars), hipMemcpyHostToDevice), "upload scalars");
    enqueue_prepare();  // stage inputs and dt partials of the new state
    synchronize();
    stepped_ = false;
}

void Engine::diagnostics(double out[5]) {
    ensure_readout();
    for (int k = 0; k < 5; ++k) out[k] = h_readout_->sums[k];
}

void Engine::light_curves(double out[4]) {
    if (dim_ != 5) throw std::invalid_argument("light curves exist for the adiabatic (dim 5) solvers only");
    ensure_readout();
    for (int k = 0; k < 4; ++k) out[k] = h_readout_->sums[5 + k];
}

void Engine::readout_noalloc(double* d_out, hipStream_t s) {
    if (!d_out) throw std::invalid_argument("d_out is null");
    if (nf_ > fp64::readout_noalloc_max_faces())
        throw std::invalid_argument("the allocation-free readout takes at most " +
                                    std::to_string(fp64::readout_noalloc_max_faces()) + " faces, this run has " +
                                    std::to_string(nf_));
    check(fp64::enqueue_readout_noalloc(args_, d_light_mask_, d_light_obs_, device_, d_out, s ? s : stream_),
          "allocation-free readout");
}

void Engine::curl(double* out) {
    if (!out) throw std::invalid_argument("out is null");
    if (!d_curl_) {
        if (geo_.nh != nh_ || geo_.ev.size() != 3 * size_t(nh_))
            throw std::logic_error("curl needs an Engine built from a mesh");
        const size_t n = 3 * size_t(nh_) + size_t(nf_) * (1 + dim_);
        std::vector<double> ev(3 * size_t(nh_));
        for (int32_t h = 0; h < nh_; ++h)
            for (int c = 0; c < 3; ++c) ev[3 * size_t(h) + c] = geo_.ev[3 * size_t(he_old_[h]) + c];
        void* p = nullptr;
        check(hipMalloc(&p, sizeof(double) * n), "hipMalloc");
        allocations_.push_back(p);
        d_curl_ = static_cast<double*>(p);
        check(hipMemcpy(d_curl_, ev.data(), sizeof(double) * ev.size(), hipMemcpyHostToDevice), "upload edges");
    }
    double* d_out = d_curl_ + 3 * size_t(nh_);
    fp64::enqueue_curl(args_, stepped_, d_curl_, d_out + nf_, d_out, stream_);
    check(hipGetLastError(), "curl launch");
    std::vector<double> v(nf_);
    check(hipMemcpyAsync(v.data(), d_out, sizeof(double) * v.size(), hipMemcpyDeviceToHost, stream_), "read curl");
    synchronize();
    for (int32_t k = 0; k < nf_; ++k) out[perm_[k]] = v[k];
}

}  // namespace splash
