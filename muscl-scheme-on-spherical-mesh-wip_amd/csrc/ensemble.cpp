#include "ensemble.hpp"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>

namespace splash {

namespace {

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
void append(std::vector<T>& dst, const std::vector<T>& src) {
    dst.insert(dst.end(), src.begin(), src.end());
}

}  // namespace

void* Ensemble::dev(size_t bytes, const void* host) {
    void* p = nullptr;
    check(hipMalloc(&p, bytes ? bytes : 16), "hipMalloc");
    allocations_.push_back(p);
    if (host && bytes) check(hipMemcpy(p, host, bytes, hipMemcpyHostToDevice), "upload");
    return p;
}

Ensemble::Ensemble(const std::vector<Run>& runs, int dim, int solver, int geometry_threads, int patch_faces,
                   int block_threads)
    : dim_(dim), solver_(solver) {
    if (runs.empty()) throw std::invalid_argument("an ensemble needs at least one run");
    if (runs.size() > 65536) throw std::invalid_argument("an ensemble holds at most 65536 runs");
    if (patch_faces < -1 || patch_faces > kPatchFacesMax) throw std::invalid_argument("patch_faces out of range");
    if (block_threads != 0 && !mixed::valid_threads(block_threads))
        throw std::invalid_argument("block_threads must be 0, 128, 192 or 256");
    const int32_t threads = block_threads ? block_threads : mixed::default_threads();
    std::vector<RunHost> hosts;
    std::vector<const double*> U_in;
    for (const Run& run : runs) {
        if (!run.mesh) throw std::invalid_argument("null mesh");
        // without a set cap each run takes its own mesh's default, falling
        // back to one face wave where its face sets overflow
        const int32_t cap = patch_faces > 0 ? patch_faces : default_patch_faces(*run.mesh, threads);
        hosts.push_back(build_run_host(*run.mesh, run.U_in, dim, solver, run.gam, run.omega_ns, run.params,
                                       mixed::slots(threads), cap, geometry_threads,
                                       patch_faces > 0 ? 0 : kPatchFacesOneWave));
        U_in.push_back(run.U_in);
    }
    if (patch_faces == -1) {
        // the fastest patch shape for the whole ensemble, measured on trial
        // ensembles with their own copies of the states
        PatchShape best{threads, kPatchFacesOneWave};
        double best_t = 0;
        std::vector<double> dt0(runs.size(), 1e-4);
        for (const PatchShape& c : kPatchShapeCandidates) {
            try {
                for (RunHost& h : hosts) relayout(h, mixed::slots(c.threads), c.faces);
            } catch (const FaceSetOverflow&) {
                continue;  // a run's face sets do not fit at this shape; the loop below re-lays every run
            }
            Ensemble trial(hosts, U_in, dim, solver);
            const double t =
                time_trial([&](int n) { trial.steps(dt0.data(), n); }, [&] { trial.synchronize(); }, 32);
            if (best_t == 0 || t < best_t) {
                best_t = t;
                best = c;
            }
        }
        for (RunHost& h : hosts)
            if (h.L.max_faces != best.faces || h.L.slots != mixed::slots(best.threads))
                relayout(h, mixed::slots(best.threads), best.faces);
    }
    init(hosts, U_in);
}

Ensemble::Ensemble(const std::vector<RunHost>& hosts, const std::vector<const double*>& U_in, int dim, int solver)
    : dim_(dim), solver_(solver) {
    init(hosts, U_in);
}

void Ensemble::init(const std::vector<RunHost>& hosts, const std::vector<const double*>& U_in) {
    const int dim = dim_;
    const int R = int(hosts.size());
    patch_faces_ = hosts[0].L.max_faces;
    block_threads_ = hosts[0].L.slots / 2;
    for (const RunHost& h : hosts) {
        if (h.L.slots != hosts[0].L.slots) throw std::invalid_argument("ensemble runs laid out for different blocks");
        if (h.L.max_faces > mixed::max_faces())
            throw std::logic_error("patch face cap exceeds the faces the stage kernel updates per patch");
    }

    // merged tables: every run's arrays appended, indices shifted by the
    // faces / half-edges / halo slots / face-set entries before it; only
    // the trailing padding blocks of the whole grid are empty
    std::vector<int32_t> foff{0}, blk_face, halo_off, fset_off, fset_face, blk_run, run_blk;
    std::vector<uint32_t> slot_a, edge_rec;
    std::vector<float> slot_b, edge32, inv_area, rate, fc32;
    std::vector<double> fcn4, area, U;
    std::vector<uint8_t> light_mask;
    std::vector<int2> seg;  // readout workgroups: each run's faces in blocks from its first face
    std::vector<int32_t> seg_run, run_part;
    const int32_t kReadoutFaces = fp64::readout_faces();
    std::vector<PhysParams> phys;
    int32_t F = 0, H = 0, halo = 0, fsets = 0, fset_max = 1, slot_max = 4;
    for (int r = 0; r < R; ++r) {
        const RunHost& rh = hosts[size_t(r)];
        if (rh.dim != dim || rh.solver != solver_) throw std::invalid_argument("ensemble runs differ in dim or solver");
        const MixedTables m = pack_mixed_tables(rh);
        const Layout& L = rh.L;
        const int32_t nf = rh.geo.nf, nh = rh.geo.nh;
        int32_t nb = 0;  // the run's blocks without its own trailing padding
        while (nb < L.n_blocks() && L.blk_face[nb + 1] > L.blk_face[nb]) ++nb;
        if (int64_t(slot_a.size()) / 2 != int64_t(H) + halo) throw std::logic_error("ensemble row offsets");
        run_blk.push_back(int32_t(blk_face.size()));
        for (int32_t b = 0; b < nb; ++b) {
            blk_face.push_back(L.blk_face[b] + F);
            halo_off.push_back(L.halo_off[b] + halo);
            fset_off.push_back(L.fset_off[b] + fsets);
            blk_run.push_back(r);
        }
        for (int32_t k = 1; k <= nf; ++k) foff.push_back(L.foff[k] + H);
        for (int32_t f : L.fset_face) fset_face.push_back(f + F);
        append(slot_a, m.slot_a);
        append(slot_b, m.slot_b);
        append(edge32, m.edge32);
        append(edge_rec, m.edge_rec);
        append(rate, m.rate);
        append(inv_area, m.inv_area);
        append(fc32, m.fc32);
        append(fcn4, rh.fcn4);
        append(area, rh.area);
        if (dim == 5) append(light_mask, rh.light_mask);
        run_part.push_back(int32_t(seg.size()));
        for (int32_t f = 0; f < nf; f += kReadoutFaces) {
            seg.push_back(make_int2(F + f, F + std::min(nf, f + kReadoutFaces)));
            seg_run.push_back(r);
        }
        for (int32_t k = 0; k < nf; ++k)
            for (int c = 0; c < dim; ++c) U.push_back(U_in[size_t(r)][size_t(L.perm[k]) * dim + c]);
        phys.push_back(rh.phys);
        fset_max = std::max(fset_max, m.fset_max);
        slot_max = std::max(slot_max, m.slot_max);
        faces_.push_back(nf);
        face_off_.push_back(F);
        perm_.push_back(L.perm);
        F += nf;
        H += nh;
        halo += int32_t(L.halo_he.size());
        fsets += int32_t(L.fset_face.size());
    }
    run_part.push_back(int32_t(seg.size()));
    run_blk.push_back(int32_t(blk_face.size()));
    blk_face.push_back(F);
    halo_off.push_back(halo);
    fset_off.push_back(fsets);
    while ((blk_face.size() - 1) % 8) {  // the XCD-aware block numbering needs a multiple of 8
        blk_face.push_back(F);
        halo_off.push_back(halo);
        fset_off.push_back(fsets);
        blk_run.push_back(R - 1);
    }
    if (16 * (size_t(H) + size_t(halo)) >= (size_t(1) << 32) || 8 * size_t(F) * size_t(dim) >= (size_t(1) << 32))
        throw std::runtime_error("ensemble too large for the mixed-precision path's 32-bit table offsets");

    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    try {
        GeomPtrs& d = args_.geom;
        d = GeomPtrs{};
        d.nf = F;
        d.nh = H;
        d.foff = static_cast<const int32_t*>(dev(sizeof(int32_t) * foff.size(), foff.data()));
        d.area = static_cast<const double*>(dev(sizeof(double) * area.size(), area.data()));
        d.fcn = static_cast<const double*>(dev(sizeof(double) * fcn4.size(), fcn4.data()));
        d.blk_face = static_cast<const int32_t*>(dev(sizeof(int32_t) * blk_face.size(), blk_face.data()));
        d.halo_off = static_cast<const int32_t*>(dev(sizeof(int32_t) * halo_off.size(), halo_off.data()));
        d.slot_a = static_cast<const uint32_t*>(dev(sizeof(uint32_t) * slot_a.size(), slot_a.data()));
        d.slot_b = static_cast<const float*>(dev(sizeof(float) * slot_b.size(), slot_b.data()));
        d.edge32 = static_cast<const float*>(dev(sizeof(float) * edge32.size(), edge32.data()));
        d.edge_rec = static_cast<const uint32_t*>(dev(sizeof(uint32_t) * edge_rec.size(), edge_rec.data()));
        d.acc_rate = static_cast<const float*>(dev(sizeof(float) * rate.size(), rate.data()));
        d.inv_area = static_cast<const float*>(dev(sizeof(float) * inv_area.size(), inv_area.data()));
        d.fc32 = static_cast<const float*>(dev(sizeof(float) * fc32.size(), fc32.data()));
        d.fset_off = static_cast<const int32_t*>(dev(sizeof(int32_t) * fset_off.size(), fset_off.data()));
        d.fset_face = static_cast<const int32_t*>(dev(sizeof(int32_t) * fset_face.size(), fset_face.data()));
        d.fset_max = fset_max;
        d.slot_max = slot_max;
        d.n_blocks = int32_t(blk_face.size()) - 1;
        // the mixed path's per-run dt accumulators (no dt kernel)
        d.dt_partial = d.dt_partial2 = nullptr;
        d.dt_ticket = nullptr;
        const std::vector<unsigned long long> keys(size_t(R) * kDtRunWords, ~0ull);
        d.dt_acc = static_cast<unsigned long long*>(dev(sizeof(unsigned long long) * keys.size(), keys.data()));
        d.blk_run = static_cast<const int32_t*>(dev(sizeof(int32_t) * blk_run.size(), blk_run.data()));
        d.run_blk = static_cast<const int32_t*>(dev(sizeof(int32_t) * run_blk.size(), run_blk.data()));
        d.n_runs = R;

        args_.phys = phys[0];
        args_.phys_dev = static_cast<const PhysParams*>(dev(sizeof(PhysParams) * phys.size(), phys.data()));
        dt0_host_.assign(size_t(R), 0.0);
        dt0_dev_ = static_cast<double*>(dev(sizeof(double) * size_t(R), dt0_host_.data()));
        args_.dt0_dev = dt0_dev_;
        std::vector<Scalars> sc(static_cast<size_t>(R));
        for (auto& s : sc) {
            s = Scalars{};
            s.first_err = ~0ull;
        }
        args_.sc = static_cast<Scalars*>(dev(sizeof(Scalars) * sc.size(), sc.data()));
        args_.U = static_cast<double*>(dev(sizeof(double) * U.size(), U.data()));
        args_.W = args_.W2 = nullptr;
        n_seg_ = int(seg.size());
        d_seg_ = static_cast<const int2*>(dev(sizeof(int2) * seg.size(), seg.data()));
        d_seg_run_ = static_cast<const int*>(dev(sizeof(int32_t) * seg_run.size(), seg_run.data()));
        d_run_part_ = static_cast<const int*>(dev(sizeof(int32_t) * run_part.size(), run_part.data()));
        if (dim == 5) {
            double obs[16];
            light_observers(obs);
            d_light_mask_ = static_cast<uint8_t*>(dev(light_mask.size(), light_mask.data()));
            d_light_obs_ = static_cast<double*>(dev(sizeof(obs), obs));
        }
        d_readout_part_ = static_cast<double*>(dev(sizeof(double) * 9 * seg.size(), nullptr));
        check(hipHostMalloc(reinterpret_cast<void**>(&h_sums_), sizeof(double) * 9 * size_t(R),
                            hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void**>(&h_scalars_), sizeof(Scalars) * size_t(R),
                            hipHostMallocMapped | hipHostMallocCoherent),
              "hipHostMalloc");
        check(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_sums_view_), h_sums_, 0), "hipHostGetDevicePointer");
        check(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_scalars_view_), h_scalars_, 0),
              "hipHostGetDevicePointer");
        const size_t row = dim == 5 ? 5 : 4;  // float4 (Sigma, L) rows, then the pressures
        args_.W32 = static_cast<float*>(dev(sizeof(float) * row * size_t(F), nullptr));
        args_.W32b = static_cast<float*>(dev(sizeof(float) * row * size_t(F), nullptr));
        args_.block_threads = block_threads_;

        mixed::enqueue_prepare(args_, stream_);
        check(hipGetLastError(), "prepare launch");
        check(hipStreamSynchronize(stream_), "prepare");
    } catch (...) {
        (void)hipStreamSynchronize(stream_);
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_exec_ = nullptr;
        graph_ = nullptr;
        for (void* p : allocations_) (void)hipFree(p);
        if (h_sums_) (void)hipHostFree(h_sums_);
        if (h_scalars_) (void)hipHostFree(h_scalars_);
        (void)hipStreamDestroy(stream_);
        throw;
    }
}

Ensemble::~Ensemble() {
    // graph launches still in flight use the executable graph and every
    // buffer: wait for them before anything is released
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (void* p : allocations_) (void)hipFree(p);
    if (h_sums_) (void)hipHostFree(h_sums_);
    if (h_scalars_) (void)hipHostFree(h_scalars_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Ensemble::upload_dt0(const double* dt0) {
    if (!dt0) throw std::invalid_argument("dt0 is null");
    if (std::equal(dt0_host_.begin(), dt0_host_.end(), dt0)) return;
    // the previous steps still read the old values: wait for them first
    check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    dt0_host_.assign(dt0, dt0 + dt0_host_.size());
    check(hipMemcpy(dt0_dev_, dt0_host_.data(), sizeof(double) * dt0_host_.size(), hipMemcpyHostToDevice),
          "upload dt0");
}

void Ensemble::enqueue_kernel(int which, const KernelEvents* ev) {
    // stage 1 forms each run's dt (dt0 from the per-run array): no dt kernel
    if (which > 0) mixed::enqueue_stage(args_, which, 0.0, stream_, ev);
}

void Ensemble::enqueue_step() {
    for (int k = 0; k < kKernelsPerStep; ++k) enqueue_kernel(k);
}

void Ensemble::capture_graph() {
    // dt0 lives in device memory, so this one graph serves every call; a
    // failed capture is ended before the error propagates
    check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
        for (int k = 0; k < kGraphSteps; ++k) enqueue_step();
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

void Ensemble::steps(const double* dt0, int n) {
    upload_dt0(dt0);
    if (n <= 0) return;
    readout_valid_ = false;
    if (!graph_exec_) {
        // the first call launches its steps one by one and captures the
        // graph afterwards, so every kernel has run before it is captured
        for (int k = 0; k < n; ++k) enqueue_step();
        check(hipGetLastError(), "steps launch");
        if (use_graph_) capture_graph();
        return;
    }
    while (n >= kGraphSteps) {
        check(hipGraphLaunch(graph_exec_, stream_), "hipGraphLaunch");
        n -= kGraphSteps;
    }
    for (int k = 0; k < n; ++k) enqueue_step();
    check(hipGetLastError(), "steps launch");
}

void Ensemble::profile(const double* dt0, int n, double* ms) {
    upload_dt0(dt0);
    readout_valid_ = false;
    constexpr int K = kKernelsPerStep;
    constexpr int k0 = 1, kn = K - k0;  // no dt kernel: the two stage launches
    const size_t nk = size_t(n > 0 ? n : 0) * size_t(kn);
    std::vector<hipEvent_t> ev(2 * nk, nullptr);
    double acc[K + 1] = {0};
    try {
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
        for (size_t q = 0; q < nk; ++q) {
            const KernelEvents ke{ev[2 * q], ev[2 * q + 1]};
            enqueue_kernel(k0 + int(q % kn), &ke);
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

void Ensemble::synchronize() { check(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

Scalars Ensemble::read_scalars(int r) {
    if (r < 0 || r >= n_runs()) throw std::out_of_range("run index");
    if (readout_valid_) return h_scalars_[r];
    Scalars s;
    check(hipMemcpyAsync(&s, args_.sc + r, sizeof(Scalars), hipMemcpyDeviceToHost, stream_), "read scalars");
    synchronize();
    return s;
}

void Ensemble::ensure_readout() {
    if (readout_valid_) return;
    fp64::enqueue_readout_runs(args_, d_light_mask_, d_light_obs_, d_seg_, d_seg_run_, d_run_part_, n_seg_,
                               d_readout_part_, d_sums_view_, d_scalars_view_, stream_);
    check(hipGetLastError(), "readout launch");
    synchronize();
    readout_valid_ = true;
}

void Ensemble::diagnostics(int r, double out[5]) {
    if (r < 0 || r >= n_runs()) throw std::out_of_range("run index");
    ensure_readout();
    for (int k = 0; k < 5; ++k) out[k] = h_sums_[9 * size_t(r) + k];
}

void Ensemble::light_curves(int r, double out[4]) {
    if (r < 0 || r >= n_runs()) throw std::out_of_range("run index");
    if (dim_ != 5) throw std::invalid_argument("light curves exist for the adiabatic (dim 5) solvers only");
    ensure_readout();
    for (int k = 0; k < 4; ++k) out[k] = h_sums_[9 * size_t(r) + 5 + k];
}

double Ensemble::time(int r) { return read_scalars(r).t; }
double Ensemble::last_dt(int r) { return read_scalars(r).dt; }
int Ensemble::stop_bits(int r) { return read_scalars(r).stop; }

void Ensemble::set_state(int r, const double* U_in, double t, int stop_bits) {
    if (r < 0 || r >= n_runs()) throw std::out_of_range("run index");
    if (!U_in) throw std::invalid_argument("U_in is null");
    synchronize();  // steps in flight still read the old state
    readout_valid_ = false;
    const int32_t nf = faces_[size_t(r)];
    const std::vector<int32_t>& perm = perm_[size_t(r)];
    std::vector<double> U(size_t(nf) * dim_);
    for (int32_t k = 0; k < nf; ++k)
        for (int c = 0; c < dim_; ++c) U[size_t(k) * dim_ + c] = U_in[size_t(perm[k]) * dim_ + c];
    check(hipMemcpy(args_.U + size_t(face_off_[size_t(r)]) * dim_, U.data(), U.size() * sizeof(double),
                    hipMemcpyHostToDevice),
          "upload state");
    Scalars s = read_scalars(r);
    s.t = t;
    s.stop = stop_bits;
    s.first_err = ~0ull;
    check(hipMemcpy(args_.sc + r, &s, sizeof(Scalars), hipMemcpyHostToDevice), "upload scalars");
    // stage inputs and dt partials of every run: the prepare kernel uses the
    // stage epilogue's arithmetic, so the other runs' values do not change
    mixed::enqueue_prepare(args_, stream_);
    check(hipGetLastError(), "prepare launch");
    synchronize();
}

void Ensemble::get_state(int r, double* U_out) {
    if (r < 0 || r >= n_runs()) throw std::out_of_range("run index");
    const int32_t nf = faces_[size_t(r)];
    std::vector<double> U(size_t(nf) * dim_);
    check(hipMemcpyAsync(U.data(), args_.U + size_t(face_off_[size_t(r)]) * dim_, U.size() * sizeof(double),
                         hipMemcpyDeviceToHost, stream_),
          "read state");
    synchronize();
    const std::vector<int32_t>& perm = perm_[size_t(r)];
    for (int32_t k = 0; k < nf; ++k)
        for (int c = 0; c < dim_; ++c) U_out[size_t(perm[k]) * dim_ + c] = U[size_t(k) * dim_ + c];
}

}  // namespace splash
