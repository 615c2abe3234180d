// extern "C" boundary (include/splash.h): converts exceptions to error codes.
#include "../../include/splash.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"
#include "ensemble.hpp"
#include "geometry.hpp"
#include "launch.hpp"
#include "layout.hpp"
#include "mesh.hpp"

struct splash_mesh {
    splash::PolyMesh mesh;
};
struct splash_ctx {
    std::unique_ptr<splash::Engine> engine;
};
struct splash_geometry {
    splash::Geometry geo;
};
struct splash_ensemble {
    std::unique_ptr<splash::Ensemble> ens;
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* msg) {
    g_last_error = msg;
    return code;
}

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const splash::HipError& e) {
        return fail(SPLASH_EDEVICE, e.what());
    } catch (const std::bad_alloc&) {
        return fail(SPLASH_ENOMEM, "out of host memory");
    } catch (const std::invalid_argument& e) {
        return fail(SPLASH_EINVAL, e.what());
    } catch (const std::runtime_error& e) {
        return fail(SPLASH_EMESH, e.what());
    } catch (const std::exception& e) {
        return fail(SPLASH_EINTERNAL, e.what());
    } catch (...) {
        return fail(SPLASH_EINTERNAL, "unknown error");
    }
}

splash::RunParams run_params(const splash_params& p) {
    splash::RunParams rp;
    rp.accretion_on = p.accretion_on;
    rp.friction_on = p.friction_on;
    rp.acc_rate = p.acc_rate;
    rp.e_acc = p.e_acc;
    rp.omega_acc_abs = p.omega_acc_abs;
    rp.tilt_angle = p.tilt_angle;
    rp.acc_width = p.acc_width;
    rp.alpha = p.alpha;
    return rp;
}

splash::PolyMesh from_desc(const splash_mesh_desc* d) {
    if (!d || d->n_vertices <= 0 || d->n_faces <= 0 || !d->vertices || !d->face_offsets || !d->face_vertices)
        throw std::invalid_argument("invalid mesh descriptor");
    splash::PolyMesh m;
    m.xyz.assign(d->vertices, d->vertices + 3 * size_t(d->n_vertices));
    m.foff.assign(d->face_offsets, d->face_offsets + d->n_faces + 1);
    if (m.foff[0] != 0) throw std::invalid_argument("face_offsets[0] must be 0");
    for (int32_t f = 0; f < d->n_faces; ++f)
        if (m.foff[f + 1] - m.foff[f] < 3) throw std::invalid_argument("every face needs at least 3 vertices");
    m.fv.assign(d->face_vertices, d->face_vertices + m.foff.back());
    for (int32_t v : m.fv)
        if (v < 0 || v >= d->n_vertices) throw std::invalid_argument("face vertex index out of range");
    return m;
}

int make_mesh(splash::PolyMesh (*gen)(int), int n, splash_mesh** out) {
    if (!out) return fail(SPLASH_EINVAL, "out is null");
    return guarded([&] {
        auto m = std::make_unique<splash_mesh>();
        m->mesh = gen(n);
        *out = m.release();
        return SPLASH_OK;
    });
}

}  // namespace

extern "C" {

void splash_default_params(splash_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    splash::RunParams d;
    p->accretion_on = 0;
    p->friction_on = d.friction_on;
    p->acc_rate = d.acc_rate;
    p->e_acc = d.e_acc;
    p->omega_acc_abs = d.omega_acc_abs;
    p->tilt_angle = d.tilt_angle;
    p->acc_width = d.acc_width;
    p->alpha = d.alpha;
    p->precision = SPLASH_FP64;
    p->setup_threads = 0;
    p->patch_faces = 0;
    p->no_graph = 0;
    p->block_threads = 0;
}

int splash_mesh_uv_sphere(int n_slices, int n_stacks, splash_mesh** out) {
    if (!out) return fail(SPLASH_EINVAL, "out is null");
    return guarded([&] {
        auto m = std::make_unique<splash_mesh>();
        m->mesh = splash::uv_sphere(n_slices, n_stacks);
        *out = m.release();
        return SPLASH_OK;
    });
}
int splash_mesh_icosphere(int n, splash_mesh** out) { return make_mesh(splash::icosphere, n, out); }
int splash_mesh_icosphere_hex(int n, splash_mesh** out) { return make_mesh(splash::icosphere_hex, n, out); }
int splash_mesh_quad_sphere(int n, splash_mesh** out) { return make_mesh(splash::quad_sphere, n, out); }

int splash_mesh_view(const splash_mesh* m, splash_mesh_desc* out) {
    if (!m || !out) return fail(SPLASH_EINVAL, "null argument");
    out->n_vertices = m->mesh.n_vertices();
    out->n_faces = m->mesh.n_faces();
    out->vertices = m->mesh.xyz.data();
    out->face_offsets = m->mesh.foff.data();
    out->face_vertices = m->mesh.fv.data();
    return SPLASH_OK;
}

void splash_mesh_free(splash_mesh* m) { delete m; }

int splash_create(const splash_mesh_desc* mesh, const double* U_in, int dim, int solver, double gam, double omega_ns,
                  const splash_params* params, splash_ctx** out) {
    if (!out) return fail(SPLASH_EINVAL, "out is null");
    *out = nullptr;
    return guarded([&] {
        splash_params p;
        if (params) p = *params;
        else splash_default_params(&p);
        const splash::RunParams rp = run_params(p);
        splash::PolyMesh m = from_desc(mesh);
        auto ctx = std::make_unique<splash_ctx>();
        ctx->engine = std::make_unique<splash::Engine>(m, U_in, dim, solver, gam, omega_ns, rp, p.precision,
                                                       p.setup_threads, p.patch_faces, p.block_threads);
        ctx->engine->set_use_graph(!p.no_graph);
        *out = ctx.release();
        return SPLASH_OK;
    });
}

void splash_destroy(splash_ctx* ctx) { delete ctx; }

#define CTX_OR_FAIL                                                    \
    if (!ctx || !ctx->engine) return fail(SPLASH_EINVAL, "null context")

int splash_do_step(splash_ctx* ctx, double dt0) {
    CTX_OR_FAIL;
    return guarded([&] {
        ctx->engine->step(dt0);
        return SPLASH_OK;
    });
}

int splash_do_steps(splash_ctx* ctx, double dt0, int n) {
    CTX_OR_FAIL;
    if (n < 0) return fail(SPLASH_EINVAL, "negative step count");
    return guarded([&] {
        ctx->engine->steps(dt0, n);
        return SPLASH_OK;
    });
}

int splash_profile_step(splash_ctx* ctx, double dt0, int n, double* ms_out) {
    CTX_OR_FAIL;
    if (n < 0 || !ms_out) return fail(SPLASH_EINVAL, "bad profile arguments");
    return guarded([&] {
        ctx->engine->profile(dt0, n, ms_out);
        return SPLASH_OK;
    });
}

int splash_synchronize(splash_ctx* ctx) {
    CTX_OR_FAIL;
    return guarded([&] {
        ctx->engine->synchronize();
        return SPLASH_OK;
    });
}

int splash_time(splash_ctx* ctx, double* t) {
    CTX_OR_FAIL;
    if (!t) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        *t = ctx->engine->time();
        return SPLASH_OK;
    });
}

int splash_last_dt(splash_ctx* ctx, double* dt) {
    CTX_OR_FAIL;
    if (!dt) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        *dt = ctx->engine->last_dt();
        return SPLASH_OK;
    });
}

int splash_stop_check(splash_ctx* ctx, int* stop) {
    CTX_OR_FAIL;
    if (!stop) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        *stop = ctx->engine->stop_bits() != 0;
        return SPLASH_OK;
    });
}

int splash_stop_bits(splash_ctx* ctx, int* bits, uint64_t* first) {
    CTX_OR_FAIL;
    return guarded([&] {
        if (bits) *bits = ctx->engine->stop_bits();
        if (first) *first = ctx->engine->first_error();
        return SPLASH_OK;
    });
}

int splash_get_state(splash_ctx* ctx, double* U_out) {
    CTX_OR_FAIL;
    if (!U_out) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        ctx->engine->get_state(U_out);
        return SPLASH_OK;
    });
}

int splash_set_state(splash_ctx* ctx, const double* U_in, double t, int stop_bits) {
    CTX_OR_FAIL;
    if (!U_in) return fail(SPLASH_EINVAL, "null state");
    return guarded([&] {
        ctx->engine->set_state(U_in, t, stop_bits);
        return SPLASH_OK;
    });
}

int splash_diagnostics(splash_ctx* ctx, double out[5]) {
    CTX_OR_FAIL;
    if (!out) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        ctx->engine->diagnostics(out);
        return SPLASH_OK;
    });
}

int splash_num_faces(const splash_ctx* ctx, int32_t* n_faces, int32_t* n_halfedges) {
    CTX_OR_FAIL;
    if (n_faces) *n_faces = ctx->engine->n_faces();
    if (n_halfedges) *n_halfedges = ctx->engine->n_halfedges();
    return SPLASH_OK;
}

int splash_light_curves(splash_ctx* ctx, double out[4]) {
    CTX_OR_FAIL;
    if (!out) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        ctx->engine->light_curves(out);
        return SPLASH_OK;
    });
}

int splash_readout_noalloc(splash_ctx* ctx, double* d_out, void* stream) {
    CTX_OR_FAIL;
    if (!d_out) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        ctx->engine->readout_noalloc(d_out, static_cast<hipStream_t>(stream));
        return SPLASH_OK;
    });
}

int splash_curl(splash_ctx* ctx, double* out) {
    CTX_OR_FAIL;
    if (!out) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        ctx->engine->curl(out);
        return SPLASH_OK;
    });
}

int splash_patch_shape(const splash_ctx* ctx, int32_t* patch_faces, int32_t* block_threads) {
    CTX_OR_FAIL;
    if (!patch_faces || !block_threads) return fail(SPLASH_EINVAL, "null output");
    const bool mixed = ctx->engine->precision() == SPLASH_MIXED;
    *patch_faces = mixed ? ctx->engine->patch_faces() : 0;
    *block_threads = mixed ? ctx->engine->block_threads() : 0;
    return SPLASH_OK;
}

int splash_patch_faces(const splash_ctx* ctx, int32_t* patch_faces) {
    CTX_OR_FAIL;
    if (!patch_faces) return fail(SPLASH_EINVAL, "null output");
    *patch_faces = ctx->engine->precision() == SPLASH_MIXED ? ctx->engine->patch_faces() : 0;
    return SPLASH_OK;
}

int splash_face_geometry(const splash_ctx* ctx, double* centers_out, double* areas_out) {
    CTX_OR_FAIL;
    const splash::Geometry& g = ctx->engine->geometry();
    if (centers_out) std::memcpy(centers_out, g.fc.data(), sizeof(double) * g.fc.size());
    if (areas_out) std::memcpy(areas_out, g.area.data(), sizeof(double) * g.area.size());
    return SPLASH_OK;
}

int splash_geometry_create(const splash_mesh_desc* mesh, int threads, splash_geometry** out) {
    if (!out) return fail(SPLASH_EINVAL, "out is null");
    *out = nullptr;
    return guarded([&] {
        auto g = std::make_unique<splash_geometry>();
        g->geo = splash::build_geometry(from_desc(mesh), threads);
        *out = g.release();
        return SPLASH_OK;
    });
}

int splash_geometry_get_view(const splash_geometry* g, splash_geometry_view* v) {
    if (!g || !v) return fail(SPLASH_EINVAL, "null argument");
    const splash::Geometry& G = g->geo;
    v->nf = G.nf;
    v->nh = G.nh;
    v->fc = G.fc.data();
    v->normal = G.normal.data();
    v->area = G.area.data();
    v->nb_edge = G.nb_edge.data();
    v->twin = G.twin.data();
    v->ff_plus = G.ff_plus.data();
    v->ff_minus = G.ff_minus.data();
    v->beta_plus = G.beta_plus.data();
    v->beta_minus = G.beta_minus.data();
    v->h_plus = G.h_plus.data();
    v->h_minus = G.h_minus.data();
    v->bm_dist = G.bm_dist.data();
    v->en = G.en.data();
    v->h0 = G.h0;
    v->round_diff = G.round_diff;
    return SPLASH_OK;
}

void splash_geometry_free(splash_geometry* g) { delete g; }

namespace {

int layout_stats(const splash_mesh_desc* mesh, int32_t threads, int32_t faces, int32_t out[5]) {
    if (!out) return fail(SPLASH_EINVAL, "null output");
    if (threads == 0) threads = splash::mixed::default_threads();
    if (!splash::mixed::valid_threads(threads)) return fail(SPLASH_EINVAL, "block_threads must be 0, 128, 192 or 256");
    if (faces < 0 || faces > splash::kPatchFacesMax) return fail(SPLASH_EINVAL, "patch_faces out of range");
    return guarded([&] {
        const splash::PolyMesh m = from_desc(mesh);
        const splash::Geometry g = splash::build_geometry(m, 0);
        splash::Layout L;
        try {
            L = splash::build_layout(g, splash::mixed::slots(threads),
                                     faces ? faces : splash::default_patch_faces(m, threads));
        } catch (const splash::FaceSetOverflow&) {
            if (faces) throw;  // a defaulted cap falls back to one face wave, as the solvers do
            L = splash::build_layout(g, splash::mixed::slots(threads), splash::kPatchFacesOneWave);
        }
        splash::validate_layout(g, L);
        int32_t max_halo = 0, nonempty = 0;
        for (int32_t b = 0; b < L.n_blocks(); ++b) {
            max_halo = std::max(max_halo, L.halo_off[b + 1] - L.halo_off[b]);
            nonempty += L.blk_face[b + 1] > L.blk_face[b];
        }
        out[0] = L.n_blocks();
        out[1] = nonempty;
        out[2] = int32_t(L.halo_he.size());
        out[3] = max_halo;
        int32_t max_fset = 0;
        for (int32_t b = 0; b < L.n_blocks(); ++b) max_fset = std::max(max_fset, L.fset_off[b + 1] - L.fset_off[b]);
        out[4] = max_fset;
        return SPLASH_OK;
    });
}

}  // namespace

int splash_layout_stats(const splash_mesh_desc* mesh, int32_t out[5]) {
    return layout_stats(mesh, 0, splash::kPatchFacesOneWave, out);
}

int splash_layout_stats_shape(const splash_mesh_desc* mesh, int32_t block_threads, int32_t patch_faces,
                              int32_t out[5]) {
    return layout_stats(mesh, block_threads, patch_faces, out);
}

int splash_ensemble_create(int n_runs, const splash_mesh_desc* meshes, const double* const* U_in, int dim,
                           int solver, const double* gam, const double* omega_ns, const splash_params* params,
                           splash_ensemble** out) {
    if (!out) return fail(SPLASH_EINVAL, "out is null");
    *out = nullptr;
    if (n_runs <= 0 || !meshes || !U_in || !gam || !omega_ns || !params)
        return fail(SPLASH_EINVAL, "bad ensemble arguments");
    return guarded([&] {
        std::vector<splash::PolyMesh> m;
        m.reserve(size_t(n_runs));
        std::vector<splash::Ensemble::Run> runs;
        for (int r = 0; r < n_runs; ++r) {
            m.push_back(from_desc(&meshes[r]));
            runs.push_back({&m.back(), U_in[r], gam[r], omega_ns[r], run_params(params[r])});
        }
        auto e = std::make_unique<splash_ensemble>();
        e->ens = std::make_unique<splash::Ensemble>(runs, dim, solver, params[0].setup_threads,
                                                    params[0].patch_faces, params[0].block_threads);
        e->ens->set_use_graph(!params[0].no_graph);
        *out = e.release();
        return SPLASH_OK;
    });
}

void splash_ensemble_destroy(splash_ensemble* e) { delete e; }

#define ENS_OR_FAIL                                                    \
    if (!e || !e->ens) return fail(SPLASH_EINVAL, "null ensemble")
#define RUN_OR_FAIL                                                    \
    if (run < 0 || run >= e->ens->n_runs()) return fail(SPLASH_EINVAL, "run index out of range")

int splash_ensemble_step(splash_ensemble* e, const double* dt0, int nsteps) {
    ENS_OR_FAIL;
    if (!dt0 || nsteps < 0) return fail(SPLASH_EINVAL, "bad step arguments");
    return guarded([&] {
        e->ens->steps(dt0, nsteps);
        return SPLASH_OK;
    });
}

int splash_ensemble_profile_step(splash_ensemble* e, const double* dt0, int n, double* ms_out) {
    ENS_OR_FAIL;
    if (!dt0 || n < 0 || !ms_out) return fail(SPLASH_EINVAL, "bad profile arguments");
    return guarded([&] {
        e->ens->profile(dt0, n, ms_out);
        return SPLASH_OK;
    });
}

int splash_ensemble_synchronize(splash_ensemble* e) {
    ENS_OR_FAIL;
    return guarded([&] {
        e->ens->synchronize();
        return SPLASH_OK;
    });
}

int splash_ensemble_info(const splash_ensemble* e, int32_t* n_runs, int64_t* total_faces, int64_t* total_halfedges,
                         int32_t* n_blocks) {
    ENS_OR_FAIL;
    if (n_runs) *n_runs = e->ens->n_runs();
    if (total_faces) *total_faces = e->ens->total_faces();
    if (total_halfedges) *total_halfedges = e->ens->total_halfedges();
    if (n_blocks) *n_blocks = e->ens->n_blocks();
    return SPLASH_OK;
}

int splash_ensemble_num_faces(const splash_ensemble* e, int run, int32_t* n_faces) {
    ENS_OR_FAIL;
    RUN_OR_FAIL;
    if (!n_faces) return fail(SPLASH_EINVAL, "null output");
    *n_faces = e->ens->n_faces(run);
    return SPLASH_OK;
}

int splash_ensemble_patch_shape(const splash_ensemble* e, int32_t* patch_faces, int32_t* block_threads) {
    ENS_OR_FAIL;
    if (!patch_faces || !block_threads) return fail(SPLASH_EINVAL, "null output");
    *patch_faces = e->ens->patch_faces();
    *block_threads = e->ens->block_threads();
    return SPLASH_OK;
}

int splash_ensemble_patch_faces(const splash_ensemble* e, int32_t* patch_faces) {
    ENS_OR_FAIL;
    if (!patch_faces) return fail(SPLASH_EINVAL, "null output");
    *patch_faces = e->ens->patch_faces();
    return SPLASH_OK;
}

int splash_ensemble_time(splash_ensemble* e, int run, double* t) {
    ENS_OR_FAIL;
    RUN_OR_FAIL;
    if (!t) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        *t = e->ens->time(run);
        return SPLASH_OK;
    });
}

int splash_ensemble_last_dt(splash_ensemble* e, int run, double* dt) {
    ENS_OR_FAIL;
    RUN_OR_FAIL;
    if (!dt) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        *dt = e->ens->last_dt(run);
        return SPLASH_OK;
    });
}

int splash_ensemble_stop_check(splash_ensemble* e, int run, int* stop) {
    ENS_OR_FAIL;
    RUN_OR_FAIL;
    if (!stop) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        *stop = e->ens->stop_bits(run) != 0;
        return SPLASH_OK;
    });
}

int splash_ensemble_get_state(splash_ensemble* e, int run, double* U_out) {
    ENS_OR_FAIL;
    RUN_OR_FAIL;
    if (!U_out) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        e->ens->get_state(run, U_out);
        return SPLASH_OK;
    });
}

int splash_ensemble_set_state(splash_ensemble* e, int run, const double* U_in, double t, int stop_bits) {
    ENS_OR_FAIL;
    RUN_OR_FAIL;
    if (!U_in) return fail(SPLASH_EINVAL, "null state");
    return guarded([&] {
        e->ens->set_state(run, U_in, t, stop_bits);
        return SPLASH_OK;
    });
}

int splash_ensemble_diagnostics(splash_ensemble* e, int run, double* out) {
    ENS_OR_FAIL;
    RUN_OR_FAIL;
    if (!out) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        e->ens->diagnostics(run, out);
        return SPLASH_OK;
    });
}

int splash_ensemble_light_curves(splash_ensemble* e, int run, double* out) {
    ENS_OR_FAIL;
    RUN_OR_FAIL;
    if (!out) return fail(SPLASH_EINVAL, "null output");
    return guarded([&] {
        e->ens->light_curves(run, out);
        return SPLASH_OK;
    });
}

const char* splash_last_error_message(void) { return g_last_error.c_str(); }

}  // extern "C"
