/*
 * splash.h -- C ABI of the MI355X-native SPLASH MUSCL time step.
 *
 * The reference (TURBOLOSE/MUSCL-scheme-on-spherical-mesh-WIP) has no FFI:
 * its boundary is the C++ class API MUSCL_HLLE / MUSCL_HLLC / MUSCL_HLLCplus
 * (+ MUSCL_HLLE_p) with the methods do_step / time / get_stop_check.  Each
 * entry point below names the reference member it replaces; the C++ classes
 * with the reference signatures live in include/splash/MUSCL.hpp and call
 * only this ABI.
 *
 * Conventions: every function returns 0 (SPLASH_OK) or a negative error
 * code and never throws; the library owns device memory, the caller owns
 * every host array it passes; one context may be driven by one host thread
 * at a time.  State arrays are [n_faces][dim] row-major FP64 in the mesh's
 * face order with slot 4 holding the total energy (dim 5).
 */
#ifndef SPLASH_H
#define SPLASH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The library is built with hidden symbol visibility; these declarations
 * are its whole export list. */
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

#define SPLASH_OK 0
#define SPLASH_EINVAL (-1)   /* bad argument */
#define SPLASH_EMESH (-2)    /* mesh is not a closed manifold the stencil search accepts */
#define SPLASH_EDEVICE (-3)  /* HIP runtime error */
#define SPLASH_ENOMEM (-4)
#define SPLASH_EINTERNAL (-5)

/* solver ids: MUSCL_HLLE (HLLE.hpp:6), MUSCL_HLLE_p (HLLE_p.hpp:6),
 * MUSCL_HLLC (HLLC.hpp:6), MUSCL_HLLCplus (HLLCplus.hpp:6) */
#define SPLASH_HLLE 0
#define SPLASH_HLLE_P 1
#define SPLASH_HLLC 2
#define SPLASH_HLLCPLUS 3

/* precision modes: plain FP64, or FP32 reconstruction/flux with FP64 state */
#define SPLASH_FP64 0
#define SPLASH_MIXED 1

/* stop_check causes (bit set in splash_stop_bits) */
#define SPLASH_STOP_NAN_FLUX 1        /* MUSCL_base.hpp:264-268 */
#define SPLASH_STOP_NAN_WAVESPEED 2   /* adiabatic.hpp:671-676 */
#define SPLASH_STOP_WAVE_BRANCH 4     /* HLLC.hpp:114-119, HLLCplus.hpp:169-174, HLLE.hpp:51-55 */
#define SPLASH_STOP_NEGATIVE_DT 8     /* adiabatic.hpp:573-577 */

/* Flattened polygon mesh (what MUSCL_geometry.hpp:66-88 extracts from the
 * pmp::SurfaceMesh): vertex v at vertices[3v..3v+2]; face f has vertices
 * face_vertices[face_offsets[f] .. face_offsets[f+1]-1] in circulator order. */
typedef struct splash_mesh_desc {
    int32_t n_vertices;
    int32_t n_faces;
    const double *vertices;
    const int32_t *face_offsets;
    const int32_t *face_vertices;
} splash_mesh_desc;

/* Physics switches read by the adiabatic constructor from
 * input/parameters.json (adiabatic.hpp:59-68) plus accretion_on (ctor arg). */
typedef struct splash_params {
    int32_t accretion_on;
    int32_t friction_on;
    double acc_rate;
    double e_acc;
    double omega_acc_abs;
    double tilt_angle;
    double acc_width;
    double alpha;
    int32_t precision; /* SPLASH_FP64 or SPLASH_MIXED */
    int32_t setup_threads; /* host threads for the geometry precompute; 0 = every host thread
                            * (OpenMP's default team); the tables do not depend on it */
    int32_t patch_faces; /* SPLASH_MIXED: faces per stage-kernel patch, 1..128 (above 64 two
                          * waves update a patch's faces); 0 = 112 for 256-thread workgroups on
                          * meshes of mean valence at most 4.5, else 64 (per run in an ensemble;
                          * 64 where the larger patches' stencil face sets would overflow LDS);
                          * -1 = the fastest of a few (block_threads, patch_faces) candidates,
                          * measured at creation */
    int32_t no_graph;    /* 0: splash_do_steps replays a captured HIP graph (captured after the
                          * first call's steps); 1: every step is launched on its own */
    int32_t block_threads; /* SPLASH_MIXED: threads per stage-kernel workgroup, 128, 192 or 256
                            * (0 = 256); a patch's half-edges plus halo fit 2 * block_threads
                            * LDS slots; ignored when patch_faces = -1 (the tuner picks it) */
} splash_params;

typedef struct splash_mesh splash_mesh;   /* library-owned generated mesh */
typedef struct splash_ctx splash_ctx;     /* one resident run */

/* Fills *p with the reference's shipped input/parameters.json values,
 * accretion off, FP64. */
void splash_default_params(splash_params *p);

/* Mesh generators -- sph_gen.h uv_sphere :215, icosphere :267,
 * icosphere_hex :278, quad_sphere :286; numbering identical to pmp's. */
int splash_mesh_uv_sphere(int n_slices, int n_stacks, splash_mesh **out);
int splash_mesh_icosphere(int n_subdivisions, splash_mesh **out);
int splash_mesh_icosphere_hex(int n_subdivisions, splash_mesh **out);
int splash_mesh_quad_sphere(int n_subdivisions, splash_mesh **out);
/* View of a generated mesh; pointers stay valid until splash_mesh_free. */
int splash_mesh_view(const splash_mesh *m, splash_mesh_desc *out);
void splash_mesh_free(splash_mesh *m);

/* Replaces the constructors MUSCL_HLLE(mesh, U_in, dim, gam, threads)
 * (HLLE.hpp:13-14), MUSCL_HLLE_p / MUSCL_HLLC / MUSCL_HLLCplus(mesh, U_in,
 * dim, gam, omega_ns, accretion_on, threads) (HLLE_p.hpp:11-13,
 * HLLC.hpp:11-13, HLLCplus.hpp:11-13).  Mesh and U_in are copied.
 * params may be NULL (defaults).  Returns SPLASH_EINVAL when dim does not
 * match the solver (4 for HLLE, 5 otherwise); the reference only prints
 * "check dim" and sets stop_check (adiabatic.hpp:26-30). */
int splash_create(const splash_mesh_desc *mesh, const double *U_in, int dim, int solver, double gam,
                  double omega_ns, const splash_params *params, splash_ctx **out);
void splash_destroy(splash_ctx *ctx);

/* MUSCL_base::do_step(dt0) (MUSCL_base.hpp:97-225).  Enqueues the step on
 * the context's stream and returns; dt never leaves the GPU. */
int splash_do_step(splash_ctx *ctx, double dt0);
/* n consecutive do_step(dt0) calls.  The first call launches its steps one
 * by one and then captures 8 steps into a HIP graph (unless
 * params.no_graph); later calls replay it for each group of 8 steps. */
int splash_do_steps(splash_ctx *ctx, double dt0, int n);
/* Measurement extension: runs n do_step(dt0) calls back to back on the
 * context's stream, each kernel launched with start/stop HIP events bound
 * to its dispatch, one wait at the end; ms_out[4] receives the mean
 * duration in ms of [dt finalize, stage 1, stage 2, whole step]
 * (SPLASH_MIXED launches no dt kernel: its stage 1 forms dt, and
 * ms_out[0] = 0). */
int splash_profile_step(splash_ctx *ctx, double dt0, int n, double *ms_out);
/* Measurement extension (no counterpart in the reference): the HBM
 * bandwidth the library's own streaming copy kernels reach on the current
 * GPU, copying `bytes` (>= 1 MiB) `reps` times with 16-byte vector loads and
 * stores, plain and non-temporal; *gbs_out = the faster form's (read +
 * write) bytes per second / 1e9.  Allocates and frees its two buffers. */
int splash_copy_bandwidth(int64_t bytes, int reps, double *gbs_out);
/* Waits for every enqueued step. */
int splash_synchronize(splash_ctx *ctx);

/* MUSCL_base::time() (MUSCL_base.hpp:227-230); synchronises. */
int splash_time(splash_ctx *ctx, double *t);
/* The dt the last step used (the member dt after the clamps). */
int splash_last_dt(splash_ctx *ctx, double *dt);
/* MUSCL_base::get_stop_check() (MUSCL_base.hpp:232-235): *stop = 0/1. */
int splash_stop_check(splash_ctx *ctx, int *stop);
/* Which causes set stop_check, and the first offending half-edge/face
 * (code << 40 | index), or ~0 when none. */
int splash_stop_bits(splash_ctx *ctx, int *bits, uint64_t *first);
/* Extension (the reference keeps U private): copies the state out. */
int splash_get_state(splash_ctx *ctx, double *U_out);
/* Resume (no counterpart in the reference, which can only restart from a
 * text input file): replaces the state, the time and the stop_check bits
 * (splash_stop_bits' codes) of a context, as read back from a checkpoint;
 * U_in as for splash_create. */
int splash_set_state(splash_ctx *ctx, const double *U_in, double t, int stop_bits);
/* The do_step diagnostics sums (MUSCL_base.hpp:187-205):
 * out = {sum Sigma*A, sum Lx*A, sum Ly*A, sum Lz*A, sum E*A}. */
int splash_diagnostics(splash_ctx *ctx, double out[5]);
/* get_light_curves (adiabatic.hpp:228-333), adiabatic solvers only: the
 * pressure-weighted projected flux seen by observers along x, (x+z)/sqrt2,
 * z and -z, out = {flux_0, flux_45, flux_90, flux_180}.  It and
 * splash_diagnostics come from one device readout of the state (no state
 * transfer), made by the first of them after a step and reused by the
 * others until the next step or splash_set_state. */
int splash_light_curves(splash_ctx *ctx, double out[4]);
/* Allocation-free readout.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under their own HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Enqueues on
 * `stream` (a hipStream_t; NULL = the context's own stream, after its
 * steps) the readout of the context's current state and returns without
 * waiting: d_out, device memory of 9 doubles, receives {sum Sigma*A,
 * sum Lx*A, sum Ly*A, sum Lz*A, sum E*A, flux_0, flux_45, flux_90,
 * flux_180} (light curves 0 for dim 4), bit for bit the values of
 * splash_diagnostics and splash_light_curves.  It calls no allocator: its
 * per-workgroup partial sums live in a fixed array of 288 KiB that each GPU
 * holds from the first call until the process ends, which limits it to
 * SPLASH_READOUT_NOALLOC_MAX_FACES faces (SPLASH_EINVAL beyond).  Every call
 * on one GPU shares that array, so calls that overlap in time on one GPU
 * race; a caller passing another stream orders it after the context's steps
 * (splash_synchronize, or its own event). */
#define SPLASH_READOUT_NOALLOC_MAX_FACES 4194304
int splash_readout_noalloc(splash_ctx *ctx, double *d_out, void *stream);
/* write_t_curl's per-face values (adiabatic.hpp:135-178, isothermal.hpp:
 * 72-110), out[n_faces] in mesh face order: for face i,
 * sum_j ((ec_j x L+_j) / -Sigma_i) . (v_j - v_j1) / A_i, with L+ the face's
 * own reconstructed momenta of the last step's second stage (zero before a
 * step or after splash_set_state, as in a freshly built reference solver). */
int splash_curl(splash_ctx *ctx, double *out);
int splash_num_faces(const splash_ctx *ctx, int32_t *n_faces, int32_t *n_halfedges);
/* faces per stage-kernel patch in use (SPLASH_MIXED; the tuned value when
 * params.patch_faces = -1) */
int splash_patch_faces(const splash_ctx *ctx, int32_t *patch_faces);
/* the stage kernel's patch shape in use: faces per patch and threads per
 * workgroup (SPLASH_MIXED; both 0 for SPLASH_FP64; the tuned values when
 * params.patch_faces = -1) */
int splash_patch_shape(const splash_ctx *ctx, int32_t *patch_faces, int32_t *block_threads);
/* Face centres (MUSCL_geometry.hpp:176-187) and areas (:989-1030). */
int splash_face_geometry(const splash_ctx *ctx, double *centers_out, double *areas_out);

/* Host geometry tables for checks (MUSCL_geometry.hpp:43-575).  Arrays are
 * per face [nf] / per half-edge [nh] (h = face_offsets[f] + j); see
 * splash_geometry_view for the layout. */
typedef struct splash_geometry splash_geometry;
typedef struct splash_geometry_view {
    int32_t nf, nh;
    const double *fc;          /* [nf][3] */
    const double *normal;      /* [nf][3] */
    const double *area;        /* [nf] */
    const int32_t *nb_edge;    /* [nh] */
    const int32_t *twin;       /* [nh] */
    const int32_t *ff_plus;    /* [nh][2] */
    const int32_t *ff_minus;   /* [nh][2] */
    const double *beta_plus;   /* [nh][2] */
    const double *beta_minus;  /* [nh][2] */
    const double *h_plus;      /* [nh] */
    const double *h_minus;     /* [nh] */
    const double *bm_dist;     /* [nh] */
    const double *en;          /* [nh][3] */
    double h0;
    double round_diff;
} splash_geometry_view;
int splash_geometry_create(const splash_mesh_desc *mesh, int threads, splash_geometry **out);
int splash_geometry_get_view(const splash_geometry *g, splash_geometry_view *out);
void splash_geometry_free(splash_geometry *g);

/* ---- Batched ensemble -------------------------------------------------
 * No counterpart in the reference, which advances one run per solver
 * object.  For callers that advance many independent runs at once
 * (parameter sweeps, ensembles of initial states): n_runs runs of one
 * solver and dim, each on its own mesh with its own state, gam, omega_ns
 * and params[r], are advanced by the FP32-flux / FP64-state kernels with
 * every RK stage one launch for all runs.  Each run keeps its own dt clamp,
 * time and stop_check and evolves exactly as splash_create(...) with
 * params.precision = SPLASH_MIXED would on the same inputs.
 * params[r].precision is ignored; params[0].setup_threads sets the host
 * threads of the geometry precompute.  U_in[r] is [n_faces(r)][dim].  At
 * most 65536 runs; each run's dt accumulators take 128 KiB of device
 * memory beside its tables and state. */
typedef struct splash_ensemble splash_ensemble;
int splash_ensemble_create(int n_runs, const splash_mesh_desc *meshes, const double *const *U_in, int dim,
                           int solver, const double *gam, const double *omega_ns, const splash_params *params,
                           splash_ensemble **out);
void splash_ensemble_destroy(splash_ensemble *e);
/* nsteps do_step calls of every run, dt0[r] being run r's requested step;
 * returns without waiting.  As splash_do_steps, the first call captures a
 * HIP graph of 8 steps after launching its own (unless params[0].no_graph)
 * and later calls replay it. */
int splash_ensemble_step(splash_ensemble *e, const double *dt0, int nsteps);
/* Measurement extension, as splash_profile_step: n steps of every run with
 * dispatch-bound HIP events; ms_out[4] = mean [dt finalize (0: stage 1
 * forms dt), stage 1, stage 2, whole step] of the ensemble's launches. */
int splash_ensemble_profile_step(splash_ensemble *e, const double *dt0, int n, double *ms_out);
int splash_ensemble_synchronize(splash_ensemble *e);
/* n_runs, total faces, total half-edges, stage-kernel workgroups */
int splash_ensemble_info(const splash_ensemble *e, int32_t *n_runs, int64_t *total_faces, int64_t *total_halfedges,
                         int32_t *n_blocks);
int splash_ensemble_num_faces(const splash_ensemble *e, int run, int32_t *n_faces);
/* faces per patch in use (the tuned value when params[0].patch_faces = -1;
 * run 0's default cap when it is 0) */
int splash_ensemble_patch_faces(const splash_ensemble *e, int32_t *patch_faces);
/* faces per patch and threads per workgroup in use (params[0] sets both, or
 * the tuner when params[0].patch_faces = -1; run 0's default cap when it is 0) */
int splash_ensemble_patch_shape(const splash_ensemble *e, int32_t *patch_faces, int32_t *block_threads);
int splash_ensemble_time(splash_ensemble *e, int run, double *t);
int splash_ensemble_last_dt(splash_ensemble *e, int run, double *dt);
int splash_ensemble_stop_check(splash_ensemble *e, int run, int *stop);
int splash_ensemble_get_state(splash_ensemble *e, int run, double *U_out);
/* Resume of one run, as splash_set_state (no counterpart in the reference):
 * replaces run `run`'s state, time and stop_check bits; the other runs
 * continue exactly as they would have. */
int splash_ensemble_set_state(splash_ensemble *e, int run, const double *U_in, double t, int stop_bits);
/* Run `run`'s splash_diagnostics sums (MUSCL_base.hpp:187-205) and, for
 * adiabatic runs, splash_light_curves (adiabatic.hpp:228-333).  One device
 * readout covers every run and serves until the next splash_ensemble_step;
 * each run's values equal those of a single SPLASH_MIXED solver in the same
 * state with the same patch shape, bit for bit. */
int splash_ensemble_diagnostics(splash_ensemble *e, int run, double out[5]);
int splash_ensemble_light_curves(splash_ensemble *e, int run, double out[4]);

/* Host-side check of the internal GPU layout of a mesh that the
 * mixed-precision kernel uses (coordinate-bisection face order, here patches of at most
 * 64 faces whose half-edges plus halo fit 512 LDS slots, LDS face sets):
 * builds and validates it and returns {blocks (multiple of 8), non-empty
 * blocks, halo half-edges, max halo, max face set}. */
int splash_layout_stats(const splash_mesh_desc *mesh, int32_t out[5]);
/* The same for a given stage-kernel shape: block_threads 128, 192 or 256
 * (0 = 256), patch_faces 1..128 or 0 for the default cap of that mesh and
 * block size (splash_params.patch_faces). */
int splash_layout_stats_shape(const splash_mesh_desc *mesh, int32_t block_threads, int32_t patch_faces,
                              int32_t out[5]);

/* Human-readable text of the last error on this host thread. */
const char *splash_last_error_message(void);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif
#endif /* SPLASH_H */
