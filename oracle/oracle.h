/*
 * SPLASH MUSCL oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, scalar, FP64 restatement of the reference solver's hot path
 * (mesh generation, geometry precompute, MUSCL_base::do_step with the
 * adiabatic/isothermal physics and the HLLE/HLLE_p/HLLC/HLLC+ fluxes).
 * It exists to check the HIP product path.  Only tests/, the smoke check
 * and bench.py's cpu_baseline leg may load it; the product never links it.
 *
 * Every function follows the reference's operation order (no FMA
 * contraction: built with -ffp-contract=off) and cites the reference
 * file:line it restates.  Pinning of each part (closed forms, invariants,
 * brute force) lives in tests/test_oracle_*.py; parts with no such pin are
 * listed in the header of oracle_step.c as "parity unpinned".
 */
#ifndef SPLASH_ORACLE_H
#define SPLASH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ mesh */
/* Flattened polygon mesh in pmp numbering: vertex i at xyz[3i..3i+2];
 * face f has vertices fv[foff[f] .. foff[f+1]-1] in the order of pmp's
 * vertices-around-face circulator (starting at to_vertex(halfedge(f))). */
typedef struct or_mesh {
    int32_t nv, nf;
    double *xyz;
    int32_t *foff;
    int32_t *fv;
} or_mesh;

or_mesh *or_uv_sphere(int n_slices, int n_stacks); /* sph_gen.h:215-265 */
or_mesh *or_icosphere(int n);                      /* sph_gen.h:267-276 */
or_mesh *or_icosphere_hex(int n);                  /* sph_gen.h:278-284 */
or_mesh *or_quad_sphere(int n);                    /* sph_gen.h:286-295 */
void or_mesh_free(or_mesh *m);

/* -------------------------------------------------------------- geometry */
/* MUSCL_base_geometry tables (MUSCL_geometry.hpp:43-575), half-edge h =
 * foff[f] + j for face f, local edge j (vertices j, j+1). */
typedef struct or_geom {
    int32_t nf, nh, nv;
    const or_mesh *mesh;
    double *fc;        /* [nf][3] face centre (vertex mean, |fc|<1)       :176-187 */
    double *normal;    /* [nf][3] polygon normal                          :189-218 */
    double *area;      /* [nf]    spherical area                          :989-1030 */
    int32_t *nb_off;   /* [nf+1]  vertex-neighbour list offsets           :91-118 */
    int32_t *nb;       /*         sorted unique vertex neighbours         */
    int32_t *nb_edge;  /* [nh]    face across half-edge h                 :145-174 */
    int32_t *twin;     /* [nh]    half-edge of nb_edge[h] pointing back   MUSCL_base.hpp:383-384 */
    double *en;        /* [nh][3] outward edge normal                     :253-284 */
    double *bm_dist;   /* [nh]    great-circle centre->edge-mid distance  :286 */
    int32_t *ff_plus;  /* [nh][2] stencil faces for H+                    :288-363, :498-499 */
    int32_t *ff_minus; /* [nh][2] stencil faces for H-                    :407-408 */
    double *beta_plus; /* [nh][2]                                          :504-505 */
    double *beta_minus;/* [nh][2]                                          :484-485 */
    double *h_plus;    /* [nh]                                             :497 */
    double *h_minus;   /* [nh]                                             :406 */
    double h0;         /* min great-circle edge length        MUSCL_base.hpp:63-91 */
    double round_diff; /* arc/chord of face 0 edge (1,2)      MUSCL_base.hpp:242 */
} or_geom;

or_geom *or_geometry(const or_mesh *m);
void or_geom_free(or_geom *g);

/* ---------------------------------------------------------------- solver */
enum { OR_HLLE = 0, OR_HLLE_P = 1, OR_HLLC = 2, OR_HLLCPLUS = 3 };

typedef struct or_params {
    int32_t accretion_on;
    int32_t friction_on;
    double acc_rate, e_acc, omega_acc_abs, tilt_angle, acc_width, alpha;
} or_params;

typedef struct or_solver or_solver;

/* U_in: [nf][dim] row-major.  Returns NULL on bad arguments. */
or_solver *or_create(const or_mesh *m, const double *U_in, int dim, int solver,
                     double gam, double omega_ns, const or_params *p, int threads);
void or_do_step(or_solver *s, double dt0);                     /* MUSCL_base.hpp:97-225 */
double or_time(const or_solver *s);                            /* :227 */
int or_stop_check(const or_solver *s);                         /* :232 */
void or_get_state(const or_solver *s, double *U_out);          /* [nf][dim], slot 4 = E */
void or_curl(const or_solver *s, double *out);                 /* adiabatic.hpp:135-178, [nf] */
double or_last_dt(const or_solver *s);
void or_diagnostics(const or_solver *s, double out[5]);        /* :187-205 */
double or_area_coeff(const or_solver *s);                      /* adiabatic.hpp:70-84 */
void or_destroy(or_solver *s);

/* Single-function entry points for closed-form tests. */
/* source (adiabatic.hpp:396-541) of the p-form state u at unit face centre
 * fcn, with the solver's parameters, area_coeff and total_mass. */
void or_source(const or_solver *s, const double fcn[3], const double u[5], double res[5]);
/* find_U_edges (MUSCL_base.hpp:361-434) of the current state with step dt:
 * Uplus_out [nh][dim] (slot 4 = E_fc for dim 5); the state is unchanged. */
void or_reconstruct(or_solver *s, double dt, double *Uplus_out);
void or_limiter_adiabatic(const double r[5], double nu, double eta_minus, double eta_plus,
                          int valence, double lim_out[5]);     /* adiabatic.hpp:685-828 */
/* Riemann flux of one half-edge given the geometry of that half-edge:
 * ec = unit edge midpoint, ecn = |ec| as recomputed, en = edge normal. */
int or_flux_star(int solver, double gam, const double *uL, const double *uR,
                 const double ec[3], double ecn, const double en[3], double F_out[5]);

#ifdef __cplusplus
}
#endif
#endif
