/*
 * Oracle (test infrastructure): restatement of the reference time step.
 *
 *   MUSCL_base::do_step            src/MUSCL_base/MUSCL_base.hpp:97-225
 *   find_v_max / pressure_fc       MUSCL_base.hpp:311-343, :517-539
 *   find_U_edges (+U_H, p+-, E_fc) MUSCL_base.hpp:361-515, :541-562
 *   find_flux_var / res2d          MUSCL_base.hpp:345-359, :238-276
 *   adiabatic physics              src/physics/adiabatic.hpp:18-91, :352-828
 *   isothermal physics             src/physics/isothermal.hpp:141-320
 *   flux_star                      src/Riemann_solvers/{HLLE,HLLE_p,HLLC,HLLCplus}.hpp
 *
 * Operation order follows the reference expression by expression; the file
 * is compiled with -ffp-contract=off (the reference's GCC build emits no FMA
 * on baseline x86-64).  Dead computations of the reference (theta/phi
 * values that feed nothing, diagnostic prints) are omitted.
 *
 * Semantics chosen where the reference has undefined behaviour (its dim=4
 * isothermal path), stated once here and in the build notes:
 *   - dim=4: no E->p slot swap (the reference writes U[i][4] out of
 *     bounds); find_v_max uses the isothermal sound speed a=1 for c_s.
 *   - isothermal limiter: etha_minus (read uninitialised by the reference,
 *     isothermal.hpp:272, :307) := H_minus/BM_dist; etha_plus keeps the
 *     reference's last assignment, H_minus/BM_dist (:301-302).
 *   - the area_coeff loop (adiabatic.hpp:75, uninitialised start index)
 *     starts at face 0.
 *
 * Pins (tests/test_oracle_*.py): closed-form flux, wave-speed and limiter
 * values; conservation and rest-state invariants of do_step; for source(),
 * the accretion profile's half-maximum width and centre, the t = 0 mass
 * balance, the continuum limit of area_coeff, the accreted angular momentum
 * and energy, the cooling law on both branches, its continuity at the
 * switch and its clamps, and the friction drag law; for find_U_edges,
 * second-order convergence to smooth fields and the Courant-number
 * composition of the limiter at a large step.  Parity unpinned: only the
 * last bits of glibc libm's pow/tanh/asin/acos/exp/atan2 (no reference
 * build exists in this image to produce golden vectors -- pmp's mat_vec.h:17
 * needs the absent Eigen headers).
 */
#include "oracle.h"
#include "oracle_vec.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define DENSITY_FLOOR 1e-8
#define PRESSURE_FLOOR 1e-16
#define GM 0.217909

struct or_solver {
    or_mesh mesh;   /* private copy (the reference takes the mesh by value) */
    or_geom *g;
    int nf, nh, dim, solver, threads;
    double gam, omega_ns;
    or_params p;
    double *U, *U_temp, *Uplus, *Uminus, *flux, *src;
    double dt, t, max_vel, rho_full, E_full;
    size_t steps;
    int stop;
    double area_coeff, total_mass;
};

static int is_adiabatic(const or_solver *s) { return s->solver != OR_HLLE; }

/* unit edge midpoint as the flux/limiter routines build it:
 * edge_center = (v_j + v_j1)/2.; edge_center /= edge_center.norm(); */
static v3 edge_center(const or_solver *s, int f, int j)
{
    const or_mesh *m = &s->mesh;
    int n = m->foff[f + 1] - m->foff[f];
    int j1 = (j + 1 == n) ? 0 : j + 1;
    v3 a = v3ld(m->xyz + 3 * m->fv[m->foff[f] + j]);
    v3 b = v3ld(m->xyz + 3 * m->fv[m->foff[f] + j1]);
    v3 ec = v3divs(v3add(a, b), 2.);
    return v3divs(ec, v3norm(ec));
}

static v3 face_center_unit(const or_solver *s, int f)
{
    v3 fc = v3ld(s->g->fc + 3 * f);
    return v3divs(fc, v3norm(fc));
}

/* adiabatic.hpp:582-592 */
static double pressure(double gam, const double *u, v3 vel)
{
    double n = v3norm(vel);
    return smax(PRESSURE_FLOOR, (u[4] - u[0] * (n * n) / 2) * (gam - 1));
}

/* MUSCL_base.hpp:517-539 (u[4] is energy) */
static double pressure_fc(const or_solver *s, const double *u, int f)
{
    v3 l = v3m(u[1], u[2], u[3]);
    v3 r = face_center_unit(s, f);
    v3 vel = v3divs(v3cross(r, l), -u[0]);
    double n = v3norm(vel);
    return smax(PRESSURE_FLOOR, (u[4] - u[0] * (n * n) / 2) * (s->gam - 1));
}

/* ----------------------------------------------------------- fluxes */

/* adiabatic.hpp:352-392 */
static void flux_adiabatic(double gam, const double *u, v3 ec, double ecn, v3 en, double *res)
{
    v3 l = v3m(u[1], u[2], u[3]);
    v3 vel = v3divs(v3cross(ec, l), (-u[0]) * ecn);
    double PI = pressure(gam, u, vel);
    double ndv = v3dot(en, vel);
    v3 nxR = v3cross(en, v3divs(ec, ecn));
    res[0] = u[0] * v3dot(vel, en);
    res[1] = (u[1] * ndv - nxR.x[0] * PI);
    res[2] = (u[2] * ndv - nxR.x[1] * PI);
    res[3] = (u[3] * ndv - nxR.x[2] * PI);
    res[4] = (u[4] + PI) * v3dot(vel, en);
}

/* isothermal.hpp:141-195, a = 1 */
static void flux_isothermal(const double *u, v3 ec, v3 en, double *res)
{
    const double a = 1;
    double PI = a * a * u[0];
    v3 l = v3m(u[1], u[2], u[3]);
    v3 vel = v3divs(v3cross(ec, l), -u[0]);
    double ndv = v3dot(en, vel);
    v3 nxR = v3cross(en, ec);
    res[0] = u[0] * v3dot(vel, en);
    res[1] = u[1] * ndv - nxR.x[0] * PI;
    res[2] = u[2] * ndv - nxR.x[1] * PI;
    res[3] = u[3] * ndv - nxR.x[2] * PI;
}

/* adiabatic.hpp:594-683 (PVRS wave speeds); returns 1 on NaN */
static int char_vel_adiabatic(double gam, const double *uL, const double *uR, v3 ec, double ecn,
                              v3 en, double *SL, double *SR)
{
    v3 vel_l = v3divs(v3cross(ec, v3m(uL[1], uL[2], uL[3])), (-uL[0]) * ecn);
    v3 vel_r = v3divs(v3cross(ec, v3m(uR[1], uR[2], uR[3])), (-uR[0]) * ecn);
    double p_L = pressure(gam, uL, vel_l);
    double p_R = pressure(gam, uR, vel_r);
    double a_L = sqrt(gam * p_L / uL[0]);
    double a_R = sqrt(gam * p_R / uR[0]);
    double p_pvrs = 1 / 2. * (p_L + p_R) -
                    1 / 8. * (v3dot(en, vel_r) - v3dot(en, vel_l)) * (uL[0] + uR[0]) * (a_L + a_R);
    double p_star = smax(0., p_pvrs);
    double q_R, q_L;
    if (p_star <= p_R) {
        q_R = 1;
    } else {
        q_R = sqrt(1 + (gam + 1) / (2 * gam) * (p_star / p_R - 1));
        if (isnan(q_R) || isinf(q_R)) q_R = 1;
    }
    if (p_star <= p_L) {
        q_L = 1;
    } else {
        q_L = sqrt(1 + (gam + 1) / (2 * gam) * (p_star / p_L - 1));
        if (isnan(q_L) || isinf(q_L)) q_L = 1;
    }
    *SL = v3dot(vel_l, en) - a_L * q_L;
    *SR = v3dot(vel_r, en) + a_R * q_R;
    return isnan(*SL) || isnan(*SR);
}

/* isothermal.hpp:211-260 */
static void char_vel_isothermal(const double *uL, const double *uR, v3 ec, double ecn, v3 en,
                                double *SL, double *SR)
{
    const double a = 1;
    v3 vel_l = v3divs(v3cross(ec, v3m(uL[1], uL[2], uL[3])), (-uL[0]) * ecn);
    v3 vel_r = v3divs(v3cross(ec, v3m(uR[1], uR[2], uR[3])), (-uR[0]) * ecn);
    double dl = v3dot(vel_l, en), dr = v3dot(vel_r, en);
    *SL = smin(dl, dr) - smax(a, a);
    *SR = smax(dl, dr) + smax(a, a);
}

/* HLLE.hpp:17-59 / HLLE_p.hpp:19-57; returns 1 on the fallthrough branch */
static int hlle_combine(int dim, double S_L, double S_R, const double *ul, const double *ur,
                        const double *FL, const double *FR, double *F)
{
    if (S_L >= 0) {
        for (int i = 0; i < dim; i++) F[i] = FL[i];
    } else if (S_L < 0 && S_R > 0) {
        for (int i = 0; i < dim; i++)
            F[i] = (S_R * FL[i] - S_L * FR[i] + S_R * S_L * (ur[i] - ul[i])) / (S_R - S_L);
    } else if (S_R <= 0) {
        for (int i = 0; i < dim; i++) F[i] = FR[i];
    } else {
        for (int i = 0; i < dim; i++) F[i] = FR[i];
        return 1;
    }
    return 0;
}

/* HLLC.hpp:18-122 and HLLCplus.hpp:17-177 (plus != 0) */
static int hllc_flux(double gam, int plus, const double *u_L, const double *u_R, v3 ec, double ecn,
                     v3 en, double *F)
{
    int err = 0;
    double F_L[5], F_R[5], F_L_star[5], F_R_star[5], D[5];
    v3 vel_L = v3divs(v3cross(ec, v3m(u_L[1], u_L[2], u_L[3])), (-u_L[0]) * ecn);
    v3 vel_R = v3divs(v3cross(ec, v3m(u_R[1], u_R[2], u_R[3])), (-u_R[0]) * ecn);
    double S_L, S_R;
    err |= char_vel_adiabatic(gam, u_L, u_R, ec, ecn, en, &S_L, &S_R);
    v3 nxR = v3cross(en, v3divs(ec, ecn));
    D[0] = 0;
    D[1] = -nxR.x[0];
    D[2] = -nxR.x[1];
    D[3] = -nxR.x[2];
    flux_adiabatic(gam, u_L, ec, ecn, en, F_L);
    flux_adiabatic(gam, u_R, ec, ecn, en, F_R);
    double rho_R = u_R[0], rho_L = u_L[0];
    double p_L = pressure(gam, u_L, vel_L);
    double p_R = pressure(gam, u_R, vel_R);
    double V_L = v3dot(en, vel_L);
    double V_R = v3dot(en, vel_R);
    double S_star = (p_R - p_L + rho_L * V_L * (S_L - V_L) - rho_R * V_R * (S_R - V_R)) /
                    (rho_L * (S_L - V_L) - rho_R * (S_R - V_R));
    if (isnan(S_star) || isinf(S_star)) S_star = (V_L + V_R) / 2;
    D[4] = S_star;
    double P_LR = (p_L + p_R + u_L[0] * (S_L - v3dot(en, vel_L)) * (S_star - v3dot(en, vel_L)) +
                   u_R[0] * (S_R - v3dot(en, vel_R)) * (S_star - v3dot(en, vel_R))) / 2;
    double fix_L[5] = {0, 0, 0, 0, 0}, fix_R[5] = {0, 0, 0, 0, 0};
    if (plus) {
        double phi_L = u_L[0] * (S_L - V_L);
        double phi_R = u_R[0] * (S_R - V_R);
        double a_L = sqrt(gam * p_L / u_L[0]);
        double a_R = sqrt(gam * p_R / u_R[0]);
        double M = smin(1., smax(v3norm(vel_L) / a_L, v3norm(vel_R) / a_R));
        int lambd_L = 0, lambd_R = 0;
        if ((V_L - a_L) > 0 && (V_R - a_R) < 0) lambd_L = 1;
        if ((V_L + a_L) > 0 && (V_R + a_R) < 0) lambd_L = 1;
        if ((V_R - a_R) > 0 && (V_L - a_L) < 0) lambd_R = 1;
        if ((V_R + a_R) > 0 && (V_L + a_L) < 0) lambd_R = 1;
        double f_star = 1;
        if (lambd_L == 0 && lambd_R == 0) f_star = M * sqrt(4 + (1 - M * M) * (1 - M * M)) / ((1 + M * M));
        double h = smin(p_L / p_R, p_R / p_L);
        double g = 1 - pow(h, M);
        v3 vfL, vfR;
        for (int c = 0; c < 3; ++c) {
            vfL.x[c] = (f_star - 1) * (V_R - V_L) * en.x[c] +
                       S_L / (S_L - S_star) * g * (vel_R.x[c] - vel_L.x[c] - (V_R - V_L) * en.x[c]);
            vfR.x[c] = (f_star - 1) * (V_R - V_L) * en.x[c] +
                       S_R / (S_R - S_star) * g * (vel_R.x[c] - vel_L.x[c] - (V_R - V_L) * en.x[c]);
        }
        v3 amL = v3cross(ec, vfL);
        v3 amR = v3cross(ec, vfR);
        for (int c = 0; c < 3; ++c) {
            fix_L[1 + c] = amL.x[c] * phi_L * phi_R / (phi_R - phi_L);
            fix_R[1 + c] = amR.x[c] * phi_L * phi_R / (phi_R - phi_L);
        }
        fix_L[4] = (f_star - 1) * (V_R - V_L) * S_star * phi_L * phi_R / (phi_R - phi_L);
        fix_R[4] = (f_star - 1) * (V_R - V_L) * S_star * phi_L * phi_R / (phi_R - phi_L);
    }
    for (int i = 0; i < 5; i++) {
        F_L_star[i] = (S_star * (S_L * u_L[i] - F_L[i]) + S_L * P_LR * D[i]) / (S_L - S_star);
        F_R_star[i] = (S_star * (S_R * u_R[i] - F_R[i]) + S_R * P_LR * D[i]) / (S_R - S_star);
        if (plus) {
            F_L_star[i] = F_L_star[i] + fix_L[i];
            F_R_star[i] = F_R_star[i] + fix_R[i];
        }
    }
    const double *sel;
    if (S_L >= 0) sel = F_L;
    else if (S_L < 0 && S_star >= 0) sel = F_L_star;
    else if (S_star < 0 && S_R >= 0) sel = F_R_star;
    else if (S_R < 0) sel = F_R;
    else { sel = F_R; err |= 1; }
    for (int i = 0; i < 5; i++) F[i] = sel[i];
    return err;
}

static int flux_star_geom(int solver, double gam, const double *uL, const double *uR, v3 ec,
                          double ecn, v3 en, double *F)
{
    double FL[5], FR[5], SL, SR;
    int err = 0;
    switch (solver) {
    case OR_HLLE:
        char_vel_isothermal(uL, uR, ec, ecn, en, &SL, &SR);
        flux_isothermal(uL, ec, en, FL);
        flux_isothermal(uR, ec, en, FR);
        return hlle_combine(4, SL, SR, uL, uR, FL, FR, F);
    case OR_HLLE_P:
        err |= char_vel_adiabatic(gam, uL, uR, ec, ecn, en, &SL, &SR);
        flux_adiabatic(gam, uL, ec, ecn, en, FL);
        flux_adiabatic(gam, uR, ec, ecn, en, FR);
        err |= hlle_combine(5, SL, SR, uL, uR, FL, FR, F);
        return err;
    case OR_HLLC: return hllc_flux(gam, 0, uL, uR, ec, ecn, en, F);
    default: return hllc_flux(gam, 1, uL, uR, ec, ecn, en, F);
    }
}

int or_flux_star(int solver, double gam, const double *uL, const double *uR, const double ec[3],
                 double ecn, const double en[3], double F_out[5])
{
    return flux_star_geom(solver, gam, uL, uR, v3ld(ec), ecn, v3ld(en), F_out);
}

/* ---------------------------------------------------------- limiters */

/* adiabatic.hpp:693-702 switch function h(r) */
static double h_switch(double r)
{
    double a = 4, b = 2, c = 0.1, d = 10, e = 3, f = 6;
    double res = 0;
    if (r < 1 && r > 0) res = (1 - tanh(a * pow(r, b) * pow(1 - r, c)));
    if (r >= 1) res = pow(tanh(d * pow(r - 1, e)), f);
    return res;
}

/* adiabatic.hpp:685-828 with the face-local Courant number nu and
 * eta+- = H+-/BM_dist already evaluated */
void or_limiter_adiabatic(const double r[5], double nu, double eta_minus, double eta_plus, int valence,
                          double lim[5])
{
    for (int i = 0; i < 5; i++) {
        double sb = smax(0., smax(smin(1., eta_minus * r[i] / ((double)(2 * (size_t)valence) * nu)),
                                  smin(r[i], eta_plus)));
        if (isnan(r[i])) sb = 0;
        double to = smax(0., smin(sb, 1 + (1 + nu) / 3 * (r[i] - 1)));
        if (isnan(r[i])) to = 0;
        double hr = h_switch(r[i]);
        lim[i] = ((1 - hr) * to + hr * sb);
        if (isnan(r[i])) lim[i] = 0;
    }
}

static void limiter(const or_solver *s, const double *W, int f, int j, int h, const double *r, double *lim)
{
    const or_mesh *m = &s->mesh;
    int n = m->foff[f + 1] - m->foff[f];
    int j1 = (j + 1 == n) ? 0 : j + 1;
    v3 ec = edge_center(s, f, j);
    v3 en = v3ld(s->g->en + 3 * h);
    const double *Wf = W + (size_t)f * s->dim;
    v3 l = v3m(Wf[1], Wf[2], Wf[3]);
    double arcA = gc_distance(v3ld(m->xyz + 3 * m->fv[m->foff[f] + j]),
                              v3ld(m->xyz + 3 * m->fv[m->foff[f] + j1])) / s->g->area[f];
    if (is_adiabatic(s)) {
        /* adiabatic.hpp:741-758 and :792-813 (rho_an = p_an = 0) */
        v3 vel = v3divs(v3cross(ec, l), (-Wf[0] - 0.0) * v3norm(ec));
        double p = Wf[4] + 0.0;
        double c = sqrt(s->gam * p / (Wf[0] + 0.0));
        double nu = (c + v3dot(vel, en)) * s->dt * arcA;
        double eta_plus = s->g->h_plus[h] / s->g->bm_dist[h];
        double eta_minus = s->g->h_minus[h] / s->g->bm_dist[h];
        or_limiter_adiabatic(r, nu, eta_minus, eta_plus, n, lim);
    } else {
        /* isothermal.hpp:269-320 */
        const double a = 1;
        v3 vel = v3divs(v3cross(ec, l), -Wf[0]);
        double c = a;
        double nu = (c + v3dot(vel, en)) * s->dt * arcA;
        double eta_plus = s->g->h_minus[h] / s->g->bm_dist[h];
        double eta_minus = s->g->h_minus[h] / s->g->bm_dist[h];
        for (int i = 0; i < s->dim; i++) {
            lim[i] = smax(0., smax(smin(1., eta_minus * r[i] * 2 / ((double)(2 * (size_t)n) * nu)),
                                   smin(r[i], eta_plus)));
            if (isnan(r[i])) lim[i] = 0;
        }
    }
}

/* ---------------------------------------------------------- sources */

/* adiabatic.hpp:396-541 (u[4] is pressure) at the unit face centre fcn */
static void source_at(const or_solver *s, const double *u, v3 fcn, double *res)
{
    const int dim = s->dim;
    for (int i = 0; i < dim; i++) res[i] = 0;
    if (!is_adiabatic(s)) return; /* isothermal.hpp:199-208 */
    const double gam = s->gam;
    v3 l = v3m(u[1], u[2], u[3]);
    v3 vel = v3divs(v3cross(fcn, l), -u[0]);
    double mu = M_PI / 2;
    double sigma = s->p.acc_width / 2.355;
    double tilt = s->p.tilt_angle;
    double theta_acc = acos(fcn.x[1] * sin(tilt) + fcn.x[2] * cos(tilt));
    if (s->p.accretion_on) {
        res[0] = 1 / s->area_coeff * s->p.acc_rate / sqrt(2 * M_PI * sigma * sigma) *
                 exp(-(theta_acc - mu) * (theta_acc - mu) / (2 * sigma * sigma));
        v3 omega_acc = v3m(0, sin(tilt) * s->p.omega_acc_abs, cos(tilt) * s->p.omega_acc_abs);
        v3 omxr = v3cross(omega_acc, fcn);
        v3 rxv = v3cross(fcn, omxr);
        res[1] = res[0] * rxv.x[0];
        res[2] = res[0] * rxv.x[1];
        res[3] = res[0] * rxv.x[2];
        double dv = v3norm(v3sub(omxr, vel));
        res[4] = res[0] * (s->p.e_acc + (dv * dv) / 2.);
        double fall_eff = 1;
        rxv = v3cross(fcn, vel);
        double vn = v3norm(vel);
        if (!s->p.friction_on) {
            double dmdt = -(fall_eff * s->p.acc_rate * 4 * M_PI) / s->total_mass * u[0];
            res[0] += dmdt;
            res[1] += dmdt * rxv.x[0];
            res[2] += dmdt * rxv.x[1];
            res[3] += dmdt * rxv.x[2];
            res[4] += dmdt * (gam / (gam - 1) * u[4] / u[0] + (vn * vn) / 2.);
        }
        double g_eff = GM - vn * vn;
        double c_sigma = 4.85e36, k_m = 1.6e-13, kappa = 3.4e6;
        double C = 12. / 5 * k_m * u[0] / u[4] * pow(3. / 4 * c_sigma * g_eff * u[0], 1. / 4);
        double beta_switch = 0.736194670678821;
        double C_switch = -1 - 1 / (beta_switch - 1);
        double beta;
        if (C <= C_switch) beta = 1 - 1 / (1 + C);
        else beta = 1 - pow(2 / C, 4);
        double beta_ceil = 1 - 1e-8, beta_floor = 1e-8;
        if (beta < beta_floor || isnan(beta)) beta = beta_floor;
        if (beta > beta_ceil) beta = beta_ceil;
        res[4] -= g_eff / kappa * (1 - beta);
        if (s->p.friction_on) {
            double gam3d = 1 / (2 - gam);
            double rho = gam3d / (2 * gam3d - 1) * g_eff * u[0] * u[0] / u[4];
            v3 omega0 = v3m(0, 0, s->omega_ns);
            v3 vel0 = v3cross(omega0, fcn);
            double dn = v3norm(v3sub(vel, vel0));
            v3 wr = v3muls(v3sub(vel, vel0), -s->p.alpha * rho * dn);
            v3 l_fr = v3cross(fcn, wr);
            res[1] += l_fr.x[0];
            res[2] += l_fr.x[1];
            res[3] += l_fr.x[2];
            res[4] += v3dot(wr, vel0);
            res[0] -= s->p.alpha * rho * dn;
            res[1] -= s->p.alpha * rho * dn / u[0] * u[1];
            res[2] -= s->p.alpha * rho * dn / u[0] * u[2];
            res[3] -= s->p.alpha * rho * dn / u[0] * u[3];
            res[4] -= s->p.alpha * rho * dn * (vn * vn / 2. + u[4] * gam / ((gam - 1) * u[0]));
        }
    }
}

static void source(const or_solver *s, const double *u, int f, double *res)
{
    source_at(s, u, face_center_unit(s, f), res);
}

void or_source(const or_solver *s, const double fcn[3], const double u[5], double res[5])
{
    source_at(s, u, v3ld(fcn), res);
}

/* adiabatic.hpp:543-580 (U[i][4] is energy here); isothermal: 1e20 */
static double extra_dt_constr(or_solver *s)
{
    if (!is_adiabatic(s)) return 1e20;
    double dt_new = 1e20;
    const double gam = s->gam;
    for (int i = 0; i < s->nf; i++) {
        const double *u = s->U + (size_t)i * s->dim;
        v3 fcn = face_center_unit(s, i);
        v3 vel = v3divs(v3cross(fcn, v3m(u[1], u[2], u[3])), -u[0]);
        v3 omega0 = v3m(0, 0, s->omega_ns);
        v3 vel0 = v3cross(omega0, fcn);
        double press = pressure(gam, u, vel);
        double gam3d = 1 / (2 - gam);
        double vn = v3norm(vel);
        double g_eff = GM - vn * vn;
        double rho = gam3d / (2 * gam3d - 1) * g_eff * u[0] * u[0] / press;
        double cand = 0.01 * u[0] / (s->p.alpha * rho * v3norm(v3sub(vel, vel0)));
        if (dt_new > cand) dt_new = cand;
    }
    if (dt_new < 0) {
        s->stop = 1;
        dt_new = 1e20;
    }
    return dt_new;
}

/* MUSCL_base.hpp:311-343 */
static void find_v_max(or_solver *s)
{
    double mx = 1e-8;
    for (int i = 0; i < s->nf; i++) {
        const double *u = s->U + (size_t)i * s->dim;
        v3 vel = v3divs(v3cross(face_center_unit(s, i), v3m(u[1], u[2], u[3])), u[0]);
        double c_s;
        if (is_adiabatic(s)) {
            double p = pressure_fc(s, u, i);
            c_s = sqrt(s->gam * p / u[0]);
        } else {
            c_s = 1; /* isothermal sound speed a */
        }
        if (v3norm(vel) > mx) mx = v3norm(vel);
        if (c_s > mx) mx = c_s;
    }
    s->max_vel = mx;
}

/* MUSCL_base.hpp:541-562 (u[4] is pressure) */
static double E_fc(const or_solver *s, const double *u, int f, int j)
{
    const or_mesh *m = &s->mesh;
    int n = m->foff[f + 1] - m->foff[f];
    int j1 = (j == n - 1) ? 0 : j + 1;
    v3 r = v3add(v3ld(m->xyz + 3 * m->fv[m->foff[f] + j]), v3ld(m->xyz + 3 * m->fv[m->foff[f] + j1]));
    r = v3divs(r, v3norm(r));
    v3 vel = v3divs(v3cross(r, v3m(u[1], u[2], u[3])), (-u[0]));
    double vn = v3norm(vel);
    return 1 / (s->gam - 1) * u[4] + u[0] * (vn * vn) / 2;
}

/* MUSCL_base.hpp:361-434 */
static void find_U_edges(or_solver *s)
{
    const int dim = s->dim, nf = s->nf;
    const or_mesh *m = &s->mesh;
    const or_geom *g = s->g;
    if (dim == 5) {
        for (int i = 0; i < nf; i++) {
            double *u = s->U + (size_t)i * dim;
            u[4] = pressure_fc(s, u, i);
            u[0] -= 0.0;
            u[4] -= 0.0;
        }
    }
    double *U = s->U;
#pragma omp parallel for num_threads(s->threads) schedule(static)
    for (int i = 0; i < nf; ++i) {
        int n = m->foff[i + 1] - m->foff[i];
        for (int j = 0; j < n; ++j) {
            int h = m->foff[i] + j;
            int tw = g->twin[h];
            double pp[5], pm[5], r[5], lim[5];
            const double *Ui = U + (size_t)i * dim;
            const double *Up0 = U + (size_t)g->ff_plus[2 * h] * dim;
            const double *Up1 = U + (size_t)g->ff_plus[2 * h + 1] * dim;
            const double *Um0 = U + (size_t)g->ff_minus[2 * h] * dim;
            const double *Um1 = U + (size_t)g->ff_minus[2 * h + 1] * dim;
            for (int k = 0; k < dim; k++) {
                double up = g->beta_plus[2 * h] * Up0[k] + g->beta_plus[2 * h + 1] * Up1[k];
                pp[k] = (up - Ui[k]) / g->h_plus[h];
            }
            for (int k = 0; k < dim; k++) {
                double um = g->beta_minus[2 * h] * Um0[k] + g->beta_minus[2 * h + 1] * Um1[k];
                pm[k] = (-um + Ui[k]) / g->h_minus[h];
            }
            for (int k = 0; k < dim; k++) r[k] = pm[k] / pp[k];
            limiter(s, U, i, j, h, r, lim);
            double *UP = s->Uplus + (size_t)h * dim;
            double *UM = s->Uminus + (size_t)tw * dim;
            for (int k = 0; k < dim; k++) {
                if (k == 0) {
                    UP[k] = Ui[k] + 0.0 + pp[k] * lim[k] * g->bm_dist[h];
                } else if (k == 4) {
                    UP[k] = Ui[k] + 0.0 + pp[k] * lim[k] * g->bm_dist[h];
                    UP[k] = E_fc(s, UP, i, j);
                } else {
                    UP[k] = Ui[k] + pp[k] * lim[k] * g->bm_dist[h];
                }
                UM[k] = UP[k];
            }
        }
    }
}

/* MUSCL_base.hpp:345-359 */
static void find_flux_var(or_solver *s)
{
    const int dim = s->dim, nf = s->nf;
    const or_mesh *m = &s->mesh;
    int stop = 0;
#pragma omp parallel for num_threads(s->threads) schedule(static) reduction(| : stop)
    for (int i = 0; i < nf; i++) {
        int n = m->foff[i + 1] - m->foff[i];
        for (int j = 0; j < n; j++) {
            int h = m->foff[i] + j;
            v3 ec = edge_center(s, i, j);
            stop |= flux_star_geom(s->solver, s->gam, s->Uplus + (size_t)h * dim,
                                   s->Uminus + (size_t)h * dim, ec, v3norm(ec), v3ld(s->g->en + 3 * h),
                                   s->flux + (size_t)h * dim);
        }
        source(s, s->U + (size_t)i * dim, i, s->src + (size_t)i * dim);
    }
    if (stop) s->stop = 1;
}

/* MUSCL_base.hpp:238-276 */
static void res2d(or_solver *s, double dt_here)
{
    const int dim = s->dim;
    const or_mesh *m = &s->mesh;
    double round_diff = s->g->round_diff;
    for (int i = 0; i < s->nf; i++) {
        double *u = s->U + (size_t)i * dim;
        for (int k = 0; k < dim; k++) u[k] = dt_here * s->src[(size_t)i * dim + k];
        int n = m->foff[i + 1] - m->foff[i];
        for (int j = 0; j < n; j++) {
            int j1 = j + 1;
            if (j == n - 1) j1 = 0;
            int h = m->foff[i] + j;
            double chord = v3norm(v3sub(v3ld(m->xyz + 3 * m->fv[m->foff[i] + j]),
                                        v3ld(m->xyz + 3 * m->fv[m->foff[i] + j1])));
            for (int k = 0; k < dim; k++) {
                double F = s->flux[(size_t)h * dim + k];
                if (isnan(F)) s->stop = 1;
                u[k] -= round_diff * dt_here * chord / s->g->area[i] * F;
            }
        }
    }
}

static void combine(or_solver *s)
{
    for (int i = 0; i < s->nf; i++) {
        double *u = s->U + (size_t)i * s->dim;
        const double *u0 = s->U_temp + (size_t)i * s->dim;
        for (int k = 0; k < s->dim; k++) u[k] += u0[k];
        if (u[0] < DENSITY_FLOOR) u[0] = DENSITY_FLOOR;
    }
}

void or_do_step(or_solver *s, double dt0)
{
    const size_t n = (size_t)s->nf * s->dim;
    s->dt = dt0;
    memcpy(s->U_temp, s->U, sizeof(double) * n);
    find_v_max(s);
    if (s->dt > s->g->h0 * 0.1 / s->max_vel) s->dt = s->g->h0 * 0.1 / s->max_vel;
    double extra_dt = extra_dt_constr(s);
    if (s->dt > extra_dt) s->dt = extra_dt;

    find_U_edges(s);
    find_flux_var(s);
    res2d(s, s->dt / 2);
    combine(s);

    find_U_edges(s);
    find_flux_var(s);
    res2d(s, s->dt);
    combine(s);

    double d[5];
    or_diagnostics(s, d);
    if (s->steps == 0) {
        s->rho_full = d[0];
        s->E_full = d[4];
    }
    s->t += s->dt;
    s->steps++;
}

void or_diagnostics(const or_solver *s, double out[5])
{
    double temp = 0, l1 = 0, l2 = 0, l3 = 0, temp_E = 0;
    for (int i = 0; i < s->nf; i++) {
        const double *u = s->U + (size_t)i * s->dim;
        double A = s->g->area[i];
        temp += u[0] * A;
        l1 += u[1] * A;
        l2 += u[2] * A;
        l3 += u[3] * A;
        if (s->dim == 5) temp_E += u[4] * A;
    }
    out[0] = temp; out[1] = l1; out[2] = l2; out[3] = l3; out[4] = temp_E;
}

or_solver *or_create(const or_mesh *m, const double *U_in, int dim, int solver, double gam,
                     double omega_ns, const or_params *p, int threads)
{
    if (!m || !U_in || solver < OR_HLLE || solver > OR_HLLCPLUS) return NULL;
    if ((solver == OR_HLLE && dim != 4) || (solver != OR_HLLE && dim != 5)) return NULL;
    or_solver *s = calloc(1, sizeof(or_solver));
    s->mesh.nv = m->nv;
    s->mesh.nf = m->nf;
    s->mesh.xyz = malloc(sizeof(double) * 3 * m->nv);
    memcpy(s->mesh.xyz, m->xyz, sizeof(double) * 3 * m->nv);
    s->mesh.foff = malloc(sizeof(int32_t) * (m->nf + 1));
    memcpy(s->mesh.foff, m->foff, sizeof(int32_t) * (m->nf + 1));
    s->mesh.fv = malloc(sizeof(int32_t) * m->foff[m->nf]);
    memcpy(s->mesh.fv, m->fv, sizeof(int32_t) * m->foff[m->nf]);
    s->g = or_geometry(&s->mesh);
    s->nf = m->nf;
    s->nh = m->foff[m->nf];
    s->dim = dim;
    s->solver = solver;
    s->threads = threads > 0 ? threads : 1;
    s->gam = gam;
    s->omega_ns = omega_ns;
    if (p) s->p = *p;
    size_t nU = (size_t)s->nf * dim, nH = (size_t)s->nh * dim;
    s->U = malloc(sizeof(double) * nU);
    memcpy(s->U, U_in, sizeof(double) * nU);
    s->U_temp = malloc(sizeof(double) * nU);
    s->src = calloc(nU, sizeof(double));
    s->Uplus = calloc(nH, sizeof(double));
    s->Uminus = calloc(nH, sizeof(double));
    s->flux = calloc(nH, sizeof(double));
    if (is_adiabatic(s)) {
        /* adiabatic.hpp:70-90 */
        double area_coeff = 0;
        double mu = M_PI / 2;
        double sigma = s->p.acc_width / 2.355;
        for (int f = 0; f < s->nf; f++) {
            v3 fc = v3ld(s->g->fc + 3 * f);
            double theta = acos(fc.x[2] / v3norm(fc));
            area_coeff += s->g->area[f] / sqrt(2 * M_PI * sigma * sigma) *
                          exp(-(theta - mu) * (theta - mu) / (2 * sigma * sigma));
        }
        area_coeff /= 4 * M_PI;
        s->area_coeff = area_coeff;
        double tm = 0;
        for (int f = 0; f < s->nf; f++) tm += s->U[(size_t)f * dim] * s->g->area[f];
        s->total_mass = tm;
    }
    return s;
}

double or_time(const or_solver *s) { return s->t; }
int or_stop_check(const or_solver *s) { return s->stop; }
double or_last_dt(const or_solver *s) { return s->dt; }
double or_area_coeff(const or_solver *s) { return s->area_coeff; }
/* write_t_curl (adiabatic.hpp:135-178, isothermal.hpp:72-110): per face,
 * sum over its edges of ((edge_center x L+) / (-Sigma)) . (v_j - v_j1) / A,
 * with L+ the face's own reconstructed state of the last find_U_edges. */
void or_curl(const or_solver *s, double *out)
{
    const or_mesh *m = &s->mesh;
    for (int f = 0; f < s->nf; f++) {
        const int n = m->foff[f + 1] - m->foff[f];
        double vort = 0;
        for (int j = 0; j < n; j++) {
            const double *up = s->Uplus + (size_t)(m->foff[f] + j) * s->dim;
            v3 l = v3m(up[1], up[2], up[3]);
            const int j1 = (j == n - 1) ? 0 : j + 1;
            v3 ec = edge_center(s, f, j);
            v3 vel = v3divs(v3cross(ec, l), -s->U[(size_t)f * s->dim]);
            v3 r = v3sub(v3ld(m->xyz + 3 * m->fv[m->foff[f] + j]), v3ld(m->xyz + 3 * m->fv[m->foff[f] + j1]));
            vort += v3dot(vel, r);
        }
        out[f] = vort / s->g->area[f];
    }
}

void or_reconstruct(or_solver *s, double dt, double *Uplus_out)
{
    const size_t n = (size_t)s->nf * s->dim;
    double dt_saved = s->dt;
    memcpy(s->U_temp, s->U, sizeof(double) * n);
    s->dt = dt;
    find_U_edges(s);
    memcpy(Uplus_out, s->Uplus, sizeof(double) * (size_t)s->nh * s->dim);
    memcpy(s->U, s->U_temp, sizeof(double) * n);
    s->dt = dt_saved;
}

void or_get_state(const or_solver *s, double *U_out)
{
    memcpy(U_out, s->U, sizeof(double) * (size_t)s->nf * s->dim);
}

void or_destroy(or_solver *s)
{
    if (!s) return;
    or_geom_free(s->g);
    free(s->mesh.xyz); free(s->mesh.foff); free(s->mesh.fv);
    free(s->U); free(s->U_temp); free(s->src); free(s->Uplus); free(s->Uminus); free(s->flux);
    free(s);
}
