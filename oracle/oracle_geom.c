/*
 * Oracle (test infrastructure): restatement of MUSCL_base_geometry
 * (src/geometry/MUSCL_geometry.hpp:43-575, :981-1030) plus the h0 scan of
 * MUSCL_base (src/MUSCL_base/MUSCL_base.hpp:62-91) and round_diff
 * (MUSCL_base.hpp:242).
 */
#include "oracle.h"
#include "oracle_vec.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static int face_has(const or_mesh *m, int f, int v)
{
    for (int k = m->foff[f]; k < m->foff[f + 1]; ++k)
        if (m->fv[k] == v) return 1;
    return 0;
}

/* MUSCL_geometry.hpp:846-870 */
static v3 line_surf_intersection(v3 xl, v3 nl, v3 xs, v3 vs1, v3 vs2)
{
    double A = vs1.x[1] * vs2.x[2] - vs1.x[2] * vs2.x[1];
    double B = vs1.x[2] * vs2.x[0] - vs1.x[0] * vs2.x[2];
    double C = vs1.x[0] * vs2.x[1] - vs1.x[1] * vs2.x[0];
    double D = -xs.x[0] * A - xs.x[1] * B - xs.x[2] * C;
    double t = -(D + A * xl.x[0] + B * xl.x[1] + C * xl.x[2]) / (A * nl.x[0] + B * nl.x[1] + C * nl.x[2]);
    return v3add(xl, v3muls(nl, t));
}

or_geom *or_geometry(const or_mesh *m)
{
    const int nf = m->nf, nv = m->nv, nh = m->foff[nf];
    or_geom *g = calloc(1, sizeof(or_geom));
    g->nf = nf; g->nh = nh; g->nv = nv; g->mesh = m;
    g->fc = malloc(sizeof(double) * 3 * nf);
    g->normal = malloc(sizeof(double) * 3 * nf);
    g->area = malloc(sizeof(double) * nf);
    g->nb_off = malloc(sizeof(int32_t) * (nf + 1));
    g->nb_edge = malloc(sizeof(int32_t) * nh);
    g->twin = malloc(sizeof(int32_t) * nh);
    g->en = malloc(sizeof(double) * 3 * nh);
    g->bm_dist = malloc(sizeof(double) * nh);
    g->ff_plus = malloc(sizeof(int32_t) * 2 * nh);
    g->ff_minus = malloc(sizeof(int32_t) * 2 * nh);
    g->beta_plus = malloc(sizeof(double) * 2 * nh);
    g->beta_minus = malloc(sizeof(double) * 2 * nh);
    g->h_plus = malloc(sizeof(double) * nh);
    g->h_minus = malloc(sizeof(double) * nh);

    /* vertex -> faces incidence (the reference walks pmp's face circulator;
     * the list is sorted and de-duplicated below, so only the set matters) */
    int *vf_cnt = calloc(nv + 1, sizeof(int));
    for (int k = 0; k < nh; ++k) vf_cnt[m->fv[k] + 1]++;
    for (int v = 0; v < nv; ++v) vf_cnt[v + 1] += vf_cnt[v];
    int *vf = malloc(sizeof(int) * (nh ? nh : 1));
    int *fill = calloc(nv, sizeof(int));
    for (int f = 0; f < nf; ++f)
        for (int k = m->foff[f]; k < m->foff[f + 1]; ++k) {
            int v = m->fv[k];
            vf[vf_cnt[v] + fill[v]++] = f;
        }

    /* :91-118 neighbors = sorted unique faces sharing a vertex, minus self */
    int cap = 1024, total = 0;
    int32_t *nb = malloc(sizeof(int32_t) * cap);
    int *tmp = malloc(sizeof(int) * 4096);
    for (int f = 0; f < nf; ++f) {
        int n = 0;
        for (int k = m->foff[f]; k < m->foff[f + 1]; ++k) {
            int v = m->fv[k];
            for (int q = vf_cnt[v]; q < vf_cnt[v + 1]; ++q) {
                if (n >= 4096) abort();
                tmp[n++] = vf[q];
            }
        }
        qsort(tmp, n, sizeof(int), cmp_int);
        g->nb_off[f] = total;
        for (int q = 0; q < n; ++q) {
            if (q > 0 && tmp[q] == tmp[q - 1]) continue;
            if (tmp[q] == f) continue;
            if (total == cap) {
                cap *= 2;
                nb = realloc(nb, sizeof(int32_t) * cap);
            }
            nb[total++] = tmp[q];
        }
    }
    g->nb_off[nf] = total;
    g->nb = nb;
    free(tmp);
    free(vf);
    free(vf_cnt);
    free(fill);

    /* :145-174 neighbors_edge: the (single) neighbour holding both edge ends */
    for (int f = 0; f < nf; ++f) {
        int n = m->foff[f + 1] - m->foff[f];
        for (int j = 0; j < n; ++j) {
            int f1 = m->fv[m->foff[f] + j];
            int f2 = m->fv[m->foff[f] + (j != n - 1 ? j + 1 : 0)];
            int found = -1, count = 0;
            for (int q = g->nb_off[f]; q < g->nb_off[f + 1]; ++q)
                if (face_has(m, nb[q], f1) && face_has(m, nb[q], f2)) {
                    if (found < 0) found = nb[q];
                    count++;
                }
            if (count != 1) {
                fprintf(stderr, "oracle geometry: face %d edge %d has %d edge neighbours\n", f, j, count);
                abort();
            }
            g->nb_edge[m->foff[f] + j] = found;
        }
    }
    /* MUSCL_base.hpp:383-384: j0 = position of f in neighbors_edge[nb] */
    for (int f = 0; f < nf; ++f)
        for (int h = m->foff[f]; h < m->foff[f + 1]; ++h) {
            int o = g->nb_edge[h], j0 = -1;
            for (int q = m->foff[o]; q < m->foff[o + 1]; ++q)
                if (g->nb_edge[q] == f) { j0 = q; break; }
            if (j0 < 0) abort();
            g->twin[h] = j0;
        }

    /* :176-187 face centres */
    for (int f = 0; f < nf; ++f) {
        v3 c = v3m(0, 0, 0);
        for (int k = m->foff[f]; k < m->foff[f + 1]; ++k) c = v3add(c, v3ld(m->xyz + 3 * m->fv[k]));
        c = v3divs(c, (double)(m->foff[f + 1] - m->foff[f]));
        memcpy(g->fc + 3 * f, c.x, sizeof(c.x));
    }
    /* :189-218 normals */
    for (int f = 0; f < nf; ++f) {
        const int32_t *F = m->fv + m->foff[f];
        int n = m->foff[f + 1] - m->foff[f];
        v3 nr = v3m(0, 0, 0);
        for (int j = 0; j < n; ++j) {
            v3 r1, r2;
            if (j + 2 == n) {
                r2 = v3sub(v3ld(m->xyz + 3 * F[0]), v3ld(m->xyz + 3 * F[j + 1]));
                r1 = v3sub(v3ld(m->xyz + 3 * F[j + 1]), v3ld(m->xyz + 3 * F[j]));
            } else if (j + 1 == n) {
                r1 = v3sub(v3ld(m->xyz + 3 * F[0]), v3ld(m->xyz + 3 * F[j]));
                r2 = v3sub(v3ld(m->xyz + 3 * F[1]), v3ld(m->xyz + 3 * F[0]));
            } else {
                r2 = v3sub(v3ld(m->xyz + 3 * F[j + 2]), v3ld(m->xyz + 3 * F[j + 1]));
                r1 = v3sub(v3ld(m->xyz + 3 * F[j + 1]), v3ld(m->xyz + 3 * F[j]));
            }
            nr = v3add(nr, v3cross(r1, r2));
        }
        nr = v3divs(nr, v3norm(nr));
        memcpy(g->normal + 3 * f, nr.x, sizeof(nr.x));
    }
    /* :989-1030 spherical-excess areas */
    for (int f = 0; f < nf; ++f) {
        const int32_t *F = m->fv + m->foff[f];
        int n = m->foff[f + 1] - m->foff[f];
        v3 fc = v3ld(g->fc + 3 * f);
        v3 fcn = v3divs(fc, v3norm(fc));
        double S = 0;
        for (int j = 0; j < n; ++j) {
            int j1 = (j != n - 1) ? j + 1 : 0;
            double a = gc_distance(fcn, v3ld(m->xyz + 3 * F[j]));
            double b = gc_distance(fcn, v3ld(m->xyz + 3 * F[j1]));
            double c = gc_distance(v3ld(m->xyz + 3 * F[j]), v3ld(m->xyz + 3 * F[j1]));
            double A = acos((cos(a) - cos(b) * cos(c)) / (sin(b) * sin(c)));
            double B = acos((cos(b) - cos(a) * cos(c)) / (sin(a) * sin(c)));
            double C = acos((cos(c) - cos(b) * cos(a)) / (sin(b) * sin(a)));
            S += A + B + C - M_PI;
        }
        g->area[f] = S;
    }

    /* :227-575 process_mesh */
    for (int f = 0; f < nf; ++f) {
        const int32_t *F = m->fv + m->foff[f];
        int n = m->foff[f + 1] - m->foff[f];
        v3 fc = v3ld(g->fc + 3 * f);
        v3 nrm = v3ld(g->normal + 3 * f);
        for (int i = 0; i < n; ++i) {
            int h = m->foff[f] + i;
            int i1 = (i == n - 1) ? 0 : i + 1;
            v3 vi = v3ld(m->xyz + 3 * F[i]), vi1 = v3ld(m->xyz + 3 * F[i1]);
            v3 BM = v3sub(v3divs(v3add(vi, vi1), 2), fc);
            v3 r = v3sub(vi, vi1);
            v3 r1 = v3add(vi, vi1);
            r1 = v3divs(r1, v3norm(r1));
            v3 en = v3cross(r1, r);
            en = v3divs(en, v3norm(en));
            if (v3norm(v3add(BM, v3muls(en, 0.01))) < v3norm(BM)) en = v3muls(en, -1);
            memcpy(g->en + 3 * h, en.x, sizeof(en.x));
            g->bm_dist[h] = gc_distance(fc, r1);

            double max_l = -1, max_r = -1;
            int left1 = -1, right1 = -1, left2, right2;
            for (int q = g->nb_off[f]; q < g->nb_off[f + 1]; ++q) {
                int o = nb[q];
                v3 fo = v3ld(g->fc + 3 * o);
                v3 B_nB = v3sub(fc, fo);
                v3 B_nB_face = v3sub(B_nB, v3muls(nrm, v3dot(B_nB, nrm)));
                double ce = v3dot(BM, B_nB_face) / (v3norm(BM) * v3norm(B_nB_face));
                if (ce > max_l) { max_l = ce; left1 = o; }
                B_nB = v3sub(fo, fc);
                B_nB_face = v3sub(B_nB, v3muls(nrm, v3dot(B_nB, nrm)));
                ce = v3dot(BM, B_nB_face) / (v3norm(BM) * v3norm(B_nB_face));
                if (ce > max_r) { max_r = ce; right1 = o; }
            }
            if (left1 < 0 || right1 < 0) {
                fprintf(stderr, "oracle geometry: no stencil face for face %d edge %d\n", f, i);
                abort();
            }
            max_l = -1; max_r = -1;
            left2 = left1; right2 = right1;
            v3 B_left = v3sub(fc, v3ld(g->fc + 3 * left1));
            v3 B_right = v3sub(v3ld(g->fc + 3 * right1), fc);
            v3 B_right_face = v3sub(B_right, v3muls(nrm, v3dot(B_right, nrm)));
            v3 cr = v3cross(fc, v3divs(r1, 2));
            for (int q = g->nb_off[f]; q < g->nb_off[f + 1]; ++q) {
                int o = nb[q];
                v3 fo = v3ld(g->fc + 3 * o);
                v3 B_nB = v3sub(fc, fo);
                v3 B_nB_face = v3sub(B_nB, v3muls(nrm, v3dot(B_nB, nrm)));
                double ce = v3dot(BM, B_nB_face) / (v3norm(BM) * v3norm(B_nB_face));
                double lcc = v3dot(cr, B_nB) * v3dot(cr, B_left);
                if ((ce > max_l) && (o != left1) && (lcc <= 0)) { max_l = ce; left2 = o; }
                B_nB = v3sub(fo, fc);
                B_nB_face = v3sub(B_nB, v3muls(nrm, v3dot(B_nB, nrm)));
                ce = v3dot(BM, B_nB_face) / (v3norm(BM) * v3norm(B_nB_face));
                lcc = v3dot(nrm, v3cross(B_nB_face, BM)) * v3dot(nrm, v3cross(B_right_face, BM));
                if ((ce > max_r) && (o != right1) && (lcc <= 0)) { max_r = ce; right2 = o; }
            }
            v3 fl1 = v3ld(g->fc + 3 * left1), fl2 = v3ld(g->fc + 3 * left2);
            v3 fr1 = v3ld(g->fc + 3 * right1), fr2 = v3ld(g->fc + 3 * right2);
            v3 p3 = line_surf_intersection(fl1, v3sub(fl2, fl1), fc, BM, fc);
            v3 Hm = v3divs(p3, v3norm(p3));
            p3 = line_surf_intersection(fr1, v3sub(fr2, fr1), fc, BM, fc);
            v3 Hp = v3divs(p3, v3norm(p3));

            double Hm_dist = gc_distance(fc, Hm);
            double B1B2_d = gc_distance(fl1, fl2);
            g->h_minus[h] = Hm_dist;
            g->ff_minus[2 * h] = left1;
            g->ff_minus[2 * h + 1] = left2;
            double hm2 = gc_distance(fl2, Hm);
            double hm1 = gc_distance(fl1, Hm);
            g->beta_minus[2 * h] = hm2 / B1B2_d;
            g->beta_minus[2 * h + 1] = hm1 / B1B2_d;

            double Hp_dist = gc_distance(fc, Hp);
            double B1B2p_d = gc_distance(fr1, fr2);
            g->h_plus[h] = Hp_dist;
            g->ff_plus[2 * h] = right1;
            g->ff_plus[2 * h + 1] = right2;
            double hp2 = gc_distance(fr2, Hp);
            double hp1 = gc_distance(fr1, Hp);
            g->beta_plus[2 * h] = hp2 / B1B2p_d;
            g->beta_plus[2 * h + 1] = hp1 / B1B2p_d;
        }
    }

    /* MUSCL_base.hpp:62-91 h0 = min great-circle edge length */
    double h0 = 10;
    for (int f = 0; f < nf; ++f) {
        const int32_t *F = m->fv + m->foff[f];
        int n = m->foff[f + 1] - m->foff[f];
        for (int j = 0; j < n; ++j) {
            double d = (j < n - 1) ? gc_distance(v3ld(m->xyz + 3 * F[j]), v3ld(m->xyz + 3 * F[j + 1]))
                                   : gc_distance(v3ld(m->xyz + 3 * F[j]), v3ld(m->xyz + 3 * F[0]));
            if (d < h0) h0 = d;
        }
    }
    g->h0 = h0;
    /* MUSCL_base.hpp:242 */
    {
        v3 a = v3ld(m->xyz + 3 * m->fv[m->foff[0] + 1]);
        v3 b = v3ld(m->xyz + 3 * m->fv[m->foff[0] + 2]);
        g->round_diff = gc_distance(a, b) / v3norm(v3sub(a, b));
    }
    return g;
}

void or_geom_free(or_geom *g)
{
    if (!g) return;
    free(g->fc); free(g->normal); free(g->area); free(g->nb_off); free(g->nb);
    free(g->nb_edge); free(g->twin); free(g->en); free(g->bm_dist);
    free(g->ff_plus); free(g->ff_minus); free(g->beta_plus); free(g->beta_minus);
    free(g->h_plus); free(g->h_minus);
    free(g);
}
