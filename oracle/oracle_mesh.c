/*
 * Oracle (test infrastructure): restatement of the reference's mesh
 * generators -- src/geometry/sph_gen.h and the parts of the vendored pmp
 * library they drive (src/pmp/surface_mesh.cpp add_face / insert_vertex /
 * insert_edge, src/pmp/algorithms/subdivision.cpp loop / Catmull-Clark,
 * the dual at sph_gen.h:72-96).  Face/vertex numbering and coordinate bits
 * follow pmp's construction order operation for operation, because the
 * stencil choice downstream is an argmax with strict '>' over face ids.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define INV (-1)

typedef struct hm {
    /* halfedge connectivity (surface_mesh.h HalfedgeConnectivity) */
    int *h_to, *h_face, *h_next, *h_prev;
    int nh, caph;
    /* vertices */
    int *v_he;
    double *vp;
    int nv, capv;
    /* faces */
    int *f_he;
    int nf, capf;
} hm;

static void *xrealloc(void *p, size_t n)
{
    void *q = realloc(p, n);
    if (!q && n) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return q;
}

static void hm_init(hm *m) { memset(m, 0, sizeof(*m)); }

static void hm_free(hm *m)
{
    free(m->h_to); free(m->h_face); free(m->h_next); free(m->h_prev);
    free(m->v_he); free(m->vp); free(m->f_he);
    hm_init(m);
}

/* surface_mesh.cpp:198-206 add_vertex */
static int hm_add_vertex(hm *m, double x, double y, double z)
{
    if (m->nv == m->capv) {
        m->capv = m->capv ? 2 * m->capv : 64;
        m->v_he = xrealloc(m->v_he, sizeof(int) * m->capv);
        m->vp = xrealloc(m->vp, sizeof(double) * 3 * m->capv);
    }
    int v = m->nv++;
    m->v_he[v] = INV;
    m->vp[3 * v] = x; m->vp[3 * v + 1] = y; m->vp[3 * v + 2] = z;
    return v;
}

/* surface_mesh.h:1943-1963 new_edge(start, end): returns h0 (even), to(h0)=end */
static int hm_new_edge(hm *m, int start, int end)
{
    if (m->nh + 2 > m->caph) {
        m->caph = m->caph ? 2 * m->caph : 128;
        m->h_to = xrealloc(m->h_to, sizeof(int) * m->caph);
        m->h_face = xrealloc(m->h_face, sizeof(int) * m->caph);
        m->h_next = xrealloc(m->h_next, sizeof(int) * m->caph);
        m->h_prev = xrealloc(m->h_prev, sizeof(int) * m->caph);
    }
    int h0 = m->nh, h1 = m->nh + 1;
    m->nh += 2;
    m->h_face[h0] = m->h_face[h1] = INV;
    m->h_next[h0] = m->h_next[h1] = INV;
    m->h_prev[h0] = m->h_prev[h1] = INV;
    m->h_to[h0] = end;
    m->h_to[h1] = start;
    return h0;
}

/* surface_mesh.h:1968-1977 new_face */
static int hm_new_face(hm *m)
{
    if (m->nf == m->capf) {
        m->capf = m->capf ? 2 * m->capf : 64;
        m->f_he = xrealloc(m->f_he, sizeof(int) * m->capf);
    }
    m->f_he[m->nf] = INV;
    return m->nf++;
}

static inline int opp(int h) { return (h & 1) ? h - 1 : h + 1; }
static inline void set_next(hm *m, int h, int nh)
{
    m->h_next[h] = nh;
    m->h_prev[nh] = h;
}
static inline int cw(const hm *m, int h) { return m->h_next[opp(h)]; }
static inline int ccw(const hm *m, int h) { return opp(m->h_prev[h]); }
static inline int is_bnd_h(const hm *m, int h) { return m->h_face[h] == INV; }
static inline int is_bnd_v(const hm *m, int v)
{
    int h = m->v_he[v];
    return !(h != INV && m->h_face[h] != INV);
}

/* surface_mesh.cpp:172-190 find_halfedge */
static int hm_find_halfedge(const hm *m, int start, int end)
{
    int h = m->v_he[start];
    const int hh = h;
    if (h != INV) {
        do {
            if (m->h_to[h] == end) return h;
            h = cw(m, h);
        } while (h != hh);
    }
    return INV;
}

/* surface_mesh.cpp:198-214 adjust_outgoing_halfedge */
static void hm_adjust_outgoing(hm *m, int v)
{
    int h = m->v_he[v];
    const int hh = h;
    if (h != INV) {
        do {
            if (is_bnd_h(m, h)) {
                m->v_he[v] = h;
                return;
            }
            h = cw(m, h);
        } while (h != hh);
    }
}

/* surface_mesh.cpp:244-423 add_face (topology errors abort: the generators
 * only build valid closed meshes). */
static int hm_add_face(hm *m, const int *vs, int n)
{
    int hs[64], is_new[64], needs_adjust[64];
    int cache[6 * 64][2];
    int nc = 0;
    int i, ii;
    if (n > 64) abort();
    for (i = 0; i < n; ++i) needs_adjust[i] = 0;

    for (i = 0, ii = 1; i < n; ++i, ++ii, ii %= n) {
        if (!is_bnd_v(m, vs[i])) {
            fprintf(stderr, "oracle add_face: complex vertex\n");
            abort();
        }
        hs[i] = hm_find_halfedge(m, vs[i], vs[ii]);
        is_new[i] = hs[i] == INV;
        if (!is_new[i] && !is_bnd_h(m, hs[i])) {
            fprintf(stderr, "oracle add_face: complex edge\n");
            abort();
        }
    }

    /* re-link patches if necessary (:287-332) */
    for (i = 0, ii = 1; i < n; ++i, ++ii, ii %= n) {
        if (!is_new[i] && !is_new[ii]) {
            int inner_prev = hs[i], inner_next = hs[ii];
            if (m->h_next[inner_prev] != inner_next) {
                int outer_prev = opp(inner_next);
                int boundary_prev = outer_prev;
                do {
                    boundary_prev = opp(m->h_next[boundary_prev]);
                } while (!is_bnd_h(m, boundary_prev) || boundary_prev == inner_prev);
                int boundary_next = m->h_next[boundary_prev];
                if (boundary_next == inner_next) {
                    fprintf(stderr, "oracle add_face: patch re-linking failed\n");
                    abort();
                }
                int patch_start = m->h_next[inner_prev];
                int patch_end = m->h_prev[inner_next];
                cache[nc][0] = boundary_prev; cache[nc][1] = patch_start; nc++;
                cache[nc][0] = patch_end; cache[nc][1] = boundary_next; nc++;
                cache[nc][0] = inner_prev; cache[nc][1] = inner_next; nc++;
            }
        }
    }

    /* create missing edges (:334-341) */
    for (i = 0, ii = 1; i < n; ++i, ++ii, ii %= n)
        if (is_new[i]) hs[i] = hm_new_edge(m, vs[i], vs[ii]);

    /* create the face (:343-345) */
    int f = hm_new_face(m);
    m->f_he[f] = hs[n - 1];

    /* setup halfedges (:347-404) */
    for (i = 0, ii = 1; i < n; ++i, ++ii, ii %= n) {
        int v = vs[ii];
        int inner_prev = hs[i], inner_next = hs[ii];
        int id = 0;
        if (is_new[i]) id |= 1;
        if (is_new[ii]) id |= 2;
        if (id) {
            int outer_prev = opp(inner_next);
            int outer_next = opp(inner_prev);
            switch (id) {
            case 1: {
                int boundary_prev = m->h_prev[inner_next];
                cache[nc][0] = boundary_prev; cache[nc][1] = outer_next; nc++;
                m->v_he[v] = outer_next;
                break;
            }
            case 2: {
                int boundary_next = m->h_next[inner_prev];
                cache[nc][0] = outer_prev; cache[nc][1] = boundary_next; nc++;
                m->v_he[v] = boundary_next;
                break;
            }
            case 3:
                if (m->v_he[v] == INV) {
                    m->v_he[v] = outer_next;
                    cache[nc][0] = outer_prev; cache[nc][1] = outer_next; nc++;
                } else {
                    int boundary_next = m->v_he[v];
                    int boundary_prev = m->h_prev[boundary_next];
                    cache[nc][0] = boundary_prev; cache[nc][1] = outer_next; nc++;
                    cache[nc][0] = outer_prev; cache[nc][1] = boundary_next; nc++;
                }
                break;
            }
            cache[nc][0] = inner_prev; cache[nc][1] = inner_next; nc++;
        } else {
            needs_adjust[ii] = (m->v_he[v] == inner_next);
        }
        m->h_face[hs[i]] = f;
    }

    /* process next-halfedge cache (:406-411) */
    for (i = 0; i < nc; ++i) set_next(m, cache[i][0], cache[i][1]);

    /* adjust vertices' halfedge handle (:413-420) */
    for (i = 0; i < n; ++i)
        if (needs_adjust[i]) hm_adjust_outgoing(m, vs[i]);
    return f;
}

/* surface_mesh.cpp:598-649 insert_vertex(h0, v) */
static int hm_insert_vertex(hm *m, int h0, int v)
{
    int h2 = m->h_next[h0];
    int o0 = opp(h0);
    int o2 = m->h_prev[o0];
    int v2 = m->h_to[h0];
    int fh = m->h_face[h0];
    int fo = m->h_face[o0];

    int h1 = hm_new_edge(m, v, v2);
    int o1 = opp(h1);

    set_next(m, h1, h2);
    set_next(m, h0, h1);
    m->h_to[h0] = v;
    m->h_to[h1] = v2;
    m->h_face[h1] = fh;

    set_next(m, o1, o0);
    set_next(m, o2, o1);
    m->h_to[o1] = v;
    m->h_face[o1] = fo;

    m->v_he[v2] = o1;
    hm_adjust_outgoing(m, v2);
    m->v_he[v] = h1;
    hm_adjust_outgoing(m, v);

    if (fh != INV) m->f_he[fh] = h0;
    if (fo != INV) m->f_he[fo] = o1;
    return o1;
}

/* surface_mesh.h:1759-1762 insert_vertex(Edge, Point) */
static void hm_insert_vertex_on_edge(hm *m, int e, const double p[3])
{
    int v = hm_add_vertex(m, p[0], p[1], p[2]);
    hm_insert_vertex(m, e << 1, v);
}

/* surface_mesh.cpp:651-685 insert_edge */
static int hm_insert_edge(hm *m, int h0, int h1)
{
    int v0 = m->h_to[h0];
    int v1 = m->h_to[h1];
    int h2 = m->h_next[h0];
    int h3 = m->h_next[h1];
    int h4 = hm_new_edge(m, v0, v1);
    int h5 = opp(h4);
    int f0 = m->h_face[h0];
    int f1 = hm_new_face(m);
    m->f_he[f0] = h0;
    m->f_he[f1] = h1;
    set_next(m, h0, h4);
    set_next(m, h4, h3);
    m->h_face[h4] = f0;
    set_next(m, h1, h5);
    set_next(m, h5, h2);
    int h = h2;
    do {
        m->h_face[h] = f1;
        h = m->h_next[h];
    } while (h != h2);
    return h4;
}

/* pmp sqrnorm/norm (mat_vec.h:512-525): s = 0; s += x*x; ... */
static double pnorm(const double *p)
{
    double s = 0.0;
    s += p[0] * p[0];
    s += p[1] * p[1];
    s += p[2] * p[2];
    return sqrt(s);
}

/* sph_gen.h:140-148 project_to_unit_sphere: p = (1./n) * p */
static void hm_project(hm *m)
{
    for (int v = 0; v < m->nv; ++v) {
        double *p = m->vp + 3 * v;
        double s = 1. / pnorm(p);
        p[0] *= s; p[1] *= s; p[2] *= s;
    }
}

/* differential_geometry.cpp:150-161 / sph_gen.h:59-70 centroid(face) */
static void hm_centroid(const hm *m, int f, double c[3])
{
    double n = 0;
    c[0] = c[1] = c[2] = 0;
    int h = m->f_he[f], hh = h;
    do {
        const double *p = m->vp + 3 * m->h_to[h];
        c[0] += p[0]; c[1] += p[1]; c[2] += p[2];
        n += 1;
        h = m->h_next[h];
    } while (h != hh);
    c[0] /= n; c[1] /= n; c[2] /= n;
}

/* subdivision.cpp:180-342 loop_subdivision (closed, no features) */
static void hm_loop_subdivision(hm *m)
{
    const int nv = m->nv, ne = m->nh / 2, nf = m->nf;
    double *vpoint = malloc(sizeof(double) * 3 * nv);
    double *epoint = malloc(sizeof(double) * 3 * ne);

    for (int v = 0; v < nv; ++v) {
        double p[3] = {0, 0, 0};
        double k = 0;
        int h = m->v_he[v], hh = h;
        do {
            const double *q = m->vp + 3 * m->h_to[h];
            p[0] += q[0]; p[1] += q[1]; p[2] += q[2];
            k += 1;
            h = ccw(m, h);
        } while (h != hh);
        p[0] /= k; p[1] /= k; p[2] /= k;
        double beta = (0.625 - pow(0.375 + 0.25 * cos(2.0 * M_PI / k), 2.0));
        double omb = 1.0 - beta;
        const double *x = m->vp + 3 * v;
        for (int c = 0; c < 3; ++c) vpoint[3 * v + c] = x[c] * omb + p[c] * beta;
    }
    for (int e = 0; e < ne; ++e) {
        int h0 = 2 * e, h1 = 2 * e + 1;
        double p[3];
        const double *a = m->vp + 3 * m->h_to[h0];
        const double *b = m->vp + 3 * m->h_to[h1];
        const double *c0 = m->vp + 3 * m->h_to[m->h_next[h0]];
        const double *c1 = m->vp + 3 * m->h_to[m->h_next[h1]];
        for (int c = 0; c < 3; ++c) {
            p[c] = a[c];
            p[c] += b[c];
            p[c] *= 3.0;
            p[c] += c0[c];
            p[c] += c1[c];
            p[c] *= 0.125;
            epoint[3 * e + c] = p[c];
        }
    }
    memcpy(m->vp, vpoint, sizeof(double) * 3 * nv);
    for (int e = 0; e < ne; ++e) hm_insert_vertex_on_edge(m, e, epoint + 3 * e);
    for (int f = 0; f < nf; ++f) {
        int h = m->f_he[f];
        hm_insert_edge(m, h, m->h_next[m->h_next[h]]);
        h = m->h_next[h];
        hm_insert_edge(m, h, m->h_next[m->h_next[h]]);
        h = m->h_next[h];
        hm_insert_edge(m, h, m->h_next[m->h_next[h]]);
    }
    free(vpoint);
    free(epoint);
}

/* subdivision.cpp:9-178 catmull_clark_subdivision (closed, no features) */
static void hm_catmull_clark(hm *m)
{
    const int nv = m->nv, ne = m->nh / 2, nf = m->nf;
    double *vpoint = malloc(sizeof(double) * 3 * nv);
    double *epoint = malloc(sizeof(double) * 3 * ne);
    double *fpoint = malloc(sizeof(double) * 3 * nf);

    for (int f = 0; f < nf; ++f) hm_centroid(m, f, fpoint + 3 * f);
    for (int e = 0; e < ne; ++e) {
        const double *a = m->vp + 3 * m->h_to[2 * e];
        const double *b = m->vp + 3 * m->h_to[2 * e + 1];
        const double *fa = fpoint + 3 * m->h_face[2 * e];
        const double *fb = fpoint + 3 * m->h_face[2 * e + 1];
        for (int c = 0; c < 3; ++c) {
            double p = 0;
            p += a[c];
            p += b[c];
            p += fa[c];
            p += fb[c];
            p *= 0.25;
            epoint[3 * e + c] = p;
        }
    }
    for (int v = 0; v < nv; ++v) {
        /* valence(v): number of vertices around v */
        double k = 0;
        int h = m->v_he[v], hh = h;
        do { k += 1; h = ccw(m, h); } while (h != hh);
        double p[3] = {0, 0, 0};
        h = hh;
        do {
            const double *q = m->vp + 3 * m->h_to[h];
            p[0] += q[0]; p[1] += q[1]; p[2] += q[2];
            h = ccw(m, h);
        } while (h != hh);
        h = hh; /* closed mesh: the face circulator starts at halfedge(v) */
        do {
            const double *q = fpoint + 3 * m->h_face[h];
            p[0] += q[0]; p[1] += q[1]; p[2] += q[2];
            h = ccw(m, h);
        } while (h != hh);
        double kk = k * k;
        p[0] /= kk; p[1] /= kk; p[2] /= kk;
        double w = (k - 2.0) / k;
        const double *x = m->vp + 3 * v;
        for (int c = 0; c < 3; ++c) vpoint[3 * v + c] = p[c] + x[c] * w;
    }
    memcpy(m->vp, vpoint, sizeof(double) * 3 * nv);
    for (int e = 0; e < ne; ++e) hm_insert_vertex_on_edge(m, e, epoint + 3 * e);
    for (int f = 0; f < nf; ++f) {
        int h0 = m->f_he[f];
        hm_insert_edge(m, h0, m->h_next[m->h_next[h0]]);
        int h1 = m->h_next[h0];
        hm_insert_vertex_on_edge(m, h1 >> 1, fpoint + 3 * f);
        int h = m->h_next[m->h_next[m->h_next[h1]]];
        while (h != h0) {
            hm_insert_edge(m, h1, h);
            h = m->h_next[m->h_next[m->h_next[h1]]];
        }
    }
    free(vpoint);
    free(epoint);
    free(fpoint);
}

/* sph_gen.h:72-96 dual, then mesh.assign(tmp) */
static void hm_dual(hm *m)
{
    hm t;
    hm_init(&t);
    int *fvertex = malloc(sizeof(int) * m->nf);
    for (int f = 0; f < m->nf; ++f) {
        double c[3];
        hm_centroid(m, f, c);
        fvertex[f] = hm_add_vertex(&t, c[0], c[1], c[2]);
    }
    int buf[64];
    for (int v = 0; v < m->nv; ++v) {
        int n = 0;
        int h = m->v_he[v];
        if (h != INV && is_bnd_h(m, h)) {
            do { h = ccw(m, h); } while (is_bnd_h(m, h));
        }
        int hh = h;
        do {
            if (n >= 64) abort();
            buf[n++] = fvertex[m->h_face[h]];
            do { h = ccw(m, h); } while (is_bnd_h(m, h));
        } while (h != hh);
        hm_add_face(&t, buf, n);
    }
    free(fvertex);
    hm_free(m);
    *m = t;
}

/* sph_gen.h:158-205 icosahedron */
static void hm_icosahedron(hm *m)
{
    double phi = (1. + sqrt(5.)) * 0.5;
    double a = 1.0;
    double b = 1.0 / phi;
    hm_add_vertex(m, 0, b, -a);  /* v1 -> 0 */
    hm_add_vertex(m, b, a, 0);   /* v2 -> 1 */
    hm_add_vertex(m, -b, a, 0);  /* v3 -> 2 */
    hm_add_vertex(m, 0, b, a);   /* v4 -> 3 */
    hm_add_vertex(m, 0, -b, a);  /* v5 -> 4 */
    hm_add_vertex(m, -a, 0, b);  /* v6 -> 5 */
    hm_add_vertex(m, 0, -b, -a); /* v7 -> 6 */
    hm_add_vertex(m, a, 0, -b);  /* v8 -> 7 */
    hm_add_vertex(m, a, 0, b);   /* v9 -> 8 */
    hm_add_vertex(m, -a, 0, -b); /* v10 -> 9 */
    hm_add_vertex(m, b, -a, 0);  /* v11 -> 10 */
    hm_add_vertex(m, -b, -a, 0); /* v12 -> 11 */
    hm_project(m);
    /* triangles in sph_gen.h:183-202 order, 1-based names -> 0-based ids */
    static const int tri[20][3] = {
        {3, 2, 1},  {2, 3, 4},  {6, 5, 4},  {5, 9, 4},   {8, 7, 1},
        {7, 10, 1}, {12, 11, 5}, {11, 12, 7}, {10, 6, 3}, {6, 10, 12},
        {9, 8, 2},  {8, 9, 11}, {3, 6, 4},  {9, 2, 4},   {10, 3, 1},
        {2, 8, 1},  {12, 10, 7}, {8, 11, 7}, {6, 12, 5}, {11, 9, 5}};
    for (int t = 0; t < 20; ++t) {
        int vs[3] = {tri[t][0] - 1, tri[t][1] - 1, tri[t][2] - 1};
        hm_add_face(m, vs, 3);
    }
}

/* sph_gen.h:31-57 hexahedron */
static void hm_hexahedron(hm *m)
{
    double a = 1. / sqrt(3.);
    hm_add_vertex(m, -a, -a, -a);
    hm_add_vertex(m, a, -a, -a);
    hm_add_vertex(m, a, a, -a);
    hm_add_vertex(m, -a, a, -a);
    hm_add_vertex(m, -a, -a, a);
    hm_add_vertex(m, a, -a, a);
    hm_add_vertex(m, a, a, a);
    hm_add_vertex(m, -a, a, a);
    static const int q[6][4] = {{3, 2, 1, 0}, {2, 6, 5, 1}, {5, 6, 7, 4},
                                {0, 4, 7, 3}, {3, 7, 6, 2}, {1, 5, 4, 0}};
    for (int f = 0; f < 6; ++f) hm_add_face(m, q[f], 4);
}

static or_mesh *hm_export(const hm *m)
{
    or_mesh *o = calloc(1, sizeof(or_mesh));
    o->nv = m->nv;
    o->nf = m->nf;
    o->xyz = malloc(sizeof(double) * 3 * m->nv);
    memcpy(o->xyz, m->vp, sizeof(double) * 3 * m->nv);
    o->foff = malloc(sizeof(int32_t) * (m->nf + 1));
    /* MUSCL_geometry.hpp:79-88: vertices(face) from to_vertex(halfedge(f)) */
    int total = 0;
    for (int f = 0; f < m->nf; ++f) {
        o->foff[f] = total;
        int h = m->f_he[f], hh = h;
        do { total++; h = m->h_next[h]; } while (h != hh);
    }
    o->foff[m->nf] = total;
    o->fv = malloc(sizeof(int32_t) * (total ? total : 1));
    for (int f = 0; f < m->nf; ++f) {
        int k = o->foff[f];
        int h = m->f_he[f], hh = h;
        do { o->fv[k++] = m->h_to[h]; h = m->h_next[h]; } while (h != hh);
    }
    return o;
}

or_mesh *or_uv_sphere(int n_slices, int n_stacks)
{
    if (n_slices < 3 || n_stacks < 2) return NULL;
    hm m;
    hm_init(&m);
    int v0 = hm_add_vertex(&m, 0, 1, 0);
    for (int i = 0; i < n_stacks - 1; i++) {
        double phi = M_PI * (double)(i + 1) / (double)n_stacks;
        for (int j = 0; j < n_slices; j++) {
            double theta = 2.0 * M_PI * (double)j / (double)n_slices;
            double x = sin(phi) * cos(theta);
            double y = cos(phi);
            double z = sin(phi) * sin(theta);
            hm_add_vertex(&m, x, y, z);
        }
    }
    int v1 = hm_add_vertex(&m, 0, -1, 0);
    for (int i = 0; i < n_slices; ++i) {
        int i0 = i + 1;
        int i1 = (i + 1) % n_slices + 1;
        int t0[3] = {v0, i1, i0};
        hm_add_face(&m, t0, 3);
        i0 = i + n_slices * (n_stacks - 2) + 1;
        i1 = (i + 1) % n_slices + n_slices * (n_stacks - 2) + 1;
        int t1[3] = {v1, i0, i1};
        hm_add_face(&m, t1, 3);
    }
    for (int j = 0; j < n_stacks - 2; j++) {
        int j0 = j * n_slices + 1;
        int j1 = (j + 1) * n_slices + 1;
        for (int i = 0; i < n_slices; i++) {
            int q[4] = {j0 + i, j0 + (i + 1) % n_slices, j1 + (i + 1) % n_slices, j1 + i};
            hm_add_face(&m, q, 4);
        }
    }
    or_mesh *o = hm_export(&m);
    hm_free(&m);
    return o;
}

static void hm_icosphere(hm *m, int n)
{
    hm_icosahedron(m);
    for (int i = 0; i < n; i++) {
        hm_loop_subdivision(m);
        hm_project(m);
    }
}

or_mesh *or_icosphere(int n)
{
    if (n < 0 || n > 12) return NULL;
    hm m;
    hm_init(&m);
    hm_icosphere(&m, n);
    or_mesh *o = hm_export(&m);
    hm_free(&m);
    return o;
}

or_mesh *or_icosphere_hex(int n)
{
    if (n < 0 || n > 12) return NULL;
    hm m;
    hm_init(&m);
    hm_icosphere(&m, n);
    hm_dual(&m);
    hm_project(&m);
    or_mesh *o = hm_export(&m);
    hm_free(&m);
    return o;
}

or_mesh *or_quad_sphere(int n)
{
    if (n < 0 || n > 12) return NULL;
    hm m;
    hm_init(&m);
    hm_hexahedron(&m);
    for (int i = 0; i < n; i++) {
        hm_catmull_clark(&m);
        hm_project(&m);
    }
    or_mesh *o = hm_export(&m);
    hm_free(&m);
    return o;
}

void or_mesh_free(or_mesh *m)
{
    if (!m) return;
    free(m->xyz);
    free(m->foff);
    free(m->fv);
    free(m);
}
