// Stencil-table construction (see geometry.hpp).  Compiled with
// -ffp-contract=off; the per-face loops are independent and run under
// OpenMP, so the tables do not depend on the thread count.
#include "geometry.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace splash {

namespace {

struct V3 {
    double x, y, z;
};
inline V3 ld(const double* p) { return {p[0], p[1], p[2]}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
// vector3d::norm (vec3d.hpp:77-85): left fold from 0
inline double norm(V3 a) {
    double s = 0.0;
    s = s + a.x * a.x;
    s = s + a.y * a.y;
    s = s + a.z * a.z;
    return std::sqrt(s);
}
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline void st(double* p, V3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
// MUSCL_geometry.hpp:981-987
inline double gc_distance(V3 a, V3 b) {
    a = a / norm(a);
    b = b / norm(b);
    return 2 * std::asin(norm(a - b) / 2);
}
// MUSCL_geometry.hpp:846-870
V3 line_plane(V3 xl, V3 nl, V3 xs, V3 vs1, V3 vs2) {
    const double A = vs1.y * vs2.z - vs1.z * vs2.y;
    const double B = vs1.z * vs2.x - vs1.x * vs2.z;
    const double C = vs1.x * vs2.y - vs1.y * vs2.x;
    const double D = -xs.x * A - xs.y * B - xs.z * C;
    const double t = -(D + A * xl.x + B * xl.y + C * xl.z) / (A * nl.x + B * nl.y + C * nl.z);
    return xl + nl * t;
}

}  // namespace

Geometry build_geometry(const PolyMesh& mesh, int threads) {
    Geometry g;
    const int32_t nf = mesh.n_faces(), nh = mesh.n_halfedges(), nv = mesh.n_vertices();
    if (nf < 4) throw std::runtime_error("mesh has fewer than 4 faces");
    g.nf = nf;
    g.nh = nh;
    g.nv = nv;
    g.foff = mesh.foff;
    const auto& X = mesh.xyz;
    const auto& FV = mesh.fv;
    auto vert = [&](int32_t v) { return ld(&X[3 * size_t(v)]); };
    auto val = [&](int32_t f) { return mesh.foff[f + 1] - mesh.foff[f]; };
    if (threads <= 0) threads = omp_get_max_threads();

    // vertex -> incident faces (CSR); the reference sorts and de-duplicates
    // the gathered list (:91-118), so only the set of faces matters.
    std::vector<int32_t> vf_off(nv + 1, 0), vf(nh);
    for (int32_t k = 0; k < nh; ++k) vf_off[FV[k] + 1]++;
    for (int32_t v = 0; v < nv; ++v) vf_off[v + 1] += vf_off[v];
    {
        std::vector<int32_t> fill(vf_off.begin(), vf_off.end() - 1);
        for (int32_t f = 0; f < nf; ++f)
            for (int32_t k = mesh.foff[f]; k < mesh.foff[f + 1]; ++k) vf[fill[FV[k]]++] = f;
    }

    // vertex neighbours (sorted, unique, without the face itself) as one
    // CSR array: each thread gathers a face's list into its own scratch
    // buffer, once to count and once to fill, so the loops allocate nothing
    // per face
    auto gather_nbs = [&](int32_t f, std::vector<int32_t>& list) {
        list.clear();
        for (int32_t k = mesh.foff[f]; k < mesh.foff[f + 1]; ++k)
            for (int32_t q = vf_off[FV[k]]; q < vf_off[FV[k] + 1]; ++q) list.push_back(vf[q]);
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
        list.erase(std::remove(list.begin(), list.end(), f), list.end());
    };
    std::vector<int32_t> nb_off(nf + 1, 0), nb_list;
#pragma omp parallel num_threads(threads)
    {
        std::vector<int32_t> list;
        list.reserve(64);
#pragma omp for schedule(static)
        for (int32_t f = 0; f < nf; ++f) {
            gather_nbs(f, list);
            nb_off[f + 1] = int32_t(list.size());
        }
#pragma omp single
        {
            for (int32_t f = 0; f < nf; ++f) nb_off[f + 1] += nb_off[f];
            nb_list.resize(size_t(nb_off[nf]));
        }
#pragma omp for schedule(static)
        for (int32_t f = 0; f < nf; ++f) {
            gather_nbs(f, list);
            std::copy(list.begin(), list.end(), nb_list.begin() + nb_off[f]);
        }
    }
    auto nbs_begin = [&](int32_t f) { return nb_list.data() + nb_off[f]; };
    auto nbs_end = [&](int32_t f) { return nb_list.data() + nb_off[f + 1]; };

    g.he_face.resize(nh);
    g.nb_edge.resize(nh);
    g.twin.resize(nh);
    int bad = 0;
    // :145-174 edge neighbour = the vertex neighbour holding both edge ends
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : bad)
    for (int32_t f = 0; f < nf; ++f) {
        const int32_t n = val(f);
        for (int32_t j = 0; j < n; ++j) {
            const int32_t a = FV[mesh.foff[f] + j], b = FV[mesh.foff[f] + (j + 1 == n ? 0 : j + 1)];
            int32_t found = -1, count = 0;
            for (const int32_t* po = nbs_begin(f); po != nbs_end(f); ++po) {
                const int32_t o = *po;
                const auto first = FV.begin() + mesh.foff[o], last = FV.begin() + mesh.foff[o + 1];
                if (std::find(first, last, a) != last && std::find(first, last, b) != last) {
                    if (found < 0) found = o;
                    ++count;
                }
            }
            if (count != 1) ++bad;
            g.he_face[mesh.foff[f] + j] = f;
            g.nb_edge[mesh.foff[f] + j] = found;
        }
    }
    if (bad) throw std::runtime_error("mesh is not a closed manifold: " + std::to_string(bad) + " edges without a unique neighbour");
    // MUSCL_base.hpp:383-384 reverse slot j0 (first match in neighbors_edge[nb])
    int asym = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : asym)
    for (int32_t f = 0; f < nf; ++f)
        for (int32_t h = mesh.foff[f]; h < mesh.foff[f + 1]; ++h) {
            const int32_t o = g.nb_edge[h];
            int32_t j0 = -1;
            for (int32_t q = mesh.foff[o]; q < mesh.foff[o + 1]; ++q)
                if (g.nb_edge[q] == f) {
                    j0 = q;
                    break;
                }
            if (j0 < 0) ++asym;
            g.twin[h] = j0;
        }
    if (asym) throw std::runtime_error("edge-neighbour relation is not symmetric");

    g.fc.resize(3 * size_t(nf));
    g.normal.resize(3 * size_t(nf));
    g.area.resize(nf);
    g.fcn.resize(3 * size_t(nf));
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int32_t f = 0; f < nf; ++f) {
        const int32_t* F = &FV[mesh.foff[f]];
        const int32_t n = val(f);
        // :176-187
        V3 c{0, 0, 0};
        for (int32_t k = 0; k < n; ++k) c = c + vert(F[k]);
        c = c / double(n);
        st(&g.fc[3 * size_t(f)], c);
        st(&g.fcn[3 * size_t(f)], c / norm(c));
        // :189-218
        V3 s{0, 0, 0};
        for (int32_t j = 0; j < n; ++j) {
            V3 r1, r2;
            if (j + 2 == n) {
                r2 = vert(F[0]) - vert(F[j + 1]);
                r1 = vert(F[j + 1]) - vert(F[j]);
            } else if (j + 1 == n) {
                r1 = vert(F[0]) - vert(F[j]);
                r2 = vert(F[1]) - vert(F[0]);
            } else {
                r2 = vert(F[j + 2]) - vert(F[j + 1]);
                r1 = vert(F[j + 1]) - vert(F[j]);
            }
            s = s + cross(r1, r2);
        }
        st(&g.normal[3 * size_t(f)], s / norm(s));
        // :989-1030 spherical excess of the (centre, v_j, v_j1) triangles
        const V3 cu = c / norm(c);
        double S = 0;
        for (int32_t j = 0; j < n; ++j) {
            const int32_t j1 = (j != n - 1) ? j + 1 : 0;
            const double a = gc_distance(cu, vert(F[j]));
            const double b = gc_distance(cu, vert(F[j1]));
            const double e = gc_distance(vert(F[j]), vert(F[j1]));
            const double A = std::acos((std::cos(a) - std::cos(b) * std::cos(e)) / (std::sin(b) * std::sin(e)));
            const double B = std::acos((std::cos(b) - std::cos(a) * std::cos(e)) / (std::sin(a) * std::sin(e)));
            const double C = std::acos((std::cos(e) - std::cos(b) * std::cos(a)) / (std::sin(b) * std::sin(a)));
            S += A + B + C - M_PI;
        }
        g.area[f] = S;
    }

    g.ff_plus.resize(2 * size_t(nh));
    g.ff_minus.resize(2 * size_t(nh));
    g.beta_plus.resize(2 * size_t(nh));
    g.beta_minus.resize(2 * size_t(nh));
    g.h_plus.resize(nh);
    g.h_minus.resize(nh);
    g.bm_dist.resize(nh);
    g.en.resize(3 * size_t(nh));
    g.ec.resize(3 * size_t(nh));
    g.ecn.resize(nh);
    g.ev.resize(3 * size_t(nh));
    g.chord.resize(nh);
    g.arc.resize(nh);
    int missing = 0;
    // :227-575 process_mesh
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256) reduction(+ : missing)
    for (int32_t f = 0; f < nf; ++f) {
        const int32_t* F = &FV[mesh.foff[f]];
        const int32_t n = val(f);
        const V3 c = ld(&g.fc[3 * size_t(f)]);
        const V3 nrm = ld(&g.normal[3 * size_t(f)]);
        for (int32_t i = 0; i < n; ++i) {
            const int32_t h = mesh.foff[f] + i;
            const V3 vi = vert(F[i]), vi1 = vert(F[i == n - 1 ? 0 : i + 1]);
            const V3 BM = (vi + vi1) / 2 - c;
            const V3 r = vi - vi1;
            V3 r1 = vi + vi1;
            r1 = r1 / norm(r1);
            V3 e = cross(r1, r);
            e = e / norm(e);
            if (norm(BM + e * 0.01) < norm(BM)) e = e * -1;
            st(&g.en[3 * size_t(h)], e);
            g.bm_dist[h] = gc_distance(c, r1);

            // first pass: most aligned neighbour on either side (strict '>')
            double best_l = -1, best_r = -1;
            int32_t l1 = -1, rr1 = -1;
            const double nBM = norm(BM);
            for (const int32_t* po = nbs_begin(f); po != nbs_end(f); ++po) {
                const int32_t o = *po;
                const V3 fo = ld(&g.fc[3 * size_t(o)]);
                V3 d = c - fo;
                V3 dp = d - nrm * dot(d, nrm);
                double cs = dot(BM, dp) / (nBM * norm(dp));
                if (cs > best_l) {
                    best_l = cs;
                    l1 = o;
                }
                d = fo - c;
                dp = d - nrm * dot(d, nrm);
                cs = dot(BM, dp) / (nBM * norm(dp));
                if (cs > best_r) {
                    best_r = cs;
                    rr1 = o;
                }
            }
            if (l1 < 0 || rr1 < 0) {
                ++missing;
                continue;
            }
            // second pass: best neighbour on the other side of the centre line
            best_l = -1;
            best_r = -1;
            int32_t l2 = l1, rr2 = rr1;
            const V3 B_left = c - ld(&g.fc[3 * size_t(l1)]);
            const V3 B_right = ld(&g.fc[3 * size_t(rr1)]) - c;
            const V3 B_right_face = B_right - nrm * dot(B_right, nrm);
            const V3 side = cross(c, r1 / 2);
            for (const int32_t* po = nbs_begin(f); po != nbs_end(f); ++po) {
                const int32_t o = *po;
                const V3 fo = ld(&g.fc[3 * size_t(o)]);
                V3 d = c - fo;
                V3 dp = d - nrm * dot(d, nrm);
                double cs = dot(BM, dp) / (nBM * norm(dp));
                double check = dot(side, d) * dot(side, B_left);
                if ((cs > best_l) && (o != l1) && (check <= 0)) {
                    best_l = cs;
                    l2 = o;
                }
                d = fo - c;
                dp = d - nrm * dot(d, nrm);
                cs = dot(BM, dp) / (nBM * norm(dp));
                check = dot(nrm, cross(dp, BM)) * dot(nrm, cross(B_right_face, BM));
                if ((cs > best_r) && (o != rr1) && (check <= 0)) {
                    best_r = cs;
                    rr2 = o;
                }
            }
            const V3 fl1 = ld(&g.fc[3 * size_t(l1)]), fl2 = ld(&g.fc[3 * size_t(l2)]);
            const V3 fr1 = ld(&g.fc[3 * size_t(rr1)]), fr2 = ld(&g.fc[3 * size_t(rr2)]);
            V3 p3 = line_plane(fl1, fl2 - fl1, c, BM, c);
            const V3 Hm = p3 / norm(p3);
            p3 = line_plane(fr1, fr2 - fr1, c, BM, c);
            const V3 Hp = p3 / norm(p3);

            const double dl = gc_distance(fl1, fl2);
            g.h_minus[h] = gc_distance(c, Hm);
            g.ff_minus[2 * size_t(h)] = l1;
            g.ff_minus[2 * size_t(h) + 1] = l2;
            g.beta_minus[2 * size_t(h)] = gc_distance(fl2, Hm) / dl;
            g.beta_minus[2 * size_t(h) + 1] = gc_distance(fl1, Hm) / dl;

            const double dr = gc_distance(fr1, fr2);
            g.h_plus[h] = gc_distance(c, Hp);
            g.ff_plus[2 * size_t(h)] = rr1;
            g.ff_plus[2 * size_t(h) + 1] = rr2;
            g.beta_plus[2 * size_t(h)] = gc_distance(fr2, Hp) / dr;
            g.beta_plus[2 * size_t(h) + 1] = gc_distance(fr1, Hp) / dr;

            // constants the flux/limiter routines rebuild on every call
            V3 m = (vi + vi1) / 2.;
            m = m / norm(m);
            st(&g.ec[3 * size_t(h)], m);
            g.ecn[h] = norm(m);
            g.chord[h] = norm(vi - vi1);
            st(&g.ev[3 * size_t(h)], vi - vi1);
            g.arc[h] = gc_distance(vi, vi1);
        }
    }
    if (missing) throw std::runtime_error("stencil search found no neighbour for " + std::to_string(missing) + " half-edges");

    // shortest edge arc (MUSCL_base.hpp:63-91): a minimum, so the thread
    // count cannot change it
    double h0 = 10;
    int32_t maxval = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(min : h0) reduction(max : maxval)
    for (int32_t f = 0; f < nf; ++f) {
        const int32_t* F = &FV[mesh.foff[f]];
        const int32_t n = val(f);
        maxval = std::max(maxval, n);
        for (int32_t j = 0; j < n; ++j) {
            const double d = gc_distance(vert(F[j]), vert(F[j < n - 1 ? j + 1 : 0]));
            if (d < h0) h0 = d;
        }
    }
    g.h0 = h0;
    g.max_valence = maxval;
    const V3 a = vert(FV[mesh.foff[0] + 1]), b = vert(FV[mesh.foff[0] + 2]);
    g.round_diff = gc_distance(a, b) / norm(a - b);
    return g;
}

}  // namespace splash
