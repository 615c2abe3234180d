#include "run_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace splash {

namespace {

inline double norm3(double x, double y, double z) {
    double s = 0.0;
    s = s + x * x;
    s = s + y * y;
    s = s + z * z;
    return std::sqrt(s);
}

}  // namespace

void light_observers(double obs[16]) {
    // observers at 9400 pc along x, (x+z), z and -z (adiabatic.hpp:228-333)
    const double d = 9.4 * 3 * 1e19 / 15000, sq = std::sqrt(9.4 * 3 * 1e19 / 15000);
    const double o[4][3] = {{d, 0, 0}, {sq, 0, sq}, {0, 0, d}, {0, 0, -d}};
    for (int q = 0; q < 4; ++q) {
        for (int c = 0; c < 3; ++c) obs[4 * q + c] = o[q][c];
        obs[4 * q + 3] = norm3(o[q][0], o[q][1], o[q][2]);
    }
}

RunHost build_run_host(const PolyMesh& mesh, const double* U_in, int dim, int solver, double gam, double omega_ns,
                       const RunParams& params, int32_t slots, int32_t max_faces, int threads,
                       int32_t fallback_faces) {
    if (solver < kHLLE || solver > kHLLCPLUS) throw std::invalid_argument("unknown solver");
    if ((solver == kHLLE && dim != 4) || (solver != kHLLE && dim != 5))
        throw std::invalid_argument("dim must be 4 for MUSCL_HLLE and 5 for the adiabatic solvers");
    if (!U_in) throw std::invalid_argument("U_in is null");
    RunHost r;
    r.dim = dim;
    r.solver = solver;
    r.geo = build_geometry(mesh, threads);
    const Geometry& g = r.geo;
    const int32_t nf = g.nf;
    if (g.max_valence >= 32) throw std::runtime_error("face valence >= 32 is not supported");
    if (dim == 5) {
        // adiabatic.hpp:70-90 (the index of the :75 loop starts at face 0)
        double area_coeff = 0;
        const double mu = M_PI / 2;
        const double sigma = params.acc_width / 2.355;
        for (int32_t f = 0; f < nf; f++) {
            const double* c = &g.fc[3 * size_t(f)];
            const double theta = std::acos(c[2] / norm3(c[0], c[1], c[2]));
            area_coeff += g.area[f] / std::sqrt(2 * M_PI * sigma * sigma) *
                          std::exp(-(theta - mu) * (theta - mu) / (2 * sigma * sigma));
        }
        area_coeff /= 4 * M_PI;
        r.area_coeff = area_coeff;
        double tm = 0;
        for (int32_t f = 0; f < nf; f++) tm += U_in[size_t(f) * dim] * g.area[f];
        r.total_mass = tm;
    }

    PhysParams& P = r.phys;
    P.dim = dim;
    P.solver = solver;
    P.accretion_on = params.accretion_on;
    P.friction_on = params.friction_on;
    P.gam = gam;
    P.omega_ns = omega_ns;
    P.h0 = g.h0;
    P.round_diff = g.round_diff;
    P.e_acc = params.e_acc;
    P.alpha = params.alpha;
    const double fall_eff = 1;
    P.dmdt_coef = dim == 5 ? -(fall_eff * params.acc_rate * 4 * M_PI) / r.total_mass : 0.0;
    P.om_acc[0] = 0;
    P.om_acc[1] = std::sin(params.tilt_angle) * params.omega_acc_abs;
    P.om_acc[2] = std::cos(params.tilt_angle) * params.omega_acc_abs;
    const double gam3d = 1 / (2 - gam);
    P.gam_h = gam / (gam - 1);
    P.gam3d_rho = gam3d / (2 * gam3d - 1);
    P.cool_k = 12. / 5 * 1.6e-13 * std::pow(3. / 4 * 4.85e36, 0.25);  // adiabatic.hpp cooling constants
    P.f.gam = float(gam);
    P.f.inv_gm1 = float(1 / (gam - 1));
    P.f.kq = float((gam + 1) / (2 * gam));
    P.f.gm1 = float(gam - 1);
    P.f.gam_h = float(P.gam_h);
    P.f.gam3d_rho = float(P.gam3d_rho);
    P.f.cool_k = float(P.cool_k);
    P.f.e_acc = float(P.e_acc);
    P.f.alpha = float(P.alpha);
    P.f.omega_ns = float(P.omega_ns);
    P.f.dmdt_coef = float(P.dmdt_coef);
    for (int c = 0; c < 3; ++c) P.f.om_acc[c] = float(P.om_acc[c]);
    r.params = params;
    try {
        relayout(r, slots, max_faces);
    } catch (const FaceSetOverflow&) {
        if (fallback_faces <= 0 || fallback_faces >= max_faces) throw;
        relayout(r, slots, fallback_faces);
    }
    return r;
}

void relayout(RunHost& r, int32_t slots, int32_t max_faces) {
    const Geometry& g = r.geo;
    const int32_t nf = g.nf, nh = g.nh;
    const RunParams& params = r.params;
    const int dim = r.dim;
    r.L = build_layout(g, slots, max_faces);
    validate_layout(g, r.L);
    const Layout& L = r.L;

    r.he_fv.resize(nh);
    r.sten.resize(4 * size_t(nh));
    r.beta.resize(4 * size_t(nh));
    r.hsc.resize(4 * size_t(nh));
    r.ec.resize(4 * size_t(nh));
    r.en.resize(4 * size_t(nh));
    for (int32_t hn = 0; hn < nh; ++hn) {
        const int32_t h = L.he_old[hn];
        const int32_t f = g.he_face[h];
        r.he_fv[hn] = L.iperm[f] | ((g.foff[f + 1] - g.foff[f]) << 27);
        r.sten[4 * size_t(hn) + 0] = L.iperm[g.ff_plus[2 * size_t(h)]];
        r.sten[4 * size_t(hn) + 1] = L.iperm[g.ff_plus[2 * size_t(h) + 1]];
        r.sten[4 * size_t(hn) + 2] = L.iperm[g.ff_minus[2 * size_t(h)]];
        r.sten[4 * size_t(hn) + 3] = L.iperm[g.ff_minus[2 * size_t(h) + 1]];
        r.beta[4 * size_t(hn) + 0] = g.beta_plus[2 * size_t(h)];
        r.beta[4 * size_t(hn) + 1] = g.beta_plus[2 * size_t(h) + 1];
        r.beta[4 * size_t(hn) + 2] = g.beta_minus[2 * size_t(h)];
        r.beta[4 * size_t(hn) + 3] = g.beta_minus[2 * size_t(h) + 1];
        r.hsc[4 * size_t(hn) + 0] = g.h_plus[h];
        r.hsc[4 * size_t(hn) + 1] = g.h_minus[h];
        r.hsc[4 * size_t(hn) + 2] = g.bm_dist[h];
        // adiabatic.hpp:757-758: distance(v_j, v_j1) / surface_area[n_face]
        r.hsc[4 * size_t(hn) + 3] = g.arc[h] / g.area[f];
        for (int c = 0; c < 3; ++c) {
            r.ec[4 * size_t(hn) + c] = g.ec[3 * size_t(h) + c];
            r.en[4 * size_t(hn) + c] = g.en[3 * size_t(h) + c];
        }
        r.ec[4 * size_t(hn) + 3] = g.ecn[h];
        r.en[4 * size_t(hn) + 3] = g.chord[h];
    }
    r.area.resize(nf);
    r.fcn4.assign(4 * size_t(nf), 0.0);
    for (int32_t k = 0; k < nf; ++k) {
        const int32_t f = L.perm[k];
        r.area[k] = g.area[f];
        for (int c = 0; c < 3; ++c) r.fcn4[4 * size_t(k) + c] = g.fcn[3 * size_t(f) + c];
        r.fcn4[4 * size_t(k) + 3] = g.round_diff / g.area[f];
    }

    // get_light_curves (adiabatic.hpp:228-333): each observer's flux is
    // sum_f Pi_f * cos(alpha) * A_f over the faces it sees.  Which faces an
    // observer sees comes from the reference's atan2 / acos tests, so it is
    // decided here once per face (bit k for observer k); the device forms the
    // weight |r.o| / |o| * A itself (r.o summed over x, y, z in that order,
    // without contraction), the value this loop would form on the host.
    r.light_mask.assign(dim == 5 ? size_t(nf) : 0, 0);
    if (dim == 5) {
        double ob[16];
        light_observers(ob);
        const double obs[4][3] = {{ob[0], ob[1], ob[2]}, {ob[4], ob[5], ob[6]}, {ob[8], ob[9], ob[10]},
                                  {ob[12], ob[13], ob[14]}};
        for (int32_t k = 0; k < nf; ++k) {
            const int32_t f = L.perm[k];
            const double* c = &g.fc[3 * size_t(f)];
            const double n = norm3(c[0], c[1], c[2]);
            const double* u = &r.fcn4[4 * size_t(k)];
            const double phi = std::atan2(c[1] / n, c[0] / n);
            const double theta = std::acos(c[2] / n);
            uint8_t m = 0;
            if (phi < M_PI / 2 && phi > -M_PI / 2) m |= 1;
            if (obs[1][0] * u[0] + obs[1][1] * u[1] + obs[1][2] * u[2] > 0) m |= 2;
            if (theta < M_PI / 2) m |= 4;
            if (theta > M_PI / 2) m |= 8;
            r.light_mask[k] = m;
        }
    }

    // accretion profile per face (adiabatic.hpp:425-449), evaluated once with
    // the expressions the reference evaluates on every source() call
    r.acc.assign(8 * size_t(nf), 0.0);
    if (dim == 5 && params.accretion_on) {
        const double mu = M_PI / 2;
        const double sigma = params.acc_width / 2.355;
        const double tilt = params.tilt_angle;
        for (int32_t k = 0; k < nf; ++k) {
            const double* p = &r.fcn4[4 * size_t(k)];
            const double theta_acc = std::acos(p[1] * std::sin(tilt) + p[2] * std::cos(tilt));
            const double rate = 1 / r.area_coeff * params.acc_rate / std::sqrt(2 * M_PI * sigma * sigma) *
                                std::exp(-(theta_acc - mu) * (theta_acc - mu) / (2 * sigma * sigma));
            const double ox = 0, oy = std::sin(tilt) * params.omega_acc_abs, oz = std::cos(tilt) * params.omega_acc_abs;
            // omxr = omega_acc x r ; rxv = r x omxr
            const double mx = oy * p[2] - oz * p[1], my = oz * p[0] - ox * p[2], mz = ox * p[1] - oy * p[0];
            double* a = &r.acc[8 * size_t(k)];
            a[0] = rate;
            a[1] = mx;
            a[2] = my;
            a[3] = mz;
            a[4] = p[1] * mz - p[2] * my;
            a[5] = p[2] * mx - p[0] * mz;
            a[6] = p[0] * my - p[1] * mx;
        }
    }

}

MixedTables pack_mixed_tables(const RunHost& r) {
    const Layout& L = r.L;
    const int32_t nf = r.geo.nf;
    const size_t n_ext = L.ext_loc.size() / 4;
    if (n_ext % 2) throw std::runtime_error("odd slot count");
    MixedTables m;
    m.slot_a.assign(2 * n_ext, 0u);
    m.slot_b.assign(4 * n_ext, 0.f);
    m.edge32.assign(4 * n_ext, 0.f);  // n_ext/2 edges of 8 floats
    m.edge_rec.assign(n_ext / 2, 0u);
    const int32_t nb = L.n_blocks();
    std::vector<int32_t> eidx, own_of_halo;
    for (int32_t b = 0; b < nb; ++b) {
        const int32_t h0 = L.foff[L.blk_face[b]], h1 = L.foff[L.blk_face[b + 1]];
        const int32_t n_own = h1 - h0;
        const int32_t n_slots = n_own + L.halo_off[b + 1] - L.halo_off[b];
        const size_t e0 = size_t(h0) + size_t(L.halo_off[b]);
        if (n_slots % 2 || e0 % 2) throw std::runtime_error("block slots do not pair into edges");
        if (n_slots > 512) throw std::runtime_error("block has more than 512 slots");
        eidx.assign(size_t(n_slots), -1);
        own_of_halo.assign(size_t(n_slots), -1);
        auto tw_of = [&](int32_t s) { return L.twin_slot[2 * size_t(h0 + s) + 1]; };
        // the side that evaluates an edge's flux: the half-edge of lower mesh
        // index, so the choice (and the FP32 rounding) is that of the mesh,
        // whatever the layout's patch shape and face order
        auto low_of = [&](int32_t s) {  // own slots s < n_own
            const int32_t h = L.he_old[h0 + s];
            return h < r.geo.twin[h];
        };
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
            const int32_t* X = &L.ext_loc[4 * e];
            const size_t hn = size_t(X[3]);
            const uint32_t owner = uint32_t(X[0] & 0xffff), valence = uint32_t(X[0] >> 16);
            const uint32_t q[4] = {uint32_t(X[1] & 0xffff), uint32_t(X[1] >> 16), uint32_t(X[2] & 0xffff),
                                   uint32_t(X[2] >> 16)};
            if (owner > 1023 || q[0] > 1023 || q[1] > 1023 || q[2] > 1023 || q[3] > 1023 || valence > 15)
                throw std::runtime_error("slot record does not fit its 10-bit local indices / 4-bit valence");
            const bool own = s < n_own;
            const bool rep = own && (low_of(s) || tw_of(s) >= n_own);
            const uint32_t ei = uint32_t(eidx[size_t(s)]);
            if (ei > 255) throw std::runtime_error("slot record does not fit its fields");
            m.slot_a[2 * e] = q[0] | q[1] << 10 | q[2] << 20 | (rep ? 0u : 1u) << 30;
            m.slot_a[2 * e + 1] = q[3] | owner << 10 | ei << 20 | valence << 28;
            if (rep)
                m.edge_rec[e0 / 2 + ei] = uint32_t(s) | uint32_t(tw_of(s)) << 9 | (low_of(s) ? 1u : 0u) << 18;
            const double h_p = r.hsc[4 * hn + 0], h_m = r.hsc[4 * hn + 1], bm_dist = r.hsc[4 * hn + 2];
            m.slot_b[4 * e + 0] = float(h_p / h_m);
            m.slot_b[4 * e + 1] = float(bm_dist / h_p);
            m.slot_b[4 * e + 2] = float(r.beta[4 * hn + 0]);
            m.slot_b[4 * e + 3] = float(r.beta[4 * hn + 2]);
            if (rep) {
                float* row = &m.edge32[8 * (e0 / 2 + ei)];
                for (int c = 0; c < 4; ++c) {
                    row[c] = float(r.ec[4 * hn + c]);
                    row[4 + c] = float(r.en[4 * hn + c]);
                }
                // the FP32 kernel takes |ec| (the recomputed norm of the unit
                // midpoint) as exactly 1; its place holds the edge's arc
                // length, the same on both sides (one great-circle distance
                // of the same two vertices)
                if (row[3] != 1.0f) throw std::runtime_error("edge midpoint norm is not 1 in FP32 (degenerate edge)");
                const int32_t h = L.he_old[hn];
                if (r.geo.arc[h] != r.geo.arc[r.geo.twin[h]]) throw std::runtime_error("edge arc differs between its sides");
                row[3] = float(r.geo.arc[h]);
            }
        }
    }
    m.rate.resize(nf);
    m.inv_area.resize(nf);
    m.fc32.resize(4 * size_t(nf));
    for (int32_t k = 0; k < nf; ++k) {
        m.rate[k] = float(r.acc[8 * size_t(k)]);
        m.inv_area[k] = float(1.0 / r.area[k]);
        for (int c = 0; c < 3; ++c) m.fc32[4 * size_t(k) + c] = float(r.fcn4[4 * size_t(k) + c]);
        m.fc32[4 * size_t(k) + 3] = m.rate[k];
    }
    for (size_t q = 0; q + 1 < L.fset_off.size(); ++q) m.fset_max = std::max(m.fset_max, L.fset_off[q + 1] - L.fset_off[q]);
    for (int32_t b = 0; b < nb; ++b) {
        const int32_t n_slots = L.foff[L.blk_face[b + 1]] - L.foff[L.blk_face[b]] + L.halo_off[b + 1] - L.halo_off[b];
        m.slot_max = std::max(m.slot_max, (n_slots + 3) / 4 * 4);
    }
    return m;
}

}  // namespace splash
