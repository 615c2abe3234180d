#include "layout.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace splash {

namespace {

// spread the low 24 bits of x to every second bit (2-D Morton interleave)
uint64_t spread2(uint64_t x) {
    x &= 0xffffff;
    x = (x | x << 16) & 0x0000ffff0000ffffull;
    x = (x | x << 8) & 0x00ff00ff00ff00ffull;
    x = (x | x << 4) & 0x0f0f0f0f0f0f0f0full;
    x = (x | x << 2) & 0x3333333333333333ull;
    x = (x | x << 1) & 0x5555555555555555ull;
    return x;
}

// Cube-map Morton code of a unit face centre: the cube face it projects to
// (dominant axis and sign) in the top bits, then the 2-D Morton interleave
// of its gnomonic coordinates on that face (the other two components over
// the dominant one, in [-1, 1], at 24 bits).  Runs of this order are compact
// on the sphere, and a patch may take any number of consecutive faces.
uint64_t cube_morton(const double* c) {
    int ax = 0;
    for (int k = 1; k < 3; ++k)
        if (std::abs(c[k]) > std::abs(c[ax])) ax = k;
    const uint64_t face = uint64_t(2 * ax + (c[ax] > 0 ? 1 : 0));
    const double d = std::abs(c[ax]);
    const double top = double((1u << 24) - 1);
    auto q = [&](double x) {
        const double v = (x / d + 1.0) * 0.5 * top;
        return uint64_t(std::min(std::max(v, 0.0), top));
    };
    return face << 48 | spread2(q(c[(ax + 1) % 3])) | spread2(q(c[(ax + 2) % 3])) << 1;
}

// Recursive coordinate bisection of the unit face centres: the range
// p[lo, hi) is split along the coordinate of its largest extent, at a
// multiple of `leaf` near its middle, until each range holds at most `leaf`
// faces.  Every range starts at a multiple of `leaf`, so every leaf but the
// last is full.  A leaf is a box of bounded aspect on the sphere, whose
// boundary (a patch's halo half-edges) is shorter than that of a run of a
// space-filling curve, which follows the curve's turns: about a fifth fewer
// halo half-edges at 64 faces.  The split orders by (coordinate, face index),
// a strict total order, and each leaf is sorted by face index, so the
// permutation does not depend on the standard library's algorithms.
void bisect(std::vector<int32_t>& p, size_t lo, size_t hi, const std::vector<double>& c, size_t leaf) {
    const size_t m = hi - lo;
    if (m <= leaf) {
        std::sort(p.begin() + lo, p.begin() + hi);
        return;
    }
    double mn[3] = {c[3 * size_t(p[lo])], c[3 * size_t(p[lo]) + 1], c[3 * size_t(p[lo]) + 2]};
    double mx[3] = {mn[0], mn[1], mn[2]};
    for (size_t k = lo + 1; k < hi; ++k)
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], c[3 * size_t(p[k]) + a]);
            mx[a] = std::max(mx[a], c[3 * size_t(p[k]) + a]);
        }
    int ax = 0;
    for (int a = 1; a < 3; ++a)
        if (mx[a] - mn[a] > mx[ax] - mn[ax]) ax = a;
    // leaf <= half <= m/2 + leaf/2 < m
    const size_t half = leaf * std::max<size_t>(1, (m + leaf) / (2 * leaf));
    std::nth_element(p.begin() + lo, p.begin() + lo + half, p.begin() + hi, [&](int32_t x, int32_t y) {
        const double cx = c[3 * size_t(x) + ax], cy = c[3 * size_t(y) + ax];
        return cx < cy || (cx == cy && x < y);
    });
    // the halves are disjoint ranges of p: large ones split into tasks
#pragma omp task default(shared) if (half > 65536)
    bisect(p, lo, lo + half, c, leaf);
    bisect(p, lo + half, hi, c, leaf);
#pragma omp taskwait
}


// Greedy patches over the face order perm (iperm its inverse): add faces
// while own half-edges plus halo fit the slots and the face count stays
// within max_faces; with leaf > 0 a patch also never spans two runs of
// `leaf` faces (a bisection leaf whose halo overflows the slots is split).
// Returns the patches' first faces and the end, nf.
std::vector<int32_t> pack(const Geometry& g, const std::vector<int32_t>& perm, const std::vector<int32_t>& iperm,
                          int32_t slots, int32_t max_faces, int32_t leaf) {
    const int32_t nf = g.nf;
    std::vector<int32_t> blk{0};
    int32_t own = 0, halo = 0;
    for (int32_t k = 0; k < nf; ++k) {
        const int32_t f = perm[k];
        const int32_t val = g.foff[f + 1] - g.foff[f];
        if (val > slots) throw std::runtime_error("face valence exceeds the block's LDS slots");
        int32_t dh = 0;
        const int32_t f0 = blk.back();
        for (int32_t h = g.foff[f]; h < g.foff[f + 1]; ++h) {
            const int32_t o = iperm[g.he_face[g.twin[h]]];
            dh += (o >= f0 && o < k) ? -1 : 1;
        }
        if (own + val + halo + dh > slots || k - f0 >= max_faces || (leaf > 0 && k > f0 && k % leaf == 0)) {
            blk.push_back(k);
            own = 0;
            halo = 0;
            dh = val;
        }
        own += val;
        halo += dh;
    }
    blk.push_back(nf);
    return blk;
}

std::vector<int32_t> inverse(const std::vector<int32_t>& perm) {
    std::vector<int32_t> inv(perm.size());
    for (size_t k = 0; k < perm.size(); ++k) inv[size_t(perm[k])] = int32_t(k);
    return inv;
}

void require(bool ok, const char* what) {
    if (!ok) throw std::logic_error(std::string("layout check failed: ") + what);
}

}  // namespace

Layout build_layout(const Geometry& g, int32_t slots, int32_t max_faces) {
    const int32_t nf = g.nf, nh = g.nh;
    Layout L;
    L.slots = slots;
    L.max_faces = max_faces;
    if (max_faces < 1) throw std::runtime_error("patch face cap must be positive");
    for (size_t k = 0; k < 3 * size_t(nf); ++k)
        if (!std::isfinite(g.fcn[k])) throw std::runtime_error("a face centre is not finite (degenerate face)");
    // Two face orders; the layout keeps the one with fewer patches, the
    // bisection on a tie.  Bisection leaves of max_faces faces have the
    // shortest boundaries and reach the fewest patches, ceil(nf / max_faces),
    // when max_faces faces and their halo fit the slots (the FP32-flux
    // layout).  Where the slots bind first (the FP64 layout: 256 slots,
    // 256 faces) runs of the cube-map curve are filled up to the slots, which
    // fixed-size leaves cannot be.
    // (Leaves whose own half-edges alone exceed the slots on average cannot
    // reach the minimum: the bisection is not tried then.)
    const bool try_bisection = int64_t(max_faces) * nh <= int64_t(slots) * nf;
    if (try_bisection) {
        L.perm.resize(nf);
        for (int32_t f = 0; f < nf; ++f) L.perm[f] = f;
#pragma omp parallel
#pragma omp single
        bisect(L.perm, 0, size_t(nf), g.fcn, size_t(max_faces));
        L.iperm = inverse(L.perm);
        L.blk_face = pack(g, L.perm, L.iperm, slots, max_faces, max_faces);
    }
    if (!try_bisection || int64_t(L.blk_face.size()) - 1 > (int64_t(nf) + max_faces - 1) / max_faces) {
        std::vector<uint64_t> code(nf);
        for (int32_t f = 0; f < nf; ++f) code[f] = cube_morton(&g.fcn[3 * size_t(f)]);
        std::vector<int32_t> perm(nf);
        for (int32_t f = 0; f < nf; ++f) perm[f] = f;
        std::stable_sort(perm.begin(), perm.end(), [&](int32_t x, int32_t y) { return code[x] < code[y]; });
        std::vector<int32_t> iperm = inverse(perm);
        std::vector<int32_t> blk = pack(g, perm, iperm, slots, max_faces, 0);
        if (!try_bisection || blk.size() < L.blk_face.size()) {
            L.perm = std::move(perm);
            L.iperm = std::move(iperm);
            L.blk_face = std::move(blk);
        }
    }
    while ((L.blk_face.size() - 1) % 8) L.blk_face.push_back(nf);

    L.foff.assign(nf + 1, 0);
    for (int32_t k = 0; k < nf; ++k) L.foff[k + 1] = L.foff[k] + (g.foff[L.perm[k] + 1] - g.foff[L.perm[k]]);
    L.he_old.resize(nh);
    L.he_new.resize(nh);
    for (int32_t k = 0; k < nf; ++k) {
        const int32_t f = L.perm[k];
        for (int32_t j = 0; j < g.foff[f + 1] - g.foff[f]; ++j) {
            L.he_old[L.foff[k] + j] = g.foff[f] + j;
            L.he_new[g.foff[f] + j] = L.foff[k] + j;
        }
    }
    auto twin_owner = [&](int32_t hn) { return L.iperm[g.he_face[g.twin[L.he_old[hn]]]]; };

    // halo (the outside twins of each block's boundary half-edges, in the
    // block's half-edge order) and the twin slots; the blocks are
    // independent, so one parallel pass counts, a prefix sum places each
    // block's halo, and a second parallel pass fills it
    const int32_t nb = L.n_blocks();
    L.twin_slot.assign(2 * size_t(nh), 0);
    L.halo_off.assign(size_t(nb) + 1, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t b = 0; b < nb; ++b) {
        const int32_t f0 = L.blk_face[b], f1 = L.blk_face[b + 1];
        int32_t n_halo = 0;
        for (int32_t hn = L.foff[f0]; hn < L.foff[f1]; ++hn) {
            const int32_t o = twin_owner(hn);
            n_halo += (o < f0 || o >= f1) ? 1 : 0;
        }
        L.halo_off[size_t(b) + 1] = n_halo;
    }
    for (int32_t b = 0; b < nb; ++b) L.halo_off[size_t(b) + 1] += L.halo_off[size_t(b)];
    L.halo_he.assign(size_t(L.halo_off[size_t(nb)]), 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t b = 0; b < nb; ++b) {
        const int32_t f0 = L.blk_face[b], f1 = L.blk_face[b + 1];
        const int32_t h0 = L.foff[f0], h1 = L.foff[f1];
        int32_t n_halo = 0;
        for (int32_t hn = h0; hn < h1; ++hn) {
            const int32_t tw = L.he_new[g.twin[L.he_old[hn]]];
            const int32_t o = twin_owner(hn);
            L.twin_slot[2 * size_t(hn)] = tw;
            if (o >= f0 && o < f1) {
                L.twin_slot[2 * size_t(hn) + 1] = tw - h0;
            } else {
                L.twin_slot[2 * size_t(hn) + 1] = (h1 - h0) + n_halo;
                L.halo_he[size_t(L.halo_off[size_t(b)] + n_halo++)] = tw;
            }
        }
    }

    // face sets and extended slot rows: each block gathers its set in its
    // own list (own faces first, then stencil faces in slot order), with a
    // per-thread face -> local index map that it clears again; the lists
    // are then concatenated in block order
    const int32_t n_ext = nh + int32_t(L.halo_he.size());
    L.ext_loc.assign(4 * size_t(n_ext), 0);
    std::vector<std::vector<int32_t>> sets(static_cast<size_t>(nb));
    int too_big = 0;
#pragma omp parallel reduction(+ : too_big)
    {
        std::vector<int32_t> local(nf, -1);
#pragma omp for schedule(dynamic, 64)
        for (int32_t b = 0; b < nb; ++b) {
            const int32_t f0 = L.blk_face[b], f1 = L.blk_face[b + 1];
            const int32_t h0 = L.foff[f0], h1 = L.foff[f1];
            const int32_t n_own = h1 - h0, n_halo = L.halo_off[b + 1] - L.halo_off[b];
            std::vector<int32_t>& fs = sets[size_t(b)];
            auto loc = [&](int32_t f) {
                if (local[f] < 0) {
                    local[f] = int32_t(fs.size());
                    fs.push_back(f);
                }
                return local[f];
            };
            for (int32_t f = f0; f < f1; ++f) loc(f);
            for (int32_t slot = 0; slot < n_own + n_halo; ++slot) {
                const int32_t he = slot < n_own ? h0 + slot : L.halo_he[L.halo_off[b] + slot - n_own];
                const int32_t h = L.he_old[he];
                const int32_t owner = L.iperm[g.he_face[h]];
                const int32_t val = g.foff[g.he_face[h] + 1] - g.foff[g.he_face[h]];
                int32_t* row = &L.ext_loc[4 * size_t(h0 + L.halo_off[b] + slot)];
                row[0] = loc(owner) | (val << 16);
                row[1] = loc(L.iperm[g.ff_plus[2 * size_t(h)]]) | (loc(L.iperm[g.ff_plus[2 * size_t(h) + 1]]) << 16);
                row[2] = loc(L.iperm[g.ff_minus[2 * size_t(h)]]) | (loc(L.iperm[g.ff_minus[2 * size_t(h) + 1]]) << 16);
                row[3] = he;
            }
            if (int32_t(fs.size()) > Layout::kFaceSetMax) ++too_big;
            for (int32_t f : fs) local[f] = -1;
        }
    }
    if (too_big) throw FaceSetOverflow("a block's stencil face set exceeds the LDS face slots");
    L.fset_off.assign(size_t(nb) + 1, 0);
    for (int32_t b = 0; b < nb; ++b) L.fset_off[size_t(b) + 1] = L.fset_off[size_t(b)] + int32_t(sets[size_t(b)].size());
    L.fset_face.assign(size_t(L.fset_off[size_t(nb)]), 0);
#pragma omp parallel for schedule(static)
    for (int32_t b = 0; b < nb; ++b)
        std::copy(sets[size_t(b)].begin(), sets[size_t(b)].end(), L.fset_face.begin() + L.fset_off[size_t(b)]);
    return L;
}

void validate_layout(const Geometry& g, const Layout& L) {
    const int32_t nf = g.nf, nh = g.nh, nb = L.n_blocks();
    require(int32_t(L.perm.size()) == nf && int32_t(L.foff.size()) == nf + 1, "sizes");
    std::vector<char> seen(nf, 0);
    for (int32_t k = 0; k < nf; ++k) {
        require(L.perm[k] >= 0 && L.perm[k] < nf && !seen[L.perm[k]], "perm is a permutation");
        seen[L.perm[k]] = 1;
        require(L.iperm[L.perm[k]] == k, "iperm inverts perm");
    }
    require(L.foff[0] == 0 && L.foff[nf] == nh, "foff covers the half-edges");
    for (int32_t hn = 0; hn < nh; ++hn) require(L.he_new[L.he_old[hn]] == hn, "half-edge maps are inverse");
    require(nb % 8 == 0 && L.blk_face[0] == 0 && L.blk_face[nb] == nf, "blocks cover the faces");
    require(int32_t(L.halo_off.size()) == nb + 1 && L.halo_off[nb] == int32_t(L.halo_he.size()), "halo offsets");
    require(int32_t(L.fset_off.size()) == nb + 1 && L.n_ext() == nh + int32_t(L.halo_he.size()), "face sets");
    // the per-block checks are independent: they run in parallel, and the
    // first failure (in no particular order) is rethrown after the loop
    std::string failure;
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t b = 0; b < nb; ++b) {
        try {
            const int32_t f0 = L.blk_face[b], f1 = L.blk_face[b + 1];
            require(f0 <= f1 && f1 - f0 <= L.max_faces, "faces per block within the limit");
            const int32_t h0 = L.foff[f0], h1 = L.foff[f1];
            const int32_t n_own = h1 - h0, n_halo = L.halo_off[b + 1] - L.halo_off[b];
            require(n_halo >= 0 && n_own + n_halo <= L.slots, "own + halo half-edges fit the LDS slots");
            for (int32_t q = L.halo_off[b]; q < L.halo_off[b + 1]; ++q) {
                const int32_t he = L.halo_he[q];
                require(he >= 0 && he < nh && (he < h0 || he >= h1), "halo half-edges are outside the block");
            }
            // one-sided fluxes of the mixed kernel: an own half-edge evaluates its
            // edge when it is the side of lower mesh index or its twin is in the
            // halo, and also fills its twin's slot when that twin is own; every
            // own slot must be written exactly once
            std::vector<int32_t> filled(size_t(n_own), 0);
            for (int32_t hn = h0; hn < h1; ++hn) {
                const int32_t tw = L.twin_slot[2 * size_t(hn)];
                const bool twin_own = tw >= h0 && tw < h1;
                const bool low = L.he_old[hn] < L.he_old[tw];
                if (low || !twin_own) filled[size_t(hn - h0)]++;
                if (low && twin_own) filled[size_t(tw - h0)]++;
            }
            for (int32_t q = 0; q < n_own; ++q) require(filled[size_t(q)] == 1, "each own flux written once");
            // face set: own faces first, every local index in range and pointing
            // at the face the geometry names
            const int32_t fs0 = L.fset_off[b], n_fs = L.fset_off[b + 1] - fs0;
            require(n_fs >= f1 - f0 && n_fs <= Layout::kFaceSetMax, "face set size");
            for (int32_t f = f0; f < f1; ++f) require(L.fset_face[fs0 + f - f0] == f, "own faces lead the face set");
            for (int32_t slot = 0; slot < n_own + n_halo; ++slot) {
                const int32_t* row = &L.ext_loc[4 * size_t(h0 + L.halo_off[b] + slot)];
                const int32_t he = slot < n_own ? h0 + slot : L.halo_he[L.halo_off[b] + slot - n_own];
                require(row[3] == he, "extended row names its half-edge");
                const int32_t h = L.he_old[he];
                const int32_t want[5] = {L.iperm[g.he_face[h]], L.iperm[g.ff_plus[2 * size_t(h)]],
                                         L.iperm[g.ff_plus[2 * size_t(h) + 1]], L.iperm[g.ff_minus[2 * size_t(h)]],
                                         L.iperm[g.ff_minus[2 * size_t(h) + 1]]};
                const int32_t got[5] = {row[0] & 0xffff, row[1] & 0xffff, row[1] >> 16, row[2] & 0xffff, row[2] >> 16};
                for (int q = 0; q < 5; ++q) {
                    require(got[q] >= 0 && got[q] < n_fs, "local face index in range");
                    require(L.fset_face[fs0 + got[q]] == want[q], "local face index names the stencil face");
                }
                require((row[0] >> 16) == g.foff[g.he_face[h] + 1] - g.foff[g.he_face[h]], "valence");
            }
            for (int32_t hn = h0; hn < h1; ++hn) {
                const int32_t tw = L.twin_slot[2 * size_t(hn)], slot = L.twin_slot[2 * size_t(hn) + 1];
                require(tw >= 0 && tw < nh && L.twin_slot[2 * size_t(tw)] == hn, "twin is an involution");
                require(slot >= 0 && slot < n_own + n_halo, "twin slot in range");
                const int32_t at = slot < n_own ? h0 + slot : L.halo_he[L.halo_off[b] + slot - n_own];
                require(at == tw, "twin slot holds the twin");
            }
        } catch (const std::exception& e) {
#pragma omp critical(validate_layout_failure)
            if (failure.empty()) failure = e.what();
        }
    }
    if (!failure.empty()) throw std::logic_error(failure);
}

}  // namespace splash
