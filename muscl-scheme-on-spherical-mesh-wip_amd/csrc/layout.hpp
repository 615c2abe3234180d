// Internal GPU layout of a mesh: faces reordered by a coordinate bisection or
// along a cube-map Morton curve of their centres, half-edges face-major in that order, and the partition of
// faces into patches ("blocks") that one workgroup advances through an RK
// stage with its half-edges plus the outside twins of its boundary
// half-edges (the halo) resident in LDS.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <vector>

#include "geometry.hpp"

namespace splash {

// A patch whose stencil faces exceed the LDS face set (Layout::kFaceSetMax):
// a smaller patch face cap lays the mesh out.
struct FaceSetOverflow : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Layout {
    int32_t slots = 0;                 // LDS half-edge slots per block
    int32_t max_faces = 0;             // faces per block (one thread per face in the update)
    std::vector<int32_t> perm, iperm;  // internal -> mesh face, mesh -> internal
    std::vector<int32_t> foff;         // [nf+1] internal face offsets
    std::vector<int32_t> he_old;       // internal -> mesh half-edge
    std::vector<int32_t> he_new;       // mesh -> internal half-edge
    std::vector<int32_t> blk_face;     // [n_blocks+1], n_blocks a multiple of 8
    std::vector<int32_t> halo_off;     // [n_blocks+1]
    std::vector<int32_t> halo_he;      // internal half-edge ids
    std::vector<int32_t> twin_slot;    // [nh][2]: internal twin, its slot in the owner's block
    // LDS face sets: block b stages the states of fset_face[fset_off[b] ..
    // fset_off[b+1]) (its own faces first, then every other face its own and
    // halo half-edges read); slot rows of block b start at foff[f0] +
    // halo_off[b] ("extended rows", own then halo, nh + halo_he.size() rows)
    static constexpr int32_t kFaceSetMax = 1024;
    std::vector<int32_t> fset_off;     // [n_blocks+1]
    std::vector<int32_t> fset_face;    // internal face ids
    std::vector<int32_t> ext_loc;      // [n_ext][4]: owner local | valence << 16, plus0 | plus1 << 16,
                                       //             minus0 | minus1 << 16, internal half-edge id
    int32_t n_blocks() const { return int32_t(blk_face.size()) - 1; }
    int32_t n_ext() const { return int32_t(ext_loc.size() / 4); }
};

// Patches of at most max_faces faces whose own half-edges plus halo fit in
// `slots` LDS slots.
Layout build_layout(const Geometry& g, int32_t slots, int32_t max_faces);

// Checks every index the stage kernel dereferences (throws std::logic_error
// on the first violation).
void validate_layout(const Geometry& g, const Layout& L);

}  // namespace splash
