// Internal interface of the symmetric self-potential kernel (sym.hip).
#pragma once
#include <vector>

#include "halma_common.h"

namespace halma {

constexpr int kSymK = 4;  // tiles of 128 members per block

// The partial slots of a halo take (nb + 1) * n * 8 bytes when its block pairs run in one
// launch, quadratic in n: the unbinding keeps haloes above this size on the chunked
// kernel (their plan has nb = 0).
constexpr int kSymMaxMembers = 65536;

// Band width (blocks of 512 members) of the banded plans: a halo of more blocks runs its
// block pairs (A, B) in tiles of bands (P, Q), one launch per tile, each tile's slots
// merged into an FP64 sum per member before the next, so that a member needs at most
// kSymBand + 1 slots.
constexpr int kSymBand = 64;

// blocks of a halo of n members
__host__ __device__ inline int sym_blocks(int n) {
    const int nt = (n + 127) / 128;
    return (nt + kSymK - 1) / kSymK;
}

// Per-halo descriptor: members src4[off, off + n); partial slot s of member i at
// part[part_off + s * n + i].  wb is the band width in blocks (> nb: one band).  The test
// side of item (A, B) writes slot B mod wb for the members of block A; its source side
// writes slot A mod wb for the members of block B when A < B, and slot min(wb, nb) (the
// diagonal slot) when A = B.  With one band the slots are 0..nb (n, nb as planned; a
// shrinking live set uses its own nb and the planned stride).
struct __attribute__((aligned(8))) SymDesc {
    int64_t off, part_off;
    int32_t n, nb, wb, pad;
};

struct SymPlan {
    std::vector<SymDesc> halos;
    std::vector<int32_t> order; // the haloes by size, largest first
    int64_t n_items = 0, n_members = 0, n_part = 0;
    // with items: launch t runs items [launch_off[t], launch_off[t + 1]) (tile t of every
    // halo that has one; tile t = (P, Q) with t = Q (Q + 1) / 2 + P) and then merges the
    // member segments [seg_off[t], seg_off[t + 1]) of segs: {halo, first member, member
    // count, role << 24 | local slots}; role 0: the test side of band P (off-diagonal
    // tile), 1: the source side of band Q, 2: a diagonal tile (and its diagonal slot).
    // The items themselves ({halo, block A, block B, 0}, A <= B, B-major within a tile,
    // largest haloes first within a launch) are written on the device from tiles:
    // {halo, P, Q, index of the tile's first item}, O(haloes x tiles) on the host.
    std::vector<int64_t> launch_off, seg_off;
    std::vector<int4> segs, tiles;
};

// Plans from HOST member offsets (H+1 entries); the item list only with with_items.
// With max_members != 0, haloes above that size (every halo when it is negative) get no
// blocks (nb = 0, no slots).  band > 0 bands haloes of more than band blocks (items then
// listed tile by tile, see halma_pot_sum_self_batched); 0 keeps every halo in one band.
int plan_sym(int64_t H, const int64_t* off, SymPlan& plan, bool with_items,
             int max_members = 0, int band = 0);

// Writes the items of plan.tiles (n_tiles, on the device) into items.
int expand_sym_items(const int4* d_tiles, int64_t n_tiles, const SymDesc* d_halos, int4* items,
                     hipStream_t st);

// Persistent launch over the items (count plan.n_items, or *n_items_dev when given);
// live_n (device, may be null) overrides each halo's member count.  head must hold 0
// at launch unless reset_head.
int launch_pot_sym(const SymPlan& plan, const float4* src4, const SymDesc* d_halos,
                   const int* live_n, const int4* d_items, const int64_t* n_items_dev,
                   float2* part, int* head, hipStream_t st, bool reset_head);

}  // namespace halma
