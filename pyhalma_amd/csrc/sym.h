// Internal interface of the symmetric self-potential kernel (sym.hip).
#pragma once
#include <vector>

#include "halma_common.h"

namespace halma {

constexpr int kSymK = 4;  // tiles of 128 members per block

// The partial slots of a halo take (nb + 1) * n * 8 bytes, quadratic in n: the
// unbinding keeps haloes above this size on the chunked kernel (their plan has nb = 0).
constexpr int kSymMaxMembers = 65536;

// blocks of a halo of n members
__host__ __device__ inline int sym_blocks(int n) {
    const int nt = (n + 127) / 128;
    return (nt + kSymK - 1) / kSymK;
}

// Per-halo descriptor: members src4[off, off + n); partial slot s of member i at
// part[part_off + s * n + i], s = 0..nb (n, nb as planned; a shrinking live set uses its
// own nb and the planned stride).
struct __attribute__((aligned(8))) SymDesc {
    int64_t off, part_off;
    int32_t n, nb;
};

struct SymPlan {
    std::vector<SymDesc> halos;
    std::vector<int4> items;    // {halo, block A, block B, 0}, A <= B, largest haloes first
    std::vector<int32_t> order; // the haloes by size, largest first
    int64_t n_items = 0, n_members = 0, n_part = 0;
};

// Plans from HOST member offsets (H+1 entries); the item list only with with_items.
// With max_members != 0, haloes above that size (every halo when it is negative) get no
// blocks (nb = 0, no slots).
int plan_sym(int64_t H, const int64_t* off, SymPlan& plan, bool with_items,
             int max_members = 0);

// Persistent launch over the items (count plan.n_items, or *n_items_dev when given);
// live_n (device, may be null) overrides each halo's member count.  head must hold 0
// at launch unless reset_head.
int launch_pot_sym(const SymPlan& plan, const float4* src4, const SymDesc* d_halos,
                   const int* live_n, const int4* d_items, const int64_t* n_items_dev,
                   float2* part, int* head, hipStream_t st, bool reset_head);

}  // namespace halma
