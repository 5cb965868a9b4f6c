// Barnes-Hut workspace (bh.hip).
#pragma once

#include "internal.hpp"

namespace bht {

// Tree in breadth-first order, as the reference's arena (src/tsne/arena.rs:36-62).
struct BhNodes {
    float4 *node;     // [2n] com.x, com.y, count (u32 bits), first_child (u32 bits, ~0 = leaf)
    uint32_t *meta;   // [2n] child_count | level << 8
    uint32_t *range;  // [4n] start, end in the sorted order
    float *hw2;       // [33] squared half-width per level
    uint32_t *counts; // [1] node count
};

struct BhDev {
    int n;
    float *y;
    uint64_t *codes, *codes_alt;
    uint32_t *idx, *idx_alt;
    void *sort_temp;
    BhNodes nodes;
    float *frep;
    EpochScalars *scal;
    double *dpart;
};

// Requires scal->amin / amax of the current y (the centring kernel writes them).  Writes frep
// (non-edge forces) and scal->zsum / z / inv_z.  Returns the point indices in Morton order.
const uint32_t *bh_repulsion(const BhDev &d, float theta, hipStream_t s, int slot);

}  // namespace bht
