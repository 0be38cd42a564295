// Geometry helpers shared by the list build and the surrogate features:
// wrapping fixed-point differences, the conservative box test and the
// periodic column reach.
#pragma once

#include "gm_internal.hpp"

namespace gm {

__device__ inline int wrapdiff(int a, int b) { return (int)((unsigned)a - (unsigned)b); }

// Conservative box-box test under the minimum image: per-axis gap between
// the boxes, compared against r_skin plus a margin that covers the f32 and
// quantisation error (the exact per-pair decision comes later, in f64).
__device__ inline bool boxes_close(int4 a, float4 ha, int4 b, float4 hb, const BuildParams &bp) {
    float gx = fabsf((float)wrapdiff(b.x, a.x) * bp.sq[0]) - ha.x - hb.x;
    float gy = fabsf((float)wrapdiff(b.y, a.y) * bp.sq[1]) - ha.y - hb.y;
    float gz = fabsf((float)wrapdiff(b.z, a.z) * bp.sq[2]) - ha.z - hb.z;
    gx = fmaxf(gx, 0.f); gy = fmaxf(gy, 0.f); gz = fmaxf(gz, 0.f);
    return gx * gx + gy * gy + gz * gz <= bp.r_search * bp.r_search;
}

// Columns of a periodic ring within `k` of `c`, as at most two ascending
// [first, last] pieces p0, p1 (an absent piece is [0, -1]), so that visiting
// them in order visits ascending indices.
__device__ inline void ring_pieces(int c, int k, int n, int2 &p0, int2 &p1) {
    p1 = make_int2(0, -1);
    const int lo = c - k, hi = c + k;
    if (2 * k + 1 >= n) p0 = make_int2(0, n - 1);
    else if (lo < 0) { p0 = make_int2(0, hi); p1 = make_int2(lo + n, n - 1); }
    else if (hi >= n) { p0 = make_int2(0, hi - n); p1 = make_int2(lo, n - 1); }
    else p0 = make_int2(lo, hi);
}

// Smallest possible distance along one axis between atoms of two columns
// `d` apart on a ring of `n` (quantisation error is far inside r_search's margin).
__device__ inline float column_gap(int d, int n, float w) {
    d = d < 0 ? -d : d;
    d = min(d, n - d);
    return (float)max(d - 1, 0) * w;
}

}  // namespace gm
