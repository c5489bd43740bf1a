// cg_tile.h -- the peak index of a query tile and its window probe, shared by
// the device kernels and the host harness (tests/native/tile_harness.cpp).
//
// A tile is kTileQueries consecutive queries.  Their peaks (m/z + shift, in
// float32) are merged into one list sorted by (value, query lane, peak index)
// and padded with +inf.  A directory of kTileBuckets buckets over [lo, hi]
// maps bucket(x) to the first list position whose bucket is >= bucket(x);
// since bucket() is monotone in x under float32 rounding, every peak >= x sits
// at or after dir[bucket(x)], so a probe for the window [x0, x1] starts there,
// skips the few peaks below x0 and collects peaks up to x1.
//
// The window [fl(r - tolc), fl(r + tolc)] of a reference peak r is a
// superset of the exact float64 matchms window for the coarse tolerance tolc
// of cg_lane.h (every rounding involved is covered by its margin); each hit
// is re-tested exactly by the caller.
#pragma once

#include "cg_lane.h"

namespace cs {

constexpr int kTileQueries = 64;
constexpr int kTileBuckets = 4096;

CS_HD int tile_bucket(float x, float lo, float inv_w) {
    const float f = (x - lo) * inv_w;
    if (!(f > 0.0f)) return 0;  // also NaN-safe
    if (f >= (float)(kTileBuckets - 1)) return kTileBuckets - 1;
    return (int)f;
}

// inv_w of a tile whose sorted values span [lo, hi].
CS_HD float tile_inv_width(float lo, float hi) {
    const float span = hi - lo;
    return span > 0.0f ? (float)kTileBuckets / span : 0.0f;
}

// Calls emit(p) for every list position p with x0 <= smz(p) <= x1.
// smz(p) must return +inf at p == n (the pad), dir(b) the directory entry.
template <class Smz, class Dir, class Emit>
CS_HD void tile_probe(float x0, float x1, float lo, float inv_w, Smz smz, Dir dir, Emit emit) {
    int p = dir(tile_bucket(x0, lo, inv_w));
    float v = smz(p);
    while (v < x0) v = smz(++p);
    while (v <= x1) {
        emit(p);
        v = smz(++p);
    }
}

CS_HD uint32_t tile_who(int lane, int j) { return (uint32_t(lane) << 16) | uint32_t(j); }

}  // namespace cs
