// LDS node tiles shared by the particle kernels (forward pass and its adjoint): the
// stencil bounding box of a workgroup's (or a wavefront's) particles staged in LDS so
// that gathers read each node once and scatters accumulate with LDS atomics before one
// global flush per node.
#ifndef MPM_TILE_H
#define MPM_TILE_H

#include <limits.h>

#include "mpm_device.h"

namespace {

// Deterministic scatter: every contribution in fixed point of scale S (a power of two),
// rounded to the nearest unit (ties to even). Integer sums do not depend on the order the
// atomics land in. The rounding goes through float64: y is exact there, and y + 1.5 *
// 2^52 lies in the binade [2^52, 2^53), whose spacing is 1, so the add rounds y to an
// integer and the difference of the two bit patterns is that integer (for |y| < 2^51;
// 2^12 such contributions to one node still fit an int64). Three instructions instead of
// the generic float -> int64 conversion, 108 times per particle in the forward scatter
// and 81 in the G2P adjoint.
MPM_DEV int mod3(int v) { return ((v % 3) + 3) % 3; }

MPM_DEV unsigned long long fix_round(float y) {
    const double magic = 6755399441055744.0;  // 1.5 * 2^52
    return (unsigned long long)(__double_as_longlong((double)y + magic) - __double_as_longlong(magic));
}

// Workgroup size kT (256 or 128, a launch-shape choice) is a template parameter; the LDS
// arena holds 3 nodes per thread: 768 nodes (12 KB gather + 12 KB scatter) at 256.
template <int kT>
constexpr int arena_of() { return 3 * kT; }

// An LDS tile: the node box [lo, lo + (d0, d1, d2)) of one scene (node offset `soff` in
// the batched grid) staged at arena offset `base`.
// Tier 1: the stencil box of all particles of the workgroup fits kArena nodes and they
// belong to one scene (the usual case for block-sorted particles). Tier 2 (`wg` false):
// each wavefront stages the box of its own 64 particles in its quarter of the arena; a
// wavefront whose box still does not fit, or that spans two scenes (`lds` false), gathers
// and scatters through global memory.
struct Tile {
    int lo[3];
    int d1, d2, vol, base;
    size_t soff;
    bool wg, lds;
};

MPM_DEV bool fits(const int lo[3], const int hi[3], int cap, int& d0, int& d1, int& d2) {
    d0 = d1 = d2 = 0;
    if (hi[0] <= lo[0] || hi[1] <= lo[1] || hi[2] <= lo[2]) return false;  // empty (sentinels)
    d0 = hi[0] - lo[0];
    d1 = hi[1] - lo[1];
    d2 = hi[2] - lo[2];
    if (d0 > cap || d1 > cap || d2 > cap) return false;
    return d0 * d1 * d2 <= cap;
}

// Wavefront sum with DPP (all 64 lanes active): row shifts by 1, 2, 4, 8 leave each row's
// inclusive prefix sum in its lanes (lanes shifted in from outside the row add 0), so
// lane 15 holds the row sum; row_bcast:15 adds rows 0 and 2 into rows 1 and 3 and
// row_bcast:31 rows 0-1 into rows 2 and 3: lane 63 ends with the total, which every lane
// reads. A fixed order, so the sum is the same in every run.
template <int kCtrl, int kRowMask>
MPM_DEV float dpp_add(float x) {
    const int y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, kRowMask, 0xf, false);
    return x + __int_as_float(y);
}
MPM_DEV float wave_sum(float v) {
    v = dpp_add<0x111, 0xf>(v);  // row_shr:1
    v = dpp_add<0x112, 0xf>(v);  // row_shr:2
    v = dpp_add<0x114, 0xf>(v);  // row_shr:4
    v = dpp_add<0x118, 0xf>(v);  // row_shr:8
    v = dpp_add<0x142, 0xa>(v);  // row_bcast:15
    v = dpp_add<0x143, 0xc>(v);  // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Wavefront min / max with DPP (all 64 lanes active): row shifts by 1, 2, 4, 8 leave the
// row's extreme in its lane 15 (min and max are idempotent, so overlapping windows are
// fine; lanes shifted in from outside the row keep the identity), row_bcast:15 and
// row_bcast:31 carry rows 0-2 into lane 63, which every lane then reads.
template <bool kMax>
MPM_DEV int wave_extreme(int x) {
    const int id = kMax ? INT_MIN : INT_MAX;
    auto op = [](int a, int b) { return kMax ? max(a, b) : min(a, b); };
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x111, 0xf, 0xf, false));  // row_shr:1
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x112, 0xf, 0xf, false));  // row_shr:2
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x114, 0xf, 0xf, false));  // row_shr:4
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x118, 0xf, 0xf, false));  // row_shr:8
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x142, 0xa, 0xf, false));  // row_bcast:15 -> rows 1, 3
    x = op(x, __builtin_amdgcn_update_dpp(id, x, 0x143, 0xc, 0xf, false));  // row_bcast:31 -> rows 2, 3
    return __builtin_amdgcn_readlane(x, 63);
}

// n = (i * d1 + j) * d2 + k -> (i, j, k) for n < 2^16 and d1, d2 <= 2^10 without integer
// division: floor((n + 1/2) / d) from a float reciprocal (the hardware v_rcp_f32, within
// 1 ulp) is exact there: the half keeps the product of an exact multiple above the
// integer, and the rounding error (< 2^-6 / 2^10 here) stays far below 1/(2d)
MPM_DEV void split3(int n, float inv_d1, float inv_d2, int d1, int d2, int& i, int& j, int& k) {
    const int r = (int)(((float)n + 0.5f) * inv_d2);
    k = n - r * d2;
    i = (int)(((float)r + 0.5f) * inv_d1);
    j = r - i * d1;
}

// per-thread stencil box [lo, hi) and scene range [slo, shi] (sentinels INT_MAX / INT_MIN
// when inactive) -> tile; contains one __syncthreads, so every thread of the workgroup
// must call it
template <int kT>
MPM_DEV Tile make_tile(const mpm_grid_t& G, int lo[3], int hi[3], int slo, int shi, int (*red)[8], bool periodic) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_extreme<false>(lo[a]);
        hi[a] = wave_extreme<true>(hi[a]);
    }
    slo = wave_extreme<false>(slo);
    shi = wave_extreme<true>(shi);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int a = 0; a < 3; ++a) {
            red[wave][a] = lo[a];
            red[wave][3 + a] = hi[a];
        }
        red[wave][6] = slo;
        red[wave][7] = shi;
    }
    __syncthreads();
    int wlo[3], whi[3];
    for (int a = 0; a < 3; ++a) {
        int l = red[0][a], h = red[0][3 + a];
        for (int w = 1; w < kT / 64; ++w) {
            l = min(l, red[w][a]);
            h = max(h, red[w][3 + a]);
        }
        wlo[a] = l;
        whi[a] = h;
    }
    int wslo = red[0][6], wshi = red[0][7];
    for (int w = 1; w < kT / 64; ++w) {
        wslo = min(wslo, red[w][6]);
        wshi = max(wshi, red[w][7]);
    }
    constexpr int kArena = arena_of<kT>();
    Tile t;
    int d0, d1, d2;
    if (!periodic && wslo == wshi && fits(wlo, whi, kArena, d0, d1, d2)) {
        t.wg = true;
        t.lds = true;
        t.base = 0;
        t.soff = scene_base(G, wslo);
        for (int a = 0; a < 3; ++a) t.lo[a] = wlo[a];
    } else {
        t.wg = false;
        t.lds = !periodic && slo == shi && fits(lo, hi, kArena / (kT / 64), d0, d1, d2);
        t.base = wave * (kArena / (kT / 64));
        t.soff = t.lds ? scene_base(G, slo) : 0;
        for (int a = 0; a < 3; ++a) t.lo[a] = lo[a];
    }
    t.d1 = d1;
    t.d2 = d2;
    t.vol = t.lds ? d0 * d1 * d2 : 0;
    return t;
}

// cooperative loop over the tile's nodes: the whole workgroup for a workgroup tile, the
// owning wavefront for a wavefront tile
template <int kT, class F>
MPM_DEV void tile_for(const Tile& t, F f) {
    if (!t.lds) return;
    const int start = t.wg ? (int)threadIdx.x : (int)(threadIdx.x & 63);
    const int step = t.wg ? kT : 64;
    const float inv1 = __builtin_amdgcn_rcpf((float)t.d1), inv2 = __builtin_amdgcn_rcpf((float)t.d2);
    for (int n = start; n < t.vol; n += step) {
        int i, j, k;
        split3(n, inv1, inv2, t.d1, t.d2, i, j, k);
        f(t.base + n, t.lo[0] + i, t.lo[1] + j, t.lo[2] + k);
    }
}

// Mark every grid block the tile's box overlaps (clamped to the grid) as scattered into,
// by the tile's participants: a plain store of 3 (bit 0 = scattered; bit 1 does not matter
// once bit 0 is set, the grid update rewrites the block and leaves it at 2), so concurrent
// markers need no atomics and nothing waits on them; over-marking only costs the grid
// update some work
template <int kT>
MPM_DEV void mark_tile_blocks(const mpm_grid_t& G, const Tile& t) {
    if (!G.block_touch || !t.lds || t.vol == 0) return;
    const int dims[3] = {t.vol / (t.d1 * t.d2), t.d1, t.d2};
    int b0[3], nbk[3];
    for (int a = 0; a < 3; ++a) {
        const int lo = max(t.lo[a], 0), hi = min(t.lo[a] + dims[a], G.n[a]) - 1;
        b0[a] = lo >> 2;
        nbk[a] = hi >= lo ? (hi >> 2) - b0[a] + 1 : 0;
    }
    const int total = nbk[0] * nbk[1] * nbk[2];
    const size_t boff = t.soff >> 6;
    const int start = t.wg ? (int)threadIdx.x : (int)(threadIdx.x & 63);
    const int step = t.wg ? kT : 64;
    const float inv1 = __builtin_amdgcn_rcpf((float)nbk[1]), inv2 = __builtin_amdgcn_rcpf((float)nbk[2]);
    for (int n = start; n < total; n += step) {
        int i, j, k;
        split3(n, inv1, inv2, nbk[1], nbk[2], i, j, k);
        G.block_touch[boff + ((size_t)(b0[0] + i) * G.nb[1] + (b0[1] + j)) * G.nb[2] + (b0[2] + k)] = 3;
    }
}

MPM_DEV bool in_grid(const mpm_grid_t& G, int i, int j, int k) {
    return (unsigned)i < (unsigned)G.n[0] && (unsigned)j < (unsigned)G.n[1] && (unsigned)k < (unsigned)G.n[2];
}

}  // namespace

#endif
