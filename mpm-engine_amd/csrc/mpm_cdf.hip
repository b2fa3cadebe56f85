// CPIC (CDF) node stamps and lane upload: the host-driven per-substep tag maintenance of
// the reference (cdf_zero_tags / cdf_copy_prev / cdf_block_mask / cdf_stamp,
// src/warpmpm/kernels/mpm_solver_warp.py:94-183, orchestrated by _stamp_cdf, :3040-3098).
// Each launch covers a node (or block) box; one thread per node, so the read-modify-write
// of a stamp needs no atomics as long as colliders stamp one after the other on a stream.
#include <hip/hip_runtime.h>

#include "mpm_device.h"

namespace {

constexpr int kThreads = 256;

struct Box {
    int lo[3], d[3];
    long total;
};

// clip [lo, hi) to [0, lim) per axis; total = 0 when empty
Box clip_box(const int32_t* lo, const int32_t* hi, const int32_t* lim) {
    Box b;
    b.total = 1;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = lo[a] < 0 ? 0 : lo[a];
        const int h = hi[a] > lim[a] ? lim[a] : hi[a];
        b.d[a] = h - b.lo[a];
        if (b.d[a] <= 0) b.total = 0;
    }
    if (b.total) b.total = (long)b.d[0] * b.d[1] * b.d[2];
    return b;
}

MPM_DEV bool box_node(const Box& b, long t, int& i, int& j, int& k) {
    if (t >= b.total) return false;
    k = b.lo[2] + (int)(t % b.d[2]);
    const long r = t / b.d[2];
    j = b.lo[1] + (int)(r % b.d[1]);
    i = b.lo[0] + (int)(r / b.d[1]);
    return true;
}

__global__ __launch_bounds__(kThreads) void k_cdf_zero(mpm_grid_t G, Box b) {
    int i, j, k;
    if (!box_node(b, (long)blockIdx.x * kThreads + threadIdx.x, i, j, k)) return;
    const int n = node_index(G, i, j, k);
    G.cdf.tag[n] = 0;
    G.cdf.dist[n] = 1.0e9f;
}

__global__ __launch_bounds__(kThreads) void k_cdf_copy_prev(mpm_grid_t G, Box b) {
    int i, j, k;
    if (!box_node(b, (long)blockIdx.x * kThreads + threadIdx.x, i, j, k)) return;
    const int n = node_index(G, i, j, k);
    G.cdf.tag_prev[n] = G.cdf.tag[n];
    G.cdf.dist_prev[n] = G.cdf.dist[n];
}

// one wavefront per block: OR of tag | tag_prev over the block's 64 nodes (nodes past
// the logical grid hold 0, as the zero / copy kernels never write them)
__global__ __launch_bounds__(kThreads) void k_cdf_block_mask(mpm_grid_t G, Box b) {
    const long wave = ((long)blockIdx.x * kThreads + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    int bi, bj, bk;
    if (!box_node(b, wave, bi, bj, bk)) return;
    const int code = (bi * G.nb[1] + bj) * G.nb[2] + bk;
    const int ci = bi * 4 + (lane >> 4), cj = bj * 4 + ((lane >> 2) & 3), ck = bk * 4 + (lane & 3);
    int occ = 0;
    if (ci < G.n[0] && cj < G.n[1] && ck < G.n[2]) {
        const int n = code * 64 + lane;
        occ = G.cdf.tag[n] | G.cdf.tag_prev[n];
    }
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) occ |= __shfl_xor(occ, s, 64);
    if (lane == 0) G.cdf.block[code] = occ;
}

// cdf_stamp, mpm_solver_warp.py:142-183
__global__ __launch_bounds__(kThreads) void k_cdf_stamp(mpm_grid_t G, mpm_cdf_collider_t c, Box b) {
    int gi, gj, gk;
    if (!box_node(b, (long)blockIdx.x * kThreads + threadIdx.x, gi, gj, gk)) return;
    if (c.time < c.start_time || c.time >= c.end_time) return;
    const v3 xw = mk3((float)gi * G.dx, (float)gj * G.dx, (float)gk * G.dx);
    const v3 body = quat_rotate(c.quat, xw - mk3(c.center[0], c.center[1], c.center[2]), true);
    const v3 f = mk3((body.x - c.origin[0]) / c.cell, (body.y - c.origin[1]) / c.cell, (body.z - c.origin[2]) / c.cell);
    const float rf = (float)c.res - 1.0f;
    if (f.x < 0.0f || f.y < 0.0f || f.z < 0.0f || f.x > rf || f.y > rf || f.z > rf) return;
    float ok, d;
    trilerp(c.valid, 1, c.res, f, &ok);
    if (ok < 0.999f) return;
    trilerp(c.val, 1, c.res, f, &d);
    const float ud = fabsf(d);
    if (ud > c.band) return;
    const int side = d >= 0.0f ? 1 : 0;
    const int n = node_index(G, gi, gj, gk);
    int t = G.cdf.tag[n];
    t = t | (1 << (2 * c.lane)) | (side << (2 * c.lane + 1));
    if (ud < G.cdf.dist[n]) {
        G.cdf.dist[n] = ud;
        t = (t & ~(3 << MPM_CDF_OWNER_SHIFT)) | (c.lane << MPM_CDF_OWNER_SHIFT);
    }
    G.cdf.tag[n] = t;
}

__global__ void k_cdf_set_lanes(mpm_cdf_lanes_t* dst, mpm_cdf_lanes_t src) {
    constexpr int kWords = (int)(sizeof(mpm_cdf_lanes_t) / 4);
    const int t = threadIdx.x;
    if (t < kWords) reinterpret_cast<int*>(dst)[t] = reinterpret_cast<const int*>(&src)[t];
}

inline int cdf_ok(const mpm_grid_t* G) {
    return G && !G->block_slot && G->n[0] > 0 && G->n[1] > 0 && G->n[2] > 0 && G->n_scenes == 1 &&
           G->nb[0] == (G->n[0] + 3) / 4 &&
           G->nb[1] == (G->n[1] + 3) / 4 && G->nb[2] == (G->n[2] + 3) / 4 && G->cdf.tag && G->cdf.tag_prev &&
           G->cdf.dist && G->cdf.dist_prev && G->cdf.block && G->cdf.lanes && G->cdf.reaction;
}

inline unsigned groups(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline int32_t status() { return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP; }

}  // namespace

extern "C" {

int32_t mpm_cdf_zero_tags(const mpm_grid_t* G, const int32_t* lo, const int32_t* hi, void* stream) {
    if (!cdf_ok(G) || !lo || !hi) return MPM_ERR_ARG;
    const Box b = clip_box(lo, hi, G->n);
    if (!b.total) return MPM_OK;
    hipLaunchKernelGGL(k_cdf_zero, dim3(groups(b.total)), dim3(kThreads), 0, (hipStream_t)stream, *G, b);
    return status();
}

int32_t mpm_cdf_copy_prev(const mpm_grid_t* G, const int32_t* lo, const int32_t* hi, void* stream) {
    if (!cdf_ok(G) || !lo || !hi) return MPM_ERR_ARG;
    const Box b = clip_box(lo, hi, G->n);
    if (!b.total) return MPM_OK;
    hipLaunchKernelGGL(k_cdf_copy_prev, dim3(groups(b.total)), dim3(kThreads), 0, (hipStream_t)stream, *G, b);
    return status();
}

int32_t mpm_cdf_stamp(const mpm_grid_t* G, const mpm_cdf_collider_t* c, const int32_t* lo, const int32_t* hi,
                      void* stream) {
    if (!cdf_ok(G) || !c || !lo || !hi || !c->val || !c->valid || c->res < 2 || c->lane < 0 || c->lane >= MPM_MAX_CDF ||
        !(c->cell > 0.0f))
        return MPM_ERR_ARG;
    const Box b = clip_box(lo, hi, G->n);
    if (!b.total) return MPM_OK;
    hipLaunchKernelGGL(k_cdf_stamp, dim3(groups(b.total)), dim3(kThreads), 0, (hipStream_t)stream, *G, *c, b);
    return status();
}

int32_t mpm_cdf_block_mask(const mpm_grid_t* G, const int32_t* blo, const int32_t* bhi, void* stream) {
    if (!cdf_ok(G) || !blo || !bhi) return MPM_ERR_ARG;
    const Box b = clip_box(blo, bhi, G->nb);
    if (!b.total) return MPM_OK;
    hipLaunchKernelGGL(k_cdf_block_mask, dim3(groups(b.total * 64)), dim3(kThreads), 0, (hipStream_t)stream, *G, b);
    return status();
}

int32_t mpm_cdf_set_lanes(const mpm_grid_t* G, const mpm_cdf_lanes_t* lanes, void* stream) {
    if (!G || !G->cdf.lanes || !lanes) return MPM_ERR_ARG;
    static_assert(sizeof(mpm_cdf_lanes_t) / 4 <= 256, "one workgroup copies the lane table");
    hipLaunchKernelGGL(k_cdf_set_lanes, dim3(1), dim3(256), 0, (hipStream_t)stream, G->cdf.lanes, *lanes);
    return status();
}

}  // extern "C"
