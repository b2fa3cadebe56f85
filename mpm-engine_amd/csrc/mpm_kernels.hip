// gfx950 kernels and C-ABI entry points of libmpm_hip.so (declared in include/mpm_hip.h).
//
// Particle state is SoA planes, the grid is a dense array of 4x4x4-node blocks holding
// float4 (momentum, mass) in and float4 velocity out. All launches go to the caller's
// stream; nothing here allocates.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mpm_device.h"
#include "mpm_tile.h"

namespace {

constexpr int kThreads = 256;

// ------------------------------------------------------------------------ colliders
// Applies collider c to one node; returns the reaction impulse in imp (zero if untouched).
MPM_DEV void apply_collider(const mpm_collider_t& c, const mpm_grid_t& G, int gi, int gj, int gk,
                            float m, v3& vo, v3& imp, v3& torque) {
    imp = mk3(0.f, 0.f, 0.f);
    torque = mk3(0.f, 0.f, 0.f);
    const float time = c.time;
    const bool live = time >= c.start_time && time < c.end_time;
    // node position rounded first: offsets are xn - point (no multiply-add fused into the
    // subtraction), the rounding the oracle's node-on-the-boundary tests see
    const v3 xn = mk3((float)gi * G.dx, (float)gj * G.dx, (float)gk * G.dx);
    switch (c.kind) {
        case MPM_COLLIDER_PLANE: {  // mpm_solver_warp.py:1929-1992
            if (!live) return;
            v3 n = mk3(c.normal[0], c.normal[1], c.normal[2]);
            v3 off = xn - mk3(c.point[0], c.point[1], c.point[2]);
            if (!(dot(off, n) < 0.0f)) return;
            if (c.surface == 0) {
                vo = mk3(0.f, 0.f, 0.f);
            } else if (c.surface == 11) {
                float z = (float)gk * G.dx;
                if (z < 0.4f || z > 0.53f) vo = mk3(0.f, 0.f, 0.f);
                else vo = mk3(vo.x, 0.0f, vo.z) * 0.3f;
            } else {
                float nc = dot(vo, n);
                v3 w = (c.surface == 1) ? vo - n * nc : vo - n * fminf(nc, 0.0f);
                if (nc < 0.0f && len(w) > 1e-20f) w = normalize(w) * fmaxf(0.0f, len(w) + nc * c.friction);
                vo = w;
            }
        } break;
        case MPM_COLLIDER_BOX: {  // mpm_solver_warp.py:2053-2088
            v3 vb = mk3(c.velocity[0], c.velocity[1], c.velocity[2]);
            if (live) {
                v3 off = xn - mk3(c.point[0], c.point[1], c.point[2]);
                if (fabsf(off.x) < c.size[0] && fabsf(off.y) < c.size[1] && fabsf(off.z) < c.size[2]) {
                    imp = (vo - vb) * m;
                    vo = vb;
                }
            } else if (c.reset == 1) {
                if (time < c.end_time + 15.0f * c.dt) vo = mk3(0.f, 0.f, 0.f);
            }
        } break;
        case MPM_COLLIDER_CUP: {  // mpm_solver_warp.py:2161-2225
            if (!live) return;
            m33 R;
#pragma unroll
            for (int q = 0; q < 9; ++q) R.a[q] = c.rot[q];
            v3 rel = xn - mk3(c.point[0], c.point[1], c.point[2]);
            v3 loc = mTvmul(R, rel);
            float sdf = glass_sdf(loc, c);
            if (!(sdf < c.contact_band)) return;
            v3 vs = mk3(c.velocity[0], c.velocity[1], c.velocity[2]) + cross(mk3(c.omega[0], c.omega[1], c.omega[2]), rel);
            v3 vin = vo;
            v3 vnew = vs;
            if (sdf > -c.sticky_depth) {
                vnew = vin;
                float h = 0.25f * G.dx;
                v3 nl = mk3(glass_sdf(loc + mk3(h, 0, 0), c) - glass_sdf(loc - mk3(h, 0, 0), c),
                            glass_sdf(loc + mk3(0, h, 0), c) - glass_sdf(loc - mk3(0, h, 0), c),
                            glass_sdf(loc + mk3(0, 0, h), c) - glass_sdf(loc - mk3(0, 0, h), c));
                float nlen = len(nl);
                if (nlen > 1.0e-12f) {
                    v3 n = mvmul(R, vdiv(nl, nlen));
                    v3 vr = vin - vs;
                    float vn = dot(vr, n);
                    if (vn < 0.0f) {
                        v3 vt = vr - n * vn;
                        float vtl = len(vt);
                        if (vtl > 1.0e-20f) vt = vdiv(vt, vtl) * fmaxf(vtl + c.friction * vn, 0.0f);
                        vnew = vs + vt;
                    }
                }
            }
            imp = (vin - vnew) * m;
            torque = cross(rel, imp);
            vo = vnew;
        } break;
        case MPM_COLLIDER_WALLS: {  // mpm_solver_warp.py:2294-2359
            if (!live) return;
            const int pad = 3;
            if (gi < pad && vo.x < 0) vo.x = 0.0f;
            if (gi >= G.n[0] - pad && vo.x > 0) vo.x = 0.0f;
            if (gj < pad && vo.y < 0) vo.y = 0.0f;
            if (gj >= G.n[1] - pad && vo.y > 0) vo.y = 0.0f;
            if (gk < pad && vo.z < 0) vo.z = 0.0f;
            if (gk >= G.n[2] - pad && vo.z > 0) vo.z = 0.0f;
        } break;
        case MPM_COLLIDER_SDF: {  // mpm_solver_warp.py:2680-2736
            if (!live) return;
            v3 rel = xn - mk3(c.point[0], c.point[1], c.point[2]);
            v3 body = quat_rotate(c.quat, rel, true);
            v3 f = mk3((body.x - c.origin[0]) / c.cell, (body.y - c.origin[1]) / c.cell, (body.z - c.origin[2]) / c.cell);
            float rf = (float)c.res - 1.0f;
            if (!(f.x >= 0.0f && f.y >= 0.0f && f.z >= 0.0f && f.x <= rf && f.y <= rf && f.z <= rf)) return;
            float sd;
            trilerp(c.sdf_val, 1, c.res, f, &sd);
            if (!(sd <= c.band)) return;
            float g[3];
            trilerp(c.sdf_grad, 3, c.res, f, g);
            v3 gr = mk3(g[0], g[1], g[2]);
            if (!(len(gr) > 1.0e-12f)) return;
            v3 nw = quat_rotate(c.quat, normalize(gr), false);
            v3 vsurf = mk3(c.velocity[0], c.velocity[1], c.velocity[2]) + cross(mk3(c.omega[0], c.omega[1], c.omega[2]), rel);
            v3 vf = vo;
            v3 vr = vf - vsurf;
            float vn = dot(vr, nw);
            v3 vnew = vf;
            if (c.surface == 0) vnew = vsurf;
            else if (c.surface == 1) vnew = vsurf + (vr - nw * vn);
            else if (vn < 0.0f) {
                v3 vt = vr - nw * vn;
                float tl = len(vt);
                if (tl > 1.0e-12f) vt = vt * (fmaxf(0.0f, tl + c.friction * vn) / tl);
                vnew = vsurf + vt;
            }
            imp = (vf - vnew) * m;
            torque = cross(rel, imp);
            vo = vnew;
        } break;
        case MPM_COLLIDER_POINTCLOUD: {  // mpm_solver_warp.py:2586-2605
            if (!live) return;
            if (c.occupancy[((size_t)gi * G.n[1] + gj) * G.n[2] + gk] == 1) vo = mk3(0.f, 0.f, 0.f);
        } break;
        default: break;
    }
}

constexpr int kMaxSlots = 64;    // wrench slots per launch (documented in mpm_hip.h)
constexpr int kGridWgs = 1024;   // persistent workgroups of the grid update (4 wavefronts per SIMD)

// Deterministic wrench: per-wavefront sums in fixed point. Force units are the momentum
// scale / 2^8 (a wavefront sums 64 nodes' impulses, each a sum of many contributions),
// torque units that over the next power of two of the domain diagonal (the lever arm
// bound). A wavefront sum beyond 2^52 units is dropped and raises error bit 64.
MPM_DEV void wrench_scales(const mpm_grid_t& G, float& sf, float& st) {
    sf = ldexpf(G.fix_scale[1], -8);
    const float diag = 1.7320508f * (float)max(G.n[0], max(G.n[1], G.n[2])) * G.dx;
    st = ldexpf(sf, -max(0, (int)ceilf(log2f(diag))));
}

MPM_DEV unsigned long long wrench_q(float x, float S, int* error_flag) {
    const float y = x * S;
    if (!(fabsf(y) < 4503599627370496.0f)) {
        if (error_flag) atomicOr(error_flag, 64);
        return 0ull;
    }
    return (unsigned long long)__float2ll_rn(y);
}

// One substep's grid update over listed blocks (one 64-node block per wavefront per
// iteration, persistent workgroups). Block codes are global over the batched scenes
// (scene * blocks-per-scene + local code); a block of scene s applies that scene's row
// of ncoll colliders. Collider reaction impulses are summed per wavefront (shuffles).
// Up to kMaxSlots wrench slots, each wavefront adds its sums into its own LDS row, the
// rows are added in a fixed order at the end and written to this workgroup's slice of
// `partials`; k_wrench_reduce then adds the slices in workgroup order. No float
// atomics: the wrench readout is bit-reproducible for a given block list. With more
// slots (one or more wrench colliders per scene of a large batch) lane 0 adds the
// wavefront sums to `wrench` with float atomics instead.
// kStage (at most kStageRows colliders per scene, 16-byte aligned rows): each wavefront
// copies its block's scene row of colliders into LDS with one round of parallel vector
// loads (again only when the scene changes), and the per-node collider loop reads them
// from there, instead of a chain of dependent scalar loads, one per collider, that miss
// the scalar cache together at the start of the launch.
constexpr int kStageRows = 12;
static_assert(sizeof(mpm_collider_t) % 16 == 0, "collider rows are copied as float4");
constexpr int kRowVec = (int)(sizeof(mpm_collider_t) / 16);

template <bool kStage>
__global__ __launch_bounds__(kThreads) void k_grid_update(mpm_grid_t G, const mpm_model_t* __restrict__ md,
                                                          const mpm_tick_ctl_t* __restrict__ ctl,
                                                          const int32_t* __restrict__ list,
                                                          const int32_t* __restrict__ count,
                                                          const mpm_collider_t* __restrict__ col, int ncoll,
                                                          int nslots, float* __restrict__ partials,
                                                          float* __restrict__ wrench, float dt, int flags) {
    // per-wavefront wrench rows: float sums, or fixed-point integer sums (deterministic)
    __shared__ unsigned long long s_wraw[kThreads / 64][kMaxSlots * 6];
    __shared__ float4 s_rows[kStage ? kThreads / 64 : 1][kStage ? kStageRows * kRowVec : 1];
    float(*s_w)[kMaxSlots * 6] = reinterpret_cast<float(*)[kMaxSlots * 6]>(&s_wraw[0][0]);
    int staged = -1;  // scene whose collider row this wavefront holds in s_rows
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const bool lds_w = nslots <= kMaxSlots;
    const bool fixed = (flags & MPM_GRID_FIXED_IN) != 0;
    if (lds_w)
        for (int t = lane; t < nslots * 6; t += 64) {
            if (fixed) s_wraw[wave][t] = 0ull;
            else s_w[wave][t] = 0.0f;
        }
    const int nblk = *count;
    const int bps = G.nb[0] * G.nb[1] * G.nb[2];
    const int ncode = bps * G.n_scenes;
    const int waves_total = gridDim.x * (kThreads / 64);
    const float g0 = md->gravity[0], g1 = md->gravity[1], g2 = md->gravity[2];
    const float damp = md->grid_v_damping_scale;
    const int blo0 = ctl->box_lo[0], blo1 = ctl->box_lo[1], blo2 = ctl->box_lo[2];
    const int bhi0 = ctl->box_hi[0], bhi1 = ctl->box_hi[1], bhi2 = ctl->box_hi[2];
    const double inv_m = fixed ? 1.0 / (double)G.fix_scale[0] : 0.0;
    const double inv_p = fixed ? 1.0 / (double)G.fix_scale[1] : 0.0;
    float sw_f = 0.f, sw_t = 0.f;  // fixed-point wrench scales (force, torque)
    if (fixed) wrench_scales(G, sw_f, sw_t);
    // Wavefront w owns list entries w, w + W, w + 2W, ... (W wavefronts in the launch) and
    // takes 64 of them at a time (one per lane) with their block states, ballots the
    // blocks that need a visit and walks only those: skipped blocks cost one parallel load
    // instead of a chain of dependent ones.
    const int wg_wave = blockIdx.x * (kThreads / 64) + wave;
    for (long it = 0; wg_wave + (long)waves_total * 64 * it < nblk; ++it) {
        const long bl = wg_wave + (long)waves_total * (lane + 64 * it);
        int my_code = bl < nblk ? list[bl] : -1;
        if ((unsigned)my_code >= (unsigned)ncode) my_code = -1;
        const int my_touch = (my_code >= 0 && G.block_touch) ? G.block_touch[my_code] : 1;
        unsigned long long todo = __ballot(my_code >= 0 && (my_touch & 3) != 0);
    while (todo) {
        const int l = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        const int code = __shfl(my_code, l, 64);
        const int scene = code / bps;
        const mpm_collider_t* __restrict__ cs = col + (size_t)scene * ncoll;
        if constexpr (kStage) {
            if (scene != staged) {  // wavefront-uniform
                const float4* __restrict__ src = reinterpret_cast<const float4*>(cs);
                for (int t = lane; t < ncoll * kRowVec; t += 64) s_rows[wave][t] = src[t];
                staged = scene;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            cs = reinterpret_cast<const mpm_collider_t*>(&s_rows[wave][0]);
        }
        int bi, bj, bk;
        decode_block(G, code - scene * bps, bi, bj, bk);
        const int gi = bi * 4 + (lane >> 4), gj = bj * 4 + ((lane >> 2) & 3), gk = bk * 4 + (lane & 3);
        const long sb = block_storage(G, code);  // wavefront-uniform
        if (sb < 0) continue;                    // a listed block without a pool slot: nothing stored
        const size_t o = 4 * ((size_t)sb * 64 + lane);
        const bool valid = gi < G.n[0] && gj < G.n[1] && gk < G.n[2];
        const bool inbox = valid && gi >= blo0 && gi < bhi0 && gj >= blo1 && gj < bhi1 && gk >= blo2 && gk < bhi2;
        if (G.block_touch) {  // wavefront-uniform: one block per wavefront
            const int touch = __shfl(my_touch, l, 64);
            if (!(touch & 1)) {  // nothing scattered here: accumulators are zero
                if (touch & 2) {  // velocities written last time go back to zero
                    if (inbox) *reinterpret_cast<float4*>(G.node_out + o) = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (lane == 0) G.block_touch[code] = 0;
                }
                continue;
            }
        }
        float4 in4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid && fixed) {  // deterministic scatter: fixed point -> fp32, then clear
            longlong2* f = reinterpret_cast<longlong2*>(G.node_fix + o);
            const longlong2 f01 = f[0], f23 = f[1];
            in4 = make_float4((float)((double)f01.x * inv_p), (float)((double)f01.y * inv_p),
                              (float)((double)f23.x * inv_p), (float)((double)f23.y * inv_m));
            if ((f01.x | f01.y | f23.x | f23.y) != 0) {  // empty nodes are already clear
                f[0] = make_longlong2(0, 0);
                f[1] = make_longlong2(0, 0);
            }
        } else if (valid) {
            in4 = *reinterpret_cast<const float4*>(G.node_in + o);
        }
        const float m = in4.w;
        v3 vo = mk3(0.f, 0.f, 0.f);
        if (m > 1e-15f) {
            const float inv = 1.0f / m;
            vo = mk3(in4.x * inv + dt * g0, in4.y * inv + dt * g1, in4.z * inv + dt * g2);
        }
        if (damp < 1.0f) vo = vo * damp;
        auto collider_step = [&](int q) {
            const mpm_collider_t& c = cs[q];
            v3 imp = mk3(0.f, 0.f, 0.f), tq = mk3(0.f, 0.f, 0.f);
            const bool act = inbox && c.active && m > 0.0f && gi >= c.lo[0] && gi < c.hi[0] && gj >= c.lo[1] &&
                             gj < c.hi[1] && gk >= c.lo[2] && gk < c.hi[2];
            if (act) apply_collider(c, G, gi, gj, gk, m, vo, imp, tq);
            if (c.slot >= 0 && c.slot < nslots && __any(act)) {
                const float s0 = wave_sum(imp.x), s1 = wave_sum(imp.y), s2 = wave_sum(imp.z);
                const float s3 = wave_sum(tq.x), s4 = wave_sum(tq.y), s5 = wave_sum(tq.z);
                // the wavefront's own LDS row takes adds that do not return the old value
                // (one lane, so the order is fixed)
                if (lane == 0 && fixed) {  // order-independent: integer sums of fixed point
                    int* ef = G.error_flag;
                    const unsigned long long q6[6] = {wrench_q(s0, sw_f, ef), wrench_q(s1, sw_f, ef),
                                                      wrench_q(s2, sw_f, ef), wrench_q(s3, sw_t, ef),
                                                      wrench_q(s4, sw_t, ef), wrench_q(s5, sw_t, ef)};
                    if (lds_w) {
                        unsigned long long* w = &s_wraw[wave][6 * c.slot];
                        for (int e = 0; e < 6; ++e) atomicAdd(w + e, q6[e]);
                    } else {
                        unsigned long long* w = reinterpret_cast<unsigned long long*>(partials) + 6 * (size_t)c.slot;
                        for (int e = 0; e < 6; ++e) atomicAdd(w + e, q6[e]);
                    }
                } else if (lane == 0) {
                    if (lds_w) {
                        float* w = &s_w[wave][6 * c.slot];
                        atomicAdd(w + 0, s0); atomicAdd(w + 1, s1); atomicAdd(w + 2, s2);
                        atomicAdd(w + 3, s3); atomicAdd(w + 4, s4); atomicAdd(w + 5, s5);
                    } else {
                        float* w = wrench + 6 * (size_t)c.slot;
                        atomicAdd(w + 0, s0); atomicAdd(w + 1, s1); atomicAdd(w + 2, s2);
                        atomicAdd(w + 3, s3); atomicAdd(w + 4, s4); atomicAdd(w + 5, s5);
                    }
                }
            }
        };
        if constexpr (kStage) {
            // colliders whose box misses this block act on none of its nodes: one round of
            // LDS reads marks the others, and only those run, in registration order
            unsigned hit = 0;
            const int gi0 = bi * 4, gj0 = bj * 4, gk0 = bk * 4;
#pragma unroll
            for (int q = 0; q < kStageRows; ++q) {
                if (q < ncoll) {
                    const mpm_collider_t& c = cs[q];
                    if (c.active && gi0 < c.hi[0] && gi0 + 4 > c.lo[0] && gj0 < c.hi[1] && gj0 + 4 > c.lo[1] &&
                        gk0 < c.hi[2] && gk0 + 4 > c.lo[2])
                        hit |= 1u << q;
                }
            }
            for (; hit; hit &= hit - 1) collider_step(__ffs(hit) - 1);
        } else {
            for (int q = 0; q < ncoll; ++q) collider_step(q);
        }
        if (inbox) *reinterpret_cast<float4*>(G.node_out + o) = make_float4(vo.x, vo.y, vo.z, 0.0f);
        if (valid && !fixed && (flags & MPM_GRID_ZERO_IN) &&
            (in4.x != 0.0f || in4.y != 0.0f || in4.z != 0.0f || in4.w != 0.0f))  // empty: already clear
            *reinterpret_cast<float4*>(G.node_in + o) = make_float4(0.f, 0.f, 0.f, 0.f);
        if (G.block_touch && lane == 0) G.block_touch[code] = 2;
    }
    }
    if (nslots > 0 && lds_w) {
        __syncthreads();
        for (int t = threadIdx.x; t < nslots * 6; t += kThreads) {
            if (fixed) {
                unsigned long long acc = 0ull;
                for (int w = 0; w < kThreads / 64; ++w) acc += s_wraw[w][t];
                unsigned long long* row = reinterpret_cast<unsigned long long*>(partials) + (size_t)blockIdx.x * nslots * 6;
                if (!(flags & MPM_GRID_WRENCH_DEFER)) row[t] = acc;
                // deferred: the row (this workgroup's own, zero after each flush) gathers the
                // tick's sums; an integer add sent without waiting for the old value, and
                // only when there is something to add
                else if (acc != 0ull) atomicAdd(row + t, acc);
            } else {
                float acc = s_w[0][t];
                for (int w = 1; w < kThreads / 64; ++w) acc += s_w[w][t];
                partials[(size_t)blockIdx.x * nslots * 6 + t] = acc;
            }
        }
    }
}

// wrench[v] += sum over workgroups of partials[wg][v]: one workgroup per value, strided
// per-thread sums then an LDS tree, both in a fixed order (bit-reproducible)
__global__ __launch_bounds__(kThreads) void k_wrench_reduce(const float* __restrict__ partials, int nwg, int nvals,
                                                            float* __restrict__ wrench) {
    __shared__ float sh[kThreads];
    const int v = blockIdx.x;
    float acc = 0.0f;
    for (int b = threadIdx.x; b < nwg; b += kThreads) acc += partials[(size_t)b * nvals + v];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) wrench[v] += sh[0];
}

// deterministic wrench: wrench[v] += (sum over the nparts integer partial rows of value v)
// / scale; with nparts == 1 the single row (global-atomic path) is cleared after use.
// Integer sums: the float result does not depend on any ordering.
__global__ __launch_bounds__(kThreads) void k_wrench_fix_flush(mpm_grid_t G, unsigned long long* __restrict__ fixed,
                                                                int nparts, int nvals, float* __restrict__ wrench) {
    __shared__ unsigned long long sh[kThreads];
    const int v = blockIdx.x;  // one workgroup per value: strided sums, then an LDS tree
    unsigned long long acc = 0ull;
    for (int b = threadIdx.x; b < nparts; b += kThreads) {
        acc += fixed[(size_t)b * nvals + v];
        fixed[(size_t)b * nvals + v] = 0ull;  // consumed: rows are zero between flushes
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float sf, st;
        wrench_scales(G, sf, st);
        const double inv = 1.0 / (double)((v % 6) < 3 ? sf : st);
        wrench[v] += (float)((double)(long long)sh[0] * inv);
    }
}

inline unsigned grid_update_wgs(int capacity) { return (unsigned)min((capacity + 3) / 4, kGridWgs); }

__global__ void k_grid_zero_box(mpm_grid_t G, int lo0, int lo1, int lo2, int d0, int d1, int d2, int which) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long vol = (long)d0 * d1 * d2;
    if (t >= vol * G.n_scenes) return;
    const int scene = (int)(t / vol);
    t -= scene * vol;
    int k = lo2 + (int)(t % d2);
    long r = t / d2;
    int j = lo1 + (int)(r % d1);
    int i = lo0 + (int)(r / d1);
    size_t o = 4 * (scene_base(G, scene) + (size_t)node_index(G, i, j, k));
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
    if (which & 1) *reinterpret_cast<float4*>(G.node_in + o) = z;
    if (which & 2) *reinterpret_cast<float4*>(G.node_out + o) = z;
}

__global__ void k_grid_zero_blocks(mpm_grid_t G, const int32_t* list, const int32_t* count, int which) {
    const int nblk = *count;
    for (int b = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); b < nblk; b += gridDim.x * (kThreads / 64)) {
        const long sb = block_storage(G, list[b]);
        if (sb < 0) continue;
        size_t o = 4 * ((size_t)sb * 64 + (threadIdx.x & 63));
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        if (which & 1) *reinterpret_cast<float4*>(G.node_in + o) = z;
        if (which & 2) *reinterpret_cast<float4*>(G.node_out + o) = z;
    }
}

__global__ void k_blocks_from_box(mpm_grid_t G, mpm_tick_ctl_t* ctl, int32_t* list) {
    const int bl0 = ctl->box_lo[0] >> 2, bl1 = ctl->box_lo[1] >> 2, bl2 = ctl->box_lo[2] >> 2;
    const int bh0 = (ctl->box_hi[0] + 3) >> 2, bh1 = (ctl->box_hi[1] + 3) >> 2, bh2 = (ctl->box_hi[2] + 3) >> 2;
    const int d0 = max(bh0 - bl0, 0), d1 = max(bh1 - bl1, 0), d2 = max(bh2 - bl2, 0);
    const int per = d0 * d1 * d2;
    const int total = per * G.n_scenes;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) ctl->nblocks = total;
    if (t >= total) return;
    const int scene = t / per;
    const int u = t - scene * per;
    int bk = bl2 + u % d2;
    int r = u / d2;
    int bj = bl1 + r % d1;
    int bi = bl0 + r / d1;
    list[t] = scene * (G.nb[0] * G.nb[1] * G.nb[2]) + (bi * G.nb[1] + bj) * G.nb[2] + bk;
}

// Launch box on the device from mpm_bounds' output: the particle box widened by
// pad + travel * max |v|, as nodes [floor(lo / dx), ceil(hi / dx) + 1) clamped to the grid
// (the host's grid_box with no halo); the whole grid when a position is not finite.
__global__ void k_box_from_bounds(const float* __restrict__ b, mpm_grid_t G, float pad, float travel,
                                  mpm_tick_ctl_t* ctl) {
    if (threadIdx.x != 0) return;
    const float w = pad + travel * b[6];
    const bool ok = b[7] == 0.0f && w == w;
    for (int a = 0; a < 3; ++a) {
        int lo = ok ? (int)floorf((b[a] - w) / G.dx) : 0;
        int hi = ok ? (int)ceilf((b[3 + a] + w) / G.dx) + 1 : G.n[a];
        lo = max(0, min(lo, G.n[a]));
        hi = max(lo, min(hi, G.n[a]));
        ctl->box_lo[a] = lo;
        ctl->box_hi[a] = hi;
    }
}

// ------------------------------------------------------------------ sparse blocks
// blocks_mark, mpm_utils.py:1506-1513 (int() truncation toward zero, >> 2, clamp)
__global__ void k_blocks_mark(mpm_particles_t P, mpm_grid_t G, int32_t* raw) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int scene = P.scene ? P.scene[p] : 0;
    if ((unsigned)scene >= (unsigned)G.n_scenes) return;
    int b[3];
    for (int a = 0; a < 3; ++a) {
        int t = (int)(P.x[a * P.stride + p] * G.inv_dx - 0.5f);
        t >>= 2;
        b[a] = min(max(t, 0), G.nb[a] - 1);
    }
    raw[(size_t)scene * (G.nb[0] * G.nb[1] * G.nb[2]) + (b[0] * G.nb[1] + b[1]) * G.nb[2] + b[2]] = 1;
}

// blocks_dilate, mpm_utils.py:1516-1529
__global__ void k_blocks_dilate(mpm_grid_t G, const int32_t* raw, int32_t* cur) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int bps = G.nb[0] * G.nb[1] * G.nb[2];
    if (t >= bps * G.n_scenes) return;
    const int scene = t / bps;
    raw += (size_t)scene * bps;  // dilation stays inside the scene
    const int u = t - scene * bps;
    int z = u % G.nb[2], r = u / G.nb[2], y = r % G.nb[1], x = r / G.nb[1];
    int on = 0;
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
            for (int k = -1; k <= 1; ++k) {
                int xi = min(max(x + i, 0), G.nb[0] - 1), yj = min(max(y + j, 0), G.nb[1] - 1),
                    zk = min(max(z + k, 0), G.nb[2] - 1);
                if (raw[(xi * G.nb[1] + yj) * G.nb[2] + zk] == 1) on = 1;
            }
    cur[t] = on;
}

// blocks_compact, mpm_utils.py:1532-1544
__global__ void k_blocks_compact(mpm_grid_t G, const int32_t* cur, const int32_t* prev, int32_t* cur_list,
                                 int32_t* union_list, int32_t* counts) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int total = G.nb[0] * G.nb[1] * G.nb[2] * G.n_scenes;
    if (t >= total) return;
    int c = cur[t];
    if (c == 1) cur_list[atomicAdd(counts, 1)] = t;
    if (c == 1 || prev[t] == 1) union_list[atomicAdd(counts + 1, 1)] = t;
}

// ---------------------------------------------------------------------- bounds
MPM_DEV int ord_key(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
MPM_DEV float ord_val(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }

__global__ void k_bounds_init(int* out) {
    int t = threadIdx.x;
    if (t < 3) out[t] = 0x7fffffff;                          // min keys start at +max
    else if (t < 6) out[t] = (int)0x80000000;                // max keys start at -max
    else if (t == 6) out[t] = 0;                             // max |v| key (>= 0)
    else if (t == 7) out[t] = 0;                             // non-finite count
}

__global__ __launch_bounds__(kThreads) void k_bounds(mpm_particles_t P, int* out) {
    __shared__ int sh[8][kThreads / 64];
    int kmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
    int kmax[3] = {(int)0x80000000, (int)0x80000000, (int)0x80000000};
    int kv = 0, bad = 0;
    // grid-stride: a bounded launch keeps the same-address global atomics at the end few
    // (one set per workgroup, all on the same seven words)
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P.n; p += gridDim.x * blockDim.x) {
        int nonfinite = 0;
        for (int a = 0; a < 3; ++a) {
            float x = P.x[a * P.stride + p];
            if (!isfinite(x)) nonfinite = 1;
            int k = ord_key(x);
            kmin[a] = min(kmin[a], k); kmax[a] = max(kmax[a], k);
            float av = fabsf(P.v[a * P.stride + p]);
            kv = max(kv, __float_as_int(av));
        }
        bad += nonfinite;  // particles with a non-finite position
    }
    for (int off = 32; off > 0; off >>= 1) {
        for (int a = 0; a < 3; ++a) {
            kmin[a] = min(kmin[a], __shfl_xor(kmin[a], off, 64));
            kmax[a] = max(kmax[a], __shfl_xor(kmax[a], off, 64));
        }
        kv = max(kv, __shfl_xor(kv, off, 64));
        bad += __shfl_xor(bad, off, 64);
    }
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        for (int a = 0; a < 3; ++a) { sh[a][wv] = kmin[a]; sh[3 + a][wv] = kmax[a]; }
        sh[6][wv] = kv; sh[7][wv] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kThreads / 64; ++w) {
            for (int a = 0; a < 3; ++a) { sh[a][0] = min(sh[a][0], sh[a][w]); sh[3 + a][0] = max(sh[3 + a][0], sh[3 + a][w]); }
            sh[6][0] = max(sh[6][0], sh[6][w]); sh[7][0] += sh[7][w];
        }
        for (int a = 0; a < 3; ++a) { atomicMin(out + a, sh[a][0]); atomicMax(out + 3 + a, sh[3 + a][0]); }
        atomicMax(out + 6, sh[6][0]);
        if (sh[7][0]) atomicAdd(out + 7, sh[7][0]);
    }
}

__global__ void k_bounds_finish(int* out) {
    int t = threadIdx.x;
    float* f = reinterpret_cast<float*>(out);
    if (t < 6) f[t] = ord_val(out[t]);
    else if (t == 6) f[t] = __int_as_float(out[t]);
    else if (t == 7) f[t] = (float)out[t];
}

// ------------------------------------------------------------------------ sorting
// key of solver.py:458-461: base = floor(x / dx - 0.5) >> 2, nb = n/4 + 2; batched
// scenes are scene-major (scene * nb^3 + key), so each scene stays one contiguous range
// cells = 1 refines the key by the stencil base cell inside the block (6 low bits), so
// particles sharing a base cell become consecutive
__global__ void k_sort_keys(mpm_particles_t P, mpm_grid_t G, int32_t* keys, int32_t* vals, int32_t* unsorted,
                            int cells) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int nb = (G.n[0] >> 2) + 2;
    auto key_of = [&](int q) {
        int c[3], b[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = (int)floorf(P.x[a * P.stride + q] / G.dx - 0.5f);
            b[a] = c[a] >> 2;
        }
        const int scene = P.scene ? min(max(P.scene[q], 0), G.n_scenes - 1) : 0;
        const int k = scene * (nb * nb * nb) + (b[0] * nb + b[1]) * nb + b[2];
        return cells ? (k << 6) | ((c[0] & 3) << 4) | ((c[1] & 3) << 2) | (c[2] & 3) : k;
    };
    int k = key_of(p);
    keys[p] = k;
    vals[p] = p;
    if (p + 1 < P.n && key_of(p + 1) < k) atomicOr(unsorted, 1);
}

// Up to kGroupPlanes 4-byte planes of n particles (any fields), gathered by perm into
// scratch and copied back: one launch each way for the group instead of one per field and
// one copy per plane.
constexpr int kGroupPlanes = 9;
struct PlaneGroup {
    uint32_t* p[kGroupPlanes];
    int count;
};

__global__ void k_gather_planes(PlaneGroup g, const int32_t* __restrict__ perm, uint32_t* __restrict__ scratch, int n,
                                int stride) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * g.count) return;
    const int q = (int)(t / n), i = (int)(t - (long)q * n);
    scratch[(size_t)q * stride + i] = g.p[q][perm[i]];
}

__global__ void k_copy_planes(PlaneGroup g, const uint32_t* __restrict__ scratch, int n, int stride) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * g.count) return;
    const int q = (int)(t / n), i = (int)(t - (long)q * n);
    g.p[q][i] = scratch[(size_t)q * stride + i];
}

// ------------------------------------------------------------------------ exports
__global__ void k_export_R(mpm_particles_t P, float* R) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    m33 F = ld9(P.F_trial, P.stride, p), U, V;
    v3 s;
    svd3(F, U, s, V);
    if (mdet(U) < 0.0f) { U.a[2] = -U.a[2]; U.a[5] = -U.a[5]; U.a[8] = -U.a[8]; }
    if (mdet(V) < 0.0f) { V.a[2] = -V.a[2]; V.a[5] = -V.a[5]; V.a[8] = -V.a[8]; }
    m33 Rt = mT(mmul(U, mT(V)));
    for (int i = 0; i < 9; ++i) R[9 * (size_t)p + i] = Rt.a[i];
}

__global__ void k_export_cov(mpm_particles_t P, const float* init_cov, float* cov) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    m33 F = ld9(P.F_trial, P.stride, p);
    const float* c = init_cov + 6 * (size_t)p;
    m33 ic;
    ic.a[0] = c[0]; ic.a[1] = c[1]; ic.a[2] = c[2];
    ic.a[3] = c[1]; ic.a[4] = c[3]; ic.a[5] = c[4];
    ic.a[6] = c[2]; ic.a[7] = c[4]; ic.a[8] = c[5];
    m33 r = mmul(mmul(F, ic), mT(F));
    float* o = cov + 6 * (size_t)p;
    o[0] = r.a[0]; o[1] = r.a[1]; o[2] = r.a[2]; o[3] = r.a[4]; o[4] = r.a[5]; o[5] = r.a[8];
}

__global__ void k_grid_to_dense(mpm_grid_t G, float* m, float* vin, float* vout) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)G.n[0] * G.n[1] * G.n[2];
    if (t >= per * G.n_scenes) return;
    const int scene = (int)(t / per);
    const long u = t - scene * per;
    int k = (int)(u % G.n[2]);
    long r = u / G.n[2];
    int j = (int)(r % G.n[1]);
    int i = (int)(r / G.n[1]);
    const long na = node_addr(G, scene_base(G, scene), i, j, k);
    if (na < 0) {  // a block without pool storage reads as empty
        if (m) m[t] = 0.0f;
        if (vin) vin[3 * t] = vin[3 * t + 1] = vin[3 * t + 2] = 0.0f;
        if (vout) vout[3 * t] = vout[3 * t + 1] = vout[3 * t + 2] = 0.0f;
        return;
    }
    size_t o = 4 * (size_t)na;
    if (m) m[t] = G.node_in[o + 3];
    if (vin) { vin[3 * t] = G.node_in[o]; vin[3 * t + 1] = G.node_in[o + 1]; vin[3 * t + 2] = G.node_in[o + 2]; }
    if (vout) { vout[3 * t] = G.node_out[o]; vout[3 * t + 1] = G.node_out[o + 1]; vout[3 * t + 2] = G.node_out[o + 2]; }
}

inline unsigned blocks_for(long n, int threads = kThreads) { return (unsigned)((n + threads - 1) / threads); }

inline int32_t launch_status() { return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP; }

inline int grid_ok(const mpm_grid_t* G) {
    return G && G->node_in && G->node_out && G->n[0] > 0 && G->n[1] > 0 && G->n[2] > 0 && G->n_scenes >= 1 &&
           G->nb[0] == (G->n[0] + 3) / 4 && G->nb[1] == (G->n[1] + 3) / 4 && G->nb[2] == (G->n[2] + 3) / 4 &&
           (!G->block_slot || G->pool_blocks > 0);
}

// entry points that address the node arrays densely
inline int dense_grid_ok(const mpm_grid_t* G) { return grid_ok(G) && !G->block_slot; }

inline int particles_ok(const mpm_particles_t* P) {
    return P && P->n >= 0 && P->stride >= P->n && P->x && P->v && P->F && P->F_trial && P->C && P->stress &&
           P->L && P->vol && P->mass && P->mu && P->lam && P->yield_stress && P->Jp && P->material && P->selection;
}

}  // namespace

// ======================================================================== C ABI
// Scratch of mpm_grid_update_noalloc: the wrench partial rows that mpm_grid_update takes
// as its `workspace`, held in a file-scope device array instead (each GPU the library runs
// on holds its own copy until the process ends). In host code `scratch` names a host-side
// stand-in, not device memory: kernels get the per-device address from scratch_ptr().
static __device__ __align__(256) float scratch[MPM_NOALLOC_SCRATCH_FLOATS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device with that
// device current (hipGetSymbolAddress) and kept.
int32_t scratch_ptr(float** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return MPM_ERR_HIP;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MPM_ERR_HIP;
        g_scratch[dev] = static_cast<float*>(p);
    }
    *out = g_scratch[dev];
    return MPM_OK;
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n) {
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < n) p[t] = 0ull;
}

}  // namespace

extern "C" {

int32_t mpm_abi_version(void) { return MPM_ABI_VERSION; }

const char* mpm_status_string(int32_t s) {
    switch (s) {
        case MPM_OK: return "ok";
        case MPM_ERR_ARG: return "invalid argument";
        case MPM_ERR_HIP: return "HIP runtime error";
        case MPM_ERR_SIZE: return "size beyond the documented maximum";
        default: return "unknown status";
    }
}

int32_t mpm_struct_size(int32_t which) {
    switch (which) {
        case 0: return (int32_t)sizeof(mpm_particles_t);
        case 1: return (int32_t)sizeof(mpm_grid_t);
        case 2: return (int32_t)sizeof(mpm_model_t);
        case 3: return (int32_t)sizeof(mpm_collider_t);
        case 4: return (int32_t)sizeof(mpm_tick_ctl_t);
        case 5: return (int32_t)sizeof(mpm_cdf_lanes_t);
        case 6: return (int32_t)sizeof(mpm_cdf_collider_t);
        case 7: return (int32_t)sizeof(mpm_adjoint_t);
        case 8: return (int32_t)sizeof(mpm_rigid_body_t);
        case 9: return (int32_t)sizeof(mpm_rigid_plane_t);
        case 10: return (int32_t)sizeof(mpm_particle_modifier_t);
        default: return -1;
    }
}

int32_t mpm_grid_zero_box(const mpm_grid_t* G, const int32_t* lo, const int32_t* hi, int32_t which, void* stream) {
    if (!dense_grid_ok(G) || !lo || !hi) return MPM_ERR_ARG;
    int l[3], d[3];
    for (int a = 0; a < 3; ++a) {
        l[a] = lo[a] < 0 ? 0 : lo[a];
        int h = hi[a] > G->n[a] ? G->n[a] : hi[a];
        d[a] = h - l[a];
        if (d[a] <= 0) return MPM_OK;
    }
    long total = (long)d[0] * d[1] * d[2] * G->n_scenes;
    hipLaunchKernelGGL(k_grid_zero_box, dim3(blocks_for(total)), dim3(kThreads), 0, (hipStream_t)stream, *G, l[0],
                       l[1], l[2], d[0], d[1], d[2], which);
    return launch_status();
}

int32_t mpm_grid_zero_blocks(const mpm_grid_t* G, const int32_t* list, const int32_t* count, int32_t capacity,
                             int32_t which, void* stream) {
    if (!grid_ok(G) || !list || !count || capacity < 0) return MPM_ERR_ARG;
    if (capacity == 0) return MPM_OK;
    unsigned nwg = (unsigned)min((capacity + 3) / 4, 8192);
    hipLaunchKernelGGL(k_grid_zero_blocks, dim3(nwg), dim3(kThreads), 0, (hipStream_t)stream, *G, list, count, which);
    return launch_status();
}

int32_t mpm_blocks_from_box(const mpm_grid_t* G, mpm_tick_ctl_t* ctl, int32_t* list, void* stream) {
    if (!dense_grid_ok(G) || !ctl || !list) return MPM_ERR_ARG;
    long total = (long)G->nb[0] * G->nb[1] * G->nb[2] * G->n_scenes;
    hipLaunchKernelGGL(k_blocks_from_box, dim3(blocks_for(total)), dim3(kThreads), 0, (hipStream_t)stream, *G, ctl,
                       list);
    return launch_status();
}

int32_t mpm_box_from_bounds(const float* bounds, const mpm_grid_t* G, float pad, float travel, mpm_tick_ctl_t* ctl,
                            void* stream) {
    if (!bounds || !dense_grid_ok(G) || !ctl || !(pad >= 0.0f) || !(travel >= 0.0f)) return MPM_ERR_ARG;
    hipLaunchKernelGGL(k_box_from_bounds, dim3(1), dim3(64), 0, (hipStream_t)stream, bounds, *G, pad, travel, ctl);
    return launch_status();
}

int32_t mpm_grid_workspace_floats(int32_t capacity, int32_t wrench_slots) {
    if (capacity < 0 || wrench_slots < 0) return -1;
    if (wrench_slots > kMaxSlots) return 12 * wrench_slots;  // atomic path; fixed-point wrench only
    return (int32_t)(grid_update_wgs(capacity) * wrench_slots * 12);  // float or int64 partial rows
}

int32_t mpm_grid_update(const mpm_grid_t* G, const mpm_model_t* model, const mpm_tick_ctl_t* ctl,
                        const int32_t* list, const int32_t* count, int32_t capacity, const mpm_collider_t* colliders,
                        int32_t ncoll, float* wrench, int32_t wrench_slots, float* workspace, float dt,
                        int32_t flags, void* stream) {
    if (!grid_ok(G) || !model || !ctl || !list || !count || capacity < 0 || ncoll < 0) return MPM_ERR_ARG;
    if (ncoll > 0 && !colliders) return MPM_ERR_ARG;
    if (wrench_slots < 0) return MPM_ERR_ARG;
    const bool fixed = (flags & MPM_GRID_FIXED_IN) != 0;
    const bool fixed_order = wrench_slots <= kMaxSlots;
    if (wrench_slots > 0 && (!wrench || ((fixed_order || fixed) && !workspace))) return MPM_ERR_ARG;
    if (fixed && (!G->node_fix || !(G->fix_scale[0] > 0.0f) || !(G->fix_scale[1] > 0.0f))) return MPM_ERR_ARG;
    if (capacity == 0) return MPM_OK;
    const unsigned nwg = grid_update_wgs(capacity);
    hipStream_t s = (hipStream_t)stream;
    const bool stage = ncoll > 0 && ncoll <= kStageRows && ((uintptr_t)colliders & 15) == 0;
    hipLaunchKernelGGL(stage ? k_grid_update<true> : k_grid_update<false>, dim3(nwg), dim3(kThreads), 0, s, *G,
                       model, ctl, list, count, colliders, ncoll, wrench_slots, workspace, wrench, dt, flags);
    if (wrench_slots > 0 && fixed && !(flags & MPM_GRID_WRENCH_DEFER))
        hipLaunchKernelGGL(k_wrench_fix_flush, dim3(wrench_slots * 6), dim3(kThreads), 0, s, *G,
                           reinterpret_cast<unsigned long long*>(workspace), fixed_order ? (int)nwg : 1,
                           wrench_slots * 6, wrench);
    else if (wrench_slots > 0 && fixed_order && !fixed)
        hipLaunchKernelGGL(k_wrench_reduce, dim3(wrench_slots * 6), dim3(kThreads), 0, s, workspace, (int)nwg,
                           wrench_slots * 6, wrench);
    return launch_status();
}

int32_t mpm_grid_update_noalloc(const mpm_grid_t* G, const mpm_model_t* model, const mpm_tick_ctl_t* ctl,
                                const int32_t* list, const int32_t* count, int32_t capacity,
                                const mpm_collider_t* colliders, int32_t ncoll, float* wrench, int32_t wrench_slots,
                                float dt, int32_t flags, void* stream) {
    if (capacity < 0 || wrench_slots < 0) return MPM_ERR_ARG;
    // the deferred integer rows would have to survive between calls in the shared array
    if (flags & MPM_GRID_WRENCH_DEFER) return MPM_ERR_ARG;
    const int64_t need = wrench_slots > kMaxSlots ? 12 * (int64_t)wrench_slots
                                                  : (int64_t)grid_update_wgs(capacity) * wrench_slots * 12;
    if (need > MPM_NOALLOC_SCRATCH_FLOATS) return MPM_ERR_SIZE;
    float* ws = nullptr;
    if (wrench_slots > 0) {
        const int32_t st = scratch_ptr(&ws);
        if (st != MPM_OK) return st;
        // above kMaxSlots the fixed-point wrench is summed with integer atomics into one
        // row, which must start at zero (the array keeps the last call's contents; a fill
        // kernel, not a memset, so the call can be captured in a graph). Up to kMaxSlots
        // every workgroup overwrites its own row.
        if ((flags & MPM_GRID_FIXED_IN) && wrench_slots > kMaxSlots) {
            const int n = 6 * wrench_slots;
            hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                               (hipStream_t)stream, reinterpret_cast<unsigned long long*>(ws), n);
        }
    }
    return mpm_grid_update(G, model, ctl, list, count, capacity, colliders, ncoll, wrench, wrench_slots, ws, dt,
                           flags, stream);
}

int32_t mpm_grid_wrench_flush(const mpm_grid_t* G, int32_t capacity, float* wrench, int32_t wrench_slots,
                              float* workspace, void* stream) {
    if (!grid_ok(G) || capacity < 0 || wrench_slots < 0 || (wrench_slots > 0 && (!wrench || !workspace)) ||
        !(G->fix_scale[0] > 0.0f) || !(G->fix_scale[1] > 0.0f))
        return MPM_ERR_ARG;
    if (wrench_slots == 0 || capacity == 0) return MPM_OK;
    const unsigned nwg = grid_update_wgs(capacity);
    hipLaunchKernelGGL(k_wrench_fix_flush, dim3(wrench_slots * 6), dim3(kThreads), 0, (hipStream_t)stream, *G,
                       reinterpret_cast<unsigned long long*>(workspace), wrench_slots <= kMaxSlots ? (int)nwg : 1,
                       wrench_slots * 6, wrench);
    return launch_status();
}

int32_t mpm_blocks_mark(const mpm_particles_t* P, const mpm_grid_t* G, int32_t* raw, void* stream) {
    if (!particles_ok(P) || !grid_ok(G) || !raw || (G->n_scenes > 1 && !P->scene)) return MPM_ERR_ARG;
    if (P->n == 0) return MPM_OK;
    hipLaunchKernelGGL(k_blocks_mark, dim3(blocks_for(P->n)), dim3(kThreads), 0, (hipStream_t)stream, *P, *G, raw);
    return launch_status();
}

int32_t mpm_blocks_dilate(const mpm_grid_t* G, const int32_t* raw, int32_t* cur, void* stream) {
    if (!grid_ok(G) || !raw || !cur) return MPM_ERR_ARG;
    long total = (long)G->nb[0] * G->nb[1] * G->nb[2] * G->n_scenes;
    hipLaunchKernelGGL(k_blocks_dilate, dim3(blocks_for(total)), dim3(kThreads), 0, (hipStream_t)stream, *G, raw, cur);
    return launch_status();
}

int32_t mpm_blocks_compact(const mpm_grid_t* G, const int32_t* cur, const int32_t* prev, int32_t* cur_list,
                           int32_t* union_list, int32_t* counts, void* stream) {
    if (!grid_ok(G) || !cur || !prev || !cur_list || !union_list || !counts) return MPM_ERR_ARG;
    long total = (long)G->nb[0] * G->nb[1] * G->nb[2] * G->n_scenes;
    hipLaunchKernelGGL(k_blocks_compact, dim3(blocks_for(total)), dim3(kThreads), 0, (hipStream_t)stream, *G, cur,
                       prev, cur_list, union_list, counts);
    return launch_status();
}

int32_t mpm_bounds(const mpm_particles_t* P, float* out, void* stream) {
    if (!particles_ok(P) || !out) return MPM_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    int* o = reinterpret_cast<int*>(out);
    hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, o);
    if (P->n > 0) hipLaunchKernelGGL(k_bounds, dim3(min(blocks_for(P->n), 512u)), dim3(kThreads), 0, s, *P, o);
    hipLaunchKernelGGL(k_bounds_finish, dim3(1), dim3(64), 0, s, o);
    return launch_status();
}

static int32_t sort_keys(const mpm_particles_t* P, const mpm_grid_t* G, int32_t* keys, int32_t* vals,
                         int32_t* unsorted, int cells, void* stream) {
    if (!particles_ok(P) || !grid_ok(G) || !keys || !vals || !unsorted) return MPM_ERR_ARG;
    const long nbs = (G->n[0] >> 2) + 2;
    if (nbs * nbs * nbs * G->n_scenes * (cells ? 64 : 1) > 0x7fffffffL) return MPM_ERR_SIZE;
    if (P->n == 0) return MPM_OK;
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks_for(P->n)), dim3(kThreads), 0, (hipStream_t)stream, *P, *G, keys, vals,
                       unsorted, cells);
    return launch_status();
}

int32_t mpm_sort_keys(const mpm_particles_t* P, const mpm_grid_t* G, int32_t* keys, int32_t* vals, int32_t* unsorted,
                      void* stream) {
    return sort_keys(P, G, keys, vals, unsorted, 0, stream);
}

int32_t mpm_sort_keys_cells(const mpm_particles_t* P, const mpm_grid_t* G, int32_t* keys, int32_t* vals,
                            int32_t* unsorted, void* stream) {
    return sort_keys(P, G, keys, vals, unsorted, 1, stream);
}

int32_t mpm_permute(const mpm_particles_t* P, const int32_t* perm, float* scratch, void* stream) {
    if (!particles_ok(P) || !perm || !scratch) return MPM_ERR_ARG;
    if (P->n == 0) return MPM_OK;
    hipStream_t s = (hipStream_t)stream;
    const int n = P->n, S = P->stride;
    uint32_t* sc = reinterpret_cast<uint32_t*>(scratch);
    PlaneGroup g{};
    auto flush = [&]() {
        if (g.count == 0) return;
        const unsigned nb = blocks_for((long)n * g.count);
        hipLaunchKernelGGL(k_gather_planes, dim3(nb), dim3(kThreads), 0, s, g, perm, sc, n, S);
        hipLaunchKernelGGL(k_copy_planes, dim3(nb), dim3(kThreads), 0, s, g, sc, n, S);
        g.count = 0;
    };
    auto add = [&](void* a, int planes) {  // a field of `planes` planes of 4-byte values
        if (!a) return;
        for (int q = 0; q < planes; ++q) {
            g.p[g.count++] = reinterpret_cast<uint32_t*>(a) + (size_t)q * S;
            if (g.count == kGroupPlanes) flush();
        }
    };
    add(P->x, 3); add(P->v, 3);
    add(P->F, 9); add(P->F_trial, 9); add(P->C, 9); add(P->stress, 9); add(P->L, 9);
    add(P->vol, 1); add(P->mass, 1); add(P->mu, 1); add(P->lam, 1);
    add(P->yield_stress, 1); add(P->Jp, 1);
    add(P->cov, 6);
    add(P->material, 1); add(P->selection, 1);
    add(P->scene, 1); add(P->cdf_tag, 1); add(P->rigid_id, 1);
    add(P->x_ref, 3); add(P->modifier_mask, 1);
    flush();
    return launch_status();
}

int32_t mpm_export_R(const mpm_particles_t* P, float* R_out, void* stream) {
    if (!particles_ok(P) || !R_out) return MPM_ERR_ARG;
    if (P->n == 0) return MPM_OK;
    hipLaunchKernelGGL(k_export_R, dim3(blocks_for(P->n)), dim3(kThreads), 0, (hipStream_t)stream, *P, R_out);
    return launch_status();
}

int32_t mpm_export_cov(const mpm_particles_t* P, const float* init_cov, float* cov_out, void* stream) {
    if (!particles_ok(P) || !init_cov || !cov_out) return MPM_ERR_ARG;
    if (P->n == 0) return MPM_OK;
    hipLaunchKernelGGL(k_export_cov, dim3(blocks_for(P->n)), dim3(kThreads), 0, (hipStream_t)stream, *P, init_cov,
                       cov_out);
    return launch_status();
}

int32_t mpm_grid_to_dense(const mpm_grid_t* G, float* m, float* v_in, float* v_out, void* stream) {
    if (!grid_ok(G)) return MPM_ERR_ARG;
    long total = (long)G->n[0] * G->n[1] * G->n[2] * G->n_scenes;
    hipLaunchKernelGGL(k_grid_to_dense, dim3(blocks_for(total)), dim3(kThreads), 0, (hipStream_t)stream, *G, m, v_in,
                       v_out);
    return launch_status();
}

}  // extern "C"
