// The fused particle pass: gather (g2p_particle, mpm_utils.py:944-1043), return map +
// Kirchhoff stress (stress_update_particle, :1056-1149) and scatter (p2g_particle,
// :819-916) for 256 (or 128) particles per workgroup, each thread taking one or more
// particles in turn (particles per thread).
//
// LDS staging (claymore-style, on block-sorted particles): the workgroup reduces the
// stencil bounding box of its particles; when the box holds at most kArena nodes it
// loads the gather box of node_out into LDS once (one 16-byte read per node instead of
// 27 per particle) and scatters into an LDS arena of 64-bit fixed-point sums
// (ds_add_u64), then flushes each touched arena node with four global atomics: integer
// ones to node_fix in the deterministic mode, float ones to node_in otherwise. Chunks whose box is
// larger (unsorted or scattered particles) fall back to direct global gathers and
// atomics, so results never depend on the particle order. Everything that decides a
// barrier is reduced first, so every branch around __syncthreads is workgroup-uniform.
#include <hip/hip_runtime.h>
#include <limits.h>

#include <type_traits>

#include "mpm_device.h"
#include "mpm_tile.h"
#include "mpm_cdf.h"

namespace {

// Stress update restricted to the viscous / rigid ids (6, 7, 8, 10, 12): the SVD-free
// branches of stress_update_particle. Other ids flag error bit 4 and keep F_trial.
MPM_DEV bool stress_update_viscous(const m33& Ft, const m33& L, int mat, const mpm_model_t* md, PState& ps,
                                   m33& F, m33& S) {
    S = mzero();
    if (mat == 6 || mat == 10 || mat == 12) {
        float Jc = powf(mdet(Ft), 1.0f / 3.0f);
        F = mdiag(Jc, Jc, Jc);
        float J = mdet(F);
        if (mat == 6) {
            float pr = -md->bulk[mat] * (powf(J, -1.1f) - 1.0f);
            S = mdiag(pr * J, pr * J, pr * J);
        } else {
            S = stress_viscous(J, md->bulk[mat], L, md->plastic_viscosity[mat], ps.yield, md->hardening[mat],
                               md->softening[mat], md, mat == 12);
        }
        S = mscale(madd(S, mT(S)), 0.5f);
        return true;
    }
    if (mat == 7 || mat == 8) {
        F = meye();
        return true;
    }
    F = Ft;
    return false;
}

MPM_DEV float sel3(int i, const float* a) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

// Tensor-product form of the 27-node gather: per (i, j) row sum over k, then over j,
// then over i. Same sums as the reference's triple loop (v = sum w v_i, C = sum w v_i
// (x_i - x_p)^T 4/dx^2, L = sum v_i grad w^T) with fewer products and live values.
// fetch(i, j, k) returns the node velocity for stencil offsets (i, j, k). kRolled keeps
// the loop over i rolled (the weights of row i picked by selects, which are exact): the
// global-memory variant then holds 9 node loads in flight instead of 27 and fits the
// register budget, and computes bit for bit what the unrolled LDS variant computes.
template <bool kRolled = false, class Fetch>
MPM_DEV void gather_tp(const Stencil& st, float inv_dx, Fetch fetch, v3& nv, m33& nC, m33& nF) {
    nv = mk3(0.f, 0.f, 0.f);
    v3 cX = nv, cY = nv, cZ = nv;  // columns of C (by dpos component)
    v3 fX = nv, fY = nv, fZ = nv;  // columns of L (by gradient component)
#pragma unroll(kRolled ? 1 : 3)
    for (int i = 0; i < 3; ++i) {
        v3 Ai = mk3(0.f, 0.f, 0.f), Ay = Ai, Bi = Ai, Cz = Ai, Dy = Ai;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            v3 A = mk3(0.f, 0.f, 0.f), B = A, Z = A;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const v3 g = fetch(i, j, k);
                A = fma3(g, st.w[2][k], A);
                B = fma3(g, st.dw[2][k], B);
                Z = fma3(g, st.w[2][k] * ((float)k - st.fx[2]), Z);
            }
            const float wj = st.w[1][j];
            Ai = fma3(A, wj, Ai);
            Ay = fma3(A, wj * ((float)j - st.fx[1]), Ay);
            Bi = fma3(B, wj, Bi);
            Cz = fma3(Z, wj, Cz);
            Dy = fma3(A, st.dw[1][j], Dy);
        }
        const float wi = kRolled ? sel3(i, st.w[0]) : st.w[0][i];
        const float dwi = kRolled ? sel3(i, st.dw[0]) : st.dw[0][i];
        nv = fma3(Ai, wi, nv);
        cX = fma3(Ai, wi * ((float)i - st.fx[0]), cX);
        cY = fma3(Ay, wi, cY);
        cZ = fma3(Cz, wi, cZ);
        fX = fma3(Ai, dwi, fX);
        fY = fma3(Dy, wi, fY);
        fZ = fma3(Bi, wi, fZ);
    }
    const float c4 = inv_dx * 4.0f;
    nC.a[0] = cX.x * c4; nC.a[1] = cY.x * c4; nC.a[2] = cZ.x * c4;
    nC.a[3] = cX.y * c4; nC.a[4] = cY.y * c4; nC.a[5] = cZ.y * c4;
    nC.a[6] = cX.z * c4; nC.a[7] = cY.z * c4; nC.a[8] = cZ.z * c4;
    nF.a[0] = fX.x * inv_dx; nF.a[1] = fY.x * inv_dx; nF.a[2] = fZ.x * inv_dx;
    nF.a[3] = fX.y * inv_dx; nF.a[4] = fY.y * inv_dx; nF.a[5] = fZ.y * inv_dx;
    nF.a[6] = fX.z * inv_dx; nF.a[7] = fY.z * inv_dx; nF.a[8] = fZ.z * inv_dx;
}

// Fixed-point scatter of one particle in tensor-product form. The reference's momentum
// contribution to stencil node (i, j, k) (p2g_particle, mpm_utils.py:880-892),
//   w m (v + C dpos) + fs grad w,   w = w_i w_j w_k,  dpos = ((i, j, k) - f) dx,
//   grad w = (dw_i w_j w_k, w_i dw_j w_k, w_i w_j dw_k) / dx,  fs = -vol dt tau,
// regroups as
//   w (a + b0 i + b1 j + b2 k) + g0 dw_i w_j w_k + g1 w_i dw_j w_k + g2 w_i w_j dw_k
// with a = m (v - C f dx), b_c = m dx C e_c, g_c = fs e_c / dx, evaluated row by row:
// about ten operations per node instead of the ~60 of the node-by-node form. The
// coefficients carry the fixed-point unit S (a power of two, so scaling them is exact),
// and one bound per particle replaces the per-contribution range test: with |w| < 1 and
// |dw| <= 1 every contribution is below |a| + 2 sum_c |b_c| + sum_c |g_c| per component.
// Beyond 2^51 units (or NaN) the particle adds nothing and false is returned (error bit
// 64). add(i, j, k, q_momentum[3], q_mass) receives the rounded integers.
// si, sj: the caller's node strides along i and j (k has stride 1); add also receives the
// node's offset i si + j sj + k, built from per-particle row offsets (two adds per node).
template <class Add>
MPM_DEV bool scatter_tp_fix(const Stencil& st, float dx, float inv_dx, float mass, v3 v, const m33& Cm,
                            const m33& fs, float Sm, float Sp, int rot, int si, int sj, Add add) {
    const float cb = mass * dx * Sp, cg = inv_dx * Sp;
    v3 b[3], g[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        b[c] = mk3(Cm.a[c], Cm.a[3 + c], Cm.a[6 + c]) * cb;
        g[c] = mk3(fs.a[c], fs.a[3 + c], fs.a[6 + c]) * cg;
    }
    const v3 a = v * (mass * Sp) - (b[0] * st.fx[0] + b[1] * st.fx[1] + b[2] * st.fx[2]);
    const v3 bnd = mk3(fabsf(a.x), fabsf(a.y), fabsf(a.z)) +
                   2.0f * (mk3(fabsf(b[0].x) + fabsf(b[1].x) + fabsf(b[2].x), fabsf(b[0].y) + fabsf(b[1].y) + fabsf(b[2].y),
                               fabsf(b[0].z) + fabsf(b[1].z) + fabsf(b[2].z))) +
                   mk3(fabsf(g[0].x) + fabsf(g[1].x) + fabsf(g[2].x), fabsf(g[0].y) + fabsf(g[1].y) + fabsf(g[2].y),
                       fabsf(g[0].z) + fabsf(g[1].z) + fabsf(g[2].z));
    const float ms = mass * Sm;
    const float lim = 2251799813685248.0f;  // 2^51
    if (!(fmaxf(bnd.x, fmaxf(bnd.y, bnd.z)) < lim) || !(fabsf(ms) < lim)) return false;
    // Conflict-free node order. The 27 stencil nodes of any particle hold exactly one node
    // of each residue class (x mod 3, y mod 3, z mod 3). At step (si, sj, sk) the lane of
    // rank rho = rot mod 27 targets class ((si, sj, sk) + decode(rho)) mod 3, i.e. stencil
    // offset ((s + decode(rho)) - base) mod 3 per axis: lanes of different rank hit
    // different nodes in the same instruction whichever cells their particles sit in (only
    // lanes 27 apart, particles about 3 cells apart after a cell sort, can meet). Integer
    // sums, so the order changes no bit of the result.
    const int rho = rot % 27;
    const int i0 = mod3(rho / 9 - st.base[0]), j0 = mod3((rho / 3) % 3 - st.base[1]), k0 = mod3(rho % 3 - st.base[2]);
    int jr[3], kr[3], jo[3];
    float wjr[3], djr[3], wkr[3], dkr[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const int j = j0 + t >= 3 ? j0 + t - 3 : j0 + t, k = k0 + t >= 3 ? k0 + t - 3 : k0 + t;
        jr[t] = j;
        kr[t] = k;
        jo[t] = j * sj;
        wjr[t] = sel3(j, st.w[1]);
        djr[t] = sel3(j, st.dw[1]);
        wkr[t] = sel3(k, st.w[2]);
        dkr[t] = sel3(k, st.dw[2]);
    }
#pragma unroll
    for (int ti = 0; ti < 3; ++ti) {
        const int i = i0 + ti >= 3 ? i0 + ti - 3 : i0 + ti;
        const int io = i * si;
        const float wi = sel3(i, st.w[0]);
        const v3 Pi = fma3(fma3(b[0], (float)i, a), wi, g[0] * sel3(i, st.dw[0]));
        const v3 Qi = b[1] * wi, Ri = g[1] * wi, Ti = b[2] * wi, Ui = g[2] * wi;
        const float mi = ms * wi;
#pragma unroll
        for (int tj = 0; tj < 3; ++tj) {
            const float wj = wjr[tj];
            const v3 Pij = fma3(fma3(Qi, (float)jr[tj], Pi), wj, Ri * djr[tj]);
            const v3 Tij = Ti * wj, Uij = Ui * wj;
            const float mij = mi * wj;
            const int ijo = io + jo[tj];
#pragma unroll
            for (int tk = 0; tk < 3; ++tk) {
                const v3 mom = fma3(fma3(Tij, (float)kr[tk], Pij), wkr[tk], Uij * dkr[tk]);
                add(i, jr[tj], kr[tk], ijo + kr[tk], fix_round(mom.x), fix_round(mom.y), fix_round(mom.z),
                    fix_round(mij * wkr[tk]));
            }
        }
    }
    return true;
}

// Global-memory scatter for chunks whose stencil box does not fit the LDS arena
// (unsorted or spread-out particles): a rolled loop over the 27 nodes with the weights
// picked by selects (the fixed-point contributions come from scatter_tp_fix, as in the
// LDS path), so this rare path stays small next to the LDS fast path it shares a kernel
// (and a register budget) with.
template <bool kDet, class Mask>
MPM_DEV void scatter_global(const mpm_grid_t& G, size_t soff, const Stencil& st, bool periodic, float mass, v3 v,
                            const m33& Cm, const m33& fs, Mask mask, int& err) {
    float* __restrict__ gin = G.node_in;
    size_t last_blk = ~(size_t)0;  // block code this particle marked last
    // storage offset (x4) of node (ix, iy, iz), -1 without storage (error bit 512); marks
    // the node's block as scattered into (a plain store, as in mark_tile_blocks)
    auto target = [&](int ix, int iy, int iz) -> long {
        const int nidx = node_index(G, ix, iy, iz);
        if (G.block_touch) {
            const size_t blk = (soff >> 6) + (size_t)(nidx >> 6);
            if (blk != last_blk) {
                last_blk = blk;
                G.block_touch[blk] = 3;
            }
        }
        const long na = node_addr(G, soff, ix, iy, iz);
        if (na < 0) err |= 512;
        return na < 0 ? -1L : 4 * na;
    };
    if constexpr (kDet) {  // the same contributions as the LDS path, bit for bit
        unsigned long long* __restrict__ gfix = reinterpret_cast<unsigned long long*>(G.node_fix);
        const bool ok = scatter_tp_fix(st, G.dx, G.inv_dx, mass, v, Cm, fs, G.fix_scale[0], G.fix_scale[1],
                                       (int)(threadIdx.x & 63u), 0, 0,
                                       [&](int i, int j, int k, int, unsigned long long q0, unsigned long long q1,
                                           unsigned long long q2, unsigned long long q3) {
            int ix = st.base[0] + i;
            if (periodic) ix = ((ix % G.n[0]) + G.n[0]) % G.n[0];
            const int iy = st.base[1] + j, iz = st.base[2] + k;
            if (!in_grid(G, ix, iy, iz)) return;
            const long o = target(ix, iy, iz);
            if (o < 0) return;
            atomicAdd(gfix + o + 0, q0);
            atomicAdd(gfix + o + 1, q1);
            atomicAdd(gfix + o + 2, q2);
            atomicAdd(gfix + o + 3, q3);
        });
        if (!ok) err |= 64;
        return;
    }
#pragma unroll 1
    for (int n = 0; n < 27; ++n) {
        const int i = n / 9, j = (n / 3) % 3, k = n % 3;
        int ix = st.base[0] + i;
        if (periodic) ix = ((ix % G.n[0]) + G.n[0]) % G.n[0];
        const int iy = st.base[1] + j, iz = st.base[2] + k;
        if (!in_grid(G, ix, iy, iz)) continue;
        const float w0 = sel3(i, st.w[0]), w1 = sel3(j, st.w[1]), w2 = sel3(k, st.w[2]);
        const float d0 = sel3(i, st.dw[0]), d1 = sel3(j, st.dw[1]), d2 = sel3(k, st.dw[2]);
        const float wm = w0 * w1 * w2 * mass;
        const v3 dpos = mk3(((float)i - st.fx[0]) * G.dx, ((float)j - st.fx[1]) * G.dx, ((float)k - st.fx[2]) * G.dx);
        const v3 dwt = mk3(d0 * w1 * w2, w0 * d1 * w2, w0 * w1 * d2) * G.inv_dx;
        v3 add = (v + mvmul(Cm, dpos)) * wm + mvmul(fs, dwt);
        float wmv = wm;
        if (!mask(i, j, k, add, wmv)) continue;
        const long o = target(ix, iy, iz);
        if (o < 0) continue;
        atomicAdd(gin + o + 0, add.x);
        atomicAdd(gin + o + 1, add.y);
        atomicAdd(gin + o + 2, add.z);
        atomicAdd(gin + o + 3, wmv);
    }
}

// One particle's 27 scatter contributions (v + C dpos) w m + fs grad w (fs = -vol dt tau)
// handed to add(i, j, k, value, w m). kResidue: the conflict-free order of scatter_tp_fix
// (the lane of rank rot mod 27 visits residue class (s + rot) mod 27 at step s, so lanes
// of different rank never add to one LDS address in the same instruction). Otherwise step
// s visits stencil node (s + rot) mod 27 (rot = 0: every lane in the same order, as the
// segmented DPP sums need).
template <bool kResidue = false, class Add>
MPM_DEV void scatter_rotated(const Stencil& st, float dx, float inv_dx, float mass, v3 v, const m33& Cm,
                             const m33& fs, int rot, Add add) {
    int i0 = 0, j0 = 0, k0 = 0;
    if constexpr (kResidue) {
        const int rho = rot % 27;
        i0 = mod3(rho / 9 - st.base[0]);
        j0 = mod3((rho / 3) % 3 - st.base[1]);
        k0 = mod3(rho % 3 - st.base[2]);
    }
#pragma unroll 3
    for (int s = 0; s < 27; ++s) {
        int i, j, k;
        if constexpr (kResidue) {
            i = s / 9 + i0;
            j = (s / 3) % 3 + j0;
            k = s % 3 + k0;
            i = i >= 3 ? i - 3 : i;
            j = j >= 3 ? j - 3 : j;
            k = k >= 3 ? k - 3 : k;
        } else {
            int n = s + rot;
            n = n >= 27 ? n - 27 : n;
            i = n / 9;
            j = (n / 3) % 3;
            k = n % 3;
        }
        const float w0 = sel3(i, st.w[0]), w1 = sel3(j, st.w[1]), w2 = sel3(k, st.w[2]);
        const float e0 = sel3(i, st.dw[0]), e1 = sel3(j, st.dw[1]), e2 = sel3(k, st.dw[2]);
        const float wm = w0 * w1 * w2 * mass;
        const v3 dpos = mk3(((float)i - st.fx[0]) * dx, ((float)j - st.fx[1]) * dx, ((float)k - st.fx[2]) * dx);
        const v3 dwt = mk3(e0 * w1 * w2, w0 * e1 * w2, w0 * w1 * e2) * inv_dx;
        add(i, j, k, (v + mvmul(Cm, dpos)) * wm + mvmul(fs, dwt), wm);
    }
}

// Segmented sum within 16-lane rows: lanes [seg_start, lane] of the caller's segment
// (consecutive lanes scattering to the same stencil) are summed into `x` of the last
// lane, with four DPP row shifts; c1..c8 are the per-step "source lane is in my
// segment" conditions.
MPM_DEV float seg_row_sum(float x, bool c1, bool c2, bool c4, bool c8) {
    float y;
    y = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xf, 0xf, true));
    x += c1 ? y : 0.0f;
    y = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x112, 0xf, 0xf, true));
    x += c2 ? y : 0.0f;
    y = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xf, 0xf, true));
    x += c4 ? y : 0.0f;
    y = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x118, 0xf, 0xf, true));
    x += c8 ? y : 0.0f;
    return x;
}

MPM_DEV bool stencil_in_grid(const mpm_grid_t& G, const Stencil& st, bool periodic) {
    bool ok = true;
#pragma unroll
    for (int a = periodic ? 1 : 0; a < 3; ++a) ok = ok && st.base[a] >= 0 && st.base[a] + 2 < G.n[a];
    return ok;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2). With G
// groups of `ppt` consecutive chunks of block-sorted particles (G from the device's chunk
// count) and q = ceil(G / 8), group g = (b % 8) * q + y is handed to workgroup b for
// y = b / 8, b / 8 + Q, ... < q (Q = launch / 8), so every XCD sweeps one contiguous eighth
// of the chunks: spatial neighbours, whose scatter flushes overlap on the same grid
// nodes, share one L2 instead of bouncing lines between L2s, and the XCDs stay balanced
// whatever the launch size (a launch larger than the count leaves workgroups that exit at
// once, a smaller one takes further sweeps). A workgroup processes the ppt chunks of a
// group one after the other (one particle per thread per chunk).
constexpr int kXcds = 8;

// amdgpu_waves_per_eu(3): keep the allocation within 168 VGPRs (3 wavefronts per SIMD; left
// alone the scheduler lands anywhere between 110 and 200 depending on unrelated edits). At
// 128 VGPRs (4 per SIMD) the fixed-point lean pass fits without spilling but runs ~4 %
// slower on the pour: the extra wavefront is not missed.
#ifndef MPM_PASS_WAVES_PER_EU
#define MPM_PASS_WAVES_PER_EU 3
#endif
template <int MODE, bool kSolid, int kT, bool kCdf, bool kDet>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(MPM_PASS_WAVES_PER_EU))) void k_particle_pass(mpm_particles_t P, mpm_grid_t G,
                                                      const mpm_model_t* __restrict__ md, float dt,
                                                      int ppt) {
    constexpr int kArena = arena_of<kT>();
    constexpr bool kG2P = (MODE & MPM_PASS_G2P) != 0;
    constexpr bool kStress = (MODE & MPM_PASS_STRESS) != 0;
    constexpr bool kP2G = (MODE & MPM_PASS_P2G) != 0;
    constexpr bool kLean = (MODE & MPM_PASS_LEAN) != 0;
    constexpr bool kKeepStress = (MODE & MPM_PASS_KEEP_STRESS) != 0;
    __shared__ float4 s_gather[kG2P ? kArena : 1];
    // scatter arena: 64-bit fixed-point sums, in the grid's global unit in the deterministic
    // mode, in a per-workgroup unit otherwise (whose flush adds floats to the grid)
    using Acc = unsigned long long;
    __shared__ Acc s_scatter[kP2G ? 4 : 1][kP2G ? kArena : 1];
    __shared__ int s_bnd[kDet ? 1 : kT / 64][2];
    __shared__ int s_red0[kT / 64][8];
    __shared__ int s_red1[kT / 64][8];
    __shared__ float s_react[kCdf && kP2G ? 6 * MPM_MAX_CDF : 1];  // blocked deposits per lane

    const int S = P.stride;
    // chunks: group-aligned runs of sorted particles (chunk_start), or kT consecutive ones
    const bool chunked = P.chunk_start != nullptr;
    const int nchunks = chunked ? *P.chunk_count : (P.n + kT - 1) / kT;
    const int xcd = (int)(blockIdx.x % kXcds), Q = (int)(gridDim.x / kXcds);
    const int qe = ((nchunks + ppt - 1) / ppt + kXcds - 1) / kXcds;  // groups per XCD
    bool first = true;
    for (int it = 0;; ++it) {
        const int y = (int)(blockIdx.x / kXcds) + (it / ppt) * Q;
        if (y >= qe) return;  // this workgroup's share is done (uniform, before any barrier)
        const int chunk = (xcd * qe + y) * ppt + it % ppt;
        if (chunk >= nchunks) continue;  // the last XCD's short tail (uniform)
        if (!first) __syncthreads();     // the previous chunk's LDS reads are done
        first = false;
        const int p0 = chunked ? P.chunk_start[chunk] : chunk * kT;
        const int p1 = chunked ? P.chunk_start[chunk + 1] : min(p0 + kT, P.n);
        const int p = p0 + threadIdx.x;
        if constexpr (kP2G && !kCdf) {
            // The whole scatter arena is cleared up front (six 16-byte stores per lane), so
            // the scatter box reduction's barrier orders the clearing before any scatter:
            // no clearing phase and barrier of its own between the two.
            for (int a = threadIdx.x; a < kArena; a += kT) {
                s_scatter[0][a] = Acc(0);
                s_scatter[1][a] = Acc(0);
                s_scatter[2][a] = Acc(0);
                s_scatter[3][a] = Acc(0);
            }
        }
        // The particle's selection, scene, material and position, and the inputs the pass
        // reads after the gather (F, the per-particle material state, mass and volume), in
        // one round of loads at the chunk's start: not one round per condition that
        // depends on the last, and not one after each of the chain's later barriers. (The
        // CPIC pass, which has less register room, loads the latter where it uses them.)
        constexpr bool kEarly = !kCdf;
        int sel_v = 1, scene_v = 0, mat_v = 0;
        v3 x_v = mk3(0.f, 0.f, 0.f);
        m33 F_v = mzero();
        float mass_v = 0.f, vol_v = 0.f, mu_v = 0.f, lam_v = 0.f, ys_v = 0.f, jp_v = 0.f;
        if (p < p1) {
            sel_v = P.selection[p];
            if (P.scene) scene_v = P.scene[p];
            mat_v = P.material[p];
            x_v = ld3(P.x, S, p);
            if constexpr (kEarly && kG2P) F_v = ld9(P.F, S, p);
            if constexpr (kEarly && kStress) {
                mu_v = P.mu[p];
                lam_v = P.lam[p];
                ys_v = P.yield_stress[p];
                jp_v = P.Jp[p];
            }
            if constexpr (kEarly && kP2G) {
                mass_v = P.mass[p];
                vol_v = P.vol[p];
            }
        }
        const bool sel = p < p1 && sel_v == 0;
        const int scene = sel ? scene_v : 0;
        const bool alive = sel && (unsigned)scene < (unsigned)G.n_scenes;  // bad scene ids: error bit 8
        const int mat = alive ? mat_v : 0;
        const size_t soff = scene_base(G, alive ? scene : 0);
        const bool periodic = md->periodic_x != 0;
        int err = (sel && !alive) ? 8 : 0;

        v3 x = alive ? x_v : mk3(0.f, 0.f, 0.f);
        v3 v = mk3(0.f, 0.f, 0.f);
        m33 Cm = mzero(), L = mzero(), Ft = mzero();

        int pc_g = 0;  // CPIC tag of the gather half (stored), 0 without CDF colliders
        if constexpr (kG2P) {
            const bool gather = alive && mat != 7 && mat != 8;
            const Stencil st = make_stencil(x, G.inv_dx);
            v3 ghost = mk3(0.f, 0.f, 0.f);
            if constexpr (kCdf) {
                if (gather) {
                    pc_g = cdf_particle_tag(G, st, P.cdf_tag[p], true);
                    P.cdf_tag[p] = pc_g;
                    if (pc_g != 0) ghost = cdf_ghost(G, md, dt, x, ld3(P.v, S, p), pc_g, st);
                }
            }
            // ghost fill (mpm_utils.py:1005-1017): incompatible nodes read the ghost
            auto subst = [&](int i, int j, int k, v3 g) {
                if constexpr (kCdf) {
                    if (pc_g != 0) {
                        const int gi = st.base[0] + i, gj = st.base[1] + j, gk = st.base[2] + k;
                        if (in_nodes(G, gi, gj, gk) && cdf_incompat(pc_g, G.cdf.tag_prev[node_index(G, gi, gj, gk)]))
                            g = ghost;
                    }
                }
                return g;
            };
            int lo[3], hi[3];
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = gather ? st.base[a] : INT_MAX;
                hi[a] = gather ? st.base[a] + 3 : INT_MIN;
            }
            const Tile tg = make_tile<kT>(G, lo, hi, gather ? scene : INT_MAX, gather ? scene : INT_MIN, s_red0, periodic);
            {
                const float4* __restrict__ out = reinterpret_cast<const float4*>(G.node_out);
                tile_for<kT>(tg, [&](int a, int gi, int gj, int gk) {
                    const long o = in_grid(G, gi, gj, gk) ? node_addr(G, tg.soff, gi, gj, gk) : -1L;
                    s_gather[a] = o >= 0 ? out[o] : make_float4(0.f, 0.f, 0.f, 0.f);
                });
            }
            __syncthreads();
            if (gather) {
                if (!stencil_in_grid(G, st, periodic)) err |= 1;
                v3 nv;
                m33 nC, nF;
                if (tg.lds) {
                    const int o0 = st.base[0] - tg.lo[0], o1 = st.base[1] - tg.lo[1], o2 = st.base[2] - tg.lo[2];
                    const int d1 = tg.d1, d2 = tg.d2, b0 = tg.base;
                    gather_tp(st, G.inv_dx,
                              [&](int i, int j, int k) {
                                  const float4 g = s_gather[b0 + ((o0 + i) * d1 + (o1 + j)) * d2 + (o2 + k)];
                                  return subst(i, j, k, mk3(g.x, g.y, g.z));
                              },
                              nv, nC, nF);
                } else {
                    // the same tensor-product sums, fed from global memory: a particle's
                    // gather is bitwise the same whichever path its chunk takes
                    const float4* __restrict__ gout = reinterpret_cast<const float4*>(G.node_out);
                    gather_tp<true>(st, G.inv_dx,
                              [&](int i, int j, int k) {
                                  int ix = st.base[0] + i;
                                  if (periodic) ix = ((ix % G.n[0]) + G.n[0]) % G.n[0];
                                  const int iy = st.base[1] + j, iz = st.base[2] + k;
                                  v3 g = mk3(0.f, 0.f, 0.f);
                                  const long na = in_grid(G, ix, iy, iz) ? node_addr(G, soff, ix, iy, iz) : -1L;
                                  if (na >= 0) {
                                      const float4 o4 = gout[na];
                                      g = mk3(o4.x, o4.y, o4.z);
                                  }
                                  return subst(i, j, k, g);
                              },
                              nv, nC, nF);
                }
                v = nv;
                Cm = nC;
                L = nF;
                x = x + v * dt;
                if (periodic) {
                    const float lx = (float)G.n[0] * G.dx;
                    if (x.x < 0.0f) x.x = x.x + lx;
                    if (x.x >= lx) x.x = x.x - lx;
                }
                m33 A = mscale(nF, dt);
                A.a[0] += 1.f; A.a[4] += 1.f; A.a[8] += 1.f;
                Ft = mmul(A, kEarly ? F_v : ld9(P.F, S, p));
                if (md->update_cov_with_F && P.cov) {  // update_cov, mpm_utils.py:613-633
                    float c[6];
    #pragma unroll
                    for (int q = 0; q < 6; ++q) c[q] = P.cov[q * S + p];
                    m33 cn;
                    cn.a[0] = c[0]; cn.a[1] = c[1]; cn.a[2] = c[2];
                    cn.a[3] = c[1]; cn.a[4] = c[3]; cn.a[5] = c[4];
                    cn.a[6] = c[2]; cn.a[7] = c[4]; cn.a[8] = c[5];
                    m33 n1 = madd(cn, mscale(madd(mmul(nF, cn), mmul(cn, mT(nF))), dt));
                    P.cov[p] = n1.a[0]; P.cov[S + p] = n1.a[1]; P.cov[2 * S + p] = n1.a[2];
                    P.cov[3 * S + p] = n1.a[4]; P.cov[4 * S + p] = n1.a[5]; P.cov[5 * S + p] = n1.a[8];
                }
                st3(P.x, S, p, x);
                if (!kLean) {
                    st3(P.v, S, p, v);
                    st9(P.C, S, p, Cm);
                    st9(P.L, S, p, L);
                    st9(P.F_trial, S, p, Ft);
                }
            } else if (alive) {
                if (kP2G) {
                    v = ld3(P.v, S, p);
                    Cm = ld9(P.C, S, p);
                }
                if (kStress) {
                    Ft = ld9(P.F_trial, S, p);
                    if (mat == 10 || mat == 12) L = ld9(P.L, S, p);
                }
            }
        } else if (alive) {
            if (kP2G) {
                v = ld3(P.v, S, p);
                Cm = ld9(P.C, S, p);
            }
            if (kStress) {
                Ft = ld9(P.F_trial, S, p);
                if (mat == 10 || mat == 12) L = ld9(P.L, S, p);
            }
        }

        m33 stress = mzero();
        if constexpr (kStress) {
            if (alive) {
                PState ps;
                ps.mu = kEarly ? mu_v : P.mu[p];
                ps.lam = kEarly ? lam_v : P.lam[p];
                ps.yield = kEarly ? ys_v : P.yield_stress[p];
                ps.Jp = kEarly ? jp_v : P.Jp[p];
                ps.dirty_mu = ps.dirty_yield = ps.dirty_Jp = false;
                m33 F;
                if constexpr (kSolid) {
                    stress_update(Ft, L, mat, md, dt, ps, F, stress);
                } else {
                    if (!stress_update_viscous(Ft, L, mat, md, ps, F, stress)) err |= 4;
                }
                st9(P.F, S, p, F);
                if (!kLean || kKeepStress) st9(P.stress, S, p, stress);
                if (ps.dirty_yield) P.yield_stress[p] = ps.yield;
                if (ps.dirty_mu) { P.mu[p] = ps.mu; P.lam[p] = ps.lam; }
                if (ps.dirty_Jp) P.Jp[p] = ps.Jp;
            }
        } else if (kP2G && alive) {
            stress = ld9(P.stress, S, p);
        }

        if constexpr (kP2G) {
            const Stencil st = make_stencil(x, G.inv_dx);
            int lo[3], hi[3];
    #pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = alive ? st.base[a] : INT_MAX;
                hi[a] = alive ? st.base[a] + 3 : INT_MIN;
            }
            const Tile ts = make_tile<kT>(G, lo, hi, alive ? scene : INT_MAX, alive ? scene : INT_MIN, s_red1, periodic);
            int pc_s = 0;  // CPIC tag of the scatter half (mpm_utils.py:851-860)
            if constexpr (kCdf) {
                tile_for<kT>(ts, [&](int a, int, int, int) {
                    s_scatter[0][a] = Acc(0);
                    s_scatter[1][a] = Acc(0);
                    s_scatter[2][a] = Acc(0);
                    s_scatter[3][a] = Acc(0);
                });
                if (threadIdx.x < 6 * MPM_MAX_CDF) s_react[threadIdx.x] = 0.f;
                if (alive) {
                    const int old = (kG2P && mat != 7 && mat != 8) ? pc_g : P.cdf_tag[p];
                    pc_s = cdf_particle_tag(G, st, old, false);
                }
                __syncthreads();
            }
            // blocked deposits (mpm_utils.py:894-912): a contribution to an incompatible
            // node goes to its owner lane's reaction wrench instead of the grid. Each lane
            // sums its own in registers (force rf, torque rt, for owner r_own; a change of
            // owner, rare, adds the running sum to LDS first), and after the scatter each
            // wavefront adds one DPP-summed wrench per owner: every lane adding to the same
            // six LDS words per node serialised the wavefront's float atomics.
            int r_own = -1;
            v3 rf = mk3(0.f, 0.f, 0.f), rt = mk3(0.f, 0.f, 0.f);
            auto react_to_lds = [&]() {
                float* r = s_react + 6 * r_own;
                atomicAdd(r + 0, rf.x);
                atomicAdd(r + 1, rf.y);
                atomicAdd(r + 2, rf.z);
                atomicAdd(r + 3, rt.x);
                atomicAdd(r + 4, rt.y);
                atomicAdd(r + 5, rt.z);
            };
            auto mask = [&](int i, int j, int k, v3& add, float& wm) {
                if constexpr (kCdf) {
                    if (pc_s != 0) {
                        const int gi = st.base[0] + i, gj = st.base[1] + j, gk = st.base[2] + k;
                        if (in_nodes(G, gi, gj, gk)) {
                            const int nt = G.cdf.tag[node_index(G, gi, gj, gk)];
                            if (cdf_incompat(pc_s, nt)) {
                                const int own = (nt >> MPM_CDF_OWNER_SHIFT) & 3;
                                const float* c = G.cdf.lanes->center[own];
                                const v3 xw = mk3((float)gi * G.dx, (float)gj * G.dx, (float)gk * G.dx);
                                const v3 tq = cross(xw - mk3(c[0], c[1], c[2]), add);
                                if (own != r_own) {
                                    if (r_own >= 0) react_to_lds();
                                    r_own = own;
                                    rf = mk3(0.f, 0.f, 0.f);
                                    rt = mk3(0.f, 0.f, 0.f);
                                }
                                rf = rf + add;
                                rt = rt + tq;
                                add = mk3(0.f, 0.f, 0.f);
                                wm = 0.f;
                                return false;
                            }
                        }
                    }
                }
                return true;
            };
            m33 fs = mzero();
            float mass = 0.0f;
            if (alive) {
                const float r = md->rpic_damping;
                if (r != 0.0f) {
                    const m33 Ct = mT(Cm);
                    Cm = madd(mscale(Cm, 1.0f - r), mscale(msub(Cm, Ct), r / 2.0f));
                    if (r < -0.001f) Cm = mzero();
                }
                fs = mscale(stress, -(kEarly ? vol_v : P.vol[p]) * dt);
                mass = kEarly ? mass_v : P.mass[p];
                if (!stencil_in_grid(G, st, periodic)) err |= 1;
            }
            // Float mode: the LDS sums are fixed point too (float atomic adds on LDS are
            // several times slower than 64-bit integer ones), in units 2^(46 - e) chosen per
            // workgroup from bounds below 2^e on its particles' contributions, one for
            // momentum and one for mass: a contribution stays below 2^46 units, a 16-lane
            // segment sum below 2^50 and the workgroup's sum on a node below 2^54. With
            // |w|, |dpos| / (1.5 dx), |grad w| dx all at most 1, every momentum component
            // is below m (|v|_1 + 1.5 dx sum|C|) + sum|fs| / dx. A workgroup with a bound
            // that is not finite scatters in float through global memory.
            float unit_p = 0.f, unit_m = 0.f;
            double inv_p = 0.0, inv_m = 0.0;
            bool fin = true;
            if constexpr (!kDet) {
                float bp = 0.f, bm = 0.f;
                if (alive) {
                    float sc = 0.f, sf = 0.f;
#pragma unroll
                    for (int e = 0; e < 9; ++e) {
                        sc += fabsf(Cm.a[e]);
                        sf += fabsf(fs.a[e]);
                    }
                    bm = fabsf(mass);
                    bp = bm * (fabsf(v.x) + fabsf(v.y) + fabsf(v.z) + 1.5f * G.dx * sc) + G.inv_dx * sf;
                }
                const int wp = wave_extreme<true>(__float_as_int(bp)), wq = wave_extreme<true>(__float_as_int(bm));
                if ((threadIdx.x & 63u) == 0) {
                    s_bnd[threadIdx.x >> 6][0] = wp;
                    s_bnd[threadIdx.x >> 6][1] = wq;
                }
                __syncthreads();
                int mp = s_bnd[0][0], mq = s_bnd[0][1];
#pragma unroll
                for (int w = 1; w < kT / 64; ++w) {
                    mp = max(mp, s_bnd[w][0]);
                    mq = max(mq, s_bnd[w][1]);
                }
                fin = mp < 0x7f800000 && mq < 0x7f800000;
                const int ep = max(((mp >> 23) & 0xff) - 126, -80), em = max(((mq >> 23) & 0xff) - 126, -80);
                unit_p = __int_as_float((173 - ep) << 23);
                unit_m = __int_as_float((173 - em) << 23);
                inv_p = __longlong_as_double((long long)(977 + ep) << 52);
                inv_m = __longlong_as_double((long long)(977 + em) << 52);
            }
            // Lanes whose particles share a stencil base cell (consecutive after a cell
            // sort) form segments; when a wavefront has enough of them, each segment is
            // summed with DPP row shifts and only its last lane (per 16-lane row) adds
            // to LDS. Otherwise every lane adds, in lane-rotated node order.
            const int lane = (int)(threadIdx.x & 63);
            const int o0 = st.base[0] - ts.lo[0], o1 = st.base[1] - ts.lo[1], o2 = st.base[2] - ts.lo[2];
            const int cellkey = (ts.lds && alive) ? (o0 * ts.d1 + o1) * ts.d2 + o2 : -1 - lane;
            const int prevkey = __shfl_up(cellkey, 1, 64);
            const bool head = lane == 0 || prevkey != cellkey;
            const unsigned long long heads = __ballot(head);
            // (never in the deterministic mode: a float pre-sum would depend on the grouping)
            const bool segmented = !kDet && ts.lds && fin && __popcll(~heads) >= 16;  // wavefront-uniform
            if (segmented) {
                const unsigned long long upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
                const int seg0 = 63 - __clzll(heads & upto);
                const int rl = lane & 15;
                const bool c1 = rl >= 1 && lane - 1 >= seg0, c2 = rl >= 2 && lane - 2 >= seg0;
                const bool c4 = rl >= 4 && lane - 4 >= seg0, c8 = rl >= 8 && lane - 8 >= seg0;
                const bool tail = rl == 15 || lane == 63 || ((heads >> (lane + 1)) & 1ull);
                const int d1 = ts.d1, d2 = ts.d2, b0 = ts.base;
                scatter_rotated(st, G.dx, G.inv_dx, mass, v, Cm, fs, 0, [&](int i, int j, int k, v3 add, float wm) {
                    if constexpr (kCdf) {
                        if (alive) mask(i, j, k, add, wm);
                    }
                    const float s0 = seg_row_sum(add.x, c1, c2, c4, c8);
                    const float s1 = seg_row_sum(add.y, c1, c2, c4, c8);
                    const float s2 = seg_row_sum(add.z, c1, c2, c4, c8);
                    const float s3 = seg_row_sum(wm, c1, c2, c4, c8);
                    if (tail && alive) {
                        const int a = b0 + ((o0 + i) * d1 + (o1 + j)) * d2 + (o2 + k);
                        if constexpr (!kDet) {
                            atomicAdd(&s_scatter[0][a], fix_round(s0 * unit_p));
                            atomicAdd(&s_scatter[1][a], fix_round(s1 * unit_p));
                            atomicAdd(&s_scatter[2][a], fix_round(s2 * unit_p));
                            atomicAdd(&s_scatter[3][a], fix_round(s3 * unit_m));
                        }
                    }
                });
            } else if (alive && kDet && ts.lds) {
                // tile address of node (i, j, k): row base per (i, j), k in the offset
                const int d2 = ts.d2, d12 = ts.d1 * ts.d2;
                const int r0 = ts.base + (o0 * ts.d1 + o1) * d2 + o2;
                const bool ok = scatter_tp_fix(st, G.dx, G.inv_dx, mass, v, Cm, fs, G.fix_scale[0], G.fix_scale[1],
                                               lane, d12, d2, [&](int, int, int, int off, unsigned long long q0,
                                                   unsigned long long q1, unsigned long long q2, unsigned long long q3) {
                    const int a = r0 + off;
                    Acc* sc = &s_scatter[0][0];
                    atomicAdd(sc + a, q0);
                    atomicAdd(sc + kArena + a, q1);
                    atomicAdd(sc + 2 * kArena + a, q2);
                    atomicAdd(sc + 3 * kArena + a, q3);
                });
                if (!ok) err |= 64;
            } else if (alive) {
                if (ts.lds && fin) {
                    const int d1 = ts.d1, d2 = ts.d2, b0 = ts.base;
                    scatter_rotated<true>(st, G.dx, G.inv_dx, mass, v, Cm, fs, (int)(threadIdx.x % 27u),
                                    [&](int i, int j, int k, v3 add, float wm) {
                                        if (!mask(i, j, k, add, wm)) return;
                                        const int a = b0 + ((o0 + i) * d1 + (o1 + j)) * d2 + (o2 + k);
                                        atomicAdd(&s_scatter[0][a], fix_round(add.x * unit_p));
                                        atomicAdd(&s_scatter[1][a], fix_round(add.y * unit_p));
                                        atomicAdd(&s_scatter[2][a], fix_round(add.z * unit_p));
                                        atomicAdd(&s_scatter[3][a], fix_round(wm * unit_m));
                                    });
                } else {
                    scatter_global<kDet>(G, soff, st, periodic, mass, v, Cm, fs, mask, err);
                }
            }
            if constexpr (kCdf) {  // all lanes converged: the wavefront's wrench per owner
                if (__ballot(r_own >= 0)) {
#pragma unroll 1
                    for (int o = 0; o < MPM_MAX_CDF; ++o) {
                        const bool mine = r_own == o;
                        if (!__ballot(mine)) continue;
                        const float w6[6] = {wave_sum(mine ? rf.x : 0.f), wave_sum(mine ? rf.y : 0.f),
                                             wave_sum(mine ? rf.z : 0.f), wave_sum(mine ? rt.x : 0.f),
                                             wave_sum(mine ? rt.y : 0.f), wave_sum(mine ? rt.z : 0.f)};
                        if ((threadIdx.x & 63u) == 0) {
#pragma unroll
                            for (int q = 0; q < 6; ++q) atomicAdd(s_react + 6 * o + q, w6[q]);
                        }
                    }
                }
            }
            __syncthreads();
            {
                float* __restrict__ gin = G.node_in;
                unsigned long long* __restrict__ gfix = reinterpret_cast<unsigned long long*>(G.node_fix);
                tile_for<kT>(ts, [&](int a, int gi, int gj, int gk) {
                    const Acc a0 = s_scatter[0][a], a1 = s_scatter[1][a], a2 = s_scatter[2][a], a3 = s_scatter[3][a];
                    if (a0 == Acc(0) && a1 == Acc(0) && a2 == Acc(0) && a3 == Acc(0)) return;
                    if (!in_grid(G, gi, gj, gk)) return;
                    const long na = node_addr(G, ts.soff, gi, gj, gk);
                    if (na < 0) {  // a block without a pool slot: the contribution is dropped
                        err |= 512;
                        return;
                    }
                    const size_t o = 4 * (size_t)na;
                    if constexpr (kDet) {
                        atomicAdd(gfix + o + 0, a0);
                        atomicAdd(gfix + o + 1, a1);
                        atomicAdd(gfix + o + 2, a2);
                        atomicAdd(gfix + o + 3, a3);
                    } else {
                        atomicAdd(gin + o + 0, (float)((double)(long long)a0 * inv_p));
                        atomicAdd(gin + o + 1, (float)((double)(long long)a1 * inv_p));
                        atomicAdd(gin + o + 2, (float)((double)(long long)a2 * inv_p));
                        atomicAdd(gin + o + 3, (float)((double)(long long)a3 * inv_m));
                    }
                });
                mark_tile_blocks<kT>(G, ts);
            }
            if constexpr (kCdf) {
                if (threadIdx.x < 6 * MPM_MAX_CDF && s_react[threadIdx.x] != 0.f)
                    atomicAdd(G.cdf.reaction + threadIdx.x, s_react[threadIdx.x]);
            }
        }
        if (alive && !(x.x == x.x && x.y == x.y && x.z == x.z)) err |= 2;
        if (err && G.error_flag) atomicOr(G.error_flag, err);
    }  // chunk loop
}

// Paired lean pass (MPM_PASS_PAIR, the deterministic interior substep): a workgroup takes
// a group-aligned chunk of up to 512 sorted particles, two per lane (p0 + t and p0 + 256 +
// t). The chunk's latency-bound phases (particle loads, the two box reductions with their
// barriers, the tile load and the flush) are paid once for both particles, and the ~700
// chunks of a 340k-particle pour fit one round of resident workgroups (3 per CU) instead
// of two. The first particle's scatter inputs wait in an LDS stash while the second one
// gathers; the gather tile and the scatter arena are never live together and share one
// buffer, so LDS stays at ~50 KB per workgroup. The per-particle arithmetic is that of
// k_particle_pass (the same helpers), and the fixed-point contributions are integers, so
// the result is bit-identical to the single-particle pass.
constexpr int kPairStash = 25;  // x 3, v 3, C 9, fs 9, mass 1
template <bool kSolid, bool kKeepStress>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MPM_PASS_WAVES_PER_EU))) void k_particle_pass_pair(
    mpm_particles_t P, mpm_grid_t G, const mpm_model_t* __restrict__ md, float dt) {
    constexpr int kT = 256;
    constexpr int kArena = arena_of<kT>();
    // the gather tile (kArena float4) lives in the first half of the scatter arena
    __shared__ unsigned long long s_arena[4 * kArena];
    __shared__ float s_stash[kPairStash][kT];
    __shared__ int s_red0[kT / 64][8];
    __shared__ int s_red1[kT / 64][8];
    float4* s_gather = reinterpret_cast<float4*>(s_arena);

    const int S = P.stride;
    const int nchunks = *P.chunk_count;
    // XCD-contiguous shares of the chunks, swept as in k_particle_pass
    const int xcd = (int)(blockIdx.x % kXcds), Q = (int)(gridDim.x / kXcds);
    const int qe = (nchunks + kXcds - 1) / kXcds;
    for (int y = (int)(blockIdx.x / kXcds); y < qe; y += Q) {
    const int chunk = xcd * qe + y;
    if (chunk >= nchunks) break;  // the last XCD's short tail (uniform)
    if (y != (int)(blockIdx.x / kXcds)) __syncthreads();  // the previous chunk's LDS reads are done
    const int p0 = P.chunk_start[chunk], p1 = P.chunk_start[chunk + 1];
    const bool periodic = md->periodic_x != 0;
    int err = 0;

    int pp[2], mat[2], scene[2];
    bool alive[2];
    v3 x[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int p = p0 + h * kT + (int)threadIdx.x;
        int sel_v = 1, scene_v = 0, mat_v = 0;  // one round of loads, as in the single pass
        v3 x_v = mk3(0.f, 0.f, 0.f);
        if (p < p1) {
            sel_v = P.selection[p];
            if (P.scene) scene_v = P.scene[p];
            mat_v = P.material[p];
            x_v = ld3(P.x, S, p);
        }
        const bool sel = p < p1 && sel_v == 0;
        scene[h] = sel ? scene_v : 0;
        alive[h] = sel && (unsigned)scene[h] < (unsigned)G.n_scenes;  // bad scene ids: error bit 8
        if (sel && !alive[h]) err |= 8;
        mat[h] = alive[h] ? mat_v : 0;
        x[h] = alive[h] ? x_v : mk3(0.f, 0.f, 0.f);
        pp[h] = p;
    }

    // gather tile over both particles' stencils
    int lo[3], hi[3], slo = INT_MAX, shi = INT_MIN;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = INT_MAX;
        hi[a] = INT_MIN;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (alive[h] && mat[h] != 7 && mat[h] != 8) {
            const Stencil st = make_stencil(x[h], G.inv_dx);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], st.base[a]);
                hi[a] = max(hi[a], st.base[a] + 3);
            }
            slo = min(slo, scene[h]);
            shi = max(shi, scene[h]);
        }
    }
    const Tile tg = make_tile<kT>(G, lo, hi, slo, shi, s_red0, periodic);
    {
        const float4* __restrict__ out = reinterpret_cast<const float4*>(G.node_out);
        tile_for<kT>(tg, [&](int a, int gi, int gj, int gk) {
            const long o = in_grid(G, gi, gj, gk) ? node_addr(G, tg.soff, gi, gj, gk) : -1L;
            s_gather[a] = o >= 0 ? out[o] : make_float4(0.f, 0.f, 0.f, 0.f);
        });
    }
    __syncthreads();

    // gather + return map + stress of one particle (k_particle_pass's lean fused mode);
    // leaves its scatter inputs (new x, v, C after RPIC damping, fs = -vol dt tau, mass)
    auto front = [&](int h, v3& v, m33& Cm, m33& fs, float& mass) __attribute__((always_inline)) {
        const int p = pp[h];
        v = mk3(0.f, 0.f, 0.f);
        Cm = mzero();
        fs = mzero();
        mass = 0.0f;
        if (!alive[h]) return;
        const int mt = mat[h];
        m33 Ft, L = mzero();
        if (mt != 7 && mt != 8) {
            const Stencil st = make_stencil(x[h], G.inv_dx);
            if (!stencil_in_grid(G, st, periodic)) err |= 1;
            v3 nv;
            m33 nC, nF;
            if (tg.lds) {
                const int o0 = st.base[0] - tg.lo[0], o1 = st.base[1] - tg.lo[1], o2 = st.base[2] - tg.lo[2];
                const int d1 = tg.d1, d2 = tg.d2, b0 = tg.base;
                gather_tp(st, G.inv_dx,
                          [&](int i, int j, int k) {
                              const float4 g = s_gather[b0 + ((o0 + i) * d1 + (o1 + j)) * d2 + (o2 + k)];
                              return mk3(g.x, g.y, g.z);
                          },
                          nv, nC, nF);
            } else {
                const float4* __restrict__ gout = reinterpret_cast<const float4*>(G.node_out);
                const size_t soff = scene_base(G, scene[h]);
                gather_tp<true>(st, G.inv_dx,
                                [&](int i, int j, int k) {
                                    int ix = st.base[0] + i;
                                    if (periodic) ix = ((ix % G.n[0]) + G.n[0]) % G.n[0];
                                    const int iy = st.base[1] + j, iz = st.base[2] + k;
                                    const long na = in_grid(G, ix, iy, iz) ? node_addr(G, soff, ix, iy, iz) : -1L;
                                    v3 g = mk3(0.f, 0.f, 0.f);
                                    if (na >= 0) {
                                        const float4 o4 = gout[na];
                                        g = mk3(o4.x, o4.y, o4.z);
                                    }
                                    return g;
                                },
                                nv, nC, nF);
            }
            v = nv;
            Cm = nC;
            L = nF;
            v3 xn = x[h] + v * dt;
            if (periodic) {
                const float lx = (float)G.n[0] * G.dx;
                if (xn.x < 0.0f) xn.x = xn.x + lx;
                if (xn.x >= lx) xn.x = xn.x - lx;
            }
            x[h] = xn;
            m33 A = mscale(nF, dt);
            A.a[0] += 1.f; A.a[4] += 1.f; A.a[8] += 1.f;
            Ft = mmul(A, ld9(P.F, S, p));
            if (md->update_cov_with_F && P.cov) {  // update_cov, mpm_utils.py:613-633
                float c[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) c[q] = P.cov[q * S + p];
                m33 cn;
                cn.a[0] = c[0]; cn.a[1] = c[1]; cn.a[2] = c[2];
                cn.a[3] = c[1]; cn.a[4] = c[3]; cn.a[5] = c[4];
                cn.a[6] = c[2]; cn.a[7] = c[4]; cn.a[8] = c[5];
                const m33 n1 = madd(cn, mscale(madd(mmul(nF, cn), mmul(cn, mT(nF))), dt));
                P.cov[p] = n1.a[0]; P.cov[S + p] = n1.a[1]; P.cov[2 * S + p] = n1.a[2];
                P.cov[3 * S + p] = n1.a[4]; P.cov[4 * S + p] = n1.a[5]; P.cov[5 * S + p] = n1.a[8];
            }
            st3(P.x, S, p, xn);
        } else {  // stationary / rigid ids: no gather; their stored v, C, F_trial scatter
            v = ld3(P.v, S, p);
            Cm = ld9(P.C, S, p);
            Ft = ld9(P.F_trial, S, p);
        }
        PState ps;
        ps.mu = P.mu[p];
        ps.lam = P.lam[p];
        ps.yield = P.yield_stress[p];
        ps.Jp = P.Jp[p];
        ps.dirty_mu = ps.dirty_yield = ps.dirty_Jp = false;
        m33 F, stress;
        if constexpr (kSolid) {
            stress_update(Ft, L, mt, md, dt, ps, F, stress);
        } else {
            if (!stress_update_viscous(Ft, L, mt, md, ps, F, stress)) err |= 4;
        }
        st9(P.F, S, p, F);
        if constexpr (kKeepStress) st9(P.stress, S, p, stress);
        if (ps.dirty_yield) P.yield_stress[p] = ps.yield;
        if (ps.dirty_mu) { P.mu[p] = ps.mu; P.lam[p] = ps.lam; }
        if (ps.dirty_Jp) P.Jp[p] = ps.Jp;
        const float r = md->rpic_damping;
        if (r != 0.0f) {
            const m33 Ct = mT(Cm);
            Cm = madd(mscale(Cm, 1.0f - r), mscale(msub(Cm, Ct), r / 2.0f));
            if (r < -0.001f) Cm = mzero();
        }
        fs = mscale(stress, -P.vol[p] * dt);
        mass = P.mass[p];
    };

    const int t = (int)threadIdx.x;
    {  // first particle: its scatter inputs wait in the stash (read back by this lane only)
        v3 v;
        m33 Cm, fs;
        float mass;
        front(0, v, Cm, fs, mass);
        s_stash[0][t] = x[0].x; s_stash[1][t] = x[0].y; s_stash[2][t] = x[0].z;
        s_stash[3][t] = v.x; s_stash[4][t] = v.y; s_stash[5][t] = v.z;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            s_stash[6 + e][t] = Cm.a[e];
            s_stash[15 + e][t] = fs.a[e];
        }
        s_stash[24][t] = mass;
    }
    v3 vB;
    m33 CB, fsB;
    float mB;
    front(1, vB, CB, fsB, mB);

    // scatter tile over both particles' new stencils (its barrier also ends every read of
    // the gather tile, whose memory the arena reuses)
    slo = INT_MAX;
    shi = INT_MIN;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = INT_MAX;
        hi[a] = INT_MIN;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (alive[h]) {
            const Stencil st = make_stencil(x[h], G.inv_dx);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = min(lo[a], st.base[a]);
                hi[a] = max(hi[a], st.base[a] + 3);
            }
            slo = min(slo, scene[h]);
            shi = max(shi, scene[h]);
        }
    }
    const Tile ts = make_tile<kT>(G, lo, hi, slo, shi, s_red1, periodic);
    tile_for<kT>(ts, [&](int a, int, int, int) {
        s_arena[a] = 0ull;
        s_arena[kArena + a] = 0ull;
        s_arena[2 * kArena + a] = 0ull;
        s_arena[3 * kArena + a] = 0ull;
    });
    __syncthreads();
    const int lane = t & 63;
    auto scatter = [&](int h, v3 v, const m33& Cm, const m33& fs, float mass) __attribute__((always_inline)) {
        if (!alive[h]) return;
        const Stencil st = make_stencil(x[h], G.inv_dx);
        if (!stencil_in_grid(G, st, periodic)) err |= 1;
        if (ts.lds) {
            const int o0 = st.base[0] - ts.lo[0], o1 = st.base[1] - ts.lo[1], o2 = st.base[2] - ts.lo[2];
            const int d2 = ts.d2, d12 = ts.d1 * ts.d2;
            const int r0 = ts.base + (o0 * ts.d1 + o1) * d2 + o2;
            const bool ok = scatter_tp_fix(st, G.dx, G.inv_dx, mass, v, Cm, fs, G.fix_scale[0], G.fix_scale[1], lane,
                                           d12, d2, [&](int, int, int, int off, unsigned long long q0,
                                                        unsigned long long q1, unsigned long long q2,
                                                        unsigned long long q3) {
                                               const int a = r0 + off;
                                               atomicAdd(s_arena + a, q0);
                                               atomicAdd(s_arena + kArena + a, q1);
                                               atomicAdd(s_arena + 2 * kArena + a, q2);
                                               atomicAdd(s_arena + 3 * kArena + a, q3);
                                           });
            if (!ok) err |= 64;
        } else {
            scatter_global<true>(G, scene_base(G, scene[h]), st, periodic, mass, v, Cm, fs,
                                 [](int, int, int, v3&, float&) { return true; }, err);
        }
    };
    scatter(1, vB, CB, fsB, mB);
    {
        v3 v = mk3(s_stash[3][t], s_stash[4][t], s_stash[5][t]);
        m33 Cm, fs;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            Cm.a[e] = s_stash[6 + e][t];
            fs.a[e] = s_stash[15 + e][t];
        }
        scatter(0, v, Cm, fs, s_stash[24][t]);
    }
    __syncthreads();
    {
        unsigned long long* __restrict__ gfix = reinterpret_cast<unsigned long long*>(G.node_fix);
        tile_for<kT>(ts, [&](int a, int gi, int gj, int gk) {
            const unsigned long long a0 = s_arena[a], a1 = s_arena[kArena + a], a2 = s_arena[2 * kArena + a],
                                     a3 = s_arena[3 * kArena + a];
            if ((a0 | a1 | a2 | a3) == 0ull) return;
            if (!in_grid(G, gi, gj, gk)) return;
            const long na = node_addr(G, ts.soff, gi, gj, gk);
            if (na < 0) {  // a block without a pool slot: the contribution is dropped
                err |= 512;
                return;
            }
            const size_t o = 4 * (size_t)na;
            atomicAdd(gfix + o + 0, a0);
            atomicAdd(gfix + o + 1, a1);
            atomicAdd(gfix + o + 2, a2);
            atomicAdd(gfix + o + 3, a3);
        });
        mark_tile_blocks<kT>(G, ts);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
        if (alive[h] && !(x[h].x == x[h].x && x[h].y == x[h].y && x[h].z == x[h].z)) err |= 2;
    if (err && G.error_flag) atomicOr(G.error_flag, err);
    }  // chunk sweep
}

inline int32_t launch_status() { return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP; }

}  // namespace

extern "C" int32_t mpm_particle_pass(const mpm_particles_t* P, const mpm_grid_t* G, const mpm_model_t* model,
                                     float dt, int32_t mode, void* stream) {
    if (!P || !G || !model || P->n < 0 || P->stride < P->n || !P->x || !P->v || !P->F || !P->F_trial || !P->C ||
        !P->stress || !P->L || !P->vol || !P->mass || !P->mu || !P->lam || !P->yield_stress || !P->Jp ||
        !P->material || !P->selection || !G->node_in || !G->node_out || G->n[0] <= 0 || G->n[1] <= 0 ||
        G->n[2] <= 0 || G->n_scenes < 1 || (G->n_scenes > 1 && !P->scene) || G->nb[0] != (G->n[0] + 3) / 4 || G->nb[1] != (G->n[1] + 3) / 4 ||
        G->nb[2] != (G->n[2] + 3) / 4)
        return MPM_ERR_ARG;
    const bool cdf = G->cdf.n_cdf > 0;
    const bool det = (mode & MPM_PASS_DETERMINISTIC) != 0;
    if (det && (cdf || !G->node_fix || !(G->fix_scale[0] > 0.0f) || !(G->fix_scale[1] > 0.0f)))
        return MPM_ERR_ARG;
    if (G->block_slot && (cdf || G->pool_blocks < 1)) return MPM_ERR_ARG;
    if (cdf && (G->n_scenes != 1 || !P->cdf_tag || !G->cdf.tag || !G->cdf.tag_prev || !G->cdf.dist ||
                !G->cdf.dist_prev || !G->cdf.block || !G->cdf.lanes || !G->cdf.reaction || (mode & MPM_PASS_LEAN)))
        return MPM_ERR_ARG;
    if (P->n == 0) return MPM_OK;
    if (mode & MPM_PASS_PAIR) {  // two particles per lane over chunks of up to 512
        const int lean = MPM_PASS_G2P | MPM_PASS_STRESS | MPM_PASS_P2G | MPM_PASS_LEAN;
        const int mm = mode & (lean | MPM_PASS_KEEP_STRESS);
        if (!det || cdf || (mm != lean && mm != (lean | MPM_PASS_KEEP_STRESS)) ||
            (mode & (MPM_PASS_WG128 | MPM_PASS_PPT2 | MPM_PASS_PPT4)) || !P->chunk_start || !P->chunk_count ||
            P->chunk_capacity < 1)
            return MPM_ERR_ARG;
        const int q = (P->chunk_capacity + kXcds - 1) / kXcds;
        const dim3 grid((unsigned)(q * kXcds)), block(256);
        const bool solid = (mode & MPM_PASS_VISCOUS_ONLY) == 0, keep = (mm & MPM_PASS_KEEP_STRESS) != 0;
        hipStream_t s = (hipStream_t)stream;
        if (solid && keep)
            hipLaunchKernelGGL((k_particle_pass_pair<true, true>), grid, block, 0, s, *P, *G, model, dt);
        else if (solid)
            hipLaunchKernelGGL((k_particle_pass_pair<true, false>), grid, block, 0, s, *P, *G, model, dt);
        else if (keep)
            hipLaunchKernelGGL((k_particle_pass_pair<false, true>), grid, block, 0, s, *P, *G, model, dt);
        else
            hipLaunchKernelGGL((k_particle_pass_pair<false, false>), grid, block, 0, s, *P, *G, model, dt);
        return launch_status();
    }
    const int wg = (!cdf && !det && (mode & MPM_PASS_WG128)) ? 128 : 256;
    const int ppt = (mode & MPM_PASS_PPT4) ? 4 : ((mode & MPM_PASS_PPT2) ? 2 : 1);
    hipStream_t s = (hipStream_t)stream;
    // chunk_start holds chunks of up to 256 particles: used by 256-lane workgroups only
    mpm_particles_t Pl = *P;
    if (wg != 256 || (Pl.chunk_start && (!Pl.chunk_count || Pl.chunk_capacity < 1))) {
        if (Pl.chunk_start && wg == 256) return MPM_ERR_ARG;
        Pl.chunk_start = nullptr;
    }
    const int nchunks = Pl.chunk_start ? Pl.chunk_capacity : (P->n + wg - 1) / wg;
    const int ngroups = (nchunks + ppt - 1) / ppt;
    const int q = (ngroups + kXcds - 1) / kXcds;
    const dim3 grid((unsigned)(q * kXcds)), block(wg);
    const bool solid = (mode & MPM_PASS_VISCOUS_ONLY) == 0;
    const int m = mode & (MPM_PASS_G2P | MPM_PASS_STRESS | MPM_PASS_P2G | MPM_PASS_LEAN | MPM_PASS_KEEP_STRESS);
#define LAUNCH(M, T, C, D)                                                                                     \
    if (solid)                                                                                                 \
        hipLaunchKernelGGL((k_particle_pass<(M), true, T, C, D>), grid, block, 0, s, Pl, *G, model, dt, ppt); \
    else                                                                                                       \
        hipLaunchKernelGGL((k_particle_pass<(M), false, T, C, D>), grid, block, 0, s, Pl, *G, model, dt, ppt);
#define PLAIN(M)                         \
    if (wg == 128) {                     \
        LAUNCH(M, 128, false, false)     \
    } else {                             \
        LAUNCH(M, 256, false, false)     \
    }
// CPIC colliders: non-lean modes, 256-thread workgroups; deterministic scatter: the
// modes that scatter, 256-thread workgroups (each variant instantiated only where used)
#define CDF_(M) if (cdf) { LAUNCH(M, 256, true, false) break; }
#define DET_(M) if (det) { LAUNCH(M, 256, false, true) break; }
#define CASE_CD(M) case (M): CDF_(M) DET_(M) PLAIN(M) break;
#define CASE_C(M) case (M): CDF_(M) PLAIN(M) break;  /* no scatter: the deterministic bit is moot */
#define CASE_D(M) case (M): if (cdf) return MPM_ERR_ARG; DET_(M) PLAIN(M) break;
    switch (m) {
        CASE_C(MPM_PASS_G2P)
        CASE_C(MPM_PASS_STRESS)
        CASE_CD(MPM_PASS_P2G)
        CASE_CD(MPM_PASS_STRESS | MPM_PASS_P2G)
        CASE_CD(MPM_PASS_G2P | MPM_PASS_STRESS | MPM_PASS_P2G)
        CASE_D(MPM_PASS_G2P | MPM_PASS_STRESS | MPM_PASS_P2G | MPM_PASS_LEAN)
        CASE_D(MPM_PASS_G2P | MPM_PASS_STRESS | MPM_PASS_P2G | MPM_PASS_LEAN | MPM_PASS_KEEP_STRESS)
        default: return MPM_ERR_ARG;
    }
#undef CDF_
#undef DET_
#undef CASE_CD
#undef CASE_C
#undef CASE_D
#undef LAUNCH
#undef PLAIN
    return launch_status();
}
