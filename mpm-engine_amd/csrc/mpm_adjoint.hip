// Hand-written adjoint of the explicit MLS-MPM/APIC substep (gfx950).
//
// The forward substep k maps the particle state S_k = (x, v, C, F_trial) to S_{k+1}:
//   stress:  (F_k, tau_k) = return map of F_trial_k        (stress_update_particle,
//                                                           mpm_utils.py:1056-1149)
//   P2G:     m_i, mv_i += w m (v + C (x_i - x)) - dt V tau grad w   (p2g, :819-916)
//   grid:    u_i = mv_i / m_i + dt g, then the colliders in order    (:927-941, BCs)
//   G2P:     v' = sum w u, C' = 4/dx^2 sum w u (x_i - x)^T, L = sum u grad w^T,
//            x' = x + dt v', F_trial' = (I + dt L) F_k               (g2p, :944-1043)
// The reverse substep runs the transposes in the opposite order: the G2P adjoint is a
// scatter into the grid adjoint (global float atomics), the grid adjoint is per node,
// and the P2G adjoint is a gather, fused with the return-map adjoint. Derivatives through
// the B-spline weights use the weight Hessian, so gradients with respect to positions
// are exact for the discrete map (between cell crossings of the stencil base).
//
// Return-map adjoints (von Mises, material 1, with or without hardening, and local mu(I),
// material 9, both with Hencky elasticity): F_trial = U S V^T; the stress tau = U diag(h) U^T and the
// corrected F = U diag(g) V^T are isotropic functions of F_trial, and both maps scale
// the deviatoric log-strain by one factor k, so they share the assembly below. Their
// derivatives are assembled in the SVD frame from the diagonal Jacobians dh/dsigma, dg/dsigma and, per pair (a, b), the divided differences
// (h_a - h_b)/(s_a - s_b) and (g_a - g_b)/(s_a - s_b) plus the sums over (s_a + s_b).
// The divided differences are evaluated through log-ratios with expm1 and a series
// below |log(s_a/s_b)| < 1e-3, so repeated or nearly repeated singular values (F near a
// rotation, the undeformed state) give the finite limit instead of 1/(s_a^2 - s_b^2).
// Drucker-Prager sand (material 2) shares it too (k = -c tr / |e - me| on the cone, 0 in
// tension). Fixed-corotated elasticity (material 0) has no return map; its stress adjoint goes
// through the left stretch U S U^T, whose SVD-frame derivative has only the sums
// (s_a + s_b) in its denominators. The generalized-Newtonian fluid (material 10) adds
// the adjoint of the velocity gradient L, which flows to the previous substep's G2P.
#include <hip/hip_runtime.h>

#include "mpm_device.h"
#include "mpm_tile.h"
#include "mpm_adjoint_math.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxAdjColliders = 16;

// error bits of the adjoint (OR-ed into grid.error_flag)
constexpr int kErrAdjMaterial = 16;  // a material without an adjoint
constexpr int kErrAdjCollider = 32;  // a collider kind without an adjoint

// dimensionless B-spline data of one stencil node: weight, d/dfx and the Hessian
struct NodeW {
    float w;
    v3 dw;      // gradient of w with respect to fx (x = dx * fx)
    m33 H;      // Hessian of w with respect to fx
    v3 q;       // node offset (i - fx), (x_i - x_p) / dx
};

MPM_DEV float sel3f(int i, const float* a) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

MPM_DEV NodeW node_w(const Stencil& st, int i, int j, int k) {
    const float w0 = sel3f(i, st.w[0]), w1 = sel3f(j, st.w[1]), w2 = sel3f(k, st.w[2]);
    const float d0 = sel3f(i, st.dw[0]), d1 = sel3f(j, st.dw[1]), d2 = sel3f(k, st.dw[2]);
    const float dd0 = i == 1 ? -2.0f : 1.0f, dd1 = j == 1 ? -2.0f : 1.0f, dd2 = k == 1 ? -2.0f : 1.0f;
    NodeW n;
    n.w = w0 * w1 * w2;
    n.dw = mk3(d0 * w1 * w2, w0 * d1 * w2, w0 * w1 * d2);
    n.H.a[0] = dd0 * w1 * w2; n.H.a[4] = w0 * dd1 * w2; n.H.a[8] = w0 * w1 * dd2;
    n.H.a[1] = n.H.a[3] = d0 * d1 * w2;
    n.H.a[2] = n.H.a[6] = d0 * w1 * d2;
    n.H.a[5] = n.H.a[7] = w0 * d1 * d2;
    n.q = mk3((float)i - st.fx[0], (float)j - st.fx[1], (float)k - st.fx[2]);
    return n;
}

MPM_DEV bool node_of(const mpm_grid_t& G, const Stencil& st, int n, int& i, int& j, int& k, size_t& o) {
    i = n / 9;
    j = (n / 3) % 3;
    k = n % 3;
    const int gi = st.base[0] + i, gj = st.base[1] + j, gk = st.base[2] + k;
    if ((unsigned)gi >= (unsigned)G.n[0] || (unsigned)gj >= (unsigned)G.n[1] || (unsigned)gk >= (unsigned)G.n[2])
        return false;
    o = 4 * (size_t)node_index(G, gi, gj, gk);
    return true;
}

MPM_DEV m33 outer(v3 a, v3 b) {
    m33 r;
    r.a[0] = a.x * b.x; r.a[1] = a.x * b.y; r.a[2] = a.x * b.z;
    r.a[3] = a.y * b.x; r.a[4] = a.y * b.y; r.a[5] = a.y * b.z;
    r.a[6] = a.z * b.x; r.a[7] = a.z * b.y; r.a[8] = a.z * b.z;
    return r;
}

// ---------------------------------------------------------------- G2P adjoint
// One workgroup of kAT particles stages the grid velocity over its stencil box in LDS
// (read once per node instead of 27 times per particle) and accumulates the scattered
// grid-velocity adjoint in an LDS arena, flushed with one global float atomic per node
// and component; chunks whose box does not fit (unsorted particles) gather and scatter
// through global memory. L_k comes from the forward's G2P (P.L).
constexpr int kAT = 256;

// G2P adjoint of one particle over its 27 stencil nodes, in tensor-product form. With
// per-axis weights w_a, derivatives d_a, second derivatives h_a and offsets q_a (node
// index minus fx) of the node (i, j, k), the forward gather's weight is w = w0 w1 w2, its
// gradient dw = (d0 w1 w2, w0 d1 w2, w0 w1 d2) and Hessian H (H00 = h0 w1 w2, H01 = d0 d1 w2,
// ...); the adjoint of the node velocity u and of the particle position are
//   ub  = w (vt + c4 Cb q) + Lb dw / dx,
//   xb += dw (u.vt + c4 u.(Cb q)) / dx - c4 / dx Cb^T (sum_nodes w u) + H (Lb^T u) / dx^2,
// where the second term, linear in u, is summed over the nodes first and multiplied by
// Cb^T once, and the products shared by a row (i, j) of nodes are formed once per row.
// Nodes go in the residue-class order (at step (ti, tj, tk) the lane of rank rho visits
// class ((ti, tj, tk) + decode(rho)) mod 3), so lanes of different rank add to
// different LDS addresses in the same instruction. The LDS sums of ub are 64-bit fixed
// point (ds_add_u64) in a unit chosen per workgroup from a bound on its particles' ub:
// float atomic adds on LDS cost about as much as the rest of the node loop three times
// over, the integer adds a small part of it, and the workgroup's sums no longer depend on
// the order of its adds. A workgroup whose bound is not finite (a NaN or infinity in
// the adjoints) adds in float through global memory, so the NaN reaches the grid.
__global__ __launch_bounds__(kAT) void k_adj_g2p(mpm_particles_t P, mpm_grid_t G, mpm_adjoint_t A,
                                                 const mpm_model_t* __restrict__ md, float dt) {
    constexpr int kArena = arena_of<kAT>();
    __shared__ float4 s_u[kArena];
    __shared__ unsigned long long s_ub[3][kArena];
    __shared__ int s_red[kAT / 64][8];
    __shared__ int s_bound[kAT / 64];
    (void)md;
    const int p = blockIdx.x * kAT + threadIdx.x;
    const int S = P.stride;
    const bool sel = p < P.n && P.selection[p] == 0;
    const int mat = sel ? P.material[p] : 0;
    const bool moving = sel && mat != 7 && mat != 8;
    if (sel && !moving) {  // no gather for stationary / rigid ids: x, F pass through
        st9(A.F, S, p, ld9(A.F_trial, S, p));
        st3(A.v, S, p, mk3(0.f, 0.f, 0.f));
        st9(A.C, S, p, mzero());
    }
    const float inv_dx = G.inv_dx;
    const v3 x = moving ? ld3(P.x, S, p) : mk3(0.f, 0.f, 0.f);
    const Stencil st = make_stencil(x, inv_dx);
    int lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = moving ? st.base[a] : INT_MAX;
        hi[a] = moving ? st.base[a] + 3 : INT_MIN;
    }
    const Tile tl = make_tile<kAT>(G, lo, hi, moving ? 0 : INT_MAX, moving ? 0 : INT_MIN, s_red, false);
    const float c4 = 4.0f * inv_dx;
    m33 Cb1 = mzero(), Lb = mzero(), Fb = mzero();
    v3 xb1 = mk3(0.f, 0.f, 0.f), vt = mk3(0.f, 0.f, 0.f);
    float bound = 0.f;
    if (moving) {
        const m33 F = ld9(P.F, S, p);
        const m33 L = ld9(P.L, S, p);
        xb1 = ld3(A.x, S, p);
        const v3 vb1 = ld3(A.v, S, p);
        Cb1 = ld9(A.C, S, p);
        const m33 Ftb1 = ld9(A.F_trial, S, p);
        vt = vb1 + xb1 * dt;  // x' = x + dt v'
        // F_trial' = (I + dt L) F, plus the adjoint of L from the next substep's stress
        Lb = madd(mscale(mmul(Ftb1, mT(F)), dt), ld9(A.L, S, p));
        m33 IL = mscale(L, dt);
        IL.a[0] += 1.f; IL.a[4] += 1.f; IL.a[8] += 1.f;
        Fb = mmul(mT(IL), Ftb1);
        // every component of every ub of this particle is below this (w, |dw| <= 1, |q| <= 1.5)
        float sc = 0.f, sl = 0.f;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            sc += fabsf(Cb1.a[e]);
            sl += fabsf(Lb.a[e]);
        }
        bound = fabsf(vt.x) + fabsf(vt.y) + fabsf(vt.z) + 1.5f * c4 * sc + inv_dx * sl;
    }
    // the workgroup's largest bound, as float bits (order-preserving for values >= 0; NaN
    // and infinity sort above every finite value)
    const int wmax = wave_extreme<true>(__float_as_int(bound));
    if ((threadIdx.x & 63u) == 0) s_bound[threadIdx.x >> 6] = wmax;
    const float4* __restrict__ uo = reinterpret_cast<const float4*>(G.node_out);
    tile_for<kAT>(tl, [&](int a, int gi, int gj, int gk) {
        s_u[a] = in_grid(G, gi, gj, gk) ? uo[node_index(G, gi, gj, gk)] : make_float4(0.f, 0.f, 0.f, 0.f);
        s_ub[0][a] = 0ull;
        s_ub[1][a] = 0ull;
        s_ub[2][a] = 0ull;
    });
    __syncthreads();
    int bmax = s_bound[0];
#pragma unroll
    for (int w = 1; w < kAT / 64; ++w) bmax = max(bmax, s_bound[w]);
    // fixed-point unit 2^(51 - e) with bound < 2^e: each contribution stays below 2^51 units
    // and 256 particles' sum on one node below 2^59
    const int ex = max(((bmax >> 23) & 0xff) - 126, -75);
    const float unit = __int_as_float((178 - ex) << 23);
    const double inv_unit = __longlong_as_double((long long)(972 + ex) << 52);
    const bool fixed = tl.lds && bmax < 0x7f800000;
    if (moving) {
        const float k1 = inv_dx, k2 = c4 * inv_dx;
        const m33 Lbs = mscale(Lb, inv_dx), Lb2 = mscale(Lb, inv_dx * inv_dx);
        float* __restrict__ gub = A.grid_v;
        // rotated per-axis stencil values: slot t of axis a is stencil index (t + r0_a) mod 3
        const int rho = (int)(threadIdx.x % 27u);
        const int r0[3] = {((rho / 9 - st.base[0]) % 3 + 3) % 3, (((rho / 3) % 3 - st.base[1]) % 3 + 3) % 3,
                           ((rho % 3 - st.base[2]) % 3 + 3) % 3};
        int ix[3][3];
        float wv[3][3], dv[3][3], hv[3][3], qv[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                const int r = r0[a] + t >= 3 ? r0[a] + t - 3 : r0[a] + t;
                ix[a][t] = r;
                wv[a][t] = sel3f(r, st.w[a]);
                dv[a][t] = sel3f(r, st.dw[a]);
                hv[a][t] = r == 1 ? -2.0f : 1.0f;
                qv[a][t] = (float)r - st.fx[a];
            }
        }
        v3 xb = mk3(0.f, 0.f, 0.f), su = mk3(0.f, 0.f, 0.f);
        auto nodes = [&](auto lds_tag) {
            constexpr bool kLds = decltype(lds_tag)::value;
            const int o0 = st.base[0] - tl.lo[0], o1 = st.base[1] - tl.lo[1], o2 = st.base[2] - tl.lo[2];
#pragma unroll
            for (int ti = 0; ti < 3; ++ti) {
                const int i = ix[0][ti];
                const float w0 = wv[0][ti], d0 = dv[0][ti], h0 = hv[0][ti];
                const v3 Cq0 = mk3(Cb1.a[0], Cb1.a[3], Cb1.a[6]) * qv[0][ti];
#pragma unroll
                for (int tj = 0; tj < 3; ++tj) {
                    const int j = ix[1][tj];
                    const float w1 = wv[1][tj], d1 = dv[1][tj], h1 = hv[1][tj];
                    const float w01 = w0 * w1, d0w1 = d0 * w1, w0d1 = w0 * d1, d0d1 = d0 * d1, h0w1 = h0 * w1,
                                w0h1 = w0 * h1;
                    const v3 Cq01 = fma3(mk3(Cb1.a[1], Cb1.a[4], Cb1.a[7]), qv[1][tj], Cq0);
                    // Lb dw / dx = (Lbs e0 d0w1 + Lbs e1 w0d1) w2 + Lbs e2 w01 d2
                    const v3 La = mk3(Lbs.a[0], Lbs.a[3], Lbs.a[6]) * d0w1 + mk3(Lbs.a[1], Lbs.a[4], Lbs.a[7]) * w0d1;
                    const v3 Lc = mk3(Lbs.a[2], Lbs.a[5], Lbs.a[8]) * w01;
                    const int row = kLds ? tl.base + ((o0 + i) * tl.d1 + (o1 + j)) * tl.d2 + o2 : 0;
#pragma unroll
                    for (int tk = 0; tk < 3; ++tk) {
                        const int k = ix[2][tk];
                        const float w2 = wv[2][tk], d2 = dv[2][tk], h2 = hv[2][tk];
                        float4 u4;
                        int a = 0;
                        size_t o = 0;
                        if constexpr (kLds) {
                            a = row + k;
                            u4 = s_u[a];
                        } else {
                            const int gi = st.base[0] + i, gj = st.base[1] + j, gk = st.base[2] + k;
                            if (!in_grid(G, gi, gj, gk)) continue;
                            o = 4 * (size_t)node_index(G, gi, gj, gk);
                            u4 = uo[o / 4];
                        }
                        const v3 u = mk3(u4.x, u4.y, u4.z);
                        const float w = w01 * w2;
                        const v3 dw = mk3(d0w1 * w2, w0d1 * w2, w01 * d2);
                        const v3 Cq = fma3(mk3(Cb1.a[2], Cb1.a[5], Cb1.a[8]), qv[2][tk], Cq01);
                        const float T = k1 * dot(u, vt) + k2 * dot(u, Cq);
                        xb = fma3(dw, T, xb);
                        su = fma3(u, w, su);
                        // H (Lb^T u) / dx^2, H symmetric
                        const v3 lv = mTvmul(Lb2, u);
                        const float H00 = h0w1 * w2, H11 = w0h1 * w2, H22 = w01 * h2;
                        const float H01 = d0d1 * w2, H02 = d0w1 * d2, H12 = w0d1 * d2;
                        xb = xb + mk3(H00 * lv.x + H01 * lv.y + H02 * lv.z, H01 * lv.x + H11 * lv.y + H12 * lv.z,
                                      H02 * lv.x + H12 * lv.y + H22 * lv.z);
                        const v3 ub = fma3(La, w2, Lc * d2) + fma3(Cq, c4, vt) * w;
                        if constexpr (kLds) {
                            atomicAdd(&s_ub[0][a], fix_round(ub.x * unit));
                            atomicAdd(&s_ub[1][a], fix_round(ub.y * unit));
                            atomicAdd(&s_ub[2][a], fix_round(ub.z * unit));
                        } else {
                            atomicAdd(gub + o + 0, ub.x);
                            atomicAdd(gub + o + 1, ub.y);
                            atomicAdd(gub + o + 2, ub.z);
                        }
                    }
                }
            }
        };
        if (fixed) nodes(std::true_type{});
        else nodes(std::false_type{});
        // the weights' term linear in u, summed first: - c4 / dx Cb^T (sum w u)
        xb = xb - mTvmul(Cb1, su) * k2;
        st3(A.x, S, p, xb1 + xb);
        st9(A.F, S, p, Fb);
    }
    __syncthreads();
    float* __restrict__ gub = A.grid_v;
    tile_for<kAT>(tl, [&](int a, int gi, int gj, int gk) {
        const long long q0 = (long long)s_ub[0][a], q1 = (long long)s_ub[1][a], q2 = (long long)s_ub[2][a];
        if ((q0 | q1 | q2) == 0 || !in_grid(G, gi, gj, gk)) return;
        const size_t o = 4 * (size_t)node_index(G, gi, gj, gk);
        atomicAdd(gub + o + 0, (float)((double)q0 * inv_unit));
        atomicAdd(gub + o + 1, (float)((double)q1 * inv_unit));
        atomicAdd(gub + o + 2, (float)((double)q2 * inv_unit));
    });
}

// --------------------------------------------------------------- grid adjoint
// One node: recompute the forward chain (normalise + gravity, damping, colliders in
// order), then transpose it. Plate / box impulses add their adjoint through `impb`.
// kMaxC bounds ncoll: at kMaxC = 4 both collider loops unroll and the per-collider
// velocities stay in registers; the 16-collider variant keeps them in scratch.
template <int kMaxC>
MPM_DEV void adj_node(const mpm_grid_t& G, const mpm_model_t* md, const mpm_collider_t* col, int ncoll,
                      const float* impb, int nslots, int gi, int gj, int gk, float4 in4, v3 ub, float dt, float4& out,
                      int& err) {
    constexpr int kUnroll = kMaxC <= 4 ? kMaxC : 1;
    const v3 xn = mk3((float)gi * G.dx, (float)gj * G.dx, (float)gk * G.dx);  // rounded, as the forward
    const float m = in4.w;
    v3 us[kMaxC];
    v3 uo = mk3(0.f, 0.f, 0.f);
    const v3 mv = mk3(in4.x, in4.y, in4.z);
    if (m > 1e-15f) uo = mk3(mv.x / m + dt * md->gravity[0], mv.y / m + dt * md->gravity[1], mv.z / m + dt * md->gravity[2]);
    const float damp = md->grid_v_damping_scale;
    if (damp < 1.0f) uo = uo * damp;
    bool act[kMaxC];
#pragma unroll kUnroll
    for (int q = 0; q < kMaxC; ++q) {
        if (q >= ncoll) break;
        const mpm_collider_t& c = col[q];
        us[q] = uo;
        act[q] = c.active && m > 0.0f && gi >= c.lo[0] && gi < c.hi[0] && gj >= c.lo[1] && gj < c.hi[1] &&
                 gk >= c.lo[2] && gk < c.hi[2];
        if (act[q]) {
            // forward collider rules of the grid update (mpm_kernels.hip apply_collider)
            const float time = c.time;
            const bool live = time >= c.start_time && time < c.end_time;
            if (c.kind == MPM_COLLIDER_BOX) {
                const v3 vb = mk3(c.velocity[0], c.velocity[1], c.velocity[2]);
                if (live) {
                    const float ox = xn.x - c.point[0], oy = xn.y - c.point[1], oz = xn.z - c.point[2];
                    if (fabsf(ox) < c.size[0] && fabsf(oy) < c.size[1] && fabsf(oz) < c.size[2]) uo = vb;
                } else if (c.reset == 1 && time < c.end_time + 15.0f * c.dt) {
                    uo = mk3(0.f, 0.f, 0.f);
                }
            } else if (c.kind == MPM_COLLIDER_PLANE) {
                if (live) {
                    const v3 n = mk3(c.normal[0], c.normal[1], c.normal[2]);
                    const v3 off = xn - mk3(c.point[0], c.point[1], c.point[2]);
                    if (dot(off, n) < 0.0f) {
                        if (c.surface == 0) uo = mk3(0.f, 0.f, 0.f);
                        else if (c.surface == 11) err |= kErrAdjCollider;
                        else {
                            const float nc = dot(uo, n);
                            v3 w = (c.surface == 1) ? uo - n * nc : uo - n * fminf(nc, 0.0f);
                            if (nc < 0.0f && len(w) > 1e-20f) w = normalize(w) * fmaxf(0.0f, len(w) + nc * c.friction);
                            uo = w;
                        }
                    }
                }
            } else if (c.kind == MPM_COLLIDER_WALLS) {
                if (live) {
                    const int pad = 3;
                    if (gi < pad && uo.x < 0) uo.x = 0.0f;
                    if (gi >= G.n[0] - pad && uo.x > 0) uo.x = 0.0f;
                    if (gj < pad && uo.y < 0) uo.y = 0.0f;
                    if (gj >= G.n[1] - pad && uo.y > 0) uo.y = 0.0f;
                    if (gk < pad && uo.z < 0) uo.z = 0.0f;
                    if (gk >= G.n[2] - pad && uo.z > 0) uo.z = 0.0f;
                }
            } else if (c.kind == MPM_COLLIDER_POINTCLOUD) {
                if (live && c.occupancy[((size_t)gi * G.n[1] + gj) * G.n[2] + gk] == 1) uo = mk3(0.f, 0.f, 0.f);
            } else if (c.kind == MPM_COLLIDER_CUP || c.kind == MPM_COLLIDER_SDF) {
                SurfaceHit h;
                if (live && surface_hit(c, G, gi, gj, gk, h)) uo = surface_apply(h, c.friction, uo);
            } else {
                err |= kErrAdjCollider;
            }
        }
    }
    // reverse sweep
    float mb = 0.0f;
#pragma unroll kUnroll
    for (int q = kMaxC - 1; q >= 0; --q) {
        if (q >= ncoll || !act[q]) continue;
        const mpm_collider_t& c = col[q];
        const v3 ui = us[q];
        const float time = c.time;
        const bool live = time >= c.start_time && time < c.end_time;
        if (c.kind == MPM_COLLIDER_BOX) {
            const v3 vb = mk3(c.velocity[0], c.velocity[1], c.velocity[2]);
            if (live) {
                const float ox = xn.x - c.point[0], oy = xn.y - c.point[1], oz = xn.z - c.point[2];
                if (fabsf(ox) < c.size[0] && fabsf(oy) < c.size[1] && fabsf(oz) < c.size[2]) {
                    v3 ib = mk3(0.f, 0.f, 0.f);
                    if (c.slot >= 0 && c.slot < nslots)
                        ib = mk3(impb[6 * c.slot], impb[6 * c.slot + 1], impb[6 * c.slot + 2]);
                    mb += dot(ib, ui - vb);  // impulse = m (u_in - v_box)
                    ub = ib * m;
                }
            } else if (c.reset == 1 && time < c.end_time + 15.0f * c.dt) {
                ub = mk3(0.f, 0.f, 0.f);
            }
        } else if (c.kind == MPM_COLLIDER_PLANE) {
            if (!live) continue;
            const v3 n = mk3(c.normal[0], c.normal[1], c.normal[2]);
            const v3 off = xn - mk3(c.point[0], c.point[1], c.point[2]);
            if (!(dot(off, n) < 0.0f)) continue;
            if (c.surface == 0) {
                ub = mk3(0.f, 0.f, 0.f);
            } else {
                // w = u - n (u.n) (slip) or u - n min(u.n, 0) (separable); for u.n < 0
                // both are the tangential part, then Coulomb: w_hat max(0, |w| + f u.n)
                const float nc = dot(ui, n);
                const v3 Pub = ub - n * dot(n, ub);
                if (nc >= 0.0f) {
                    if (c.surface == 1) ub = Pub;
                } else {
                    const v3 w = ui - n * nc;
                    const float wl = len(w);
                    if (!(wl > 1e-20f)) {
                        ub = Pub;
                    } else if (wl + nc * c.friction <= 0.0f) {
                        ub = mk3(0.f, 0.f, 0.f);
                    } else {
                        const float sc = 1.0f + c.friction * nc / wl;
                        const float a = dot(w, ub);
                        ub = Pub * sc + n * (c.friction * a / wl) - w * (c.friction * nc * a / (wl * wl * wl));
                    }
                }
            }
        } else if (c.kind == MPM_COLLIDER_WALLS) {
            if (!live) continue;
            const int pad = 3;
            if ((gi < pad && ui.x < 0) || (gi >= G.n[0] - pad && ui.x > 0)) ub.x = 0.0f;
            if ((gj < pad && ui.y < 0) || (gj >= G.n[1] - pad && ui.y > 0)) ub.y = 0.0f;
            if ((gk < pad && ui.z < 0) || (gk >= G.n[2] - pad && ui.z > 0)) ub.z = 0.0f;
        } else if (c.kind == MPM_COLLIDER_POINTCLOUD) {
            if (live && c.occupancy[((size_t)gi * G.n[1] + gj) * G.n[2] + gk] == 1) ub = mk3(0.f, 0.f, 0.f);
        } else if (c.kind == MPM_COLLIDER_CUP || c.kind == MPM_COLLIDER_SDF) {
            SurfaceHit h;
            if (!live || !surface_hit(c, G, gi, gj, gk, h)) continue;
            // impulse m (u_in - u_new) into the force slot, r x impulse into the torque slot
            v3 ib = mk3(0.f, 0.f, 0.f), tb = mk3(0.f, 0.f, 0.f);
            if (c.slot >= 0 && c.slot < nslots) {
                ib = mk3(impb[6 * c.slot], impb[6 * c.slot + 1], impb[6 * c.slot + 2]);
                tb = mk3(impb[6 * c.slot + 3], impb[6 * c.slot + 4], impb[6 * c.slot + 5]);
            }
            const v3 pb = ib + cross(tb, h.rel);  // adjoint of the impulse
            mb += dot(pb, ui - surface_apply(h, c.friction, ui));
            ub = pb * m + surface_apply_T(h, c.friction, ui, ub - pb * m);
        }
    }
    if (damp < 1.0f) ub = ub * damp;
    out = make_float4(0.f, 0.f, 0.f, 0.f);
    if (m > 1e-15f) {
        const float inv = 1.0f / m;
        out.x = ub.x * inv;
        out.y = ub.y * inv;
        out.z = ub.z * inv;
        out.w = mb - dot(ub, mv) * inv * inv;
    }
}

template <int kMaxC>
__global__ __launch_bounds__(kThreads) void k_adj_grid(mpm_grid_t G, const mpm_model_t* __restrict__ md,
                                                       const mpm_collider_t* __restrict__ col, int ncoll,
                                                       const float* __restrict__ impb, int nslots, mpm_adjoint_t A,
                                                       const int32_t* __restrict__ list,
                                                       const int32_t* __restrict__ count, float dt) {
    const int nall = G.nb[0] * G.nb[1] * G.nb[2];
    const int nblk = list ? *count : nall;
    const int lane = threadIdx.x & 63;
    int err = 0;
    for (int e = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); e < nblk; e += gridDim.x * (kThreads / 64)) {
        const int b = list ? list[e] : e;
        if ((unsigned)b >= (unsigned)nall) continue;
        int bi, bj, bk;
        decode_block(G, b, bi, bj, bk);
        const int gi = bi * 4 + (lane >> 4), gj = bj * 4 + ((lane >> 2) & 3), gk = bk * 4 + (lane & 3);
        const size_t o = 4 * ((size_t)b * 64 + lane);
        if (gi >= G.n[0] || gj >= G.n[1] || gk >= G.n[2]) continue;
        const float4 in4 = *reinterpret_cast<const float4*>(G.node_in + o);
        float4 ub4 = *reinterpret_cast<const float4*>(A.grid_v + o);
        // An empty node (m = 0) has no collider acting and a zero adjoint: skip the
        // chain (and its per-collider spill) there; whole empty wavefronts skip it.
        float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
        if (in4.w > 0.0f)
            adj_node<kMaxC>(G, md, col, ncoll, impb, nslots, gi, gj, gk, in4, mk3(ub4.x, ub4.y, ub4.z), dt, out,
                            err);
        *reinterpret_cast<float4*>(A.grid_mv + o) = out;
        *reinterpret_cast<float4*>(A.grid_v + o) = make_float4(0.f, 0.f, 0.f, 0.f);  // consumed
    }
    if (err && G.error_flag) atomicOr(G.error_flag, err);
}

// ----------------------------------------------------- P2G + return-map adjoint
__global__ __launch_bounds__(kThreads) void k_adj_p2g_stress(mpm_particles_t P, mpm_grid_t G, mpm_adjoint_t A,
                                                             const mpm_model_t* __restrict__ md, float dt) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n || P.selection[p] != 0) return;
    const int S = P.stride;
    const int mat = P.material[p];
    const float inv_dx = G.inv_dx, dx = G.dx;
    const v3 x = ld3(P.x, S, p);
    const v3 v = ld3(P.v, S, p);
    m33 Cm = ld9(P.C, S, p);
    const float r = md->rpic_damping;
    if (r != 0.0f) {
        Cm = madd(mscale(Cm, 1.0f - r), mscale(msub(Cm, mT(Cm)), r / 2.0f));
        if (r < -0.001f) Cm = mzero();
    }
    const m33 tau = ld9(P.stress, S, p);
    const float mp = P.mass[p], V0 = P.vol[p];
    const m33 fs = mscale(tau, -V0 * dt);
    const Stencil st = make_stencil(x, inv_dx);
    const float4* __restrict__ gmv = reinterpret_cast<const float4*>(A.grid_mv);
    v3 vb = mk3(0.f, 0.f, 0.f), xb = ld3(A.x, S, p), wa = vb;
    m33 Cb = mzero(), fsb = mzero();
#pragma unroll
    for (int n = 0; n < 27; ++n) {
        int i, j, k;
        size_t o;
        if (!node_of(G, st, n, i, j, k, o)) continue;
        const float4 ab = gmv[o / 4];
        const v3 a = mk3(ab.x, ab.y, ab.z);
        const NodeW nw = node_w(st, i, j, k);
        wa = wa + a * nw.w;
        Cb = madd(Cb, outer(a * nw.w, nw.q));
        fsb = madd(fsb, outer(a, nw.dw));
        const v3 vi = v + mvmul(Cm, nw.q * dx);
        const float s = mp * dot(a, vi) + ab.w * mp;
        xb = xb + nw.dw * (inv_dx * s) + mvmul(nw.H, mTvmul(fs, a)) * (inv_dx * inv_dx);
    }
    vb = wa * mp;
    xb = xb - mTvmul(Cm, wa) * mp;
    Cb = mscale(Cb, mp * dx);
    if (r != 0.0f) Cb = r < -0.001f ? mzero() : msub(mscale(Cb, 1.0f - 0.5f * r), mscale(mT(Cb), 0.5f * r));
    const m33 tb = mscale(fsb, -V0 * dt * inv_dx);
    st3(A.x, S, p, xb);
    st3(A.v, S, p, vb);
    st9(A.C, S, p, Cb);
    // return map + stress adjoint
    const m33 Fb = ld9(A.F, S, p);
    m33 Ftb;
    m33 Lb = mzero();
    if (mat == 0) {
        float mub = 0.f, lamb = 0.f;
        Ftb = adj_fcr(ld9(P.F_trial, S, p), P.mu[p], P.lam[p], Fb, tb, mub, lamb);
        A.mu[p] += mub;
        A.lam[p] += lamb;
    } else if (mat == 2) {
        float mub = 0.f, lamb = 0.f;
        Ftb = adj_sand(ld9(P.F_trial, S, p), P.mu[p], P.lam[p], md->alpha[mat], Fb, tb, mub, lamb);
        A.mu[p] += mub;
        A.lam[p] += lamb;
    } else if (mat == 1) {
        // A.yield_stress holds the adjoint of the yield stress after this substep (hardened
        // or passed on) and leaves with the adjoint of the one this substep read
        float mub = 0.f, lamb = 0.f, yb = 0.f;
        const float xi = md->hardening[mat] > 0.5f ? md->xi[mat] : 0.0f;
        Ftb = adj_von_mises(ld9(P.F_trial, S, p), P.mu[p], P.lam[p], P.yield_stress[p], xi, A.yield_stress[p], Fb,
                            tb, mub, lamb, yb);
        A.mu[p] += mub;
        A.lam[p] += lamb;
        A.yield_stress[p] = yb;
    } else if (mat == 9) {
        float mub = 0.f, lamb = 0.f, pb[4] = {0.f, 0.f, 0.f, 0.f};
        Ftb = adj_mu_i(ld9(P.F_trial, S, p), md, mat, dt, P.mu[p], P.lam[p], Fb, tb, mub, lamb, pb);
        A.mu[p] += mub;
        A.lam[p] += lamb;
        A.mui_mu_s[p] += pb[0];
        A.mui_dmu[p] += pb[1];
        A.mui_I0[p] += pb[2];
        A.mui_d[p] += pb[3];
    } else if (mat == 10) {
        float eb = 0.f, bb = 0.f, yb = 0.f, kb = 0.f, nb = 0.f;
        Ftb = adj_viscous(ld9(P.F_trial, S, p), ld9(P.L, S, p), md->bulk[mat], md->plastic_viscosity[mat],
                          P.yield_stress[p], md->hardening[mat], md->softening[mat], Fb, tb, Lb, eb, bb, yb, kb, nb);
        A.visc[p] += eb;
        A.bulk[p] += bb;
        A.yield_stress[p] += yb;
        A.hb_K[p] += kb;
        A.hb_n[p] += nb;
    } else {
        Ftb = mzero();
        if (G.error_flag) atomicOr(G.error_flag, kErrAdjMaterial);
    }
    st9(A.F_trial, S, p, Ftb);
    st9(A.L, S, p, Lb);
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }
inline int32_t launch_status() { return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP; }

inline bool adj_ok(const mpm_adjoint_t* A) {
    return A && A->x && A->v && A->C && A->F_trial && A->F && A->stress && A->mu && A->lam && A->yield_stress &&
           A->grid_v && A->grid_mv && A->L && A->visc && A->bulk && A->hb_K && A->hb_n && A->mui_mu_s &&
           A->mui_dmu && A->mui_I0 && A->mui_d;
}

inline bool grid_ok1(const mpm_grid_t* G) {
    return G && G->node_in && G->node_out && G->n_scenes == 1 && !G->block_slot && G->n[0] > 0 && G->n[1] > 0 &&
           G->n[2] > 0 && G->nb[0] == (G->n[0] + 3) / 4 && G->nb[1] == (G->n[1] + 3) / 4 &&
           G->nb[2] == (G->n[2] + 3) / 4;
}

}  // namespace

extern "C" {

int32_t mpm_adjoint_g2p(const mpm_particles_t* P, const mpm_grid_t* G, const mpm_adjoint_t* A,
                        const mpm_model_t* model, float dt, void* stream) {
    if (!P || !grid_ok1(G) || !adj_ok(A) || !model || P->n < 0 || P->stride < P->n || !P->x || !P->F || !P->L ||
        !P->material || !P->selection)
        return MPM_ERR_ARG;
    if (P->n == 0) return MPM_OK;
    hipLaunchKernelGGL(k_adj_g2p, dim3((unsigned)((P->n + kAT - 1) / kAT)), dim3(kAT), 0, (hipStream_t)stream, *P, *G,
                       *A, model, dt);
    return launch_status();
}

int32_t mpm_adjoint_grid(const mpm_grid_t* G, const mpm_model_t* model, const int32_t* block_list,
                         const int32_t* block_count, int32_t capacity, const mpm_collider_t* colliders,
                         int32_t ncoll, const float* impulse_adj, int32_t wrench_slots, const mpm_adjoint_t* A,
                         float dt, void* stream) {
    if (!grid_ok1(G) || !model || !adj_ok(A) || ncoll < 0 || (ncoll > 0 && !colliders) || wrench_slots < 0 ||
        (wrench_slots > 0 && !impulse_adj) || (block_list && (!block_count || capacity < 1)))
        return MPM_ERR_ARG;
    if (ncoll > kMaxAdjColliders) return MPM_ERR_SIZE;
    const long nblk = block_list ? (long)capacity : (long)G->nb[0] * G->nb[1] * G->nb[2];
    const unsigned nwg = (unsigned)min((nblk + 3) / 4, 2048L);
    if (ncoll <= 4)
        hipLaunchKernelGGL(k_adj_grid<4>, dim3(nwg), dim3(kThreads), 0, (hipStream_t)stream, *G, model, colliders,
                           ncoll, impulse_adj, wrench_slots, *A, block_list, block_count, dt);
    else
        hipLaunchKernelGGL(k_adj_grid<kMaxAdjColliders>, dim3(nwg), dim3(kThreads), 0, (hipStream_t)stream, *G,
                           model, colliders, ncoll, impulse_adj, wrench_slots, *A, block_list, block_count, dt);
    return launch_status();
}

int32_t mpm_adjoint_p2g_stress(const mpm_particles_t* P, const mpm_grid_t* G, const mpm_adjoint_t* A,
                               const mpm_model_t* model, float dt, void* stream) {
    if (!P || !grid_ok1(G) || !adj_ok(A) || !model || P->n < 0 || P->stride < P->n || !P->x || !P->v || !P->C ||
        !P->F_trial || !P->stress || !P->L || !P->vol || !P->mass || !P->mu || !P->lam || !P->yield_stress ||
        !P->material || !P->selection)
        return MPM_ERR_ARG;
    if (P->n == 0) return MPM_OK;
    hipLaunchKernelGGL(k_adj_p2g_stress, dim3(blocks_for(P->n)), dim3(kThreads), 0, (hipStream_t)stream, *P, *G, *A,
                       model, dt);
    return launch_status();
}

}  // extern "C"
