// CPIC thin-boundary (CDF) transfer functions for the particle pass: the particle tag
// reconstructed from node tags, colour persistence, the pair compatibility test and the
// ghost velocity used at incompatible nodes (src/warpmpm/kernels/mpm_utils.py:648-816).
//
// Node tag bits: 2l = lane l valid, 2l+1 = lane l side, [8, 10) = owner lane. Tag and
// owner-distance grids share the blocked node layout of node_in; the occupancy mask has
// one int per grid block (4x4x4 nodes), so the pretest reads at most 8 ints.
#pragma once
#include "mpm_device.h"

// cdf_pair_incompatible, mpm_utils.py:656-662: some lane valid on both, sides differ
MPM_DEV bool cdf_incompat(int pc, int nt) { return ((pc & nt & 0x55) & ((pc ^ nt) >> 1)) != 0; }

MPM_DEV bool in_nodes(const mpm_grid_t& G, int i, int j, int k) {
    return (unsigned)i < (unsigned)G.n[0] && (unsigned)j < (unsigned)G.n[1] && (unsigned)k < (unsigned)G.n[2];
}

// cdf_stencil_near, mpm_utils.py:665-683: is any block under the 3x3x3 stencil tagged
MPM_DEV bool cdf_near(const mpm_grid_t& G, const Stencil& st) {
    bool near = false;
    for (int bi = st.base[0] >> 2; bi <= (st.base[0] + 2) >> 2; ++bi)
        for (int bj = st.base[1] >> 2; bj <= (st.base[1] + 2) >> 2; ++bj)
            for (int bk = st.base[2] >> 2; bk <= (st.base[2] + 2) >> 2; ++bk) {
                if ((unsigned)bi >= (unsigned)G.nb[0] || (unsigned)bj >= (unsigned)G.nb[1] ||
                    (unsigned)bk >= (unsigned)G.nb[2])
                    continue;
                near = near || G.cdf.block[(bi * G.nb[1] + bj) * G.nb[2] + bk] != 0;
            }
    return near;
}

MPM_DEV float pick3(int i, const float* a) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

// cdf_particle_tag, mpm_utils.py:686-724: lane affinity = a valid node in the stencil;
// side = sign of sum w d side over the nodes valid for that lane (>= 0 sets the bit)
MPM_DEV int cdf_vote(const mpm_grid_t& G, const Stencil& st, bool prev) {
    const int* __restrict__ tg = prev ? G.cdf.tag_prev : G.cdf.tag;
    const float* __restrict__ dg = prev ? G.cdf.dist_prev : G.cdf.dist;
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3_ = 0.f;
    int seen = 0;
#pragma unroll 1
    for (int n = 0; n < 27; ++n) {
        const int i = n / 9, j = (n / 3) % 3, k = n % 3;
        const int gi = st.base[0] + i, gj = st.base[1] + j, gk = st.base[2] + k;
        if (!in_nodes(G, gi, gj, gk)) continue;
        const int idx = node_index(G, gi, gj, gk);
        const int t = tg[idx];
        if (t == 0) continue;
        const float d = dg[idx];
        const float wd = pick3(i, st.w[0]) * pick3(j, st.w[1]) * pick3(k, st.w[2]) * d;
        seen |= t & 0x55;
        if (t & 1) v0 += (t & 2) ? wd : -wd;
        if (t & 4) v1 += (t & 8) ? wd : -wd;
        if (t & 16) v2 += (t & 32) ? wd : -wd;
        if (t & 64) v3_ += (t & 128) ? wd : -wd;
    }
    int pc = seen;
    if ((seen & 1) && v0 >= 0.0f) pc |= 2;
    if ((seen & 4) && v1 >= 0.0f) pc |= 8;
    if ((seen & 16) && v2 >= 0.0f) pc |= 32;
    if ((seen & 64) && v3_ >= 0.0f) pc |= 128;
    return pc;
}

// cdf_persist_tag, mpm_utils.py:727-741: lanes still in the kernel keep their old side,
// new affinities take the fresh vote's side, lanes out of reach are dropped
MPM_DEV int cdf_persist(int old, int fresh) {
    const int valid = fresh & 0x55;
    const int kept = valid & old;                   // lanes valid before and now
    const int side = ((kept << 1) & old) | (((valid & ~kept) << 1) & fresh);
    return valid | side;
}

// The particle's tag for one half of the pass (mpm_utils.py:851-860 and :971-982): the
// block pretest is exact (an empty stencil votes 0 through the full path as well).
MPM_DEV int cdf_particle_tag(const mpm_grid_t& G, const Stencil& st, int old, bool prev) {
    if (G.cdf.mask_off == 0 && !cdf_near(G, st)) return 0;
    return cdf_persist(old, cdf_vote(G, st, prev));
}

// cdf_particle_ghost, mpm_utils.py:744-816: the particle's free velocity v + dt g
// projected against its primary surface (owner lane of the heaviest valid node, normal
// from sum grad w s d over the prev grids): sticky / slip / separable + Coulomb, plus a
// push out along the normal when the reconstructed distance on the particle's side is
// below 0.05 dx (capped at half a cell per substep).
MPM_DEV v3 cdf_ghost(const mpm_grid_t& G, const mpm_model_t* md, float dt, v3 x, v3 v, int pc, const Stencil& st) {
    float d_hat = 0.0f, best_w = 0.0f;
    v3 n_acc = mk3(0.f, 0.f, 0.f);
    int lane_p = 0;
#pragma unroll 1
    for (int n = 0; n < 27; ++n) {
        const int i = n / 9, j = (n / 3) % 3, k = n % 3;
        const int gi = st.base[0] + i, gj = st.base[1] + j, gk = st.base[2] + k;
        if (!in_nodes(G, gi, gj, gk)) continue;
        const int idx = node_index(G, gi, gj, gk);
        const int t = G.cdf.tag_prev[idx];
        if (t == 0) continue;
        const int own = (t >> MPM_CDF_OWNER_SHIFT) & 3;
        if (((t >> (2 * own)) & 1) == 0) continue;
        const float w0 = pick3(i, st.w[0]), w1 = pick3(j, st.w[1]), w2 = pick3(k, st.w[2]);
        const float weight = w0 * w1 * w2;
        const float sg = ((t >> (2 * own + 1)) & 1) ? 1.0f : -1.0f;
        const float d = G.cdf.dist_prev[idx];
        d_hat = d_hat + weight * sg * d;
        const v3 dwt = mk3(pick3(i, st.dw[0]) * w1 * w2, w0 * pick3(j, st.dw[1]) * w2, w0 * w1 * pick3(k, st.dw[2])) *
                       G.inv_dx;
        n_acc = n_acc + (dwt * sg) * d;
        if (weight > best_w) {
            best_w = weight;
            lane_p = own;
        }
    }
    const v3 v_free = v + mk3(md->gravity[0], md->gravity[1], md->gravity[2]) * dt;
    v3 ghost = v_free;
    if (len(n_acc) > 1.0e-12f) {
        const mpm_cdf_lanes_t* __restrict__ ln = G.cdf.lanes;
        const v3 n_hat = normalize(n_acc);
        const float side = ((pc >> (2 * lane_p + 1)) & 1) ? 1.0f : -1.0f;
        const v3 n_par = n_hat * side;
        const v3 rel = x - mk3(ln->center_prev[lane_p][0], ln->center_prev[lane_p][1], ln->center_prev[lane_p][2]);
        const v3 om = mk3(ln->omega_prev[lane_p][0], ln->omega_prev[lane_p][1], ln->omega_prev[lane_p][2]);
        const v3 v_surf =
            mk3(ln->velocity_prev[lane_p][0], ln->velocity_prev[lane_p][1], ln->velocity_prev[lane_p][2]) +
            cross(om, rel);
        const v3 v_rel = v_free - v_surf;
        const float vn = dot(v_rel, n_par);
        v3 v_tan = v_rel - n_par * vn;
        float c = 0.0f;
        const float d_side = d_hat * side;
        const float eps = 0.05f * G.dx;
        if (d_side < eps) c = fminf((eps - d_side) / dt, 0.5f * G.dx / dt);
        const int stype = ln->type[lane_p];
        if (stype == 0) {
            ghost = v_surf + n_par * c;
        } else if (stype == 1) {
            ghost = (v_surf + v_tan) + n_par * c;
        } else {
            const float dvn = fmaxf(-vn, 0.0f) + c;
            const float tlen = len(v_tan);
            if (tlen > 1.0e-12f && dvn > 0.0f) {
                const float scale = fmaxf(0.0f, tlen - ln->friction[lane_p] * dvn) / tlen;
                v_tan = v_tan * scale;
            }
            ghost = ((v_surf + v_tan) + n_par * fmaxf(vn, 0.0f)) + n_par * c;
        }
    }
    return ghost;
}
