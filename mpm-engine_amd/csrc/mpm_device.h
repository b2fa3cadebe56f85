// Device-side math for the MLS-MPM/APIC substep on gfx950: 3-vectors and row-major
// 3x3 matrices in registers, the quadratic B-spline stencil, a float32 3x3 SVD and the
// constitutive models (return maps + Kirchhoff stress) of warpmpm's material switch.
//
// Reference semantics: src/warpmpm/kernels/mpm_utils.py (cited per function).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpm_hip.h"

// Device functions; a host-side check of the constitutive math defines MPM_DEV as
// __host__ __device__ before including this header.
#ifndef MPM_DEV
#define MPM_DEV __device__ __forceinline__
#endif

struct v3 { float x, y, z; };
struct m33 { float a[9]; };  // row-major: a[3*r + c]

MPM_DEV v3 mk3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
MPM_DEV v3 operator+(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
MPM_DEV v3 operator-(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
MPM_DEV v3 operator*(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
MPM_DEV v3 operator*(float s, v3 a) { return mk3(a.x * s, a.y * s, a.z * s); }
MPM_DEV v3 vdiv(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
// a s + c with one rounding per component (fused multiply-add, stated in the source so the
// rounding does not depend on what the compiler contracts)
MPM_DEV v3 fma3(v3 a, float s, v3 c) { return mk3(fmaf(a.x, s, c.x), fmaf(a.y, s, c.y), fmaf(a.z, s, c.z)); }
MPM_DEV float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MPM_DEV float len(v3 a) { return sqrtf(dot(a, a)); }
MPM_DEV v3 cross(v3 a, v3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
MPM_DEV v3 normalize(v3 a) {
    float l = len(a);
    return l > 0.0f ? vdiv(a, l) : mk3(0.f, 0.f, 0.f);
}

MPM_DEV m33 mzero() { m33 r; for (int i = 0; i < 9; ++i) r.a[i] = 0.f; return r; }
MPM_DEV m33 meye() { m33 r = mzero(); r.a[0] = r.a[4] = r.a[8] = 1.f; return r; }
MPM_DEV m33 mdiag(float a, float b, float c) { m33 r = mzero(); r.a[0] = a; r.a[4] = b; r.a[8] = c; return r; }
MPM_DEV m33 mT(const m33& m) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[3 * i + j] = m.a[3 * j + i];
    return r;
}
MPM_DEV m33 mmul(const m33& x, const m33& y) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.a[3 * i + j] = x.a[3 * i] * y.a[j] + x.a[3 * i + 1] * y.a[3 + j] + x.a[3 * i + 2] * y.a[6 + j];
    return r;
}
MPM_DEV v3 mvmul(const m33& m, v3 v) {
    return mk3(m.a[0] * v.x + m.a[1] * v.y + m.a[2] * v.z, m.a[3] * v.x + m.a[4] * v.y + m.a[5] * v.z,
               m.a[6] * v.x + m.a[7] * v.y + m.a[8] * v.z);
}
MPM_DEV v3 mTvmul(const m33& m, v3 v) {  // m^T v
    return mk3(m.a[0] * v.x + m.a[3] * v.y + m.a[6] * v.z, m.a[1] * v.x + m.a[4] * v.y + m.a[7] * v.z,
               m.a[2] * v.x + m.a[5] * v.y + m.a[8] * v.z);
}
MPM_DEV m33 madd(const m33& x, const m33& y) { m33 r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] + y.a[i]; return r; }
MPM_DEV m33 msub(const m33& x, const m33& y) { m33 r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] - y.a[i]; return r; }
MPM_DEV m33 mscale(const m33& x, float s) { m33 r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] * s; return r; }
MPM_DEV float mdet(const m33& m) {
    const float* a = m.a;
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
           a[2] * (a[3] * a[7] - a[4] * a[6]);
}
// U diag(d) V^T
MPM_DEV m33 udvt(const m33& U, float d0, float d1, float d2, const m33& V) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.a[3 * i + j] = U.a[3 * i] * d0 * V.a[3 * j] + U.a[3 * i + 1] * d1 * V.a[3 * j + 1] +
                             U.a[3 * i + 2] * d2 * V.a[3 * j + 2];
    return r;
}

// ------------------------------------------------------------------------------ SVD
// A = U diag(s) V^T, U and V proper rotations, s sorted by decreasing magnitude with
// s0, s1 >= 0 and sign(s2) = sign(det A) (the convention of McAdams et al. 2011, which is
// what warp's svd3 implements). Method: cyclic Jacobi on S = A^T A with exact rotations
// (Golub & Van Loan Alg. 8.5.2), descending sort of the eigenpairs, then B = A V and a
// Gram-Schmidt orthonormalisation of B's columns whose third axis is u0 x u1.
MPM_DEV void jacobi_rot(float& spp, float& sqq, float& spq, float& sr_p, float& sr_q, m33& V,
                        int p, int q) {
    // rotate the symmetric pair (p,q); (sr_*, ss_*) are the two off-diagonal entries that
    // couple p and q to the remaining index r.
    if (spq == 0.0f) return;
    float theta = (sqq - spp) / (2.0f * spq);
    float t = (theta >= 0.0f ? 1.0f : -1.0f) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
    float c = 1.0f / sqrtf(t * t + 1.0f), s = t * c;
    spp = spp - t * spq;
    sqq = sqq + t * spq;
    spq = 0.0f;
    float a = sr_p, b = sr_q;  // S[r][p], S[r][q]
    sr_p = c * a - s * b;
    sr_q = s * a + c * b;
    for (int k = 0; k < 3; ++k) {
        float vkp = V.a[3 * k + p], vkq = V.a[3 * k + q];
        V.a[3 * k + p] = c * vkp - s * vkq;
        V.a[3 * k + q] = s * vkp + c * vkq;
    }
}

MPM_DEV void svd3(const m33& A, m33& U, v3& s, m33& V) {
    float s00 = 0.f, s11 = 0.f, s22 = 0.f, s01 = 0.f, s02 = 0.f, s12 = 0.f;
    for (int k = 0; k < 3; ++k) {
        float a0 = A.a[3 * k], a1 = A.a[3 * k + 1], a2 = A.a[3 * k + 2];
        s00 += a0 * a0; s11 += a1 * a1; s22 += a2 * a2;
        s01 += a0 * a1; s02 += a0 * a2; s12 += a1 * a2;
    }
    V = meye();
#pragma unroll
    for (int sweep = 0; sweep < 5; ++sweep) {
        jacobi_rot(s00, s11, s01, s02, s12, V, 0, 1);  // r = 2
        jacobi_rot(s00, s22, s02, s01, s12, V, 0, 2);  // r = 1
        jacobi_rot(s11, s22, s12, s01, s02, V, 1, 2);  // r = 0
    }
    // sort eigenpairs descending; each swap negates one column to keep det V = +1
    float e0 = s00, e1 = s11, e2 = s22;
#define SWAPCOL(i, j, ei, ej)                                              \
    {                                                                      \
        float te = ei; ei = ej; ej = te;                                   \
        for (int k = 0; k < 3; ++k) {                                      \
            float t0 = V.a[3 * k + i];                                     \
            V.a[3 * k + i] = V.a[3 * k + j];                               \
            V.a[3 * k + j] = -t0;                                          \
        }                                                                  \
    }
    if (e0 < e1) SWAPCOL(0, 1, e0, e1)
    if (e0 < e2) SWAPCOL(0, 2, e0, e2)
    if (e1 < e2) SWAPCOL(1, 2, e1, e2)
#undef SWAPCOL
    m33 B = mmul(A, V);
    v3 b0 = mk3(B.a[0], B.a[3], B.a[6]);
    v3 b1 = mk3(B.a[1], B.a[4], B.a[7]);
    v3 b2 = mk3(B.a[2], B.a[5], B.a[8]);
    float n0 = len(b0);
    v3 u0 = n0 > 1e-30f ? vdiv(b0, n0) : mk3(1.f, 0.f, 0.f);
    v3 b1p = b1 - u0 * dot(u0, b1);
    float n1 = len(b1p);
    v3 u1;
    if (n1 > 1e-30f && n1 > 1e-7f * n0) {
        u1 = vdiv(b1p, n1);
    } else {  // any unit vector orthogonal to u0
        v3 t = fabsf(u0.x) < 0.577f ? mk3(1.f, 0.f, 0.f) : (fabsf(u0.y) < 0.577f ? mk3(0.f, 1.f, 0.f) : mk3(0.f, 0.f, 1.f));
        t = t - u0 * dot(u0, t);
        u1 = normalize(t);
        n1 = 0.0f;
    }
    v3 u2 = cross(u0, u1);
    s = mk3(n0, n1, dot(u2, b2));
    U.a[0] = u0.x; U.a[3] = u0.y; U.a[6] = u0.z;
    U.a[1] = u1.x; U.a[4] = u1.y; U.a[7] = u1.z;
    U.a[2] = u2.x; U.a[5] = u2.y; U.a[8] = u2.z;
}

// ------------------------------------------------------------------ B-spline stencil
// quadratic B-spline, mpm_utils.py:826-849 / 947-970; w[axis][node], fx = x/dx - base
struct Stencil {
    int base[3];
    float fx[3];
    float w[3][3];
    float dw[3][3];
};

MPM_DEV Stencil make_stencil(v3 x, float inv_dx) {
    Stencil s;
    float gp[3] = {x.x * inv_dx, x.y * inv_dx, x.z * inv_dx};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float fb = floorf(gp[a] - 0.5f);
        s.base[a] = (int)fb;
        float f = gp[a] - fb;
        s.fx[a] = f;
        float wa = 1.5f - f, wb = f - 1.0f, wc = f - 0.5f;
        s.w[a][0] = wa * wa * 0.5f;
        s.w[a][1] = 0.75f - wb * wb;
        s.w[a][2] = wc * wc * 0.5f;
        s.dw[a][0] = f - 1.5f;
        s.dw[a][1] = -2.0f * (f - 1.0f);
        s.dw[a][2] = f - 0.5f;
    }
    return s;
}

// --------------------------------------------------------------------- materials
MPM_DEV float lerp_table(const float* tab, int n, float smin, float smax, float s) {
    float idx = (s - smin) / (smax - smin) * (float)(n - 1);
    float hi = (float)(n - 1) - 1.0e-6f;
    idx = fminf(fmaxf(idx, 0.0f), hi);
    int i0 = (int)floorf(idx);
    float fr = idx - (float)i0;
    return tab[i0] * (1.0f - fr) + tab[i0 + 1] * fr;
}

// generalized-Newtonian Kirchhoff stress (mat 10 and tabulated mat 12),
// mpm_utils.py:27-54 and :56-85
MPM_DEV m33 stress_viscous(float J, float bulk, const m33& L, float eta, float tau_y, float pk,
                           float pn, const mpm_model_t* md, bool tab) {
    const float eps = 0.02f;
    float pressure = -bulk * (powf(J, -1.1f) - 1.0f);
    m33 D = mscale(madd(L, mT(L)), 0.5f);
    float trd = (D.a[0] + D.a[4] + D.a[8]) / 3.0f;
    D.a[0] -= trd; D.a[4] -= trd; D.a[8] -= trd;
    float dd = 0.f;
    for (int i = 0; i < 9; ++i) dd += D.a[i] * D.a[i];
    float gd = sqrtf(2.0f * dd + eps * eps);
    float eta_app;
    if (tab)
        eta_app = lerp_table(md->eta_table, md->eta_table_n, md->eta_table_smin, md->eta_table_smax,
                             logf(gd) * 0.43429448190325176f);
    else
        eta_app = eta + tau_y / gd + (pk != 0.0f ? pk * powf(gd, pn - 1.0f) : 0.0f);  // Newtonian: K = 0
    m33 c = mscale(D, 2.0f * eta_app);
    c.a[0] += pressure; c.a[4] += pressure; c.a[8] += pressure;
    return mscale(c, J);
}

// Hencky Kirchhoff stress U diag(2 mu eps + lam tr eps) U^T, mpm_utils.py:228-243
MPM_DEV m33 stress_hencky(const m33& U, v3 sig, float mu, float lam) {
    float e0 = logf(fmaxf(sig.x, 1e-6f)), e1 = logf(fmaxf(sig.y, 1e-6f)), e2 = logf(fmaxf(sig.z, 1e-6f));
    float tr = e0 + e1 + e2;
    return udvt(U, 2.0f * mu * e0 + lam * tr, 2.0f * mu * e1 + lam * tr, 2.0f * mu * e2 + lam * tr, U);
}

// per-particle mutable material state carried through the return map
struct PState {
    float mu, lam, yield, Jp;
    bool dirty_mu, dirty_yield, dirty_Jp;
};

// von Mises radial return (mat 1) and its damage variant (mat 5),
// mpm_utils.py:127-170 and :172-225
MPM_DEV m33 rm_von_mises(const m33& Ft, const mpm_model_t* md, int mat, PState& ps, bool damage) {
    m33 U, V; v3 so;
    svd3(Ft, U, so, V);
    float e0 = logf(fmaxf(so.x, 0.01f)), e1 = logf(fmaxf(so.y, 0.01f)), e2 = logf(fmaxf(so.z, 0.01f));
    float temp = (e0 + e1 + e2) / 3.0f;
    float trl = ps.lam * (e0 + e1 + e2);
    float t0 = 2.0f * ps.mu * e0 + trl, t1 = 2.0f * ps.mu * e1 + trl, t2 = 2.0f * ps.mu * e2 + trl;
    float st = (t0 + t1 + t2) / 3.0f;
    float c0 = t0 - st, c1 = t1 - st, c2 = t2 - st;
    if (sqrtf(c0 * c0 + c1 * c1 + c2 * c2) > ps.yield) {
        if (damage && ps.yield <= 0.0f) return Ft;
        float h0 = e0 - temp, h1 = e1 - temp, h2 = e2 - temp;
        float hn = sqrtf(h0 * h0 + h1 * h1 + h2 * h2) + 1e-6f;
        float dg = hn - ps.yield / (2.0f * ps.mu);
        float r = dg / hn;
        e0 -= r * h0; e1 -= r * h1; e2 -= r * h2;
        if (damage) {
            ps.yield = ps.yield - md->softening[mat] * sqrtf((r * h0) * (r * h0) + (r * h1) * (r * h1) + (r * h2) * (r * h2));
            ps.dirty_yield = true;
            if (ps.yield <= 0.0f) {
                ps.mu = 0.0f; ps.lam = 0.0f; ps.dirty_mu = true;
            }
        }
        m33 Fe = udvt(U, expf(e0), expf(e1), expf(e2), V);
        if (md->hardening[mat] > 0.5f) {
            ps.yield = ps.yield + 2.0f * ps.mu * md->xi[mat] * dg;
            ps.dirty_yield = true;
        }
        return Fe;
    }
    return Ft;
}

// bisection on g(gd) = tau_bar - G dt gd - mu(I) p, shared by mats 9 / 11 / 13
MPM_DEV float mu_i_bisect(float tbK, float G, float dt, float pK, float Ic, const mpm_model_t* md,
                          int mat, bool tab) {
    float mu_s = md->muI_mu_s[mat], dmu = md->muI_delta_mu[mat], I0 = md->muI_I0[mat];
    float lo = 0.0f, hi = tbK / (G * dt);
    for (int it = 0; it < 48; ++it) {
        float mid = 0.5f * (lo + hi);
        float Im = mid * Ic;
        float mum;
        if (tab)
            mum = lerp_table(md->eta_table, md->eta_table_n, md->eta_table_smin, md->eta_table_smax,
                             logf(fmaxf(Im, 1e-12f)) * 0.43429448190325176f);
        else
            mum = mu_s + dmu * Im / (Im + I0);
        float g = tbK - G * dt * mid - mum * pK;
        if (g > 0.0f) lo = mid; else hi = mid;
    }
    return 0.5f * (lo + hi);
}

// local mu(I) return map (mat 9) and tabulated variant (mat 13),
// mpm_utils.py:246-324 and :327-407
MPM_DEV m33 rm_mu_i(const m33& Ft, const mpm_model_t* md, int mat, float dt, PState& ps, bool tab) {
    m33 U, V; v3 sg;
    svd3(Ft, U, sg, V);
    float c0 = fmaxf(fabsf(sg.x), 1e-6f), c1 = fmaxf(fabsf(sg.y), 1e-6f), c2 = fmaxf(fabsf(sg.z), 1e-6f);
    float e0 = logf(c0), e1 = logf(c1), e2 = logf(c2);
    float tr = e0 + e1 + e2;
    if (tr >= 0.0f) return mmul(U, mT(V));
    float G = ps.mu, lam = ps.lam;
    float J = c0 * c1 * c2;
    float me = tr / 3.0f;
    float d0 = e0 - me, d1 = e1 - me, d2 = e2 - me;
    float dn = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
    float pK = -(2.0f * G / 3.0f + lam) * tr;
    float tbK = 2.0f * G * dn / sqrtf(2.0f);
    float pC = pK / J;
    float thr = tab ? md->eta_table[0] : md->muI_mu_s[mat];
    if (tbK > thr * pK) {
        float Ic = md->muI_d[mat] * sqrtf(md->muI_rho_s[mat] / pC);
        float gdp = mu_i_bisect(tbK, G, dt, pK, Ic, md, mat, tab);
        float scale = (tbK - G * dt * gdp) / fmaxf(tbK, 1e-20f);
        return udvt(U, expf(d0 * scale + me), expf(d1 * scale + me), expf(d2 * scale + me), V);
    }
    return Ft;
}

// mu_i_phi_pressure, mpm_utils.py:421-430
MPM_DEV float mu_i_phi_pressure(const mpm_model_t* md, int mat, float J, float Jp_ref) {
    float K = md->E[mat] / (3.0f * (1.0f - 2.0f * md->nu[mat]));
    return fmaxf(K * (Jp_ref / fmaxf(J, 1e-6f) - 1.0f), 0.0f);
}

// compressible mu(I)-Phi(I) return map (mat 11), mpm_utils.py:433-522
MPM_DEV m33 rm_mu_i_phi(const m33& Ft, const mpm_model_t* md, int mat, float dt, PState& ps) {
    m33 U, V; v3 sg;
    svd3(Ft, U, sg, V);
    float c0 = fmaxf(fabsf(sg.x), 1e-6f), c1 = fmaxf(fabsf(sg.y), 1e-6f), c2 = fmaxf(fabsf(sg.z), 1e-6f);
    float e0 = logf(c0), e1 = logf(c1), e2 = logf(c2);
    float J = c0 * c1 * c2;
    float tr = e0 + e1 + e2;
    float me = tr / 3.0f;
    float d0 = e0 - me, d1 = e1 - me, d2 = e2 - me;
    float dn = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
    float G = ps.mu;
    float Jp_ref = ps.Jp < 0.5f ? 1.0f : ps.Jp;
    float p_comp = mu_i_phi_pressure(md, mat, J, Jp_ref);
    ps.Jp = Jp_ref;
    ps.dirty_Jp = true;
    if (p_comp <= 0.0f) {
        float si = expf(me);
        return udvt(U, si, si, si, V);
    }
    float pK = p_comp * J;
    float tbK = 2.0f * G * dn / sqrtf(2.0f);
    if (tbK > md->muI_mu_s[mat] * pK) {
        float Ic = md->muI_d[mat] * sqrtf(md->muI_rho_s[mat] / p_comp);
        float gdp = mu_i_bisect(tbK, G, dt, pK, Ic, md, mat, false);
        float I_real = gdp * Ic;
        float dgam = fminf(8.0f * gdp * dt, 1.0f);
        float denom = fmaxf(1.0f - md->muI_phi_chi[mat] * I_real / (I_real + md->muI_I0[mat]), 1e-3f);
        ps.Jp = Jp_ref + dgam * (1.0f / denom - Jp_ref);
        float scale = (tbK - G * dt * gdp) / fmaxf(tbK, 1e-20f);
        return udvt(U, expf(d0 * scale + me), expf(d1 * scale + me), expf(d2 * scale + me), V);
    }
    return Ft;
}

// Drucker-Prager sand return map (mat 2), mpm_utils.py:543-580
MPM_DEV m33 rm_sand(const m33& Ft, const mpm_model_t* md, int mat, const PState& ps) {
    m33 U, V; v3 sg;
    svd3(Ft, U, sg, V);
    float e0 = logf(fmaxf(fabsf(sg.x), 1e-14f)), e1 = logf(fmaxf(fabsf(sg.y), 1e-14f)),
          e2 = logf(fmaxf(fabsf(sg.z), 1e-14f));
    float tr = e0 + e1 + e2;
    float h0 = e0 - tr / 3.0f, h1 = e1 - tr / 3.0f, h2 = e2 - tr / 3.0f;
    float hn = sqrtf(h0 * h0 + h1 * h1 + h2 * h2);
    float dg = hn + (3.0f * ps.lam + 2.0f * ps.mu) / (2.0f * ps.mu) * tr * md->alpha[mat];
    if (dg > 0.0f && tr > 0.0f) return mmul(U, mT(V));
    if (dg > 0.0f && tr <= 0.0f) {
        float r = dg / hn;
        return udvt(U, expf(e0 - h0 * r), expf(e1 - h1 * r), expf(e2 - h2 * r), V);
    }
    return Ft;
}

// stress_update_particle, mpm_utils.py:1056-1149: return map F_trial -> F, then the
// symmetrised Kirchhoff stress of the returned F.
MPM_DEV void stress_update(const m33& Ft, const m33& L, int mat, const mpm_model_t* md, float dt,
                           PState& ps, m33& F, m33& S) {
    switch (mat) {
        case 1: F = rm_von_mises(Ft, md, mat, ps, false); break;
        case 2: F = rm_sand(Ft, md, mat, ps); break;
        case 5: F = rm_von_mises(Ft, md, mat, ps, true); break;
        case 9: F = rm_mu_i(Ft, md, mat, dt, ps, false); break;
        case 13: F = rm_mu_i(Ft, md, mat, dt, ps, true); break;
        case 11: F = rm_mu_i_phi(Ft, md, mat, dt, ps); break;
        case 6: case 10: case 12: {
            float Jc = powf(mdet(Ft), 1.0f / 3.0f);
            F = mdiag(Jc, Jc, Jc);
        } break;
        case 7: case 8: F = meye(); break;
        default: F = Ft; break;
    }
    float J = mdet(F);
    S = mzero();
    if (mat == 6 || mat == 10 || mat == 12) {
        if (mat == 6) {
            float pr = -md->bulk[mat] * (powf(J, -1.1f) - 1.0f);
            S = mdiag(pr * J, pr * J, pr * J);
        } else {
            S = stress_viscous(J, md->bulk[mat], L, md->plastic_viscosity[mat], ps.yield,
                               md->hardening[mat], md->softening[mat], md, mat == 12);
        }
    } else if (mat == 0 || mat == 5 || mat == 1 || mat == 9 || mat == 13 || mat == 2 || mat == 11) {
        m33 U, V; v3 sg;
        svd3(F, U, sg, V);
        if (mat == 0 || mat == 5) {  // FCR, mpm_utils.py:8-15
            m33 R = mmul(U, mT(V));
            S = mmul(mscale(msub(F, R), 2.0f * ps.mu), mT(F));
            float lj = ps.lam * J * (J - 1.0f);
            S.a[0] += lj; S.a[4] += lj; S.a[8] += lj;
        } else if (mat == 1 || mat == 9 || mat == 13) {
            S = stress_hencky(U, sg, ps.mu, ps.lam);
        } else if (mat == 2) {  // Drucker-Prager stress, mpm_utils.py:109-124
            float l0 = logf(sg.x), l1 = logf(sg.y), l2 = logf(sg.z);
            float ls = l0 + l1 + l2;
            float c0 = 2.0f * ps.mu * l0 * (1.0f / sg.x) + ps.lam * ls * (1.0f / sg.x);
            float c1 = 2.0f * ps.mu * l1 * (1.0f / sg.y) + ps.lam * ls * (1.0f / sg.y);
            float c2 = 2.0f * ps.mu * l2 * (1.0f / sg.z) + ps.lam * ls * (1.0f / sg.z);
            S = mmul(udvt(U, c0, c1, c2, V), mT(F));
        } else {  // mat 11, mpm_utils.py:525-540 with the particle's Jp
            float Jp_ref = ps.Jp < 0.5f ? 1.0f : ps.Jp;
            float pc = mu_i_phi_pressure(md, mat, J, Jp_ref);
            float c0 = fmaxf(sg.x, 1e-6f), c1 = fmaxf(sg.y, 1e-6f), c2 = fmaxf(sg.z, 1e-6f);
            float e0 = logf(c0), e1 = logf(c1), e2 = logf(c2);
            float Js = c0 * c1 * c2;
            float me = (e0 + e1 + e2) / 3.0f;
            float iso = -pc * Js;
            S = udvt(U, 2.0f * ps.mu * (e0 - me) + iso, 2.0f * ps.mu * (e1 - me) + iso,
                     2.0f * ps.mu * (e2 - me) + iso, U);
        }
    }
    // enforce symmetry
    S = mscale(madd(S, mT(S)), 0.5f);
}

// ------------------------------------------------------------- voxel fields
// sdf_trilerp(_vec), mpm_solver_warp.py:60-91
MPM_DEV void trilerp(const float* __restrict__ vals, int nc, int res, v3 f, float* out) {
    float hi = (float)res - 1.0001f;
    float fi = fminf(fmaxf(f.x, 0.0f), hi), fj = fminf(fmaxf(f.y, 0.0f), hi), fk = fminf(fmaxf(f.z, 0.0f), hi);
    int i = (int)floorf(fi), j = (int)floorf(fj), k = (int)floorf(fk);
    float tx = fi - (float)i, ty = fj - (float)j, tz = fk - (float)k;
    auto V = [&](int a, int b, int c, int ch) { return vals[(((size_t)a * res + b) * res + c) * nc + ch]; };
    for (int ch = 0; ch < nc; ++ch) {
        float c00 = V(i, j, k, ch) * (1.0f - tx) + V(i + 1, j, k, ch) * tx;
        float c01 = V(i, j, k + 1, ch) * (1.0f - tx) + V(i + 1, j, k + 1, ch) * tx;
        float c10 = V(i, j + 1, k, ch) * (1.0f - tx) + V(i + 1, j + 1, k, ch) * tx;
        float c11 = V(i, j + 1, k + 1, ch) * (1.0f - tx) + V(i + 1, j + 1, k + 1, ch) * tx;
        float c0 = c00 * (1.0f - ty) + c10 * ty;
        float c1 = c01 * (1.0f - ty) + c11 * ty;
        out[ch] = c0 * (1.0f - tz) + c1 * tz;
    }
}

// v' = v(2w^2 - 1) + 2(q.v) q +/- 2w (q x v): rotation by a unit quaternion (x, y, z, w)
MPM_DEV v3 quat_rotate(const float* q, v3 v, bool inverse) {
    v3 qv = mk3(q[0], q[1], q[2]);
    float w = q[3];
    return v * (2.0f * w * w - 1.0f) + qv * (2.0f * dot(qv, v)) + cross(qv, v) * ((inverse ? -w : w) * 2.0f);
}

// ------------------------------------------------------------------ collider SDFs
// capped_cylinder_sdf / revolved_glass_sdf, mpm_solver_warp.py:29-57 (the revolved cup:
// grid update and its adjoint)
MPM_DEV float capped_cylinder_sdf(float r, float z, float radius, float z0, float z1) {
    float zc = 0.5f * (z0 + z1), hh = 0.5f * (z1 - z0);
    float dr = r - radius, dz = fabsf(z - zc) - hh;
    float dro = fmaxf(dr, 0.0f), dzo = fmaxf(dz, 0.0f);
    return sqrtf(dro * dro + dzo * dzo) + fminf(fmaxf(dr, dz), 0.0f);
}

MPM_DEV float glass_sdf(v3 l, const mpm_collider_t& c) {
    float r = sqrtf(l.x * l.x + l.y * l.y);
    float d_outer = capped_cylinder_sdf(r, l.z, c.outer_radius, -c.half_height, c.half_height);
    float d_cav = capped_cylinder_sdf(r, l.z, c.inner_radius - c.fillet_radius,
                                      c.inner_floor_z + c.fillet_radius, c.half_height + c.outer_radius) -
                  c.fillet_radius;
    return fmaxf(d_outer, -d_cav);
}

// ------------------------------------------------------------- layout helpers
MPM_DEV int node_index(const mpm_grid_t& G, int i, int j, int k) {
    return ((((i >> 2) * G.nb[1] + (j >> 2)) * G.nb[2] + (k >> 2)) << 6) + ((i & 3) << 4) +
           ((j & 3) << 2) + (k & 3);
}

MPM_DEV size_t scene_base(const mpm_grid_t& G, int scene) {
    return (size_t)scene * (size_t)G.nb[0] * G.nb[1] * G.nb[2] * 64;
}

// Storage address (in nodes) of node (i, j, k) of the scene whose dense node offset is
// soff (scene_base): dense storage keeps the block code's own 64 nodes; with a block pool
// the block's slot holds them. -1: the block has no storage this tick.
MPM_DEV long node_addr(const mpm_grid_t& G, size_t soff, int i, int j, int k) {
    const int nidx = node_index(G, i, j, k);
    if (!G.block_slot) return (long)soff + nidx;
    const int s = G.block_slot[(soff >> 6) + (size_t)(nidx >> 6)];
    return (s >= 0 && s < G.pool_blocks) ? ((long)s << 6) + (nidx & 63) : -1L;
}

// storage block of global block code `code` (dense: the code itself), -1 without storage
MPM_DEV long block_storage(const mpm_grid_t& G, long code) {
    if (!G.block_slot) return code;
    const int s = G.block_slot[code];
    return (s >= 0 && s < G.pool_blocks) ? (long)s : -1L;
}

MPM_DEV void decode_block(const mpm_grid_t& G, int code, int& bi, int& bj, int& bk) {
    bk = code % G.nb[2];
    int r = code / G.nb[2];
    bj = r % G.nb[1];
    bi = r / G.nb[1];
}

MPM_DEV v3 ld3(const float* base, int stride, int p) {
    return mk3(base[p], base[stride + p], base[2 * stride + p]);
}
MPM_DEV void st3(float* base, int stride, int p, v3 v) {
    base[p] = v.x; base[stride + p] = v.y; base[2 * stride + p] = v.z;
}
MPM_DEV m33 ld9(const float* base, int stride, int p) {
    m33 m;
#pragma unroll
    for (int i = 0; i < 9; ++i) m.a[i] = base[i * stride + p];
    return m;
}
MPM_DEV void st9(float* base, int stride, int p, const m33& m) {
#pragma unroll
    for (int i = 0; i < 9; ++i) base[i * stride + p] = m.a[i];
}

