// Adjoints of the constitutive updates (return maps + Kirchhoff stress) used by the
// reverse substep (mpm_adjoint.hip): fixed-corotated elasticity (material 0), von Mises
// (material 1), Drucker-Prager sand (material 2), local mu(I) (material 9) and the
// generalized-Newtonian fluid (material 10); and the contact maps of the surface colliders
// (revolved cup, mesh SDF) with their transposed Jacobians. Pure register math on the types of mpm_device.h; a host-side check compiles it
// with MPM_DEV = __host__ __device__ and compares it with float64 finite differences of
// the forward maps.
#pragma once

#include "mpm_device.h"

// u / expm1(u), its series below |u| < 1e-3
MPM_DEV float u_over_expm1(float u) { return fabsf(u) < 1e-3f ? 1.0f - 0.5f * u + u * u * (1.0f / 12.0f) : u / expm1f(u); }
// expm1(k u) / expm1(u)
MPM_DEV float expm1_ratio(float k, float u) {
    return fabsf(u) < 1e-3f ? k * (1.0f + 0.5f * (k - 1.0f) * u) : expm1f(k * u) / expm1f(u);
}

// SVD-frame part of the adjoint shared by the return maps whose corrected log-strain is
// en = me + k (e - me) (e = log of the clamped singular values s of F_trial, me their
// mean): the Hencky stress tau = U diag(h) U^T, h = 2 mu en + lam tr(en), and, when fmap,
// the corrected F = U diag(exp en) V^T (otherwise F = F_trial passes through). In: Fb,
// tb. Out: Al = the off-diagonal SVD-frame adjoint, enb = adjoint of en; the mu and lam
// partials of tau are added.
MPM_DEV void iso_adjoint(const m33& U, const m33& V, const float s[3], const float e[3], const float en[3], float k,
                         float mu, float lam, bool fmap, const m33& Fb, const m33& tb_in, m33& Al, float enb[3],
                         float& mub, float& lamb) {
    const float tre = en[0] + en[1] + en[2];
    float h[3];
    for (int a = 0; a < 3; ++a) h[a] = 2.0f * mu * en[a] + lam * tre;
    const m33 tb = mscale(madd(tb_in, mT(tb_in)), 0.5f);
    const m33 T = mmul(mT(U), mmul(tb, U));
    const float trT = T.a[0] + T.a[4] + T.a[8];
    for (int a = 0; a < 3; ++a) enb[a] = 2.0f * mu * T.a[4 * a] + lam * trT;
    mub += 2.0f * (en[0] * T.a[0] + en[1] * T.a[4] + en[2] * T.a[8]);
    lamb += tre * trT;
    Al = mzero();
    const int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
    for (int q = 0; q < 3; ++q) {
        const int a = pa[q], b = pb[q];
        const float u = e[a] - e[b];
        const float d1 = 2.0f * mu * k * u_over_expm1(u) / s[b];
        const float ssum = s[a] + s[b];
        const float d2 = fabsf(ssum) > 1e-12f ? (h[b] - h[a]) / ssum : 0.0f;
        const float Tab = T.a[3 * a + b] + T.a[3 * b + a];
        Al.a[3 * a + b] += 0.5f * Tab * (d1 + d2);
        Al.a[3 * b + a] += 0.5f * Tab * (d1 - d2);
    }
    if (fmap) {
        const m33 M = mmul(mT(U), mmul(Fb, V));
        float g[3];
        for (int a = 0; a < 3; ++a) {
            g[a] = expf(en[a]);
            enb[a] += M.a[4 * a] * g[a];
        }
        for (int q = 0; q < 3; ++q) {
            const int a = pa[q], b = pb[q];
            const float u = e[a] - e[b];
            const float cA = g[b] / s[b] * expm1_ratio(k, u);
            const float ssum = s[a] + s[b];
            const float cB = fabsf(ssum) > 1e-12f ? (g[a] + g[b]) / ssum : 0.0f;
            const float S = M.a[3 * a + b] + M.a[3 * b + a], D = M.a[3 * a + b] - M.a[3 * b + a];
            Al.a[3 * a + b] += 0.5f * (cA * S + cB * D);
            Al.a[3 * b + a] += 0.5f * (cA * S - cB * D);
        }
    }
}

// Adjoint of the von Mises return map + Hencky stress (mat 1) at F_trial. In: Fb = adjoint
// of the corrected F, tb = adjoint of tau, ysb_next = adjoint of the yield stress the map
// leaves behind, which with hardening (xi > 0; xi = 0 without) is ys + 2 mu xi dg when
// yielding (mpm_utils.py:164-167) and ys otherwise. Out: adjoint of F_trial; the mu and
// lam adjoints are added, and yb receives the adjoint of the yield stress the map read
// (ysb_next carried through plus this substep's own dependence).
MPM_DEV m33 adj_von_mises(const m33& Ft, float mu, float lam, float ys, float xi, float ysb_next, const m33& Fb,
                          const m33& tb_in, float& mub, float& lamb, float& yb) {
    m33 U, V;
    v3 s3;
    svd3(Ft, U, s3, V);
    const float s[3] = {s3.x, s3.y, s3.z};
    float e[3];
    for (int a = 0; a < 3; ++a) e[a] = logf(fmaxf(s[a], 0.01f));
    const float me = (e[0] + e[1] + e[2]) / 3.0f;
    float eh[3] = {e[0] - me, e[1] - me, e[2] - me};
    // forward's yield test, same expression and order
    const float trl = lam * (e[0] + e[1] + e[2]);
    const float t0 = 2.0f * mu * e[0] + trl, t1 = 2.0f * mu * e[1] + trl, t2 = 2.0f * mu * e[2] + trl;
    const float stv = (t0 + t1 + t2) / 3.0f;
    const float c0 = t0 - stv, c1 = t1 - stv, c2 = t2 - stv;
    const bool yield = sqrtf(c0 * c0 + c1 * c1 + c2 * c2) > ys;
    float k = 1.0f, hn = 0.0f;
    if (yield) {
        hn = sqrtf(eh[0] * eh[0] + eh[1] * eh[1] + eh[2] * eh[2]) + 1e-6f;
        const float dg = hn - ys / (2.0f * mu);
        k = 1.0f - dg / hn;
    }
    float en[3];
    for (int a = 0; a < 3; ++a) en[a] = me + k * eh[a];
    m33 Al;
    float enb[3];
    iso_adjoint(U, V, s, e, en, k, mu, lam, yield, Fb, tb_in, Al, enb, mub, lamb);
    float eb[3];
    yb += ysb_next;  // the yield stress passes on (the whole of it below the yield surface)
    if (yield) {
        // en = me 1 + (ys / (2 mu hn)) eh with hn = |eh| + 1e-6
        const float ehn = hn - 1e-6f;
        const float kk = ys / (2.0f * mu * hn);
        const float dotv = eh[0] * enb[0] + eh[1] * enb[1] + eh[2] * enb[2];
        const float c = ehn > 0.0f ? dotv / (ehn * hn) : 0.0f;
        float tmp[3];
        for (int a = 0; a < 3; ++a) tmp[a] = enb[a] - eh[a] * c;
        const float mt = (tmp[0] + tmp[1] + tmp[2]) / 3.0f;
        const float mb = (enb[0] + enb[1] + enb[2]) / 3.0f;
        for (int a = 0; a < 3; ++a) eb[a] = mb + kk * (tmp[a] - mt);
        yb += dotv / (2.0f * mu * hn);
        mub += -ys * dotv / (2.0f * mu * mu * hn);
        if (xi != 0.0f && ysb_next != 0.0f) {
            // hardening: ys' = ys + 2 mu xi dg, dg = |eh| + 1e-6 - ys / (2 mu), so
            // d ys'/d ys = 1 - xi, d ys'/d mu = 2 xi dg + xi ys / mu and
            // d ys'/d e = 2 mu xi eh / |eh| (eh is deviatoric: no mean to remove)
            const float dg = hn - ys / (2.0f * mu);
            yb -= xi * ysb_next;
            mub += ysb_next * (2.0f * xi * dg + xi * ys / mu);
            const float ce = ehn > 0.0f ? 2.0f * mu * xi * ysb_next / ehn : 0.0f;
            for (int a = 0; a < 3; ++a) eb[a] += ce * eh[a];
        }
    } else {
        for (int a = 0; a < 3; ++a) eb[a] = enb[a];
    }
    for (int a = 0; a < 3; ++a) Al.a[4 * a] += s[a] > 0.01f ? eb[a] / s[a] : 0.0f;
    m33 Ftb = mmul(U, mmul(Al, mT(V)));
    if (!yield) Ftb = madd(Ftb, Fb);  // F = F_trial passes through
    return Ftb;
}

// Adjoint of the local mu(I) return map + Hencky stress (mat 9, rm_mu_i; reference
// mpm_utils.py:246-324) at F_trial. Tension (tr e >= 0): F = U V^T, tau = 0 (k = 0).
// Yield (tau_K > mu_s p_K): en = me + k (e - me) with k = 1 - G dt gam / tau_K, where gam
// is the root of g(gam) = tau_K - G dt gam - mu(gam I_c) p_K (mu(I) = mu_s + dmu I /
// (I + I0), p_K = -(2 G / 3 + lam) tr e, tau_K = sqrt(2) G |e - me|, I_c = d sqrt(rho_s /
// p_C), p_C = p_K / J); the root's derivative comes from the implicit function theorem,
// gam_theta = -g_theta / g_gam. Out: adjoint of F_trial; the partials of mu, lam, mu_s,
// dmu, I0 and the grain diameter d are added.
MPM_DEV m33 adj_mu_i(const m33& Ft, const mpm_model_t* md, int mat, float dt, float mu, float lam, const m33& Fb,
                     const m33& tb_in, float& mub, float& lamb, float* pb) {
    m33 U, V;
    v3 s3;
    svd3(Ft, U, s3, V);
    const float s[3] = {s3.x, s3.y, s3.z};
    float c[3], e[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = fmaxf(fabsf(s[a]), 1e-6f);
        e[a] = logf(c[a]);
    }
    const float tr = e[0] + e[1] + e[2];
    const float me = tr / 3.0f;
    const float d[3] = {e[0] - me, e[1] - me, e[2] - me};
    const bool tension = tr >= 0.0f;
    bool yield = false;
    float k = 1.0f, J = 0.f, dn = 0.f, pK = 0.f, tbK = 0.f, pC = 0.f, Ic = 0.f, gam = 0.f;
    const float G = mu;
    if (tension) {
        k = 0.0f;
    } else {
        J = c[0] * c[1] * c[2];
        dn = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        pK = -(2.0f * G / 3.0f + lam) * tr;
        tbK = 2.0f * G * dn / sqrtf(2.0f);
        pC = pK / J;
        if (tbK > md->muI_mu_s[mat] * pK) {
            yield = true;
            Ic = md->muI_d[mat] * sqrtf(md->muI_rho_s[mat] / pC);
            gam = mu_i_bisect(tbK, G, dt, pK, Ic, md, mat, false);
            k = (tbK - G * dt * gam) / fmaxf(tbK, 1e-20f);
        }
    }
    float en[3];
    for (int a = 0; a < 3; ++a) en[a] = tension ? 0.0f : me + k * d[a];
    m33 Al;
    float enb[3];
    iso_adjoint(U, V, s, e, en, k, mu, lam, tension || yield, Fb, tb_in, Al, enb, mub, lamb);
    float eb[3] = {0.f, 0.f, 0.f};
    if (yield) {
        const float mu_s = md->muI_mu_s[mat], dmu = md->muI_delta_mu[mat], I0 = md->muI_I0[mat];
        const float meb = enb[0] + enb[1] + enb[2];
        float db[3];
        for (int a = 0; a < 3; ++a) db[a] = k * enb[a];
        const float kb = enb[0] * d[0] + enb[1] * d[1] + enb[2] * d[2];
        // k = 1 - G dt gam / tau_K
        const float gamb = -kb * G * dt / tbK;
        float tbKb = kb * G * dt * gam / (tbK * tbK);
        float Gb = -kb * dt * gam / tbK;
        // implicit root: multiplier lm = -gamb / g_gam
        const float I = gam * Ic;
        const float mu_I = mu_s + dmu * I / (I + I0);
        const float dmu_dI = dmu * I0 / ((I + I0) * (I + I0));
        const float g_gam = -G * dt - dmu_dI * Ic * pK;
        const float lm = -gamb / g_gam;
        tbKb += lm;
        Gb += -lm * dt * gam;
        float pKb = -lm * mu_I;
        const float Icb = -lm * dmu_dI * gam * pK;
        pb[0] += -lm * pK;
        pb[1] += -lm * pK * I / (I + I0);
        pb[2] += lm * dmu * I * pK / ((I + I0) * (I + I0));
        pb[3] += Icb * Ic / md->muI_d[mat];
        // I_c = d sqrt(rho_s / p_C), p_C = p_K / J, J = exp(tr)
        const float pCb = -Icb * Ic / (2.0f * pC);
        pKb += pCb / J;
        float trb = -pCb * pK / J;
        // p_K = -(2 G / 3 + lam) tr
        Gb += -pKb * (2.0f / 3.0f) * tr;
        lamb += -pKb * tr;
        trb += -pKb * (2.0f * G / 3.0f + lam);
        // tau_K = sqrt(2) G dn, dn = |d|
        Gb += tbKb * sqrtf(2.0f) * dn;
        const float dnb = tbKb * sqrtf(2.0f) * G;
        for (int a = 0; a < 3; ++a) db[a] += dn > 0.0f ? dnb * d[a] / dn : 0.0f;
        const float mdb = (db[0] + db[1] + db[2]) / 3.0f;
        for (int a = 0; a < 3; ++a) eb[a] = db[a] - mdb + meb / 3.0f + trb;
        mub += Gb;
    } else if (!tension) {
        for (int a = 0; a < 3; ++a) eb[a] = enb[a];
    }
    for (int a = 0; a < 3; ++a) Al.a[4 * a] += s[a] > 1e-6f ? eb[a] / s[a] : 0.0f;
    m33 Ftb = mmul(U, mmul(Al, mT(V)));
    if (!tension && !yield) Ftb = madd(Ftb, Fb);  // F = F_trial passes through
    return Ftb;
}

// Adjoint of the Drucker-Prager sand return map (mat 2, rm_sand; reference
// sand_return_mapping, mpm_utils.py:544-580) and its stress, the Hencky stress of the
// returned F (kirchoff_stress_drucker_prager, mpm_utils.py:109-124). With e = log of the
// clamped singular values, tr = sum e, h = e - tr / 3 and dg = |h| + c tr, c = alpha (3 lam
// + 2 mu) / (2 mu): dg <= 0 keeps F = F_trial; dg > 0 in tension (tr > 0) gives F = U V^T
// (en = 0, k = 0); dg > 0 in compression puts the log-strain on the cone, en = tr / 3 + k h
// with k = 1 - dg / |h| = -c tr / |h|. Out: adjoint of F_trial; mu and lam partials added
// (through the stress and through c).
MPM_DEV m33 adj_sand(const m33& Ft, float mu, float lam, float alpha, const m33& Fb, const m33& tb_in, float& mub,
                     float& lamb) {
    m33 U, V;
    v3 s3;
    svd3(Ft, U, s3, V);
    const float s[3] = {s3.x, s3.y, s3.z};
    float e[3];
    for (int a = 0; a < 3; ++a) e[a] = logf(fmaxf(fabsf(s[a]), 1e-14f));
    const float tr = e[0] + e[1] + e[2];
    const float me = tr / 3.0f;
    const float h[3] = {e[0] - me, e[1] - me, e[2] - me};
    const float hn = sqrtf(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    const float c = (3.0f * lam + 2.0f * mu) / (2.0f * mu) * alpha;
    const float dg = hn + c * tr;  // forward's test, same expression
    const bool tension = dg > 0.0f && tr > 0.0f;
    const bool cone = dg > 0.0f && tr <= 0.0f;  // then |h| > -c tr >= 0
    float k = 1.0f;
    if (tension) k = 0.0f;
    if (cone) k = -c * tr / hn;
    float en[3];
    for (int a = 0; a < 3; ++a) en[a] = tension ? 0.0f : (cone ? me + k * h[a] : e[a]);
    m33 Al;
    float enb[3];
    iso_adjoint(U, V, s, e, en, k, mu, lam, tension || cone, Fb, tb_in, Al, enb, mub, lamb);
    float eb[3] = {0.f, 0.f, 0.f};
    if (cone) {
        const float meb = enb[0] + enb[1] + enb[2];
        const float kb = enb[0] * h[0] + enb[1] * h[1] + enb[2] * h[2];
        // k = -c tr / |h|
        float trb = -kb * c / hn + meb / 3.0f;
        const float hnb = kb * c * tr / (hn * hn);
        const float cb = -kb * tr / hn;
        float hb[3];
        for (int a = 0; a < 3; ++a) hb[a] = k * enb[a] + hnb * h[a] / hn;
        const float mhb = (hb[0] + hb[1] + hb[2]) / 3.0f;
        for (int a = 0; a < 3; ++a) eb[a] = hb[a] - mhb + trb;
        // c = alpha (3 lam / (2 mu) + 1)
        lamb += cb * 3.0f * alpha / (2.0f * mu);
        mub += -cb * 3.0f * alpha * lam / (2.0f * mu * mu);
    } else if (!tension) {
        for (int a = 0; a < 3; ++a) eb[a] = enb[a];
    }
    for (int a = 0; a < 3; ++a) Al.a[4 * a] += fabsf(s[a]) > 1e-14f ? eb[a] / s[a] : 0.0f;
    m33 Ftb = mmul(U, mmul(Al, mT(V)));
    if (!tension && !cone) Ftb = madd(Ftb, Fb);  // F = F_trial passes through
    return Ftb;
}

// cofactor matrix of m: the derivative of det(m) with respect to m
MPM_DEV m33 mcof(const m33& m) {
    const float* a = m.a;
    m33 c;
    c.a[0] = a[4] * a[8] - a[5] * a[7];
    c.a[1] = a[5] * a[6] - a[3] * a[8];
    c.a[2] = a[3] * a[7] - a[4] * a[6];
    c.a[3] = a[2] * a[7] - a[1] * a[8];
    c.a[4] = a[0] * a[8] - a[2] * a[6];
    c.a[5] = a[1] * a[6] - a[0] * a[7];
    c.a[6] = a[1] * a[5] - a[2] * a[4];
    c.a[7] = a[2] * a[3] - a[0] * a[5];
    c.a[8] = a[0] * a[4] - a[1] * a[3];
    return c;
}

// Adjoint of the fixed-corotated stress (mat 0, kirchoff_stress_FCR, mpm_utils.py:8-15, with
// F = F_trial, mpm_utils.py:1094-1095). With F = U diag(s) V^T and R = U V^T,
//   tau = 2 mu (F - R) F^T + lam J (J - 1) I = 2 mu (F F^T - P) + lam J (J - 1) I,
// where P = R F^T = U diag(s) U^T is the left stretch. In the SVD frame (dF^ = U^T dF V)
// dP^_ab = (s_b dF^_ab + s_a dF^_ba) / (s_a + s_b), which holds for the signed
// decomposition too. In: Fb = adjoint of F, tb = adjoint of tau. Returns the adjoint of
// F_trial; the mu and lam partials are added.
MPM_DEV m33 adj_fcr(const m33& F, float mu, float lam, const m33& Fb, const m33& tb_in, float& mub, float& lamb) {
    m33 U, V;
    v3 s3;
    svd3(F, U, s3, V);
    const float s[3] = {s3.x, s3.y, s3.z};
    const m33 T = mscale(madd(tb_in, mT(tb_in)), 0.5f);  // the forward symmetrises tau
    const float J = mdet(F);
    const float trT = T.a[0] + T.a[4] + T.a[8];
    const m33 Th = mmul(mT(U), mmul(T, U));
    m33 Ph;  // adjoint of dF^ through P
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            const float ssum = s[a] + s[b];
            Ph.a[3 * a + b] = a == b ? Th.a[4 * a] : (fabsf(ssum) > 1e-12f ? 2.0f * Th.a[3 * a + b] * s[b] / ssum : 0.0f);
        }
    const m33 FFt = mmul(F, mT(F));
    const m33 P = mmul(U, mmul(mdiag(s[0], s[1], s[2]), mT(U)));
    float tp = 0.0f;
    for (int i = 0; i < 9; ++i) tp += T.a[i] * (FFt.a[i] - P.a[i]);
    mub += 2.0f * tp;
    lamb += trT * J * (J - 1.0f);
    m33 Ftb = madd(Fb, mscale(mmul(T, F), 4.0f * mu));
    Ftb = msub(Ftb, mscale(mmul(U, mmul(Ph, mT(V))), 2.0f * mu));
    return madd(Ftb, mscale(mcof(F), lam * (2.0f * J - 1.0f) * trT));
}

// Adjoint of the generalized-Newtonian fluid (mat 10, stress_update_particle's viscous
// branch, mpm_utils.py:1056-1149 with the stress of stress_viscous): F = Jc I with
// Jc = det(F_trial)^(1/3), J = Jc^3, and
//   tau = J (2 eta_app D + p I),  D = dev sym L,  p = -bulk (J^-1.1 - 1),
//   eta_app = eta + tau_y / gd + K gd^(n-1),  gd = sqrt(2 D:D + eps^2), eps = 0.02.
// In: Fb = adjoint of F, tb = adjoint of tau. Returns the adjoint of F_trial; Lb is the
// adjoint of L; the partials of eta, bulk, tau_y, K and n are added.
MPM_DEV m33 adj_viscous(const m33& Ft, const m33& L, float bulk, float eta, float ty, float K, float n,
                        const m33& Fb, const m33& tb_in, m33& Lb, float& etab, float& bulkb, float& tyb,
                        float& Kb, float& nb) {
    const float eps = 0.02f;
    const float d = mdet(Ft);
    const float Jc = powf(d, 1.0f / 3.0f);
    const float J = Jc * Jc * Jc;
    const float pJ = powf(J, -1.1f);
    const float pressure = -bulk * (pJ - 1.0f);
    m33 D = mscale(madd(L, mT(L)), 0.5f);
    const float trd = (D.a[0] + D.a[4] + D.a[8]) / 3.0f;
    D.a[0] -= trd; D.a[4] -= trd; D.a[8] -= trd;
    float dd = 0.f;
    for (int i = 0; i < 9; ++i) dd += D.a[i] * D.a[i];
    const float gd = sqrtf(2.0f * dd + eps * eps);
    const float gp = powf(gd, n - 1.0f);
    const float eta_app = eta + ty / gd + K * gp;
    // the forward symmetrises tau
    const m33 T = mscale(madd(tb_in, mT(tb_in)), 0.5f);
    float TD = 0.f;
    for (int i = 0; i < 9; ++i) TD += T.a[i] * D.a[i];
    const float trT = T.a[0] + T.a[4] + T.a[8];
    // tau = J c: Jb = T : c, cb = J T; c = 2 eta_app D + p I
    float Jb = 2.0f * eta_app * TD + pressure * trT;
    const float eab = 2.0f * J * TD;
    const float pb = J * trT;
    etab += eab;
    tyb += eab / gd;
    Kb += eab * gp;
    nb += eab * K * gp * logf(gd);
    const float gdb = eab * (-ty / (gd * gd) + K * (n - 1.0f) * gp / gd);
    // Db = 2 eta_app J T + 2 D (gdb / gd); L enters through the projector dev sym
    m33 Db = madd(mscale(T, 2.0f * eta_app * J), mscale(D, 2.0f * gdb / gd));
    const float trb = (Db.a[0] + Db.a[4] + Db.a[8]) / 3.0f;
    Db.a[0] -= trb; Db.a[4] -= trb; Db.a[8] -= trb;
    Lb = Db;
    bulkb += -pb * (pJ - 1.0f);
    Jb += pb * 1.1f * bulk * pJ / J;
    const float Jcb = 3.0f * Jc * Jc * Jb + (Fb.a[0] + Fb.a[4] + Fb.a[8]);
    const float db = Jc != 0.0f ? Jcb / (3.0f * Jc * Jc) : 0.0f;
    return mscale(mcof(Ft), db);
}

// ------------------------------------------- surface colliders (revolved cup, mesh SDF)
// The contact of a cup or mesh-SDF collider with one node, recomputed with the forward
// rule's expressions (apply_collider of the grid update, mpm_solver_warp.py:2161-2225 and
// :2680-2736) so that the branch taken is the forward's. The normal (central differences
// of the cup SDF, trilinear SDF gradient of the mesh), the surface velocity v + w x r
// and the lever arm r depend on the node position and the collider pose only: constants
// of the substep for the state adjoint.
struct SurfaceHit {
    int mode;  // 1 grab (u = vs), 2 separable Coulomb on u - vs, 3 slip, 4 pass (cup, no normal)
    bool cup;  // the cup's form of the Coulomb scaling (else the mesh collider's)
    v3 vs, n, rel;
};

MPM_DEV bool surface_hit(const mpm_collider_t& c, const mpm_grid_t& G, int gi, int gj, int gk, SurfaceHit& h) {
    const v3 xn = mk3((float)gi * G.dx, (float)gj * G.dx, (float)gk * G.dx);
    h.rel = xn - mk3(c.point[0], c.point[1], c.point[2]);
    h.vs = mk3(c.velocity[0], c.velocity[1], c.velocity[2]) + cross(mk3(c.omega[0], c.omega[1], c.omega[2]), h.rel);
    h.n = mk3(0.f, 0.f, 0.f);
    h.cup = c.kind == MPM_COLLIDER_CUP;
    if (h.cup) {
        m33 R;
#pragma unroll
        for (int q = 0; q < 9; ++q) R.a[q] = c.rot[q];
        const v3 loc = mTvmul(R, h.rel);
        const float sdf = glass_sdf(loc, c);
        if (!(sdf < c.contact_band)) return false;
        if (!(sdf > -c.sticky_depth)) {
            h.mode = 1;
            return true;
        }
        const float hd = 0.25f * G.dx;
        const v3 nl = mk3(glass_sdf(loc + mk3(hd, 0, 0), c) - glass_sdf(loc - mk3(hd, 0, 0), c),
                          glass_sdf(loc + mk3(0, hd, 0), c) - glass_sdf(loc - mk3(0, hd, 0), c),
                          glass_sdf(loc + mk3(0, 0, hd), c) - glass_sdf(loc - mk3(0, 0, hd), c));
        const float nlen = len(nl);
        if (!(nlen > 1.0e-12f)) {
            h.mode = 4;
            return true;
        }
        h.n = mvmul(R, vdiv(nl, nlen));
        h.mode = 2;
        return true;
    }
    const v3 body = quat_rotate(c.quat, h.rel, true);
    const v3 f = mk3((body.x - c.origin[0]) / c.cell, (body.y - c.origin[1]) / c.cell, (body.z - c.origin[2]) / c.cell);
    const float rf = (float)c.res - 1.0f;
    if (!(f.x >= 0.0f && f.y >= 0.0f && f.z >= 0.0f && f.x <= rf && f.y <= rf && f.z <= rf)) return false;
    float sd;
    trilerp(c.sdf_val, 1, c.res, f, &sd);
    if (!(sd <= c.band)) return false;
    float g[3];
    trilerp(c.sdf_grad, 3, c.res, f, g);
    const v3 gr = mk3(g[0], g[1], g[2]);
    if (!(len(gr) > 1.0e-12f)) return false;
    h.n = quat_rotate(c.quat, normalize(gr), false);
    h.mode = c.surface == 0 ? 1 : (c.surface == 1 ? 3 : 2);
    return true;
}

// the node velocity after the contact (forward)
MPM_DEV v3 surface_apply(const SurfaceHit& h, float fr, v3 vin) {
    if (h.mode == 1) return h.vs;
    if (h.mode == 4) return vin;
    const v3 vr = vin - h.vs;
    const float vn = dot(vr, h.n);
    if (h.mode == 3) return h.vs + (vr - h.n * vn);
    if (!(vn < 0.0f)) return vin;
    v3 vt = vr - h.n * vn;
    if (h.cup) {
        const float vtl = len(vt);
        if (vtl > 1.0e-20f) vt = vdiv(vt, vtl) * fmaxf(vtl + fr * vn, 0.0f);
    } else {
        const float tl = len(vt);
        if (tl > 1.0e-12f) vt = vt * (fmaxf(0.0f, tl + fr * vn) / tl);
    }
    return h.vs + vt;
}

// J^T g for J = d(velocity after the contact) / d(velocity before), at vin. Coulomb: with
// vr = vin - vs, vn = vr . n, vt = vr - n vn, the result is vs + vt (1 + fr vn / |vt|), so
// J^T g = (1 + fr vn / |vt|) P g + n fr (vt . g) / |vt| - vt fr vn (vt . g) / |vt|^3, P the
// tangential projector; zero where the friction stops the tangential motion.
MPM_DEV v3 surface_apply_T(const SurfaceHit& h, float fr, v3 vin, v3 g) {
    if (h.mode == 1) return mk3(0.f, 0.f, 0.f);
    if (h.mode == 4) return g;
    const v3 vr = vin - h.vs;
    const float vn = dot(vr, h.n);
    const v3 Pg = g - h.n * dot(h.n, g);
    if (h.mode == 3) return Pg;
    if (!(vn < 0.0f)) return g;
    const v3 vt = vr - h.n * vn;
    const float tl = len(vt);
    if (!(tl > (h.cup ? 1.0e-20f : 1.0e-12f))) return Pg;
    if (!(tl + fr * vn > 0.0f)) return mk3(0.f, 0.f, 0.f);
    const float sc = 1.0f + fr * vn / tl;
    const float a = dot(vt, g);
    return Pg * sc + h.n * (fr * a / tl) - vt * (fr * vn * a / (tl * tl * tl));
}
