/*
 * CPU oracle: a plain-C restatement of warpmpm's substep kernels.
 *
 * TEST INFRASTRUCTURE ONLY (the checker). Nothing in the product path links or calls it.
 * Every function cites the reference file:line it restates; paths are relative to the
 * reference repository root. Arithmetic is fp32 in the reference's expression order,
 * except the 3x3 SVD (Warp's svd3 is third-party code, warp-lang 1.14.0): here it is a
 * float64 Jacobi eigen-solve of A^T A rounded to fp32, using the same convention as the
 * McAdams et al. 2011 method Warp implements (U, V proper rotations, singular values
 * sorted by magnitude, the last one carrying the sign of det A). The SVD is pinned in
 * tests against numpy.linalg.svd and scipy.linalg.polar.
 */
#include "mpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* Planted-mistake builds for the tests that show a pin has teeth: compiled with
 * -DO_MUTANT=k, the oracle carries mistake k (a flipped sign, a transposed operand, a
 * dropped factor) at one place marked MUT(k). The checker build has O_MUTANT 0, where
 * every MUT(k) is false and the code is the restatement. */
#ifndef O_MUTANT
#define O_MUTANT 0
#endif
#define MUT(k) (O_MUTANT == (k))

/* Built twice: libmpm_oracle.so (the checker: single-threaded, particle order) and, with
 * -fopenmp -DO_OPENMP, libmpm_oracle_omp.so, used only to time the CPU baseline on every
 * host core: particle loops split over threads, scatter adds atomic (summation order then
 * varies, so that build is never the checker). */
#ifdef O_OPENMP
#define O_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#define O_ADD(lhs, val) _Pragma("omp atomic") lhs += (val)
#else
#define O_PARALLEL_FOR
#define O_ADD(lhs, val) lhs += (val)
#endif

/* ---------------------------------------------------------------- small linear algebra */
typedef struct { float a[9]; } m33;
typedef struct { float c[3]; } v3;

static inline v3 V3(float x, float y, float z) { v3 r = {{x, y, z}}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V3(a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.c[0] - b.c[0], a.c[1] - b.c[1], a.c[2] - b.c[2]); }
static inline v3 vscale(v3 a, float s) { return V3(a.c[0] * s, a.c[1] * s, a.c[2] * s); }
static inline float vdot(v3 a, v3 b) { return a.c[0] * b.c[0] + a.c[1] * b.c[1] + a.c[2] * b.c[2]; }
static inline float vlen(v3 a) { return sqrtf(vdot(a, a)); }
static inline v3 vcross(v3 a, v3 b) {
    return V3(a.c[1] * b.c[2] - a.c[2] * b.c[1], a.c[2] * b.c[0] - a.c[0] * b.c[2],
              a.c[0] * b.c[1] - a.c[1] * b.c[0]);
}
static inline v3 vdiv(v3 a, float s) { return V3(a.c[0] / s, a.c[1] / s, a.c[2] / s); }
static inline v3 vnormalize(v3 a) {
    float l = vlen(a);
    if (l > 0.0f) return vdiv(a, l);
    return V3(0.f, 0.f, 0.f);
}
static inline v3 vload(const float *p) { return V3(p[0], p[1], p[2]); }
static inline void vstore(float *p, v3 a) { p[0] = a.c[0]; p[1] = a.c[1]; p[2] = a.c[2]; }

static inline m33 mzero(void) { m33 r; memset(&r, 0, sizeof r); return r; }
static inline m33 meye(void) { m33 r = mzero(); r.a[0] = r.a[4] = r.a[8] = 1.f; return r; }
static inline m33 mload(const float *p) { m33 r; memcpy(r.a, p, sizeof r.a); return r; }
static inline void mstore(float *p, m33 m) { memcpy(p, m.a, sizeof m.a); }
static inline m33 mdiag(float a, float b, float c) { m33 r = mzero(); r.a[0] = a; r.a[4] = b; r.a[8] = c; return r; }
static inline m33 mT(m33 m) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[3 * i + j] = m.a[3 * j + i];
    return r;
}
static inline m33 mmul(m33 x, m33 y) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            r.a[3 * i + j] = x.a[3 * i + 0] * y.a[0 + j] + x.a[3 * i + 1] * y.a[3 + j] +
                             x.a[3 * i + 2] * y.a[6 + j];
    return r;
}
static inline v3 mvmul(m33 m, v3 v) {
    return V3(m.a[0] * v.c[0] + m.a[1] * v.c[1] + m.a[2] * v.c[2],
              m.a[3] * v.c[0] + m.a[4] * v.c[1] + m.a[5] * v.c[2],
              m.a[6] * v.c[0] + m.a[7] * v.c[1] + m.a[8] * v.c[2]);
}
static inline m33 madd(m33 x, m33 y) { m33 r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] + y.a[i]; return r; }
static inline m33 msub(m33 x, m33 y) { m33 r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] - y.a[i]; return r; }
static inline m33 mscale(m33 x, float s) { m33 r; for (int i = 0; i < 9; ++i) r.a[i] = x.a[i] * s; return r; }
static inline m33 mouter(v3 a, v3 b) {
    m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[3 * i + j] = a.c[i] * b.c[j];
    return r;
}
static inline float mdet(m33 m) {
    const float *a = m.a;
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
           a[2] * (a[3] * a[7] - a[4] * a[6]);
}

/* ----------------------------------------------------------------------------- SVD */
/* Symmetric 3x3 eigen-decomposition by cyclic Jacobi in float64 (Golub & Van Loan,
 * Matrix Computations, Alg. 8.5.2/8.5.3). S is overwritten; V receives the eigenvectors
 * as columns. */
static void jacobi_eig3(double S[3][3], double V[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2];
        double diag = S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2];
        if (off <= 1e-34 * diag || off == 0.0) break;
        for (int p = 0; p < 2; ++p) {
            for (int q = p + 1; q < 3; ++q) {
                if (S[p][q] == 0.0) continue;
                double theta = (S[q][q] - S[p][p]) / (2.0 * S[p][q]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) { /* S <- S J */
                    double skp = S[k][p], skq = S[k][q];
                    S[k][p] = c * skp - s * skq;
                    S[k][q] = s * skp + c * skq;
                }
                for (int k = 0; k < 3; ++k) { /* S <- J^T S */
                    double spk = S[p][k], sqk = S[q][k];
                    S[p][k] = c * spk - s * sqk;
                    S[q][k] = s * spk + c * sqk;
                }
                for (int k = 0; k < 3; ++k) {
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
        }
    }
}

/* A = U diag(s) V^T with U, V proper rotations, |s0| >= |s1| >= |s2|, s0, s1 >= 0 and
 * sign(s2) = sign(det A): the McAdams et al. 2011 convention that Warp's svd3 follows. */
EXPORT void o_svd3(const float *Af, float *Uf, float *sf, float *Vf) {
    double A[3][3], S[3][3], V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = Af[3 * i + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += A[k][i] * A[k][j];
            S[i][j] = MUT(13) ? (double)(float)acc : acc;
        }
    jacobi_eig3(S, V);
    double ev[3] = {S[0][0], S[1][1], S[2][2]};
    int ord[3] = {0, 1, 2};
    for (int a = 0; a < 3; ++a) /* sort eigenvalues descending */
        for (int b = a + 1; b < 3; ++b)
            if (ev[ord[b]] > ev[ord[a]]) { int t = ord[a]; ord[a] = ord[b]; ord[b] = t; }
    double Vs[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Vs[i][j] = V[i][ord[j]];
    double detV = Vs[0][0] * (Vs[1][1] * Vs[2][2] - Vs[1][2] * Vs[2][1]) -
                  Vs[0][1] * (Vs[1][0] * Vs[2][2] - Vs[1][2] * Vs[2][0]) +
                  Vs[0][2] * (Vs[1][0] * Vs[2][1] - Vs[1][1] * Vs[2][0]);
    if (detV < 0)
        for (int i = 0; i < 3; ++i) Vs[i][2] = -Vs[i][2];
    double B[3][3]; /* B = A V */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += A[i][k] * Vs[k][j];
            B[i][j] = acc;
        }
    /* Gram-Schmidt on the columns of B; u2 = u0 x u1 keeps U a proper rotation */
    double u0[3], u1[3], u2[3];
    double n0 = sqrt(B[0][0] * B[0][0] + B[1][0] * B[1][0] + B[2][0] * B[2][0]);
    if (n0 > 1e-300) {
        for (int i = 0; i < 3; ++i) u0[i] = B[i][0] / n0;
    } else {
        u0[0] = 1; u0[1] = 0; u0[2] = 0;
    }
    double b1[3] = {B[0][1], B[1][1], B[2][1]};
    double d01 = u0[0] * b1[0] + u0[1] * b1[1] + u0[2] * b1[2];
    for (int i = 0; i < 3; ++i) b1[i] -= d01 * u0[i];
    double n1 = sqrt(b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]);
    if (n1 > 1e-300 && n1 > 1e-14 * n0) {
        for (int i = 0; i < 3; ++i) u1[i] = b1[i] / n1;
    } else { /* any unit vector orthogonal to u0 */
        double t[3] = {0, 0, 0};
        int m = fabs(u0[0]) < fabs(u0[1]) ? (fabs(u0[0]) < fabs(u0[2]) ? 0 : 2)
                                          : (fabs(u0[1]) < fabs(u0[2]) ? 1 : 2);
        t[m] = 1.0;
        double d = u0[0] * t[0] + u0[1] * t[1] + u0[2] * t[2];
        for (int i = 0; i < 3; ++i) t[i] -= d * u0[i];
        double nt = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int i = 0; i < 3; ++i) u1[i] = t[i] / nt;
        n1 = 0.0;
    }
    u2[0] = u0[1] * u1[2] - u0[2] * u1[1];
    u2[1] = u0[2] * u1[0] - u0[0] * u1[2];
    u2[2] = u0[0] * u1[1] - u0[1] * u1[0];
    double s2 = u2[0] * B[0][2] + u2[1] * B[1][2] + u2[2] * B[2][2];
    for (int i = 0; i < 3; ++i) {
        Uf[3 * i + 0] = (float)u0[i];
        Uf[3 * i + 1] = (float)u1[i];
        Uf[3 * i + 2] = (float)u2[i];
        for (int j = 0; j < 3; ++j) Vf[3 * i + j] = (float)Vs[i][j];
    }
    sf[0] = (float)n0;
    sf[1] = (float)n1;
    sf[2] = (float)s2;
    if (MUT(13)) /* singular values read off the eigenvalues of a float32 A^T A */
        for (int a = 0; a < 3; ++a) {
            double e = ev[ord[a]] > 0.0 ? sqrt(ev[ord[a]]) : 0.0;
            sf[a] = (float)(a == 2 && s2 < 0.0 ? -e : e);
        }
    if (MUT(14) && s2 < 0.0) { /* the reflection carried by the largest axis, not the smallest */
        sf[0] = -sf[0];
        sf[2] = -sf[2];
        for (int i = 0; i < 3; ++i) {
            Uf[3 * i + 0] = -Uf[3 * i + 0];
            Uf[3 * i + 2] = -Uf[3 * i + 2];
        }
    }
}

static inline void svd3(m33 A, m33 *U, v3 *s, m33 *V) { o_svd3(A.a, U->a, s->c, V->a); }

/* ------------------------------------------------------------ constitutive functions */
/* kirchoff_stress_FCR, src/warpmpm/kernels/mpm_utils.py:8-15 */
static m33 stress_fcr(m33 F, m33 U, m33 V, float J, float mu, float lam) {
    m33 R = mmul(U, mT(V));
    return madd(mmul(mscale(msub(F, R), 2.0f * mu), mT(F)), mscale(meye(), lam * J * (J - 1.0f)));
}

/* kirchoff_stress_water, mpm_utils.py:17-25 */
static m33 stress_water(float J, float bulk) {
    float pressure = -bulk * (powf(J, -1.1f) - 1.0f);
    return mscale(mscale(meye(), pressure), J);
}

static inline float ddot_dev(m33 D) {
    float s = 0.f;
    for (int i = 0; i < 9; ++i) s += D.a[i] * D.a[i];
    return s;
}

/* kirchoff_stress_newtonian, mpm_utils.py:27-54 */
static m33 stress_newtonian(float J, float bulk, m33 L, float eta, float tau_y, float pk, float pn) {
    const float eps = 0.02f;
    float pressure = -bulk * (powf(J, -1.1f) - 1.0f);
    m33 D = mscale(madd(L, mT(L)), 0.5f);
    float trd = (D.a[0] + D.a[4] + D.a[8]) / 3.0f;
    m33 Dd = msub(D, mscale(meye(), trd));
    float dd = ddot_dev(Dd);
    float gd = sqrtf(2.0f * dd + eps * eps);
    float eta_app = eta + tau_y / gd + pk * powf(gd, pn - 1.0f);
    m33 cauchy = madd(mscale(meye(), pressure), mscale(Dd, 2.0f * eta_app));
    return mscale(cauchy, J);
}

/* kirchoff_stress_tabulated, mpm_utils.py:56-85 */
static m33 stress_tabulated(float J, float bulk, m33 L, const float *table, float smin, float smax, int n) {
    const float eps = 0.02f;
    float pressure = -bulk * (powf(J, -1.1f) - 1.0f);
    m33 D = mscale(madd(L, mT(L)), 0.5f);
    float trd = (D.a[0] + D.a[4] + D.a[8]) / 3.0f;
    m33 Dd = msub(D, mscale(meye(), trd));
    float dd = ddot_dev(Dd);
    float gd = sqrtf(2.0f * dd + eps * eps);
    float s = logf(gd) * 0.43429448190325176f;
    float idx = (s - smin) / (smax - smin) * (float)(n - 1);
    float hi = (float)(n - 1) - 1.0e-6f;
    idx = idx < 0.f ? 0.f : (idx > hi ? hi : idx);
    int i0 = (int)floorf(idx);
    float frac = idx - (float)i0;
    float eta_app = table[i0] * (1.0f - frac) + table[i0 + 1] * frac;
    m33 cauchy = madd(mscale(meye(), pressure), mscale(Dd, 2.0f * eta_app));
    return mscale(cauchy, J);
}

/* kirchoff_stress_drucker_prager, mpm_utils.py:109-124 */
static m33 stress_drucker_prager(m33 F, m33 U, m33 V, v3 sig, float mu, float lam) {
    float l0 = logf(sig.c[0]), l1 = logf(sig.c[1]), l2 = logf(sig.c[2]);
    float lsum = l0 + l1 + l2;
    float c0 = 2.0f * mu * l0 * (1.0f / sig.c[0]) + lam * lsum * (1.0f / sig.c[0]);
    float c1 = 2.0f * mu * l1 * (1.0f / sig.c[1]) + lam * lsum * (1.0f / sig.c[1]);
    float c2 = 2.0f * mu * l2 * (1.0f / sig.c[2]) + lam * lsum * (1.0f / sig.c[2]);
    return mmul(mmul(mmul(U, mdiag(c0, c1, c2)), mT(V)), mT(F));
}

/* kirchoff_stress_hencky, mpm_utils.py:228-243 */
static m33 stress_hencky(m33 U, v3 sig, float mu, float lam) {
    float e0 = logf(fmaxf(sig.c[0], 1e-6f)), e1 = logf(fmaxf(sig.c[1], 1e-6f)),
          e2 = logf(fmaxf(sig.c[2], 1e-6f));
    float tr = e0 + e1 + e2;
    m33 tau = mdiag(2.0f * mu * e0 + lam * tr, 2.0f * mu * e1 + lam * tr, 2.0f * mu * e2 + lam * tr);
    return mmul(mmul(U, tau), mT(U));
}

/* von_mises_return_mapping, mpm_utils.py:127-170 */
static m33 rm_von_mises(m33 Ft, const OModel *md, int p, int mat) {
    m33 U, V; v3 so;
    svd3(Ft, &U, &so, &V);
    v3 sig = V3(fmaxf(so.c[0], 0.01f), fmaxf(so.c[1], 0.01f), fmaxf(so.c[2], 0.01f));
    v3 eps = V3(logf(sig.c[0]), logf(sig.c[1]), logf(sig.c[2]));
    float temp = (eps.c[0] + eps.c[1] + eps.c[2]) / 3.0f;
    float mu = md->mu[p], lam = md->lam[p];
    float trl = lam * (eps.c[0] + eps.c[1] + eps.c[2]);
    v3 tau = V3(2.0f * mu * eps.c[0] + trl, 2.0f * mu * eps.c[1] + trl, 2.0f * mu * eps.c[2] + trl);
    float st = tau.c[0] + tau.c[1] + tau.c[2];
    v3 cond = V3(tau.c[0] - st / 3.0f, tau.c[1] - st / 3.0f, tau.c[2] - st / 3.0f);
    if (vlen(cond) > md->yield_stress[p]) {
        v3 eh = vsub(eps, V3(temp, temp, temp));
        float ehn = vlen(eh) + 1e-6f;
        float dg = ehn - md->yield_stress[p] / (2.0f * mu);
        eps = vsub(eps, vscale(eh, dg / ehn));
        m33 Fe = mmul(mmul(U, mdiag(expf(eps.c[0]), expf(eps.c[1]), expf(eps.c[2]))), mT(V));
        if (md->hardening[mat] > 0.5f)
            md->yield_stress[p] = md->yield_stress[p] + 2.0f * mu * md->xi[mat] * dg;
        return Fe;
    }
    return Ft;
}

/* von_mises_return_mapping_with_damage, mpm_utils.py:172-225 */
static m33 rm_von_mises_damage(m33 Ft, const OModel *md, int p, int mat) {
    m33 U, V; v3 so;
    svd3(Ft, &U, &so, &V);
    v3 sig = V3(fmaxf(so.c[0], 0.01f), fmaxf(so.c[1], 0.01f), fmaxf(so.c[2], 0.01f));
    v3 eps = V3(logf(sig.c[0]), logf(sig.c[1]), logf(sig.c[2]));
    float temp = (eps.c[0] + eps.c[1] + eps.c[2]) / 3.0f;
    float mu = md->mu[p], lam = md->lam[p];
    float trl = lam * (eps.c[0] + eps.c[1] + eps.c[2]);
    v3 tau = V3(2.0f * mu * eps.c[0] + trl, 2.0f * mu * eps.c[1] + trl, 2.0f * mu * eps.c[2] + trl);
    float st = tau.c[0] + tau.c[1] + tau.c[2];
    v3 cond = V3(tau.c[0] - st / 3.0f, tau.c[1] - st / 3.0f, tau.c[2] - st / 3.0f);
    if (vlen(cond) > md->yield_stress[p]) {
        if (md->yield_stress[p] <= 0.0f) return Ft;
        v3 eh = vsub(eps, V3(temp, temp, temp));
        float ehn = vlen(eh) + 1e-6f;
        float dg = ehn - md->yield_stress[p] / (2.0f * mu);
        eps = vsub(eps, vscale(eh, dg / ehn));
        md->yield_stress[p] = md->yield_stress[p] - md->softening[mat] * vlen(vscale(eh, dg / ehn));
        if (md->yield_stress[p] <= 0.0f) {
            md->mu[p] = 0.0f;
            md->lam[p] = 0.0f;
        }
        m33 Fe = mmul(mmul(U, mdiag(expf(eps.c[0]), expf(eps.c[1]), expf(eps.c[2]))), mT(V));
        if (md->hardening[mat] > 0.5f)
            md->yield_stress[p] = md->yield_stress[p] + 2.0f * md->mu[p] * md->xi[mat] * dg;
        return Fe;
    }
    return Ft;
}

/* mu_i_return_mapping (mat 9), mpm_utils.py:246-324; with tab != 0 the tabulated
 * variant mu_i_tabulated_return_mapping (mat 13), mpm_utils.py:327-407 */
static m33 rm_mu_i(m33 Ft, const OModel *md, int p, int mat, float dt, int tab) {
    m33 U, V; v3 sig;
    svd3(Ft, &U, &sig, &V);
    v3 sc = V3(fmaxf(fabsf(sig.c[0]), 1e-6f), fmaxf(fabsf(sig.c[1]), 1e-6f), fmaxf(fabsf(sig.c[2]), 1e-6f));
    v3 eps = V3(logf(sc.c[0]), logf(sc.c[1]), logf(sc.c[2]));
    float tr = eps.c[0] + eps.c[1] + eps.c[2];
    m33 Fe = Ft;
    if (tr >= 0.0f) {
        Fe = mmul(U, mT(V));
    } else {
        float G = md->mu[p], lam = md->lam[p];
        float J = sc.c[0] * sc.c[1] * sc.c[2];
        float me = tr / 3.0f;
        float d0 = eps.c[0] - me, d1 = eps.c[1] - me, d2 = eps.c[2] - me;
        float dn = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
        float pK = -(2.0f * G / 3.0f + lam) * tr;
        float tbK = 2.0f * G * dn / sqrtf(2.0f);
        float pC = pK / J;
        float mu_s = md->muI_mu_s[mat], dmu = md->muI_delta_mu[mat], I0 = md->muI_I0[mat];
        float thr = tab ? md->eta_table[0] : mu_s;
        if (tbK > thr * pK) {
            float Ic = md->muI_d[mat] * sqrtf(md->muI_rho_s[mat] / pC);
            float lo = 0.0f, hi = tbK / (G * dt);
            for (int it = 0; it < 48; ++it) {
                float mid = 0.5f * (lo + hi);
                float Im = mid * Ic;
                float mum;
                if (tab) {
                    float s = logf(fmaxf(Im, 1e-12f)) * 0.43429448190325176f;
                    float smin = md->eta_table_smin, smax = md->eta_table_smax;
                    int nt = md->eta_table_n;
                    float idx = (s - smin) / (smax - smin) * (float)(nt - 1);
                    float ih = (float)(nt - 1) - 1.0e-6f;
                    idx = idx < 0.f ? 0.f : (idx > ih ? ih : idx);
                    int i0 = (int)floorf(idx);
                    float fr = idx - (float)i0;
                    mum = md->eta_table[i0] * (1.0f - fr) + md->eta_table[i0 + 1] * fr;
                } else {
                    mum = mu_s + dmu * Im / (Im + I0);
                }
                float g = tbK - G * dt * mid - mum * pK;
                if (g > 0.0f) lo = mid; else hi = mid;
            }
            float gdp = 0.5f * (lo + hi);
            float tbn = tbK - G * dt * gdp;
            float scale = tbn / fmaxf(tbK, 1e-20f);
            float e0 = d0 * scale + me, e1 = d1 * scale + me, e2 = d2 * scale + me;
            Fe = mmul(mmul(U, mdiag(expf(e0), expf(e1), expf(e2))), mT(V));
        }
    }
    return Fe;
}

/* mu_i_phi_pressure, mpm_utils.py:421-430 */
static float mu_i_phi_pressure(const OModel *md, int mat, float J, float Jp_ref) {
    float K = md->E[mat] / (3.0f * (1.0f - 2.0f * md->nu[mat]));
    float pr = K * (Jp_ref / fmaxf(J, 1e-6f) - 1.0f);
    return fmaxf(pr, 0.0f);
}

/* mu_i_phi_return_mapping (mat 11), mpm_utils.py:433-522 (with j_cs_of_I :410-418) */
static m33 rm_mu_i_phi(m33 Ft, OState *st, const OModel *md, int p, int mat, float dt) {
    m33 U, V; v3 sig;
    svd3(Ft, &U, &sig, &V);
    v3 sc = V3(fmaxf(fabsf(sig.c[0]), 1e-6f), fmaxf(fabsf(sig.c[1]), 1e-6f), fmaxf(fabsf(sig.c[2]), 1e-6f));
    v3 eps = V3(logf(sc.c[0]), logf(sc.c[1]), logf(sc.c[2]));
    float J = sc.c[0] * sc.c[1] * sc.c[2];
    float tr = eps.c[0] + eps.c[1] + eps.c[2];
    float me = tr / 3.0f;
    float d0 = eps.c[0] - me, d1 = eps.c[1] - me, d2 = eps.c[2] - me;
    float dn = sqrtf(d0 * d0 + d1 * d1 + d2 * d2);
    float G = md->mu[p];
    float Jp_ref = st->Jp[p];
    if (Jp_ref < 0.5f) Jp_ref = 1.0f;
    float p_comp = mu_i_phi_pressure(md, mat, J, Jp_ref);
    const float beta = 8.0f;
    m33 Fe = Ft;
    st->Jp[p] = Jp_ref;
    if (p_comp <= 0.0f) {
        float s_iso = expf(me);
        Fe = mmul(mmul(U, mdiag(s_iso, s_iso, s_iso)), mT(V));
    } else {
        float pK = p_comp * J;
        float tbK = 2.0f * G * dn / sqrtf(2.0f);
        float mu_s = md->muI_mu_s[mat], dmu = md->muI_delta_mu[mat], I0 = md->muI_I0[mat];
        if (tbK > mu_s * pK) {
            float Ic = md->muI_d[mat] * sqrtf(md->muI_rho_s[mat] / p_comp);
            float lo = 0.0f, hi = tbK / (G * dt);
            for (int it = 0; it < 48; ++it) {
                float mid = 0.5f * (lo + hi);
                float Im = mid * Ic;
                float mum = mu_s + dmu * Im / (Im + I0);
                float g = tbK - G * dt * mid - mum * pK;
                if (g > 0.0f) lo = mid; else hi = mid;
            }
            float gdp = 0.5f * (lo + hi);
            float I_real = gdp * Ic;
            float dgam = fminf(beta * gdp * dt, 1.0f);
            float chi = md->muI_phi_chi[mat];
            float denom = fmaxf(1.0f - chi * I_real / (I_real + I0), 1e-3f);
            float j_cs = 1.0f / denom;
            st->Jp[p] = Jp_ref + dgam * (j_cs - Jp_ref);
            float tbn = tbK - G * dt * gdp;
            float scale = tbn / fmaxf(tbK, 1e-20f);
            float e0 = d0 * scale + me, e1 = d1 * scale + me, e2 = d2 * scale + me;
            Fe = mmul(mmul(U, mdiag(expf(e0), expf(e1), expf(e2))), mT(V));
        }
    }
    return Fe;
}

/* kirchoff_stress_mu_i_phi, mpm_utils.py:525-540 */
static m33 stress_mu_i_phi(m33 U, v3 sig, float G, float p_comp) {
    v3 sc = V3(fmaxf(sig.c[0], 1e-6f), fmaxf(sig.c[1], 1e-6f), fmaxf(sig.c[2], 1e-6f));
    v3 eps = V3(logf(sc.c[0]), logf(sc.c[1]), logf(sc.c[2]));
    float J = sc.c[0] * sc.c[1] * sc.c[2];
    float me = (eps.c[0] + eps.c[1] + eps.c[2]) / 3.0f;
    float iso = -p_comp * J;
    m33 tau = mdiag(2.0f * G * (eps.c[0] - me) + iso, 2.0f * G * (eps.c[1] - me) + iso,
                    2.0f * G * (eps.c[2] - me) + iso);
    return mmul(mmul(U, tau), mT(U));
}

/* sand_return_mapping (mat 2), mpm_utils.py:543-580 */
static m33 rm_sand(m33 Ft, const OModel *md, int p, int mat) {
    m33 U, V; v3 sig;
    svd3(Ft, &U, &sig, &V);
    v3 eps = V3(logf(fmaxf(fabsf(sig.c[0]), 1e-14f)), logf(fmaxf(fabsf(sig.c[1]), 1e-14f)),
                logf(fmaxf(fabsf(sig.c[2]), 1e-14f)));
    float tr = eps.c[0] + eps.c[1] + eps.c[2];
    v3 eh = vsub(eps, V3(tr / 3.0f, tr / 3.0f, tr / 3.0f));
    float ehn = vlen(eh);
    float mu = md->mu[p], lam = md->lam[p];
    float dg = ehn + (3.0f * lam + 2.0f * mu) / (2.0f * mu) * tr * md->alpha[mat];
    m33 Fe = Ft;
    if (dg <= 0.0f) Fe = Ft;
    if (dg > 0.0f && tr > 0.0f) Fe = mmul(U, mT(V));
    if (dg > 0.0f && tr <= 0.0f) {
        v3 H = vsub(eps, vscale(eh, dg / ehn));
        Fe = mmul(mmul(U, mdiag(expf(H.c[0]), expf(H.c[1]), expf(H.c[2]))), mT(V));
    }
    return Fe;
}

/* stress_update_particle, mpm_utils.py:1056-1149 */
static void stress_update_particle(OState *st, OModel *md, float dt, int p) {
    int mat = st->material[p];
    m33 Ft = mload(st->F_trial + 9 * p);
    m33 F;
    if (mat == 1) F = rm_von_mises(Ft, md, p, mat);
    else if (mat == 2) F = rm_sand(Ft, md, p, mat);
    else if (mat == 5) F = rm_von_mises_damage(Ft, md, p, mat);
    else if (mat == 9) F = rm_mu_i(Ft, md, p, mat, dt, 0);
    else if (mat == 13) F = rm_mu_i(Ft, md, p, mat, dt, 1);
    else if (mat == 11) F = rm_mu_i_phi(Ft, st, md, p, mat, dt);
    else if (mat == 6 || mat == 10 || mat == 12) {
        float J = mdet(Ft);
        float Jc = powf(J, 1.0f / 3.0f);
        F = mdiag(Jc, Jc, Jc);
    } else if (mat == 7 || mat == 8) F = meye();
    else F = Ft;
    mstore(st->F + 9 * p, F);

    float J = mdet(F);
    m33 U = mzero(), V = mzero(); v3 sig = V3(0, 0, 0);
    m33 S = mzero();
    if (mat != 6 && mat != 7 && mat != 8 && mat != 10 && mat != 12) svd3(F, &U, &sig, &V);
    if (mat == 0 || mat == 5) S = stress_fcr(F, U, V, J, md->mu[p], md->lam[p]);
    if (mat == 1) S = stress_hencky(U, sig, md->mu[p], md->lam[p]);
    if (mat == 2) S = stress_drucker_prager(F, U, V, sig, md->mu[p], md->lam[p]);
    if (mat == 6) S = stress_water(J, md->bulk[mat]);
    if (mat == 10)
        S = stress_newtonian(J, md->bulk[mat], mload(st->L + 9 * p), md->plastic_viscosity[mat],
                             md->yield_stress[p], md->hardening[mat], md->softening[mat]);
    if (mat == 12)
        S = stress_tabulated(J, md->bulk[mat], mload(st->L + 9 * p), md->eta_table,
                             md->eta_table_smin, md->eta_table_smax, md->eta_table_n);
    if (mat == 9 || mat == 13) S = stress_hencky(U, sig, md->mu[p], md->lam[p]);
    if (mat == 11) {
        float Jp_ref = st->Jp[p];
        if (Jp_ref < 0.5f) Jp_ref = 1.0f;
        float pc = mu_i_phi_pressure(md, mat, J, Jp_ref);
        S = stress_mu_i_phi(U, sig, md->mu[p], pc);
    }
    S = mscale(madd(S, mT(S)), 0.5f);
    mstore(st->stress + 9 * p, S);
}

/* ------------------------------------------------------------------- B-spline stencil */
/* quadratic B-spline weights, mpm_utils.py:826-849 (p2g) and :947-970 (g2p); w[axis][node] */
typedef struct { int base[3]; v3 fx; float w[3][3]; float dw[3][3]; } Stencil;

static Stencil stencil(const OModel *md, v3 x) {
    Stencil s;
    v3 gp = vscale(x, md->inv_dx);
    for (int a = 0; a < 3; ++a) {
        s.base[a] = (int)floorf(gp.c[a] - 0.5f);
        s.fx.c[a] = gp.c[a] - (float)s.base[a];
        float f = s.fx.c[a];
        float wa = 1.5f - f, wb = f - 1.0f, wc = f - 0.5f;
        s.w[a][0] = wa * wa * 0.5f;
        s.w[a][1] = (0.0f - wb * wb) + 0.75f;
        s.w[a][2] = wc * wc * 0.5f;
        s.dw[a][0] = f - 1.5f;
        s.dw[a][1] = -2.0f * (f - 1.0f);
        s.dw[a][2] = f - 0.5f;
    }
    return s;
}

/* compute_dweight, mpm_utils.py:601-610 */
static inline v3 dweight(const OModel *md, const Stencil *s, int i, int j, int k) {
    v3 d = V3(s->dw[0][i] * s->w[1][j] * s->w[2][k], s->w[0][i] * s->dw[1][j] * s->w[2][k],
              s->w[0][i] * s->w[1][j] * s->dw[2][k]);
    return vscale(d, md->inv_dx);
}

static inline long gidx(const OModel *md, int i, int j, int k) {
    return ((long)i * md->ny + j) * md->nz + k;
}

static inline int wrap_x(const OModel *md, int ix) {
    if (md->periodic_x) ix = ((ix % md->nx) + md->nx) % md->nx;
    return ix;
}

/* ------------------------------------------------------- CPIC transfer functions */
/* cdf_pair_incompatible, mpm_utils.py:656-662: a lane valid on both with differing sides */
static inline int cdf_incompat(int pc, int nt) {
    int both_valid = pc & nt & 0x55;
    int sides_differ = (pc ^ nt) >> 1;
    return (both_valid & sides_differ) != 0;
}

static inline long cdf_bidx(const OCdf *c, int bi, int bj, int bk) {
    return ((long)bi * c->nb[1] + bj) * c->nb[2] + bk;
}

/* cdf_stencil_near, mpm_utils.py:665-683: any 4^3 block under the 3x3x3 stencil tagged */
static int cdf_near(const OCdf *c, const Stencil *s) {
    for (int bi = s->base[0] >> 2; bi <= (s->base[0] + 2) >> 2; ++bi)
        for (int bj = s->base[1] >> 2; bj <= (s->base[1] + 2) >> 2; ++bj)
            for (int bk = s->base[2] >> 2; bk <= (s->base[2] + 2) >> 2; ++bk)
                if (c->block[cdf_bidx(c, bi, bj, bk)] != 0) return 1;
    return 0;
}

/* cdf_particle_tag, mpm_utils.py:686-724: affinity = any valid node in the stencil, side =
 * sign of the weighted vote sum w * d * side per lane (>= 0 -> side bit set) */
static int cdf_vote(const OCdf *c, const OModel *md, const Stencil *s, int use_prev) {
    float votes[O_MAX_CDF] = {0.f, 0.f, 0.f, 0.f};
    int seen = 0;
    const int32_t *tg = use_prev ? c->tag_prev : c->tag;
    const float *dg = use_prev ? c->d_prev : c->d;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                long g = gidx(md, s->base[0] + i, s->base[1] + j, s->base[2] + k);
                int t = tg[g];
                float d = dg[g];
                if (t == 0) continue;
                float weight = s->w[0][i] * s->w[1][j] * s->w[2][k];
                for (int l = 0; l < O_MAX_CDF; ++l)
                    if ((t >> (2 * l)) & 1) {
                        seen |= 1 << (2 * l);
                        float sg = ((t >> (2 * l + 1)) & 1) ? 1.0f : -1.0f;
                        votes[l] = votes[l] + weight * d * sg;
                    }
            }
    int pc = 0;
    for (int l = 0; l < O_MAX_CDF; ++l)
        if ((seen >> (2 * l)) & 1) {
            pc |= 1 << (2 * l);
            if (votes[l] >= 0.0f) pc |= 1 << (2 * l + 1);
        }
    return pc;
}

/* cdf_persist_tag, mpm_utils.py:727-741: a lane kept in the kernel keeps its old side */
static int cdf_persist(int old, int fresh) {
    int pc = 0;
    for (int l = 0; l < O_MAX_CDF; ++l)
        if ((fresh >> (2 * l)) & 1) {
            pc |= 1 << (2 * l);
            if ((old >> (2 * l)) & 1) pc |= old & (1 << (2 * l + 1));
            else pc |= fresh & (1 << (2 * l + 1));
        }
    return pc;
}

/* particle tag for one half of the pass (mpm_utils.py:851-860 scatter, :971-982 gather):
 * block pretest (skipped with mask_off), vote, persistence against the stored tag */
static int cdf_particle_pc(const OState *st, const OModel *md, const Stencil *s, int p, int use_prev) {
    const OCdf *c = st->cdf;
    if (!c || c->n_cdf <= 0) return 0;
    if (c->mask_off != 0 || cdf_near(c, s)) return cdf_persist(c->ptag[p], cdf_vote(c, md, s, use_prev));
    return 0;
}

/* cdf_particle_ghost, mpm_utils.py:744-816: the particle's free velocity projected against
 * its primary surface (reconstructed from the prev tag grids), plus a separation push */
static v3 cdf_ghost(const OState *st, const OModel *md, float dt, int p, int pc, const Stencil *s) {
    const OCdf *c = st->cdf;
    float d_hat = 0.0f, best_w = 0.0f;
    v3 n_acc = V3(0.f, 0.f, 0.f);
    int lane_p = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                long g = gidx(md, s->base[0] + i, s->base[1] + j, s->base[2] + k);
                int t = c->tag_prev[g];
                if (t == 0) continue;
                int own = (t >> O_CDF_OWNER_SHIFT) & 3;
                if (((t >> (2 * own)) & 1) == 0) continue;
                float weight = s->w[0][i] * s->w[1][j] * s->w[2][k];
                float sg = ((t >> (2 * own + 1)) & 1) ? 1.0f : -1.0f;
                float d = c->d_prev[g];
                d_hat = d_hat + weight * sg * d;
                v3 dwt = dweight(md, s, i, j, k);
                n_acc = vadd(n_acc, vscale(vscale(dwt, sg), d));
                if (weight > best_w) {
                    best_w = weight;
                    lane_p = own;
                }
            }
    v3 v_free = vadd(vload(st->v + 3 * p), vscale(V3(md->gravity[0], md->gravity[1], md->gravity[2]), dt));
    v3 ghost = v_free;
    if (vlen(n_acc) > 1.0e-12f) {
        v3 n_hat = vnormalize(n_acc);
        float side = ((pc >> (2 * lane_p + 1)) & 1) ? 1.0f : -1.0f;
        v3 n_par = vscale(n_hat, side);
        v3 rel = vsub(vload(st->x + 3 * p), vload(c->lane_center_prev[lane_p]));
        v3 v_surf = vadd(vload(c->lane_velocity_prev[lane_p]), vcross(vload(c->lane_omega_prev[lane_p]), rel));
        v3 v_rel = vsub(v_free, v_surf);
        float vn = vdot(v_rel, n_par);
        v3 v_tan = vsub(v_rel, vscale(n_par, vn));
        float cpush = 0.0f;
        float d_side = d_hat * side;
        float eps = 0.05f * md->dx;
        if (d_side < eps) cpush = fminf((eps - d_side) / dt, 0.5f * md->dx / dt);
        int stype = c->lane_type[lane_p];
        if (stype == 0) {
            ghost = vadd(v_surf, vscale(n_par, cpush));
        } else if (stype == 1) {
            ghost = vadd(vadd(v_surf, v_tan), vscale(n_par, cpush));
        } else {
            float dvn = fmaxf(-vn, 0.0f) + cpush;
            float tlen = vlen(v_tan);
            if (tlen > 1.0e-12f && dvn > 0.0f) {
                float scale = fmaxf(0.0f, tlen - c->lane_friction[lane_p] * dvn) / tlen;
                v_tan = vscale(v_tan, scale);
            }
            ghost = vadd(vadd(vadd(v_surf, v_tan), vscale(n_par, fmaxf(vn, 0.0f))), vscale(n_par, cpush));
        }
    }
    return ghost;
}

/* p2g_particle, mpm_utils.py:819-916, with the CPIC blocked deposits of :851-912 */
static void p2g_particle(OState *st, const OModel *md, float dt, int p) {
    m33 stress = mload(st->stress + 9 * p);
    Stencil s = stencil(md, vload(st->x + 3 * p));
    const int pc = cdf_particle_pc(st, md, &s, p, 0);
    OCdf *cd = st->cdf;
    m33 C = mload(st->C + 9 * p);
    C = madd(mscale(C, 1.0f - md->rpic_damping), mscale(msub(C, mT(C)), md->rpic_damping / 2.0f));
    if (md->rpic_damping < -0.001f) C = mzero();
    float vol = st->vol[p], mass = st->mass[p];
    v3 vp = vload(st->v + 3 * p);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                v3 dpos = vscale(vsub(V3((float)i, (float)j, (float)k), s.fx), md->dx);
                int ix = wrap_x(md, s.base[0] + i), iy = s.base[1] + j, iz = s.base[2] + k;
                float weight = s.w[0][i] * s.w[1][j] * s.w[2][k];
                v3 dwt = dweight(md, &s, i, j, k);
                v3 ef = mvmul(mscale(stress, -vol), dwt);
                v3 add = vadd(vscale(vadd(vp, mvmul(C, dpos)), weight * mass), vscale(ef, dt));
                long g = gidx(md, ix, iy, iz);
                if (pc != 0) {
                    int nt = cd->tag[g];
                    if (cdf_incompat(pc, nt)) {
                        int own = (nt >> O_CDF_OWNER_SHIFT) & 3;
                        v3 xw = vscale(V3((float)ix, (float)iy, (float)iz), md->dx);
                        v3 tq = vcross(vsub(xw, vload(cd->lane_center[own])), add);
                        for (int c = 0; c < 3; ++c) {
                            O_ADD(cd->reaction_force[own][c], add.c[c]);
                            O_ADD(cd->reaction_torque[own][c], tq.c[c]);
                        }
                        continue;
                    }
                }
                O_ADD(st->grid_v_in[3 * g + 0], add.c[0]);
                O_ADD(st->grid_v_in[3 * g + 1], add.c[1]);
                O_ADD(st->grid_v_in[3 * g + 2], add.c[2]);
                O_ADD(st->grid_m[g], weight * mass);
            }
}

/* update_cov, mpm_utils.py:613-633 */
static void update_cov(OState *st, int p, m33 gv, float dt) {
    float *c = st->cov + 6 * p;
    m33 cn = mzero();
    cn.a[0] = c[0]; cn.a[1] = c[1]; cn.a[2] = c[2];
    cn.a[3] = c[1]; cn.a[4] = c[3]; cn.a[5] = c[4];
    cn.a[6] = c[2]; cn.a[7] = c[4]; cn.a[8] = c[5];
    if (MUT(10)) gv = mT(gv);
    m33 n1 = madd(cn, mscale(madd(mmul(gv, cn), mmul(cn, mT(gv))), dt));
    c[0] = n1.a[0]; c[1] = n1.a[1]; c[2] = n1.a[2];
    c[3] = n1.a[4]; c[4] = n1.a[5]; c[5] = n1.a[8];
}

/* g2p_particle, mpm_utils.py:944-1043, with the CPIC tag store and ghost fill of :971-1017 */
static void g2p_particle(OState *st, const OModel *md, float dt, int p) {
    v3 xp = vload(st->x + 3 * p);
    Stencil s = stencil(md, xp);
    int pc = 0;
    v3 ghost = V3(0.f, 0.f, 0.f);
    OCdf *cd = st->cdf;
    if (cd && cd->n_cdf > 0) {
        pc = cdf_particle_pc(st, md, &s, p, 1);
        cd->ptag[p] = pc;
        if (pc != 0) ghost = cdf_ghost(st, md, dt, p, pc, &s);
    }
    v3 nv = V3(0, 0, 0);
    m33 nC = mzero(), nF = mzero();
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                int ix = wrap_x(md, s.base[0] + i), iy = s.base[1] + j, iz = s.base[2] + k;
                v3 dpos = vsub(V3((float)i, (float)j, (float)k), s.fx);
                float weight = s.w[0][i] * s.w[1][j] * s.w[2][k];
                v3 gv = vload(st->grid_v_out + 3 * gidx(md, ix, iy, iz));
                if (pc != 0 && cdf_incompat(pc, cd->tag_prev[gidx(md, ix, iy, iz)])) gv = ghost;
                nv = vadd(nv, vscale(gv, weight));
                nC = madd(nC, mscale(mouter(gv, dpos), weight * md->inv_dx * 4.0f));
                v3 dwt = dweight(md, &s, i, j, k);
                nF = madd(nF, mouter(gv, dwt));
            }
    vstore(st->v + 3 * p, nv);
    v3 xn = vadd(xp, vscale(nv, dt));
    if (md->periodic_x) {
        float lx = (float)md->nx * md->dx;
        if (xn.c[0] < 0.0f) xn.c[0] = xn.c[0] + lx;
        if (xn.c[0] >= lx) xn.c[0] = xn.c[0] - lx;
    }
    vstore(st->x + 3 * p, xn);
    mstore(st->C + 9 * p, nC);
    mstore(st->L + 9 * p, nF);
    m33 Ftmp = mmul(madd(meye(), mscale(nF, dt)), mload(st->F + 9 * p));
    mstore(st->F_trial + 9 * p, Ftmp);
    if (md->update_cov_with_F) update_cov(st, p, nF, dt);
}

/* ------------------------------------------------------------------ particle kernels */
/* compute_mu_lam_from_E_nu, mpm_utils.py:583-588 */
EXPORT void o_compute_mu_lam(OState *st, OModel *md) {
    for (int p = 0; p < st->n; ++p) {
        int mat = st->material[p];
        md->mu[p] = md->E[mat] / (2.0f * (1.0f + md->nu[mat]));
        md->lam[p] = md->E[mat] * md->nu[mat] / ((1.0f + md->nu[mat]) * (1.0f - 2.0f * md->nu[mat]));
    }
}

/* compute_stress_from_F_trial, mpm_utils.py:1152-1158 */
EXPORT void o_compute_stress(OState *st, OModel *md, float dt) {
    O_PARALLEL_FOR
    for (int p = 0; p < st->n; ++p)
        if (st->selection[p] == 0) stress_update_particle(st, md, dt, p);
}

/* p2g_apic_with_stress, mpm_utils.py:919-923 */
EXPORT void o_p2g(OState *st, OModel *md, float dt) {
    O_PARALLEL_FOR
    for (int p = 0; p < st->n; ++p)
        if (st->selection[p] == 0) p2g_particle(st, md, dt, p);
}

/* g2p, mpm_utils.py:1046-1052 (stationary/rigid 7, 8 skipped) */
EXPORT void o_g2p(OState *st, OModel *md, float dt) {
    O_PARALLEL_FOR
    for (int p = 0; p < st->n; ++p) {
        if (st->selection[p] != 0) continue;
        int mat = st->material[p];
        if (mat == 7 || mat == 8) continue;
        g2p_particle(st, md, dt, p);
    }
}

/* g2p_stress_p2g, mpm_utils.py:1161-1178 */
EXPORT void o_g2p_stress_p2g(OState *st, OModel *md, float dt) {
    O_PARALLEL_FOR
    for (int p = 0; p < st->n; ++p) {
        if (st->selection[p] != 0) continue;
        int mat = st->material[p];
        if (mat != 7 && mat != 8) g2p_particle(st, md, dt, p);
        stress_update_particle(st, md, dt, p);
        p2g_particle(st, md, dt, p);
    }
}

/* ---------------------------------------------------------------------- grid kernels */
#define BOX_LOOP(lo, dims)                                  \
    for (int gi = lo[0]; gi < lo[0] + dims[0]; ++gi)        \
        for (int gj = lo[1]; gj < lo[1] + dims[1]; ++gj)    \
            for (int gk = lo[2]; gk < lo[2] + dims[2]; ++gk)

/* zero_grid / zero_grid_m_vin / zero_grid_vout, mpm_utils.py:590-598, 1181-1202.
 * which: bit 1 = grid_m, bit 2 = grid_v_in, bit 4 = grid_v_out */
EXPORT void o_zero_grid(OState *st, OModel *md, const int32_t *lo, const int32_t *dims, int which) {
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        if (which & 1) st->grid_m[g] = 0.0f;
        if (which & 2) st->grid_v_in[3 * g] = st->grid_v_in[3 * g + 1] = st->grid_v_in[3 * g + 2] = 0.0f;
        if (which & 4) st->grid_v_out[3 * g] = st->grid_v_out[3 * g + 1] = st->grid_v_out[3 * g + 2] = 0.0f;
    }
}

static inline void normalize_node(OState *st, const OModel *md, float dt, long g) {
    float m = st->grid_m[g];
    if (m > 1e-15f) {
        float inv = 1.0f / m;
        for (int c = 0; c < 3; ++c)
            st->grid_v_out[3 * g + c] = st->grid_v_in[3 * g + c] * inv + dt * md->gravity[c];
    }
}

/* grid_normalization_and_gravity, mpm_utils.py:927-941 */
EXPORT void o_grid_normalize(OState *st, OModel *md, float dt, const int32_t *lo, const int32_t *dims) {
    BOX_LOOP(lo, dims) normalize_node(st, md, dt, gidx(md, gi, gj, gk));
}

/* add_damping_via_grid, mpm_utils.py:1287-1295 */
EXPORT void o_grid_damping(OState *st, OModel *md, float scale, const int32_t *lo, const int32_t *dims) {
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        for (int c = 0; c < 3; ++c)
            st->grid_v_out[3 * g + c] = MUT(9) ? st->grid_v_out[3 * g + c] / scale : st->grid_v_out[3 * g + c] * scale;
    }
}

/* ------------------------------------------------------------------- collider kernels */
/* plane collide kernel, mpm_solver_warp.py:1929-1992 */
EXPORT void o_collide_plane(float time, float dt, OState *st, OModel *md, const OPlane *pm,
                            const int32_t *lo, const int32_t *dims) {
    (void)dt;
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        if (st->grid_m[g] <= 0.0f) continue;
        if (!(time >= pm->start_time && time < pm->end_time)) continue;
        v3 off = V3((float)gi * md->dx - pm->point[0], (float)gj * md->dx - pm->point[1],
                    (float)gk * md->dx - pm->point[2]);
        v3 n = V3(pm->normal[0], pm->normal[1], pm->normal[2]);
        float dp = vdot(off, n);
        if (MUT(2)) dp = -dp;
        if (!(dp < 0.0f)) continue;
        float *vo = st->grid_v_out + 3 * g;
        if (pm->surface_type == 0) {
            vo[0] = vo[1] = vo[2] = 0.0f;
        } else if (pm->surface_type == 11) {
            float z = (float)gk * md->dx;
            if (z < 0.4f || z > 0.53f) {
                vo[0] = vo[1] = vo[2] = 0.0f;
            } else {
                v3 vi = vload(vo);
                vstore(vo, vscale(V3(vi.c[0], 0.0f, vi.c[2]), 0.3f));
            }
        } else {
            v3 v = vload(vo);
            float nc = vdot(v, n);
            if (pm->surface_type == 1) v = vsub(v, vscale(n, nc));
            else v = vsub(v, vscale(n, fminf(nc, 0.0f)));
            if (nc < 0.0f && vlen(v) > 1e-20f)
                v = vscale(vnormalize(v), fmaxf(0.0f, vlen(v) + (MUT(1) ? -nc : nc) * pm->friction));
            vstore(vo, v);
        }
    }
}

/* cuboid collide kernel, mpm_solver_warp.py:2053-2088 */
EXPORT void o_collide_box(float time, float dt, OState *st, OModel *md, const OBox *pm,
                          const int32_t *lo, const int32_t *dims) {
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        if (st->grid_m[g] <= 0.0f) continue;
        float *vo = st->grid_v_out + 3 * g;
        if (time >= pm->start_time && time < pm->end_time) {
            v3 off = V3((float)gi * md->dx - pm->point[0], (float)gj * md->dx - pm->point[1],
                        (float)gk * md->dx - pm->point[2]);
            if (fabsf(off.c[0]) < pm->size[0] && fabsf(off.c[1]) < pm->size[1] &&
                fabsf(off.c[2]) < pm->size[2]) {
                float m = st->grid_m[g];
                v3 vf = vload(vo);
                v3 imp = vscale(vsub(vf, vload(pm->velocity)), m);
                pm->force[0] += imp.c[0];
                pm->force[1] += imp.c[1];
                pm->force[2] += imp.c[2];
                vstore(vo, vload(pm->velocity));
            }
        } else if (pm->reset == 1) {
            if (time < pm->end_time + 15.0f * dt) vo[0] = vo[1] = vo[2] = 0.0f;
        }
    }
}

/* capped_cylinder_sdf, mpm_solver_warp.py:29-40 */
static float capped_cylinder_sdf(float r, float z, float radius, float z0, float z1) {
    float zc = 0.5f * (z0 + z1), hh = 0.5f * (z1 - z0);
    float dr = r - radius, dz = fabsf(z - zc) - hh;
    float dro = fmaxf(dr, 0.0f), dzo = fmaxf(dz, 0.0f);
    float outside = sqrtf(dro * dro + dzo * dzo);
    float inside = fminf(fmaxf(dr, dz), 0.0f);
    return outside + inside;
}

/* revolved_glass_sdf, mpm_solver_warp.py:43-57 */
static float glass_sdf(v3 l, const OCup *pm) {
    float r = sqrtf(l.c[0] * l.c[0] + l.c[1] * l.c[1]);
    float d_outer = capped_cylinder_sdf(r, l.c[2], pm->outer_radius, -pm->half_height, pm->half_height);
    float d_cav = capped_cylinder_sdf(r, l.c[2], pm->inner_radius - pm->fillet_radius,
                                      pm->inner_floor_z + pm->fillet_radius,
                                      pm->half_height + pm->outer_radius) - pm->fillet_radius;
    return fmaxf(d_outer, -d_cav);
}

/* revolved-SDF cup collide kernel, mpm_solver_warp.py:2161-2225 */
EXPORT void o_collide_cup(float time, float dt, OState *st, OModel *md, const OCup *pm,
                          const int32_t *lo, const int32_t *dims) {
    (void)dt;
    m33 R = mload(pm->rot);
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        if (st->grid_m[g] <= 0.0f) continue;
        if (!(time >= pm->start_time && time < pm->end_time)) continue;
        v3 xn = V3((float)gi * md->dx, (float)gj * md->dx, (float)gk * md->dx);
        v3 rel = vsub(xn, vload(pm->point));
        v3 loc = mvmul(mT(R), rel);
        float sdf = glass_sdf(loc, pm);
        if (!(sdf < pm->contact_band)) continue;
        v3 vs = vadd(vload(pm->velocity), vcross(vload(pm->omega), rel));
        float *vo = st->grid_v_out + 3 * g;
        v3 vin = vload(vo);
        v3 vnew = vs;
        if (MUT(7) ? sdf <= -pm->sticky_depth : sdf > -pm->sticky_depth) {
            vnew = vin;
            float h = 0.25f * md->dx;
            v3 nl = V3(glass_sdf(vadd(loc, V3(h, 0, 0)), pm) - glass_sdf(vsub(loc, V3(h, 0, 0)), pm),
                       glass_sdf(vadd(loc, V3(0, h, 0)), pm) - glass_sdf(vsub(loc, V3(0, h, 0)), pm),
                       glass_sdf(vadd(loc, V3(0, 0, h)), pm) - glass_sdf(vsub(loc, V3(0, 0, h)), pm));
            float nlen = vlen(nl);
            if (nlen > 1.0e-12f) {
                v3 n = mvmul(R, vdiv(nl, nlen));
                v3 vr = vsub(vin, vs);
                float vn = vdot(vr, n);
                if (vn < 0.0f) {
                    v3 vt = vsub(vr, vscale(n, vn));
                    float vtl = vlen(vt);
                    if (vtl > 1.0e-20f)
                        vt = vscale(vdiv(vt, vtl), fmaxf(vtl + (MUT(6) ? -pm->friction : pm->friction) * vn, 0.0f));
                    vnew = vadd(vs, vt);
                }
            }
        }
        v3 imp = vscale(vsub(vin, vnew), st->grid_m[g]);
        v3 tq = MUT(11) ? vcross(imp, rel) : vcross(rel, imp);
        for (int c = 0; c < 3; ++c) { pm->force[c] += imp.c[c]; pm->torque[c] += tq.c[c]; }
        vstore(vo, vnew);
    }
}

/* bounding-box (domain walls) collide kernel, mpm_solver_warp.py:2294-2359 */
EXPORT void o_collide_walls(float time, float dt, OState *st, OModel *md, const OWalls *pm,
                            const int32_t *lo, const int32_t *dims) {
    (void)dt;
    const int pad = 3;
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        if (st->grid_m[g] <= 0.0f) continue;
        if (!(time >= pm->start_time && time < pm->end_time)) continue;
        float *vo = st->grid_v_out + 3 * g;
        if (gi < pad && (MUT(3) ? vo[0] > 0 : vo[0] < 0)) vo[0] = 0.0f;
        if (gi >= md->nx - pad && vo[0] > 0) vo[0] = 0.0f;
        if (gj < pad && vo[1] < 0) vo[1] = 0.0f;
        if (gj >= md->ny - pad && vo[1] > 0) vo[1] = 0.0f;
        if (gk < pad && vo[2] < 0) vo[2] = 0.0f;
        if (gk >= md->nz - pad && vo[2] > 0) vo[2] = 0.0f;
    }
}

/* sdf_trilerp / sdf_trilerp_vec, mpm_solver_warp.py:60-91 (nc = 1 or 3 channels) */
static void trilerp(const float *vals, int nc, int res, v3 f, float *out) {
    float hi = (float)res - 1.0001f;
    float fi = fminf(fmaxf(f.c[0], 0.0f), hi), fj = fminf(fmaxf(f.c[1], 0.0f), hi),
          fk = fminf(fmaxf(f.c[2], 0.0f), hi);
    int i = (int)floorf(fi), j = (int)floorf(fj), k = (int)floorf(fk);
    float tx = fi - (float)i, ty = fj - (float)j, tz = fk - (float)k;
#define VAL(a, b, c, ch) vals[(((long)(a) * res + (b)) * res + (c)) * nc + (ch)]
    for (int ch = 0; ch < nc; ++ch) {
        float c00 = VAL(i, j, k, ch) * (1.0f - tx) + VAL(i + 1, j, k, ch) * tx;
        float c01 = VAL(i, j, k + 1, ch) * (1.0f - tx) + VAL(i + 1, j, k + 1, ch) * tx;
        float c10 = VAL(i, j + 1, k, ch) * (1.0f - tx) + VAL(i + 1, j + 1, k, ch) * tx;
        float c11 = VAL(i, j + 1, k + 1, ch) * (1.0f - tx) + VAL(i + 1, j + 1, k + 1, ch) * tx;
        float c0 = c00 * (1.0f - ty) + c10 * ty;
        float c1 = c01 * (1.0f - ty) + c11 * ty;
        out[ch] = c0 * (1.0f - tz) + c1 * tz;
    }
#undef VAL
}

/* rotate a vector by a unit quaternion (x, y, z, w): v' = v(2w^2-1) + 2(q.v)q + 2w(q x v);
 * inverse = the same with -w in the cross term (Warp's quat_rotate / quat_rotate_inv). */
static v3 quat_rotate(const float *q, v3 v, int inverse) {
    v3 qv = V3(q[0], q[1], q[2]);
    float w = q[3];
    float c = 2.0f * w * w - 1.0f;
    float d = 2.0f * vdot(qv, v);
    v3 cr = vscale(vcross(qv, v), (inverse ? -w : w) * 2.0f);
    return vadd(vadd(vscale(v, c), vscale(qv, d)), cr);
}

/* ------------------------------------------------------------------ CPIC stamps */
/* cdf_zero_tags, mpm_solver_warp.py:94-102 */
EXPORT void o_cdf_zero_tags(OState *st, OModel *md, const int32_t *lo, const int32_t *dims) {
    OCdf *c = st->cdf;
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        c->tag[g] = 0;
        c->d[g] = 1.0e9f;
    }
}

/* cdf_copy_prev, mpm_solver_warp.py:105-115 */
EXPORT void o_cdf_copy_prev(OState *st, OModel *md, const int32_t *lo, const int32_t *dims) {
    OCdf *c = st->cdf;
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        c->tag_prev[g] = c->tag[g];
        c->d_prev[g] = c->d[g];
    }
}

/* cdf_block_mask, mpm_solver_warp.py:118-139 (blo / bdims in 4^3 blocks) */
EXPORT void o_cdf_block_mask(OState *st, OModel *md, const int32_t *blo, const int32_t *bdims) {
    OCdf *c = st->cdf;
    BOX_LOOP(blo, bdims) {
        int occ = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    int ci = (gi << 2) + i, cj = (gj << 2) + j, ck = (gk << 2) + k;
                    if (ci < md->nx && cj < md->ny && ck < md->nz) {
                        long g = gidx(md, ci, cj, ck);
                        occ = occ | c->tag[g] | c->tag_prev[g];
                    }
                }
        c->block[cdf_bidx(c, gi, gj, gk)] = occ;
    }
}

/* cdf_stamp, mpm_solver_warp.py:142-183: side / validity bits and owner distance of one
 * collider under its live pose; a node is tagged only when all eight trilerp corners are
 * valid (trilerp >= 0.999) and |d| <= band */
EXPORT void o_cdf_stamp(float time, OState *st, OModel *md, const OCdfCollider *pm, const int32_t *lo,
                        const int32_t *dims) {
    OCdf *c = st->cdf;
    BOX_LOOP(lo, dims) {
        if (time < pm->start_time || time >= pm->end_time) return;
        v3 xw = V3((float)gi * md->dx, (float)gj * md->dx, (float)gk * md->dx);
        v3 rel = vsub(xw, vload(pm->center));
        v3 body = quat_rotate(pm->quat, rel, 1);
        v3 fidx = V3((body.c[0] - pm->origin[0]) / pm->cell, (body.c[1] - pm->origin[1]) / pm->cell,
                     (body.c[2] - pm->origin[2]) / pm->cell);
        float rf = (float)pm->res - 1.0f;
        if (fidx.c[0] < 0.0f || fidx.c[1] < 0.0f || fidx.c[2] < 0.0f || fidx.c[0] > rf || fidx.c[1] > rf ||
            fidx.c[2] > rf)
            continue;
        float ok, d;
        trilerp(pm->cdf_valid, 1, pm->res, fidx, &ok);
        if (ok < 0.999f) continue;
        trilerp(pm->cdf_val, 1, pm->res, fidx, &d);
        float ud = fabsf(d);
        if (ud > pm->band) continue;
        int side = d >= 0.0f ? 1 : 0;
        long g = gidx(md, gi, gj, gk);
        int t = c->tag[g];
        t = t | (1 << (2 * pm->lane)) | (side << (2 * pm->lane + 1));
        if (ud < c->d[g]) {
            c->d[g] = ud;
            t = (t ^ (t & (3 << O_CDF_OWNER_SHIFT))) | (pm->lane << O_CDF_OWNER_SHIFT);
        }
        c->tag[g] = t;
    }
}

/* mesh-SDF collide kernel, mpm_solver_warp.py:2680-2736 */
EXPORT void o_collide_sdf(float time, float dt, OState *st, OModel *md, const OSdf *pm,
                          const int32_t *lo, const int32_t *dims) {
    (void)dt;
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        if (st->grid_m[g] <= 0.0f) continue;
        if (!(time >= pm->start_time && time < pm->end_time)) continue;
        v3 xw = V3((float)gi * md->dx, (float)gj * md->dx, (float)gk * md->dx);
        v3 rel = vsub(xw, vload(pm->center));
        v3 body = quat_rotate(pm->quat, rel, 1);
        v3 fidx = V3((body.c[0] - pm->origin[0]) / pm->cell, (body.c[1] - pm->origin[1]) / pm->cell,
                     (body.c[2] - pm->origin[2]) / pm->cell);
        float rf = (float)pm->res - 1.0f;
        if (!(fidx.c[0] >= 0.0f && fidx.c[1] >= 0.0f && fidx.c[2] >= 0.0f && fidx.c[0] <= rf &&
              fidx.c[1] <= rf && fidx.c[2] <= rf))
            continue;
        float sd;
        trilerp(pm->sdf_val, 1, pm->res, fidx, &sd);
        if (!(sd <= pm->band)) continue;
        v3 gr;
        trilerp(pm->sdf_grad, 3, pm->res, fidx, gr.c);
        if (!(vlen(gr) > 1.0e-12f)) continue;
        v3 nw = quat_rotate(pm->quat, vnormalize(gr), 0);
        if (MUT(4)) nw = vscale(nw, -1.0f);
        v3 vsurf = vadd(vload(pm->velocity), vcross(vload(pm->omega), rel));
        float *vo = st->grid_v_out + 3 * g;
        v3 vf = vload(vo);
        v3 vr = vsub(vf, vsurf);
        float vn = vdot(vr, nw);
        v3 vnew = vf;
        if (pm->surface_type == 0) vnew = vsurf;
        else if (pm->surface_type == 1) vnew = vadd(vsurf, MUT(12) ? vr : vsub(vr, vscale(nw, vn)));
        else if (vn < 0.0f) {
            v3 vt = vsub(vr, vscale(nw, vn));
            float tl = vlen(vt);
            if (tl > 1.0e-12f) vt = vscale(vt, fmaxf(0.0f, tl + (MUT(5) ? -pm->friction : pm->friction) * vn) / tl);
            vnew = vadd(vsurf, vt);
        }
        v3 imp = vscale(vsub(vf, vnew), st->grid_m[g]);
        v3 tq = vcross(rel, imp);
        for (int c = 0; c < 3; ++c) { pm->force[c] += imp.c[c]; pm->torque[c] += tq.c[c]; }
        vstore(vo, vnew);
    }
}

/* point-cloud collide kernel, mpm_solver_warp.py:2586-2605 */
EXPORT void o_collide_pointcloud(float time, float dt, OState *st, OModel *md, const OPointCloud *pm,
                                 const int32_t *lo, const int32_t *dims) {
    (void)dt;
    BOX_LOOP(lo, dims) {
        long g = gidx(md, gi, gj, gk);
        if (st->grid_m[g] <= 0.0f) continue;
        if (!(time >= pm->start_time && time < pm->end_time)) continue;
        if (pm->occupancy[g] == (MUT(8) ? 0 : 1)) {
            float *vo = st->grid_v_out + 3 * g;
            vo[0] = vo[1] = vo[2] = 0.0f;
        }
    }
}

/* ------------------------------------------------------------------- sparse blocks */
/* blocks_mark, mpm_utils.py:1506-1513 (int() truncates toward zero, then >> 2, clamp) */
EXPORT void o_blocks_mark(const float *x, int n, float inv_dx, const int32_t *bd, int32_t *mask) {
    for (int p = 0; p < n; ++p) {
        int b[3];
        for (int a = 0; a < 3; ++a) {
            int t = (int)(x[3 * p + a] * inv_dx - 0.5f);
            t >>= 2;
            b[a] = t < 0 ? 0 : (t > bd[a] - 1 ? bd[a] - 1 : t);
        }
        mask[((long)b[0] * bd[1] + b[1]) * bd[2] + b[2]] = 1;
    }
}

/* blocks_dilate, mpm_utils.py:1516-1529 */
EXPORT void o_blocks_dilate(const int32_t *src, const int32_t *bd, int32_t *dst) {
    for (int x = 0; x < bd[0]; ++x)
        for (int y = 0; y < bd[1]; ++y)
            for (int z = 0; z < bd[2]; ++z) {
                int on = 0;
                for (int i = -1; i <= 1; ++i)
                    for (int j = -1; j <= 1; ++j)
                        for (int k = -1; k <= 1; ++k) {
                            int xi = x + i, yj = y + j, zk = z + k;
                            xi = xi < 0 ? 0 : (xi > bd[0] - 1 ? bd[0] - 1 : xi);
                            yj = yj < 0 ? 0 : (yj > bd[1] - 1 ? bd[1] - 1 : yj);
                            zk = zk < 0 ? 0 : (zk > bd[2] - 1 ? bd[2] - 1 : zk);
                            if (src[((long)xi * bd[1] + yj) * bd[2] + zk] == 1) on = 1;
                        }
                dst[((long)x * bd[1] + y) * bd[2] + z] = on;
            }
}

/* blocks_compact, mpm_utils.py:1532-1544 (sequential order on CPU) */
EXPORT void o_blocks_compact(const int32_t *cur, const int32_t *prev, const int32_t *bd,
                             int32_t *cur_list, int32_t *union_list, int32_t *counts) {
    for (int x = 0; x < bd[0]; ++x)
        for (int y = 0; y < bd[1]; ++y)
            for (int z = 0; z < bd[2]; ++z) {
                int code = (x * bd[1] + y) * bd[2] + z;
                int c = cur[code];
                if (c == 1) cur_list[counts[0]++] = code;
                if (c == 1 || prev[code] == 1) union_list[counts[1]++] = code;
            }
}

/* _block_node, mpm_utils.py:1547-1553 */
static inline void block_node(int code, int cell, const int32_t *bd, int *n) {
    int bz = code % bd[2];
    int r = code / bd[2];
    int by = r % bd[1];
    int bx = r / bd[1];
    n[0] = bx * 4 + (cell >> 4);
    n[1] = by * 4 + ((cell >> 2) & 3);
    n[2] = bz * 4 + (cell & 3);
}

/* zero_grid_blocks / grid_normalization_and_gravity_blocks / add_damping_via_grid_blocks,
 * mpm_utils.py:1556-1588. op: 0 zero, 1 normalize, 2 damping */
EXPORT void o_blocks_grid_op(OState *st, OModel *md, int op, float dt, float scale,
                             const int32_t *blist, int nblocks, const int32_t *bd) {
    for (int t = 0; t < nblocks * 64; ++t) {
        int n[3];
        block_node(blist[t >> 6], t & 63, bd, n);
        if (!(n[0] < md->nx && n[1] < md->ny && n[2] < md->nz)) continue;
        long g = gidx(md, n[0], n[1], n[2]);
        if (op == 0) {
            st->grid_m[g] = 0.0f;
            for (int c = 0; c < 3; ++c) st->grid_v_in[3 * g + c] = st->grid_v_out[3 * g + c] = 0.0f;
        } else if (op == 1) {
            normalize_node(st, md, dt, g);
        } else {
            for (int c = 0; c < 3; ++c) st->grid_v_out[3 * g + c] = st->grid_v_out[3 * g + c] * scale;
        }
    }
}

/* ------------------------------------------------------------------- rigid bodies */
/* rigid_g2p_accumulate, mpm_utils.py:1369-1418: per rigid particle (material 8, selected)
 * the B-spline-weighted grid_v_out, summed as m v and (x - x_cm) x m v into its body;
 * sequential in particle order (the checker's sum order) */
EXPORT void o_rigid_gather(OState *st, OModel *md, const float *x_cm, int nb, float *lin, float *ang) {
    for (int b = 0; b < 3 * nb; ++b) lin[b] = ang[b] = 0.0f;
    for (int p = 0; p < st->n; ++p) {
        if (st->selection[p] != 0 || st->material[p] != 8) continue;
        int bid = st->rigid_id[p];
        if (bid < 0 || bid >= nb) continue;
        v3 xp = vload(st->x + 3 * p);
        Stencil s = stencil(md, xp);
        v3 vi = V3(0, 0, 0);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) {
                    float w = s.w[0][i] * s.w[1][j] * s.w[2][k];
                    vi = vadd(vi, vscale(vload(st->grid_v_out + 3 * gidx(md, wrap_x(md, s.base[0] + i),
                                                                          s.base[1] + j, s.base[2] + k)), w));
                }
        v3 mv = vscale(vi, st->mass[p]);
        v3 r = vsub(xp, vload(x_cm + 3 * bid));
        v3 am = vcross(r, mv);
        for (int c = 0; c < 3; ++c) {
            lin[3 * bid + c] += mv.c[c];
            ang[3 * bid + c] += am.c[c];
        }
    }
}

/* rigid_body_integrate, mpm_utils.py:1421-1466: v = P / M, omega = R I^-1 R^T L,
 * x += v dt, R <- polar factor of R + skew(omega) R dt */
EXPORT void o_rigid_integrate(float *x_cm, float *v_cm, float *omega, float *R, const float *mass,
                              const float *inv_I, const float *lin, const float *ang, int nb, float dt) {
    for (int b = 0; b < nb; ++b) {
        m33 Rb = mload(R + 9 * b);
        float M = mass[b];
        v3 v = vdiv(vload(lin + 3 * b), M);
        m33 Iw = mmul(mmul(Rb, mload(inv_I + 9 * b)), mT(Rb));
        v3 om = mvmul(Iw, vload(ang + 3 * b));
        v3 x = vadd(vload(x_cm + 3 * b), vscale(v, dt));
        m33 sk = mzero();
        sk.a[1] = -om.c[2]; sk.a[2] = om.c[1];
        sk.a[3] = om.c[2];  sk.a[5] = -om.c[0];
        sk.a[6] = -om.c[1]; sk.a[7] = om.c[0];
        m33 Rn = madd(Rb, mscale(mmul(sk, Rb), dt));
        m33 U, V; v3 sg;
        svd3(Rn, &U, &sg, &V);
        vstore(v_cm + 3 * b, v);
        vstore(omega + 3 * b, om);
        vstore(x_cm + 3 * b, x);
        mstore(R + 9 * b, mmul(U, mT(V)));
    }
}

/* rigid_particle_update, mpm_utils.py:1469-1493: x = x_cm + R x_ref, v = v_cm + omega x r,
 * C = skew(omega), F = I */
EXPORT void o_rigid_update(OState *st, const float *x_cm, const float *v_cm, const float *omega,
                           const float *R, int nb) {
    for (int p = 0; p < st->n; ++p) {
        if (st->selection[p] != 0 || st->material[p] != 8) continue;
        int bid = st->rigid_id[p];
        if (bid < 0 || bid >= nb) continue;
        v3 xc = vload(x_cm + 3 * bid), om = vload(omega + 3 * bid);
        v3 xp = vadd(xc, mvmul(mload(R + 9 * bid), vload(st->x_ref + 3 * p)));
        v3 r = vsub(xp, xc);
        vstore(st->v + 3 * p, vadd(vload(v_cm + 3 * bid), vcross(om, r)));
        vstore(st->x + 3 * p, xp);
        m33 sk = mzero();
        sk.a[1] = -om.c[2]; sk.a[2] = om.c[1];
        sk.a[3] = om.c[2];  sk.a[5] = -om.c[0];
        sk.a[6] = -om.c[1]; sk.a[7] = om.c[0];
        mstore(st->C + 9 * p, sk);
        mstore(st->F + 9 * p, meye());
    }
}

/* ------------------------------------------------------------------------- exports */
/* compute_R_from_F, mpm_utils.py:1261-1285 (stores R^T, as the reference does) */
EXPORT void o_compute_R(OState *st) {
    for (int p = 0; p < st->n; ++p) {
        m33 U, V; v3 s;
        svd3(mload(st->F_trial + 9 * p), &U, &s, &V);
        if (mdet(U) < 0.0f) { U.a[2] = -U.a[2]; U.a[5] = -U.a[5]; U.a[8] = -U.a[8]; }
        if (mdet(V) < 0.0f) { V.a[2] = -V.a[2]; V.a[5] = -V.a[5]; V.a[8] = -V.a[8]; }
        mstore(st->R + 9 * p, mT(mmul(U, mT(V))));
    }
}

/* compute_cov_from_F, mpm_utils.py:1234-1258 */
EXPORT void o_compute_cov(OState *st) {
    for (int p = 0; p < st->n; ++p) {
        m33 F = mload(st->F_trial + 9 * p);
        const float *c = st->init_cov + 6 * p;
        m33 ic = mzero();
        ic.a[0] = c[0]; ic.a[1] = c[1]; ic.a[2] = c[2];
        ic.a[3] = c[1]; ic.a[4] = c[3]; ic.a[5] = c[4];
        ic.a[6] = c[2]; ic.a[7] = c[4]; ic.a[8] = c[5];
        m33 cv = mmul(mmul(F, ic), mT(F));
        float *o = st->cov + 6 * p;
        o[0] = cv.a[0]; o[1] = cv.a[1]; o[2] = cv.a[2]; o[3] = cv.a[4]; o[4] = cv.a[5]; o[5] = cv.a[8];
    }
}

/* single-particle constitutive probe for unit tests: F_trial -> (F, stress) of particle p */
EXPORT void o_stress_particle(OState *st, OModel *md, float dt, int p) { stress_update_particle(st, md, dt, p); }
