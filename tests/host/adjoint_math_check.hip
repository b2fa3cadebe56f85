// Host-side driver for the constitutive adjoints of mpm_adjoint_math.h (compiled by
// tests/test_adjoint_math.py with MPM_DEV = __host__ __device__; it makes no HIP call and
// runs without a GPU). One case per input line:
//   mat dt mu lam ys eta bulk K n mu_s dmu I0 d rho_s Ft[9] L[9] Fb[9] tb[9]
// (for material 2 the ys slot carries the Drucker-Prager alpha; for material 1 the eta and
// bulk slots carry the hardening xi and the adjoint of the yield stress after the map)
// One output line per case:
//   Ftb[9] Lb[9] mu_b lam_b ys_b eta_b bulk_b K_b n_b mu_s_b dmu_b I0_b d_b
#define MPM_DEV __host__ __device__ inline
#include "mpm_adjoint_math.h"

#include <cstdio>
#include <cstring>

int main() {
    static mpm_model_t md;
    float in[14 + 36];
    for (;;) {
        if (scanf("%f", &in[0]) != 1) return 0;  // end of input
        for (int i = 1; i < 50; ++i)
            if (scanf("%f", &in[i]) != 1) return 1;  // truncated case
        memset(&md, 0, sizeof(md));
        const int mat = (int)in[0];
        const float dt = in[1], mu = in[2], lam = in[3], ys = in[4];
        md.plastic_viscosity[mat] = in[5];
        md.bulk[mat] = in[6];
        md.hardening[mat] = in[7];
        md.softening[mat] = in[8];
        md.muI_mu_s[mat] = in[9];
        md.muI_delta_mu[mat] = in[10];
        md.muI_I0[mat] = in[11];
        md.muI_d[mat] = in[12];
        md.muI_rho_s[mat] = in[13];
        m33 Ft, L, Fb, tb;
        memcpy(Ft.a, in + 14, 36);
        memcpy(L.a, in + 23, 36);
        memcpy(Fb.a, in + 32, 36);
        memcpy(tb.a, in + 41, 36);
        float mub = 0.f, lamb = 0.f, yb = 0.f, eb = 0.f, bb = 0.f, kb = 0.f, nb = 0.f;
        float pb[4] = {0.f, 0.f, 0.f, 0.f};
        m33 Lb = mzero(), Ftb = mzero();
        if (mat == 0) {
            Ftb = adj_fcr(Ft, mu, lam, Fb, tb, mub, lamb);
        } else if (mat == 2) {
            Ftb = adj_sand(Ft, mu, lam, ys, Fb, tb, mub, lamb);
        } else if (mat == 1) {  // eta / bulk slots: hardening xi, adjoint of the hardened ys
            Ftb = adj_von_mises(Ft, mu, lam, ys, in[5], in[6], Fb, tb, mub, lamb, yb);
        } else if (mat == 9) {
            Ftb = adj_mu_i(Ft, &md, mat, dt, mu, lam, Fb, tb, mub, lamb, pb);
        } else if (mat == 10) {
            Ftb = adj_viscous(Ft, L, md.bulk[mat], md.plastic_viscosity[mat], ys, md.hardening[mat],
                              md.softening[mat], Fb, tb, Lb, eb, bb, yb, kb, nb);
        } else {
            return 2;
        }
        for (int i = 0; i < 9; ++i) printf("%.9e ", Ftb.a[i]);
        for (int i = 0; i < 9; ++i) printf("%.9e ", Lb.a[i]);
        printf("%.9e %.9e %.9e %.9e %.9e %.9e %.9e %.9e %.9e %.9e %.9e\n", mub, lamb, yb, eb, bb, kb, nb, pb[0],
               pb[1], pb[2], pb[3]);
    }
}
