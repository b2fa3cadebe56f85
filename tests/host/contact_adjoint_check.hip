// Host-side driver for the surface-collider contact maps of mpm_adjoint_math.h (compiled
// by tests/test_contact_adjoint.py with MPM_DEV = __host__ __device__; no HIP call, runs
// without a GPU). One case per input line:
//   mode cup vs[3] n[3] friction vin[3] g[3]
// One output line per case: the velocity after the contact, then J^T g (J its derivative
// with respect to the velocity before).
#define MPM_DEV __host__ __device__ inline
#include "mpm_adjoint_math.h"

#include <cstdio>

int main() {
    float in[15];
    for (;;) {
        if (scanf("%f", &in[0]) != 1) return 0;  // end of input
        for (int i = 1; i < 15; ++i)
            if (scanf("%f", &in[i]) != 1) return 1;  // truncated case
        SurfaceHit h;
        h.mode = (int)in[0];
        h.cup = in[1] != 0.0f;
        h.vs = mk3(in[2], in[3], in[4]);
        h.n = mk3(in[5], in[6], in[7]);
        h.rel = mk3(0.f, 0.f, 0.f);
        const float fr = in[8];
        const v3 vin = mk3(in[9], in[10], in[11]), g = mk3(in[12], in[13], in[14]);
        const v3 out = surface_apply(h, fr, vin), jt = surface_apply_T(h, fr, vin, g);
        printf("%.9e %.9e %.9e %.9e %.9e %.9e\n", out.x, out.y, out.z, jt.x, jt.y, jt.z);
    }
}
