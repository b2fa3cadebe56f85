// Particle modifiers on the GPU: pre-P2G impulses and particle velocity modifiers
// (add_impulse_on_particles, enforce_particle_velocity_translation / _rotation,
// mpm_solver_warp.py:2385-2539, with the selection kernels of mpm_utils.py:1320-1363).
// A modifier selects its particles once, into one bit of a per-particle mask word that
// travels with the particle through the block sort; one kernel per substep then applies
// every active modifier to every selected particle, impulses before velocity overrides.
#include <hip/hip_runtime.h>

#include "mpm_device.h"

namespace {

constexpr int kThreads = 256;

MPM_DEV v3 mvec(const float* a) { return mk3(a[0], a[1], a[2]); }

// selection_add_impulse_on_particles / selection_enforce_particle_velocity_translation
// (box) and selection_enforce_particle_velocity_cylinder (cylinder about vec)
MPM_DEV bool selected(const mpm_particle_modifier_t& m, v3 x) {
    const v3 off = x - mvec(m.point);
    if (m.kind == 2) {
        const v3 n = mvec(m.vec);
        const float ax = dot(off, n);
        const float hd = len(off - n * ax);
        return fabsf(ax) < m.half_height && hd < m.radius;
    }
    return fabsf(off.x) < m.size[0] && fabsf(off.y) < m.size[1] && fabsf(off.z) < m.size[2];
}

__global__ __launch_bounds__(kThreads) void k_modifier_select(mpm_particles_t P, mpm_particle_modifier_t m) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= P.n) return;
    const unsigned bit = 1u << m.bit;
    const unsigned old = (unsigned)P.modifier_mask[p];
    P.modifier_mask[p] = (int32_t)(selected(m, ld3(P.x, P.stride, p)) ? (old | bit) : (old & ~bit));
}

// the rotation field of enforce_particle_velocity_rotation's modify kernel
MPM_DEV v3 rotation_velocity(const mpm_particle_modifier_t& m, v3 x) {
    const v3 n = mvec(m.vec), a1 = mvec(m.axis1), a2 = mvec(m.axis2);
    const v3 off = x - mvec(m.point);
    const float hd = len(off - n * dot(off, n));
    const float c = dot(off, a1) / hd;
    float th = acosf(fminf(fmaxf(c, -1.0f), 1.0f));
    if (!(dot(off, a2) > 0.0f)) th = -th;
    const float s1 = -hd * sinf(th) * m.rotation_scale;
    const float s2 = hd * cosf(th) * m.rotation_scale;
    return a1 * s1 + a2 * s2 + n * m.translation_scale;
}

__global__ __launch_bounds__(kThreads) void k_modifiers_apply(mpm_particles_t P,
                                                              const mpm_particle_modifier_t* __restrict__ mods,
                                                              int n_mods, const int32_t* __restrict__ active,
                                                              float dt) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (p >= P.n) return;
    const unsigned mask = (unsigned)P.modifier_mask[p];
    if (mask == 0u) return;
    const unsigned act = (unsigned)*active;
    const int S = P.stride;
    v3 v = ld3(P.v, S, p);
    bool changed = false;
    for (int pass = 0; pass < 2; ++pass) {  // impulses, then velocity modifiers
        for (int k = 0; k < n_mods; ++k) {
            const mpm_particle_modifier_t& m = mods[k];
            if ((m.kind == 0) != (pass == 0) || !((mask >> m.bit) & 1u)) continue;
            if (!((act >> k) & 1u)) continue;
            if (m.kind == 0) {
                const float mp = P.mass[p];
                v = v + mk3(m.vec[0] / mp, m.vec[1] / mp, m.vec[2] / mp) * dt;
            } else if (m.kind == 1) {
                v = mvec(m.vec);
            } else {
                v = rotation_velocity(m, ld3(P.x, S, p));
            }
            changed = true;
        }
    }
    if (changed) st3(P.v, S, p, v);
}

}  // namespace

extern "C" {

int32_t mpm_modifier_select(const mpm_particles_t* P, const mpm_particle_modifier_t* mod, void* stream) {
    if (!P || !mod || P->n < 0 || P->stride < P->n || !P->x || !P->modifier_mask || mod->bit < 0 ||
        mod->bit >= MPM_MAX_MODIFIERS || mod->kind < 0 || mod->kind > 2)
        return MPM_ERR_ARG;
    if (P->n == 0) return MPM_OK;
    hipLaunchKernelGGL(k_modifier_select, dim3((P->n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       (hipStream_t)stream, *P, *mod);
    return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

int32_t mpm_modifiers_apply(const mpm_particles_t* P, const mpm_particle_modifier_t* mods, int32_t n_mods,
                            const int32_t* active, float dt, void* stream) {
    if (!P || n_mods < 0 || n_mods > MPM_MAX_MODIFIERS || (n_mods > 0 && !mods) || !active || P->n < 0 ||
        P->stride < P->n || !P->x || !P->v || !P->mass || !P->modifier_mask)
        return MPM_ERR_ARG;
    if (P->n == 0 || n_mods == 0) return MPM_OK;
    hipLaunchKernelGGL(k_modifiers_apply, dim3((P->n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       (hipStream_t)stream, *P, mods, n_mods, active, dt);
    return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // extern "C"
