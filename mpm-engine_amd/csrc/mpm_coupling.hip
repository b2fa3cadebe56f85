// Device-side coupling readout: the stress-integral reaction on a box end-effector
// (box_contact_wrench, src/warpmpm/coupling/wrench.py:15-46) evaluated on the particle
// state in HBM, so a force-controlled press loop reads 8 floats per tick instead of
// copying positions, F and stress of every particle to the host.
//
// Per particle: J = det F; volume = vol0 * |J| and Cauchy = tau / max(|J|, 1e-9) (the
// reference Solver's default inversion policy, solver.py:559-598). Contact layer of
// thickness T (given, or 3 * mean(volume)^(1/3)) under the box bottom; traction
// -(sigma . e_z) * volume / T summed into force and torque about the box centre.
// Reductions are per-workgroup sums in a fixed order, then one workgroup over the
// partials in workgroup order: the readout is bit-reproducible.
#include <hip/hip_runtime.h>

#include "mpm_device.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxGroups = 1024;  // workgroups of the strided sweeps (partials capacity)

MPM_DEV float block_sum(float v, float* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const float r = sh[0];
    __syncthreads();
    return r;
}

MPM_DEV float jacobian(const mpm_particles_t& P, int p, int& inverted) {
    const m33 F = ld9(P.F, P.stride, p);
    const float J = mdet(F);
    inverted = !(J > 0.0f) || !isfinite(J);
    return fabsf(J);
}

// partial[g] = sum of current volumes (and inverted count) over workgroup g's stride
__global__ __launch_bounds__(kThreads) void k_vol_partials(mpm_particles_t P, float* __restrict__ partial) {
    __shared__ float sh[kThreads];
    float v = 0.0f, bad = 0.0f;
    for (int p = blockIdx.x * kThreads + threadIdx.x; p < P.n; p += gridDim.x * kThreads) {
        int inv;
        const float J = jacobian(P, p, inv);
        v += P.vol[p] * J;
        bad += (float)inv;
    }
    const float a = block_sum(v, sh);
    const float b = block_sum(bad, sh);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

// out[6] = layer thickness T, out[7] = inverted count (partials in workgroup order)
__global__ __launch_bounds__(kThreads) void k_vol_finish(const float* __restrict__ partial, int ngroups, int n,
                                                         float layer, float* __restrict__ out) {
    __shared__ float sh[kThreads];
    float a = 0.0f, b = 0.0f;
    for (int g = threadIdx.x; g < ngroups; g += kThreads) {
        a += partial[2 * g];
        b += partial[2 * g + 1];
    }
    const float s = block_sum(a, sh);
    const float c = block_sum(b, sh);
    if (threadIdx.x == 0) {
        out[6] = layer > 0.0f ? layer : 3.0f * cbrtf(s / (float)max(n, 1));
        out[7] = c;
    }
}

// 7 partial sums per workgroup: force xyz, torque xyz, contact count
__global__ __launch_bounds__(kThreads) void k_contact_partials(mpm_particles_t P, float cx, float cy, float cz,
                                                               float hx, float hy, float hz, int compressive,
                                                               const float* __restrict__ out,
                                                               float* __restrict__ partial) {
    __shared__ float sh[kThreads];
    const float T = out[6];
    const float bottom = cz - hz;
    float acc[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int p = blockIdx.x * kThreads + threadIdx.x; p < P.n; p += gridDim.x * kThreads) {
        const v3 x = ld3(P.x, P.stride, p);
        if (!(fabsf(x.x - cx) < hx && fabsf(x.y - cy) < hy && x.z > bottom - T && x.z < bottom + 0.5f * T)) continue;
        int inv;
        const float J = jacobian(P, p, inv);
        const float Jc = fmaxf(J, 1e-9f);
        const int S = P.stride;
        const float s02 = P.stress[2 * S + p] / Jc, s12 = P.stress[5 * S + p] / Jc, s22 = P.stress[8 * S + p] / Jc;
        if (compressive && !(s22 < 0.0f)) continue;
        const float w = P.vol[p] * J / T;
        const v3 f = mk3(-s02 * w, -s12 * w, -s22 * w);
        const v3 t = cross(x - mk3(cx, cy, cz), f);
        acc[0] += f.x; acc[1] += f.y; acc[2] += f.z;
        acc[3] += t.x; acc[4] += t.y; acc[5] += t.z;
        acc[6] += 1.0f;
    }
    for (int c = 0; c < 7; ++c) {
        const float r = block_sum(acc[c], sh);
        if (threadIdx.x == 0) partial[8 * blockIdx.x + c] = r;
    }
}

__global__ __launch_bounds__(kThreads) void k_contact_finish(const float* __restrict__ partial, int ngroups,
                                                             float* __restrict__ out) {
    __shared__ float sh[kThreads];
    for (int c = 0; c < 7; ++c) {
        float a = 0.0f;
        for (int g = threadIdx.x; g < ngroups; g += kThreads) a += partial[8 * g + c];
        const float r = block_sum(a, sh);
        if (threadIdx.x == 0) out[c] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0 && out[6] < 3.0f)  // fewer than 3 contacts: no reading
        for (int c = 0; c < 6; ++c) out[c] = 0.0f;
}

}  // namespace

extern "C" {

int32_t mpm_box_contact_workspace_floats(void) { return 8 * kMaxGroups; }

int32_t mpm_box_contact_wrench(const mpm_particles_t* P, const float* box, float layer, int32_t compressive_only,
                               float* out, float* workspace, void* stream) {
    if (!P || !box || !out || !workspace || P->n < 0 || P->stride < P->n || !P->x || !P->F || !P->stress ||
        !P->vol || !(box[3] > 0.0f) || !(box[4] > 0.0f) || !(box[5] >= 0.0f))
        return MPM_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int groups = P->n == 0 ? 1 : (int)min((long)kMaxGroups, ((long)P->n + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_vol_partials, dim3(groups), dim3(kThreads), 0, s, *P, workspace);
    hipLaunchKernelGGL(k_vol_finish, dim3(1), dim3(kThreads), 0, s, workspace, groups, P->n, layer, out);
    hipLaunchKernelGGL(k_contact_partials, dim3(groups), dim3(kThreads), 0, s, *P, box[0], box[1], box[2], box[3],
                       box[4], box[5], (int)(compressive_only != 0), out, workspace);
    hipLaunchKernelGGL(k_contact_finish, dim3(1), dim3(kThreads), 0, s, workspace, groups, out);
    return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // extern "C"
