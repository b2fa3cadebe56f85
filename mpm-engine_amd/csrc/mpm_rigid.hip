// Rigid bodies (material 8) on the GPU: the per-substep rigid step of the split pipeline
// (_p2g2p_tail, mpm_solver_warp.py:1338-1367): momentum gather, body integration, the
// restitution planes and the particle update. The reference runs the restitution on the
// host in float64 numpy (_apply_rigid_restitution, mpm_solver_warp.py:874-962) with a
// device -> host -> device round trip per substep; here it is a device step as well, so a
// rigid tick has no host synchronisation and replays as a hipGraph.
//
// Contact point: per (body, plane) the deepest particle is found in two passes, an
// atomicMin over the order-preserving bit pattern of its float64 signed distance, then an
// atomicMin over the indices of the particles at that distance: the lowest index wins a
// tie, as numpy's argmin over the body's particles in array order does. Pushes along
// earlier planes translate the whole body, which moves every distance to a later plane by
// the same amount and keeps its argmin; the push is added to the minimum afterwards.
#include <hip/hip_runtime.h>

#include "mpm_device.h"

namespace {

constexpr int kThreads = 256;
constexpr int kAccum = 6;  // linear momentum xyz, angular momentum xyz

struct RigidWs {
    float* accum;             // [n_bodies][6]
    unsigned long long* dmin;  // [n_bodies * n_planes] ordered float64 bits of the minimum
    unsigned int* imin;        // [n_bodies * n_planes] lowest particle index at the minimum
};

MPM_DEV RigidWs rigid_ws(void* ws, int nb, int np) {
    RigidWs w;
    w.dmin = reinterpret_cast<unsigned long long*>(ws);
    w.imin = reinterpret_cast<unsigned int*>(w.dmin + (size_t)nb * np);
    w.accum = reinterpret_cast<float*>(w.imin + (size_t)nb * np + ((nb * np) & 1));
    return w;
}

// float64 -> unsigned with the same order (negative values: all bits flipped)
MPM_DEV unsigned long long ordered_bits(double d) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

MPM_DEV bool rigid_particle(const mpm_particles_t& P, int p, int nb, int& bid) {
    if (p >= P.n || P.selection[p] != 0 || P.material[p] != 8) return false;
    bid = P.rigid_id[p];
    return (unsigned)bid < (unsigned)nb;
}

MPM_DEV m33 body_R(const mpm_rigid_body_t& b) {
    m33 R;
    for (int i = 0; i < 9; ++i) R.a[i] = b.R[i];
    return R;
}

// x_cm + R x_ref in float32, the order numpy's float32 product takes
MPM_DEV v3 world_pos(const mpm_rigid_body_t& b, v3 xr) {
    const m33 R = body_R(b);
    const v3 rx = mvmul(R, xr);
    return mk3(b.x_cm[0] + rx.x, b.x_cm[1] + rx.y, b.x_cm[2] + rx.z);
}

MPM_DEV bool plane_active(unsigned active, int k) { return (active >> k) & 1u; }

MPM_DEV double plane_dist(const mpm_rigid_plane_t& pl, v3 xw) {
    return ((double)xw.x - (double)pl.point[0]) * (double)pl.normal[0] +
           ((double)xw.y - (double)pl.point[1]) * (double)pl.normal[1] +
           ((double)xw.z - (double)pl.point[2]) * (double)pl.normal[2];
}

__global__ void k_rigid_clear(void* ws, int nb, int np) {
    const RigidWs w = rigid_ws(ws, nb, np);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb * np; i += gridDim.x * blockDim.x) {
        w.dmin[i] = ~0ull;
        w.imin[i] = 0xffffffffu;
    }
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb * kAccum; i += gridDim.x * blockDim.x) w.accum[i] = 0.0f;
}

// rigid_g2p_accumulate, mpm_utils.py:1369-1418: grid velocity gathered with the quadratic
// B-spline weights, summed into m v and (x - x_cm) x m v per body
__global__ __launch_bounds__(kThreads) void k_rigid_gather(mpm_particles_t P, mpm_grid_t G,
                                                           const mpm_rigid_body_t* __restrict__ bodies, int nb,
                                                           int np, void* ws) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    int bid;
    if (!rigid_particle(P, p, nb, bid)) return;
    const RigidWs w = rigid_ws(ws, nb, np);
    const int S = P.stride;
    const v3 x = ld3(P.x, S, p);
    const Stencil st = make_stencil(x, G.inv_dx);
    const float4* __restrict__ uo = reinterpret_cast<const float4*>(G.node_out);
    v3 vi = mk3(0.f, 0.f, 0.f);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                const int gi = st.base[0] + i, gj = st.base[1] + j, gk = st.base[2] + k;
                if ((unsigned)gi >= (unsigned)G.n[0] || (unsigned)gj >= (unsigned)G.n[1] ||
                    (unsigned)gk >= (unsigned)G.n[2])
                    continue;
                const float wt = st.w[0][i] * st.w[1][j] * st.w[2][k];
                const float4 u = uo[node_index(G, gi, gj, gk)];
                vi = vi + mk3(u.x, u.y, u.z) * wt;
            }
    const float m = P.mass[p];
    const mpm_rigid_body_t& b = bodies[bid];
    const v3 r = mk3(x.x - b.x_cm[0], x.y - b.x_cm[1], x.z - b.x_cm[2]);
    const v3 mv = vi * m;
    const v3 am = cross(r, mv);
    float* a = w.accum + (size_t)bid * kAccum;
    atomicAdd(a + 0, mv.x);
    atomicAdd(a + 1, mv.y);
    atomicAdd(a + 2, mv.z);
    atomicAdd(a + 3, am.x);
    atomicAdd(a + 4, am.y);
    atomicAdd(a + 5, am.z);
}

// rigid_body_integrate, mpm_utils.py:1421-1466, one thread per body
__global__ void k_rigid_integrate(mpm_rigid_body_t* bodies, int nb, int np, void* ws, float dt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const RigidWs w = rigid_ws(ws, nb, np);
    mpm_rigid_body_t body = bodies[b];
    const float* a = w.accum + (size_t)b * kAccum;
    const m33 R = body_R(body);
    m33 Ii;
    for (int i = 0; i < 9; ++i) Ii.a[i] = body.inv_inertia[i];
    const float M = body.mass;
    const v3 v = mk3(a[0] / M, a[1] / M, a[2] / M);
    const m33 Iw = mmul(mmul(R, Ii), mT(R));
    const v3 om = mvmul(Iw, mk3(a[3], a[4], a[5]));
    m33 sk = mzero();
    sk.a[1] = -om.z; sk.a[2] = om.y;
    sk.a[3] = om.z;  sk.a[5] = -om.x;
    sk.a[6] = -om.y; sk.a[7] = om.x;
    const m33 Rn = madd(R, mscale(mmul(sk, R), dt));
    m33 U, V;
    v3 sg;
    svd3(Rn, U, sg, V);
    const m33 Ro = mmul(U, mT(V));
    body.v_cm[0] = v.x; body.v_cm[1] = v.y; body.v_cm[2] = v.z;
    body.omega[0] = om.x; body.omega[1] = om.y; body.omega[2] = om.z;
    body.x_cm[0] += v.x * dt; body.x_cm[1] += v.y * dt; body.x_cm[2] += v.z * dt;
    for (int i = 0; i < 9; ++i) body.R[i] = Ro.a[i];
    bodies[b] = body;
}

// pass 1 of the deepest particle: minimum signed distance per (body, plane)
__global__ __launch_bounds__(kThreads) void k_rigid_contact_min(mpm_particles_t P,
                                                                const mpm_rigid_body_t* __restrict__ bodies, int nb,
                                                                const mpm_rigid_plane_t* __restrict__ planes, int np,
                                                                const int32_t* __restrict__ active, void* ws) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    int bid;
    if (!rigid_particle(P, p, nb, bid)) return;
    const RigidWs w = rigid_ws(ws, nb, np);
    const unsigned act = (unsigned)*active;
    const v3 xw = world_pos(bodies[bid], ld3(P.x_ref, P.stride, p));
    for (int k = 0; k < np; ++k) {
        if (!plane_active(act, k)) continue;
        atomicMin(w.dmin + (size_t)bid * np + k, ordered_bits(plane_dist(planes[k], xw)));
    }
}

// pass 2: the lowest particle index at that distance
__global__ __launch_bounds__(kThreads) void k_rigid_contact_index(mpm_particles_t P,
                                                                  const mpm_rigid_body_t* __restrict__ bodies, int nb,
                                                                  const mpm_rigid_plane_t* __restrict__ planes, int np,
                                                                  const int32_t* __restrict__ active, void* ws) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    int bid;
    if (!rigid_particle(P, p, nb, bid)) return;
    const RigidWs w = rigid_ws(ws, nb, np);
    const unsigned act = (unsigned)*active;
    const v3 xw = world_pos(bodies[bid], ld3(P.x_ref, P.stride, p));
    for (int k = 0; k < np; ++k) {
        if (!plane_active(act, k)) continue;
        const size_t o = (size_t)bid * np + k;
        if (ordered_bits(plane_dist(planes[k], xw)) == w.dmin[o]) atomicMin(w.imin + o, (unsigned int)p);
    }
}

MPM_DEV void dcross(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

MPM_DEV void dmatvec(const float* M, const double* v, double* out) {
    for (int i = 0; i < 3; ++i) out[i] = (double)M[3 * i] * v[0] + (double)M[3 * i + 1] * v[1] + (double)M[3 * i + 2] * v[2];
}

// _apply_rigid_restitution, mpm_solver_warp.py:874-962, one thread per body, planes in order;
// float64 arithmetic with the body state rounded to float32 between planes, as the
// reference's numpy round trip does
__global__ void k_rigid_restitution(mpm_particles_t P, mpm_rigid_body_t* bodies, int nb,
                                    const mpm_rigid_plane_t* __restrict__ planes, int np,
                                    const int32_t* __restrict__ active, void* ws) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const RigidWs w = rigid_ws(ws, nb, np);
    const unsigned act = (unsigned)*active;
    const mpm_rigid_body_t body0 = bodies[b];  // world positions start from this pose
    mpm_rigid_body_t body = body0;
    const m33 R = body_R(body);
    m33 Ii;
    for (int i = 0; i < 9; ++i) Ii.a[i] = body.inv_inertia[i];
    const m33 Iw = mmul(mmul(R, Ii), mT(R));  // float32, as numpy's R @ I_inv @ R.T
    const double M = (double)body.mass;
    double shift[3] = {0.0, 0.0, 0.0};  // accumulated push (the world positions move with it)
    bool modified = false;
    for (int k = 0; k < np; ++k) {
        const mpm_rigid_plane_t& pl = planes[k];
        const size_t o = (size_t)b * np + k;
        if (!plane_active(act, k) || w.imin[o] == 0xffffffffu) continue;
        const int pc = (int)w.imin[o];
        const double n[3] = {(double)pl.normal[0], (double)pl.normal[1], (double)pl.normal[2]};
        const v3 xw0 = world_pos(body0, ld3(P.x_ref, P.stride, pc));  // before any push
        const double dmin = plane_dist(pl, xw0) + shift[0] * n[0] + shift[1] * n[1] + shift[2] * n[2];
        if (dmin >= 0.0) continue;
        double push[3];
        for (int i = 0; i < 3; ++i) {
            push[i] = -dmin * n[i];
            body.x_cm[i] = (float)((double)body.x_cm[i] + push[i]);
            shift[i] += push[i];
        }
        modified = true;
        const double xc[3] = {(double)xw0.x + shift[0], (double)xw0.y + shift[1], (double)xw0.z + shift[2]};
        const double r[3] = {xc[0] - (double)body.x_cm[0], xc[1] - (double)body.x_cm[1], xc[2] - (double)body.x_cm[2]};
        const double om[3] = {(double)body.omega[0], (double)body.omega[1], (double)body.omega[2]};
        double wr[3];
        dcross(om, r, wr);
        const double vc[3] = {(double)body.v_cm[0] + wr[0], (double)body.v_cm[1] + wr[1], (double)body.v_cm[2] + wr[2]};
        const double vn = vc[0] * n[0] + vc[1] * n[1] + vc[2] * n[2];
        if (vn >= 0.0) continue;  // separating
        double rxn[3], Irxn[3], q[3];
        dcross(r, n, rxn);
        dmatvec(Iw.a, rxn, Irxn);
        dcross(Irxn, r, q);
        const double denom = 1.0 / M + (q[0] * n[0] + q[1] * n[1] + q[2] * n[2]);
        const double Jn = -(1.0 + (double)pl.restitution) * vn / denom;
        double vcm[3], omn[3];
        for (int i = 0; i < 3; ++i) {
            vcm[i] = (double)body.v_cm[i] + (Jn / M) * n[i];
            omn[i] = (double)body.omega[i] + Jn * Irxn[i];
        }
        for (int i = 0; i < 3; ++i) {
            body.v_cm[i] = (float)vcm[i];
            body.omega[i] = (float)omn[i];
        }
        if (pl.friction > 0.0f) {
            double vt[3];
            for (int i = 0; i < 3; ++i) vt[i] = vc[i] - vn * n[i];
            const double vtm = sqrt(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
            if (vtm > 1e-10) {
                const double th[3] = {vt[0] / vtm, vt[1] / vtm, vt[2] / vtm};
                double rxt[3], Irxt[3], qt[3];
                dcross(r, th, rxt);
                dmatvec(Iw.a, rxt, Irxt);
                dcross(Irxt, r, qt);
                const double dt_ = 1.0 / M + (qt[0] * th[0] + qt[1] * th[1] + qt[2] * th[2]);
                const double Jt = fmin(vtm / dt_, (double)pl.friction * Jn);
                for (int i = 0; i < 3; ++i) {
                    body.v_cm[i] = (float)((double)body.v_cm[i] - (Jt / M) * th[i]);
                    body.omega[i] = (float)((double)body.omega[i] - Jt * Irxt[i]);
                }
            }
        }
    }
    if (modified) bodies[b] = body;
}

// rigid_particle_update, mpm_utils.py:1469-1493
__global__ __launch_bounds__(kThreads) void k_rigid_update(mpm_particles_t P,
                                                           const mpm_rigid_body_t* __restrict__ bodies, int nb) {
    const int p = blockIdx.x * kThreads + threadIdx.x;
    int bid;
    if (!rigid_particle(P, p, nb, bid)) return;
    const int S = P.stride;
    const mpm_rigid_body_t& b = bodies[bid];
    const v3 xp = world_pos(b, ld3(P.x_ref, S, p));
    const v3 r = mk3(xp.x - b.x_cm[0], xp.y - b.x_cm[1], xp.z - b.x_cm[2]);
    const v3 om = mk3(b.omega[0], b.omega[1], b.omega[2]);
    const v3 v = mk3(b.v_cm[0], b.v_cm[1], b.v_cm[2]) + cross(om, r);
    st3(P.v, S, p, v);
    st3(P.x, S, p, xp);
    m33 sk = mzero();
    sk.a[1] = -om.z; sk.a[2] = om.y;
    sk.a[3] = om.z;  sk.a[5] = -om.x;
    sk.a[6] = -om.y; sk.a[7] = om.x;
    st9(P.C, S, p, sk);
    st9(P.F, S, p, meye());
}

inline unsigned blocks_for(long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

extern "C" {

int64_t mpm_rigid_workspace_bytes(int32_t n_bodies, int32_t n_planes) {
    if (n_bodies < 0 || n_planes < 0) return -1;
    const long pairs = (long)n_bodies * n_planes;
    return (int64_t)(pairs * 8 + (pairs + (pairs & 1)) * 4 + (long)n_bodies * kAccum * 4);
}

int32_t mpm_rigid_step(const mpm_particles_t* P, const mpm_grid_t* G, mpm_rigid_body_t* bodies, int32_t n_bodies,
                       const mpm_rigid_plane_t* planes, int32_t n_planes, const int32_t* active, float dt,
                       void* workspace, void* stream) {
    if (!P || !G || !bodies || n_bodies < 1 || n_planes < 0 || n_planes > 32 || (n_planes > 0 && !planes) ||
        !active || !workspace ||
        P->n < 0 || P->stride < P->n || !P->x || !P->v || !P->C || !P->F || !P->mass || !P->material ||
        !P->selection || !P->rigid_id || !P->x_ref || !G->node_out || G->n_scenes != 1 || G->block_slot)
        return MPM_ERR_ARG;
    hipStream_t s = (hipStream_t)stream;
    const int nb = n_bodies, np = n_planes;
    const unsigned pb = P->n > 0 ? blocks_for(P->n) : 0;
    hipLaunchKernelGGL(k_rigid_clear, dim3(1), dim3(kThreads), 0, s, workspace, nb, np);
    if (pb) hipLaunchKernelGGL(k_rigid_gather, dim3(pb), dim3(kThreads), 0, s, *P, *G, bodies, nb, np, workspace);
    hipLaunchKernelGGL(k_rigid_integrate, dim3((nb + 63) / 64), dim3(64), 0, s, bodies, nb, np, workspace, dt);
    if (np > 0 && pb) {
        hipLaunchKernelGGL(k_rigid_contact_min, dim3(pb), dim3(kThreads), 0, s, *P, bodies, nb, planes, np, active,
                           workspace);
        hipLaunchKernelGGL(k_rigid_contact_index, dim3(pb), dim3(kThreads), 0, s, *P, bodies, nb, planes, np, active,
                           workspace);
        hipLaunchKernelGGL(k_rigid_restitution, dim3((nb + 63) / 64), dim3(64), 0, s, *P, bodies, nb, planes, np,
                           active, workspace);
    }
    if (pb) hipLaunchKernelGGL(k_rigid_update, dim3(pb), dim3(kThreads), 0, s, *P, bodies, nb);
    return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // extern "C"
