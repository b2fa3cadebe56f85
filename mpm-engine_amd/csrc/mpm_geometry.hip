// Mesh -> distance-field voxelisation on the GPU: the input producer of the mesh-SDF and
// CDF colliders (the reference builds these on the host, src/warpmpm/geometry/
// mesh_sdf.py:328-366 and :412-481, O(voxels x faces)).
//
// One thread per voxel of a cubic res^3 grid (origin, cell; z fastest). Triangles stream
// through LDS in tiles shared by the workgroup; for each the thread evaluates the exact
// squared point-triangle distance (closest point by Voronoi-region tests, Ericson,
// "Real-Time Collision Detection", 2005, sec. 5.1.5) and, for the SDF, the solid angle the
// triangle subtends (Van Oosterom & Strackee, IEEE Trans. Biomed. Eng. 30, 1983), whose sum
// over a closed mesh is 4 pi times the winding number. For the CDF it keeps the side of
// the nearest triangle (sign of (p - closest) . face normal) and, in a second kernel, the
// distance to the open rim (boundary edges).
#include <hip/hip_runtime.h>

#include "mpm_device.h"

namespace {

constexpr int kThreads = 256;
constexpr int kTile = 256;  // triangles per LDS tile

struct Tri {
    float a[3], b[3], c[3];
};

MPM_DEV v3 ld(const float* p) { return mk3(p[0], p[1], p[2]); }

// closest point on triangle abc to p (Ericson 5.1.5)
MPM_DEV v3 closest_on_triangle(v3 p, v3 a, v3 b, v3 c) {
    const v3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    const v3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + ab * (d1 / (d1 - d3));
    const v3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + ac * (d2 / (d2 - d6));
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
        return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
    const float den = 1.0f / (va + vb + vc);
    return a + ab * (vb * den) + ac * (vc * den);
}

// solid angle of triangle abc seen from p (Van Oosterom & Strackee)
MPM_DEV float solid_angle(v3 p, v3 a, v3 b, v3 c) {
    const v3 x = a - p, y = b - p, z = c - p;
    const float lx = len(x), ly = len(y), lz = len(z);
    const float num = dot(x, cross(y, z));
    const float den = lx * ly * lz + dot(x, y) * lz + dot(x, z) * ly + dot(y, z) * lx;
    return 2.0f * atan2f(num, den);
}

// mode 0: dist = unsigned distance, aux = winding number; mode 1: dist = unsigned
// distance, aux = side of the nearest triangle (+1 / -1)
__global__ __launch_bounds__(kThreads) void k_mesh_field(const float* __restrict__ verts, int nv,
                                                         const int32_t* __restrict__ faces, int nf, float o0,
                                                         float o1, float o2, float cell, int res, int mode,
                                                         float* __restrict__ dist, float* __restrict__ aux) {
    __shared__ Tri s_tri[kTile];
    const long nvox = (long)res * res * res;
    const long t = (long)blockIdx.x * kThreads + threadIdx.x;
    const bool live = t < nvox;
    const int k = (int)(t % res), j = (int)((t / res) % res), i = (int)(t / ((long)res * res));
    const v3 p = mk3(o0 + cell * (float)i, o1 + cell * (float)j, o2 + cell * (float)k);
    float best = 3.0e38f, wind = 0.0f, side = 1.0f;
    for (int base = 0; base < nf; base += kTile) {
        __syncthreads();
        for (int q = threadIdx.x; q < kTile && base + q < nf; q += kThreads) {
            const int f = base + q;
            const int ia = faces[3 * f], ib = faces[3 * f + 1], ic = faces[3 * f + 2];
            Tri tr;
            if ((unsigned)ia < (unsigned)nv && (unsigned)ib < (unsigned)nv && (unsigned)ic < (unsigned)nv) {
                for (int d = 0; d < 3; ++d) {
                    tr.a[d] = verts[3 * ia + d];
                    tr.b[d] = verts[3 * ib + d];
                    tr.c[d] = verts[3 * ic + d];
                }
            } else {  // bad vertex index: a degenerate far-away triangle (no distance, no angle)
                for (int d = 0; d < 3; ++d) tr.a[d] = tr.b[d] = tr.c[d] = 1.0e30f;
            }
            s_tri[q] = tr;
        }
        __syncthreads();
        const int m = min(kTile, nf - base);
        if (!live) continue;
        for (int q = 0; q < m; ++q) {
            const v3 a = ld(s_tri[q].a), b = ld(s_tri[q].b), c = ld(s_tri[q].c);
            const v3 cp = closest_on_triangle(p, a, b, c);
            const v3 dvec = p - cp;
            const float d2 = dot(dvec, dvec);
            if (mode == 0) {
                wind += solid_angle(p, a, b, c);
                best = fminf(best, d2);
            } else if (d2 < best) {
                best = d2;
                side = dot(dvec, cross(b - a, c - a)) >= 0.0f ? 1.0f : -1.0f;
            }
        }
    }
    if (!live) return;
    dist[t] = sqrtf(best);
    aux[t] = mode == 0 ? wind * 0.07957747154594767f : side;  // 1 / (4 pi)
}

// distance from every voxel to the nearest segment (the open rim's boundary edges)
__global__ __launch_bounds__(kThreads) void k_segment_distance(const float* __restrict__ seg, int ns, float o0,
                                                               float o1, float o2, float cell, int res,
                                                               float* __restrict__ dist) {
    __shared__ float s_seg[kTile][6];
    const long nvox = (long)res * res * res;
    const long t = (long)blockIdx.x * kThreads + threadIdx.x;
    const bool live = t < nvox;
    const int k = (int)(t % res), j = (int)((t / res) % res), i = (int)(t / ((long)res * res));
    const v3 p = mk3(o0 + cell * (float)i, o1 + cell * (float)j, o2 + cell * (float)k);
    float best = 3.0e38f;
    for (int base = 0; base < ns; base += kTile) {
        __syncthreads();
        for (int q = threadIdx.x; q < kTile && base + q < ns; q += kThreads)
            for (int d = 0; d < 6; ++d) s_seg[q][d] = seg[6 * (base + q) + d];
        __syncthreads();
        const int m = min(kTile, ns - base);
        if (!live) continue;
        for (int q = 0; q < m; ++q) {
            const v3 a = mk3(s_seg[q][0], s_seg[q][1], s_seg[q][2]);
            const v3 ab = mk3(s_seg[q][3], s_seg[q][4], s_seg[q][5]) - a;
            const float l2 = dot(ab, ab);
            const float u = l2 > 0.0f ? fminf(fmaxf(dot(p - a, ab) / l2, 0.0f), 1.0f) : 0.0f;
            const v3 d = p - (a + ab * u);
            best = fminf(best, dot(d, d));
        }
    }
    if (live) dist[t] = sqrtf(best);
}

}  // namespace

extern "C" {

int32_t mpm_mesh_field(const float* verts, int32_t nv, const int32_t* faces, int32_t nf, const float* origin,
                       float cell, int32_t res, int32_t mode, float* dist, float* aux, void* stream) {
    if (!verts || nv < 3 || !faces || nf < 1 || !origin || !(cell > 0.0f) || res < 2 || res > 1024 || (mode != 0 && mode != 1) ||
        !dist || !aux)
        return MPM_ERR_ARG;
    const long nvox = (long)res * res * res;
    hipLaunchKernelGGL(k_mesh_field, dim3((unsigned)((nvox + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       (hipStream_t)stream, verts, nv, faces, nf, origin[0], origin[1], origin[2], cell, res, mode,
                       dist, aux);
    return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

int32_t mpm_segment_distance(const float* segments, int32_t ns, const float* origin, float cell, int32_t res,
                             float* dist, void* stream) {
    if (!segments || ns < 1 || !origin || !(cell > 0.0f) || res < 2 || res > 1024 || !dist) return MPM_ERR_ARG;
    const long nvox = (long)res * res * res;
    hipLaunchKernelGGL(k_segment_distance, dim3((unsigned)((nvox + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                       (hipStream_t)stream, segments, ns, origin[0], origin[1], origin[2], cell, res, dist);
    return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // extern "C"
