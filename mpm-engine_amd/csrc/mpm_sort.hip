// Stable device radix sort of (block key, particle index) pairs for the block sort
// (Solver._sort_particles, src/warpmpm/core/solver.py:454-467). rocPRIM's radix sort is
// stable, so equal keys keep their index order exactly like the reference's
// np.argsort(kind="stable").
//
// The sort is called inside captured hipGraphs (the simulators replay the whole
// permutation as one graph), so it must launch kernels only. rocPRIM's default radix sort
// switches above 2^20 items to its onesweep algorithm, which clears its digit histogram,
// lookback states and block counter with hipMemsetAsync: memset nodes in a captured graph,
// which this platform does not replay reliably (a stale lookback state sends keys to
// wrong offsets). The configuration below keeps rocPRIM's block-sort + merge path, which
// launches only kernels, for every size.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "../../include/mpm_hip.h"

namespace {
using SortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                              rocprim::default_config, (size_t)1 << 32>;
}

extern "C" {

int64_t mpm_sort_temp_bytes(int32_t n) {
    size_t bytes = 0;
    if (n <= 0) return 0;
    if (rocprim::radix_sort_pairs<SortConfig>(nullptr, bytes, (const int32_t*)nullptr, (int32_t*)nullptr,
                                              (const int32_t*)nullptr, (int32_t*)nullptr, (size_t)n) != hipSuccess)
        return -1;
    return (int64_t)bytes;
}

int32_t mpm_sort_pairs(const int32_t* keys, int32_t* keys_out, const int32_t* vals, int32_t* vals_out, int32_t n,
                       void* temp, int64_t temp_bytes, void* stream) {
    if (n < 0 || (n > 0 && (!keys || !keys_out || !vals || !vals_out || !temp))) return MPM_ERR_ARG;
    if (n == 0) return MPM_OK;
    size_t bytes = (size_t)temp_bytes;
    hipError_t e = rocprim::radix_sort_pairs<SortConfig>(temp, bytes, keys, keys_out, vals, vals_out, (size_t)n, 0,
                                                         (unsigned)(sizeof(int32_t) * 8), (hipStream_t)stream);
    return e == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // extern "C"

// ------------------------------------------------------------------ particle-pass chunks
// Chunk boundaries over block-sorted particles: runs of particles that share a group of gz
// consecutive blocks of one block column (scene, bx, by, bz / gz), each cut into chunks of
// at most `chunk` particles. A fixed split into 256 consecutive particles lets a chunk
// straddle a jump in key order (the end of one block column and the start of the next):
// its stencil box then spans two distant clusters, overflows the particle pass's LDS tile
// and sends the whole chunk through global memory; group-aligned chunks keep every box
// within a column segment of gz blocks.
namespace {

constexpr int kSortThreads = 256;

__device__ inline int group_of(int key, int shift, int nb, int gz) {
    const int bk = key >> shift;                  // scene * nb^3 + (bx * nb + by) * nb + bz
    const int col = bk / nb, bz = bk - col * nb;  // col = scene * nb^2 + bx * nb + by
    return col * ((nb + gz - 1) / gz) + bz / gz;
}

__global__ void k_group_flags(const int32_t* __restrict__ keys, int n, int shift, int nb, int gz,
                              int32_t* __restrict__ flags) {
    const int p = blockIdx.x * kSortThreads + threadIdx.x;
    if (p >= n) return;
    flags[p] = (p == 0 || group_of(keys[p], shift, nb, gz) != group_of(keys[p - 1], shift, nb, gz)) ? 1 : 0;
}

// gs[g] = first particle of group g; gs[G] = n and *ng = G (written by the last thread)
__global__ void k_group_starts(const int32_t* __restrict__ flags, const int32_t* __restrict__ gidx, int n,
                               int32_t* __restrict__ gs, int32_t* __restrict__ ng) {
    const int p = blockIdx.x * kSortThreads + threadIdx.x;
    if (p >= n) return;
    if (flags[p]) gs[gidx[p]] = p;
    if (p == n - 1) {
        const int G = gidx[p] + flags[p];
        gs[G] = n;
        *ng = G;
    }
}

__global__ void k_group_chunks(const int32_t* __restrict__ gs, const int32_t* __restrict__ ng, int n, int chunk,
                               int32_t* __restrict__ nck) {
    const int g = blockIdx.x * kSortThreads + threadIdx.x;
    if (g >= n) return;
    nck[g] = g < *ng ? (gs[g + 1] - gs[g] + chunk - 1) / chunk : 0;
}

__global__ void k_chunk_fill(const int32_t* __restrict__ gs, const int32_t* __restrict__ nck,
                             const int32_t* __restrict__ coff, const int32_t* __restrict__ ng, int n, int chunk,
                             int32_t* __restrict__ chunk_start, int32_t* __restrict__ count) {
    const int g = blockIdx.x * kSortThreads + threadIdx.x;
    const int G = *ng;
    if (g >= G) return;
    for (int k = 0; k < nck[g]; ++k) chunk_start[coff[g] + k] = gs[g] + k * chunk;
    if (g == G - 1) {
        const int total = coff[g] + nck[g];
        chunk_start[total] = n;
        *count = total;
    }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int64_t mpm_sort_chunks_temp_bytes(int32_t n) {
    if (n <= 0) return 0;
    size_t scan = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const int32_t*)nullptr, (int32_t*)nullptr, n) != hipSuccess)
        return -1;
    // scan temp, then flags, gidx, gs (n + 1), nck, coff, ng
    return (int64_t)(align256(scan) + 5 * align256(sizeof(int32_t) * ((size_t)n + 1)) + 256);
}

int32_t mpm_sort_chunks(const int32_t* keys, int32_t n, int32_t key_shift, int32_t nb, int32_t gz, int32_t chunk,
                        int32_t* chunk_start, int32_t* count, void* temp, int64_t temp_bytes, void* stream) {
    if (n < 0 || (n > 0 && (!keys || !chunk_start || !count || !temp)) || nb <= 0 || gz <= 0 || chunk <= 0 ||
        (key_shift != 0 && key_shift != 6))
        return MPM_ERR_ARG;
    if (n == 0) return MPM_OK;
    if (temp_bytes < mpm_sort_chunks_temp_bytes(n)) return MPM_ERR_SIZE;
    hipStream_t s = (hipStream_t)stream;
    size_t scan = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const int32_t*)nullptr, (int32_t*)nullptr, n) != hipSuccess)
        return MPM_ERR_HIP;
    char* base = static_cast<char*>(temp);
    const size_t arr = align256(sizeof(int32_t) * ((size_t)n + 1));
    void* cub_tmp = base;
    int32_t* flags = reinterpret_cast<int32_t*>(base + align256(scan));
    int32_t* gidx = reinterpret_cast<int32_t*>(base + align256(scan) + arr);
    int32_t* gs = reinterpret_cast<int32_t*>(base + align256(scan) + 2 * arr);
    int32_t* nck = reinterpret_cast<int32_t*>(base + align256(scan) + 3 * arr);
    int32_t* coff = reinterpret_cast<int32_t*>(base + align256(scan) + 4 * arr);
    int32_t* ng = reinterpret_cast<int32_t*>(base + align256(scan) + 5 * arr);
    const dim3 grid((unsigned)((n + kSortThreads - 1) / kSortThreads)), block(kSortThreads);
    hipLaunchKernelGGL(k_group_flags, grid, block, 0, s, keys, n, key_shift, nb, gz, flags);
    size_t b = scan;
    if (hipcub::DeviceScan::ExclusiveSum(cub_tmp, b, flags, gidx, n, s) != hipSuccess) return MPM_ERR_HIP;
    hipLaunchKernelGGL(k_group_starts, grid, block, 0, s, flags, gidx, n, gs, ng);
    hipLaunchKernelGGL(k_group_chunks, grid, block, 0, s, gs, ng, n, chunk, nck);
    b = scan;
    if (hipcub::DeviceScan::ExclusiveSum(cub_tmp, b, nck, coff, n, s) != hipSuccess) return MPM_ERR_HIP;
    hipLaunchKernelGGL(k_chunk_fill, grid, block, 0, s, gs, nck, coff, ng, n, chunk, chunk_start, count);
    return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // extern "C"

// ------------------------------------------------------------------------ block pool
// Slots of the active blocks: an exclusive scan of the 0/1 active flags gives each
// active code its rank among the active codes (its slot, so slots follow code order and
// spatial neighbours within a block column sit next to each other in the node arrays);
// inactive codes get -1.
namespace {

__global__ void k_pool_fill(const int32_t* __restrict__ cur, int n, int32_t* __restrict__ slot,
                            int32_t* __restrict__ list, int32_t* __restrict__ count) {
    const int t = blockIdx.x * kSortThreads + threadIdx.x;
    if (t >= n) return;
    const int s = slot[t];  // rank among the active codes below t
    const int on = cur[t] == 1;
    if (on) list[s] = t;
    slot[t] = on ? s : -1;
    if (t == n - 1) *count = s + on;
}

}  // namespace

extern "C" {

int64_t mpm_pool_temp_bytes(int32_t n_codes) {
    if (n_codes <= 0) return 0;
    size_t scan = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const int32_t*)nullptr, (int32_t*)nullptr, n_codes) !=
        hipSuccess)
        return -1;
    return (int64_t)align256(scan);
}

int32_t mpm_pool_assign(const mpm_grid_t* G, const int32_t* cur, int32_t* slot, int32_t* list, int32_t* count,
                        void* temp, int64_t temp_bytes, void* stream) {
    if (!G || !cur || !slot || !list || !count || !temp || G->n_scenes < 1 || G->nb[0] <= 0 || G->nb[1] <= 0 ||
        G->nb[2] <= 0)
        return MPM_ERR_ARG;
    const long total = (long)G->nb[0] * G->nb[1] * G->nb[2] * G->n_scenes;
    if (total > 0x7fffffffL) return MPM_ERR_SIZE;
    const int n = (int)total;
    if (temp_bytes < mpm_pool_temp_bytes(n)) return MPM_ERR_SIZE;
    hipStream_t s = (hipStream_t)stream;
    size_t b = (size_t)temp_bytes;
    if (hipcub::DeviceScan::ExclusiveSum(temp, b, cur, slot, n, s) != hipSuccess) return MPM_ERR_HIP;
    hipLaunchKernelGGL(k_pool_fill, dim3((unsigned)((n + kSortThreads - 1) / kSortThreads)), dim3(kSortThreads), 0, s,
                       cur, n, slot, list, count);
    return hipGetLastError() == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // extern "C"
