// Stable device radix sort of (block key, particle index) pairs for the block sort
// (Solver._sort_particles, src/warpmpm/core/solver.py:454-467). rocPRIM's radix sort
// (through hipCUB) is stable, so equal keys keep their index order exactly like the
// reference's np.argsort(kind="stable").
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "../../include/mpm_hip.h"

extern "C" {

int64_t mpm_sort_temp_bytes(int32_t n) {
    size_t bytes = 0;
    if (n <= 0) return 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const int32_t*)nullptr, (int32_t*)nullptr,
                                           (const int32_t*)nullptr, (int32_t*)nullptr, n) != hipSuccess)
        return -1;
    return (int64_t)bytes;
}

int32_t mpm_sort_pairs(const int32_t* keys, int32_t* keys_out, const int32_t* vals, int32_t* vals_out, int32_t n,
                       void* temp, int64_t temp_bytes, void* stream) {
    if (n < 0 || (n > 0 && (!keys || !keys_out || !vals || !vals_out || !temp))) return MPM_ERR_ARG;
    if (n == 0) return MPM_OK;
    size_t bytes = (size_t)temp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys, keys_out, vals, vals_out, n, 0,
                                                      (int)(sizeof(int32_t) * 8), (hipStream_t)stream);
    return e == hipSuccess ? MPM_OK : MPM_ERR_HIP;
}

}  // extern "C"
