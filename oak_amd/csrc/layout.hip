// State layout on the GPU: the zone permutation of ensembles when they are
// loaded and saved (loadVector/loadEnsemble and saveVector with
// ModML%permute, assimilation.F90:869-880, 901, 1268-1359): rows gathered
// from the full (land and sea) field vectors into the permuted sea-only
// state, and scattered back.  HBM-bound copies: each thread moves one row
// element of one member, consecutive threads consecutive rows of the
// permuted state (coalesced on that side; the full-field side follows the
// index runs, which are long for column zones).
#include <stdint.h>

#include "common.hpp"
#include "internal.hpp"

namespace oakg {

template <typename T, bool GATHER>
__global__ __launch_bounds__(256) void k_permute_rows(int64_t n, int N, const int64_t* __restrict__ idx,
                                                      const T* __restrict__ in, int64_t ldi,
                                                      T* __restrict__ out, int64_t ldo,
                                                      int64_t lim, unsigned* __restrict__ err) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (l >= n || k >= N) return;
  const int64_t j = idx[l];
  if (j < 0 || j >= lim) {
    if (k == 0) atomicOr(err, OAKG_DERR_IDX);
    return;
  }
  if (GATHER)
    out[l + (int64_t)k * ldo] = in[j + (int64_t)k * ldi];
  else
    out[j + (int64_t)k * ldo] = in[l + (int64_t)k * ldi];
}

void launch_permute_rows(bool gather, int elem, int64_t n, int N, const int64_t* idx,
                         const void* in, int64_t ldi, void* out, int64_t ldo, int64_t lim,
                         unsigned* err, hipStream_t st) {
  if (n <= 0 || N <= 0) return;
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)N), blk(256);
  if (elem == 8) {
    if (gather)
      hipLaunchKernelGGL((k_permute_rows<double, true>), grid, blk, 0, st, n, N, idx,
                         (const double*)in, ldi, (double*)out, ldo, lim, err);
    else
      hipLaunchKernelGGL((k_permute_rows<double, false>), grid, blk, 0, st, n, N, idx,
                         (const double*)in, ldi, (double*)out, ldo, lim, err);
  } else {
    if (gather)
      hipLaunchKernelGGL((k_permute_rows<float, true>), grid, blk, 0, st, n, N, idx,
                         (const float*)in, ldi, (float*)out, ldo, lim, err);
    else
      hipLaunchKernelGGL((k_permute_rows<float, false>), grid, blk, 0, st, n, N, idx,
                         (const float*)in, ldi, (float*)out, ldo, lim, err);
  }
}

}  // namespace oakg
