// Shared helpers for the gfx950 kernels of matcha_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#define MCH_OK 0
#define MCH_ERR_ARG 1001
#define MCH_ERR_SIZE 1002
#define MCH_ERR_HIP 1003
#define MCH_ERR_UNSUPPORTED 1004

namespace mch {

// Last error message, per thread (the library is not re-entrant per plan).
void set_error(const char* fmt, ...);

inline int hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    set_error("%s failed: %s", what, hipGetErrorString(e));
    return MCH_ERR_HIP;
  }
  return MCH_OK;
}

constexpr int kWave = 64;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// D layout size for 0 <= l <= L: sum (2l+1)^2.
__host__ __device__ __forceinline__ int dsize(int L) {
  return L < 0 ? 0 : (L + 1) * (2 * L + 1) * (2 * L + 3) / 3;
}
__host__ __device__ __forceinline__ int dindex(int l, int mp, int m) {
  return dsize(l - 1) + (mp + l) * (2 * l + 1) + (m + l);
}
// Repacked sigma: for shell l, fundamental pairs p < (l+1)^2, 4 images each.
// Offset of shell l in float2 units: 4 * sum_{k=1}^{l} k^2.
__host__ __device__ __forceinline__ int pk_offset(int l) {
  return 4 * (l * (l + 1) * (2 * l + 1) / 6);
}
__host__ __device__ __forceinline__ int pk_size(int L) { return pk_offset(L + 1); }

// Trilinear sample of an N^3 volume (z, y, x index order) at index
// coordinates, corners outside the volume counting as zero (grid_sample with
// align_corners=True and zero padding, in index units).
__device__ __forceinline__ float tri_sample_real(const float* __restrict__ v, int N, float z, float y,
                                                 float x) {
  const float fz0 = floorf(z), fy0 = floorf(y), fx0 = floorf(x);
  const int z0 = (int)fz0, y0 = (int)fy0, x0 = (int)fx0;
  const float fz = z - fz0, fy = y - fy0, fx = x - fx0;
  float acc = 0.f;
#pragma unroll
  for (int dz = 0; dz < 2; ++dz)
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
      for (int dx = 0; dx < 2; ++dx) {
        const int zi = z0 + dz, yi = y0 + dy, xi = x0 + dx;
        if (zi < 0 || zi >= N || yi < 0 || yi >= N || xi < 0 || xi >= N) continue;
        const float w = (dz ? fz : 1.f - fz) * (dy ? fy : 1.f - fy) * (dx ? fx : 1.f - fx);
        acc += w * v[((long long)zi * N + yi) * N + xi];
      }
  return acc;
}

// XCD-aware block remap: the hardware hands consecutive workgroups to the 8
// XCDs round-robin; map them so that a contiguous range of logical blocks
// (e.g. all candidates of one particle) lands on one XCD and shares its L2.
// Bijective for any grid size: XCDs 0..r-1 take q+1 logical blocks, the rest
// q (n = 8q + r).
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7;
  const int x = bid & 7, i = bid >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

}  // namespace mch
