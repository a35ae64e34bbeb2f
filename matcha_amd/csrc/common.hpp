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

// Sum of v over the wave, returned to every lane (as a wave-uniform value):
// within each row of 16 lanes four DPP additions (xor 1, xor 2, the half-row
// and the row mirror: after them every lane holds its row's sum), then the
// four row sums read as scalars and added in row order.  One VALU instruction
// per exchange and no LDS round trips (a shuffle goes through ds_bpermute and
// its address arithmetic).  Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
  // a plain DPP move (no old value): the compiler folds it into the addition
  return v + __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v = dpp_add<0xB1>(v);   // quad_perm [1, 0, 3, 2]
  v = dpp_add<0x4E>(v);   // quad_perm [2, 3, 0, 1]
  v = dpp_add<0x141>(v);  // row_half_mirror
  v = dpp_add<0x140>(v);  // row_mirror
  const int b = __float_as_int(v);
  return ((__int_as_float(__builtin_amdgcn_readlane(b, 0)) + __int_as_float(__builtin_amdgcn_readlane(b, 16))) +
          __int_as_float(__builtin_amdgcn_readlane(b, 32))) +
         __int_as_float(__builtin_amdgcn_readlane(b, 48));
}

// floor(sqrt(x)) for 0 <= x < 2^24: the hardware square root (within 1 ulp)
// and one correction each way.
__device__ __forceinline__ int isqrt_small(int x) {
  int r = (int)__builtin_amdgcn_sqrtf((float)x);
  if (r * r > x) --r;
  if ((r + 1) * (r + 1) <= x) ++r;
  return r;
}

// Maximum of w over the wave, returned to every lane.  Within each row of 16
// lanes four DPP exchanges (xor 1, xor 2, the half-row and the row mirror:
// after them every lane holds its row's maximum), then the four row maxima
// read as scalars: no LDS round trips, unlike a shuffle through ds_bpermute.
// Every lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_max_u64(unsigned long long w) {
  const int lo = (int)(unsigned)w, hi = (int)(unsigned)(w >> 32);
  const unsigned olo = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const unsigned ohi = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  const unsigned long long o = ((unsigned long long)ohi << 32) | olo;
  return o > w ? o : w;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long w) {
  w = dpp_max_u64<0xB1>(w);   // quad_perm [1, 0, 3, 2]
  w = dpp_max_u64<0x4E>(w);   // quad_perm [2, 3, 0, 1]
  w = dpp_max_u64<0x141>(w);  // row_half_mirror
  w = dpp_max_u64<0x140>(w);  // row_mirror
  const int lo = (int)(unsigned)w, hi = (int)(unsigned)(w >> 32);
  unsigned long long r = 0ull;
#pragma unroll
  for (int row = 0; row < 4; ++row) {
    const unsigned long long x = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane(hi, 16 * row) << 32) |
                                 (unsigned)__builtin_amdgcn_readlane(lo, 16 * row);
    r = x > r ? x : r;
  }
  return r;
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

// Raw buffer resource over nbytes at p: loads at byte offsets >= nbytes return 0
// (hardware range check), so out-of-volume samples need no branch.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)nbytes, 0x00020000);
}
constexpr unsigned kOutOfRange = 0x80000000u;
__device__ __forceinline__ float4 load_f4(__amdgpu_buffer_rsrc_t r, unsigned off) {
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  const v4u v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
  return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z), __uint_as_float(v.w));
}
__device__ __forceinline__ float load_f1(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// Shifted patch value at array index (a0, a1, a2): vol[a + s] with trilinear
// sampling for fractional s and the border slice zeroed as the reference does.
struct ShiftSpec {
  float s[3];
  int integral;       // all components integral
  int zero_idx[3];    // index of the slice to zero along each axis (-1: none)
  int si[3];
};

__device__ __forceinline__ float patch_value(const float* __restrict__ v, int N, const ShiftSpec& sp,
                                             int a0, int a1, int a2) {
  if ((sp.zero_idx[0] >= 0 && a0 == sp.zero_idx[0]) || (sp.zero_idx[1] >= 0 && a1 == sp.zero_idx[1]) ||
      (sp.zero_idx[2] >= 0 && a2 == sp.zero_idx[2]))
    return 0.f;
  if (sp.integral) {
    const int z = a0 + sp.si[0], y = a1 + sp.si[1], x = a2 + sp.si[2];
    if (z < 0 || z >= N || y < 0 || y >= N || x < 0 || x >= N) return 0.f;
    return v[((long long)z * N + y) * N + x];
  }
  return tri_sample_real(v, N, a0 + sp.s[0], a1 + sp.s[1], a2 + sp.s[2]);
}

__device__ __forceinline__ void make_shift_spec(const float* __restrict__ shift_zyx, int b, int N, ShiftSpec& sp) {
  // the patch shift in array-axis order is the flipped (z, y, x) shift
  sp.integral = 1;
  for (int d = 0; d < 3; ++d) {
    const float s = shift_zyx ? shift_zyx[b * 3 + (2 - d)] : 0.f;
    sp.s[d] = s;
    const float a = fabsf(s);
    const float fl = floorf(a);
    sp.si[d] = (int)rintf(s);
    sp.zero_idx[d] = -1;
    if (a - fl > 0.f) {
      sp.integral = 0;
      // idx = floor(|s|) if s < 0 else -ceil(|s|) (python negative index)
      sp.zero_idx[d] = (s < 0.f) ? (int)fl : N - (int)ceilf(a);
    }
  }
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
