// FFT shift match on gfx950.
//
// Reference behaviour (swing-research/Matcha, src/matcha/core/ShiftMatcher.py):
//   rotate_complex_volume :491-537, mask_full :442-453, extract_shifted_patch_batch
//   :558-616, apply_cosine_mask :540-555, rfftn(roll(., -N/2), "forward") :651-656,
//   compute_shift (complex space, no phase normalisation, score normalised,
//   upsample 1) :94-282, estimate_shift_like_relion :619-674.
//
// Per batch, integer lags (4 launches):
//   fwd_xy : one (b, z) slice: shifted, rolled patch split by linearity of the
//            cosine mask into A = (1 - w) v and bg * w (w = background weight);
//            FFT of A along x (two real rows per complex line) and along y,
//            written as T[b][kx][z][ky]; the slice's share of sum(w v)
//   z      : one workgroup per (b, kx, block of ky columns): T + bg * XMy (XMy =
//            the (x, y) spectrum of w, precomputed), FFT along z (1/N^3),
//            ||S||^2, the rotated reference sampled on the fly (trilinear,
//            ifftshift, mask_full, ctf), ||R||^2, P = conj(R) S, inverse FFT
//            along z
//   inv_xy : one (b, z) slice: inverse FFT along y, then along x of the half
//            axis zero-padded to N, real part, first argmax
//   final  : first argmax per particle, wrap to (-N/2, N/2], flip to the
//            reference's (z,y,x) labels, score / (||S|| ||R||)
// Sub-voxel lags (upsample > 1): fwd_x (x only, T[b][kx][z][y]), a plane pass
// (T + bg * XM, FFT along y and z, the same cross power, then the local matrix
// DFT over y and z), a kx pass and the final argmax.
// All 1-D transforms are compile-time four-step FFTs in LDS (fft_lds.hpp).
#include "common.hpp"
#include "fft_lds.hpp"

#include <math.h>
#include <type_traits>
#include <vector>

namespace {

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }
// threads of the row passes (N/2 lines): whole waves (the block reductions
// assume NT % 64 == 0) and a multiple of N (a thread keeps one column x)
template <int N> constexpr int row_threads() {
  return (2 * N) % 64 == 0 ? 2 * N : N / gcd_c(N, 64) * 64;
}
template <int N> constexpr int plane_threads() { return 4 * N; }
// loads issued back to back before their values are used
template <int PER> constexpr int load_batch() { return PER % 8 == 0 ? 8 : (PER % 4 == 0 ? 4 : (PER % 2 == 0 ? 2 : 1)); }

constexpr int kMaxLut = 2048;  // background-weight table entries (the 5-voxel cosine shell up to N = 192)

struct ShiftPlan {
  int N, H, max_batch;
  float radius, radius_p, width, wsum;  // cosine mask geometry, sum of background weights
  int lut_lo, lut_n;  // w(r2) tabulated for lut_lo <= r2 < lut_lo + lut_n; 0 below, 1 above
  float4* refpair;  // FP32 pairs of the centred reference (table_index)
  uint4* refh16;     // FP16 quads of the same reference, scaled by 2^e (table_index)
  float h16_unscale;  // 2^-e
  int reduced;      // 1: reduced precision (refh16, BF16 work spectrum), 0 (default): FP32
  float2* twid;     // (N) e^{-2 pi i j / N}
  float* lut;       // (lut_n)
  float2* XM;       // (H, N, N): x-spectrum of the rolled background weight, [kx][z][y]
  float2* XMy;      // (H, N, N): its (x, y) spectrum, [kx][z][ky]
  float2* T;        // (B, H, N, N) work spectrum: [b][kx][z][y] (sub-voxel path), [b][kx][z][ky]
  float* rnorm_part;  // (B, zparts): per column block on the integer path; (B, H) on the sub-voxel path
  float* snorm_part;  // same
  float* bg_part;     // (B, N)
  float* amax_val;    // (B, N)
  int* amax_idx;      // (B, N)
  int ntile_x;
  // sub-voxel refinement (upsample factor > 1): R = ceil(up (2 r + 1))
  int up, R;
  float2* t01;        // (R, N) e^{2 pi i (u - floor(R/2)) fftfreq(N, up)[k]}
  float2* t2;         // (R, H) the same for the half axis, fftfreq(H, up)
  float2* Qg;         // (B, H, R, R)
  float* up_val;      // (B, R)
  int* up_idx;        // (B, R)
  // optional per-kernel timing: events around the four launches of a search
  int timing;
  hipEvent_t ev[5];
};

using mch::kOutOfRange;
using mch::load_f1;
using mch::load_f4;
using mch::make_rsrc;
using mch::make_shift_spec;
using mch::patch_value;
using mch::ShiftSpec;

// Background weight of the cosine mask at squared distance r2 from the centre:
// 0 for r < radius (the patch is kept), 1 for r > radius_p (replaced by the
// background average), 0.5 + 0.5 cos(pi (radius_p - r) / width) in the shell.
// The masked patch is (1 - w) v + w bg.  Tabulated per integer r2 at plan time.
float bg_weight_host(int r2, float radius, float radius_p, float width) {
  const float r = sqrtf((float)r2);
  if (r < radius) return 0.f;
  if (r > radius_p) return 1.f;
  return 0.5f + 0.5f * cosf(3.14159265358979323846f * (radius_p - r) / width);
}
__device__ __forceinline__ float bg_weight(int r2, const float* lut, int lut_lo, int lut_n) {
  const int i = r2 - lut_lo;
  return i < 0 ? 0.f : (i >= lut_n ? 1.f : lut[i]);
}

template <int NT>
__device__ __forceinline__ float block_sum(float v, float* red) {
  static_assert(NT % 64 == 0, "whole waves");
  v = mch::wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float s = 0.f;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) s += red[w];
  return s;
}

// (i + n/2) mod n for 0 <= i < n
template <int N>
__device__ __forceinline__ int roll_half(int i) { return i + N / 2 < N ? i + N / 2 : i - N / 2; }

}  // namespace

// ---------------------------------------------------------------------------
// forward FFT along x of one (b, z) slice, two real rows per complex line
// ---------------------------------------------------------------------------
// The rolled, shifted patch is split by linearity of the mask: masked =
// A + bg * w with A = (1 - w) v.  This pass transforms A and accumulates the
// background numerator sum(w v) of the slice; the plane pass adds bg * XM.
// (Sub-voxel path; the integer path uses shift_fwd_xy_kernel.)
// Complex line j carries rows j (real part) and j + N/2 (imaginary part).
// Each thread keeps one column x (loads) and one row y (spectrum stores).
template <int N>
__global__ __launch_bounds__(row_threads<N>()) void shift_fwd_x_kernel(
    const float* __restrict__ vols, long long ldv, const float* __restrict__ shift_zyx,
    const float* __restrict__ lutg, int lut_lo, int lut_n, const float2* __restrict__ twg,
    float2* __restrict__ T, float* __restrict__ bg_part) {
  constexpr int NT = row_threads<N>(), NL = N / 2, S = N + 1, H = N / 2 + 1, YSTEP = NT / N;
  static_assert(NT % N == 0 && N % YSTEP == 0, "column split");
  extern __shared__ float4 smem4[];
  float2* buf = reinterpret_cast<float2*>(smem4);  // [line][x], stride S
  float2* tw = buf + NL * S;
  float* lut = reinterpret_cast<float*>(tw + N);
  __shared__ float red[NT / 64];
  const int z = blockIdx.x, b = blockIdx.y;
  for (int k = threadIdx.x; k < N; k += NT) tw[k] = twg[k];
  for (int k = threadIdx.x; k < lut_n; k += NT) lut[k] = lutg[k];
  __syncthreads();
  ShiftSpec sp;
  make_shift_spec(shift_zyx, b, N, sp);
  const float* v = vols + (long long)b * ldv;
  // roll by -N/2: rolled[i] = patch[(i + N/2) mod N]
  const int x = threadIdx.x % N, yb = threadIdx.x / N;
  const int a0 = roll_half<N>(z), a2 = roll_half<N>(x);
  const int r2zx = (a0 - N / 2) * (a0 - N / 2) + (a2 - N / 2) * (a2 - N / 2);
  float bgs = 0.f;
  auto deposit = [&](int y, float pv) {
    const int a1 = roll_half<N>(y);
    const float w = bg_weight(r2zx + (a1 - N / 2) * (a1 - N / 2), lut, lut_lo, lut_n);
    bgs += w * pv;
    float* dst = reinterpret_cast<float*>(buf + (y < NL ? y : y - NL) * S + x);
    dst[y >= NL ? 1 : 0] = (1.f - w) * pv;
  };
  if (sp.integral) {
    // integral shift: one buffer load per voxel, out-of-volume reads return 0
    const __amdgpu_buffer_rsrc_t rv = make_rsrc(v, (unsigned)(N * N * N) * 4u);
    const int zs = a0 + sp.si[0], xs = a2 + sp.si[2];
    const bool zxin = zs >= 0 && zs < N && xs >= 0 && xs < N;
    const int zxoff = zs * N * N + xs;
    constexpr int PER = N / YSTEP;
    constexpr int U = load_batch<PER>();
    for (int base = 0; base < PER; base += U) {
      float pv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int y = yb + (base + u) * YSTEP;
        const int ys = roll_half<N>(y) + sp.si[1];
        const bool ok = zxin && ys >= 0 && ys < N;
        pv[u] = load_f1(rv, ok ? (unsigned)(zxoff + ys * N) * 4u : kOutOfRange);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) deposit(yb + (base + u) * YSTEP, pv[u]);
    }
  } else {
    for (int y = yb; y < N; y += YSTEP) deposit(y, patch_value(v, N, sp, a0, roll_half<N>(y), a2));
  }
  bgs = block_sum<NT>(bgs, red);  // its barrier also publishes buf
  if (threadIdx.x == 0) bg_part[b * N + z] = bgs;
  mch::fft::forward<N, NL, NT, -1>(buf, S, 1, tw);
  // unpack A = (Z[k] + conj Z[N-k]) / 2 (row j), B = (Z[k] - conj Z[N-k]) / (2i) (row j + N/2)
  const int y = threadIdx.x % N;
  const float2* line = buf + (y < NL ? y : y - NL) * S;
  float2* out = T + (long long)b * H * N * N + (long long)z * N + y;
  for (int kx = threadIdx.x / N; kx < H; kx += YSTEP) {
    const float2 Zk = line[mch::fft::slot_of_freq<N>(kx)];
    const float2 Zn = line[mch::fft::slot_of_freq<N>(kx == 0 ? 0 : N - kx)];
    float2 o;
    if (y < NL) {
      o = make_float2(0.5f * (Zk.x + Zn.x), 0.5f * (Zk.y - Zn.y));
    } else {
      const float dx = Zk.x - Zn.x, dy = Zk.y + Zn.y;
      o = make_float2(0.5f * dy, -0.5f * dx);
    }
    out[(long long)kx * N * N] = o;
  }
}

// ---------------------------------------------------------------------------
// rotated reference sample
// ---------------------------------------------------------------------------
// Two tables of the centred reference, one 16-byte entry per (z', y', x'),
// 0 <= z', y' < N + 2, 0 <= x' <= N, for the point (z, y, x) = (z' - 1, y' - 1,
// x' - 1), with a zero border around the volume: a sample whose lower corner
// (z0, y0, x0) lies in [-1, N)^3 reads entry (z0 + 1, y0 + 1, x0 + 1) and its
// neighbours, all inside the table.
//   FP32 pairs  (ref[z][y][x], ref[z][y][x + 1]) as complex FP32 -- the
//               reference's complex64 values; a trilinear sample reads the four
//               (z, y) corners;
//   FP16 quads  rows y, y + 1 and columns x, x + 1 of plane z, scaled by 2^e,
//               as four FP16 pairs over x: (re, y), (im, y), (re, y + 1),
//               (im, y + 1) -- reduced precision; a sample reads two entries
//               (z0 and z0 + 1), half the gathers and half the bytes through
//               the texture path (the column pass: 1.60 -> 1.45 ms per batch of
//               512 at N = 96; a compact FP16 table of one 4-byte point per
//               entry, read as eight 4-byte loads, kept 2.4 GB more of the
//               gathers in L2 but took 1.65 ms).
// Entries in bricks of 2 x 2 (z', y'): the four entries of one brick and
// column fill one 64-byte segment, so that the gathers of samples that are
// close in z and y (the 4 x 4 x 4 blocks of frequencies a wave samples) share
// segments (the column pass took 1.77 ms per batch of 512 with rows of pairs,
// 1.61 with these bricks; 128-byte bricks of 2 x 2 x 2 gave no further gain).
__host__ __device__ constexpr unsigned table_index(int N, int zp, int yp, int xp) {
  return (((unsigned)((zp >> 1) * ((N + 2) >> 1) + (yp >> 1)) * (unsigned)(N + 1) + (unsigned)xp) << 2) |
         ((unsigned)(zp & 1) << 1) | (unsigned)(yp & 1);
}
// entries of either table (N even)
constexpr size_t table_entries(int N) { return (size_t)(N + 2) * (N + 2) * (N + 1); }
template <int N>
constexpr unsigned table_bytes() { return (unsigned)(table_entries(N) * 16u); }
// byte steps to the entry of z' + 1 / y' + 1 from (z', y'): the next entry of
// the brick for even z' / y', else the next brick along z / y
template <int N>
constexpr unsigned table_zbrick() { return (unsigned)((N + 2) / 2) * (unsigned)(N + 1) * 64u; }
template <int N>
constexpr unsigned table_ybrick() { return (unsigned)(N + 1) * 64u; }

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Work-spectrum elements: complex FP32, or (RED) complex BF16 in 4 bytes,
// rounded to nearest even on store and widened exactly on load.
template <bool RED>
using SpecElem = typename std::conditional<RED, unsigned, float2>::type;
template <bool RED>
__device__ __forceinline__ SpecElem<RED> spec_pack(float2 v) {
  if constexpr (RED) {
    const bf16x2 h = {(__bf16)v.x, (__bf16)v.y};
    return __builtin_bit_cast(unsigned, h);
  } else {
    return v;
  }
}
template <bool RED>
__device__ __forceinline__ float2 spec_unpack(SpecElem<RED> e) {
  if constexpr (RED)
    return make_float2(__uint_as_float(e << 16), __uint_as_float(e & 0xffff0000u));
  else
    return e;
}

// Byte offset of entry (z', y', x') from 24-bit products (every index of the
// table stays below 2^24).
template <int N>
__device__ __forceinline__ unsigned table_off24(unsigned zp, unsigned yp, unsigned xp) {
  const unsigned brick = __umul24(zp >> 1, (unsigned)((N + 2) >> 1)) + (yp >> 1);
  return ((__umul24(brick, (unsigned)(N + 1)) + xp) << 6) | ((zp & 1u) << 5) | ((yp & 1u) << 4);
}

// The corners of one trilinear sample of the centred reference spectrum,
// loaded (FP32: four pairs over x for the (z, y) corners; FP16: two quads for
// z0 and z0 + 1), and the fractional position.  A sample goes in three steps,
// so that a caller can keep the next sample's loads in flight while it
// combines the current one, and skip the geometry of samples it knows to be
// zero: ref_place (offset and fractions), ref_load (the loads, unconditional:
// corners outside the volume and samples outside the ball take the
// out-of-range offset and read as zero), ref_combine.
template <bool H16>
struct RefCorners {
  v4u q[H16 ? 2 : 4];
  float fx, fy, fz;
  unsigned o0, dz, dy;  // offset of the lower corner's entry (kOutOfRange: all read zero), steps
};

// Sample at the rotated position (rx, ry, rz) (x the last axis).  FP32 table:
// through the reference's grid_sample normalisation with size_e = N + 1
// (align_corners), i = ((r sc - 1) + 1) / 2 (N - 1), rounded as the reference
// rounds it.  FP16 table: that normalisation is the identity up to rounding
// and is skipped (i = r).
template <int N, bool H16>
__device__ __forceinline__ void ref_place(RefCorners<H16>& r, float rx, float ry, float rz, bool inball) {
  float ix = rx, iy = ry, iz = rz;
  if constexpr (!H16) {
    const float sc = 2.0f / (float)(N - 1);  // 2 / (size_e - 2) with size_e = N + 1
    const float um = (float)(N - 1);
    ix = ((rx * sc - 1.0f) + 1.0f) * 0.5f * um;
    iy = ((ry * sc - 1.0f) + 1.0f) * 0.5f * um;
    iz = ((rz * sc - 1.0f) + 1.0f) * 0.5f * um;
  }
  const float fz0 = floorf(iz), fy0 = floorf(iy), fx0 = floorf(ix);
  const int z0 = (int)fz0, y0 = (int)fy0, x0 = (int)fx0;
  r.fz = iz - fz0;
  r.fy = iy - fy0;
  r.fx = ix - fx0;
  // one range check for all corners: the lower corner in [-1, N)^3 (the
  // border entries read as zero); the other corners are steps from it
  const bool ok = inball && (unsigned)(z0 + 1) <= (unsigned)N && (unsigned)(y0 + 1) <= (unsigned)N &&
                  (unsigned)(x0 + 1) <= (unsigned)N;
  r.o0 = ok ? table_off24<N>((unsigned)(z0 + 1), (unsigned)(y0 + 1), (unsigned)(x0 + 1)) : kOutOfRange;
  r.dz = ((z0 + 1) & 1) ? table_zbrick<N>() - 32u : 32u;
  r.dy = ((y0 + 1) & 1) ? table_ybrick<N>() - 16u : 16u;
}

template <bool H16>
__device__ __forceinline__ void ref_load(RefCorners<H16>& r, __amdgpu_buffer_rsrc_t rref) {
  const bool ok = r.o0 != kOutOfRange;
  if constexpr (H16) {
    r.q[0] = __builtin_amdgcn_raw_buffer_load_b128(rref, r.o0, 0, 0);
    r.q[1] = __builtin_amdgcn_raw_buffer_load_b128(rref, ok ? r.o0 + r.dz : kOutOfRange, 0, 0);
  } else {
    r.q[0] = __builtin_amdgcn_raw_buffer_load_b128(rref, r.o0, 0, 0);
    r.q[1] = __builtin_amdgcn_raw_buffer_load_b128(rref, ok ? r.o0 + r.dy : kOutOfRange, 0, 0);
    r.q[2] = __builtin_amdgcn_raw_buffer_load_b128(rref, ok ? r.o0 + r.dz : kOutOfRange, 0, 0);
    r.q[3] = __builtin_amdgcn_raw_buffer_load_b128(rref, ok ? r.o0 + r.dz + r.dy : kOutOfRange, 0, 0);
  }
}

// The trilinear value from loaded corners.  FP32: the reference's order of
// operations (x, then the four (z, y) weights).  FP16: the eight weights
// rounded to FP16 in x pairs, each (z, y) corner pair through one FP16 dot
// product accumulated in FP32; the result is still scaled by the table's 2^e.
template <bool H16>
__device__ __forceinline__ float2 ref_combine(const RefCorners<H16>& r) {
  const float fx = r.fx, fy = r.fy, fz = r.fz;
  if constexpr (H16) {
    const float a00 = (1.f - fz) * (1.f - fy), a01 = (1.f - fz) * fy, a10 = fz * (1.f - fy), a11 = fz * fy;
    const float wx0 = 1.f - fx, wx1 = fx;
    const half2v h00 = {(_Float16)(a00 * wx0), (_Float16)(a00 * wx1)};
    const half2v h01 = {(_Float16)(a01 * wx0), (_Float16)(a01 * wx1)};
    const half2v h10 = {(_Float16)(a10 * wx0), (_Float16)(a10 * wx1)};
    const half2v h11 = {(_Float16)(a11 * wx0), (_Float16)(a11 * wx1)};
    auto h2 = [](unsigned u) { return __builtin_bit_cast(half2v, u); };
    float re = __builtin_amdgcn_fdot2(h00, h2(r.q[0].x), 0.f, false);
    float im = __builtin_amdgcn_fdot2(h00, h2(r.q[0].y), 0.f, false);
    re = __builtin_amdgcn_fdot2(h01, h2(r.q[0].z), re, false);
    im = __builtin_amdgcn_fdot2(h01, h2(r.q[0].w), im, false);
    re = __builtin_amdgcn_fdot2(h10, h2(r.q[1].x), re, false);
    im = __builtin_amdgcn_fdot2(h10, h2(r.q[1].y), im, false);
    re = __builtin_amdgcn_fdot2(h11, h2(r.q[1].z), re, false);
    im = __builtin_amdgcn_fdot2(h11, h2(r.q[1].w), im, false);
    return make_float2(re, im);
  } else {
    float4 q[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      q[i] = make_float4(__uint_as_float(r.q[i].x), __uint_as_float(r.q[i].y), __uint_as_float(r.q[i].z),
                         __uint_as_float(r.q[i].w));
    const float s00 = (1.f - fz) * (1.f - fy), s01 = (1.f - fz) * fy, s10 = fz * (1.f - fy), s11 = fz * fy;
    float2 rv;
    rv.x = s00 * fmaf(fx, q[0].z - q[0].x, q[0].x) + s01 * fmaf(fx, q[1].z - q[1].x, q[1].x) +
           s10 * fmaf(fx, q[2].z - q[2].x, q[2].x) + s11 * fmaf(fx, q[3].z - q[3].x, q[3].x);
    rv.y = s00 * fmaf(fx, q[0].w - q[0].y, q[0].y) + s01 * fmaf(fx, q[1].w - q[1].y, q[1].y) +
           s10 * fmaf(fx, q[2].w - q[2].y, q[2].y) + s11 * fmaf(fx, q[3].w - q[3].y, q[3].y);
    return rv;
  }
}

// One sample, fetched and combined, in true units (the FP16 table's 2^e removed).
template <int N, bool H16>
__device__ __forceinline__ float2 ref_sample(__amdgpu_buffer_rsrc_t rref, float unscale, float rx, float ry,
                                             float rz, bool inball) {
  RefCorners<H16> r;
  ref_place<N, H16>(r, rx, ry, rz, inball);
  ref_load<H16>(r, rref);
  float2 v = ref_combine<H16>(r);
  if constexpr (H16) {
    v.x *= unscale;
    v.y *= unscale;
  }
  return v;
}

// ---------------------------------------------------------------------------
// one (b, kx) plane: FFT y, z; cross power with the rotated reference
// ---------------------------------------------------------------------------
// Leaves P = conj(R) S of plane kx of particle b in LDS, slot (sz, sy) holding
// frequency (freq(sz), freq(sy)), the plane's shares of ||S||^2 and ||R||^2 in
// snorm_part / rnorm_part, and ends on a barrier.  `red` holds 2 * NT / 64 floats.
template <int N, bool H16>
__device__ __forceinline__ void plane_cross_power(
    float2* __restrict__ plane, float2* __restrict__ tw, float* __restrict__ red, int kx, int b,
    const float2* __restrict__ T, const float2* __restrict__ XM, const float* __restrict__ bg_part, float wsum,
    const void* __restrict__ reftab, float unscale, const float* __restrict__ rot, const float* __restrict__ ctf,
    const float2* __restrict__ twg, float* __restrict__ rnorm_part, float* __restrict__ snorm_part) {
  constexpr int NT = plane_threads<N>(), S = N + 1, H = N / 2 + 1, ZSTEP = NT / N;
  for (int k = threadIdx.x; k < N; k += NT) tw[k] = twg[k];
  // background average of particle b; every wave sums the partials in the same order
  float bgs = 0.f;
  for (int zz = threadIdx.x & 63; zz < N; zz += 64) bgs += bg_part[b * N + zz];
  bgs = mch::wave_sum(bgs);
  const float bg = wsum > 0.f ? bgs / wsum : 0.f;
  const float2* P = T + ((long long)b * H + kx) * N * N;
  const float2* M = XM + (long long)kx * N * N;
  // thread column y (fixed), rows z = zb, zb + ZSTEP, ...
  const int cy = threadIdx.x % N, zb = threadIdx.x / N;
  {
    constexpr int PER = N / ZSTEP;
    constexpr int U = load_batch<PER>();
    for (int base = 0; base < PER; base += U) {
      float2 a[U], mm[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int zz = zb + (base + u) * ZSTEP;
        a[u] = P[zz * N + cy];
        mm[u] = M[zz * N + cy];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        plane[(zb + (base + u) * ZSTEP) * S + cy] = make_float2(a[u].x + bg * mm[u].x, a[u].y + bg * mm[u].y);
    }
  }
  __syncthreads();
  mch::fft::forward<N, N, NT, -1>(plane, S, 1, tw);  // along y: lines are rows z
  mch::fft::forward<N, N, NT, -1>(plane, 1, S, tw);  // along z: lines are columns y
  // slot (sz, sy) now holds frequency (freq(sz), freq(sy)); cross power in place.
  // The thread keeps frequency ky = cy (natural order across lanes, so that
  // neighbouring lanes sample neighbouring points of the reference) in slot
  // column sy = slot(ky): everything that depends on ky is hoisted.
  const float* m = rot + b * 9;
  const float m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3], m4 = m[4], m5 = m[5], m6 = m[6], m7 = m[7],
              m8 = m[8];
  const float c = (float)N * 0.5f;
  const float inv = 1.0f / ((float)N * (float)N * (float)N);
  const int jp = kx < N / 2 ? kx : kx - H;
  const float bx = (float)(kx + N / 2) - c;
  const int ky = cy, sy = mch::fft::slot_of_freq<N>(cy);
  const int ip = ky < N / 2 ? ky : ky - N;
  const int r2xy = ip * ip + jp * jp;
  const float by = (float)roll_half<N>(ky) - c;
  const float px = m0 * bx + m1 * by + c, py = m3 * bx + m4 * by + c, pz = m6 * bx + m7 * by + c;
  const __amdgpu_buffer_rsrc_t rref = make_rsrc(reftab, table_bytes<N>());
  const float* ctfb = ctf ? ctf + (long long)b * N * N * H : nullptr;
  float ns = 0.f, nr = 0.f;
#pragma unroll 2
  for (int i = 0; i < N / ZSTEP; ++i) {
    const int sz = zb + i * ZSTEP;
    const int kz = mch::fft::freq_of_slot<N>(sz);
    const int kp = kz < N / 2 ? kz : kz - N;
    const bool inball = kp * kp + r2xy <= (N / 2) * (N / 2);
    const float bz = (float)roll_half<N>(kz) - c;
    float2 rv = ref_sample<N, H16>(rref, unscale, fmaf(m2, bz, px), fmaf(m5, bz, py), fmaf(m8, bz, pz), inball);
    float cv = 1.f;
    if (ctfb && inball) cv = ctfb[(kz * N + ky) * H + kx];
    rv.x *= cv;
    rv.y *= cv;
    float2* slot = plane + sz * S + sy;
    float2 s = *slot;
    s.x *= inv;
    s.y *= inv;
    ns += s.x * s.x + s.y * s.y;
    nr += rv.x * rv.x + rv.y * rv.y;
    // conj(R) * S
    *slot = make_float2(rv.x * s.x + rv.y * s.y, rv.x * s.y - rv.y * s.x);
  }
  ns = mch::wave_sum(ns);
  nr = mch::wave_sum(nr);
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = ns;
    red[NT / 64 + (threadIdx.x >> 6)] = nr;
  }
  __syncthreads();  // also publishes the cross power to every thread
  if (threadIdx.x == 0) {
    float a = 0.f, bsum = 0.f;
    for (int w = 0; w < NT / 64; ++w) { a += red[w]; bsum += red[NT / 64 + w]; }
    snorm_part[b * H + kx] = a;
    rnorm_part[b * H + kx] = bsum;
  }
}

// ---------------------------------------------------------------------------
// integer lags, separable passes: (x, y) per slice -> z per column block ->
// (y, x) per slice
// ---------------------------------------------------------------------------
// The forward pass transforms each (b, z) slice along x and y and stores
// T[b][kx][z][ky] (ky in natural order); the column pass adds bg * XMy (the
// (x, y) spectrum of the background weight), transforms along z, forms the
// cross power with the rotated reference, transforms back along z and stores
// in place; the inverse pass transforms each slice back along y and x and
// takes the argmax.  A column block of KB ky columns needs N (KB + 1) complex
// values of LDS, so several blocks share a CU, where a whole (y, z) plane needs
// N (N + 1).
// threads of the (x, y) slice passes: 4 per slice column, so that a CU holds
// enough waves besides the slice's LDS
template <int N> constexpr int slice_threads() { return 4 * N; }
constexpr int kZThreads = 256, kZKx = 4, kZKy = 8;  // column block: 4 kx x 8 ky
template <int N> constexpr int zgroups() { return (N / 2 + 1 + kZKx - 1) / kZKx; }
// norm partials per particle on the integer path: one per column block
template <int N> constexpr int zparts() { return zgroups<N>() * (N / kZKy); }

// Forward transform along x and y of one (b, z) slice.  Loads and the mask
// split as shift_fwd_x_kernel; the x spectra are unpacked in registers,
// re-laid out as lines kx, transformed along y and stored with ky natural.
template <int N, bool RED>
__global__ __launch_bounds__(slice_threads<N>()) void shift_fwd_xy_kernel(
    const float* __restrict__ vols, long long ldv, const float* __restrict__ shift_zyx,
    const float* __restrict__ lutg, int lut_lo, int lut_n, const float2* __restrict__ twg,
    SpecElem<RED>* __restrict__ T, float* __restrict__ bg_part) {
  constexpr int NT = slice_threads<N>(), NL = N / 2, S = N + 1, H = N / 2 + 1, YSTEP = NT / N;
  constexpr int KXP = (H + YSTEP - 1) / YSTEP;
  static_assert(NT % N == 0 && N % YSTEP == 0, "column split");
  extern __shared__ float4 smem4[];
  float2* buf = reinterpret_cast<float2*>(smem4);  // [line][x], then [kx][y]; stride S
  float2* tw = buf + H * S;
  float* lut = reinterpret_cast<float*>(tw + N);
  __shared__ float red[NT / 64];
  const int z = blockIdx.x, b = blockIdx.y;
  for (int k = threadIdx.x; k < N; k += NT) tw[k] = twg[k];
  for (int k = threadIdx.x; k < lut_n; k += NT) lut[k] = lutg[k];
  __syncthreads();
  ShiftSpec sp;
  make_shift_spec(shift_zyx, b, N, sp);
  const float* v = vols + (long long)b * ldv;
  // roll by -N/2: rolled[i] = patch[(i + N/2) mod N]
  const int x = threadIdx.x % N, yb = threadIdx.x / N;
  const int a0 = roll_half<N>(z), a2 = roll_half<N>(x);
  const int r2zx = (a0 - N / 2) * (a0 - N / 2) + (a2 - N / 2) * (a2 - N / 2);
  float bgs = 0.f;
  auto deposit = [&](int y, float pv) {
    const int a1 = roll_half<N>(y);
    const float w = bg_weight(r2zx + (a1 - N / 2) * (a1 - N / 2), lut, lut_lo, lut_n);
    bgs += w * pv;
    float* dst = reinterpret_cast<float*>(buf + (y < NL ? y : y - NL) * S + x);
    dst[y >= NL ? 1 : 0] = (1.f - w) * pv;
  };
  if (sp.integral) {
    const __amdgpu_buffer_rsrc_t rv = make_rsrc(v, (unsigned)(N * N * N) * 4u);
    const int zs = a0 + sp.si[0], xs = a2 + sp.si[2];
    const bool zxin = zs >= 0 && zs < N && xs >= 0 && xs < N;
    const int zxoff = zs * N * N + xs;
    constexpr int PER = N / YSTEP;
    // the whole share of the slice's loads in flight at once (one HBM round trip)
    constexpr int U = PER <= 32 ? PER : load_batch<PER>();
    for (int base = 0; base < PER; base += U) {
      float pv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int y = yb + (base + u) * YSTEP;
        const int ys = roll_half<N>(y) + sp.si[1];
        const bool ok = zxin && ys >= 0 && ys < N;
        pv[u] = load_f1(rv, ok ? (unsigned)(zxoff + ys * N) * 4u : kOutOfRange);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) deposit(yb + (base + u) * YSTEP, pv[u]);
    }
  } else {
    for (int y = yb; y < N; y += YSTEP) deposit(y, patch_value(v, N, sp, a0, roll_half<N>(y), a2));
  }
  bgs = block_sum<NT>(bgs, red);  // its barrier also publishes buf
  if (threadIdx.x == 0) bg_part[b * N + z] = bgs;
  mch::fft::forward<N, NL, NT, -1>(buf, S, 1, tw);
  // unpack A = (Z[k] + conj Z[N-k]) / 2 (row j), B = (Z[k] - conj Z[N-k]) / (2i) (row j + N/2)
  const int y = threadIdx.x % N, kx0 = threadIdx.x / N;
  float2 o[KXP];
  {
    const float2* line = buf + (y < NL ? y : y - NL) * S;
#pragma unroll
    for (int i = 0; i < KXP; ++i) {
      const int kx = kx0 + i * YSTEP;
      if (kx < H) {
        const float2 Zk = line[mch::fft::slot_of_freq<N>(kx)];
        const float2 Zn = line[mch::fft::slot_of_freq<N>(kx == 0 ? 0 : N - kx)];
        if (y < NL) {
          o[i] = make_float2(0.5f * (Zk.x + Zn.x), 0.5f * (Zk.y - Zn.y));
        } else {
          const float dx = Zk.x - Zn.x, dy = Zk.y + Zn.y;
          o[i] = make_float2(0.5f * dy, -0.5f * dx);
        }
      }
    }
  }
  __syncthreads();  // every read of the x spectra is done
#pragma unroll
  for (int i = 0; i < KXP; ++i) {
    const int kx = kx0 + i * YSTEP;
    if (kx < H) buf[kx * S + y] = o[i];
  }
  __syncthreads();
  mch::fft::forward<N, H, NT, -1>(buf, S, 1, tw);  // along y: lines are kx
  // stored two adjacent ky per lane (half the store requests)
  using Pair = typename std::conditional<RED, uint2, float4>::type;
  SpecElem<RED>* out = T + (long long)b * H * N * N + (long long)z * N;
  constexpr int PSTEP = NT / (N / 2);
  const int j2 = 2 * (threadIdx.x % (N / 2)), kxp = threadIdx.x / (N / 2);
  const int s0 = mch::fft::slot_of_freq<N>(j2), s1 = mch::fft::slot_of_freq<N>(j2 + 1);
  for (int kx = kxp; kx < H; kx += PSTEP) {
    const float2 v0 = buf[kx * S + s0], v1 = buf[kx * S + s1];
    Pair o;
    if constexpr (RED)
      o = make_uint2(spec_pack<true>(v0), spec_pack<true>(v1));
    else
      o = make_float4(v0.x, v0.y, v1.x, v1.y);
    *reinterpret_cast<Pair*>(out + (long long)kx * N * N + j2) = o;
  }
}

// One workgroup per (b, group of 4 kx, block of 8 ky): T + bg * XMy, FFT along z
// (1/N^3), ||S||^2, the rotated reference sampled on the fly (trilinear,
// ifftshift, mask_full, ctf), ||R||^2, P = conj(R) S, inverse FFT along z,
// back to T in place.  The cross power runs over wave tiles of 4 kx x 4 ky x
// 4 kz, whose samples fall in a compact region of the reference (lanes share
// its cache lines).  Norm partials per workgroup.
template <int N, bool H16>
__global__ __launch_bounds__(kZThreads) void shift_z_kernel(
    SpecElem<H16>* __restrict__ T, const float2* __restrict__ XMy, const float* __restrict__ bg_part, float wsum,
    const void* __restrict__ reftab, float unscale, const float* __restrict__ rot, const float* __restrict__ ctf,
    const float2* __restrict__ twg, float* __restrict__ rnorm_part, float* __restrict__ snorm_part) {
  constexpr int NT = kZThreads, NC = kZKx * kZKy, SC = NC + 1, H = N / 2 + 1, NG = zgroups<N>(),
                NKB = N / kZKy, ZSTEP = NT / NC;
  static_assert(NT == 256 && kZKx == 4 && kZKy == 8, "wave tiles of 4 x 4 x 4 in a 4 x 8 column block");
  static_assert(N % kZKy == 0 && N % ZSTEP == 0 && N % 4 == 0, "column block split");
  extern __shared__ float4 smem4[];
  float2* tile = reinterpret_cast<float2*>(smem4);  // [z][column], stride SC; column = 8 (kx - kx0) + (ky - ky0)
  float2* tw = tile + N * SC;
  __shared__ float red[2 * NT / 64];
  // XCD-aware order: neighbouring column blocks of one particle, which sample
  // neighbouring regions of the reference, run on one XCD
  const int lb = mch::xcd_remap(blockIdx.x, gridDim.x);
  const int kb = lb % NKB, g = (lb / NKB) % NG, b = lb / (NKB * NG);
  const int kx0 = g * kZKx, ky0 = kb * kZKy;
  for (int k = threadIdx.x; k < N; k += NT) tw[k] = twg[k];
  float bgs = 0.f;
  for (int zz = threadIdx.x & 63; zz < N; zz += 64) bgs += bg_part[b * N + zz];
  bgs = mch::wave_sum(bgs);
  const float bg = wsum > 0.f ? bgs / wsum : 0.f;
  // loads and stores of T and XMy move two adjacent ky columns per lane (8
  // bytes of BF16 or 16 of FP32 T): half the requests of one column per lane
  constexpr int NP = NC / 2, PSTEP = NT / NP;  // 16 column pairs, rows z = zp + PSTEP k
  static_assert(N % PSTEP == 0 && kZKy % 2 == 0, "paired column split");
  using Pair = typename std::conditional<H16, uint2, float4>::type;
  const int pc = threadIdx.x % NP, zp = threadIdx.x / NP;
  const int pkx = kx0 + pc / (kZKy / 2), pky = ky0 + 2 * (pc % (kZKy / 2));
  const bool pok = pkx < H;  // the last group of kx can be partial
  const int lc0 = (pc / (kZKy / 2)) * kZKy + 2 * (pc % (kZKy / 2));  // tile column of the pair's first
  SpecElem<H16>* col = T + ((long long)b * H + (pok ? pkx : 0)) * N * N + pky;
  const float2* mcol = XMy + (long long)(pok ? pkx : 0) * N * N + pky;
  {
    constexpr int PER = N / PSTEP;
    Pair a[PER];
    float4 mm[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int zz = zp + u * PSTEP;
      if (pok) {
        a[u] = *reinterpret_cast<const Pair*>(col + zz * N);
        mm[u] = *reinterpret_cast<const float4*>(mcol + zz * N);
      } else {
        a[u] = Pair{};
        mm[u] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      float2 a0, a1;
      if constexpr (H16) {
        a0 = spec_unpack<true>(a[u].x);
        a1 = spec_unpack<true>(a[u].y);
      } else {
        a0 = make_float2(a[u].x, a[u].y);
        a1 = make_float2(a[u].z, a[u].w);
      }
      float2* t = tile + (zp + u * PSTEP) * SC + lc0;
      t[0] = make_float2(a0.x + bg * mm[u].x, a0.y + bg * mm[u].y);
      t[1] = make_float2(a1.x + bg * mm[u].z, a1.y + bg * mm[u].w);
    }
  }
  __syncthreads();
  mch::fft::forward<N, NC, NT, -1>(tile, 1, SC, tw);  // along z: lines are columns
  // slot sz of a column holds frequency kz = freq(sz)
  const float* m = rot + b * 9;
  const float m0 = m[0], m1 = m[1], m2 = m[2], m3 = m[3], m4 = m[4], m5 = m[5], m6 = m[6], m7 = m[7],
              m8 = m[8];
  const float c = (float)N * 0.5f;
  const float inv = 1.0f / ((float)N * (float)N * (float)N);
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int tx = l & 3, ty = (l >> 2) & 3, tz = l >> 4;
  const int kx = kx0 + tx;
  const bool kxok = kx < H;
  const int jp = kx < N / 2 ? kx : kx - H;
  const float bx = (float)(kx + N / 2) - c;
  const __amdgpu_buffer_rsrc_t rref = make_rsrc(reftab, table_bytes<N>());
  const float* ctfb = ctf ? ctf + (long long)b * N * N * H : nullptr;
  float ns = 0.f, nr = 0.f;
  constexpr int TILES = 2 * (N / 4);  // (half of the ky block, group of 4 kz)
  constexpr int WAVES = NT / 64;
  static_assert(TILES % WAVES == 0, "every wave runs the same number of tiles");
  // FP16 table: the products and ||R||^2 are taken on the scaled samples; the
  // scales come back once, 1/N^3 and 2^-e on the cross power, 2^-2e on ||R||^2
  const float pscale = H16 ? inv * unscale : 1.f;
  struct Tile {
    RefCorners<H16> rc;
    float cv;
    int ly, kz;
    bool any;  // some lane of the wave samples inside the ball (wave-uniform)
  };
  // the CTF factor is loaded before the corners, so that waiting for a tile's
  // loads never waits for the next tile's
  auto fetch = [&](int t) {
    Tile T;
    T.ly = 4 * (t & 1) + ty;
    T.kz = 4 * (t >> 1) + tz;
    const int ky = ky0 + T.ly;
    const int ip = ky < N / 2 ? ky : ky - N;
    const int kp = T.kz < N / 2 ? T.kz : T.kz - N;
    const bool inball = kxok && kp * kp + ip * ip + jp * jp <= (N / 2) * (N / 2);
    // tiles wholly outside the ball (38% of them at N = 96) sample nothing
    T.any = __builtin_amdgcn_ballot_w64(inball) != 0;
    T.cv = 1.f;
    T.rc.o0 = kOutOfRange;
    if (T.any) {
      if (ctfb && inball) T.cv = ctfb[(T.kz * N + ky) * H + kx];
      const float by = (float)roll_half<N>(ky) - c, bz = (float)roll_half<N>(T.kz) - c;
      const float px = m0 * bx + m1 * by + c, py = m3 * bx + m4 * by + c, pz = m6 * bx + m7 * by + c;
      ref_place<N, H16>(T.rc, fmaf(m2, bz, px), fmaf(m5, bz, py), fmaf(m8, bz, pz), inball);
    }
    ref_load<H16>(T.rc, rref);  // outside the branch: the loads stay in flight across it
    return T;
  };
  auto combine = [&](const Tile& T) {
    if (!T.any) {  // R = 0: the cross power is zero, ||S||^2 still counts
      if (kxok) {
        float2* slot = tile + mch::fft::slot_of_freq<N>(T.kz) * SC + tx * kZKy + T.ly;
        float2 s = *slot;
        if constexpr (!H16) {
          s.x *= inv;
          s.y *= inv;
        }
        ns += s.x * s.x + s.y * s.y;
        *slot = make_float2(0.f, 0.f);
      }
      return;
    }
    float2 rv = ref_combine<H16>(T.rc);
    if (ctfb) {
      rv.x *= T.cv;
      rv.y *= T.cv;
    }
    if (kxok) {
      float2* slot = tile + mch::fft::slot_of_freq<N>(T.kz) * SC + tx * kZKy + T.ly;
      float2 s = *slot;
      if constexpr (H16) {
        ns += s.x * s.x + s.y * s.y;
        nr += rv.x * rv.x + rv.y * rv.y;
        // conj(R) * S
        *slot = make_float2((rv.x * s.x + rv.y * s.y) * pscale, (rv.x * s.y - rv.y * s.x) * pscale);
      } else {
        s.x *= inv;
        s.y *= inv;
        ns += s.x * s.x + s.y * s.y;
        nr += rv.x * rv.x + rv.y * rv.y;
        // conj(R) * S
        *slot = make_float2(rv.x * s.x + rv.y * s.y, rv.x * s.y - rv.y * s.x);
      }
    }
  };
  // two tiles per iteration from two buffers: one tile's loads are in flight
  // while the other is combined, and no in-flight registers are copied (the
  // last fetch repeats a tile and is not used)
  static_assert((TILES / WAVES) % 2 == 0, "tiles per wave in pairs");
  Tile ta = fetch(w);
  for (int t = w; t < TILES; t += 2 * WAVES) {
    const Tile tb = fetch(t + WAVES);
    combine(ta);
    ta = fetch(t + 2 * WAVES < TILES ? t + 2 * WAVES : t);
    combine(tb);
  }
  if constexpr (H16) {
    ns *= inv * inv;
    nr *= unscale * unscale;
  }
  ns = mch::wave_sum(ns);
  nr = mch::wave_sum(nr);
  if (l == 0) {
    red[w] = ns;
    red[NT / 64 + w] = nr;
  }
  __syncthreads();  // also publishes the cross power to every thread
  if (threadIdx.x == 0) {
    float a = 0.f, bsum = 0.f;
    for (int q = 0; q < NT / 64; ++q) { a += red[q]; bsum += red[NT / 64 + q]; }
    snorm_part[((long long)b * NG + g) * NKB + kb] = a;
    rnorm_part[((long long)b * NG + g) * NKB + kb] = bsum;
  }
  mch::fft::inverse<N, NC, NT, 1>(tile, 1, SC, tw);  // along z, natural order out
  if (pok)
    for (int zz = zp; zz < N; zz += PSTEP) {
      const float2 v0 = tile[zz * SC + lc0], v1 = tile[zz * SC + lc0 + 1];
      Pair o;
      if constexpr (H16)
        o = make_uint2(spec_pack<true>(v0), spec_pack<true>(v1));
      else
        o = make_float4(v0.x, v0.y, v1.x, v1.y);
      *reinterpret_cast<Pair*>(col + zz * N) = o;
    }
}

// Inverse transform of one (b, z) slice along y, then along x of the
// zero-padded half axis (real part), and the slice's first argmax.
// For a row X (kx = 0..N/2, zero above), Re(sum_{k<=N/2} X_k e^{2 pi i k n/N}) =
// y[n] / 2 with y the inverse DFT of the Hermitian row Y (Y_0 = 2 Re X_0,
// Y_{N/2} = 2 Re X_{N/2}, Y_k = X_k, Y_{N-k} = conj X_k).  Two such real rows
// go through one complex FFT as Y1 + i Y2: rows j and j + N/2 of the slice.
template <int N, bool RED>
__global__ __launch_bounds__(slice_threads<N>()) void shift_inv_xy_kernel(
    const SpecElem<RED>* __restrict__ T, const float2* __restrict__ twg, float* __restrict__ amax_val,
    int* __restrict__ amax_idx) {
  constexpr int NT = slice_threads<N>(), NL = N / 2, S = N + 1, H = N / 2 + 1, YSTEP = NT / N;
  constexpr int KXP = (H + YSTEP - 1) / YSTEP, ITEMS = NL * H, IP = (ITEMS + NT - 1) / NT;
  static_assert(NT % 64 == 0 && NT % N == 0, "whole waves, column split");
  extern __shared__ float4 smem4[];
  float2* buf = reinterpret_cast<float2*>(smem4);  // [kx][y], then [line][x]; stride S
  float2* tw = buf + H * S;
  __shared__ float rv[NT / 64];
  __shared__ int ri[NT / 64];
  const int z = blockIdx.x, b = blockIdx.y;
  for (int k = threadIdx.x; k < N; k += NT) tw[k] = twg[k];
  const SpecElem<RED>* in = T + (long long)b * H * N * N + (long long)z * N;
  {
    // rows kx of the slice, ky natural -> slot order for the inverse along y;
    // two adjacent ky per lane (half the load requests)
    using Pair = typename std::conditional<RED, uint2, float4>::type;
    constexpr int PSTEP = NT / (N / 2), KP = (H + PSTEP - 1) / PSTEP;
    const int j2 = 2 * (threadIdx.x % (N / 2)), kxp = threadIdx.x / (N / 2);
    const int s0 = mch::fft::slot_of_freq<N>(j2), s1 = mch::fft::slot_of_freq<N>(j2 + 1);
    Pair v[KP];
#pragma unroll
    for (int i = 0; i < KP; ++i) {
      const int kx = kxp + i * PSTEP;
      if (kx < H) v[i] = *reinterpret_cast<const Pair*>(in + (long long)kx * N * N + j2);
    }
#pragma unroll
    for (int i = 0; i < KP; ++i) {
      const int kx = kxp + i * PSTEP;
      if (kx < H) {
        if constexpr (RED) {
          buf[kx * S + s0] = spec_unpack<true>(v[i].x);
          buf[kx * S + s1] = spec_unpack<true>(v[i].y);
        } else {
          buf[kx * S + s0] = make_float2(v[i].x, v[i].y);
          buf[kx * S + s1] = make_float2(v[i].z, v[i].w);
        }
      }
    }
  }
  __syncthreads();
  mch::fft::inverse<N, H, NT, 1>(buf, S, 1, tw);  // along y: natural y out
  // lines j: Y1 + i Y2 of rows j and j + N/2, built in registers, then stored
  float2 ek[IP], en[IP];
#pragma unroll
  for (int i = 0; i < IP; ++i) {
    const int t = i * NT + threadIdx.x;
    if (t < ITEMS) {
      const int j = t % NL, k = t / NL;
      const float2 x1 = buf[k * S + j], x2 = buf[k * S + j + NL];
      if (k == 0 || k == N / 2) {
        ek[i] = make_float2(2.f * x1.x, 2.f * x2.x);
      } else {
        ek[i] = make_float2(x1.x - x2.y, x1.y + x2.x);
        en[i] = make_float2(x1.x + x2.y, x2.x - x1.y);
      }
    }
  }
  __syncthreads();  // every read of the y transforms is done
#pragma unroll
  for (int i = 0; i < IP; ++i) {
    const int t = i * NT + threadIdx.x;
    if (t < ITEMS) {
      const int j = t % NL, k = t / NL;
      float2* line = buf + j * S;
      line[k] = ek[i];
      if (k != 0 && k != N / 2) line[N - k] = en[i];
    }
  }
  __syncthreads();
  mch::fft::forward<N, NL, NT, 1>(buf, S, 1, tw);
  // NT is a multiple of N: a thread keeps slot s (x = freq_of_slot(s)) and
  // visits rows j = j0, j0 + NT / N, ...; the flat ids of each of its two
  // sequences (rows j and rows j + N/2) rise, so a strict comparison keeps
  // the first maximum of each, and the two merge by (value, id)
  const int sx = threadIdx.x % N, j0 = threadIdx.x / N;
  const int xx = mch::fft::freq_of_slot<N>(sx);
  float b1 = -INFINITY, b2 = -INFINITY;
  int i1 = 0x7fffffff, i2 = 0x7fffffff;
  for (int j = j0; j < NL; j += NT / N) {
    const float2 zv = buf[j * S + sx];
    const float v1 = 0.5f * zv.x, v2 = 0.5f * zv.y;
    const int f1 = (z * N + j) * N + xx;
    if (v1 > b1) { b1 = v1; i1 = f1; }
    if (v2 > b2) { b2 = v2; i2 = f1 + NL * N; }
  }
  float bv = b1;
  int bi = i1;
  if (b2 > bv || (b2 == bv && i2 < bi)) { bv = b2; bi = i2; }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(bv, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
  }
  if ((threadIdx.x & 63) == 0) { rv[threadIdx.x >> 6] = bv; ri[threadIdx.x >> 6] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    float v = rv[0];
    int i = ri[0];
    for (int w = 1; w < NT / 64; ++w)
      if (rv[w] > v || (rv[w] == v && ri[w] < i)) { v = rv[w]; i = ri[w]; }
    amax_val[(long long)b * N + z] = v;
    amax_idx[(long long)b * N + z] = i;
  }
}

// ---------------------------------------------------------------------------
// sub-voxel lags (shift_upsample_factor > 1): local matrix DFT of the cross power
// ---------------------------------------------------------------------------
// The reference (compute_shift, mode "local") zeroes the coarse shift and
// evaluates Re sum_k P(k) prod_d e^{2 pi i (u_d - R/2) f_d(k_d)} on an R^3 grid
// (R = ceil(up (2 r + 1)), f_d = fftfreq(n_d, up), n = N, N, N/2+1).  Here the
// plane kernel contracts y and z of each kx plane to R x R (Q), and
// shift_up_kx_kernel contracts kx and takes the first argmax.
// t01: (R, N) phases of the y / z axes in natural frequency order.
template <int N, bool H16>
__global__ __launch_bounds__(plane_threads<N>()) void shift_yz_up_kernel(
    const float2* __restrict__ T, const float2* __restrict__ XM, const float* __restrict__ bg_part,
    float wsum, const void* __restrict__ reftab, float unscale, const float* __restrict__ rot,
    const float* __restrict__ ctf, const float2* __restrict__ twg, const float2* __restrict__ t01, int R,
    float2* __restrict__ Qg, float* __restrict__ rnorm_part, float* __restrict__ snorm_part) {
  constexpr int NT = plane_threads<N>(), S = N + 1, H = N / 2 + 1;
  static_assert(NT % 64 == 0 && NT % N == 0, "whole waves, column split");
  extern __shared__ float4 smem4[];
  float2* plane = reinterpret_cast<float2*>(smem4);  // [z][y], stride S
  float2* tw = plane + N * S;
  float2* ts = tw + N;                                // (R, N): phases in slot order
  __shared__ float red[2 * NT / 64];
  const int lb = mch::xcd_remap(blockIdx.x, gridDim.x);
  const int kx = lb % H, b = lb / H;
  for (int q = threadIdx.x; q < R * N; q += NT) {
    const int u = q / N, k = q - (q / N) * N;
    ts[u * N + mch::fft::slot_of_freq<N>(k)] = t01[q];
  }
  plane_cross_power<N, H16>(plane, tw, red, kx, b, T, XM, bg_part, wsum, reftab, unscale, rot, ctf, twg,
                            rnorm_part, snorm_part);
  // Y[sz][u1] = sum_sy P[sz][sy] e[u1][sy], 4 x 4 register blocks
  const int nbz = N / 4, nbu = (R + 3) / 4;
  constexpr int MAXB = 2;  // blocks per thread (N / 4 * ceil(R / 4) <= 2 NT for R <= 2 N)
  float2 y[MAXB][16];
#pragma unroll
  for (int bi = 0; bi < MAXB; ++bi) {
    const int blk = threadIdx.x + bi * NT;
    if (blk >= nbz * nbu) continue;
    const int z4 = (blk / nbu) * 4, u4 = (blk - (blk / nbu) * nbu) * 4;
#pragma unroll
    for (int e = 0; e < 16; ++e) y[bi][e] = make_float2(0.f, 0.f);
    for (int sy = 0; sy < N; ++sy) {
      float2 pv[4], ev[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        pv[r] = plane[(z4 + r) * S + sy];
        ev[r] = (u4 + r < R) ? ts[(u4 + r) * N + sy] : make_float2(0.f, 0.f);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          y[bi][r * 4 + c].x += pv[r].x * ev[c].x - pv[r].y * ev[c].y;
          y[bi][r * 4 + c].y += pv[r].x * ev[c].y + pv[r].y * ev[c].x;
        }
    }
  }
  __syncthreads();
  // Y replaces P in LDS: [sz][u1], stride R + 1 (R <= N)
  const int YS = R + 1;
#pragma unroll
  for (int bi = 0; bi < MAXB; ++bi) {
    const int blk = threadIdx.x + bi * NT;
    if (blk >= nbz * nbu) continue;
    const int z4 = (blk / nbu) * 4, u4 = (blk - (blk / nbu) * nbu) * 4;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (u4 + c < R) plane[(z4 + r) * YS + u4 + c] = y[bi][r * 4 + c];
  }
  __syncthreads();
  // Q[u0][u1] = sum_sz e[u0][sz] Y[sz][u1], 2 x 2 register blocks
  const int nb2 = (R + 1) / 2;
  float2* Q = Qg + ((long long)b * H + kx) * R * R;
  for (int blk = threadIdx.x; blk < nb2 * nb2; blk += NT) {
    const int u0 = (blk / nb2) * 2, u1 = (blk - (blk / nb2) * nb2) * 2;
    float2 q[4] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f),
                   make_float2(0.f, 0.f)};
    for (int sz = 0; sz < N; ++sz) {
      const float2 e0 = u0 < R ? ts[u0 * N + sz] : make_float2(0.f, 0.f);
      const float2 e1 = u0 + 1 < R ? ts[(u0 + 1) * N + sz] : make_float2(0.f, 0.f);
      const float2 y0 = plane[sz * YS + u1];
      const float2 y1 = u1 + 1 < R ? plane[sz * YS + u1 + 1] : make_float2(0.f, 0.f);
      q[0].x += e0.x * y0.x - e0.y * y0.y; q[0].y += e0.x * y0.y + e0.y * y0.x;
      q[1].x += e0.x * y1.x - e0.y * y1.y; q[1].y += e0.x * y1.y + e0.y * y1.x;
      q[2].x += e1.x * y0.x - e1.y * y0.y; q[2].y += e1.x * y0.y + e1.y * y0.x;
      q[3].x += e1.x * y1.x - e1.y * y1.y; q[3].y += e1.x * y1.y + e1.y * y1.x;
    }
    if (u0 < R && u1 < R) Q[u0 * R + u1] = q[0];
    if (u0 < R && u1 + 1 < R) Q[u0 * R + u1 + 1] = q[1];
    if (u0 + 1 < R && u1 < R) Q[(u0 + 1) * R + u1] = q[2];
    if (u0 + 1 < R && u1 + 1 < R) Q[(u0 + 1) * R + u1 + 1] = q[3];
  }
}

// One workgroup per (particle, u0): Re sum_kx e2[u2][kx] Q[kx][u0][u1] over the
// R x R (u1, u2) slab, first argmax (flat index (u0 R + u1) R + u2).
__global__ __launch_bounds__(256) void shift_up_kx_kernel(const float2* __restrict__ Qg, int H, int R,
                                                          const float2* __restrict__ t2,
                                                          float* __restrict__ part_val,
                                                          int* __restrict__ part_idx) {
  extern __shared__ float4 smem4[];
  float2* qs = reinterpret_cast<float2*>(smem4);  // (H, R): Q[kx][u0][.]
  float2* es = qs + H * R;                        // (R, H)
  __shared__ float rv[4];
  __shared__ int ri[4];
  const int u0 = blockIdx.x, b = blockIdx.y;
  for (int q = threadIdx.x; q < H * R; q += 256) {
    const int kx = q / R, u1 = q - (q / R) * R;
    qs[q] = Qg[(((long long)b * H + kx) * R + u0) * R + u1];
  }
  for (int q = threadIdx.x; q < R * H; q += 256) es[q] = t2[q];
  __syncthreads();
  float bv = -INFINITY;
  int bi = 0x7fffffff;
  for (int q = threadIdx.x; q < R * R; q += 256) {
    const int u1 = q / R, u2 = q - (q / R) * R;
    float acc = 0.f;
    for (int kx = 0; kx < H; ++kx) {
      const float2 e = es[u2 * H + kx], v = qs[kx * R + u1];
      acc += e.x * v.x - e.y * v.y;
    }
    const int flat = (u0 * R + u1) * R + u2;
    if (acc > bv || (acc == bv && flat < bi)) { bv = acc; bi = flat; }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(bv, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
  }
  if ((threadIdx.x & 63) == 0) { rv[threadIdx.x >> 6] = bv; ri[threadIdx.x >> 6] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w)
      if (rv[w] > rv[0] || (rv[w] == rv[0] && ri[w] < ri[0])) { rv[0] = rv[w]; ri[0] = ri[w]; }
    part_val[(long long)b * R + u0] = rv[0];
    part_idx[(long long)b * R + u0] = ri[0];
  }
}

// argmax over the u0 partials; shift = (index - floor(R / 2)) / up per axis,
// flipped to the reference's (z, y, x) labels; score / (||S|| ||R||)
__global__ __launch_bounds__(64) void shift_up_final_kernel(const float* __restrict__ part_val,
                                                            const int* __restrict__ part_idx, int R,
                                                            float up, const float* __restrict__ rn,
                                                            const float* __restrict__ sn, int H,
                                                            float* __restrict__ out_shift,
                                                            float* __restrict__ out_score) {
  const int b = blockIdx.x;
  float bv = -INFINITY;
  int bi = 0x7fffffff;
  float r2 = 0.f, s2 = 0.f;
  for (int t = threadIdx.x; t < R; t += 64) {
    const float v = part_val[(long long)b * R + t];
    const int i = part_idx[(long long)b * R + t];
    if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
  }
  for (int t = threadIdx.x; t < H; t += 64) {
    r2 += rn[(long long)b * H + t];
    s2 += sn[(long long)b * H + t];
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(bv, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
  }
  r2 = mch::wave_sum(r2);
  s2 = mch::wave_sum(s2);
  if (threadIdx.x == 0) {
    const int u0 = bi / (R * R), u1 = (bi / R) % R, u2 = bi % R;
    const float half = floorf((float)R * 0.5f);
    // the reference returns flip(shift) of the array-axis shift
    out_shift[b * 3 + 0] = ((float)u2 - half) / up;
    out_shift[b * 3 + 1] = ((float)u1 - half) / up;
    out_shift[b * 3 + 2] = ((float)u0 - half) / up;
    out_score[b] = bv / (sqrtf(s2) * sqrtf(r2));
  }
}

__global__ __launch_bounds__(256) void shift_final_kernel(const float* __restrict__ amax_val,
                                                          const int* __restrict__ amax_idx,
                                                          int ntile_x, const float* __restrict__ rn,
                                                          const float* __restrict__ sn, int nparts,
                                                          int N, float* __restrict__ out_shift,
                                                          float* __restrict__ out_score) {
  const int b = blockIdx.x;
  __shared__ float rv[4], ra[4], rs[4];
  __shared__ int ri[4];
  float bv = -INFINITY;
  int bi = 0x7fffffff;
  float r2 = 0.f, s2 = 0.f;
  for (int t = threadIdx.x; t < ntile_x; t += blockDim.x) {
    const float v = amax_val[(long long)b * ntile_x + t];
    const int i = amax_idx[(long long)b * ntile_x + t];
    if (v > bv || (v == bv && i < bi)) { bv = v; bi = i; }
  }
  for (int t = threadIdx.x; t < nparts; t += blockDim.x) {
    r2 += rn[(long long)b * nparts + t];
    s2 += sn[(long long)b * nparts + t];
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(bv, off, 64);
    const int oi = __shfl_xor(bi, off, 64);
    if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
  }
  r2 = mch::wave_sum(r2);
  s2 = mch::wave_sum(s2);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { rv[w] = bv; ri[w] = bi; ra[w] = r2; rs[w] = s2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    float v = rv[0];
    int i = ri[0];
    float R2 = ra[0], S2 = rs[0];
    for (int k = 1; k < 4; ++k) {
      if (rv[k] > v || (rv[k] == v && ri[k] < i)) { v = rv[k]; i = ri[k]; }
      R2 += ra[k];
      S2 += rs[k];
    }
    const int t0 = i / (N * N), t1 = (i / N) % N, t2 = i % N;
    const int mid = N / 2;
    const int s0 = t0 > mid ? t0 - N : t0;
    const int s1 = t1 > mid ? t1 - N : t1;
    const int s2i = t2 > mid ? t2 - N : t2;
    // the reference returns flip(shift) of the array-axis shift
    out_shift[b * 3 + 0] = (float)s2i;
    out_shift[b * 3 + 1] = (float)s1;
    out_shift[b * 3 + 2] = (float)s0;
    out_score[b] = v / (sqrtf(S2) * sqrtf(R2));
  }
}

// patch[a] = vol[a + flip(shift_zyx)] (trilinear, zeros outside, border slice
// zeroed for fractional components): extract_shifted_patch_batch without
// normalisation (ShiftMatcher.py:558-616)
__global__ __launch_bounds__(256) void extract_patch_kernel(const float* __restrict__ vols,
                                                            long long ldv, int N,
                                                            const float* __restrict__ shift_zyx,
                                                            float* __restrict__ out,
                                                            long long ldo) {
  const int z = blockIdx.x, b = blockIdx.y;
  ShiftSpec sp;
  make_shift_spec(shift_zyx, b, N, sp);
  const float* v = vols + (long long)b * ldv;
  float* o = out + (long long)b * ldo + (long long)z * N * N;
  if (sp.integral && (N & 3) == 0 && (reinterpret_cast<uintptr_t>(o) & 15) == 0) {
    // integral shift: a shifted copy, four outputs per thread and store;
    // (y, q) advanced incrementally (no division in the loop)
    // out-of-volume voxels load through an out-of-range buffer offset, which
    // reads zero without a branch, so the loads of several iterations overlap
    const int nq = N >> 2;
    const int zs = z + sp.si[0];
    const bool zok = zs >= 0 && zs < N;
    const __amdgpu_buffer_rsrc_t rv = make_rsrc(v, (unsigned)(N * N * N) * 4u);
    const int dy = blockDim.x / nq, dq = blockDim.x - dy * nq;
    int y = threadIdx.x / nq, q = threadIdx.x - (threadIdx.x / nq) * nq;
    float4* o4 = reinterpret_cast<float4*>(o);
    const int n4 = N * nq;
    constexpr int U = 4;  // iterations whose loads are all issued before their stores
    for (int t0 = threadIdx.x; t0 < n4; t0 += U * (int)blockDim.x) {
      float4 r[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int ys = y + sp.si[1], x0 = 4 * q + sp.si[2];
        const bool rok = zok && ys >= 0 && ys < N && t0 + u * (int)blockDim.x < n4;
        const int rowoff = (zs * N + ys) * N;
        r[u].x = load_f1(rv, rok && x0 >= 0 && x0 < N ? (unsigned)(rowoff + x0) * 4u : kOutOfRange);
        r[u].y = load_f1(rv, rok && x0 + 1 >= 0 && x0 + 1 < N ? (unsigned)(rowoff + x0 + 1) * 4u : kOutOfRange);
        r[u].z = load_f1(rv, rok && x0 + 2 >= 0 && x0 + 2 < N ? (unsigned)(rowoff + x0 + 2) * 4u : kOutOfRange);
        r[u].w = load_f1(rv, rok && x0 + 3 >= 0 && x0 + 3 < N ? (unsigned)(rowoff + x0 + 3) * 4u : kOutOfRange);
        q += dq;
        y += dy;
        if (q >= nq) { q -= nq; ++y; }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int t = t0 + u * (int)blockDim.x;
        if (t < n4) o4[t] = r[u];
      }
    }
    return;
  }
  for (int t = threadIdx.x; t < N * N; t += blockDim.x) {
    const int y = t / N, x = t - y * N;
    o[t] = patch_value(v, N, sp, z, y, x);
  }
}

// Rotation for the shift search from search_orientations' Euler angles:
// R = Rz(a) Ry(b) Rz(g) acts on (axis 0, 1, 2); the shift search samples the
// reference with rb acting on (axis 2, 1, 0), rb = P R^T P (P reverses axes).
__global__ void euler_to_shift_rot_kernel(const float* __restrict__ al, const float* __restrict__ be,
                                          const float* __restrict__ ga, int B,
                                          float* __restrict__ rot) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double ca = cos((double)al[b]), sa = sin((double)al[b]);
  const double cb = cos((double)be[b]), sb = sin((double)be[b]);
  const double cg = cos((double)ga[b]), sg = sin((double)ga[b]);
  double R[3][3];
  // Rz(a) Ry(b) Rz(g)
  R[0][0] = ca * cb * cg - sa * sg; R[0][1] = -ca * cb * sg - sa * cg; R[0][2] = ca * sb;
  R[1][0] = sa * cb * cg + ca * sg; R[1][1] = -sa * cb * sg + ca * cg; R[1][2] = sa * sb;
  R[2][0] = -sb * cg;               R[2][1] = sb * sg;                 R[2][2] = cb;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) rot[b * 9 + i * 3 + j] = (float)R[2 - j][2 - i];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

template <int N> constexpr size_t row_lds() { return ((size_t)(N / 2) * (N + 1) + N) * sizeof(float2); }
template <int N> constexpr size_t plane_lds() { return ((size_t)N * (N + 1) + N) * sizeof(float2); }
// (x, y) slice passes: H lines of N + 1 and the twiddles; z column block: N rows of KB + 1
template <int N> constexpr size_t slice_lds() { return ((size_t)(N / 2 + 1) * (N + 1) + N) * sizeof(float2); }
template <int N> constexpr size_t zcol_lds() { return ((size_t)N * (kZKx * kZKy + 1) + N) * sizeof(float2); }

// Checks the LDS the kernels for box N need against the device and raises the
// dynamic-LDS limit of those that exceed the default 64 KiB.
template <int N>
int configure_kernels() {
  int dev = 0, maxlds = 0;
  int rc = mch::hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  rc = mch::hip_check(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
                      "hipDeviceGetAttribute");
  if (rc) return rc;
  const size_t fwd_lds = row_lds<N>() + kMaxLut * sizeof(float);
  const size_t fwd_xy_lds = slice_lds<N>() + kMaxLut * sizeof(float);
  if ((size_t)maxlds < fwd_lds + 1024 || (size_t)maxlds < fwd_xy_lds + 1024 || (size_t)maxlds < zcol_lds<N>()) {
    mch::set_error("mch_shift_plan_create: box %d needs %zu B of LDS per slice, device allows %d", N, fwd_xy_lds,
                   maxlds);
    return MCH_ERR_UNSUPPORTED;
  }
  auto raise = [&](const void* fn, size_t lds, const char* what) {
    if (lds <= 64 * 1024) return (int)MCH_OK;
    return mch::hip_check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds), what);
  };
  if ((rc = raise((const void*)shift_fwd_x_kernel<N>, fwd_lds, "hipFuncSetAttribute(shift_fwd_x_kernel)")))
    return rc;
  if ((rc = raise((const void*)shift_fwd_xy_kernel<N, false>, fwd_xy_lds, "hipFuncSetAttribute(shift_fwd_xy_kernel)")))
    return rc;
  if ((rc = raise((const void*)shift_fwd_xy_kernel<N, true>, fwd_xy_lds, "hipFuncSetAttribute(shift_fwd_xy_kernel)")))
    return rc;
  if ((rc = raise((const void*)shift_z_kernel<N, false>, zcol_lds<N>(), "hipFuncSetAttribute(shift_z_kernel)")))
    return rc;
  if ((rc = raise((const void*)shift_z_kernel<N, true>, zcol_lds<N>(), "hipFuncSetAttribute(shift_z_kernel)")))
    return rc;
  if ((rc = raise((const void*)shift_inv_xy_kernel<N, false>, slice_lds<N>(), "hipFuncSetAttribute(shift_inv_xy_kernel)")))
    return rc;
  if ((rc = raise((const void*)shift_inv_xy_kernel<N, true>, slice_lds<N>(), "hipFuncSetAttribute(shift_inv_xy_kernel)")))
    return rc;
  return MCH_OK;
}

void free_plan(ShiftPlan* p) {
  if (!p) return;
  if (p->timing)
    for (int i = 0; i < 5; ++i) (void)hipEventDestroy(p->ev[i]);
  (void)hipFree(p->refpair); (void)hipFree(p->refh16); (void)hipFree(p->twid); (void)hipFree(p->lut);
  (void)hipFree(p->XM);
  (void)hipFree(p->XMy);
  (void)hipFree(p->T); (void)hipFree(p->rnorm_part); (void)hipFree(p->snorm_part);
  (void)hipFree(p->bg_part); (void)hipFree(p->amax_val); (void)hipFree(p->amax_idx);
  (void)hipFree(p->t01); (void)hipFree(p->t2); (void)hipFree(p->Qg); (void)hipFree(p->up_val);
  (void)hipFree(p->up_idx);
  delete p;
}

template <int N> size_t plane_up_lds(int R) { return plane_lds<N>() + (size_t)R * N * sizeof(float2); }

// LDS of the refinement plane kernel for R against the device, attribute raised
template <int N>
int configure_up(int R) {
  int dev = 0, maxlds = 0;
  int rc = mch::hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  rc = mch::hip_check(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
                      "hipDeviceGetAttribute");
  if (rc) return rc;
  if (R > N || plane_up_lds<N>(R) + 1024 > (size_t)maxlds) {
    mch::set_error("mch_shift_plan_set_upsample: refinement grid R=%d too large for box %d (R <= N and "
                   "%zu B of LDS)", R, N, plane_up_lds<N>(R));
    return MCH_ERR_UNSUPPORTED;
  }
  rc = mch::hip_check(hipFuncSetAttribute((const void*)shift_yz_up_kernel<N, false>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)plane_up_lds<N>(R)),
                     "hipFuncSetAttribute(shift_yz_up_kernel)");
  if (rc) return rc;
  return mch::hip_check(hipFuncSetAttribute((const void*)shift_yz_up_kernel<N, true>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            (int)plane_up_lds<N>(R)),
                        "hipFuncSetAttribute(shift_yz_up_kernel)");
}

struct SearchArgs {
  const float* subtomos;
  long long ld_vol;
  int B;
  const float* shift_zyx;
  const float* rot;
  const float* ctf;
  float* out_shift;
  float* out_score;
};

// fwd_x -> yz -> inv_x -> final, all on one stream
template <int N>
int launch_search(ShiftPlan* p, const SearchArgs& a, hipStream_t stream) {
  constexpr int H = N / 2 + 1;
  int rc;
  auto mark = [&](int i) {
    if (p->timing) (void)hipEventRecord(p->ev[i], stream);
  };
  mark(0);
  if (p->up > 1) {
    const int R = p->R;
    hipLaunchKernelGGL(shift_fwd_x_kernel<N>, dim3(N, a.B), dim3(row_threads<N>()),
                       row_lds<N>() + (size_t)p->lut_n * sizeof(float), stream, a.subtomos, a.ld_vol,
                       a.shift_zyx, p->lut, p->lut_lo, p->lut_n, p->twid, p->T, p->bg_part);
    if ((rc = mch::hip_check(hipGetLastError(), "shift_fwd_x_kernel"))) return rc;
    mark(1);
    if (p->reduced)
      hipLaunchKernelGGL((shift_yz_up_kernel<N, true>), dim3(H * a.B), dim3(plane_threads<N>()),
                         plane_up_lds<N>(R), stream, p->T, p->XM, p->bg_part, p->wsum, p->refh16,
                         p->h16_unscale, a.rot, a.ctf, p->twid, p->t01, R, p->Qg, p->rnorm_part, p->snorm_part);
    else
      hipLaunchKernelGGL((shift_yz_up_kernel<N, false>), dim3(H * a.B), dim3(plane_threads<N>()),
                         plane_up_lds<N>(R), stream, p->T, p->XM, p->bg_part, p->wsum, p->refpair, 1.0f, a.rot,
                         a.ctf, p->twid, p->t01, R, p->Qg, p->rnorm_part, p->snorm_part);
    if ((rc = mch::hip_check(hipGetLastError(), "shift_yz_up_kernel"))) return rc;
    mark(2);
    hipLaunchKernelGGL(shift_up_kx_kernel, dim3(R, a.B), dim3(256), (size_t)2 * H * R * sizeof(float2), stream,
                       p->Qg, H, R, p->t2, p->up_val, p->up_idx);
    if ((rc = mch::hip_check(hipGetLastError(), "shift_up_kx_kernel"))) return rc;
    mark(3);
    hipLaunchKernelGGL(shift_up_final_kernel, dim3(a.B), dim3(64), 0, stream, p->up_val, p->up_idx, R,
                       (float)p->up, p->rnorm_part, p->snorm_part, H, a.out_shift, a.out_score);
    rc = mch::hip_check(hipGetLastError(), "shift_up_final_kernel");
    mark(4);
    return rc;
  }
  // reduced precision: the FP16 reference table and a BF16 work spectrum
  auto passes = [&](auto red) -> int {
    constexpr bool RED = decltype(red)::value;
    SpecElem<RED>* T = reinterpret_cast<SpecElem<RED>*>(p->T);
    hipLaunchKernelGGL((shift_fwd_xy_kernel<N, RED>), dim3(N, a.B), dim3(slice_threads<N>()),
                       slice_lds<N>() + (size_t)p->lut_n * sizeof(float), stream, a.subtomos, a.ld_vol,
                       a.shift_zyx, p->lut, p->lut_lo, p->lut_n, p->twid, T, p->bg_part);
    int e;
    if ((e = mch::hip_check(hipGetLastError(), "shift_fwd_xy_kernel"))) return e;
    mark(1);
    hipLaunchKernelGGL((shift_z_kernel<N, RED>), dim3(zparts<N>() * a.B), dim3(kZThreads), zcol_lds<N>(), stream,
                       T, p->XMy, p->bg_part, p->wsum, RED ? (const void*)p->refh16 : (const void*)p->refpair,
                       RED ? p->h16_unscale : 1.0f, a.rot, a.ctf, p->twid, p->rnorm_part, p->snorm_part);
    if ((e = mch::hip_check(hipGetLastError(), "shift_z_kernel"))) return e;
    mark(2);
    hipLaunchKernelGGL((shift_inv_xy_kernel<N, RED>), dim3(N, a.B), dim3(slice_threads<N>()), slice_lds<N>(),
                       stream, T, p->twid, p->amax_val, p->amax_idx);
    if ((e = mch::hip_check(hipGetLastError(), "shift_inv_xy_kernel"))) return e;
    mark(3);
    return MCH_OK;
  };
  rc = p->reduced ? passes(std::true_type()) : passes(std::false_type());
  if (rc) return rc;
  hipLaunchKernelGGL(shift_final_kernel, dim3(a.B), dim3(256), 0, stream, p->amax_val, p->amax_idx,
                     p->ntile_x, p->rnorm_part, p->snorm_part, zparts<N>(), N, a.out_shift, a.out_score);
  rc = mch::hip_check(hipGetLastError(), "shift_final_kernel");
  mark(4);
  return rc;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int mch_shift_set_precision(void* plan, int reduced) {
  ShiftPlan* p = (ShiftPlan*)plan;
  if (!p) { mch::set_error("mch_shift_set_precision: NULL plan"); return MCH_ERR_ARG; }
  if (reduced != 0 && reduced != 1) {
    mch::set_error("mch_shift_set_precision: reduced must be 0 or 1 (got %d)", reduced);
    return MCH_ERR_ARG;
  }
  p->reduced = reduced;
  return MCH_OK;
}

int mch_shift_plan_set_timing(void* plan, int enable) {
  ShiftPlan* p = (ShiftPlan*)plan;
  if (!p) { mch::set_error("mch_shift_plan_set_timing: NULL plan"); return MCH_ERR_ARG; }
  if (enable && !p->timing) {
    for (int i = 0; i < 5; ++i) {
      const int rc = mch::hip_check(hipEventCreate(&p->ev[i]), "hipEventCreate");
      if (rc) {
        for (int k = 0; k < i; ++k) (void)hipEventDestroy(p->ev[k]);
        return rc;
      }
    }
    p->timing = 1;
  } else if (!enable && p->timing) {
    for (int i = 0; i < 5; ++i) (void)hipEventDestroy(p->ev[i]);
    p->timing = 0;
  }
  return MCH_OK;
}

int mch_shift_kernel_ms(void* plan, float* ms4) {
  ShiftPlan* p = (ShiftPlan*)plan;
  if (!p || !ms4 || !p->timing) {
    mch::set_error("mch_shift_kernel_ms: timing is not enabled on this plan");
    return MCH_ERR_ARG;
  }
  int rc = mch::hip_check(hipEventSynchronize(p->ev[4]), "hipEventSynchronize");
  for (int i = 0; i < 4 && !rc; ++i)
    rc = mch::hip_check(hipEventElapsedTime(&ms4[i], p->ev[i], p->ev[i + 1]), "hipEventElapsedTime");
  return rc;
}

int mch_extract_patches(const float* vols, long long ld_vol, int B, int N, const float* shift_zyx,
                        float* out, long long ld_out, hipStream_t stream) {
  if (!vols || !out || B < 1 || N < 1 || ld_vol < (long long)N * N * N || ld_out < (long long)N * N * N) {
    mch::set_error("mch_extract_patches: bad arguments");
    return MCH_ERR_ARG;
  }
  hipLaunchKernelGGL(extract_patch_kernel, dim3(N, B), dim3(256), 0, stream, vols, ld_vol, N,
                     shift_zyx, out, ld_out);
  return mch::hip_check(hipGetLastError(), "extract_patch_kernel");
}

int mch_euler_to_shift_rot(const float* alpha, const float* beta, const float* gamma, int B,
                           float* rot, hipStream_t stream) {
  if (!alpha || !beta || !gamma || !rot || B < 1) {
    mch::set_error("mch_euler_to_shift_rot: bad arguments");
    return MCH_ERR_ARG;
  }
  hipLaunchKernelGGL(euler_to_shift_rot_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, alpha,
                     beta, gamma, B, rot);
  return mch::hip_check(hipGetLastError(), "euler_to_shift_rot_kernel");
}

int mch_shift_plan_create(int N, const void* ref_host /* complex64 (N,N,N) */, float box_radius,
                          int max_batch, void** out_plan) {
  if (!out_plan || !ref_host) { mch::set_error("mch_shift_plan_create: NULL argument"); return MCH_ERR_ARG; }
  *out_plan = nullptr;
  if (max_batch < 1) {
    mch::set_error("mch_shift_plan_create: max_batch=%d", max_batch);
    return MCH_ERR_ARG;
  }
  int rc = MCH_OK;
  switch (N) {
    case 16: rc = configure_kernels<16>(); break;
    case 32: rc = configure_kernels<32>(); break;
    case 48: rc = configure_kernels<48>(); break;
    case 64: rc = configure_kernels<64>(); break;
    case 80: rc = configure_kernels<80>(); break;
    case 96: rc = configure_kernels<96>(); break;
    case 112: rc = configure_kernels<112>(); break;
    case 128: rc = configure_kernels<128>(); break;
    case 144: rc = configure_kernels<144>(); break;
    case 160: rc = configure_kernels<160>(); break;
    case 192: rc = configure_kernels<192>(); break;
    default:
      mch::set_error("mch_shift_plan_create: box N=%d unsupported (16, 32, 48, 64, 80, 96, 112, 128, 144, 160 "
                     "or 192)", N);
      return MCH_ERR_UNSUPPORTED;
  }
  if (rc) return rc;
  ShiftPlan* p = new ShiftPlan();
  memset(p, 0, sizeof(*p));
  p->N = N;
  p->H = N / 2 + 1;
  p->max_batch = max_batch;
  p->radius = box_radius;
  p->width = 5.0f;
  p->radius_p = box_radius + 5.0f;
  p->ntile_x = N;
  p->up = 1;
  const int H = p->H;
  // background weight per integer r2 (monotone: 0, shell, 1) and its table
  const int r2max = 3 * (N / 2) * (N / 2);
  std::vector<float> wr2(r2max + 1);
  int lo = r2max + 1, hi = -1;
  for (int r2 = 0; r2 <= r2max; ++r2) {
    wr2[r2] = bg_weight_host(r2, p->radius, p->radius_p, p->width);
    if (wr2[r2] != 0.f && lo > r2max) lo = r2;
    if (wr2[r2] != 1.f) hi = r2;
  }
  if (lo > hi) { lo = hi + 1; }  // no shell: everything 0 below lo, 1 above hi
  p->lut_lo = lo;
  p->lut_n = hi >= lo ? hi - lo + 1 : 0;
  if (p->lut_n > kMaxLut) {
    mch::set_error("mch_shift_plan_create: cosine mask shell of %d radii exceeds %d", p->lut_n, kMaxLut);
    delete p;
    return MCH_ERR_UNSUPPORTED;
  }
  std::vector<float> lut(p->lut_n > 0 ? p->lut_n : 1, 0.f);
  for (int i = 0; i < p->lut_n; ++i) lut[i] = wr2[lo + i];
  // weights of the rolled patch (offset (i + N/2) mod N - N/2), their sum and
  // x-spectrum XM[kx][z][y] = sum_x w(z, y, x) e^{-2 pi i kx x / N}
  std::vector<float2> xm((size_t)H * N * N), xmy((size_t)H * N * N);
  {
    std::vector<double> xre((size_t)H * N * N), xim((size_t)H * N * N);
    std::vector<double> c(N), s(N);
    for (int k = 0; k < N; ++k) {
      c[k] = cos(-2.0 * M_PI * k / N);
      s[k] = sin(-2.0 * M_PI * k / N);
    }
    std::vector<float> row(N);
    double wsum = 0.0;
    for (int z = 0; z < N; ++z)
      for (int y = 0; y < N; ++y) {
        const int d0 = (z + N / 2) % N - N / 2, d1 = (y + N / 2) % N - N / 2;
        for (int x = 0; x < N; ++x) {
          const int d2 = (x + N / 2) % N - N / 2;
          row[x] = wr2[d0 * d0 + d1 * d1 + d2 * d2];
          wsum += row[x];
        }
        for (int kx = 0; kx < H; ++kx) {
          double re = 0.0, im = 0.0;
          for (int x = 0; x < N; ++x) {
            if (row[x] == 0.f) continue;
            const int j = (kx * x) % N;
            re += row[x] * c[j];
            im += row[x] * s[j];
          }
          const size_t o = ((size_t)kx * N + z) * N + y;
          xm[o] = make_float2((float)re, (float)im);
          xre[o] = re;
          xim[o] = im;
        }
      }
    p->wsum = (float)wsum;
    // XMy[kx][z][ky] = sum_y XM[kx][z][y] e^{-2 pi i ky y / N}, from the FP64 values
    for (size_t kz = 0; kz < (size_t)H * N; ++kz) {
      const double* r = &xre[kz * N];
      const double* q = &xim[kz * N];
      for (int ky = 0; ky < N; ++ky) {
        double re = 0.0, im = 0.0;
        for (int y = 0; y < N; ++y) {
          const int j = (ky * y) % N;
          re += r[y] * c[j] - q[y] * s[j];
          im += r[y] * s[j] + q[y] * c[j];
        }
        xmy[kz * N + ky] = make_float2((float)re, (float)im);
      }
    }
  }
  std::vector<float2> tw(N);
  for (int k = 0; k < N; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)N;
    tw[k] = make_float2((float)cos(a), (float)sin(a));
  }
  const size_t half = (size_t)N * N * H;
  // both tables over (z', y', x') = (z + 1, y + 1, x + 1) with a zero border
  const size_t nent = table_entries(N);
  const float* rf = (const float*)ref_host;
  // FP32 pairs: (ref[z][y][x], ref[z][y][x + 1]), (z, y, x) = (z' - 1, y' - 1, x' - 1)
  std::vector<float4> pair(nent, make_float4(0.f, 0.f, 0.f, 0.f));
  {
    auto at = [&](int z, int y, int x) -> float2 {
      if (z < 0 || z >= N || y < 0 || y >= N || x < 0 || x >= N) return make_float2(0.f, 0.f);
      const size_t o = (((size_t)z * N + y) * N + x) * 2;
      return make_float2(rf[o], rf[o + 1]);
    };
    for (int zp = 0; zp < N + 2; ++zp)
      for (int yp = 0; yp < N + 2; ++yp)
        for (int xp = 0; xp <= N; ++xp) {
          const float2 a = at(zp - 1, yp - 1, xp - 1), b = at(zp - 1, yp - 1, xp);
          pair[table_index(N, zp, yp, xp)] = make_float4(a.x, a.y, b.x, b.y);
        }
  }
  // FP16 quads: rows y, y + 1 x columns x, x + 1 of plane z, scaled by 2^e so
  // that the largest component stays below 2^14 (FP16 keeps 11 significant
  // bits down to 2^-14)
  std::vector<uint4> quad(nent, make_uint4(0u, 0u, 0u, 0u));
  {
    float maxabs = 0.f;
    for (size_t i = 0; i < (size_t)2 * N * N * N; ++i) maxabs = fmaxf(maxabs, fabsf(rf[i]));
    int ex = 0;
    if (maxabs > 0.f) (void)frexpf(maxabs, &ex);  // maxabs < 2^ex
    const int e = 14 - ex;
    const float up = ldexpf(1.f, e);
    p->h16_unscale = ldexpf(1.f, -e);
    auto comp = [&](int z, int y, int x, int c) -> _Float16 {
      if (z < 0 || z >= N || y < 0 || y >= N || x < 0 || x >= N) return (_Float16)0.f;
      return (_Float16)(rf[(((size_t)z * N + y) * N + x) * 2 + c] * up);
    };
    auto pack = [](_Float16 lo, _Float16 hi) {
      return (unsigned)__builtin_bit_cast(unsigned short, lo) |
             ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
    };
    for (int zp = 0; zp < N + 2; ++zp)
      for (int yp = 0; yp < N + 2; ++yp)
        for (int xp = 0; xp <= N; ++xp) {
          const int z = zp - 1, y = yp - 1, x = xp - 1;
          uint4 q;
          q.x = pack(comp(z, y, x, 0), comp(z, y, x + 1, 0));
          q.y = pack(comp(z, y, x, 1), comp(z, y, x + 1, 1));
          q.z = pack(comp(z, y + 1, x, 0), comp(z, y + 1, x + 1, 0));
          q.w = pack(comp(z, y + 1, x, 1), comp(z, y + 1, x + 1, 1));
          quad[table_index(N, zp, yp, xp)] = q;
        }
  }
#define MCH_TRY(expr, what) do { if ((rc = mch::hip_check((expr), what)) != MCH_OK) goto fail; } while (0)
  MCH_TRY(hipMalloc(&p->refpair, nent * sizeof(float4)), "hipMalloc(refpair)");
  MCH_TRY(hipMemcpy(p->refpair, pair.data(), nent * sizeof(float4), hipMemcpyHostToDevice), "copy ref");
  MCH_TRY(hipMalloc(&p->refh16, nent * sizeof(uint4)), "hipMalloc(refh16)");
  MCH_TRY(hipMemcpy(p->refh16, quad.data(), nent * sizeof(uint4), hipMemcpyHostToDevice), "copy ref quads");
  MCH_TRY(hipMalloc(&p->twid, N * sizeof(float2)), "hipMalloc(twid)");
  MCH_TRY(hipMemcpy(p->twid, tw.data(), N * sizeof(float2), hipMemcpyHostToDevice), "copy twid");
  MCH_TRY(hipMalloc(&p->lut, lut.size() * sizeof(float)), "hipMalloc(lut)");
  MCH_TRY(hipMemcpy(p->lut, lut.data(), lut.size() * sizeof(float), hipMemcpyHostToDevice), "copy lut");
  MCH_TRY(hipMalloc(&p->XM, half * sizeof(float2)), "hipMalloc(XM)");
  MCH_TRY(hipMemcpy(p->XM, xm.data(), half * sizeof(float2), hipMemcpyHostToDevice), "copy XM");
  MCH_TRY(hipMalloc(&p->XMy, half * sizeof(float2)), "hipMalloc(XMy)");
  MCH_TRY(hipMemcpy(p->XMy, xmy.data(), half * sizeof(float2), hipMemcpyHostToDevice), "copy XMy");
  MCH_TRY(hipMalloc(&p->T, (size_t)max_batch * half * sizeof(float2)), "hipMalloc(T)");
  // norm partials: per column block on the integer path, per kx on the sub-voxel path
  {
    const int zp = ((H + 3) / 4) * (N / 8);
    const size_t nparts = (size_t)max_batch * (zp > H ? zp : H);
    MCH_TRY(hipMalloc(&p->rnorm_part, nparts * sizeof(float)), "hipMalloc");
    MCH_TRY(hipMalloc(&p->snorm_part, nparts * sizeof(float)), "hipMalloc");
  }
  MCH_TRY(hipMalloc(&p->bg_part, (size_t)max_batch * N * sizeof(float)), "hipMalloc");
  MCH_TRY(hipMalloc(&p->amax_val, (size_t)max_batch * p->ntile_x * sizeof(float)), "hipMalloc");
  MCH_TRY(hipMalloc(&p->amax_idx, (size_t)max_batch * p->ntile_x * sizeof(int)), "hipMalloc");
#undef MCH_TRY
  *out_plan = p;
  return MCH_OK;
fail:
  free_plan(p);
  return rc;
}

int mch_shift_plan_destroy(void* plan) {
  free_plan((ShiftPlan*)plan);
  return MCH_OK;
}

// Sub-voxel refinement of the shift search: the reference's local matrix-DFT
// with upsample factor `up` over a window of 2 * radius + 1 voxels (up == 1:
// integer lags of the full cross-correlation, the default).
int mch_shift_plan_set_upsample(void* plan, int up, int radius) {
  ShiftPlan* p = (ShiftPlan*)plan;
  if (!p || up < 1 || radius < 0) {
    mch::set_error("mch_shift_plan_set_upsample: bad arguments (up=%d radius=%d)", up, radius);
    return MCH_ERR_ARG;
  }
  if (up == 1) {
    p->up = 1;
    return MCH_OK;
  }
  const int N = p->N, H = p->H;
  const int R = (int)ceil((double)up * (double)(2 * radius + 1));
  int rc = MCH_OK;
  switch (N) {
    case 16: rc = configure_up<16>(R); break;
    case 32: rc = configure_up<32>(R); break;
    case 48: rc = configure_up<48>(R); break;
    case 64: rc = configure_up<64>(R); break;
    case 80: rc = configure_up<80>(R); break;
    case 96: rc = configure_up<96>(R); break;
    case 112: rc = configure_up<112>(R); break;
    case 128: rc = configure_up<128>(R); break;
    default:
      // the refinement's plane pass holds a whole (y, z) plane in LDS
      mch::set_error("mch_shift_plan_set_upsample: sub-voxel refinement supports boxes up to 128 (N=%d)", N);
      rc = MCH_ERR_UNSUPPORTED;
      break;
  }
  if (rc) return rc;
  // phases in FP64: e^{2 pi i (u - floor(R/2)) f}, f = fftfreq(n, d=up)
  auto fftfreq = [](int k, int n, int d) {
    const int kk = (k <= (n - 1) / 2) ? k : k - n;
    return (double)kk / ((double)n * (double)d);
  };
  const double off = floor(R / 2.0);
  std::vector<float2> t01((size_t)R * N), t2((size_t)R * H);
  for (int u = 0; u < R; ++u) {
    for (int k = 0; k < N; ++k) {
      const double a = 2.0 * M_PI * ((double)u - off) * fftfreq(k, N, up);
      t01[(size_t)u * N + k] = make_float2((float)cos(a), (float)sin(a));
    }
    for (int k = 0; k < H; ++k) {
      const double a = 2.0 * M_PI * ((double)u - off) * fftfreq(k, H, up);
      t2[(size_t)u * H + k] = make_float2((float)cos(a), (float)sin(a));
    }
  }
  (void)hipFree(p->t01); (void)hipFree(p->t2); (void)hipFree(p->Qg); (void)hipFree(p->up_val);
  (void)hipFree(p->up_idx);
  p->t01 = p->t2 = p->Qg = nullptr;
  p->up_val = nullptr;
  p->up_idx = nullptr;
#define MCH_TRY(expr, what) do { if ((rc = mch::hip_check((expr), what)) != MCH_OK) return rc; } while (0)
  MCH_TRY(hipMalloc(&p->t01, t01.size() * sizeof(float2)), "hipMalloc(t01)");
  MCH_TRY(hipMemcpy(p->t01, t01.data(), t01.size() * sizeof(float2), hipMemcpyHostToDevice), "copy t01");
  MCH_TRY(hipMalloc(&p->t2, t2.size() * sizeof(float2)), "hipMalloc(t2)");
  MCH_TRY(hipMemcpy(p->t2, t2.data(), t2.size() * sizeof(float2), hipMemcpyHostToDevice), "copy t2");
  MCH_TRY(hipMalloc(&p->Qg, (size_t)p->max_batch * H * R * R * sizeof(float2)), "hipMalloc(Qg)");
  MCH_TRY(hipMalloc(&p->up_val, (size_t)p->max_batch * R * sizeof(float)), "hipMalloc(up_val)");
  MCH_TRY(hipMalloc(&p->up_idx, (size_t)p->max_batch * R * sizeof(int)), "hipMalloc(up_idx)");
#undef MCH_TRY
  p->up = up;
  p->R = R;
  return MCH_OK;
}

// subtomos (B, N^3) float32; shift_zyx (B, 3) float32 or NULL; rot (B, 9) float32
// (row-major matrices acting on (x, y, z) = (axis 2, 1, 0)); ctf (B, N, N, N/2+1)
// float32 or NULL.  Outputs shift (B, 3) and score (B,).
int mch_shift_search(void* plan, const float* subtomos, long long ld_vol, int B,
                     const float* shift_zyx, const float* rot, const float* ctf, float* out_shift,
                     float* out_score, hipStream_t stream) {
  ShiftPlan* p = (ShiftPlan*)plan;
  if (!p || !subtomos || !rot || !out_shift || !out_score) {
    mch::set_error("mch_shift_search: NULL argument");
    return MCH_ERR_ARG;
  }
  const int N = p->N;
  if (B < 1 || B > p->max_batch || ld_vol < (long long)N * N * N) {
    mch::set_error("mch_shift_search: bad sizes (B=%d max %d)", B, p->max_batch);
    return MCH_ERR_SIZE;
  }
  const SearchArgs a{subtomos, ld_vol, B, shift_zyx, rot, ctf, out_shift, out_score};
  switch (N) {
    case 16: return launch_search<16>(p, a, stream);
    case 32: return launch_search<32>(p, a, stream);
    case 48: return launch_search<48>(p, a, stream);
    case 64: return launch_search<64>(p, a, stream);
    case 80: return launch_search<80>(p, a, stream);
    case 96: return launch_search<96>(p, a, stream);
    case 112: return launch_search<112>(p, a, stream);
    case 128: return launch_search<128>(p, a, stream);
    case 144: return launch_search<144>(p, a, stream);
    case 160: return launch_search<160>(p, a, stream);
    case 192: return launch_search<192>(p, a, stream);
    default: break;
  }
  mch::set_error("mch_shift_search: plan with unsupported N=%d", N);
  return MCH_ERR_UNSUPPORTED;
}

}  // extern "C"
