// Ball-harmonic projection a = B^H f and correlation coefficients sigma on gfx950.
//
// Reference behaviour (swing-research/Matcha):
//   FLEBasis3D.evaluate_t      src/matcha/core/FLEBasis3D.py:945-976,1026-1267
//   compute_sigma / get_sigma  src/matcha/core/CrossCorrelationMatcher.py:183-249
//
// Pipeline per batch:
//   1. fle_symmetrize_kernel: each fundamental voxel's 16 orbit values ->
//      4 combinations per class (8 classes), written as the GEMM A operand
//      X[class][row][k] with rows (b: real part, B+b: imaginary part) and
//      k = v0 (cos-weighted) or Kpad/2 + v0 (sin-weighted); for the classes
//      m = 0, 2 mod 4 the real rows' sin half and the imaginary rows' cos
//      half are zero, and GEMM row tiles wholly of one kind skip that half.
//   2. the projection GEMM Out[row][col] = sum_k X[class(col)][row][k] W[k][col],
//      split over K, in the plan's precision:
//        FP32 (plain path): fle_gemm_kernel, v_mfma_f32_16x16x4f32;
//        BF16x3 / FP16x3 (tuned path): fle_gemm16_kernel on 16-bit hi/lo
//        operands, Xhi Whi + Xhi Wlo + Xlo Whi with FP32 accumulation
//        (v_mfma_f32_16x16x32_{bf16,f16}); BF16 X comes split out of step 1,
//        FP16 X is split by fle_xsplit_kernel with power-of-2 row scales.
//   3. fle_finalize_kernel: sum the K splits (undo the FP16 scales),
//      a_{lkm} = Re + i Im and a_{lk,-m} = (-1)^m conj(a_{lkm}), scattered
//      to the reference order.
#include "common.hpp"
#include "special.hpp"

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

namespace {

using mch::kOutOfRange;
using mch::load_f1;
using mch::make_rsrc;
using mch::make_shift_spec;
using mch::patch_value;
using mch::ShiftSpec;

constexpr int kClasses = 8;
constexpr int BM = 128, BN = 64, BK = 16;
constexpr int kGemmThreads = 256;

// 16-bit split GEMM (tuned path): tiles of 128 rows x 128 columns x 32 k
constexpr int BM16 = 128, BN16 = 128, BK16 = 32, LDA16 = BK16 + 8;

// compute precision of the projection GEMM
enum FlePrecision { kFp32 = 0, kBf16x3 = 1, kFp16x3 = 2 };

struct FlePlan {
  int N, V0, Kpad, Ntot, ne, max_batch;
  int col_off[kClasses + 1];  // padded column offsets per class
  int* orbit;                 // (V0, 16)
  int* fcoord;                // (V0) fundamental voxel (x, y, z) >= 0, packed x | y << 10 | z << 20
  float* W;                   // (Kpad, Ntot)
  int* chan_plus;             // (Ntot)
  int* chan_minus;            // (Ntot)
  int* chan_m;                // (Ntot)
  float* X;                   // (kClasses, Mpad, Kpad)
  float* part;                // (splits, Mpad, Ntot)
  int Mpad, splits;
  // tuned path: W^T split into 16-bit hi + lo with a power-of-2 scale per column
  int precision;              // FlePrecision
  int table_precision;        // precision the 16-bit tables below were built for (0: none)
  uint16_t* Xhi;              // (kClasses, Mpad, Kpad) 16-bit split X operand
  uint16_t* Xlo;              // (kClasses, Mpad, Kpad)
  int nst16;                  // pipeline depth of the 16-bit GEMM (autotuned: 2 or 3)
  int nw16;                   // waves per workgroup of the 16-bit GEMM (autotuned: 2, 4 or 8)
  int ct16;                   // 128-column tiles per workgroup (autotuned: 1, or 2 with 8 waves)
  uint16_t* Wt16;             // W^T hi/lo in DMA order, Ntot * Kpad * 2 (see fle_wsplit_kernel)
  float* colscale;            // (Ntot)
  float* rowmax;              // (Mpad) max |X| of each GEMM row (FP16 scaling)
  float* part16;              // (splits16, Mpad, Ntot)
  int splits16;               // K splits part16 holds (the largest autotune candidate)
  int ksplit16;               // K splits in use (autotuned, <= splits16)
  int cus;                    // compute units of the plan's device
  int wide_slots;             // resident workgroups of the wide 16-bit GEMM on the device
};

struct CorrPlan {
  int lmax, kdim, max_batch;
  int* idx;   // (lmax+1, 2*lmax+1, kdim): coefficient index of (l, m, k) or -1
};

// The operand slab a class reads.  Class (eps, 3) is class (eps, 1) with X2
// and X3 negated (rho = 3 is the conjugate of rho = 1): it reads the (eps, 1)
// slab against a basis whose sin rows are negated, and its imaginary output
// is negated back in the finalize.  Slabs 3 and 7 are never written.
__host__ __device__ __forceinline__ int x_slab(int cls) { return (cls & 3) == 3 ? cls - 2 : cls; }

// Class of padded column c: the count of class starts at or below it, summed
// with constant indices (a loop that indexes the by-value offsets with a
// running class would copy them to scratch memory in every thread).
__device__ __forceinline__ int column_class(const int (&col_off)[kClasses + 1], int c) {
  int cls = 0;
#pragma unroll
  for (int i = 1; i < kClasses; ++i) cls += c >= col_off[i] ? 1 : 0;
  return cls;
}

// K range a GEMM row tile [row0, row0 + rows) of class cls needs: the classes
// m = 0, 2 mod 4 (cls even) have zero sin-weighted operands in their real
// rows (< B) and zero cos-weighted ones in their imaginary rows, so a tile
// wholly of one kind reads one half of K; other tiles read all of it.
__host__ __device__ __forceinline__ void k_range(int cls, int row0, int rows, int B, int Kpad, int& klo,
                                                 int& khi) {
  // selects, not branches: branchy stores to the two outputs were turned into
  // a store through a computed address into scratch memory
  const bool half = (cls & 1) == 0;
  const bool real_only = half && row0 + rows <= B;
  const bool imag_only = half && !real_only && row0 >= B;
  klo = imag_only ? Kpad / 2 : 0;
  khi = real_only ? Kpad / 2 : Kpad;
}

// Whether operand row `row` lies in the one GEMM row tile that holds both real
// and imaginary rows (tiles are 128 or 256 rows, aligned to their height, so
// it lies inside the aligned 256 rows around B).  Only there is the zero half
// of a class m = 0, 2 mod 4 row read, so only there is it stored; elsewhere it
// keeps whatever an earlier batch left.
__device__ __forceinline__ bool straddle_row(int row, int B) {
  return (B & 255) != 0 && row >= (B & ~255) && row < ((B + 255) & ~255);
}

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// v = hi + lo with hi, lo 16-bit (round to nearest even), and the 16x16x32 MFMA
// of that format.  BF16 keeps the FP32 exponent range; FP16 needs scaled data.
template <bool FP16>
struct Split16;
template <>
struct Split16<false> {
  static __device__ __forceinline__ void split(float v, uint16_t& hi, uint16_t& lo) {
    const __bf16 h = (__bf16)v;
    hi = __builtin_bit_cast(uint16_t, h);
    lo = __builtin_bit_cast(uint16_t, (__bf16)(v - (float)h));
  }
  static __device__ __forceinline__ floatx4 mfma(u32x4 a, u32x4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  }
};
template <>
struct Split16<true> {
  static __device__ __forceinline__ void split(float v, uint16_t& hi, uint16_t& lo) {
    const _Float16 h = (_Float16)v;
    hi = __builtin_bit_cast(uint16_t, h);
    lo = __builtin_bit_cast(uint16_t, (_Float16)(v - (float)h));
  }
  static __device__ __forceinline__ floatx4 mfma(u32x4 a, u32x4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  }
};

// Power-of-2 scale that brings max |x| = m just below 2^14 (FP16 range with
// room for the FP32 accumulation of the products); 1 for m == 0.
__host__ __device__ __forceinline__ float pow2_scale(float m) {
  if (!(m > 0.f)) return 1.f;
  int e;
  frexpf(m, &e);
  return ldexpf(1.f, 14 - e);
}

}  // namespace

// ---------------------------------------------------------------------------
// 1. orbit symmetrisation
// ---------------------------------------------------------------------------
// For class (eps, rho): G_t = sum_q (-1)^{q eps} sum_p (-i)^{rho p} f[p,q,t],
// X1 = A0+A1, X2 = B0-B1, X3 = B0+B1, X4 = A1-A0 with G_t = A_t + i B_t.
// X in FP32; rowmax (Mpad floats, zeroed by the caller; NULL to skip)
// receives max |X| of each GEMM row, for the power-of-2 row scale of the FP16
// split GEMM.  shift_zyx (B, 3) or NULL: the volumes are read as the patches
// extract_shifted_patches would cut from them (vol[a + flip(shift)], trilinear
// for fractional shifts).
__global__ void fle_symmetrize_kernel(const float* __restrict__ vols, long long ld_vol, int B,
                                      const int* __restrict__ orbit, int V0, int N, int Kpad, int Mpad,
                                      const float* __restrict__ shift_zyx, float* __restrict__ X,
                                      float* __restrict__ rowmax) {
  const long long tot = (long long)B * V0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  // block-uniform trip count so that the wave reductions see every lane
  for (long long t0 = (long long)blockIdx.x * blockDim.x; t0 < tot; t0 += stride) {
    const long long t = t0 + threadIdx.x;
    const bool active = t < tot;
    const int b = active ? (int)(t / V0) : 0;
    const int v = active ? (int)(t - (long long)b * V0) : 0;
    float mre = 0.f, mim = 0.f;
    if (active) {
      const float* f = vols + (long long)b * ld_vol;
      const int4* o4 = reinterpret_cast<const int4*>(orbit + (long long)v * 16);
      float g[16];
      if (shift_zyx) {
        ShiftSpec sp;
        make_shift_spec(shift_zyx, b, N, sp);
        for (int i = 0; i < 16; ++i) {
          const int id = orbit[(long long)v * 16 + i];
          g[i] = id < 0 ? 0.f : patch_value(f, N, sp, id / (N * N), (id / N) % N, id % N);
        }
      } else {
        // index -1: an image outside the array (radius >= N/2), read as zero
        // through the buffer range check
        const __amdgpu_buffer_rsrc_t rv = make_rsrc(f, (unsigned)(N * N * N) * 4u);
        auto at = [&](int id) { return load_f1(rv, id < 0 ? kOutOfRange : (unsigned)id * 4u); };
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int4 id = o4[i];
          g[4 * i + 0] = at(id.x);
          g[4 * i + 1] = at(id.y);
          g[4 * i + 2] = at(id.z);
          g[4 * i + 3] = at(id.w);
        }
      }
      // g[p + 4q + 8t]
#pragma unroll
      for (int eps = 0; eps < 2; ++eps) {
        // H_t(q) sums over p for each rho, then combine q with (-1)^{q eps}
        float A[2][4], Bi[2][4];  // [t][rho]
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
          float s0 = 0.f, s2 = 0.f, s1r = 0.f, s1i = 0.f;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const float sg = (eps && q) ? -1.f : 1.f;
            const float* h = g + 4 * q + 8 * tt;
            // rho = 0: sum_p f ; rho = 2: sum (-1)^p f ; rho = 1: sum (-i)^p f
            s0 += sg * (h[0] + h[1] + h[2] + h[3]);
            s2 += sg * (h[0] - h[1] + h[2] - h[3]);
            s1r += sg * (h[0] - h[2]);
            s1i += sg * (h[3] - h[1]);
          }
          A[tt][0] = s0;  Bi[tt][0] = 0.f;
          A[tt][1] = s1r; Bi[tt][1] = s1i;
          A[tt][2] = s2;  Bi[tt][2] = 0.f;
          A[tt][3] = s1r; Bi[tt][3] = -s1i;  // rho = 3: conjugate of rho = 1
        }
#pragma unroll
        for (int rho = 0; rho < 3; ++rho) {  // rho = 3 reads the rho = 1 slab
          const int cls = eps * 4 + rho;
          const float X1 = A[0][rho] + A[1][rho];
          const float X2 = Bi[0][rho] - Bi[1][rho];
          const float X3 = Bi[0][rho] + Bi[1][rho];
          const float X4 = A[1][rho] - A[0][rho];
          // cos-weighted operand at k = v, sin-weighted at k = Kpad / 2 + v
          const long long orow = ((long long)cls * Mpad + b) * Kpad + v;
          const long long oim = ((long long)cls * Mpad + B + b) * Kpad + v;
          // (X2 and X3 are zero for rho = 0, 2: stored only where a GEMM tile reads them)
          const bool odd = rho & 1;
          X[orow] = X1;
          if (odd || straddle_row(b, B)) X[orow + Kpad / 2] = X2;
          if (odd || straddle_row(B + b, B)) X[oim] = X3;
          X[oim + Kpad / 2] = X4;
          mre = fmaxf(mre, fmaxf(fabsf(X1), fabsf(X2)));
          mim = fmaxf(mim, fmaxf(fabsf(X3), fabsf(X4)));
        }
      }
    }
    if (rowmax) {
      // consecutive t: a wave spans at most two particles (V0 >= 64)
      const int bfirst = __shfl(b, 0, 64);
      const bool first = b == bfirst;
      float r0 = first ? mre : 0.f, i0 = first ? mim : 0.f;
      float r1 = first ? 0.f : mre, i1 = first ? 0.f : mim;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        r0 = fmaxf(r0, __shfl_xor(r0, off, 64));
        i0 = fmaxf(i0, __shfl_xor(i0, off, 64));
        r1 = fmaxf(r1, __shfl_xor(r1, off, 64));
        i1 = fmaxf(i1, __shfl_xor(i1, off, 64));
      }
      // the second particle of the wave (if any) is the largest active b; lane 0
      // is active whenever any lane is (t grows with the lane)
      int blast = active ? b : bfirst;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) blast = max(blast, __shfl_xor(blast, off, 64));
      if ((threadIdx.x & 63) == 0 && active) {
        atomicMax(reinterpret_cast<unsigned*>(rowmax) + bfirst, __float_as_uint(r0));
        atomicMax(reinterpret_cast<unsigned*>(rowmax) + B + bfirst, __float_as_uint(i0));
        if (blast != bfirst) {
          atomicMax(reinterpret_cast<unsigned*>(rowmax) + blast, __float_as_uint(r1));
          atomicMax(reinterpret_cast<unsigned*>(rowmax) + B + blast, __float_as_uint(i1));
        }
      }
    }
  }
}

// BF16x3 path: the same combinations for 4 fundamental voxels per thread, 256
// apart.  Their 64 orbit values are requested at once (indices from the
// packed coordinates: image p + 4q + 8t of (x, y, z) is R^p My^t Mz^q, R the
// quarter turn (x, y) -> (-y, x)), and each row's values leave as 2-byte
// stores of hi and of lo, 128 contiguous bytes per wave.  Columns v >= V0 of
// the half's K padding receive zeros.
template <bool SHIFTED>
__global__ __launch_bounds__(256) void fle_symmetrize_bf16_kernel(
    const float* __restrict__ vols, long long ld_vol, int B, const int* __restrict__ fcoord, int V0, int N,
    int Kpad, int Mpad, const float* __restrict__ shift_zyx, uint16_t* __restrict__ Xhi,
    uint16_t* __restrict__ Xlo) {
  // a workgroup takes 1024 consecutive voxels of one particle, its thread l
  // voxels l, l + 256, l + 512, l + 768: each load instruction of a wave reads
  // 64 consecutive voxels (the z runs of two or three columns), whole cache
  // lines of each image
  const int nq = ((V0 + 1023) >> 10) << 8;  // threads per particle
  const long long tot = (long long)B * nq;
  const int c = N >> 1;
  const int KH = Kpad >> 1;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(t / nq), qd = (int)(t - (long long)b * nq);
    const int v0 = ((qd >> 8) << 10) + (qd & 255);  // voxel of u = 0; u adds 256 u
    const float* f = vols + (long long)b * ld_vol;
    // the patch vol[a + s] at the particle's shift (SHIFTED; values as
    // extract_shifted_patches copies them out)
    ShiftSpec sp;
    make_shift_spec(SHIFTED ? shift_zyx : nullptr, b, N, sp);
    int pcs[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int v = v0 + 256 * u;
      pcs[u] = fcoord[v < V0 ? v : 0];
    }
    // an integral shift under which every image of the thread's voxels stays in
    // the volume reads them directly at an offset; otherwise through the buffer
    // range check (zero outside), or trilinear for a fractional shift
    // (and of the patch: a ball of radius >= N/2 reaches past the array,
    // where images are zero)
    int m = 0, mz = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      m = max(m, max(pcs[u] & 1023, (pcs[u] >> 10) & 1023));
      mz = max(mz, pcs[u] >> 20);
    }
    const bool in_patch = c - m >= 0 && c + m < N && c - mz >= 0 && c + mz < N;
    int shoff = 0;
    bool direct = !SHIFTED && in_patch;
    if (SHIFTED && sp.integral) {
      direct = in_patch && c - m + sp.si[0] >= 0 && c + m + sp.si[0] < N && c - m + sp.si[1] >= 0 &&
               c + m + sp.si[1] < N && c - mz + sp.si[2] >= 0 && c + mz + sp.si[2] < N;
      shoff = (sp.si[0] * N + sp.si[1]) * N + sp.si[2];
    }
    const __amdgpu_buffer_rsrc_t rv = make_rsrc(f, (unsigned)(N * N * N) * 4u);
    float g[4][16];
    auto fill = [&](auto&& value_at) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const bool ok = v0 + 256 * u < V0;
        const int x = pcs[u] & 1023, y = (pcs[u] >> 10) & 1023, z = pcs[u] >> 20;
#pragma unroll
        for (int tt = 0; tt < 2; ++tt)
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const int zi = q ? c - z : c + z;
            const int y1 = tt ? -y : y;
#pragma unroll
            for (int p = 0; p < 4; ++p) {
              const int a0 = (p == 0 ? x : (p == 1 ? -y1 : (p == 2 ? -x : y1))) + c;
              const int a1 = (p == 0 ? y1 : (p == 1 ? x : (p == 2 ? -y1 : -x))) + c;
              const float val = value_at(a0, a1, zi);
              g[u][p + 4 * q + 8 * tt] = ok ? val : 0.f;
            }
          }
      }
    };
    if (direct) {
      fill([&](int a0, int a1, int a2) { return f[(a0 * N + a1) * N + a2 + shoff]; });
    } else if (sp.integral) {
      fill([&](int a0, int a1, int a2) {
        const int i0 = a0 + sp.si[0], i1 = a1 + sp.si[1], i2 = a2 + sp.si[2];
        const bool in = (unsigned)i0 < (unsigned)N && (unsigned)i1 < (unsigned)N && (unsigned)i2 < (unsigned)N &&
                        (unsigned)a0 < (unsigned)N && (unsigned)a1 < (unsigned)N && (unsigned)a2 < (unsigned)N;
        return load_f1(rv, in ? (unsigned)((i0 * N + i1) * N + i2) * 4u : kOutOfRange);
      });
    } else {
      fill([&](int a0, int a1, int a2) {
        const bool in = (unsigned)a0 < (unsigned)N && (unsigned)a1 < (unsigned)N && (unsigned)a2 < (unsigned)N;
        return in ? patch_value(f, N, sp, a0, a1, a2) : 0.f;
      });
    }
    // class (eps, rho): G_t = sum_q (-1)^{q eps} sum_p (-i)^{rho p} f[p, q, t]
    // (rho = 3 reads the rho = 1 slab)
#pragma unroll
    for (int eps = 0; eps < 2; ++eps)
#pragma unroll
      for (int rho = 0; rho < 3; ++rho) {
        uint16_t hx[4][4], lx[4][4];  // [X1..X4][voxel u] hi and lo
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          float A[2], Bv[2];
#pragma unroll
          for (int tt = 0; tt < 2; ++tt) {
            float a = 0.f, bi = 0.f;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
              const float sg = (eps && q) ? -1.f : 1.f;
              const float* h = &g[u][4 * q + 8 * tt];
              if (rho == 0) {
                a += sg * (h[0] + h[1] + h[2] + h[3]);
              } else if (rho == 2) {
                a += sg * (h[0] - h[1] + h[2] - h[3]);
              } else {
                a += sg * (h[0] - h[2]);
                bi += sg * (h[3] - h[1]);
              }
            }
            A[tt] = a;
            Bv[tt] = rho == 3 ? -bi : bi;  // rho = 3: conjugate of rho = 1
          }
          Split16<false>::split(A[0] + A[1], hx[0][u], lx[0][u]);
          Split16<false>::split(Bv[0] - Bv[1], hx[1][u], lx[1][u]);
          Split16<false>::split(Bv[0] + Bv[1], hx[2][u], lx[2][u]);
          Split16<false>::split(A[1] - A[0], hx[3][u], lx[3][u]);
        }
        // the cos-weighted operand of voxel v at k = v, the sin-weighted one at
        // k = Kpad / 2 + v; voxels V0 <= v < Kpad / 2 get zeros
        const int cls = eps * 4 + rho;
        const long long orow = ((long long)cls * Mpad + b) * Kpad + v0;
        const long long oim = ((long long)cls * Mpad + B + b) * Kpad + v0;
        // (X2 and X3 are zero for rho = 0, 2: stored only where a GEMM tile reads them)
        const bool odd = rho & 1;
        const bool st1 = odd || straddle_row(b, B), st2 = odd || straddle_row(B + b, B);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (v0 + 256 * u >= KH) break;
          const int du = 256 * u;
          Xhi[orow + du] = hx[0][u];
          Xlo[orow + du] = lx[0][u];
          if (st1) {
            Xhi[orow + KH + du] = hx[1][u];
            Xlo[orow + KH + du] = lx[1][u];
          }
          if (st2) {
            Xhi[oim + du] = hx[2][u];
            Xlo[oim + du] = lx[2][u];
          }
          Xhi[oim + KH + du] = hx[3][u];
          Xlo[oim + KH + du] = lx[3][u];
        }
      }
  }
}

// The patch V_t of each particle at a continuous shift t (the Nv^3 volume at
// a + Nv/2 - N/2 + t, a = (z, y, x) in the N^3 box, trilinear, zero outside
// the volume) and its three t-derivatives, exact inside each cell (on a cell
// face the one-sided derivative of the cell above), formed per orbit image and
// symmetrised straight into operand rows j Bp + b of a batch of 4 Bp volumes
// (j = 0: V_t; j = 1, 2, 3: d/dt_0, d/dt_1, d/dt_2; t_0 moves x, t_2 moves z):
// the four volumes never go through HBM.  FP32 operand (BF16 = false, the
// layout of fle_symmetrize_kernel) or BF16 hi/lo (that of
// fle_symmetrize_bf16_kernel, whose K padding columns receive zeros); one
// thread per (particle, column v).
template <bool BF16>
__global__ __launch_bounds__(256) void fle_symmetrize_tangent_kernel(
    const float* __restrict__ vols, long long ld_vol, int Nv, int Bp, const int* __restrict__ orbit, int V0,
    int N, int Kpad, int Mpad, const float* __restrict__ tsh, float* __restrict__ X, uint16_t* __restrict__ Xhi,
    uint16_t* __restrict__ Xlo) {
  const int KH = Kpad >> 1, Bv = 4 * Bp;
  const int nv = BF16 ? KH : V0;
  const long long tot = (long long)Bp * nv;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(t / nv), v = (int)(t - (long long)b * nv);
    float g[4][16];
    if (v < V0) {
      // corners through the buffer range check (zero outside the volume),
      // the orbit's 16 indices read at once; the image loop fully unrolled, so
      // that g stays in registers
      const __amdgpu_buffer_rsrc_t rv = make_rsrc(vols + (long long)b * ld_vol, (unsigned)Nv * Nv * Nv * 4u);
      const float base = (float)(Nv / 2 - N / 2);
      const float t0 = tsh[b * 3 + 0], t1 = tsh[b * 3 + 1], t2 = tsh[b * 3 + 2];
      int ids[16];
      const int4* o4 = reinterpret_cast<const int4*>(orbit + (long long)v * 16);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int4 q = o4[i];
        ids[4 * i] = q.x; ids[4 * i + 1] = q.y; ids[4 * i + 2] = q.z; ids[4 * i + 3] = q.w;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int id = ids[i];  // < 0: an image outside the box (radius >= N/2), zero
        const int z = id / (N * N), y = (id / N) % N, x = id % N;
        const float px = (float)x + base + t0, py = (float)y + base + t1, pz = (float)z + base + t2;
        const float fx0 = floorf(px), fy0 = floorf(py), fz0 = floorf(pz);
        const int x0 = (int)fx0, y0 = (int)fy0, z0 = (int)fz0;
        const float fx = px - fx0, fy = py - fy0, fz = pz - fz0;
        float c[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int xi = x0 + (k & 1), yi = y0 + ((k >> 1) & 1), zi = z0 + (k >> 2);
          const bool ok = id >= 0 && (unsigned)xi < (unsigned)Nv && (unsigned)yi < (unsigned)Nv &&
                          (unsigned)zi < (unsigned)Nv;
          c[k] = load_f1(rv, ok ? (unsigned)((zi * Nv + yi) * Nv + xi) * 4u : kOutOfRange);
        }
        const float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
        // c[k]: k bit 0 = x, bit 1 = y, bit 2 = z
        g[0][i] = gz * (gy * (gx * c[0] + fx * c[1]) + fy * (gx * c[2] + fx * c[3])) +
                  fz * (gy * (gx * c[4] + fx * c[5]) + fy * (gx * c[6] + fx * c[7]));
        g[1][i] = gz * (gy * (c[1] - c[0]) + fy * (c[3] - c[2])) + fz * (gy * (c[5] - c[4]) + fy * (c[7] - c[6]));
        g[2][i] = gz * (gx * (c[2] - c[0]) + fx * (c[3] - c[1])) + fz * (gx * (c[6] - c[4]) + fx * (c[7] - c[5]));
        g[3][i] = gy * (gx * (c[4] - c[0]) + fx * (c[5] - c[1])) + fy * (gx * (c[6] - c[2]) + fx * (c[7] - c[3]));
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 16; ++i) g[j][i] = 0.f;
    }
    auto store = [&](long long idx, float val) {
      if constexpr (BF16) {
        uint16_t h, l;
        Split16<false>::split(val, h, l);
        Xhi[idx] = h;
        Xlo[idx] = l;
      } else {
        X[idx] = val;
      }
    };
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int vb = j * Bp + b;
      // the class combinations of fle_symmetrize_kernel; g[j][p + 4q + 8t]
#pragma unroll
      for (int eps = 0; eps < 2; ++eps) {
        float A[2][3], Bi[2][3];  // [t][rho]
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
          float s0 = 0.f, s2 = 0.f, s1r = 0.f, s1i = 0.f;
#pragma unroll
          for (int q = 0; q < 2; ++q) {
            const float sg = (eps && q) ? -1.f : 1.f;
            const float* h = &g[j][4 * q + 8 * tt];
            s0 += sg * (h[0] + h[1] + h[2] + h[3]);
            s2 += sg * (h[0] - h[1] + h[2] - h[3]);
            s1r += sg * (h[0] - h[2]);
            s1i += sg * (h[3] - h[1]);
          }
          A[tt][0] = s0;  Bi[tt][0] = 0.f;
          A[tt][1] = s1r; Bi[tt][1] = s1i;
          A[tt][2] = s2;  Bi[tt][2] = 0.f;
        }
#pragma unroll
        for (int rho = 0; rho < 3; ++rho) {  // rho = 3 reads the rho = 1 slab
          const int cls = eps * 4 + rho;
          const long long orow = ((long long)cls * Mpad + vb) * Kpad + v;
          const long long oim = ((long long)cls * Mpad + Bv + vb) * Kpad + v;
          const bool odd = rho & 1;
          store(orow, A[0][rho] + A[1][rho]);
          if (odd || straddle_row(vb, Bv)) store(orow + KH, Bi[0][rho] - Bi[1][rho]);
          if (odd || straddle_row(Bv + vb, Bv)) store(oim, Bi[0][rho] + Bi[1][rho]);
          store(oim + KH, A[1][rho] - A[0][rho]);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 2. grouped FP32 MFMA GEMM with split-K
// ---------------------------------------------------------------------------
// grid: x = column tile (over all classes), y = row tile, z = K split.
// Each of the 4 waves computes 32 rows x 64 columns as 2 x 4 blocks of 16x16.
__global__ __launch_bounds__(kGemmThreads) void fle_gemm_kernel(
    const float* __restrict__ X, const float* __restrict__ W, float* __restrict__ part,
    int Mpad, int Kpad, int Ntot, int kchunk, int B, FlePlan plan_cols) {
  __shared__ float As[2][BK][BM + 4];
  __shared__ float Bs[2][BK][BN + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col0 = blockIdx.x * BN;
  const int row0 = blockIdx.y * BM;
  const int cls = column_class(plan_cols.col_off, col0);
  const float* A = X + (long long)x_slab(cls) * Mpad * Kpad;
  int klo, khi;
  k_range(cls, row0, BM, B, Kpad, klo, khi);
  const int kbeg = max(blockIdx.z * kchunk, klo);
  int kend = min(blockIdx.z * kchunk + kchunk, khi);

  floatx4 acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  // per-thread load slots: A tile 128 x 16 = 512 float4 (2 per thread),
  // B tile 16 x 64 = 256 float4 (1 per thread)
  const int ar0 = tid >> 2, ac = (tid & 3) * 4;          // rows ar0 and ar0 + 64
  const int br = tid >> 4, bc = (tid & 15) * 4;
  float4 ra0, ra1, rb;
  auto gload = [&](int k0) {
    ra0 = *reinterpret_cast<const float4*>(A + (long long)(row0 + ar0) * Kpad + k0 + ac);
    ra1 = *reinterpret_cast<const float4*>(A + (long long)(row0 + ar0 + 64) * Kpad + k0 + ac);
    rb = *reinterpret_cast<const float4*>(W + (long long)(k0 + br) * Ntot + col0 + bc);
  };
  auto sstore = [&](int buf) {
    As[buf][ac + 0][ar0] = ra0.x; As[buf][ac + 1][ar0] = ra0.y;
    As[buf][ac + 2][ar0] = ra0.z; As[buf][ac + 3][ar0] = ra0.w;
    As[buf][ac + 0][ar0 + 64] = ra1.x; As[buf][ac + 1][ar0 + 64] = ra1.y;
    As[buf][ac + 2][ar0 + 64] = ra1.z; As[buf][ac + 3][ar0 + 64] = ra1.w;
    *reinterpret_cast<float4*>(&Bs[buf][br][bc]) = rb;
  };
  int buf = 0;
  if (kbeg < kend) {
    gload(kbeg);
    sstore(0);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = (k0 + BK) < kend;
    if (more) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kk = ks * 4 + (lane >> 4);
      float a[2], bv[4];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[buf][kk][wave * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[buf][kk][j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], bv[j], acc[i][j], 0, 0, 0);
    }
    if (more) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // D layout of 16x16x4: lane holds rows 4*(lane/16)+r, column lane%16
  float* P = part + (long long)blockIdx.z * Mpad * Ntot;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + wave * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int col = col0 + j * 16 + (lane & 15);
        P[(long long)row * Ntot + col] = acc[i][j][r];
      }
}

// ---------------------------------------------------------------------------
// 2b. tuned path: 16-bit split GEMM, C = Xhi Whi + Xhi Wlo + Xlo Whi (FP32 acc)
// ---------------------------------------------------------------------------
// 1-D grid over (column tile, row tile, K split), XCD-aware so that the column
// tiles of one class and K range (which read the same X rows) share one L2.
// Tile 128 rows x 128 columns x 32 k per stage; a wave multiplies 64 columns
// (4 column blocks) by 64 rows (4 waves, 2 x 2) or by all 128 rows (2 waves),
// which bounds the LDS fragment reads per MFMA.  Both operands arrive by LDS-DMA
// (global_load_lds_dwordx4: lane i's 16 bytes land at slot + 16 i), laid out
// in MFMA fragment order, so every lane reads its own 16-byte fragment back:
//   X hi/lo: 8 row blocks x 64 lanes (row rb*16 + lane%16, k 8*(lane/16)),
//            shared by the 4 waves (each wave fetches 2 row blocks);
//   W^T hi/lo: 4 groups of 32 columns x 4 parts x 64 lanes, wave w fetching
//            4 / NW groups.
// NST stages rotate through LDS: while stage st is multiplied, stages
// st + 1 .. st + NST - 1 are in flight (HBM latency x bandwidth), and no
// operand passes through registers on its way in.
// 128-row tiles per workgroup (8 waves: a 256-row tile) and u32x4 per stage
// (A hi and lo 512 each per 128 rows, W^T 4 groups x 4 parts x 64)
constexpr int gemm16_row_tiles(int nw) { return nw == 8 ? 2 : 1; }
constexpr int gemm16_stage_u4(int nw) { return 1024 * gemm16_row_tiles(nw) + 1024; }

// s_waitcnt immediates (gfx9 encoding: vmcnt bits [3:0] and [15:14], expcnt
// [6:4], lgkmcnt [11:8]); the counters not named are left unconstrained
constexpr int waitcnt_vm(int n) { return (n & 0xF) | ((n >> 4) << 14) | (7 << 4) | (0xF << 8); }

// Execution barrier of the workgroup without the memory fence of
// __syncthreads (which waits for every outstanding load, LDS-DMA included).
// Callers make their own data visible first: an explicit vmcnt wait for their
// LDS-DMA, and LDS reads that have been consumed.  The empty asm statements
// keep the compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void wg_barrier_keep_loads() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// 16 B per lane into LDS at lds_wave_base + 16 * lane.  CPOL is the cache
// policy: 0 = default, 2 = non-temporal (the streamed W^T, read once, so that
// it does not evict the X slices the neighbouring column tiles re-read).
template <int CPOL = 0>
__device__ __forceinline__ void lds_dma16(const void* g, u32x4* lds_wave_base) {
  __builtin_amdgcn_global_load_lds(g, (__attribute__((address_space(3))) void*)lds_wave_base, 16, 0, CPOL);
}

template <bool FP16, int NST, int NW>
__global__ __launch_bounds__(64 * NW) void fle_gemm16_kernel(
    const uint16_t* __restrict__ Xhi, const uint16_t* __restrict__ Xlo, const u32x4* __restrict__ Wt16,
    float* __restrict__ part, int Mpad, int Kpad, int Ntot, int kchunk, int ntn, int nmt, int B, FlePlan cols) {
  static_assert(NST >= 2 && NST <= 3, "pipeline depth");
  static_assert(NW == 2 || NW == 4 || NW == 8, "waves per workgroup");
  constexpr int RT = gemm16_row_tiles(NW);  // 128-row tiles per workgroup (8 waves: 2)
  constexpr int AREG = 512 * RT;            // u32x4 of the A hi (or lo) part of a stage
  constexpr int W0 = 2 * AREG;              // start of the W^T part of a stage
  constexpr int AP = 16 * RT / NW;          // A (row block, hi/lo) parts fetched per wave
  constexpr int WP = 16 / NW;               // W^T parts (of 16) fetched per wave
  // compute layout: every wave multiplies 64 columns (4 column blocks) by 64
  // rows (4 waves as 2 x 2, 8 waves as 4 x 2 over 256 rows) or by all 128 rows
  // (2 waves)
  constexpr int RB = NW == 2 ? 8 : 4;  // 16-row blocks per wave
  constexpr int CB = 4;                // 16-column blocks per wave
  extern __shared__ u32x4 lds[];   // NST x gemm16_stage_u4(NW)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lb = mch::xcd_remap(blockIdx.x, gridDim.x);
  // row tile fastest: the row tiles that share a W^T tile run side by side on
  // one XCD, so W^T is fetched from HBM once for all of them
  const int mt = lb % nmt, rest = lb / nmt;
  const int tn = rest % ntn, ks = rest / ntn;
  const int col0 = tn * BN16, row0 = mt * BM16 * RT;
  const int cls = column_class(cols.col_off, col0);
  int klo, khi;
  k_range(cls, row0, BM16 * RT, B, Kpad, klo, khi);
  const int kbeg = max(ks * kchunk, klo);
  const int kend = min(ks * kchunk + kchunk, khi);
  const int nstage = kend > kbeg ? (kend - kbeg) / BK16 : 0;

  // this lane's DMA sources: A parts (row block rb, hi/lo) of the row-major X,
  // and this wave's WP of the 16 1-KB parts of the DMA-ordered W^T per stage
  const long long xbase = ((long long)x_slab(cls) * Mpad + row0) * Kpad + (lane >> 4) * 8;
  const uint16_t* asrc[AP];
  int adst[AP];
#pragma unroll
  for (int t = 0; t < AP; ++t) {
    const int part = wave * AP + t;            // row block part / 2, hi/lo part % 2
    const int rb = part >> 1, hl = part & 1;
    asrc[t] = (hl ? Xlo : Xhi) + xbase + (long long)(rb * 16 + (lane & 15)) * Kpad;
    adst[t] = hl * AREG + rb * 64;
  }
  const int nK = Kpad / BK16;
  const u32x4* wsrc = Wt16 + ((long long)tn * nK + kbeg / BK16) * 1024 + wave * WP * 64 + lane;
  // stage layout (u32x4): A hi row block rb at rb*64, A lo at AREG + rb*64,
  // W of 32-column group g, part q = {j * 2 + hi/lo} at W0 + g*256 + q*64
  auto issue = [&](int stage) {
    u32x4* base = lds + (stage % NST) * gemm16_stage_u4(NW);
    const int k0 = kbeg + stage * BK16;
#pragma unroll
    for (int t = 0; t < AP; ++t) lds_dma16(asrc[t] + k0, base + adst[t]);
    // one row tile: W^T is read once, stream it past L2 (non-temporal);
    // several: keep it cached for the neighbouring row tiles
    if (nmt == 1) {
#pragma unroll
      for (int q = 0; q < WP; ++q)
        lds_dma16<2>(wsrc + (long long)stage * 1024 + q * 64, base + W0 + (wave * WP + q) * 64);
    } else {
#pragma unroll
      for (int q = 0; q < WP; ++q)
        lds_dma16<0>(wsrc + (long long)stage * 1024 + q * 64, base + W0 + (wave * WP + q) * 64);
    }
  };
  constexpr int PER_STAGE = AP + WP;  // DMAs per wave per stage
  const int rb0 = NW == 8 ? RB * (wave & 3) : (NW == 4 ? RB * (wave & 1) : 0);  // first row block
  const int chalf = NW == 8 ? (wave >> 2) : (NW == 4 ? (wave >> 1) : wave);   // 64-column half
  floatx4 acc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[i][c] = floatx4{0.f, 0.f, 0.f, 0.f};

#pragma unroll
  for (int s = 0; s < NST - 1; ++s)
    if (s < nstage) issue(s);
  for (int st = 0; st < nstage; ++st) {
    // this wave's DMAs of stage st are done once at most those of the later
    // stages already issued are pending
    const int later = min(st + NST - 2, nstage - 1) - st;
    if (NST == 3 && later >= 1)
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(PER_STAGE));
    else
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    // every wave's part of stage st has landed, and every wave has finished
    // stage st - 1, whose slot is refilled next
    wg_barrier_keep_loads();
    if (st + NST - 1 < nstage) issue(st + NST - 1);
    const u32x4* base = lds + (st % NST) * gemm16_stage_u4(NW);
    // W^T parts of this wave's two 32-column groups (fetched by other waves too)
    // every fragment of the stage is requested before the first MFMA, so that
    // the LDS reads overlap the products instead of preceding each row block
    u32x4 w[2 * CB], ah[RB], al[RB];
#pragma unroll
    for (int q = 0; q < 2 * CB; ++q) w[q] = base[W0 + chalf * 512 + q * 64 + lane];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      ah[i] = base[(rb0 + i) * 64 + lane];
      al[i] = base[AREG + (rb0 + i) * 64 + lane];
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the products
#pragma unroll
    for (int i = 0; i < RB; ++i) {
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        // column block c = group c / 2, block c % 2: parts q = 4 (c / 2) + 2 (c % 2) + {0, 1}
        acc[i][c] = Split16<FP16>::mfma(ah[i], w[2 * c + 0], acc[i][c]);
        acc[i][c] = Split16<FP16>::mfma(ah[i], w[2 * c + 1], acc[i][c]);
        acc[i][c] = Split16<FP16>::mfma(al[i], w[2 * c + 0], acc[i][c]);
      }
    }
  }
  // D layout of the 16x16 MFMAs: lane holds rows 4 (lane / 16) + r, column lane % 16
  float* P = part + (long long)ks * Mpad * Ntot;
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = row0 + (rb0 + i) * 16 + 4 * (lane >> 4) + r;
        const int col = col0 + chalf * 64 + c * 16 + (lane & 15);
        P[(long long)row * Ntot + col] = acc[i][c][r];
      }
}

// Wide variant: 256 rows x 256 columns (two 128-column tiles of one class) per
// workgroup of 8 waves, two stages of 64 KB.  X is re-read from L2 once per
// 256 columns instead of once per 128 (the operand traffic into LDS, not the
// MFMAs, bounds the 256 x 128 tile at large batches).  Waves: 4 row quarters
// x 2 column halves, 64 x 128 per wave.  The work items are (K split, column
// unit, row tile) triples whose K chunk meets the row tile's K range
// (k_range: the real or imaginary row tiles of the classes m = 0, 2 mod 4
// have empty chunks), numbered K split first, row tile fastest; a column unit
// is a pair of tiles within a class, or the single last tile of a class with
// an odd count, which reads its tile twice and stores the first half only.
// The grid is persistent (at most one workgroup per resident slot) and
// workgroup g takes items g, g + grid, ...: the hardware places workgroups on
// shader engines without knowing which are empty, so a grid with empty items
// can queue busy ones behind a full engine while other engines idle.  Empty
// items store nothing; the finalize skips their partial slices.
constexpr int kWideStageU4 = 4096;  // A hi 1024, A lo 1024, W 2 tiles x 1024
constexpr int kWideMaxSplits = 16, kWideMaxRowTiles = 64;

// The work items of one wide launch (host-built, passed by value).
struct WideSched {
  int nitems;
  int splits;
  int kchunk;
  int nu[kClasses];                                  // column units per class
  int per_split[kWideMaxSplits];                     // items of each K split
  unsigned long long busy[2][kWideMaxSplits];        // row tiles with work, by class parity and split
};

template <bool FP16>
__global__ __launch_bounds__(512) void fle_gemm16w_kernel(
    const uint16_t* __restrict__ Xhi, const uint16_t* __restrict__ Xlo, const u32x4* __restrict__ Wt16,
    float* __restrict__ part, int Mpad, int Kpad, int Ntot, int B, WideSched sched, FlePlan cols) {
  constexpr int NST = 2, NW = 8, AREG = 1024, W0 = 2048, AP = 4, WP = 4, RB = 4, CB = 8;
  extern __shared__ u32x4 lds[];   // NST x kWideStageU4
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb0 = RB * (wave & 3);   // first row block of this wave
  const int chalf = wave >> 2;       // 128-column half of the unit
  const int nK = Kpad / BK16;
  for (int it = mch::xcd_remap(blockIdx.x, gridDim.x); it < sched.nitems; it += gridDim.x) {
    // item -> K split, class, column unit, row tile (the j-th busy one)
    int r = it, ks = 0;
    while (ks < sched.splits - 1 && r >= sched.per_split[ks]) r -= sched.per_split[ks++];
    int cls = 0, ubase = 0, cnt = 0;
    bool found = false;
    for (; cls < kClasses; ++cls) {
      cnt = __popcll(sched.busy[cls & 1][ks]);
      if (r < sched.nu[cls] * cnt) { found = true; break; }
      r -= sched.nu[cls] * cnt;
      ubase += sched.nu[cls];
    }
    if (!found) continue;  // not reached: the host counts the same items (uniform per workgroup)
    const int un = ubase + r / cnt;
    int j = r - (r / cnt) * cnt;
    unsigned long long m = sched.busy[cls & 1][ks];
    for (; j > 0; --j) m &= m - 1;
    const int mt = __ffsll((long long)m) - 1;
    const int col0 = cols.col_off[cls] + (un - ubase) * 2 * BN16;
    const int ntile = min(2, (cols.col_off[cls + 1] - col0) / BN16);
    const int tn0 = col0 / BN16, tn1 = ntile == 2 ? tn0 + 1 : tn0;
    const int row0 = mt * 2 * BM16;
    int klo, khi;
    k_range(cls, row0, 2 * BM16, B, Kpad, klo, khi);
    const int kbeg = max(ks * sched.kchunk, klo);
    const int kend = min(ks * sched.kchunk + sched.kchunk, khi);
    const int nstage = kend > kbeg ? (kend - kbeg) / BK16 : 0;

    const long long xbase = ((long long)x_slab(cls) * Mpad + row0) * Kpad + (lane >> 4) * 8;
    const uint16_t* asrc[AP];
    int adst[AP];
#pragma unroll
    for (int t = 0; t < AP; ++t) {
      const int part = wave * AP + t;            // row block part / 2, hi/lo part % 2
      const int rb = part >> 1, hl = part & 1;
      asrc[t] = (hl ? Xlo : Xhi) + xbase + (long long)(rb * 16 + (lane & 15)) * Kpad;
      adst[t] = hl * AREG + rb * 64;
    }
    // W^T parts P = 4 wave .. 4 wave + 3 of the unit's 32: tile P / 16, part P % 16
    const u32x4* wsrc[WP];
#pragma unroll
    for (int q = 0; q < WP; ++q) {
      const int P = wave * WP + q;
      const int tn = (P >> 4) ? tn1 : tn0;
      wsrc[q] = Wt16 + ((long long)tn * nK + kbeg / BK16) * 1024 + (P & 15) * 64 + lane;
    }
    auto issue = [&](int stage) {
      u32x4* base = lds + (stage % NST) * kWideStageU4;
      const int k0 = kbeg + stage * BK16;
#pragma unroll
      for (int t = 0; t < AP; ++t) lds_dma16(asrc[t] + k0, base + adst[t]);
#pragma unroll
      for (int q = 0; q < WP; ++q)
        lds_dma16<0>(wsrc[q] + (long long)stage * 1024, base + W0 + (wave * WP + q) * 64);
    };
    floatx4 acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[i][c] = floatx4{0.f, 0.f, 0.f, 0.f};

    if (nstage > 0) issue(0);
    for (int st = 0; st < nstage; ++st) {
      // this wave's DMAs of stage st have landed; after the barrier every
      // wave's have, and every wave has finished stage st - 1, whose slot the
      // next issue refills
      __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
      wg_barrier_keep_loads();
      if (st + 1 < nstage) issue(st + 1);
      const u32x4* base = lds + (st % NST) * kWideStageU4;
      u32x4 w[2 * CB], ah[RB], al[RB];
#pragma unroll
      for (int q = 0; q < 2 * CB; ++q) w[q] = base[W0 + chalf * 1024 + q * 64 + lane];
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        ah[i] = base[(rb0 + i) * 64 + lane];
        al[i] = base[AREG + (rb0 + i) * 64 + lane];
      }
      __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the products
#pragma unroll
      for (int i = 0; i < RB; ++i) {
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          // column block c = group c / 2, block c % 2: parts 4 (c / 2) + 2 (c % 2) + {0, 1}
          acc[i][c] = Split16<FP16>::mfma(ah[i], w[2 * c + 0], acc[i][c]);
          acc[i][c] = Split16<FP16>::mfma(ah[i], w[2 * c + 1], acc[i][c]);
          acc[i][c] = Split16<FP16>::mfma(al[i], w[2 * c + 0], acc[i][c]);
        }
      }
    }
    if (chalf < ntile) {  // a single-tile unit: the second half is a copy
      float* P = part + (long long)ks * Mpad * Ntot;
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int c = 0; c < CB; ++c)
#pragma unroll
          for (int rr = 0; rr < 4; ++rr) {
            const int row = row0 + (rb0 + i) * 16 + 4 * (lane >> 4) + rr;
            const int col = col0 + chalf * BN16 + c * 16 + (lane & 15);
            P[(long long)row * Ntot + col] = acc[i][c][rr];
          }
    }
    // every wave has read its last stage before the next item refills the ring
    __syncthreads();
  }
}

// FP16 mode: X (FP32) -> Xhi/Xlo of X * rowscale (rows of every class)
__global__ void fle_xsplit_kernel(const float* __restrict__ X, const float* __restrict__ rowmax, int Mpad,
                                  int Kpad, int rows, uint16_t* __restrict__ Xhi, uint16_t* __restrict__ Xlo) {
  const long long n4 = (long long)kClasses * Mpad * Kpad / 4;
  for (long long q = blockIdx.x * (long long)blockDim.x + threadIdx.x; q < n4;
       q += (long long)gridDim.x * blockDim.x) {
    const long long e = q * 4;
    const int row = (int)((e / Kpad) % Mpad);
    const int cls = (int)(e / ((long long)Mpad * Kpad));
    if (row >= rows || x_slab(cls) != cls) continue;
    const float rs = pow2_scale(rowmax[row]);
    const float4 v = reinterpret_cast<const float4*>(X)[q];
    uint16_t h[4], l[4];
    Split16<true>::split(v.x * rs, h[0], l[0]);
    Split16<true>::split(v.y * rs, h[1], l[1]);
    Split16<true>::split(v.z * rs, h[2], l[2]);
    Split16<true>::split(v.w * rs, h[3], l[3]);
    reinterpret_cast<uint2*>(Xhi)[q] = make_uint2((unsigned)h[0] | ((unsigned)h[1] << 16),
                                                  (unsigned)h[2] | ((unsigned)h[3] << 16));
    reinterpret_cast<uint2*>(Xlo)[q] = make_uint2((unsigned)l[0] | ((unsigned)l[1] << 16),
                                                  (unsigned)l[2] | ((unsigned)l[3] << 16));
  }
}

__global__ void fle_fill_kernel(float* __restrict__ p, int n, float v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// max |W| of each column over a K range (atomic over K ranges; values >= 0)
__global__ void fle_colmax_kernel(const float* __restrict__ W, int Kpad, int Ntot, int kchunk,
                                  float* __restrict__ colmax) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Ntot) return;
  const int k0 = blockIdx.y * kchunk, k1 = min(k0 + kchunk, Kpad);
  float m = 0.f;
  for (int k = k0; k < k1; ++k) m = fmaxf(m, fabsf(W[(long long)k * Ntot + c]));
  atomicMax(reinterpret_cast<unsigned*>(colmax) + c, __float_as_uint(m));
}

__global__ void fle_colscale_kernel(float* __restrict__ cs, int Ntot, int fp16) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < Ntot) cs[c] = fp16 ? pow2_scale(cs[c]) : 1.f;
}

// W (Kpad, Ntot) FP32 -> 16-bit hi/lo of colscale * W, stored in the exact
// order the GEMM's LDS-DMA streams it, in 16-byte chunks of 8 consecutive k:
//   chunk ((((tn * nK + S) * 4 + w) * 4 + j * 2 + h) * 64 + lane)
// for column tn * 128 + 32 w + 16 j + lane % 16, k = 32 S + 8 (lane / 16),
// h = 0 (hi) / 1 (lo), nK = Kpad / 32: each workgroup reads one contiguous
// 16 KB block per stage.  One workgroup per 64 x 64 (k, column) tile.
template <bool FP16>
__global__ __launch_bounds__(256) void fle_wsplit_kernel(const float* __restrict__ W, int Kpad, int Ntot,
                                                         const float* __restrict__ colscale,
                                                         u32x4* __restrict__ Wt16) {
  __shared__ float tile[64][65];
  const int c0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  for (int i = threadIdx.x; i < 64 * 64; i += 256) {
    const int kk = i >> 6, cc = i & 63;
    tile[kk][cc] = W[(long long)(k0 + kk) * Ntot + c0 + cc];
  }
  __syncthreads();
  const int nK = Kpad / BK16;
  // thread: column cc, 8 consecutive k (kg-th group of the 64)
  for (int t = threadIdx.x; t < 64 * 8; t += 256) {
    const int cc = t & 63, kg = t >> 6;
    const int col = c0 + cc, k = k0 + kg * 8;
    const float cs = colscale[col];
    u32x4 hv, lv;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      uint16_t h0, l0, h1, l1;
      Split16<FP16>::split(tile[kg * 8 + 2 * e][cc] * cs, h0, l0);
      Split16<FP16>::split(tile[kg * 8 + 2 * e + 1][cc] * cs, h1, l1);
      hv[e] = (unsigned)h0 | ((unsigned)h1 << 16);
      lv[e] = (unsigned)l0 | ((unsigned)l1 << 16);
    }
    const int tn = col / BN16, w = (col % BN16) / 32, j = (col % 32) / 16;
    const int S = k / BK16, lane = (col % 16) + 16 * ((k % BK16) / 8);
    const long long chunk = ((((long long)tn * nK + S) * 4 + w) * 4 + j * 2) * 64 + lane;
    Wt16[chunk] = hv;
    Wt16[chunk + 64] = lv;
  }
}

// ---------------------------------------------------------------------------
// 3. split-K reduction and scatter to the reference coefficient order
// ---------------------------------------------------------------------------
// colscale / rowmax (NULL: 1) undo the power-of-2 scales of the FP16 split GEMM.
// skip_kchunk > 0 (the wide GEMM, which stores no partial for an empty item):
// split s of a row counts only if its K chunk [s, s + 1) * skip_kchunk meets
// the K range of the row's 256-row tile.
__global__ void fle_finalize_kernel(const float* __restrict__ part, int splits, int Mpad, int Ntot,
                                    int B, const int* __restrict__ chan_plus,
                                    const int* __restrict__ chan_minus,
                                    const int* __restrict__ chan_m, float2* __restrict__ a,
                                    long long lda, const float* __restrict__ colscale,
                                    const float* __restrict__ rowmax, int4 flip, int skip_kchunk,
                                    FlePlan cols) {
  const long long tot = (long long)B * Ntot;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(t / Ntot);
    const int c = (int)(t - (long long)b * Ntot);
    const int ip = chan_plus[c];
    if (ip < 0) continue;
    float re = 0.f, im = 0.f;
    if (skip_kchunk > 0) {
      const int cls = column_class(cols.col_off, c);
      int klo_r, khi_r, klo_i, khi_i;
      k_range(cls, (b / 256) * 256, 256, B, cols.Kpad, klo_r, khi_r);
      k_range(cls, ((B + b) / 256) * 256, 256, B, cols.Kpad, klo_i, khi_i);
      for (int s = 0; s < splits; ++s) {
        const float* P = part + (long long)s * Mpad * Ntot;
        const int k0 = s * skip_kchunk, k1 = k0 + skip_kchunk;
        if (max(k0, klo_r) < min(k1, khi_r)) re += P[(long long)b * Ntot + c];
        if (max(k0, klo_i) < min(k1, khi_i)) im += P[(long long)(B + b) * Ntot + c];
      }
    } else {
      for (int s = 0; s < splits; ++s) {
        const float* P = part + (long long)s * Mpad * Ntot;
        re += P[(long long)b * Ntot + c];
        im += P[(long long)(B + b) * Ntot + c];
      }
    }
    if (colscale) {
      re /= colscale[c] * pow2_scale(rowmax[b]);
      im /= colscale[c] * pow2_scale(rowmax[B + b]);
    }
    // columns of the classes (eps, 3), [flip.x, flip.y) and [flip.z, flip.w):
    // their operand is the (eps, 1) slab, whose imaginary output is negated
    if ((c >= flip.x && c < flip.y) || (c >= flip.z && c < flip.w)) im = -im;
    a[(long long)b * lda + ip] = make_float2(re, im);
    const int imn = chan_minus[c];
    if (imn >= 0) {
      const float sg = (chan_m[c] & 1) ? -1.f : 1.f;
      a[(long long)b * lda + imn] = make_float2(sg * re, -sg * im);
    }
  }
}

// ---------------------------------------------------------------------------
// sigma and norms
// ---------------------------------------------------------------------------
// One workgroup per (l, particle): sigma[Dindex(l,u,v)] =
// (-1)^{u+v} sum_k aT[l,k,-v] conj(aV[l,k,-u]).
// One workgroup per (shell l, particle b): the shell's template and particle
// coefficients are gathered into LDS once (rows m = -l..l, kdim columns, zero
// past each row's radial count, so the k sum runs to the shorter row as the
// reference pairs them), then each thread forms 4 x 4 blocks of (u, v):
// sigma[Dindex(l,u,v)] = (-1)^{u+v} sum_k aT[l,k,-v] conj(aV[l,k,-u]).
__global__ __launch_bounds__(256) void sigma_kernel(const float2* __restrict__ aT,
                                                    const float2* __restrict__ aV, long long ldv,
                                                    const int* __restrict__ idx, int lmax,
                                                    int kdim, float2* __restrict__ sigma,
                                                    long long lds) {
  const int l = blockIdx.x;
  const int b = blockIdx.y;
  const int w = 2 * l + 1;
  extern __shared__ float2 shell[];
  float2* T = shell;              // (w, kdim): row r holds m = r - l
  float2* V = shell + w * kdim;
  const float2* av = aV + (long long)b * ldv;
  const int* il = idx + (long long)l * (2 * lmax + 1) * kdim;
  for (int q = threadIdx.x; q < w * kdim; q += blockDim.x) {
    const int j = il[q];
    T[q] = j >= 0 ? aT[j] : make_float2(0.f, 0.f);
    V[q] = j >= 0 ? av[j] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2* out = sigma + (long long)b * lds + mch::dsize(l - 1);
  const int nb = (w + 3) / 4;
  for (int blk = threadIdx.x; blk < nb * nb; blk += blockDim.x) {
    const int u0 = (blk / nb) * 4, v0 = (blk - (blk / nb) * nb) * 4;
    float2 acc[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[r][c] = make_float2(0.f, 0.f);
    // row of m = -u for u = u0 + r - l is 2l - (u0 + r); rows past the shell read row 0 (unused)
    int ru[4], rv[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      ru[r] = (u0 + r < w) ? (2 * l - u0 - r) * kdim : 0;
      rv[r] = (v0 + r < w) ? (2 * l - v0 - r) * kdim : 0;
    }
    for (int k = 0; k < kdim; ++k) {
      float2 p[4], t[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        p[r] = V[ru[r] + k];
        t[r] = T[rv[r] + k];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          // t * conj(p)
          acc[r][c].x += t[c].x * p[r].x + t[c].y * p[r].y;
          acc[r][c].y += t[c].y * p[r].x - t[c].x * p[r].y;
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (u0 + r >= w) continue;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (v0 + c >= w) continue;
        const float sg = ((u0 + r + v0 + c) & 1) ? -1.f : 1.f;
        out[(u0 + r) * w + v0 + c] = make_float2(sg * acc[r][c].x, sg * acc[r][c].y);
      }
    }
  }
}

__global__ __launch_bounds__(256) void norms_kernel(const float2* __restrict__ a, long long lda,
                                                    int ne, float* __restrict__ norms) {
  const int b = blockIdx.x;
  float s = 0.f;
  for (int i = threadIdx.x; i < ne; i += blockDim.x) {
    const float2 v = a[(long long)b * lda + i];
    s += v.x * v.x + v.y * v.y;
  }
  s = mch::wave_sum(s);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    norms[b] = sqrtf(t);
  }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
namespace {

void free_fle_plan(FlePlan* p) {
  if (!p) return;
  (void)hipFree(p->orbit); (void)hipFree(p->fcoord); (void)hipFree(p->W); (void)hipFree(p->chan_plus);
  (void)hipFree(p->chan_minus); (void)hipFree(p->chan_m); (void)hipFree(p->X); (void)hipFree(p->part);
  (void)hipFree(p->Wt16); (void)hipFree(p->colscale); (void)hipFree(p->rowmax);
  (void)hipFree(p->part16); (void)hipFree(p->Xhi); (void)hipFree(p->Xlo);
  delete p;
}

// Autotune candidates of the 16-bit GEMM: (pipeline depth, waves per
// workgroup, rounds of resident workgroups the K split fills).  8 waves hold a
// 256-row tile (candidates only when the plan's rows come in 256s).
// rounds: the K split fills 1, 2 or 3 full rounds of resident workgroups, or
// (0, wide tile only) the split of the fewest k-stages per workgroup under the
// wide kernel's busy-item schedule (gemm16_splits)
struct Gemm16Cand { int nst, nw, rounds, ct; };
constexpr Gemm16Cand kGemm16Cands[] = {
    {2, 4, 1, 1}, {3, 4, 1, 1}, {2, 2, 1, 1}, {3, 2, 1, 1}, {2, 4, 2, 1}, {3, 4, 2, 1}, {2, 2, 2, 1},
    {3, 2, 2, 1}, {2, 4, 3, 1}, {3, 4, 3, 1}, {2, 2, 3, 1}, {3, 2, 3, 1}, {2, 8, 1, 1}, {3, 8, 1, 1},
    {2, 8, 2, 1}, {3, 8, 2, 1}, {2, 8, 1, 2}, {2, 8, 2, 2}, {2, 8, 3, 2}, {2, 8, 0, 2}};
constexpr int kGemm16Candidates = (int)(sizeof(kGemm16Cands) / sizeof(kGemm16Cands[0]));
inline bool gemm16_usable(const FlePlan* p, int cand) {
  if (kGemm16Cands[cand].ct == 2 && p->Mpad / (2 * BM16) > kWideMaxRowTiles) return false;
  return kGemm16Cands[cand].nw != 8 || p->Mpad % (2 * BM16) == 0;
}

inline size_t gemm16_lds(int nst, int nw, int ct = 1) {
  return (size_t)nst * (ct == 2 ? kWideStageU4 : gemm16_stage_u4(nw)) * sizeof(u32x4);
}

// column units of the wide tile: pairs of 128-column tiles within a class
int gemm16_units(const FlePlan* p) {
  int u = 0;
  for (int c = 0; c < kClasses; ++c) u += ((p->col_off[c + 1] - p->col_off[c]) / BN16 + 1) / 2;
  return u;
}

template <bool FP16>
const void* gemm16_fn(int nst, int nw, int ct = 1) {
  if (ct == 2) return (const void*)fle_gemm16w_kernel<FP16>;
  if (nw == 8) return nst == 3 ? (const void*)fle_gemm16_kernel<FP16, 3, 8> : (const void*)fle_gemm16_kernel<FP16, 2, 8>;
  if (nw == 4) return nst == 3 ? (const void*)fle_gemm16_kernel<FP16, 3, 4> : (const void*)fle_gemm16_kernel<FP16, 2, 4>;
  return nst == 3 ? (const void*)fle_gemm16_kernel<FP16, 3, 2> : (const void*)fle_gemm16_kernel<FP16, 2, 2>;
}

// Resident workgroups per CU of one tile configuration.
int gemm16_occupancy(int nst, int nw, int ct, int* occ) {
  int n = 0;
  const int rc = mch::hip_check(
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, gemm16_fn<false>(nst, nw, ct), 64 * nw,
                                                   gemm16_lds(nst, nw, ct)),
      "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  *occ = n > 0 ? n : 1;
  return rc;
}

// The work items of a wide launch at batch B (see fle_gemm16w_kernel): for
// every K split and class parity, the 256-row tiles whose K range meets the
// split's chunk.  False if the launch does not fit the schedule's limits.
bool wide_schedule(const FlePlan* p, int B, int splits, int kchunk, WideSched& ws) {
  const int nmt = (2 * B + 2 * BM16 - 1) / (2 * BM16);
  if (splits > kWideMaxSplits || nmt > kWideMaxRowTiles) return false;
  memset(&ws, 0, sizeof(ws));
  ws.splits = splits;
  ws.kchunk = kchunk;
  for (int c = 0; c < kClasses; ++c) ws.nu[c] = ((p->col_off[c + 1] - p->col_off[c]) / BN16 + 1) / 2;
  for (int ks = 0; ks < splits; ++ks) {
    for (int par = 0; par < 2; ++par)
      for (int mt = 0; mt < nmt; ++mt) {
        int klo, khi;
        k_range(par, mt * 2 * BM16, 2 * BM16, B, p->Kpad, klo, khi);
        if (std::max(ks * kchunk, klo) < std::min(ks * kchunk + kchunk, khi)) ws.busy[par][ks] |= 1ull << mt;
      }
    for (int c = 0; c < kClasses; ++c) ws.per_split[ks] += ws.nu[c] * __builtin_popcountll(ws.busy[c & 1][ks]);
    ws.nitems += ws.per_split[ks];
  }
  return true;
}

// K split that fills `rounds` rounds of `slots` resident workgroups with
// the (column tile x row tile) grid, rounded so that every split but the
// last has the same number of K blocks.
int gemm16_splits(const FlePlan* p, int nw, int ct, int slots, int rounds) {
  const int ncol = ct == 2 ? gemm16_units(p) : p->Ntot / BN16;
  const int ntiles = ncol * (p->Mpad / (BM16 * gemm16_row_tiles(nw)));
  const int kblocks = p->Kpad / BK16;
  const int smax = std::min(kblocks, ct == 2 ? kWideMaxSplits : kblocks);
  if (rounds == 0 && ct == 2) {
    // fewest k-stages per workgroup: rounds of busy items x chunk length (plus
    // a few stages for each item's pipeline fill and partial-tile store)
    int best = 1;
    long long tbest = -1;
    for (int s = 1; s <= smax; ++s) {
      const int per = (kblocks + s - 1) / s, se = (kblocks + per - 1) / per;
      WideSched ws;
      if (!wide_schedule(p, p->max_batch, se, per * BK16, ws)) continue;
      const long long t = (long long)((ws.nitems + slots - 1) / std::max(slots, 1)) * (per + 4);
      if (tbest < 0 || t < tbest) { tbest = t; best = se; }
    }
    return best;
  }
  int s = rounds * slots / ntiles;
  s = std::min(std::max(s, 1), smax);
  const int per = (kblocks + s - 1) / s;
  return (kblocks + per - 1) / per;
}

int launch_gemm(FlePlan* p, int B, hipStream_t stream, int* splits_out, int* skip_kchunk = nullptr) {
  if (skip_kchunk) *skip_kchunk = 0;
  const int mt = (2 * B + BM - 1) / BM;
  if (p->precision == kFp32) {
    const int kblocks = p->Kpad / BK;
    const int kchunk = ((kblocks + p->splits - 1) / p->splits) * BK;
    const int splits = (p->Kpad + kchunk - 1) / kchunk;
    hipLaunchKernelGGL(fle_gemm_kernel, dim3(p->Ntot / BN, mt, splits), dim3(kGemmThreads), 0,
                       stream, p->X, p->W, p->part, p->Mpad, p->Kpad, p->Ntot, kchunk, B, *p);
    *splits_out = splits;
    return mch::hip_check(hipGetLastError(), "fle_gemm_kernel launch");
  }
  const int kblocks = p->Kpad / BK16;
  const int kchunk = ((kblocks + p->ksplit16 - 1) / p->ksplit16) * BK16;
  const int splits = (p->Kpad + kchunk - 1) / kchunk;
  const int ntn = p->Ntot / BN16;
  const int rt = gemm16_row_tiles(p->nw16);
  const int mt16 = (2 * B + BM16 * rt - 1) / (BM16 * rt);
  const bool fp16 = p->precision == kFp16x3;
  if (p->ct16 == 2) {
    WideSched ws;
    if (!wide_schedule(p, B, splits, kchunk, ws)) {
      mch::set_error("wide GEMM tile: %d K splits or %d row tiles beyond its schedule", splits, mt16);
      return MCH_ERR_UNSUPPORTED;
    }
    const dim3 gridw(std::max(1, std::min(ws.nitems, p->wide_slots)));
    const size_t ldsw = gemm16_lds(2, 8, 2);
    if (fp16)
      hipLaunchKernelGGL(fle_gemm16w_kernel<true>, gridw, dim3(512), ldsw, stream, p->Xhi, p->Xlo,
                         (const u32x4*)p->Wt16, p->part16, p->Mpad, p->Kpad, p->Ntot, B, ws, *p);
    else
      hipLaunchKernelGGL(fle_gemm16w_kernel<false>, gridw, dim3(512), ldsw, stream, p->Xhi, p->Xlo,
                         (const u32x4*)p->Wt16, p->part16, p->Mpad, p->Kpad, p->Ntot, B, ws, *p);
    *splits_out = splits;
    if (skip_kchunk) *skip_kchunk = kchunk;
    return mch::hip_check(hipGetLastError(), "fle_gemm16w_kernel launch");
  }
  const dim3 grid(ntn * mt16 * splits);
  const size_t lds = gemm16_lds(p->nst16, p->nw16);
#define MCH_GEMM16(F, S, W)                                                                        \
  hipLaunchKernelGGL((fle_gemm16_kernel<F, S, W>), grid, dim3(64 * W), lds, stream, p->Xhi, p->Xlo, \
                     (const u32x4*)p->Wt16, p->part16, p->Mpad, p->Kpad, p->Ntot, kchunk, ntn, mt16, B, *p)
#define MCH_GEMM16_NW(W)                               \
  do {                                                 \
    if (fp16 && p->nst16 == 3) MCH_GEMM16(true, 3, W); \
    else if (fp16) MCH_GEMM16(true, 2, W);             \
    else if (p->nst16 == 3) MCH_GEMM16(false, 3, W);   \
    else MCH_GEMM16(false, 2, W);                      \
  } while (0)
  if (p->nw16 == 8) MCH_GEMM16_NW(8);
  else if (p->nw16 == 2) MCH_GEMM16_NW(2);
  else MCH_GEMM16_NW(4);
#undef MCH_GEMM16_NW
#undef MCH_GEMM16
  *splits_out = splits;
  return mch::hip_check(hipGetLastError(), "fle_gemm16_kernel launch");
}

}  // namespace

// The projection table W (Kpad, Ntot) in FP64, rounded to FP32: entry
// (v, col) of column (l, k, m) is h w_v j_l(lam_lk r_v / R) P_l^m(cos theta_v)
// c_lk cos(m phi_v) in the cos half (row v) and ssign times the same with
// sin(m phi_v) in the sin half (row KH + v); vox = (r, cos theta, phi, h w)
// per fundamental voxel (FP64, from the host), chan_lm = (l, m) per column
// (l < 0: padding, left zero), chan_val = (lam_lk, c_lk, ssign) per column.
// Rows beyond V0 in either half stay zero (the caller clears W first).
__global__ __launch_bounds__(256) void fle_basis_kernel(float* __restrict__ W, int V0, int KH, int Ntot,
                                                        const double4* __restrict__ vox,
                                                        const int2* __restrict__ chan_lm,
                                                        const double4* __restrict__ chan_val, double radius) {
  const long long total = (long long)V0 * Ntot;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int v = (int)(e / Ntot), col = (int)(e - (long long)v * Ntot);
    const int2 lm = chan_lm[col];
    if (lm.x < 0) continue;
    const double4 q = vox[v], cv = chan_val[col];
    const double J = mch::special::sph_jn(lm.x, cv.x * q.x / radius);
    const double P = mch::special::legendre_on(lm.x, lm.y, q.y);
    const double rad = q.w * J * P * cv.y;
    double sn, cs;
    sincos(q.z * (double)lm.y, &sn, &cs);
    W[(long long)v * Ntot + col] = (float)(rad * cs);
    W[(long long)(KH + v) * Ntot + col] = (float)(cv.z * rad * sn);
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int mch_fle_plan_create(int N, int V0, int Kpad, int Ntot, const int* class_colpad,
                        const int* orbit, const float* basis, const int* chan_plus,
                        const int* chan_minus, const int* chan_m, int ne, int max_batch,
                        void** out_plan) {
  if (!out_plan || !class_colpad || !orbit || !chan_plus || !chan_minus || !chan_m) {
    mch::set_error("mch_fle_plan_create: NULL argument");
    return MCH_ERR_ARG;
  }
  *out_plan = nullptr;
  int sum = 0;
  for (int c = 0; c < kClasses; ++c) {
    if (class_colpad[c] < 0 || class_colpad[c] % BN) {
      mch::set_error("class column counts must be multiples of %d", BN);
      return MCH_ERR_ARG;
    }
    sum += class_colpad[c];
  }
  if (N < 2 || V0 < 1 || Kpad < 2 * V0 || Kpad % (2 * 64) || Kpad / 2 < V0 || sum != Ntot || ne < 1 ||
      max_batch < 1) {
    mch::set_error("mch_fle_plan_create: inconsistent sizes (N=%d V0=%d Kpad=%d Ntot=%d ne=%d)", N,
                   V0, Kpad, Ntot, ne);
    return MCH_ERR_SIZE;
  }
  for (long long i = 0; i < (long long)V0 * 16; ++i) {
    if (orbit[i] < -1 || orbit[i] >= N * N * N) {
      mch::set_error("orbit index out of range");
      return MCH_ERR_ARG;
    }
  }
  for (int c = 0; c < Ntot; ++c) {
    if (chan_plus[c] >= ne || chan_minus[c] >= ne) {
      mch::set_error("channel index out of range");
      return MCH_ERR_ARG;
    }
  }
  FlePlan* p = new FlePlan();
  memset(p, 0, sizeof(*p));
  p->N = N; p->V0 = V0; p->Kpad = Kpad; p->Ntot = Ntot; p->ne = ne; p->max_batch = max_batch;
  p->col_off[0] = 0;
  for (int c = 0; c < kClasses; ++c) p->col_off[c + 1] = p->col_off[c] + class_colpad[c];
  p->Mpad = ((2 * max_batch + BM - 1) / BM) * BM;
  // split K so that the grid has >= ~4 workgroups per CU
  const int ntiles = Ntot / BN * (p->Mpad / BM);
  int splits = (1024 + ntiles - 1) / ntiles;
  const int kblocks = Kpad / BK;
  if (splits > kblocks) splits = kblocks;
  if (splits < 1) splits = 1;
  p->splits = splits;
  int rc = MCH_OK;
#define MCH_TRY(expr, what) do { if ((rc = mch::hip_check((expr), what)) != MCH_OK) goto fail; } while (0)
  MCH_TRY(hipMalloc(&p->orbit, (size_t)V0 * 16 * sizeof(int)), "hipMalloc(orbit)");
  MCH_TRY(hipMemcpy(p->orbit, orbit, (size_t)V0 * 16 * sizeof(int), hipMemcpyHostToDevice), "copy orbit");
  {
    // each fundamental voxel (x >= y >= 0, z >= 0) as centred coordinates, from
    // any of its images inside the array (the images are the signed
    // permutations of (x, y) with +-z); then every entry of the table must be
    // its image p + 4q + 8t, or -1 exactly where that image is outside the array
    std::vector<int> fc(V0);
    const int c = N / 2;
    for (int v = 0; v < V0; ++v) {
      const int* o = orbit + (size_t)v * 16;
      int id = -1;
      for (int i = 0; i < 16 && id < 0; ++i) id = o[i];
      const int gx = id / (N * N) - c, gy = (id / N) % N - c, gz = id % N - c;
      const int x = std::max(abs(gx), abs(gy)), y = std::min(abs(gx), abs(gy)), z = abs(gz);
      bool ok = id >= 0 && x <= 1023 && z <= 1023;
      for (int t = 0; t < 2 && ok; ++t)
        for (int q = 0; q < 2 && ok; ++q)
          for (int pp = 0; pp < 4 && ok; ++pp) {
            const int y1 = t ? -y : y, zz = q ? -z : z;
            const int a0 = pp == 0 ? x : (pp == 1 ? -y1 : (pp == 2 ? -x : y1));
            const int a1 = pp == 0 ? y1 : (pp == 1 ? x : (pp == 2 ? -y1 : -x));
            const bool in = a0 >= -c && a0 < N - c && a1 >= -c && a1 < N - c && zz >= -c && zz < N - c;
            const int want = in ? ((a0 + c) * N + (a1 + c)) * N + (zz + c) : -1;
            ok = o[pp + 4 * q + 8 * t] == want;
          }
      if (!ok) {
        mch::set_error("orbit table: row %d is not the D4h orbit (images p + 4q + 8t, -1 outside the array) "
                       "of a voxel with some image in the array", v);
        rc = MCH_ERR_ARG;
        goto fail;
      }
      fc[v] = x | (y << 10) | (z << 20);
    }
    MCH_TRY(hipMalloc(&p->fcoord, (size_t)V0 * sizeof(int)), "hipMalloc(fcoord)");
    MCH_TRY(hipMemcpy(p->fcoord, fc.data(), (size_t)V0 * sizeof(int), hipMemcpyHostToDevice), "copy fcoord");
  }
  MCH_TRY(hipMalloc(&p->W, (size_t)Kpad * Ntot * sizeof(float)), "hipMalloc(W)");
  if (basis)
    MCH_TRY(hipMemcpy(p->W, basis, (size_t)Kpad * Ntot * sizeof(float), hipMemcpyHostToDevice), "copy W");
  else  // built on the device by mch_fle_build_basis
    MCH_TRY(hipMemset(p->W, 0, (size_t)Kpad * Ntot * sizeof(float)), "memset(W)");
  MCH_TRY(hipMalloc(&p->chan_plus, (size_t)Ntot * sizeof(int)), "hipMalloc(chan)");
  MCH_TRY(hipMemcpy(p->chan_plus, chan_plus, (size_t)Ntot * sizeof(int), hipMemcpyHostToDevice), "copy chan");
  MCH_TRY(hipMalloc(&p->chan_minus, (size_t)Ntot * sizeof(int)), "hipMalloc(chan)");
  MCH_TRY(hipMemcpy(p->chan_minus, chan_minus, (size_t)Ntot * sizeof(int), hipMemcpyHostToDevice), "copy chan");
  MCH_TRY(hipMalloc(&p->chan_m, (size_t)Ntot * sizeof(int)), "hipMalloc(chan)");
  MCH_TRY(hipMemcpy(p->chan_m, chan_m, (size_t)Ntot * sizeof(int), hipMemcpyHostToDevice), "copy chan");
  MCH_TRY(hipMalloc(&p->X, (size_t)kClasses * p->Mpad * Kpad * sizeof(float)), "hipMalloc(X)");
  MCH_TRY(hipMemset(p->X, 0, (size_t)kClasses * p->Mpad * Kpad * sizeof(float)), "memset(X)");
  MCH_TRY(hipMalloc(&p->part, (size_t)splits * p->Mpad * Ntot * sizeof(float)), "hipMalloc(part)");
#undef MCH_TRY
  *out_plan = p;
  return MCH_OK;
fail:
  free_fle_plan(p);
  return rc;
}

int mch_fle_build_basis(void* plan, const double* vox, const int* chan_lm, const double* chan_val,
                        double radius) {
  FlePlan* p = (FlePlan*)plan;
  if (!p || !vox || !chan_lm || !chan_val) {
    mch::set_error("mch_fle_build_basis: NULL argument");
    return MCH_ERR_ARG;
  }
  if (!(radius > 0.0)) {
    mch::set_error("mch_fle_build_basis: radius must be positive");
    return MCH_ERR_ARG;
  }
  for (int c = 0; c < p->Ntot; ++c) {
    const int l = chan_lm[2 * c], m = chan_lm[2 * c + 1];
    if (l >= 0 && (m < 0 || m > l || l > 1000)) {
      mch::set_error("mch_fle_build_basis: column %d has (l, m) = (%d, %d)", c, l, m);
      return MCH_ERR_ARG;
    }
  }
  int rc = MCH_OK;
  void *dv = nullptr, *dl = nullptr, *dc = nullptr;
#define MCH_TRY(expr, what) do { if ((rc = mch::hip_check((expr), what)) != MCH_OK) goto done; } while (0)
  {
    MCH_TRY(hipMalloc(&dv, (size_t)p->V0 * 4 * sizeof(double)), "hipMalloc(vox)");
    MCH_TRY(hipMalloc(&dl, (size_t)p->Ntot * 2 * sizeof(int)), "hipMalloc(chan_lm)");
    MCH_TRY(hipMalloc(&dc, (size_t)p->Ntot * 4 * sizeof(double)), "hipMalloc(chan_val)");
    MCH_TRY(hipMemcpy(dv, vox, (size_t)p->V0 * 4 * sizeof(double), hipMemcpyHostToDevice), "copy vox");
    MCH_TRY(hipMemcpy(dl, chan_lm, (size_t)p->Ntot * 2 * sizeof(int), hipMemcpyHostToDevice), "copy chan_lm");
    // (lam, c, ssign) padded to four doubles per column
    std::vector<double> cv((size_t)p->Ntot * 4, 0.0);
    for (int c = 0; c < p->Ntot; ++c)
      for (int i = 0; i < 3; ++i) cv[(size_t)c * 4 + i] = chan_val[(size_t)c * 3 + i];
    MCH_TRY(hipMemcpy(dc, cv.data(), cv.size() * sizeof(double), hipMemcpyHostToDevice), "copy chan_val");
    MCH_TRY(hipMemset(p->W, 0, (size_t)p->Kpad * p->Ntot * sizeof(float)), "memset(W)");
    const long long total = (long long)p->V0 * p->Ntot;
    const int grid = (int)std::min<long long>((total + 255) / 256, 65536LL);
    hipLaunchKernelGGL(fle_basis_kernel, dim3(grid), dim3(256), 0, 0, p->W, p->V0, p->Kpad / 2, p->Ntot,
                       (const double4*)dv, (const int2*)dl, (const double4*)dc, radius);
    MCH_TRY(hipGetLastError(), "fle_basis_kernel launch");
    MCH_TRY(hipDeviceSynchronize(), "fle_basis_kernel");
  }
#undef MCH_TRY
done:
  (void)hipFree(dv); (void)hipFree(dl); (void)hipFree(dc);
  return rc;
}

int mch_fle_copy_basis(void* plan, float* out, hipStream_t stream) {
  FlePlan* p = (FlePlan*)plan;
  if (!p || !out) {
    mch::set_error("mch_fle_copy_basis: NULL argument");
    return MCH_ERR_ARG;
  }
  return mch::hip_check(hipMemcpyAsync(out, p->W, (size_t)p->Kpad * p->Ntot * sizeof(float),
                                       hipMemcpyDeviceToDevice, stream),
                        "mch_fle_copy_basis");
}

int mch_fle_plan_destroy(void* plan) {
  free_fle_plan((FlePlan*)plan);
  return MCH_OK;
}

// Selects the compute precision of the projection GEMM: 0 = FP32 (the plain
// path), 1 = BF16 hi/lo split (three MFMA products, FP32 accumulation),
// 2 = FP16 hi/lo split with power-of-2 row and column scales.  The first
// switch to a 16-bit mode builds W^T hi/lo on the device (synchronises).
int mch_fle_set_precision(void* plan, int precision) {
  FlePlan* p = (FlePlan*)plan;
  if (!p || precision < kFp32 || precision > kFp16x3) {
    mch::set_error("mch_fle_set_precision: bad arguments (precision %d)", precision);
    return MCH_ERR_ARG;
  }
  if (precision == kFp32) {
    p->precision = kFp32;
    return MCH_OK;
  }
  if (p->Kpad % BK16 || p->Ntot % BN16) {
    mch::set_error("mch_fle_set_precision: 16-bit path needs Kpad %% %d == 0 and class widths %% %d == 0",
                   BK16, BN16);
    return MCH_ERR_UNSUPPORTED;
  }
  for (int c = 0; c < kClasses; ++c) {
    if ((p->col_off[c + 1] - p->col_off[c]) % BN16) {
      mch::set_error("mch_fle_set_precision: class column counts must be multiples of %d", BN16);
      return MCH_ERR_UNSUPPORTED;
    }
  }
  int rc = MCH_OK;
  const size_t nw = (size_t)p->Ntot * p->Kpad;
#define MCH_TRY(expr, what) do { if ((rc = mch::hip_check((expr), what)) != MCH_OK) return rc; } while (0)
  if (!p->Wt16) {
    MCH_TRY(hipMalloc(&p->Wt16, 2 * nw * sizeof(uint16_t)), "hipMalloc(Wt16)");
    MCH_TRY(hipMalloc(&p->colscale, (size_t)p->Ntot * sizeof(float)), "hipMalloc(colscale)");
    MCH_TRY(hipMalloc(&p->rowmax, (size_t)p->Mpad * sizeof(float)), "hipMalloc(rowmax)");
    // dynamic LDS above 64 KB for every tile configuration that needs it
    for (int cand = 0; cand < kGemm16Candidates; ++cand) {
      const int nst = kGemm16Cands[cand].nst, nw = kGemm16Cands[cand].nw, ct = kGemm16Cands[cand].ct;
      const int lds = (int)gemm16_lds(nst, nw, ct);
      if (lds <= 64 * 1024) continue;
      MCH_TRY(hipFuncSetAttribute(gemm16_fn<false>(nst, nw, ct), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  lds), "hipFuncSetAttribute");
      MCH_TRY(hipFuncSetAttribute(gemm16_fn<true>(nst, nw, ct), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  lds), "hipFuncSetAttribute");
    }
    // partials for the largest K split any autotune candidate uses
    int cus = 0, dev = 0;
    MCH_TRY(hipGetDevice(&dev), "hipGetDevice");
    MCH_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    p->cus = cus > 0 ? cus : 256;
    {
      int occw = 0;
      if ((rc = gemm16_occupancy(2, 8, 2, &occw))) return rc;
      p->wide_slots = occw * p->cus;
    }
    int smax = 1;
    for (int cand = 0; cand < kGemm16Candidates; ++cand) {
      if (!gemm16_usable(p, cand)) continue;
      const Gemm16Cand& c = kGemm16Cands[cand];
      int occ = 0;
      if ((rc = gemm16_occupancy(c.nst, c.nw, c.ct, &occ))) return rc;
      smax = std::max(smax, gemm16_splits(p, c.nw, c.ct, occ * p->cus, c.rounds));
    }
    p->splits16 = smax;
    p->ksplit16 = smax;
    MCH_TRY(hipMalloc(&p->part16, (size_t)smax * p->Mpad * p->Ntot * sizeof(float)), "hipMalloc(part16)");
    const size_t nx = (size_t)kClasses * p->Mpad * p->Kpad;
    MCH_TRY(hipMalloc(&p->Xhi, nx * sizeof(uint16_t)), "hipMalloc(Xhi)");
    MCH_TRY(hipMalloc(&p->Xlo, nx * sizeof(uint16_t)), "hipMalloc(Xlo)");
    MCH_TRY(hipMemset(p->Xhi, 0, nx * sizeof(uint16_t)), "memset(Xhi)");
    MCH_TRY(hipMemset(p->Xlo, 0, nx * sizeof(uint16_t)), "memset(Xlo)");
    p->nst16 = 2;
    p->nw16 = 4;
    p->ct16 = 1;
  }
  if (p->table_precision != precision) {
    const bool fp16 = precision == kFp16x3;
    const int nb = (p->Ntot + 255) / 256;
    hipLaunchKernelGGL(fle_fill_kernel, dim3(nb), dim3(256), 0, 0, p->colscale, p->Ntot, 0.f);
    if (fp16) {
      const int kchunk = (p->Kpad + 63) / 64;
      hipLaunchKernelGGL(fle_colmax_kernel, dim3(nb, 64), dim3(256), 0, 0, p->W, p->Kpad, p->Ntot, kchunk,
                         p->colscale);
    }
    hipLaunchKernelGGL(fle_colscale_kernel, dim3(nb), dim3(256), 0, 0, p->colscale, p->Ntot, fp16 ? 1 : 0);
    if (fp16)
      hipLaunchKernelGGL(fle_wsplit_kernel<true>, dim3(p->Ntot / 64, p->Kpad / 64), dim3(256), 0, 0, p->W,
                         p->Kpad, p->Ntot, p->colscale, (u32x4*)p->Wt16);
    else
      hipLaunchKernelGGL(fle_wsplit_kernel<false>, dim3(p->Ntot / 64, p->Kpad / 64), dim3(256), 0, 0, p->W,
                         p->Kpad, p->Ntot, p->colscale, (u32x4*)p->Wt16);
    MCH_TRY(hipGetLastError(), "fle_wsplit_kernel launch");
    MCH_TRY(hipDeviceSynchronize(), "building the 16-bit projection tables");
    p->table_precision = precision;
    p->precision = precision;
    // autotune on the full batch (timing does not depend on the operand
    // values): pipeline depth (more stages keep more traffic in flight per
    // workgroup), waves per workgroup (8: a 256-row tile, which halves the
    // W^T bytes per MFMA but holds one workgroup per CU), and the K split as
    // 1, 2 or 3 full rounds of resident workgroups (a partial last round
    // leaves most CUs idle while it runs)
    hipEvent_t e0, e1;
    MCH_TRY(hipEventCreate(&e0), "hipEventCreate");
    MCH_TRY(hipEventCreate(&e1), "hipEventCreate");
    float best = 1e30f;
    int best_nst = 2, best_nw = 4, best_ct = 1, best_split = p->splits16;
    for (int cand = 0; cand < kGemm16Candidates; ++cand) {
      if (!gemm16_usable(p, cand)) continue;
      const int nst = kGemm16Cands[cand].nst, nw = kGemm16Cands[cand].nw, ct = kGemm16Cands[cand].ct;
      int occ = 0;
      if ((rc = gemm16_occupancy(nst, nw, ct, &occ))) break;
      p->nst16 = nst;
      p->nw16 = nw;
      p->ct16 = ct;
      p->ksplit16 = gemm16_splits(p, nw, ct, occ * p->cus, kGemm16Cands[cand].rounds);
      int splits = 0;
      for (int rep = 0; rep < 2; ++rep)
        if ((rc = launch_gemm(p, p->max_batch, 0, &splits))) break;
      if (rc) break;
      (void)hipEventRecord(e0, 0);
      for (int rep = 0; rep < 5; ++rep)
        if ((rc = launch_gemm(p, p->max_batch, 0, &splits))) break;
      if (rc) break;
      (void)hipEventRecord(e1, 0);
      MCH_TRY(hipEventSynchronize(e1), "autotune");
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; best_nst = nst; best_nw = nw; best_ct = ct; best_split = p->ksplit16; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    p->nst16 = best_nst;
    p->nw16 = best_nw;
    p->ct16 = best_ct;
    p->ksplit16 = best_split;
  }
#undef MCH_TRY
  p->precision = precision;
  return MCH_OK;
}

// Pin the 16-bit GEMM to one tile configuration (after the autotune of the
// first switch to a 16-bit mode): stages, waves per workgroup, 128-column tiles
// per workgroup and K splits, which must be an autotune candidate usable for
// the plan with splits within the plan's partial buffer.
int mch_fle_set_gemm_tiles(void* plan, int stages, int waves, int column_tiles, int k_splits) {
  FlePlan* p = (FlePlan*)plan;
  if (!p || !p->nst16) {
    mch::set_error("mch_fle_set_gemm_tiles: switch the plan to a 16-bit mode first");
    return MCH_ERR_ARG;
  }
  bool found = false;
  for (int cand = 0; cand < kGemm16Candidates && !found; ++cand)
    found = gemm16_usable(p, cand) && kGemm16Cands[cand].nst == stages && kGemm16Cands[cand].nw == waves &&
            kGemm16Cands[cand].ct == column_tiles;
  if (!found || k_splits < 1 || k_splits > p->splits16 || k_splits > p->Kpad / BK16 ||
      (column_tiles == 2 && k_splits > kWideMaxSplits)) {
    mch::set_error("mch_fle_set_gemm_tiles: (%d stages, %d waves, %d column tiles, %d splits) is not a usable "
                   "configuration (splits <= %d)", stages, waves, column_tiles, k_splits, p->splits16);
    return MCH_ERR_ARG;
  }
  p->nst16 = stages;
  p->nw16 = waves;
  p->ct16 = column_tiles;
  p->ksplit16 = k_splits;
  return MCH_OK;
}

// Tile configuration the 16-bit GEMM was tuned to: 10000 * (column tiles per
// workgroup - 1) + 100 * K splits + 10 * waves per workgroup + pipeline depth
// (0 before any 16-bit mode).
int mch_fle_tuned_depth(void* plan) {
  FlePlan* p = (FlePlan*)plan;
  return p && p->nst16 ? 10000 * (p->ct16 - 1) + 100 * p->ksplit16 + 10 * p->nw16 + p->nst16 : 0;
}

// a (B, ne) complex64 rows of stride lda = B^H f for volumes (B, N^3) float32.
int mch_fle_project_shifted(void* plan, const float* vols, long long ld_vol, int B, const float* shift_zyx,
                            void* coeffs, long long lda, hipStream_t stream);

int mch_fle_project(void* plan, const float* vols, long long ld_vol, int B, void* coeffs,
                    long long lda, hipStream_t stream) {
  return mch_fle_project_shifted(plan, vols, ld_vol, B, nullptr, coeffs, lda, stream);
}

// The same for the patches vol[a + flip(shift_zyx[b])] (extract_shifted_patches
// semantics), read straight from the volumes: no patch copy.
int mch_fle_project_shifted(void* plan, const float* vols, long long ld_vol, int B, const float* shift_zyx,
                            void* coeffs, long long lda, hipStream_t stream) {
  FlePlan* p = (FlePlan*)plan;
  if (!p || !vols || !coeffs) { mch::set_error("mch_fle_project: NULL argument"); return MCH_ERR_ARG; }
  if (B < 1 || B > p->max_batch || ld_vol < (long long)p->N * p->N * p->N || lda < p->ne) {
    mch::set_error("mch_fle_project: bad sizes (B=%d max %d, ld_vol=%lld, lda=%lld)", B,
                   p->max_batch, ld_vol, lda);
    return MCH_ERR_SIZE;
  }
  const long long tot = (long long)B * p->V0;
  int nb = (int)((tot + 255) / 256);
  if (nb > 8192) nb = 8192;
  const bool fp16 = p->precision == kFp16x3;
  int rc;
  if (fp16) {
    hipLaunchKernelGGL(fle_fill_kernel, dim3((p->Mpad + 255) / 256), dim3(256), 0, stream, p->rowmax,
                       p->Mpad, 0.f);
    if ((rc = mch::hip_check(hipGetLastError(), "fle_fill_kernel launch"))) return rc;
  }
  if (p->precision == kBf16x3) {
    const long long nq = (long long)B * (((p->V0 + 1023) / 1024) * 256);
    const int nb4 = (int)std::min<long long>((nq + 255) / 256, 16384);
    if (shift_zyx)
      hipLaunchKernelGGL(fle_symmetrize_bf16_kernel<true>, dim3(nb4), dim3(256), 0, stream, vols, ld_vol, B,
                         p->fcoord, p->V0, p->N, p->Kpad, p->Mpad, shift_zyx, p->Xhi, p->Xlo);
    else
      hipLaunchKernelGGL(fle_symmetrize_bf16_kernel<false>, dim3(nb4), dim3(256), 0, stream, vols, ld_vol, B,
                         p->fcoord, p->V0, p->N, p->Kpad, p->Mpad, nullptr, p->Xhi, p->Xlo);
  } else {
    hipLaunchKernelGGL(fle_symmetrize_kernel, dim3(nb), dim3(256), 0, stream, vols, ld_vol, B, p->orbit, p->V0,
                       p->N, p->Kpad, p->Mpad, shift_zyx, p->X, fp16 ? p->rowmax : nullptr);
  }
  rc = mch::hip_check(hipGetLastError(), "fle_symmetrize_kernel launch");
  if (rc) return rc;
  if (fp16) {
    hipLaunchKernelGGL(fle_xsplit_kernel, dim3(4096), dim3(256), 0, stream, p->X, p->rowmax, p->Mpad,
                       p->Kpad, 2 * B, p->Xhi, p->Xlo);
    if ((rc = mch::hip_check(hipGetLastError(), "fle_xsplit_kernel launch"))) return rc;
  }
  int splits = 0, skip_kchunk = 0;
  if ((rc = launch_gemm(p, B, stream, &splits, &skip_kchunk))) return rc;
  const long long ftot = (long long)B * p->Ntot;
  int fb = (int)((ftot + 255) / 256);
  if (fb > 8192) fb = 8192;
  hipLaunchKernelGGL(fle_finalize_kernel, dim3(fb), dim3(256), 0, stream,
                     p->precision == kFp32 ? p->part : p->part16, splits, p->Mpad, p->Ntot, B,
                     p->chan_plus, p->chan_minus, p->chan_m, (float2*)coeffs, lda,
                     fp16 ? p->colscale : nullptr, fp16 ? p->rowmax : nullptr,
                     make_int4(p->col_off[3], p->col_off[4], p->col_off[7], p->col_off[8]), skip_kchunk, *p);
  return mch::hip_check(hipGetLastError(), "fle_finalize_kernel launch");
}

// Coefficients of the patches V_t and their t-derivatives (the four volumes
// of fle_symmetrize_tangent_kernel) into rows b, Bp + b, 2 Bp + b, 3 Bp + b of
// coeffs: one pass over the particles instead of four volumes written and
// read back.  FP32 and BF16x3 precision.
int mch_fle_project_tangents(void* plan, const float* vols, long long ld_vol, int Nv, int Bp, const float* t,
                             void* coeffs, long long lda, hipStream_t stream) {
  FlePlan* p = (FlePlan*)plan;
  if (!p || !vols || !t || !coeffs) { mch::set_error("mch_fle_project_tangents: NULL argument"); return MCH_ERR_ARG; }
  const int B = 4 * Bp;
  if (Bp < 1 || B > p->max_batch || Nv < p->N || Nv > 1024 || ld_vol < (long long)Nv * Nv * Nv || lda < p->ne) {
    mch::set_error("mch_fle_project_tangents: bad sizes (4 x %d volumes, max %d; Nv=%d N=%d)", Bp, p->max_batch,
                   Nv, p->N);
    return MCH_ERR_SIZE;
  }
  if (p->precision == kFp16x3) {
    mch::set_error("mch_fle_project_tangents: FP16x3 precision is not supported (use FP32 or BF16x3)");
    return MCH_ERR_UNSUPPORTED;
  }
  const bool bf16 = p->precision == kBf16x3;
  const long long tot = (long long)Bp * (bf16 ? p->Kpad / 2 : p->V0);
  const int nb = (int)std::min<long long>((tot + 255) / 256, 16384);
  if (bf16)
    hipLaunchKernelGGL(fle_symmetrize_tangent_kernel<true>, dim3(nb), dim3(256), 0, stream, vols, ld_vol, Nv, Bp,
                       p->orbit, p->V0, p->N, p->Kpad, p->Mpad, t, nullptr, p->Xhi, p->Xlo);
  else
    hipLaunchKernelGGL(fle_symmetrize_tangent_kernel<false>, dim3(nb), dim3(256), 0, stream, vols, ld_vol, Nv, Bp,
                       p->orbit, p->V0, p->N, p->Kpad, p->Mpad, t, p->X, nullptr, nullptr);
  int rc = mch::hip_check(hipGetLastError(), "fle_symmetrize_tangent_kernel launch");
  if (rc) return rc;
  int splits = 0, skip_kchunk = 0;
  if ((rc = launch_gemm(p, B, stream, &splits, &skip_kchunk))) return rc;
  const long long ftot = (long long)B * p->Ntot;
  int fb = (int)((ftot + 255) / 256);
  if (fb > 8192) fb = 8192;
  hipLaunchKernelGGL(fle_finalize_kernel, dim3(fb), dim3(256), 0, stream,
                     p->precision == kFp32 ? p->part : p->part16, splits, p->Mpad, p->Ntot, B, p->chan_plus,
                     p->chan_minus, p->chan_m, (float2*)coeffs, lda, nullptr, nullptr,
                     make_int4(p->col_off[3], p->col_off[4], p->col_off[7], p->col_off[8]), skip_kchunk, *p);
  return mch::hip_check(hipGetLastError(), "fle_finalize_kernel launch");
}

// Re-run only the projection GEMM on the operand left by the last
// mch_fle_project call (for timing the dominant kernel in isolation).
int mch_fle_gemm_only(void* plan, int B, hipStream_t stream) {
  FlePlan* p = (FlePlan*)plan;
  if (!p || B < 1 || B > p->max_batch) { mch::set_error("mch_fle_gemm_only: bad arguments"); return MCH_ERR_ARG; }
  int splits = 0;
  return launch_gemm(p, B, stream, &splits);
}

int mch_corr_plan_create(int lmax, int kdim, const int* idx_host, int max_batch, void** out_plan) {
  if (!out_plan || !idx_host || lmax < 0 || kdim < 1 || max_batch < 1) {
    mch::set_error("mch_corr_plan_create: bad arguments");
    return MCH_ERR_ARG;
  }
  CorrPlan* p = new CorrPlan();
  p->lmax = lmax; p->kdim = kdim; p->max_batch = max_batch; p->idx = nullptr;
  const size_t n = (size_t)(lmax + 1) * (2 * lmax + 1) * kdim;
  int rc = mch::hip_check(hipMalloc(&p->idx, n * sizeof(int)), "hipMalloc(idx)");
  if (!rc) rc = mch::hip_check(hipMemcpy(p->idx, idx_host, n * sizeof(int), hipMemcpyHostToDevice), "copy idx");
  if (rc) { (void)hipFree(p->idx); delete p; return rc; }
  *out_plan = p;
  return MCH_OK;
}

int mch_corr_plan_destroy(void* plan) {
  CorrPlan* p = (CorrPlan*)plan;
  if (!p) return MCH_OK;
  (void)hipFree(p->idx);
  delete p;
  return MCH_OK;
}

// sigma rows (B, dsize(L)) with row stride lds (entries above L untouched),
// and ||aV|| (B,) when norms != NULL.
int mch_compute_sigma_degree(void* plan, const void* aT, const void* aV, long long ldv, int ne, int B, int L,
                             void* sigma, long long lds, float* norms, hipStream_t stream) {
  CorrPlan* p = (CorrPlan*)plan;
  if (!p || !aT || !aV || !sigma) { mch::set_error("mch_compute_sigma: NULL argument"); return MCH_ERR_ARG; }
  if (B < 1 || L < 0 || L > p->lmax || lds < mch::dsize(L) || ldv < ne) {
    mch::set_error("mch_compute_sigma: bad sizes (B=%d L=%d lds=%lld)", B, L, lds);
    return MCH_ERR_SIZE;
  }
  const size_t shm = (size_t)2 * (2 * L + 1) * p->kdim * sizeof(float2);
  if (shm > 64 * 1024) {
    int rc0 = mch::hip_check(hipFuncSetAttribute((const void*)sigma_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm),
                             "hipFuncSetAttribute(sigma_kernel)");
    if (rc0) return rc0;
  }
  hipLaunchKernelGGL(sigma_kernel, dim3(L + 1, B), dim3(256), shm, stream, (const float2*)aT,
                     (const float2*)aV, ldv, p->idx, p->lmax, p->kdim, (float2*)sigma, lds);
  int rc = mch::hip_check(hipGetLastError(), "sigma_kernel launch");
  if (rc || !norms) return rc;
  hipLaunchKernelGGL(norms_kernel, dim3(B), dim3(256), 0, stream, (const float2*)aV, ldv, ne, norms);
  return mch::hip_check(hipGetLastError(), "norms_kernel launch");
}

// sigma (B, dsize(lmax)) with row stride lds, and ||aV|| (B,) when norms != NULL.
int mch_compute_sigma(void* plan, const void* aT, const void* aV, long long ldv, int ne, int B,
                      void* sigma, long long lds, float* norms, hipStream_t stream) {
  CorrPlan* p = (CorrPlan*)plan;
  if (!p) { mch::set_error("mch_compute_sigma: NULL argument"); return MCH_ERR_ARG; }
  return mch_compute_sigma_degree(plan, aT, aV, ldv, ne, B, p->lmax, sigma, lds, norms, stream);
}

}  // extern "C"
