// SO(3) orientation search on gfx950: SOFFT grid + top-C, and the fused
// frequency-marched Wigner-D Newton refinement.
//
// Reference behaviour (swing-research/Matcha):
//   grid search      src/matcha/core/SOFFT.py:30-124, Matcha.py:27-58,207-236
//   Newton march     src/matcha/core/Matcha.py:11-24,150-205,239-403
//   derivatives      src/matcha/core/WignerDGradients.py:155-302
//   Wigner D         src/matcha/core/WignerDMatrices.py:1291-1375
//
// Design: the reference materialises D (B*C x Dsize) and a 15-column
// coefficient bank every step.  Here one workgroup owns one candidate for the
// whole march.  For every fundamental index pair (m' >= |m|) a lane marches l
// upward with the three-term l-recurrence of d^l_{m'm}(beta) and of its first
// and second beta-derivatives, accumulating U = sum_l sigma d (and U', U'')
// for the four symmetry images (m',m), (-m',-m), (m,m'), (-m,-m').  Score,
// gradient and Hessian are then closed-form contractions of U with the
// alpha/gamma phases; nothing of size Dsize is ever written.
#include "common.hpp"

#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <vector>

namespace mch {

static thread_local char g_err[512];
void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* get_error() { return g_err; }

// ---------------------------------------------------------------------------
// host FP64 Wigner d table (l-recurrence, textbook convention)
// ---------------------------------------------------------------------------
// d^l_{m'm}(beta) for all 0 <= l <= L, D layout.  Seeds at l0 = max(|m'|,|m|)
// come from the closed forms of the four edge cases of Wigner's sum.
static void host_wigner_d(double beta, int L, double* out /* dsize(L) */) {
  const double c = cos(0.5 * beta), s = sin(0.5 * beta);
  const double x = cos(beta);
  auto lf = [](int n) { return lgamma((double)n + 1.0); };
  auto ipow = [](double v, int p) {
    double r = 1.0;
    for (int i = 0; i < p; ++i) r *= v;
    return r;
  };
  for (int mp = -L; mp <= L; ++mp) {
    for (int m = -L; m <= L; ++m) {
      const int j = abs(mp) > abs(m) ? abs(mp) : abs(m);
      int a, b, k;
      double sign;
      if (mp == j) {
        a = mp + m; b = mp - m; k = mp + m; sign = ((mp - m) & 1) ? -1.0 : 1.0;
      } else if (mp == -j) {
        a = j - m; b = j + m; k = j + m; sign = 1.0;
      } else if (m == j) {
        a = j + mp; b = j - mp; k = j + mp; sign = 1.0;
      } else {  // m == -j
        a = j - mp; b = j + mp; k = j - mp; sign = ((j + mp) & 1) ? -1.0 : 1.0;
      }
      const double K = sign * exp(0.5 * (lf(2 * j) - lf(k) - lf(2 * j - k)));
      double d = K * ipow(c, a) * ipow(s, b);
      double pd = 0.0, Cprev = 0.0;
      out[dindex(j, mp, m)] = d;
      for (int l = j + 1; l <= L; ++l) {
        const double P = (double)(l * l - m * m) * (double)(l * l - mp * mp);
        const double A = (double)l * (2 * l - 1) / sqrt(P);
        const double Bc = (l > 1) ? (double)(m * mp) / ((double)(l - 1) * l) : 0.0;
        const double nd = A * ((x - Bc) * d - Cprev * pd);
        pd = d;
        d = nd;
        Cprev = sqrt(P) / ((double)l * (2 * l + 1));
        out[dindex(l, mp, m)] = d;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct So3Plan {
  int Lmax, L0, K, n, max_batch, max_cand;
  int nsz;               // compute_size(L0) = dsize(L0-1): entries used by the grid
  float* d_grid;         // (n, nsz)
  float2* twiddle;       // (n) e^{2 pi i k / n}
  float2* etab;          // (2 L0 - 1, 16 R) e^{2 pi i u x / n} of the blocked slice kernel (R > 0)
  float2* sig_pk;        // (max_batch, pk_size(Lmax))
  float* part_val;       // (max_batch, n, max_cand)
  int* part_idx;         // (max_batch, n, max_cand)
  float* cand;           // (max_batch, max_cand, 4)
  float* lgf;            // log2 factorial table, size 2*Lmax+2
  int stop_early;        // batch-global early exit of the march (stepped launches)
  double* state;         // (max_batch * max_cand, 16) optimiser state between stepped launches
  int* accept;           // (kMaxLevels * (kMaxStopSteps + 1)) accept counts of the stepped march
};
constexpr int kMaxLevels = 16;
constexpr int kMaxStopSteps = 63;

}  // namespace mch

using mch::So3Plan;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// sigma (reference D layout, row stride ldsig) -> 4-image repacked layout.
__global__ void so3_repack_kernel(const float2* __restrict__ sig, long long ldsig,
                                  float2* __restrict__ pk, int Lmax, int B) {
  const int per = mch::pk_size(Lmax) / 4;  // number of (l, pair) slots
  const long long total = (long long)per * B;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(t / per);
    int r = (int)(t - (long long)b * per);
    // find shell l: slots of shell l are (l+1)^2, offset sum_{k<=l} k^2
    int l = 0;
    while (mch::pk_offset(l + 1) / 4 <= r) ++l;
    const int p = r - mch::pk_offset(l) / 4;
    int mp = (int)sqrtf((float)p);
    while (mp * mp > p) --mp;
    while ((mp + 1) * (mp + 1) <= p) ++mp;
    const int m = p - mp * mp - mp;
    const float2* s = sig + (long long)b * ldsig;
    const float sg = ((m - mp) & 1) ? -1.f : 1.f;
    float2 v0 = s[mch::dindex(l, mp, m)];
    float2 v1 = s[mch::dindex(l, -mp, -m)];
    float2 v2 = s[mch::dindex(l, m, mp)];
    float2 v3 = s[mch::dindex(l, -m, -mp)];
    v1.x *= sg; v1.y *= sg;
    v2.x *= sg; v2.y *= sg;
    if (m == mp || m == -mp) { v2 = make_float2(0.f, 0.f); v3 = v2; }
    if (mp == 0) v1 = make_float2(0.f, 0.f);
    // images (mp, m) and (-mp, -m) carry conjugate phases e^{+-i(mp a + m g)},
    // and every derivative weight (u, v, u^2, uv, v^2) of the second equals
    // that of the first after conjugation: they enter as sigma_0 + conj sigma_1
    // (likewise (m, mp) and (-m, -mp)).  16 bytes per (l, pair).
    float4* o = reinterpret_cast<float4*>(pk + (long long)b * mch::pk_size(Lmax) +
                                          mch::pk_offset(l) / 2 + 2 * p);
    o[0] = make_float4(v0.x + v1.x, v0.y - v1.y, v2.x + v3.x, v2.y - v3.y);
  }
}

// Order-preserving 32-bit image of a float: a > b iff key(a) > key(b), with
// -0 and +0 mapped alike (a tie, as under a float compare).  0 is the image of
// the all-ones NaN only and marks an empty (taken or padding) slot.
__device__ __forceinline__ unsigned float_key(float v) {
  const unsigned u = __float_as_uint(v + 0.0f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// C rounds of argmax over the values of a 256-thread workgroup, value
// descending and flat index ascending among ties (the order torch.topk gives
// the grid in SOFFT.ids_to_angles / Matcha.py:27-58 up to the tie order, which
// is fixed here).  hk[i] holds float_key of the thread's i-th value, at flat
// index qof(i) of the slice, increasing with i; round c writes
// part_val/part_idx[orow + c] (index qbase + q, or -1 once the slice is
// exhausted).  A thread's best is one 64-bit word (key, ~q), so a round is one
// wave max-reduction and one four-way merge; each thread keeps its best two
// words, and the owner of the round's winner rescans its values only when it
// has no second left.
template <int NV, class QOf>
__device__ __forceinline__ void block_top_c(unsigned (&hk)[NV], QOf qof, int C, long long orow,
                                            int qbase, float* __restrict__ part_val,
                                            int* __restrict__ part_idx) {
  __shared__ unsigned long long red[2][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the thread's best word and the best of the rest (first maxima over i =
  // smallest q among equal keys; empty slots (0) never win): a win promotes
  // the second, and only a win without a known second rescans the values
  int ib = 0, ib2 = 0;
  unsigned long long best = 0ull, second = 0ull;
  bool has2 = false;
  auto word = [&](unsigned k, int i) {
    return k == 0u ? 0ull : ((unsigned long long)k << 32) | (unsigned)~qof(i);
  };
  auto local_best2 = [&]() {
    unsigned b1 = 0u, b2 = 0u;
    int i1 = 0, i2 = 0;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const unsigned k = hk[i];
      if (k > b1) {
        b2 = b1;
        i2 = i1;
        b1 = k;
        i1 = i;
      } else if (k > b2) {
        b2 = k;
        i2 = i;
      }
    }
    ib = i1;
    ib2 = i2;
    best = word(b1, i1);
    second = word(b2, i2);
    has2 = true;
  };
  local_best2();
  for (int c = 0; c < C; ++c) {
    unsigned long long w = mch::wave_max_u64(best);
    // two buffers: round c + 1 writes the other one, and round c + 2 starts
    // after every thread has passed round c + 1's barrier, i.e. read round c's
    if (lane == 0) red[c & 1][wave] = w;
    __syncthreads();
    w = red[c & 1][0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
      const unsigned long long o = red[c & 1][k];
      w = o > w ? o : w;
    }
    if (w != 0ull && w == best) {  // the (unique) owner of the winner
#pragma unroll
      for (int i = 0; i < NV; ++i)
        if (i == ib) hk[i] = 0u;
      if (has2) {
        best = second;
        ib = ib2;
        has2 = false;
      } else {
        local_best2();
      }
    }
    if (threadIdx.x == 0) {
      part_val[orow + c] = w != 0ull ? key_float((unsigned)(w >> 32)) : -INFINITY;
      part_idx[orow + c] = w != 0ull ? qbase + (int)~(unsigned)w : -1;
    }
  }
}

// F(u, v) = sum over max(|u|, |v|) <= l < L0 of d^l_{uv}(beta_j) sigma^l_{uv},
// the terms' loads issued eight at a time (surplus terms read the last
// shell's entry and are masked), so that a thread waits for one round trip
// per eight terms.  The indices: each from dindex with the shell clamped, or
// (STEP) stepped from shell l to l + 1 by dindex(l + 1, u, v) -
// dindex(l, u, v) = (2l + 1)^2 + 4l + 2u + 4.  Measured in the blocked slice
// kernel (batch 1024 at n = 64, L0 = 8; 64 at n = 120, L0 = 30): stepping
// 0.75-0.77 against 0.81 ms where an entry has one batch of terms, 2.28
// against 1.93 ms where it has up to four.
template <bool STEP = false>
__device__ __forceinline__ float2 sofft_f_entry(const float2* __restrict__ s, const float* __restrict__ dj,
                                                int L0, int u, int v) {
  const int l0 = abs(u) > abs(v) ? abs(u) : abs(v);
  float ax = 0.f, ay = 0.f;
  int id = mch::dindex(l0, u, v);
  for (int lb = l0; lb < L0; lb += 8) {
    float dv[8];
    float2 sv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (!STEP) id = mch::dindex(min(lb + k, L0 - 1), u, v);
      dv[k] = dj[id];
      sv[k] = s[id];
      const int l = lb + k;
      if (STEP && l + 1 < L0) id += (2 * l + 1) * (2 * l + 1) + 4 * l + 2 * u + 4;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float w = lb + k < L0 ? dv[k] : 0.f;
      ax += w * sv[k].x;
      ay += w * sv[k].y;
    }
  }
  return make_float2(ax, ay);
}

// One workgroup per (particle, beta row): F(u,v) = sum_l d^l_{uv}(beta_j) sigma,
// 2-D inverse DFT over (alpha, gamma), local top-C of the n*n slice.
template <int MAXPER>
__global__ __launch_bounds__(256) void sofft_slice_kernel(
    const float2* __restrict__ sig, long long ldsig, const float* __restrict__ dgrid,
    const float2* __restrict__ twiddle, int L0, int n, int nsz, int C,
    float* __restrict__ part_val, int* __restrict__ part_idx, float* __restrict__ grid_out) {
  extern __shared__ float4 smem4[];
  float2* smem = reinterpret_cast<float2*>(smem4);
  const int j = blockIdx.x;
  const int b = blockIdx.y;
  const int W = 2 * L0 - 1;
  float2* F = smem;            // W*W
  float2* G = F + W * W;       // W*n
  float2* tw = G + W * n;      // n
  const int tid = threadIdx.x;
  for (int k = tid; k < n; k += blockDim.x) tw[k] = twiddle[k];
  const float2* s = sig + (long long)b * ldsig;
  const float* dj = dgrid + (long long)j * nsz;
  for (int q = tid; q < W * W; q += blockDim.x) F[q] = sofft_f_entry(s, dj, L0, q / W - (L0 - 1), q % W - (L0 - 1));
  __syncthreads();
  // stage A: G(u, g) = sum_v F(u, v) e^{2 pi i v g / n}
  for (int q = tid; q < W * n; q += blockDim.x) {
    const int ui = q / n, g = q % n;
    float ax = 0.f, ay = 0.f;
    for (int vi = 0; vi < W; ++vi) {
      const int v = vi - (L0 - 1);
      int k = (v * g) % n;
      if (k < 0) k += n;
      const float2 f = F[ui * W + vi], t = tw[k];
      ax += f.x * t.x - f.y * t.y;
      ay += f.x * t.y + f.y * t.x;
    }
    G[q] = make_float2(ax, ay);
  }
  __syncthreads();
  // stage B: f(a, g) = Re sum_u G(u, g) e^{2 pi i u a / n} / n^2 ; values stay
  // in registers (static indexing: MAXPER >= ceil(n*n / 256))
  const float inv = 1.0f / ((float)n * (float)n);
  const int nn = n * n;
  unsigned vals[MAXPER];  // float_key of the values (0: padding)
#pragma unroll
  for (int r = 0; r < MAXPER; ++r) {
    const int q = tid + r * 256;
    float acc = -INFINITY;
    if (q < nn) {
      const int a = q / n, g = q % n;
      acc = 0.f;
      for (int ui = 0; ui < W; ++ui) {
        const int u = ui - (L0 - 1);
        int k = (u * a) % n;
        if (k < 0) k += n;
        const float2 gv = G[ui * n + g], t = tw[k];
        acc += gv.x * t.x - gv.y * t.y;
      }
      acc *= inv;
    }
    vals[r] = q < nn ? float_key(acc) : 0u;
    if (grid_out != nullptr && q < nn) grid_out[((long long)b * n + j) * nn + q] = acc;
  }
  block_top_c(vals, [&](int r) { return tid + r * 256; }, C, ((long long)b * n + j) * C, j * nn,
              part_val, part_idx);
}

// Same slice as sofft_slice_kernel for n <= 16 R: the phase table
// E[u][x] = e^{2 pi i u x / n} (|u| < L0) is built once in LDS and serves both
// DFT stages, and each thread owns R x R (alpha, gamma) points of stage B, so
// every pair of E / G loads feeds R outputs.  The 16 x 16 threads cover
// NP = 16 R >= n points per axis: for n < NP the tables carry NP columns (the
// surplus ones repeat phases mod n) and the points at alpha or gamma >= n
// are computed but neither stored nor ranked.
template <int R>
__global__ __launch_bounds__(256) void sofft_slice_blk_kernel(
    const float2* __restrict__ sig, long long ldsig, const float* __restrict__ dgrid,
    const float2* __restrict__ etab, int n, int L0, int nsz, int C, float* __restrict__ part_val,
    int* __restrict__ part_idx, float* __restrict__ grid_out) {
  constexpr int NP = 16 * R;
  const int nn = n * n;
  extern __shared__ float4 smem4[];
  float2* smem = reinterpret_cast<float2*>(smem4);
  const int j = blockIdx.x;
  const int b = blockIdx.y;
  const int W = 2 * L0 - 1;
  float2* F = smem;            // W*W
  float2* G = F + W * W;       // W*NP  G[ui][g]
  float2* E = G + W * NP;      // W*NP  E[ui][x]
  const int tid = threadIdx.x;
  // the plan's phase table, eight loads in flight per thread
  for (int q0 = tid; q0 < W * NP; q0 += 8 * 256) {
    float2 t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = etab[min(q0 + 256 * k, W * NP - 1)];
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (q0 + 256 * k < W * NP) E[q0 + 256 * k] = t[k];
  }
  const float2* s = sig + (long long)b * ldsig;
  const float* dj = dgrid + (long long)j * nsz;
  // stepped indices where every entry has at most one batch of terms
  if (L0 <= 8) {
    for (int q = tid; q < W * W; q += 256) F[q] = sofft_f_entry<true>(s, dj, L0, q / W - (L0 - 1), q % W - (L0 - 1));
  } else {
    for (int q = tid; q < W * W; q += 256) F[q] = sofft_f_entry<false>(s, dj, L0, q / W - (L0 - 1), q % W - (L0 - 1));
  }
  __syncthreads();
  if constexpr (R % 4 == 0) {
    // both DFT stages on the FP32 matrix cores (v_mfma_f32_16x16x4_f32: lane
    // l holds A[l % 16][l / 16] and B[l / 16][l % 16] of a 16 x 4 by 4 x 16
    // product, and rows 4 (l / 16) + j, column l % 16 of the 16 x 16 sum);
    // a complex product is four real ones, K runs over 4 terms per step
    // (rows beyond W read the last row, their terms zeroed)
    typedef float v4f __attribute__((ext_vector_type(4)));
    const int lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
    const int KS = (W + 3) >> 2, RT = (W + 15) >> 4;
    // stage A: G(u, g) over (16-row tile rt of u, 16-column tile tc of g) pairs
    for (int pr = wave; pr < RT * R; pr += 4) {
      const int rt = pr / R, tc = pr - (pr / R) * R;
      const int uu = min(16 * rt + lr, W - 1);
      v4f gr = {0.f, 0.f, 0.f, 0.f}, gi = {0.f, 0.f, 0.f, 0.f};
      for (int ks = 0; ks < KS; ++ks) {
        const int v = 4 * ks + lk, vv = min(v, W - 1);
        float2 f = F[uu * W + vv];
        if (v >= W) f = make_float2(0.f, 0.f);
        const float2 e = E[vv * NP + 16 * tc + lr];
        gr = __builtin_amdgcn_mfma_f32_16x16x4f32(f.x, e.x, gr, 0, 0, 0);
        gr = __builtin_amdgcn_mfma_f32_16x16x4f32(-f.y, e.y, gr, 0, 0, 0);
        gi = __builtin_amdgcn_mfma_f32_16x16x4f32(f.x, e.y, gi, 0, 0, 0);
        gi = __builtin_amdgcn_mfma_f32_16x16x4f32(f.y, e.x, gi, 0, 0, 0);
      }
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int u = 16 * rt + 4 * lk + jj;
        if (u < W) G[u * NP + 16 * tc + lr] = make_float2(gr[jj], gi[jj]);
      }
    }
    __syncthreads();
    // stage B: f(a, g) = sum_u Er(u, a) Gr(u, g) - Ei(u, a) Gi(u, g); the wave
    // owns alpha tiles wave + 4 t (t < R / 4) and every gamma tile
    constexpr int TR = R / 4;
    v4f acc[TR][R];
#pragma unroll
    for (int t = 0; t < TR; ++t)
#pragma unroll
      for (int c = 0; c < R; ++c) acc[t][c] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < KS; ++ks) {
      const int u = 4 * ks + lk, uu = min(u, W - 1);
      const bool ok = u < W;
      float ar[TR], ai[TR];
#pragma unroll
      for (int t = 0; t < TR; ++t) {
        const float2 e = E[uu * NP + 16 * (wave + 4 * t) + lr];
        ar[t] = ok ? e.x : 0.f;
        ai[t] = ok ? -e.y : 0.f;
      }
#pragma unroll
      for (int c = 0; c < R; ++c) {
        const float2 gv = G[uu * NP + 16 * c + lr];
#pragma unroll
        for (int t = 0; t < TR; ++t) {
          acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[t], gv.x, acc[t][c], 0, 0, 0);
          acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[t], gv.y, acc[t][c], 0, 0, 0);
        }
      }
    }
    const float inv = 1.0f / ((float)n * (float)n);
    unsigned keys[R * R];  // float_key of the values, i = (4 t + jj) R + c
#pragma unroll
    for (int t = 0; t < TR; ++t)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
#pragma unroll
        for (int c = 0; c < R; ++c) {
          const float v = acc[t][c][jj] * inv;
          const int a = 16 * (wave + 4 * t) + 4 * lk + jj, g = 16 * c + lr;
          const bool in = NP == n || (a < n && g < n);
          keys[(4 * t + jj) * R + c] = in ? float_key(v) : 0u;
          if (grid_out != nullptr && in) grid_out[((long long)b * n + j) * nn + a * n + g] = v;
        }
    // flat ids a n + g rise with i: alpha with (t, jj), gamma with c
    block_top_c(keys,
                [&](int i) {
                  const int tj = i / R;
                  return (16 * (wave + 4 * (tj >> 2)) + 4 * lk + (tj & 3)) * n + 16 * (i % R) + lr;
                },
                C, ((long long)b * n + j) * C, j * nn, part_val, part_idx);
    return;
  } else {
  // stage A: G(u, g) = sum_v F(u, v) e^{2 pi i v g / n}
  constexpr int SA = 256 % NP == 0 ? 256 / NP : 1, KR = 8;
  if (256 % NP == 0 && W <= KR * SA) {
    // a thread owns column g of rows ui0, ui0 + S, ... (S = 256 / NP, at most
    // KR rows): one E read per term serves all its rows
    constexpr int S = SA;
    const int g = tid % NP, ui0 = tid / NP;
    float2 acc[KR];
#pragma unroll
    for (int k = 0; k < KR; ++k) acc[k] = make_float2(0.f, 0.f);
    for (int vi = 0; vi < W; ++vi) {
      const float2 t = E[vi * NP + g];
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        const int ui = ui0 + k * S;
        if (ui < W) {
          const float2 f = F[ui * W + vi];
          acc[k].x += f.x * t.x - f.y * t.y;
          acc[k].y += f.x * t.y + f.y * t.x;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < KR; ++k)
      if (ui0 + k * S < W) G[(ui0 + k * S) * NP + g] = acc[k];
  } else {
    // register blocks: thread (ty, tx) owns rows ty + 16 k (four at a time)
    // of columns tx + 16 c, c < R; per term 4 F and R E reads feed 4 R
    // complex products
    const int ty = tid >> 4, tx = tid & 15;
    for (int kb = 0; kb < W; kb += 64) {
      float2 acc[4][R];
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int c = 0; c < R; ++c) acc[k][c] = make_float2(0.f, 0.f);
      int urow[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) urow[k] = min(kb + ty + 16 * k, W - 1) * W;
      for (int vi = 0; vi < W; ++vi) {
        float2 f[4], t[R];
#pragma unroll
        for (int k = 0; k < 4; ++k) f[k] = F[urow[k] + vi];
#pragma unroll
        for (int c = 0; c < R; ++c) t[c] = E[vi * NP + tx + 16 * c];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
          for (int c = 0; c < R; ++c) {
            acc[k][c].x += f[k].x * t[c].x - f[k].y * t[c].y;
            acc[k][c].y += f[k].x * t[c].y + f[k].y * t[c].x;
          }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (kb + ty + 16 * k < W)
#pragma unroll
          for (int c = 0; c < R; ++c) G[(kb + ty + 16 * k) * NP + tx + 16 * c] = acc[k][c];
    }
  }
  __syncthreads();
  // stage B: f(a, g) = Re sum_u e^{2 pi i u a / n} G(u, g) / n^2 on the
  // thread's points a = ty + 16 r, g = tx + 16 c (r, c < R): the G reads of a
  // wave are consecutive, its E reads four broadcasts
  const int ty = tid >> 4, tx = tid & 15;
  float vals[R * R];
#pragma unroll
  for (int i = 0; i < R * R; ++i) vals[i] = 0.f;
  for (int ui = 0; ui < W; ++ui) {
    float2 e[R], gv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      e[r] = E[ui * NP + ty + 16 * r];
      gv[r] = G[ui * NP + tx + 16 * r];
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int c = 0; c < R; ++c) vals[r * R + c] += e[r].x * gv[c].x - e[r].y * gv[c].y;
  }
  const float inv = 1.0f / ((float)n * (float)n);
  unsigned keys[R * R];  // float_key of the values
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int c = 0; c < R; ++c) {
      const float v = vals[r * R + c] * inv;
      const int a = ty + 16 * r, g = tx + 16 * c;
      const bool in = NP == n || (a < n && g < n);
      keys[r * R + c] = in ? float_key(v) : 0u;
      if (grid_out != nullptr && in) grid_out[((long long)b * n + j) * nn + a * n + g] = v;
    }
  // the thread's flat ids a n + g rise with r R + c, so its first maximum is
  // the smallest id among equal values
  block_top_c(keys, [&](int i) { return (ty + 16 * (i / R)) * n + tx + 16 * (i % R); }, C,
              ((long long)b * n + j) * C, j * nn, part_val, part_idx);
  }
}

// Merge the per-row candidates into the global top-C per particle and convert
// flat ids to Euler angles (SOFFT.ids_to_angles).
__global__ __launch_bounds__(256) void sofft_merge_kernel(
    const float* __restrict__ part_val, const int* __restrict__ part_idx, int n, int C,
    float* __restrict__ out_a, float* __restrict__ out_b, float* __restrict__ out_g,
    float* __restrict__ out_s, int out_ld) {
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int tot = n * C;
  __shared__ float sv[4];
  __shared__ int si[4];
  __shared__ int taken_slot;
  extern __shared__ unsigned char used[];  // tot bytes
  for (int q = tid; q < tot; q += blockDim.x) used[q] = 0;
  __syncthreads();
  const float* pv = part_val + (long long)b * tot;
  const int* pi = part_idx + (long long)b * tot;
  const int lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  const float bstep = (float)(M_PI / (double)(n - 1));
  const float astep = (float)(2.0 * M_PI / (double)n);
  for (int c = 0; c < C; ++c) {
    float bv = -INFINITY;
    int bi = 0x7fffffff, bslot = -1;
    for (int q = tid; q < tot; q += blockDim.x) {
      if (used[q] || pi[q] < 0) continue;
      const float v = pv[q];
      const int id = pi[q];
      if (v > bv || (v == bv && id < bi)) { bv = v; bi = id; bslot = q; }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const float ov = __shfl_xor(bv, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      const int os = __shfl_xor(bslot, off, 64);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; bslot = os; }
    }
    __shared__ int sslot[4];
    if (lane == 0) { sv[wave] = bv; si[wave] = bi; sslot[wave] = bslot; }
    __syncthreads();
    if (tid == 0) {
      float v = sv[0];
      int id = si[0], sl = sslot[0];
      for (int w = 1; w < nw; ++w)
        if (sv[w] > v || (sv[w] == v && si[w] < id)) { v = sv[w]; id = si[w]; sl = sslot[w]; }
      taken_slot = sl;
      const long long o = (long long)b * out_ld + c;
      if (id == 0x7fffffff) {
        // fewer than C grid points: mirror the reference's -1 fill
        out_b[o] = -bstep; out_a[o] = -astep; out_g[o] = -astep; out_s[o] = -1.f;
      } else {
        const int nn = n * n;
        const int ib = id / nn, ia = (id / n) % n, ig = id % n;
        out_b[o] = (float)ib * bstep;
        out_a[o] = (float)ia * astep;
        out_g[o] = (float)ig * astep;
        out_s[o] = v;
      }
    }
    __syncthreads();
    if (tid == 0 && taken_slot >= 0) used[taken_slot] = 1;
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// fused Newton march
// ---------------------------------------------------------------------------
struct MarchArgs {
  const float2* sig_pk;  // (B, pk_stride)
  int pk_stride;
  int B, C;
  float* alpha;  // (B*C) in/out
  float* beta;
  float* gamma;
  float* score;  // (B*C) out
  int n_levels;
  int levels[16];
  int num_steps;
  int newton;    // 1: Newton step, 0: gradient step
  float base_lr;
  int Lmax;
  // stepped launches (stop_early): this launch runs level li0, step st0 only,
  // with the optimiser state kept between launches in state (B*C x 16
  // doubles) and accept[li * (num_steps + 1) + st] counting the candidates of
  // the whole batch that step st of level li accepted (zeroed per search)
  int li0, st0;
  double* state;
  int* accept;
  // per-particle march ends (B ints on the device, NULL: every particle
  // marches `levels`): particle b marches the levels below lend[b], then
  // lend[b] itself -- the schedule that sorted(Li) + [L_end] gives a Matcha of
  // L_max = lend[b] (Matcha.py:76,381-386); a value outside [1, Lmax] makes the
  // particle's candidates NaN
  const int* lend;
};

constexpr int kMarchThreads = 256;
constexpr int kMarchWaves = kMarchThreads / 64;
constexpr int kMaxLevelsMarch = 16;
constexpr int kMaxL = 96;

struct EvalOut {
  float v[10];  // s, ga, gb, gg, Haa, Hab, Hag, Hbb, Hbg, Hgg
};

// Per-l constants of the three-term l-recurrence of d^l_{m'm}:
// (l^2, l (2l - 1), 1 / ((l - 1) l) (0 for l <= 1), 1 / (l (2l + 1))),
// rounded from FP64; built once per workgroup in LDS.
__device__ void fill_lrec(float4* lrec, int L) {
  for (int l = threadIdx.x; l <= L + 1; l += blockDim.x) {
    const double dl = (double)l;
    lrec[l] = make_float4((float)(dl * dl), (float)(dl * (2.0 * dl - 1.0)),
                          l > 1 ? (float)(1.0 / ((dl - 1.0) * dl)) : 0.f,
                          l > 0 ? (float)(1.0 / (dl * (2.0 * dl + 1.0))) : 0.f);
  }
}

// Evaluate score, gradient and Hessian of Re sum_{l<=Li} D(a,b,g) sigma.
// Lanes own fundamental pairs p = mp^2 + mp + m (0 <= |m| <= mp <= Li) and
// march l upwards from the task's smallest mp; the repacked sigma of shell l
// starts at byte 4 pk_offset(l), so one wave-uniform offset (advanced by
// 16 (l + 1)^2 bytes) plus the lane's 16 p addresses both merged image pairs.
// The evaluation's functions of beta, computed by one thread and read by all
// from LDS: cos b, sin b, c = cos b/2, s = sin b/2, log2 |c|, log2 |s|, c / s,
// s / c.
constexpr int kBetaScalars = 8;
__device__ __forceinline__ void beta_scalars(float beta, float* bs) {
  const float hc = cosf(0.5f * beta), hs = sinf(0.5f * beta);
  bs[0] = cosf(beta);
  bs[1] = sinf(beta);
  bs[2] = hc;
  bs[3] = hs;
  bs[4] = log2f(fabsf(hc));
  bs[5] = log2f(fabsf(hs));
  bs[6] = hc / hs;
  bs[7] = hs / hc;
}

// SCORE_ONLY: the score alone (out[0]; out[1..9] are zero): only the d
// recurrence and its sums, in the same order of operations as the full
// evaluation, so the score is bit for bit the same.
template <bool SCORE_ONLY = false>
__device__ void march_eval(const float2* __restrict__ sp, unsigned sp_bytes, int Li, const float* bs,
                           const float2* __restrict__ za, const float2* __restrict__ zg,
                           const float* __restrict__ lgf, const float4* __restrict__ lrec, float* red,
                           float out[10]) {
  // the wave index through readfirstlane: the compiler then keeps every
  // per-task quantity (t, lstart, l, the sigma offset) in scalar registers
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  float acc[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) acc[i] = 0.f;
  const float x = bs[0], sx = bs[1], hc = bs[2], hs = bs[3], lc = bs[4], ls = bs[5];
  const bool cneg = hc < 0.f, sneg = hs < 0.f;
  const bool ratios = fabsf(hc) >= 0x1p-20f && fabsf(hs) >= 0x1p-20f;
  const float rcs = bs[6], rsc = bs[7];
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(sp), (short)0, (int)sp_bytes, 0x00020000);
  const int P = (Li + 1) * (Li + 1);
  const int ntasks = (P + 63) / 64;
  // snake order over the waves: early tasks (small mp) march the longest
  for (int k = 0; k * kMarchWaves < ntasks; ++k) {
    const int t = k * kMarchWaves + ((k & 1) ? (kMarchWaves - 1 - wave) : wave);
    if (t >= ntasks) continue;
    const int p = t * 64 + lane;
    const bool valid = p < P;
    const int pc = valid ? p : P - 1;
    const int mp = mch::isqrt_small(pc);
    const int m = pc - mp * mp - mp;
    // mp = floor(sqrt(p)) rises with p: the task's smallest start degree is its
    // first lane's, the largest among its valid lanes its last valid lane's
    // (both wave-uniform)
    const int lstart = __builtin_amdgcn_readfirstlane(mch::isqrt_small(t * 64));
    const int mplast = __builtin_amdgcn_readfirstlane(mch::isqrt_small(min(t * 64 + 63, P - 1)));
    // seed at l0 = mp: d = (-1)^{mp-m} sqrt(C(2mp, mp+m)) c^{mp+m} s^{mp-m}
    const int a = mp + m, bb = mp - m;
    const float l2K = 0.5f * (lgf[2 * mp] - lgf[a] - lgf[2 * mp - a]);
    auto term = [&](int pa, int pb) -> float {
      if (pa < 0 || pb < 0) return 0.f;
      float e = l2K;
      if (pa) e += (float)pa * lc;
      if (pb) e += (float)pb * ls;
      float v = exp2f(e);
      if ((cneg && (pa & 1)) != (sneg && (pb & 1))) v = -v;
      return v;
    };
    const float sgn = (bb & 1) ? -1.f : 1.f;
    // the derivative seeds need term(a +- k, bb -+ k) = term(a, bb) (c / s)^{+-k},
    // k = 1, 2: away from beta = 0 and pi (|c|, |s| >= 2^-20, a workgroup-uniform
    // test) by ratios, so that a task costs one exponential; every product then
    // stays below 2^40 (|term(a, bb)| <= 1), also where its coefficient is zero
    const float t0 = term(a, bb);
    float tp1 = 0.f, tm1 = 0.f, tp2 = 0.f, tm2 = 0.f;  // term(a +- 1, bb -+ 1), term(a +- 2, bb -+ 2)
    if (SCORE_ONLY) {
    } else if (ratios) {
      tp1 = t0 * rcs;
      tm1 = t0 * rsc;
      tp2 = tp1 * rcs;
      tm2 = tm1 * rsc;
    } else {
      tp1 = term(a + 1, bb - 1);
      tm1 = term(a - 1, bb + 1);
      tp2 = term(a + 2, bb - 2);
      tm2 = term(a - 2, bb + 2);
    }
    float d = sgn * t0;
    float dp = sgn * 0.5f * ((float)bb * tp1 - (float)a * tm1);
    float dpp = sgn * 0.25f *
                ((float)(bb * (bb - 1)) * tp2 - (float)(bb * (a + 1) + a * (bb + 1)) * t0 +
                 (float)(a * (a - 1)) * tm2);
    float pd = 0.f, pdp = 0.f, pdpp = 0.f, Cprev = 0.f;
    // U[i] = sum_l sigma'_i d, U1 with d', U2 with d'' (i: merged images (mp, m), (m, mp))
    float2 U[2], U1[2], U2[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      U[i] = make_float2(0.f, 0.f);
      U1[i] = U[i];
      U2[i] = U[i];
    }
    const float fm = (float)m, fmp = (float)mp;
    const float fm2 = fm * fm, fmp2 = fmp * fmp, mmp = fm * fmp;
    const unsigned voff = 16u * (unsigned)pc;
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    // shell l: the recurrence step to d^l and the accumulation of sigma^l; the
    // guarded form for shells that some lanes of the wave have not reached
    // (l <= mp), the plain form once every lane has (invalid tail lanes then
    // compute values that are never used)
    auto shell = [&](int l, const v4u q, bool guarded) {
      if (!guarded || l > mp) {
        const float4 c = lrec[l];
        const float Pl = (c.x - fm2) * (c.x - fmp2);  // >= 1 for l > mp >= |m|
        const float r = __builtin_amdgcn_rsqf(Pl);
        const float A = c.y * r;
        const float xb = x - mmp * c.z;
        const float nd = A * (xb * d - Cprev * pd);
        if (!SCORE_ONLY) {
          const float ndp = A * (xb * dp - sx * d - Cprev * pdp);
          const float ndpp = A * (xb * dpp - 2.f * sx * dp - x * d - Cprev * pdpp);
          pdp = dp; pdpp = dpp;
          dp = ndp; dpp = ndpp;
        }
        pd = d;
        d = nd;
        Cprev = Pl * r * c.w;
      }
      if (!guarded || (valid && l >= mp)) {
        const float2 sv[2] = {make_float2(__uint_as_float(q.x), __uint_as_float(q.y)),
                              make_float2(__uint_as_float(q.z), __uint_as_float(q.w))};
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          U[i].x += sv[i].x * d;   U[i].y += sv[i].y * d;
          if (!SCORE_ONLY) {
            U1[i].x += sv[i].x * dp; U1[i].y += sv[i].y * dp;
            U2[i].x += sv[i].x * dpp; U2[i].y += sv[i].y * dpp;
          }
        }
      }
    };
    // the largest start degree among the wave's valid lanes
    const int l1 = min(mplast, Li);
    // sigma of shells l + 1 and l + 2 in flight while shell l runs (the offsets
    // are wave-uniform; reads past the particle's sigma return 0)
    unsigned soff = 4u * (unsigned)mch::pk_offset(lstart);
    v4u q0 = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
    soff += 16u * (unsigned)((lstart + 1) * (lstart + 1));
    v4u q1 = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
    soff += 16u * (unsigned)((lstart + 2) * (lstart + 2));
    for (int l = lstart; l <= l1; ++l) {
      const v4u q = q0;
      q0 = q1;
      if (l + 2 <= Li) {
        q1 = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
        soff += 16u * (unsigned)((l + 3) * (l + 3));
      }
      shell(l, q, true);
    }
    // every lane active: two shells per iteration, each shell's sigma loaded
    // two shells ahead into one of two registers, unconditionally (loads past
    // the particle's sigma return 0 and are not used), so that neither a
    // register rotation nor a wait on the newest load is needed; soff is the
    // offset of shell l1 + 3
    v4u qa = q0, qb = q1;
    for (int l = l1 + 1; l <= Li; l += 2) {
      shell(l, qa, false);
      qa = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
      soff += 16u * (unsigned)((l + 3) * (l + 3));
      if (l + 1 > Li) break;
      shell(l + 1, qb, false);
      qb = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
      soff += 16u * (unsigned)((l + 4) * (l + 4));
    }
    if (valid) {
      const int us[2] = {mp, m};
      const int vs[2] = {m, mp};
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int u = us[i], v = vs[i];
        float2 pa = za[abs(u)];
        if (u < 0) pa.y = -pa.y;
        float2 pg = zg[abs(v)];
        if (v < 0) pg.y = -pg.y;
        const float2 ph = make_float2(pa.x * pg.x - pa.y * pg.y, pa.x * pg.y + pa.y * pg.x);
        const float q0r = ph.x * U[i].x - ph.y * U[i].y;
        if (SCORE_ONLY) {
          acc[0] += q0r;
          continue;
        }
        const float q0i = ph.x * U[i].y + ph.y * U[i].x;
        const float q1r = ph.x * U1[i].x - ph.y * U1[i].y;
        const float q1i = ph.x * U1[i].y + ph.y * U1[i].x;
        const float q2r = ph.x * U2[i].x - ph.y * U2[i].y;
        const float fu = (float)u, fv = (float)v;
        acc[0] += q0r;
        acc[1] += -fu * q0i;
        acc[2] += q1r;
        acc[3] += -fv * q0i;
        acc[4] += -fu * fu * q0r;
        acc[5] += -fu * q1i;
        acc[6] += -fu * fv * q0r;
        acc[7] += q2r;
        acc[8] += -fv * q1i;
        acc[9] += -fv * fv * q0r;
      }
    }
  }
  // (the score-only evaluation's other sums are zero and not reduced)
  constexpr int NSUM = SCORE_ONLY ? 1 : 10;
#pragma unroll
  for (int i = 0; i < NSUM; ++i) acc[i] = mch::wave_sum_dpp(acc[i]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 10; ++i) red[i * kMarchWaves + wave] = i < NSUM ? acc[i] : 0.f;
  }
  __syncthreads();
  if (tid == 0) {  // the callers read the sums in thread 0 only
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      float v = 0.f;
      for (int w = 0; w < kMarchWaves; ++w) v += red[i * kMarchWaves + w];
      out[i] = v;
    }
  }
}

// e^{i k theta} for k = 0..L into z (FP64 argument reduction).
__device__ void phase_table(double theta, int L, float2* z) {
  for (int k = threadIdx.x; k <= L; k += blockDim.x) {
    double ang = (double)k * theta;
    ang = ang - 6.283185307179586476925 * rint(ang / 6.283185307179586476925);
    float sn, cs;
    sincosf((float)ang, &sn, &cs);
    z[k] = make_float2(cs, sn);
  }
}

// 3x3 solve by LU with partial pivoting in FP64; returns false if a pivot is
// exactly zero (torch.linalg.solve raises "singular" there, and the reference
// then skips that candidate, Matcha.py:315-343).
__device__ bool solve3(const double Hm[9], const double g[3], double x[3]) {
  double A[3][4];
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) A[r][c] = Hm[r * 3 + c];
    A[r][3] = g[r];
  }
  for (int k = 0; k < 3; ++k) {
    int piv = k;
    double best = fabs(A[k][k]);
    for (int r = k + 1; r < 3; ++r)
      if (fabs(A[r][k]) > best) { best = fabs(A[r][k]); piv = r; }
    if (best == 0.0) return false;
    if (piv != k)
      for (int c = 0; c < 4; ++c) { double t = A[k][c]; A[k][c] = A[piv][c]; A[piv][c] = t; }
    for (int r = k + 1; r < 3; ++r) {
      const double f = A[r][k] / A[k][k];
      for (int c = k; c < 4; ++c) A[r][c] -= f * A[k][c];
    }
  }
  for (int r = 2; r >= 0; --r) {
    double v = A[r][3];
    for (int c = r + 1; c < 3; ++c) v -= A[r][c] * x[c];
    x[r] = v / A[r][r];
  }
  return true;
}

// At least five waves per SIMD (at most 96 VGPRs, a few values spilled): the
// march's evaluation chains hide their latencies better with five resident
// workgroups per CU than with the four that 124 VGPRs allowed (1.24 against
// 1.35 ms per batch of 512; six waves spill 40 VGPRs and lose again).
//
// STEPPED (stop_early, Matcha.py:198-199): one launch per (level, step); a
// level's remaining steps are skipped once one of its steps accepted no
// candidate of the batch, read from the counts of the previous launches (the
// launches are ordered on the stream, so no host synchronisation is needed).
template <bool STEPPED>
__global__ __launch_bounds__(kMarchThreads, 5) void so3_march_kernel(MarchArgs args,
                                                                   const float* __restrict__ lgf_g) {
  if (STEPPED && args.st0 >= 2) {
    const int* acc_l = args.accept + args.li0 * (args.num_steps + 1);
    for (int st = 1; st < args.st0; ++st)
      if (acc_l[st] == 0) return;  // the whole workgroup (uniform): the level has stopped
  }
  __shared__ float2 za[kMaxL + 1], zg[kMaxL + 1];
  __shared__ float lgf[2 * kMaxL + 2];
  __shared__ float red[10 * kMarchWaves];
  __shared__ float ang[4];  // trial alpha, beta, gamma
  __shared__ float bsc[kBetaScalars];
  __shared__ float4 lrec[kMaxL + 2];
  const int nblk = args.B * args.C;
  const int q = mch::xcd_remap(blockIdx.x, nblk);
  const int b = q / args.C;
  const int tid = threadIdx.x;
  for (int k = tid; k <= 2 * args.Lmax + 1; k += blockDim.x) lgf[k] = lgf_g[k];
  fill_lrec(lrec, args.Lmax);
  const unsigned sp_bytes = (unsigned)args.pk_stride * 8u;
  const float2* sp = args.sig_pk + (long long)b * args.pk_stride;
  // the optimiser state lives in LDS and only thread 0 touches it, so it
  // holds no registers in the other lanes during the evaluation
  __shared__ double gcur[3], Hcur[9];
  __shared__ float opt[5];  // current alpha, beta, gamma, best score, lr
  double* stq = STEPPED ? args.state + (long long)q * 16 : nullptr;
  if (tid == 0) {
    if (STEPPED && (args.li0 > 0 || args.st0 > 0)) {
      for (int i = 0; i < 5; ++i) opt[i] = (float)stq[i];
      for (int i = 0; i < 3; ++i) gcur[i] = stq[5 + i];
      Hcur[0] = stq[8]; Hcur[1] = Hcur[3] = stq[9]; Hcur[2] = Hcur[6] = stq[10];
      Hcur[4] = stq[11]; Hcur[5] = Hcur[7] = stq[12]; Hcur[8] = stq[13];
    } else {
      opt[0] = args.alpha[q]; opt[1] = args.beta[q]; opt[2] = args.gamma[q];
      opt[3] = 0.f; opt[4] = 0.f;
      for (int i = 0; i < 3; ++i) gcur[i] = 0.0;
      for (int i = 0; i < 9; ++i) Hcur[i] = 0.0;
    }
  }
  float ev[10];
  // the workgroup's schedule: the shared one, or the levels below the
  // particle's own march end followed by that end
  __shared__ int slev[kMaxLevelsMarch + 1];
  __shared__ int snlev;
  if (tid == 0) {
    int n = 0;
    if (!STEPPED && args.lend) {
      const int le = args.lend[b];
      if (le >= 1 && le <= args.Lmax) {
        for (int i = 0; i < args.n_levels; ++i)
          if (args.levels[i] < le) slev[n++] = args.levels[i];
        slev[n++] = le;
      }
    } else {
      for (int i = 0; i < args.n_levels; ++i) slev[n++] = args.levels[i];
    }
    snlev = n;
  }
  __syncthreads();
  const int nlev = snlev;
  if (nlev == 0) {  // invalid march end: no level to march
    if (tid == 0) {
      const float nan = __builtin_nanf("");
      args.alpha[q] = nan; args.beta[q] = nan; args.gamma[q] = nan; args.score[q] = nan;
    }
    return;
  }
  const int li_begin = STEPPED ? args.li0 : 0, li_end = STEPPED ? args.li0 + 1 : nlev;
  for (int li = li_begin; li < li_end; ++li) {
    const int Li = slev[li];
    const int st_begin = STEPPED ? args.st0 : 0, st_end = STEPPED ? args.st0 : args.num_steps;
    for (int st = st_begin; st <= st_end; ++st) {
      // thread 0 proposes the point to evaluate; everybody reads it from LDS
      if (tid == 0) {
        const float ca = opt[0], cb = opt[1], cg = opt[2], lr = opt[4];
        float ta = ca, tb = cb, tg = cg;
        if (st > 0) {
          double Hl[9], gl[3], stepd[3];
          for (int i = 0; i < 9; ++i) Hl[i] = Hcur[i];
          for (int i = 0; i < 3; ++i) gl[i] = gcur[i];
          if (args.newton) {
            if (!solve3(Hl, gl, stepd)) { stepd[0] = stepd[1] = stepd[2] = 0.0; }
          } else {
            stepd[0] = -gl[0]; stepd[1] = -gl[1]; stepd[2] = -gl[2];
          }
          ta = ca - lr * (float)stepd[0];
          tb = cb - lr * (float)stepd[1];
          tg = cg - lr * (float)stepd[2];
        }
        ang[0] = ta; ang[1] = tb; ang[2] = tg;
      }
      __syncthreads();
      const float ta = ang[0], tb = ang[1], tg = ang[2];
      // the last wave computes the beta scalars while the first fills the phases
      if (tid == kMarchThreads - 1) beta_scalars(tb, bsc);
      phase_table((double)ta, Li, za);
      phase_table((double)tg, Li, zg);
      __syncthreads();
      // the last trial of a level needs its score only (its gradient and
      // Hessian would seed a step that is not taken: the next level starts by
      // evaluating its own)
      if (st == args.num_steps && st > 0)
        march_eval<true>(sp, sp_bytes, Li, bsc, za, zg, lgf, lrec, red, ev);
      else
        march_eval<false>(sp, sp_bytes, Li, bsc, za, zg, lgf, lrec, red, ev);
      if (tid == 0) {
        const bool first = (st == 0);
        const bool acc = first || (ev[0] > opt[3]);
        float lr = opt[4];
        if (!first) lr = acc ? lr * 1.1f : lr * 0.5f;
        if (acc) {
          opt[0] = ta; opt[1] = tb; opt[2] = tg;
          opt[3] = ev[0];
          gcur[0] = ev[1]; gcur[1] = ev[2]; gcur[2] = ev[3];
          Hcur[0] = ev[4]; Hcur[1] = ev[5]; Hcur[2] = ev[6];
          Hcur[3] = ev[5]; Hcur[4] = ev[7]; Hcur[5] = ev[8];
          Hcur[6] = ev[6]; Hcur[7] = ev[8]; Hcur[8] = ev[9];
        }
        if (first) lr = args.base_lr;
        opt[4] = lr;
        if (STEPPED && !first && acc) atomicAdd(args.accept + li * (args.num_steps + 1) + st, 1);
      }
      __syncthreads();
    }
  }
  if (STEPPED && tid == 0) {
    for (int i = 0; i < 5; ++i) stq[i] = (double)opt[i];
    for (int i = 0; i < 3; ++i) stq[5 + i] = gcur[i];
    stq[8] = Hcur[0]; stq[9] = Hcur[1]; stq[10] = Hcur[2];
    stq[11] = Hcur[4]; stq[12] = Hcur[5]; stq[13] = Hcur[8];
  }
  if (tid == 0) {
    args.alpha[q] = opt[0];
    args.beta[q] = opt[1];
    args.gamma[q] = opt[2];
    args.score[q] = opt[3];
  }
}

// Score, gradient and Hessian at given angles (one workgroup per candidate);
// out[q*10 + k] = s, ga, gb, gg, Haa, Hab, Hag, Hbb, Hbg, Hgg.
__global__ __launch_bounds__(kMarchThreads) void so3_eval_kernel(
    const float2* __restrict__ sig_pk, int pk_stride, int C, int L, const float* __restrict__ al,
    const float* __restrict__ be, const float* __restrict__ ga, const float* __restrict__ lgf_g,
    float* __restrict__ out) {
  __shared__ float2 za[kMaxL + 1], zg[kMaxL + 1];
  __shared__ float lgf[2 * kMaxL + 2];
  __shared__ float red[10 * kMarchWaves];
  __shared__ float4 lrec[kMaxL + 2];
  const int q = blockIdx.x;
  const int b = q / C;
  for (int k = threadIdx.x; k <= 2 * L + 1; k += blockDim.x) lgf[k] = lgf_g[k];
  fill_lrec(lrec, L);
  __shared__ float bsc[kBetaScalars];
  if (threadIdx.x == kMarchThreads - 1) beta_scalars(be[q], bsc);
  phase_table((double)al[q], L, za);
  phase_table((double)ga[q], L, zg);
  __syncthreads();
  float ev[10];
  march_eval(sig_pk + (long long)b * pk_stride, (unsigned)pk_stride * 8u, L, bsc, za, zg, lgf, lrec, red,
             ev);
  if (threadIdx.x == 0)
    for (int k = 0; k < 10; ++k) out[(long long)q * 10 + k] = ev[k];
}

__global__ void so3_zero_kernel(int* __restrict__ p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// argmax over candidates (first maximum), Matcha.reduce_over_candidates.
__global__ void reduce_candidates_kernel(const float* __restrict__ a, const float* __restrict__ bt,
                                         const float* __restrict__ g, const float* __restrict__ s,
                                         int B, int C, float* oa, float* ob, float* og, float* os) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  int bi = 0;
  float bv = s[(long long)b * C];
  for (int c = 1; c < C; ++c) {
    const float v = s[(long long)b * C + c];
    if (v > bv) { bv = v; bi = c; }
  }
  oa[b] = a[(long long)b * C + bi];
  ob[b] = bt[(long long)b * C + bi];
  og[b] = g[(long long)b * C + bi];
  os[b] = bv;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* mch_last_error(void) { return mch::get_error(); }

// The blocked slice kernel's R for an n-point grid (0: the grid is left to
// sofft_slice_kernel), and its LDS bytes: F, then G and E of 16 R columns.
static int sofft_blk_r(int L0, int n) {
  const int R = (n + 15) / 16;
  const size_t W = 2 * L0 - 1;
  return R <= 8 && (W * W + 2 * W * 16 * R) * sizeof(float2) <= 160 * 1024 ? R : 0;
}
static size_t sofft_blk_lds(int L0, int R) {
  const size_t W = 2 * L0 - 1;
  return (W * W + 2 * W * 16 * R) * sizeof(float2);
}
static const void* sofft_blk_fn(int R) {
  switch (R) {
    case 1: return (const void*)sofft_slice_blk_kernel<1>;
    case 2: return (const void*)sofft_slice_blk_kernel<2>;
    case 3: return (const void*)sofft_slice_blk_kernel<3>;
    case 4: return (const void*)sofft_slice_blk_kernel<4>;
    case 5: return (const void*)sofft_slice_blk_kernel<5>;
    case 6: return (const void*)sofft_slice_blk_kernel<6>;
    case 7: return (const void*)sofft_slice_blk_kernel<7>;
    default: return (const void*)sofft_slice_blk_kernel<8>;
  }
}

int mch_so3_plan_create(int L_max, int L0, int K, int max_batch, int max_candidates,
                        void** out_plan) {
  if (!out_plan) { mch::set_error("out_plan is NULL"); return MCH_ERR_ARG; }
  *out_plan = nullptr;
  if (L_max < 1 || L_max > kMaxL || L0 < 1 || L0 > L_max + 1 || K < 1 || max_batch < 1 ||
      max_candidates < 1) {
    mch::set_error("invalid so3 plan arguments (L_max=%d L0=%d K=%d batch=%d C=%d)", L_max, L0, K,
                   max_batch, max_candidates);
    return MCH_ERR_ARG;
  }
  const int n = 2 * L0 * K;
  const int W = 2 * L0 - 1;
  const size_t lds = (size_t)(W * W + W * n + n) * sizeof(float2);
  if (lds > 160 * 1024 || (n * n + 255) / 256 > 64) {
    mch::set_error("SOFFT grid too large for one workgroup (L0=%d, n=%d)", L0, n);
    return MCH_ERR_UNSUPPORTED;
  }
  So3Plan* p = new So3Plan();
  memset(p, 0, sizeof(*p));
  p->Lmax = L_max; p->L0 = L0; p->K = K; p->n = n;
  p->max_batch = max_batch; p->max_cand = max_candidates;
  p->nsz = mch::dsize(L0 - 1);
  std::vector<double> dd(mch::dsize(L0 - 1));
  std::vector<float> dg((size_t)n * p->nsz);
  for (int jb = 0; jb < n; ++jb) {
    const double beta = (n == 1) ? 0.0 : M_PI * (double)jb / (double)(n - 1);
    mch::host_wigner_d(beta, L0 - 1, dd.data());
    for (int k = 0; k < p->nsz; ++k) dg[(size_t)jb * p->nsz + k] = (float)dd[k];
  }
  std::vector<float2> tw(n);
  for (int k = 0; k < n; ++k) {
    const double a = 2.0 * M_PI * (double)k / (double)n;
    tw[k] = make_float2((float)cos(a), (float)sin(a));
  }
  std::vector<float> lg(2 * kMaxL + 2);
  for (int k = 0; k < (int)lg.size(); ++k) lg[k] = (float)(lgamma((double)k + 1.0) / M_LN2);
  int rc = MCH_OK;
#define MCH_TRY(expr, what) do { if ((rc = mch::hip_check((expr), what)) != MCH_OK) goto fail; } while (0)
  if (sofft_blk_r(L0, n) > 0 && sofft_blk_lds(L0, sofft_blk_r(L0, n)) > 64 * 1024)
    MCH_TRY(hipFuncSetAttribute(sofft_blk_fn(sofft_blk_r(L0, n)), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)sofft_blk_lds(L0, sofft_blk_r(L0, n))),
            "hipFuncSetAttribute(sofft_slice_blk_kernel)");
  MCH_TRY(hipMalloc(&p->d_grid, dg.size() * sizeof(float)), "hipMalloc(d_grid)");
  MCH_TRY(hipMemcpy(p->d_grid, dg.data(), dg.size() * sizeof(float), hipMemcpyHostToDevice), "copy d_grid");
  MCH_TRY(hipMalloc(&p->twiddle, n * sizeof(float2)), "hipMalloc(twiddle)");
  MCH_TRY(hipMemcpy(p->twiddle, tw.data(), n * sizeof(float2), hipMemcpyHostToDevice), "copy twiddle");
  if (sofft_blk_r(L0, n) > 0) {
    // E[u][x] = e^{2 pi i (u - (L0 - 1)) x / n}, x < 16 R (the columns past n
    // repeat phases mod n), from the same twiddles
    const int NP = 16 * sofft_blk_r(L0, n);
    std::vector<float2> et((size_t)W * NP);
    for (int ui = 0; ui < W; ++ui)
      for (int x = 0; x < NP; ++x) {
        int k = ((ui - (L0 - 1)) * x) % n;
        if (k < 0) k += n;
        et[(size_t)ui * NP + x] = tw[k];
      }
    MCH_TRY(hipMalloc(&p->etab, et.size() * sizeof(float2)), "hipMalloc(etab)");
    MCH_TRY(hipMemcpy(p->etab, et.data(), et.size() * sizeof(float2), hipMemcpyHostToDevice), "copy etab");
  }
  MCH_TRY(hipMalloc(&p->lgf, lg.size() * sizeof(float)), "hipMalloc(lgf)");
  MCH_TRY(hipMemcpy(p->lgf, lg.data(), lg.size() * sizeof(float), hipMemcpyHostToDevice), "copy lgf");
  MCH_TRY(hipMalloc(&p->sig_pk, (size_t)max_batch * mch::pk_size(L_max) * sizeof(float2)), "hipMalloc(sig_pk)");
  MCH_TRY(hipMalloc(&p->part_val, (size_t)max_batch * n * max_candidates * sizeof(float)), "hipMalloc(part_val)");
  MCH_TRY(hipMalloc(&p->part_idx, (size_t)max_batch * n * max_candidates * sizeof(int)), "hipMalloc(part_idx)");
  MCH_TRY(hipMalloc(&p->cand, (size_t)max_batch * max_candidates * 4 * sizeof(float)), "hipMalloc(cand)");
  MCH_TRY(hipMalloc(&p->state, (size_t)max_batch * max_candidates * 16 * sizeof(double)), "hipMalloc(state)");
  MCH_TRY(hipMalloc(&p->accept, (size_t)mch::kMaxLevels * (mch::kMaxStopSteps + 1) * sizeof(int)),
          "hipMalloc(accept)");
  if (lds > 64 * 1024) {
    MCH_TRY(hipFuncSetAttribute((const void*)sofft_slice_kernel<16>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
            "hipFuncSetAttribute(sofft_slice_kernel<16>)");
    MCH_TRY(hipFuncSetAttribute((const void*)sofft_slice_kernel<64>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
            "hipFuncSetAttribute(sofft_slice_kernel<64>)");
  }
#undef MCH_TRY
  *out_plan = p;
  return MCH_OK;
fail:
  (void)hipFree(p->d_grid); (void)hipFree(p->twiddle); (void)hipFree(p->etab); (void)hipFree(p->lgf);
  (void)hipFree(p->sig_pk);
  (void)hipFree(p->part_val); (void)hipFree(p->part_idx); (void)hipFree(p->cand);
  (void)hipFree(p->state); (void)hipFree(p->accept);
  delete p;
  return rc;
}

int mch_so3_plan_destroy(void* plan) {
  So3Plan* p = (So3Plan*)plan;
  if (!p) return MCH_OK;
  (void)hipFree(p->d_grid); (void)hipFree(p->twiddle); (void)hipFree(p->etab); (void)hipFree(p->lgf);
  (void)hipFree(p->sig_pk);
  (void)hipFree(p->part_val); (void)hipFree(p->part_idx); (void)hipFree(p->cand);
  (void)hipFree(p->state); (void)hipFree(p->accept);
  delete p;
  return MCH_OK;
}

// Top-C grid candidates: writes (B, C) alpha/beta/gamma/score.
// One launch of the per-beta-row slice kernel (grid_out: full grid or NULL).
static int launch_sofft_slice(So3Plan* p, const void* sigma, long long ld_sigma, int B, int C,
                              float* grid_out, hipStream_t stream) {
  const int W = 2 * p->L0 - 1;
  const int n = p->n;
  const int R = sofft_blk_r(p->L0, n);
  if (R > 0) {
    const dim3 grid(n, B);
    const size_t lds_blk = sofft_blk_lds(p->L0, R);
#define MCH_BLK(R)                                                                               \
  hipLaunchKernelGGL(sofft_slice_blk_kernel<R>, grid, dim3(256), lds_blk, stream,                \
                     (const float2*)sigma, ld_sigma, p->d_grid, p->etab, n, p->L0, p->nsz, C,     \
                     p->part_val, p->part_idx, grid_out)
    switch (R) {
      case 1: MCH_BLK(1); break;
      case 2: MCH_BLK(2); break;
      case 3: MCH_BLK(3); break;
      case 4: MCH_BLK(4); break;
      case 5: MCH_BLK(5); break;
      case 6: MCH_BLK(6); break;
      case 7: MCH_BLK(7); break;
      default: MCH_BLK(8); break;
    }
#undef MCH_BLK
    return mch::hip_check(hipGetLastError(), "sofft_slice_blk_kernel launch");
  }
  const size_t lds = (size_t)(W * W + W * n + n) * sizeof(float2);
  const int per = (n * n + 255) / 256;
  if (per <= 16) {
    hipLaunchKernelGGL(sofft_slice_kernel<16>, dim3(n, B), dim3(256), lds, stream,
                       (const float2*)sigma, ld_sigma, p->d_grid, p->twiddle, p->L0, n, p->nsz,
                       C, p->part_val, p->part_idx, grid_out);
  } else {
    hipLaunchKernelGGL(sofft_slice_kernel<64>, dim3(n, B), dim3(256), lds, stream,
                       (const float2*)sigma, ld_sigma, p->d_grid, p->twiddle, p->L0, n, p->nsz,
                       C, p->part_val, p->part_idx, grid_out);
  }
  return mch::hip_check(hipGetLastError(), "sofft_slice_kernel launch");
}

int mch_sofft_topk(void* plan, const void* sigma, long long ld_sigma, int B, int C,
                   float* alpha, float* beta, float* gamma, float* score, hipStream_t stream) {
  So3Plan* p = (So3Plan*)plan;
  if (!p || !sigma || !alpha || !beta || !gamma || !score) {
    mch::set_error("mch_sofft_topk: NULL argument");
    return MCH_ERR_ARG;
  }
  if (B < 1 || B > p->max_batch || C < 1 || C > p->max_cand || ld_sigma < p->nsz) {
    mch::set_error("mch_sofft_topk: bad sizes (B=%d C=%d ld=%lld; plan batch=%d C=%d nsz=%d)", B, C,
                   ld_sigma, p->max_batch, p->max_cand, p->nsz);
    return MCH_ERR_SIZE;
  }
  int rc = launch_sofft_slice(p, sigma, ld_sigma, B, C, nullptr, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(sofft_merge_kernel, dim3(B), dim3(256), (size_t)p->n * C, stream,
                     p->part_val, p->part_idx, p->n, C, alpha, beta, gamma, score, C);
  return mch::hip_check(hipGetLastError(), "sofft_merge_kernel launch");
}

// Full score grid f[b, beta, alpha, gamma] (SOFFT.eval), float32 (B, n, n, n).
int mch_sofft_eval(void* plan, const void* sigma, long long ld_sigma, int B, float* grid,
                   hipStream_t stream) {
  So3Plan* p = (So3Plan*)plan;
  if (!p || !sigma || !grid) { mch::set_error("mch_sofft_eval: NULL argument"); return MCH_ERR_ARG; }
  if (B < 1 || B > p->max_batch || ld_sigma < p->nsz) {
    mch::set_error("mch_sofft_eval: bad sizes (B=%d ld=%lld)", B, ld_sigma);
    return MCH_ERR_SIZE;
  }
  return launch_sofft_slice(p, sigma, ld_sigma, B, 0, grid, stream);
}

// Newton march over the level schedule for (B, C) candidates (in/out);
// lend (B device ints or NULL): per-particle march ends.
static int newton_march(So3Plan* p, const void* sigma, long long ld_sigma, int B, int C,
                        const int* levels, int n_levels, const int* lend, int num_steps, int newton,
                        float base_lr, float* alpha, float* beta, float* gamma, float* score,
                        hipStream_t stream) {
  if (!p || !sigma || !levels || !alpha || !beta || !gamma || !score) {
    mch::set_error("mch_newton_march: NULL argument");
    return MCH_ERR_ARG;
  }
  if (lend && p->stop_early) {
    mch::set_error("mch_newton_march: per-particle march ends cannot be combined with stop_early "
                   "(its exit is batch-global per level)");
    return MCH_ERR_UNSUPPORTED;
  }
  for (int i = 1; lend && i < n_levels; ++i) {
    if (levels[i] <= levels[i - 1]) {
      mch::set_error("mch_newton_march: levels must increase with per-particle march ends");
      return MCH_ERR_ARG;
    }
  }
  if (B < 1 || B > p->max_batch || C < 1 || n_levels < 1 || n_levels > 16 || num_steps < 0 ||
      ld_sigma < mch::dsize(p->Lmax)) {
    mch::set_error("mch_newton_march: bad sizes (B=%d C=%d levels=%d steps=%d ld=%lld)", B, C,
                   n_levels, num_steps, ld_sigma);
    return MCH_ERR_SIZE;
  }
  MarchArgs a;
  memset(&a, 0, sizeof(a));
  for (int i = 0; i < n_levels; ++i) {
    if (levels[i] < 0 || levels[i] > p->Lmax) {
      mch::set_error("mch_newton_march: level %d outside [0, %d]", levels[i], p->Lmax);
      return MCH_ERR_ARG;
    }
    a.levels[i] = levels[i];
  }
  const int per = mch::pk_size(p->Lmax) / 4;
  const long long tot = (long long)per * B;
  int nb = (int)((tot + 255) / 256);
  if (nb > 65535 * 8) nb = 65535 * 8;
  hipLaunchKernelGGL(so3_repack_kernel, dim3(nb), dim3(256), 0, stream, (const float2*)sigma,
                     ld_sigma, p->sig_pk, p->Lmax, B);
  int rc = mch::hip_check(hipGetLastError(), "so3_repack_kernel launch");
  if (rc) return rc;
  a.sig_pk = p->sig_pk;
  a.pk_stride = mch::pk_size(p->Lmax);
  a.B = B; a.C = C;
  a.alpha = alpha; a.beta = beta; a.gamma = gamma; a.score = score;
  a.n_levels = n_levels;
  a.num_steps = num_steps;
  a.newton = newton;
  a.base_lr = base_lr;
  a.Lmax = p->Lmax;
  a.lend = lend;
  if (!p->stop_early) {
    hipLaunchKernelGGL(so3_march_kernel<false>, dim3(B * C), dim3(kMarchThreads), 0, stream, a, p->lgf);
    return mch::hip_check(hipGetLastError(), "so3_march_kernel launch");
  }
  // batch-global early exit: one launch per (level, step) after zeroing the
  // accept counts (a fill kernel, so that the sequence can be graph-captured)
  if (num_steps > mch::kMaxStopSteps) {
    mch::set_error("mch_newton_march: stop_early supports at most %d steps per level", mch::kMaxStopSteps);
    return MCH_ERR_SIZE;
  }
  const int ncount = n_levels * (num_steps + 1);
  hipLaunchKernelGGL(so3_zero_kernel, dim3(1), dim3(256), 0, stream, p->accept, ncount);
  if ((rc = mch::hip_check(hipGetLastError(), "so3_zero_kernel launch"))) return rc;
  a.state = p->state;
  a.accept = p->accept;
  for (int li = 0; li < n_levels; ++li)
    for (int st = 0; st <= num_steps; ++st) {
      a.li0 = li;
      a.st0 = st;
      hipLaunchKernelGGL(so3_march_kernel<true>, dim3(B * C), dim3(kMarchThreads), 0, stream, a, p->lgf);
      if ((rc = mch::hip_check(hipGetLastError(), "so3_march_kernel launch"))) return rc;
    }
  return MCH_OK;
}

int mch_newton_march(void* plan, const void* sigma, long long ld_sigma, int B, int C,
                     const int* levels, int n_levels, int num_steps, int newton, float base_lr,
                     float* alpha, float* beta, float* gamma, float* score, hipStream_t stream) {
  return newton_march((So3Plan*)plan, sigma, ld_sigma, B, C, levels, n_levels, nullptr, num_steps,
                      newton, base_lr, alpha, beta, gamma, score, stream);
}

int mch_so3_set_stop_early(void* plan, int stop_early) {
  So3Plan* p = (So3Plan*)plan;
  if (!p) { mch::set_error("mch_so3_set_stop_early: NULL plan"); return MCH_ERR_ARG; }
  p->stop_early = stop_early ? 1 : 0;
  return MCH_OK;
}

int mch_so3_evaluate(void* plan, const void* sigma, long long ld_sigma, int B, int C, int L,
                     const float* alpha, const float* beta, const float* gamma, float* out,
                     hipStream_t stream) {
  So3Plan* p = (So3Plan*)plan;
  if (!p || !sigma || !alpha || !beta || !gamma || !out) {
    mch::set_error("mch_so3_evaluate: NULL argument");
    return MCH_ERR_ARG;
  }
  if (B < 1 || B > p->max_batch || C < 1 || L < 0 || L > p->Lmax || ld_sigma < mch::dsize(p->Lmax)) {
    mch::set_error("mch_so3_evaluate: bad sizes (B=%d C=%d L=%d)", B, C, L);
    return MCH_ERR_SIZE;
  }
  const int per = mch::pk_size(p->Lmax) / 4;
  const long long tot = (long long)per * B;
  int nb = (int)((tot + 255) / 256);
  if (nb > 65535 * 8) nb = 65535 * 8;
  hipLaunchKernelGGL(so3_repack_kernel, dim3(nb), dim3(256), 0, stream, (const float2*)sigma,
                     ld_sigma, p->sig_pk, p->Lmax, B);
  int rc = mch::hip_check(hipGetLastError(), "so3_repack_kernel launch");
  if (rc) return rc;
  hipLaunchKernelGGL(so3_eval_kernel, dim3(B * C), dim3(kMarchThreads), 0, stream, p->sig_pk,
                     mch::pk_size(p->Lmax), C, L, alpha, beta, gamma, p->lgf, out);
  return mch::hip_check(hipGetLastError(), "so3_eval_kernel launch");
}

int mch_reduce_candidates(const float* alpha, const float* beta, const float* gamma,
                          const float* score, int B, int C, float* out_alpha, float* out_beta,
                          float* out_gamma, float* out_score, hipStream_t stream) {
  if (B < 1 || C < 1) { mch::set_error("mch_reduce_candidates: bad sizes"); return MCH_ERR_SIZE; }
  hipLaunchKernelGGL(reduce_candidates_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, alpha,
                     beta, gamma, score, B, C, out_alpha, out_beta, out_gamma, out_score);
  return mch::hip_check(hipGetLastError(), "reduce_candidates_kernel launch");
}

// Full Matcha.search_orientations: grid top-C -> march -> best candidate.
static int search_orientations(void* plan, const void* sigma, long long ld_sigma, int B, int C,
                               const int* levels, int n_levels, const int* lend, int num_steps,
                               int newton, float base_lr, float* out_alpha, float* out_beta,
                               float* out_gamma, float* out_score, hipStream_t stream) {
  So3Plan* p = (So3Plan*)plan;
  if (!p) { mch::set_error("NULL plan"); return MCH_ERR_ARG; }
  if (B > p->max_batch || C > p->max_cand) {
    mch::set_error("mch_search_orientations: B=%d C=%d exceed plan (%d, %d)", B, C, p->max_batch,
                   p->max_cand);
    return MCH_ERR_SIZE;
  }
  float* ca = p->cand;
  float* cb = ca + (size_t)p->max_batch * p->max_cand;
  float* cg = cb + (size_t)p->max_batch * p->max_cand;
  float* cs = cg + (size_t)p->max_batch * p->max_cand;
  int rc = mch_sofft_topk(plan, sigma, ld_sigma, B, C, ca, cb, cg, cs, stream);
  if (rc) return rc;
  rc = newton_march(p, sigma, ld_sigma, B, C, levels, n_levels, lend, num_steps, newton, base_lr, ca,
                    cb, cg, cs, stream);
  if (rc) return rc;
  return mch_reduce_candidates(ca, cb, cg, cs, B, C, out_alpha, out_beta, out_gamma, out_score,
                               stream);
}

int mch_search_orientations(void* plan, const void* sigma, long long ld_sigma, int B, int C,
                            const int* levels, int n_levels, int num_steps, int newton,
                            float base_lr, float* out_alpha, float* out_beta, float* out_gamma,
                            float* out_score, hipStream_t stream) {
  return search_orientations(plan, sigma, ld_sigma, B, C, levels, n_levels, nullptr, num_steps,
                             newton, base_lr, out_alpha, out_beta, out_gamma, out_score, stream);
}

int mch_search_orientations_ends(void* plan, const void* sigma, long long ld_sigma, int B, int C,
                                 const int* levels, int n_levels, const int* march_ends,
                                 int num_steps, int newton, float base_lr, float* out_alpha,
                                 float* out_beta, float* out_gamma, float* out_score,
                                 hipStream_t stream) {
  if (!march_ends) {
    mch::set_error("mch_search_orientations_ends: NULL march_ends");
    return MCH_ERR_ARG;
  }
  return search_orientations(plan, sigma, ld_sigma, B, C, levels, n_levels, march_ends, num_steps,
                             newton, base_lr, out_alpha, out_beta, out_gamma, out_score, stream);
}

// Host FP64 Wigner d table (textbook convention, D layout) -- exported so the
// host side and tests can use the library's own tables.
int mch_wigner_d_host(double beta, int L, double* out) {
  if (!out || L < 0 || L > kMaxL) { mch::set_error("mch_wigner_d_host: bad args"); return MCH_ERR_ARG; }
  mch::host_wigner_d(beta, L, out);
  return MCH_OK;
}

}  // extern "C"
