// Ball-harmonic projection a = B^H f and correlation coefficients sigma on gfx950.
//
// Reference behaviour (swing-research/Matcha):
//   FLEBasis3D.evaluate_t      src/matcha/core/FLEBasis3D.py:945-976,1026-1267
//   compute_sigma / get_sigma  src/matcha/core/CrossCorrelationMatcher.py:183-249
//
// Pipeline per batch (three launches):
//   1. fle_symmetrize_kernel: each fundamental voxel's 16 orbit values ->
//      4 combinations per class (8 classes), written as the GEMM A operand
//      X[class][row][k] with rows (b: real part, B+b: imaginary part) and
//      k = 2*v0 + {0: cos-weighted, 1: sin-weighted}.
//   2. fle_gemm_kernel: grouped FP32 MFMA GEMM (v_mfma_f32_16x16x4f32),
//      Out[row][col] = sum_k X[class(col)][row][k] W[k][col], split over K.
//   3. fle_finalize_kernel: sum the K splits, a_{lkm} = Re + i Im and
//      a_{lk,-m} = (-1)^m conj(a_{lkm}), scattered to the reference order.
#include "common.hpp"

#include <vector>

namespace {

constexpr int kClasses = 8;
constexpr int BM = 128, BN = 64, BK = 16;
constexpr int kGemmThreads = 256;

struct FlePlan {
  int N, V0, Kpad, Ntot, ne, max_batch;
  int col_off[kClasses + 1];  // padded column offsets per class
  int* orbit;                 // (V0, 16)
  float* W;                   // (Kpad, Ntot)
  int* chan_plus;             // (Ntot)
  int* chan_minus;            // (Ntot)
  int* chan_m;                // (Ntot)
  float* X;                   // (kClasses, Mpad, Kpad)
  float* part;                // (splits, Mpad, Ntot)
  int Mpad, splits;
};

struct CorrPlan {
  int lmax, kdim, max_batch;
  int* idx;   // (lmax+1, 2*lmax+1, kdim): coefficient index of (l, m, k) or -1
};

typedef float floatx4 __attribute__((ext_vector_type(4)));

}  // namespace

// ---------------------------------------------------------------------------
// 1. orbit symmetrisation
// ---------------------------------------------------------------------------
// For class (eps, rho): G_t = sum_q (-1)^{q eps} sum_p (-i)^{rho p} f[p,q,t],
// X1 = A0+A1, X2 = B0-B1, X3 = B0+B1, X4 = A1-A0 with G_t = A_t + i B_t.
__global__ void fle_symmetrize_kernel(const float* __restrict__ vols, long long ld_vol, int B,
                                      const int* __restrict__ orbit, int V0, int Kpad, int Mpad,
                                      float* __restrict__ X) {
  const long long tot = (long long)B * V0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(t / V0);
    const int v = (int)(t - (long long)b * V0);
    const float* f = vols + (long long)b * ld_vol;
    const int4* o4 = reinterpret_cast<const int4*>(orbit + (long long)v * 16);
    float g[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int4 id = o4[i];
      g[4 * i + 0] = f[id.x];
      g[4 * i + 1] = f[id.y];
      g[4 * i + 2] = f[id.z];
      g[4 * i + 3] = f[id.w];
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
      for (int rho = 0; rho < 4; ++rho) {
        const int cls = eps * 4 + rho;
        const float X1 = A[0][rho] + A[1][rho];
        const float X2 = Bi[0][rho] - Bi[1][rho];
        const float X3 = Bi[0][rho] + Bi[1][rho];
        const float X4 = A[1][rho] - A[0][rho];
        float* xr = X + ((long long)cls * Mpad + b) * Kpad + 2 * v;
        float* xi = X + ((long long)cls * Mpad + B + b) * Kpad + 2 * v;
        *reinterpret_cast<float2*>(xr) = make_float2(X1, X2);
        *reinterpret_cast<float2*>(xi) = make_float2(X3, X4);
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
    int Mpad, int Kpad, int Ntot, int kchunk, FlePlan plan_cols) {
  __shared__ float As[2][BK][BM + 4];
  __shared__ float Bs[2][BK][BN + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int col0 = blockIdx.x * BN;
  const int row0 = blockIdx.y * BM;
  int cls = 0;
  while (cls < kClasses - 1 && col0 >= plan_cols.col_off[cls + 1]) ++cls;
  const float* A = X + (long long)cls * Mpad * Kpad;
  const int kbeg = blockIdx.z * kchunk;
  int kend = kbeg + kchunk;
  if (kend > Kpad) kend = Kpad;

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
// 3. split-K reduction and scatter to the reference coefficient order
// ---------------------------------------------------------------------------
__global__ void fle_finalize_kernel(const float* __restrict__ part, int splits, int Mpad, int Ntot,
                                    int B, const int* __restrict__ chan_plus,
                                    const int* __restrict__ chan_minus,
                                    const int* __restrict__ chan_m, float2* __restrict__ a,
                                    long long lda) {
  const long long tot = (long long)B * Ntot;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot;
       t += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(t / Ntot);
    const int c = (int)(t - (long long)b * Ntot);
    const int ip = chan_plus[c];
    if (ip < 0) continue;
    float re = 0.f, im = 0.f;
    for (int s = 0; s < splits; ++s) {
      const float* P = part + (long long)s * Mpad * Ntot;
      re += P[(long long)b * Ntot + c];
      im += P[(long long)(B + b) * Ntot + c];
    }
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
__global__ __launch_bounds__(256) void sigma_kernel(const float2* __restrict__ aT,
                                                    const float2* __restrict__ aV, long long ldv,
                                                    const int* __restrict__ idx, int lmax,
                                                    int kdim, float2* __restrict__ sigma,
                                                    long long lds) {
  const int l = blockIdx.x;
  const int b = blockIdx.y;
  const int w = 2 * l + 1;
  const float2* av = aV + (long long)b * ldv;
  const int* il = idx + (long long)l * (2 * lmax + 1) * kdim;
  float2* out = sigma + (long long)b * lds + mch::dsize(l - 1);
  for (int q = threadIdx.x; q < w * w; q += blockDim.x) {
    const int u = q / w - l, v = q % w - l;
    const int* iv = il + (long long)(-u + l) * kdim;   // particle, m = -u
    const int* it = il + (long long)(-v + l) * kdim;   // template, m = -v
    float re = 0.f, im = 0.f;
    for (int k = 0; k < kdim; ++k) {
      const int a = it[k], c = iv[k];
      if (a < 0 || c < 0) break;
      const float2 x = aT[a], y = av[c];
      // x * conj(y)
      re += x.x * y.x + x.y * y.y;
      im += x.y * y.x - x.x * y.y;
    }
    const float sg = ((u + v) & 1) ? -1.f : 1.f;
    out[q] = make_float2(sg * re, sg * im);
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
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int mch_fle_plan_create(int N, int V0, int Kpad, int Ntot, const int* class_colpad,
                        const int* orbit, const float* basis, const int* chan_plus,
                        const int* chan_minus, const int* chan_m, int ne, int max_batch,
                        void** out_plan) {
  if (!out_plan || !class_colpad || !orbit || !basis || !chan_plus || !chan_minus || !chan_m) {
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
  if (N < 2 || V0 < 1 || Kpad < 2 * V0 || Kpad % BK || sum != Ntot || ne < 1 || max_batch < 1) {
    mch::set_error("mch_fle_plan_create: inconsistent sizes (N=%d V0=%d Kpad=%d Ntot=%d ne=%d)", N,
                   V0, Kpad, Ntot, ne);
    return MCH_ERR_SIZE;
  }
  for (long long i = 0; i < (long long)V0 * 16; ++i) {
    if (orbit[i] < 0 || orbit[i] >= N * N * N) {
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
  MCH_TRY(hipMalloc(&p->W, (size_t)Kpad * Ntot * sizeof(float)), "hipMalloc(W)");
  MCH_TRY(hipMemcpy(p->W, basis, (size_t)Kpad * Ntot * sizeof(float), hipMemcpyHostToDevice), "copy W");
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
  (void)hipFree(p->orbit); (void)hipFree(p->W); (void)hipFree(p->chan_plus);
  (void)hipFree(p->chan_minus); (void)hipFree(p->chan_m); (void)hipFree(p->X); (void)hipFree(p->part);
  delete p;
  return rc;
}

int mch_fle_plan_destroy(void* plan) {
  FlePlan* p = (FlePlan*)plan;
  if (!p) return MCH_OK;
  (void)hipFree(p->orbit); (void)hipFree(p->W); (void)hipFree(p->chan_plus);
  (void)hipFree(p->chan_minus); (void)hipFree(p->chan_m); (void)hipFree(p->X); (void)hipFree(p->part);
  delete p;
  return MCH_OK;
}

// a (B, ne) complex64 rows of stride lda = B^H f for volumes (B, N^3) float32.
int mch_fle_project(void* plan, const float* vols, long long ld_vol, int B, void* coeffs,
                    long long lda, hipStream_t stream) {
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
  hipLaunchKernelGGL(fle_symmetrize_kernel, dim3(nb), dim3(256), 0, stream, vols, ld_vol, B,
                     p->orbit, p->V0, p->Kpad, p->Mpad, p->X);
  int rc = mch::hip_check(hipGetLastError(), "fle_symmetrize_kernel launch");
  if (rc) return rc;
  const int mt = (2 * B + BM - 1) / BM;
  const int kblocks = p->Kpad / BK;
  const int kchunk = ((kblocks + p->splits - 1) / p->splits) * BK;
  const int splits = (p->Kpad + kchunk - 1) / kchunk;
  hipLaunchKernelGGL(fle_gemm_kernel, dim3(p->Ntot / BN, mt, splits), dim3(kGemmThreads), 0,
                     stream, p->X, p->W, p->part, p->Mpad, p->Kpad, p->Ntot, kchunk, *p);
  rc = mch::hip_check(hipGetLastError(), "fle_gemm_kernel launch");
  if (rc) return rc;
  const long long ftot = (long long)B * p->Ntot;
  int fb = (int)((ftot + 255) / 256);
  if (fb > 8192) fb = 8192;
  hipLaunchKernelGGL(fle_finalize_kernel, dim3(fb), dim3(256), 0, stream, p->part, splits,
                     p->Mpad, p->Ntot, B, p->chan_plus, p->chan_minus, p->chan_m,
                     (float2*)coeffs, lda);
  return mch::hip_check(hipGetLastError(), "fle_finalize_kernel launch");
}

// Re-run only the projection GEMM on the operand left by the last
// mch_fle_project call (for timing the dominant kernel in isolation).
int mch_fle_gemm_only(void* plan, int B, hipStream_t stream) {
  FlePlan* p = (FlePlan*)plan;
  if (!p || B < 1 || B > p->max_batch) { mch::set_error("mch_fle_gemm_only: bad arguments"); return MCH_ERR_ARG; }
  const int mt = (2 * B + BM - 1) / BM;
  const int kblocks = p->Kpad / BK;
  const int kchunk = ((kblocks + p->splits - 1) / p->splits) * BK;
  const int splits = (p->Kpad + kchunk - 1) / kchunk;
  hipLaunchKernelGGL(fle_gemm_kernel, dim3(p->Ntot / BN, mt, splits), dim3(kGemmThreads), 0,
                     stream, p->X, p->W, p->part, p->Mpad, p->Kpad, p->Ntot, kchunk, *p);
  return mch::hip_check(hipGetLastError(), "fle_gemm_kernel launch");
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

// sigma (B, dsize(lmax)) with row stride lds, and ||aV|| (B,) when norms != NULL.
int mch_compute_sigma(void* plan, const void* aT, const void* aV, long long ldv, int ne, int B,
                      void* sigma, long long lds, float* norms, hipStream_t stream) {
  CorrPlan* p = (CorrPlan*)plan;
  if (!p || !aT || !aV || !sigma) { mch::set_error("mch_compute_sigma: NULL argument"); return MCH_ERR_ARG; }
  if (B < 1 || lds < mch::dsize(p->lmax) || ldv < ne) {
    mch::set_error("mch_compute_sigma: bad sizes (B=%d lds=%lld)", B, lds);
    return MCH_ERR_SIZE;
  }
  hipLaunchKernelGGL(sigma_kernel, dim3(p->lmax + 1, B), dim3(256), 0, stream, (const float2*)aT,
                     (const float2*)aV, ldv, p->idx, p->lmax, p->kdim, (float2*)sigma, lds);
  int rc = mch::hip_check(hipGetLastError(), "sigma_kernel launch");
  if (rc || !norms) return rc;
  hipLaunchKernelGGL(norms_kernel, dim3(B), dim3(256), 0, stream, (const float2*)aV, ldv, ne, norms);
  return mch::hip_check(hipGetLastError(), "norms_kernel launch");
}

}  // extern "C"
