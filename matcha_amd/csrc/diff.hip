// Differentiable correlation score for gfx950: the reference's score
//   s(omega, t, aT) = Re sum_{l<=L} D^l(R(omega)) sigma^l(aT, aV(t))
// and its gradient with respect to the axis-angle rotation omega, the shift t
// and the template coefficients aT.
//
// Reference: the score of Matcha.optimize (src/matcha/core/Matcha.py:184,288)
// with sigma from CrossCorrelationMatcher.compute_sigma
// (src/matcha/core/CrossCorrelationMatcher.py:183-249) of the particle patch
// at the shift (extract_subtomogram_patch_batch, src/matcha/utils/
// io_utils.py:113-178); the reference differentiates it in Euler angles only
// (src/matcha/core/WignerDGradients.py:247-302).
//
// Kernels:
//   pose_kernel            omega -> z-y-z Euler angles of exp([omega]x) (FP64)
//                          and the right Jacobian J_r of exp.
//   shift_tangent_kernel   the trilinear patch V_t and its three exact
//                          derivatives d V_t / d t_i (piecewise-trilinear).
//   score_shell_kernel     one workgroup per (shell l, particle): d^l(beta)
//                          re-marched in FP64 for that shell only (the D stack
//                          is never stored) into LDS, then in one pass the
//                          shell's share of the score, of the body-frame
//                          rotation derivatives g_i = Re sum D(R) sigma X_i^T
//                          (X_i the generators of D on the m index), of the
//                          shift derivatives (sigma of the derivative volumes'
//                          coefficients), and the per-particle template adjoint
//                          G_{l k,-v} = sum_u conj(D_uv) (-1)^{u+v} aV_{l k,-u}.
//   score_reduce_kernel    sums the shells (FP64): s, J_r^T g, d s / d t.
//   score_backward_kernel  scales by the upstream gradient; the template
//                          gradient is the weighted sum of G over particles.
#include "common.hpp"

#include <math.h>

#include <mutex>

namespace {

constexpr int kThreads = 256;
constexpr int kParts = 8;  // per-shell partials: s, g_x, g_y, g_z, t_0, t_1, t_2, (pad)

__global__ void pose_kernel(const float* __restrict__ omega, int B, double* __restrict__ angles,
                            double* __restrict__ jr) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double w0 = omega[b * 3 + 0], w1 = omega[b * 3 + 1], w2 = omega[b * 3 + 2];
  const double th2 = w0 * w0 + w1 * w1 + w2 * w2, th = sqrt(th2);
  // exp([w]x) = I + a K + c K^2 (Rodrigues); J_r = I - c K + e K^2
  double a, c, e;
  if (th < 1e-6) {
    a = 1.0 - th2 / 6.0;
    c = 0.5 - th2 / 24.0;
    e = 1.0 / 6.0 - th2 / 120.0;
  } else {
    a = sin(th) / th;
    c = (1.0 - cos(th)) / th2;
    e = (th - sin(th)) / (th2 * th);
  }
  const double K[9] = {0.0, -w2, w1, w2, 0.0, -w0, -w1, w0, 0.0};
  double K2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) K2[i * 3 + j] = K[i * 3] * K[j] + K[i * 3 + 1] * K[3 + j] + K[i * 3 + 2] * K[6 + j];
  double R[9];
  for (int i = 0; i < 9; ++i) {
    const double I = (i % 4 == 0) ? 1.0 : 0.0;
    R[i] = I + a * K[i] + c * K2[i];
    jr[b * 9 + i] = I - c * K[i] + e * K2[i];
  }
  // R = Rz(alpha) Ry(beta) Rz(gamma)
  const double sb = sqrt(R[2] * R[2] + R[5] * R[5]);
  const double beta = atan2(sb, R[8]);
  double alpha, gamma;
  if (sb > 1e-12) {
    alpha = atan2(R[5], R[2]);
    gamma = atan2(R[7], -R[6]);
  } else if (R[8] > 0.0) {  // beta = 0: R = Rz(alpha + gamma)
    alpha = atan2(R[3], R[0]);
    gamma = 0.0;
  } else {  // beta = pi: upper block Rz(alpha) diag(-1, 1) with gamma = 0
    alpha = atan2(-R[3], -R[0]);
    gamma = 0.0;
  }
  angles[b * 3 + 0] = alpha;
  angles[b * 3 + 1] = beta;
  angles[b * 3 + 2] = gamma;
}

// V_t[a] = V at (x + t0, y + t1, z + t2) + Nv/2 - N/2 (trilinear, zeros
// outside), a = (z, y, x); out rows b, B + b, 2B + b, 3B + b hold V_t and its
// derivatives in t0, t1, t2 (inside each cell the interpolant is trilinear
// in t, so these are exact; on a cell face they are the one-sided derivative
// of the cell above).
__global__ __launch_bounds__(kThreads) void shift_tangent_kernel(const float* __restrict__ vols, long long ld_vol,
                                                                 int Nv, int N, const float* __restrict__ t, int B,
                                                                 float* __restrict__ out, long long ld_out) {
  const int z = blockIdx.x, b = blockIdx.y;
  const float* v = vols + (long long)b * ld_vol;
  const float base = (float)(Nv / 2 - N / 2);
  const float t0 = t[b * 3 + 0], t1 = t[b * 3 + 1], t2 = t[b * 3 + 2];
  const float pz = (float)z + base + t2;
  const float fz0 = floorf(pz);
  const int z0 = (int)fz0;
  const float fz = pz - fz0;
  float* o0 = out + (long long)b * ld_out + (long long)z * N * N;
  float* o1 = o0 + (long long)B * ld_out;
  float* o2 = o1 + (long long)B * ld_out;
  float* o3 = o2 + (long long)B * ld_out;
  for (int q = threadIdx.x; q < N * N; q += kThreads) {
    const int y = q / N, x = q - y * N;
    const float px = (float)x + base + t0, py = (float)y + base + t1;
    const float fx0 = floorf(px), fy0 = floorf(py);
    const int x0 = (int)fx0, y0 = (int)fy0;
    const float fx = px - fx0, fy = py - fy0;
    float c[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int xi = x0 + (k & 1), yi = y0 + ((k >> 1) & 1), zi = z0 + (k >> 2);
      const bool ok = (unsigned)xi < (unsigned)Nv && (unsigned)yi < (unsigned)Nv && (unsigned)zi < (unsigned)Nv;
      c[k] = ok ? v[((long long)zi * Nv + yi) * Nv + xi] : 0.f;
    }
    const float gx = 1.f - fx, gy = 1.f - fy, gz = 1.f - fz;
    // c[k]: k bit 0 = x, bit 1 = y, bit 2 = z
    const float val = gz * (gy * (gx * c[0] + fx * c[1]) + fy * (gx * c[2] + fx * c[3])) +
                      fz * (gy * (gx * c[4] + fx * c[5]) + fy * (gx * c[6] + fx * c[7]));
    const float dx = gz * (gy * (c[1] - c[0]) + fy * (c[3] - c[2])) + fz * (gy * (c[5] - c[4]) + fy * (c[7] - c[6]));
    const float dy = gz * (gx * (c[2] - c[0]) + fx * (c[3] - c[1])) + fz * (gx * (c[6] - c[4]) + fx * (c[7] - c[5]));
    const float dz = gy * (gx * (c[4] - c[0]) + fx * (c[5] - c[1])) + fy * (gx * (c[6] - c[2]) + fx * (c[7] - c[3]));
    o0[q] = val;
    o1[q] = dx;
    o2[q] = dy;
    o3[q] = dz;
  }
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Dynamic LDS of score_shell_kernel for shells up to L and radial size kdim.
// The d^l block is padded to an even float count so the float2 tables after
// it stay 8-byte aligned.
__host__ __device__ inline int dl_floats(int w) { return (w * w + 1) & ~1; }
__host__ __device__ inline size_t shell_lds_bytes(int L, int kdim) {
  const int w = 2 * L + 1;
  return (size_t)dl_floats(w) * sizeof(float) + (size_t)2 * w * sizeof(float2) + (size_t)w * kdim * sizeof(float2);
}

__global__ __launch_bounds__(kThreads) void score_shell_kernel(
    const float2* __restrict__ sigma, long long ld_sigma, int B, const float2* __restrict__ aV, long long ld_aV,
    const int* __restrict__ idx, int lmax, int kdim, int L, const double* __restrict__ angles,
    float* __restrict__ partial, float2* __restrict__ G, long long ldG) {
  const int l = blockIdx.x, b = blockIdx.y;
  const int w = 2 * l + 1;
  const int* il = idx + (long long)l * (2 * lmax + 1) * kdim;  // row (m + l), column k
  int K = 0;
  while (K < kdim && il[(long long)l * kdim + K] >= 0) ++K;  // radial size of shell l (row m = 0)
  if (l > L) {
    // the score does not involve shells above L: their template gradient is 0
    if (G)
      for (int q = threadIdx.x; q < w * K; q += kThreads) {
        const int j = il[(long long)(q / K) * kdim + (q % K)];
        if (j >= 0) G[(long long)b * ldG + j] = make_float2(0.f, 0.f);
      }
    return;
  }
  extern __shared__ float smem[];
  float* dl = smem;                                     // w x w, d^l_{uv}(beta) at [(u + l) w + (v + l)]
  float2* za = reinterpret_cast<float2*>(dl + dl_floats(w));  // e^{i u alpha}, u = -l..l
  float2* zg = za + w;                                  // e^{i v gamma}
  float2* pk = zg + w;                                  // (w, K): (-1)^u e^{-i u alpha} aV_{l k,-u}
  const double alpha = angles[b * 3 + 0], beta = angles[b * 3 + 1], gamma = angles[b * 3 + 2];
  for (int i = threadIdx.x; i < w; i += kThreads) {
    const int u = i - l;
    double s, c;
    sincos((double)u * alpha, &s, &c);
    za[i] = make_float2((float)c, (float)s);
    sincos((double)u * gamma, &s, &c);
    zg[i] = make_float2((float)c, (float)s);
  }
  // d^l(beta) from the fundamental pairs mp >= |m|: seed at l0 = mp (closed
  // form), three-term l-recurrence to l in FP64, four symmetry images
  const double hc = cos(0.5 * beta), hs = sin(0.5 * beta), x = cos(beta);
  const int npairs = (l + 1) * (l + 1);
  for (int p = threadIdx.x; p < npairs; p += kThreads) {
    int mp = (int)sqrt((double)p);
    if (mp * mp > p) --mp;
    if ((mp + 1) * (mp + 1) <= p) ++mp;
    const int m = p - mp * mp - mp;
    const int a = mp + m, bb = mp - m;
    double d = exp(0.5 * (lgamma(2.0 * mp + 1.0) - lgamma(a + 1.0) - lgamma(2.0 * mp - a + 1.0)));
    if (a) d *= pow(hc, (double)a);
    if (bb) d *= pow(hs, (double)bb);
    if (bb & 1) d = -d;
    double pd = 0.0, Cprev = 0.0;
    for (int k = mp + 1; k <= l; ++k) {
      const double P = (double)(k * k - m * m) * (double)(k * k - mp * mp);
      const double sP = sqrt(P);
      const double A = (double)k * (2 * k - 1) / sP;
      const double Bc = (k > 1) ? (double)(m * mp) / ((double)(k - 1) * k) : 0.0;
      const double nd = A * ((x - Bc) * d - Cprev * pd);
      pd = d;
      d = nd;
      Cprev = sP / ((double)k * (2 * k + 1));
    }
    const float fd = (float)d, fs = (bb & 1) ? -fd : fd;
    dl[(mp + l) * w + (m + l)] = fd;    // (mp, m)
    dl[(m + l) * w + (mp + l)] = fs;    // (m, mp)
    dl[(-m + l) * w + (-mp + l)] = fd;  // (-m, -mp)
    dl[(-mp + l) * w + (-m + l)] = fs;  // (-mp, -m)
  }
  __syncthreads();

  // score, body-frame rotation derivatives and shift derivatives of this shell
  const long long off = mch::dsize(l - 1);
  const float2* s0 = sigma + (long long)b * ld_sigma + off;
  const float2* st0 = sigma + (long long)(B + b) * ld_sigma + off;
  const float2* st1 = sigma + (long long)(2 * B + b) * ld_sigma + off;
  const float2* st2 = sigma + (long long)(3 * B + b) * ld_sigma + off;
  float acc[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int q = threadIdx.x; q < w * w; q += kThreads) {
    const int ui = q / w, vi = q - ui * w;
    const int v = vi - l;
    const float dd = dl[q];
    if (dd == 0.f) continue;
    const float2 ph = cmul(za[ui], zg[vi]);
    const float2 D = make_float2(dd * ph.x, dd * ph.y);
    const float2 sg = s0[q];
    const float2 up = (v < l) ? s0[q + 1] : make_float2(0.f, 0.f);
    const float2 dn = (v > -l) ? s0[q - 1] : make_float2(0.f, 0.f);
    const float cu = 0.5f * sqrtf((float)((l - v) * (l + v + 1)));
    const float cd = 0.5f * sqrtf((float)((l - v + 1) * (l + v)));
    // Re(D z) and Re(D i z) = -Im(D z)
    auto re = [&](float2 z) { return D.x * z.x - D.y * z.y; };
    auto im = [&](float2 z) { return D.x * z.y + D.y * z.x; };
    const float2 ladd = make_float2(cu * up.x + cd * dn.x, cu * up.y + cd * dn.y);
    const float2 lsub = make_float2(cu * up.x - cd * dn.x, cu * up.y - cd * dn.y);
    acc[0] += re(sg);
    acc[1] += -im(ladd);
    acc[2] += re(lsub);
    acc[3] += -(float)v * im(sg);
    acc[4] += re(st0[q]);
    acc[5] += re(st1[q]);
    acc[6] += re(st2[q]);
  }
  __shared__ float red[7][kThreads / 64];
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    const float s = mch::wave_sum(acc[i]);
    if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    float s = 0.f;
#pragma unroll
    for (int wv = 0; wv < kThreads / 64; ++wv) s += red[threadIdx.x][wv];
    partial[((long long)b * (L + 1) + l) * kParts + threadIdx.x] = s;
  }
  if (!G) return;

  // template adjoint of this shell:
  // G_{l k,-v} = (-1)^v e^{-i v gamma} sum_u d_uv [(-1)^u e^{-i u alpha} aV_{l k,-u}]
  const float2* av = aV + (long long)b * ld_aV;
  for (int q = threadIdx.x; q < w * K; q += kThreads) {
    const int ui = q / K, k = q - ui * K;
    const int u = ui - l;
    const int j = il[(long long)(-u + l) * kdim + k];
    float2 z = j >= 0 ? av[j] : make_float2(0.f, 0.f);
    const float2 e = make_float2(za[ui].x, -za[ui].y);
    z = cmul(e, z);
    if (u & 1) z = make_float2(-z.x, -z.y);
    pk[q] = z;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < w * K; q += kThreads) {
    const int vi = q / K, k = q - vi * K;
    const int v = vi - l;
    float2 sum = make_float2(0.f, 0.f);
    for (int ui = 0; ui < w; ++ui) {
      const float dd = dl[ui * w + vi];
      const float2 z = pk[ui * K + k];
      sum.x += dd * z.x;
      sum.y += dd * z.y;
    }
    float2 r = cmul(make_float2(zg[vi].x, -zg[vi].y), sum);
    if (v & 1) r = make_float2(-r.x, -r.y);
    const int j = il[(long long)(-v + l) * kdim + k];
    if (j >= 0) G[(long long)b * ldG + j] = r;
  }
}

// out (B, 7): s, d s / d omega (J_r^T g), d s / d t
__global__ void score_reduce_kernel(const float* __restrict__ partial, int B, int L, const double* __restrict__ jr,
                                    float* __restrict__ out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int l = 0; l <= L; ++l)
    for (int i = 0; i < 7; ++i) s[i] += partial[((long long)b * (L + 1) + l) * kParts + i];
  const double* J = jr + b * 9;
  out[b * 7 + 0] = (float)s[0];
  for (int i = 0; i < 3; ++i) out[b * 7 + 1 + i] = (float)(J[0 * 3 + i] * s[1] + J[1 * 3 + i] * s[2] + J[2 * 3 + i] * s[3]);
  for (int i = 0; i < 3; ++i) out[b * 7 + 4 + i] = (float)s[4 + i];
}

// grad_omega = w_b d s/d omega, grad_t = w_b d s/d t, grad_aT = sum_b w_b G_b
__global__ void score_backward_kernel(const float* __restrict__ weights, const float* __restrict__ grads, int B,
                                      const float2* __restrict__ G, long long ldG, int ne,
                                      float* __restrict__ d_omega, float* __restrict__ d_t,
                                      float2* __restrict__ d_aT) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < B) {
    const float wb = weights[j];
    for (int i = 0; i < 3; ++i) {
      if (d_omega) d_omega[j * 3 + i] = wb * grads[j * 7 + 1 + i];
      if (d_t) d_t[j * 3 + i] = wb * grads[j * 7 + 4 + i];
    }
  }
  if (!G || j >= ne) return;
  double re = 0.0, im = 0.0;
  for (int b = 0; b < B; ++b) {
    const float2 g = G[(long long)b * ldG + j];
    re += (double)weights[b] * g.x;
    im += (double)weights[b] * g.y;
  }
  d_aT[j] = make_float2((float)re, (float)im);
}

// Static scratch of mch_score_gradients_noalloc: the per-shell partials
// (B (L + 1) 8 floats), at most kScratchFloats.  Each GPU holds its own copy
// for the life of the process; host code reaches it only through the device
// address from hipGetSymbolAddress, looked up once per device.
constexpr size_t kScratchFloats = size_t(1) << 20;  // 4 MiB
__device__ __align__(256) float scratch[kScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices];

int scratch_address(float** out) {
  int dev = -1;
  int rc = mch::hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kMaxDevices) {
    mch::set_error("device %d beyond the %d devices the static scratch is tracked for", dev, kMaxDevices);
    return MCH_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    rc = mch::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return MCH_OK;
}

}  // namespace

extern "C" {

int mch_score_pose(const float* omega, int B, double* angles, double* jr, hipStream_t stream) {
  if (!omega || !angles || !jr || B < 1) {
    mch::set_error("mch_score_pose: bad arguments");
    return MCH_ERR_ARG;
  }
  hipLaunchKernelGGL(pose_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, omega, B, angles, jr);
  return mch::hip_check(hipGetLastError(), "pose_kernel");
}

int mch_shift_tangent(const float* vols, long long ld_vol, int Nv, int B, int N, const float* t, float* out,
                      long long ld_out, hipStream_t stream) {
  if (!vols || !t || !out || B < 1 || N < 2 || Nv < N || Nv > 1024 || ld_vol < (long long)Nv * Nv * Nv ||
      ld_out < (long long)N * N * N) {
    mch::set_error("mch_shift_tangent: bad arguments");
    return MCH_ERR_ARG;
  }
  hipLaunchKernelGGL(shift_tangent_kernel, dim3(N, B), dim3(kThreads), 0, stream, vols, ld_vol, Nv, N, t, B, out,
                     ld_out);
  return mch::hip_check(hipGetLastError(), "shift_tangent_kernel");
}

int mch_score_gradients(const int* idx, int lmax, int kdim, const void* sigma, long long ld_sigma, int B,
                        const void* aV, long long ld_aV, int L, const double* angles, const double* jr,
                        float* partial, float* out, void* G, long long ldG, hipStream_t stream) {
  if (!idx || !sigma || !angles || !jr || !partial || !out || B < 1 || L < 0 || L > lmax || kdim < 1 ||
      ld_sigma < mch::dsize(lmax) || (G && !aV)) {
    mch::set_error("mch_score_gradients: bad arguments (B=%d L=%d lmax=%d)", B, L, lmax);
    return MCH_ERR_ARG;
  }
  const size_t lds = shell_lds_bytes(L, kdim);
  if (lds > 160 * 1024) {
    mch::set_error("mch_score_gradients: L=%d with %d radial orders needs %zu B of LDS (max 160 KiB)", L, kdim, lds);
    return MCH_ERR_UNSUPPORTED;
  }
  int rc = mch::hip_check(hipFuncSetAttribute((const void*)score_shell_kernel,
                                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                          "hipFuncSetAttribute");
  if (rc) return rc;
  hipLaunchKernelGGL(score_shell_kernel, dim3(G ? lmax + 1 : L + 1, B), dim3(kThreads), lds, stream,
                     (const float2*)sigma, ld_sigma, B, (const float2*)aV, ld_aV, idx, lmax, kdim, L, angles,
                     partial, (float2*)G, ldG);
  rc = mch::hip_check(hipGetLastError(), "score_shell_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(score_reduce_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, partial, B, L, jr, out);
  return mch::hip_check(hipGetLastError(), "score_reduce_kernel");
}

int mch_score_gradients_noalloc(const int* idx, int lmax, int kdim, const void* sigma, long long ld_sigma, int B,
                                const void* aV, long long ld_aV, int L, const double* angles, const double* jr,
                                float* out, void* G, long long ldG, hipStream_t stream) {
  if (B < 1 || L < 0) {
    mch::set_error("mch_score_gradients_noalloc: bad arguments (B=%d L=%d)", B, L);
    return MCH_ERR_ARG;
  }
  const size_t need = (size_t)B * (size_t)(L + 1) * kParts;
  if (need > kScratchFloats) {
    mch::set_error("mch_score_gradients_noalloc: B (L + 1) = %zu exceeds the static scratch (at most %zu)",
                   (size_t)B * (size_t)(L + 1), kScratchFloats / kParts);
    return MCH_ERR_SIZE;
  }
  float* partial = nullptr;
  const int rc = scratch_address(&partial);
  if (rc) return rc;
  return mch_score_gradients(idx, lmax, kdim, sigma, ld_sigma, B, aV, ld_aV, L, angles, jr, partial, out, G, ldG,
                             stream);
}

int mch_score_backward(const float* weights, const float* grads, int B, const void* G, long long ldG, int ne,
                       float* d_omega, float* d_t, void* d_aT, hipStream_t stream) {
  if (!weights || !grads || B < 1 || (G && (!d_aT || ne < 1 || ldG < ne))) {
    mch::set_error("mch_score_backward: bad arguments");
    return MCH_ERR_ARG;
  }
  const int n = G ? (ne > B ? ne : B) : B;
  hipLaunchKernelGGL(score_backward_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, weights, grads, B,
                     (const float2*)G, ldG, ne, d_omega, d_t, (float2*)d_aT);
  return mch::hip_check(hipGetLastError(), "score_backward_kernel");
}

}  // extern "C"
