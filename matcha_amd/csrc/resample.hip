// Alternation resampler for gfx950: each alternation's particles are the
// patch at the running shift, normalised by its masked mean and standard
// deviation, masked, and rotated by the running rotation estimate; plus the
// device-side rotation-tracker update.
//
// Reference: BatchProcessor._extract_rotated_batch
// (src/matcha/align_subtomograms.py:240-252) = extract_subtomogram_patch_batch
// (src/matcha/utils/io_utils.py:94-178, normalize=True) followed by
// rotate_volumes (src/matcha/utils/volume_rotation.py:9-125); the tracker rule
// is update_rotation_estimate (src/matcha/utils/rotation_ops.py:23-34).
//
// Two passes per batch:
//   resample_stats_kernel   one (particle, z slice): masked sum, sum of squares
//                           and count of the shifted patch; the patch itself is
//                           written to the work buffer only for fractional
//                           shifts (integral shifts are read straight from the
//                           particle by the second pass).
//   resample_rotate_kernel  one (particle, z slice) of the output: reduce the
//                           particle's slice sums (FP64), then per voxel the
//                           trilinear sample at c + M (i - c) of the padded,
//                           normalised, masked patch, the 8 corner values
//                           normalised and masked on the fly.
// The patch is never written normalised: the normalisation is affine, so it
// is applied to the corner values inside the rotation pass.
#include "common.hpp"

#include <math.h>

namespace {

constexpr int kThreads = 256;

struct PatchSource {
  const float* v;  // the particle (integral shift) or its extracted patch (fractional)
  int n;           // edge of that volume
  int off[3];      // index offset (x, y, z) of patch coordinates in it
};

// Shift of particle b: (x, y, z) displacement; integral when every component is.
__device__ __forceinline__ bool load_shift(const float* __restrict__ shift, int b, float s[3]) {
  bool integral = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    s[d] = shift ? shift[b * 3 + d] : 0.f;
    integral &= s[d] == floorf(s[d]);
  }
  return integral;
}

__device__ __forceinline__ float fetch(const PatchSource& src, int qx, int qy, int qz) {
  const int x = qx + src.off[0], y = qy + src.off[1], z = qz + src.off[2];
  if ((unsigned)x >= (unsigned)src.n || (unsigned)y >= (unsigned)src.n || (unsigned)z >= (unsigned)src.n)
    return 0.f;
  return src.v[((long long)z * src.n + y) * src.n + x];
}

// Threads of a workgroup: 64 along x, 4 along y; one z slice per workgroup.
constexpr int kTx = 64, kTy = kThreads / kTx;

__global__ __launch_bounds__(kThreads) void resample_stats_kernel(
    const float* __restrict__ vols, long long ld_vol, int Nv, int N, const float* __restrict__ shift,
    int d2max, float* __restrict__ work, float4* __restrict__ partial) {
  const int z = blockIdx.x, b = blockIdx.y;
  const int tx = threadIdx.x % kTx, ty = threadIdx.x / kTx;
  float s[3];
  const bool integral = load_shift(shift, b, s);
  const float* v = vols + (long long)b * ld_vol;
  const int base = Nv / 2 - N / 2, c = N / 2;
  const int dz2 = (z - c) * (z - c);
  float sum = 0.f, sq = 0.f, cnt = 0.f;
  if (integral) {
    const int ox = base + (int)s[0], oy = base + (int)s[1], sz = z + base + (int)s[2];
    const bool zin = (unsigned)sz < (unsigned)Nv;
    for (int y = ty; y < N; y += kTy) {
      const int sy = y + oy;
      const bool yin = zin && (unsigned)sy < (unsigned)Nv;
      const float* row = v + ((long long)sz * Nv + sy) * Nv;
      const int dyz2 = dz2 + (y - c) * (y - c);
      for (int x = tx; x < N; x += kTx) {
        const int sx = x + ox;
        const float p = (yin && (unsigned)sx < (unsigned)Nv) ? row[sx] : 0.f;
        if (dyz2 + (x - c) * (x - c) <= d2max) {
          sum += p;
          sq += p * p;
          cnt += 1.f;
        }
      }
    }
  } else {
    float* w = work + (long long)b * N * N * N + (long long)z * N * N;
    for (int y = ty; y < N; y += kTy) {
      const int dyz2 = dz2 + (y - c) * (y - c);
      for (int x = tx; x < N; x += kTx) {
        // grid_sample of the particle at (x, y, z) + base + s, zeros outside
        const float p = mch::tri_sample_real(v, Nv, z + base + s[2], y + base + s[1], x + base + s[0]);
        w[y * N + x] = p;
        if (dyz2 + (x - c) * (x - c) <= d2max) {
          sum += p;
          sq += p * p;
          cnt += 1.f;
        }
      }
    }
  }
  __shared__ float red[3][kThreads / 64];
  sum = mch::wave_sum(sum);
  sq = mch::wave_sum(sq);
  cnt = mch::wave_sum(cnt);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sum;
    red[1][threadIdx.x >> 6] = sq;
    red[2][threadIdx.x >> 6] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = 0.f, q = 0.f, n = 0.f;
#pragma unroll
    for (int i = 0; i < kThreads / 64; ++i) {
      a += red[0][i];
      q += red[1][i];
      n += red[2][i];
    }
    partial[(long long)b * N + z] = make_float4(a, q, n, 0.f);
  }
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, mch::kWave);
  return v;
}

// Largest squared distance to the centre over the two integer positions
// a, a + 1 (offsets from the centre) of a cell edge.
__device__ __forceinline__ int far2(int a) { return max(a * a, (a + 1) * (a + 1)); }

__global__ __launch_bounds__(kThreads) void resample_rotate_kernel(
    const float* __restrict__ vols, long long ld_vol, int Nv, const float* __restrict__ work, int N,
    const float* __restrict__ shift, const float4* __restrict__ partial, const float* __restrict__ rot,
    int d2max, float* __restrict__ out, long long ld_out) {
  const int lb = mch::xcd_remap(blockIdx.x, gridDim.x);
  const int z = lb % N, b = lb / N;
  const int tx = threadIdx.x % kTx, ty = threadIdx.x / kTx;
  __shared__ float stat[2];
  if (threadIdx.x < 64) {
    double a = 0.0, q = 0.0, n = 0.0;
    for (int i = threadIdx.x; i < N; i += 64) {
      const float4 p = partial[(long long)b * N + i];
      a += p.x;
      q += p.y;
      n += p.z;
    }
    a = wave_sum_f64(a);
    q = wave_sum_f64(q);
    n = wave_sum_f64(n);
    if (threadIdx.x == 0) {
      // empty mask: every sample is masked out (the reference multiplies by it)
      const double mean = n > 0.0 ? a / n : 0.0;
      const double var = n > 0.0 ? q / n - mean * mean : 1.0;
      stat[0] = (float)mean;
      stat[1] = (float)(1.0 / sqrt(var));
    }
  }
  __syncthreads();
  const float mean = stat[0], inv_std = stat[1];

  float s[3];
  const bool integral = load_shift(shift, b, s);
  PatchSource src;
  if (integral) {
    const int base = Nv / 2 - N / 2;
    src = PatchSource{vols + (long long)b * ld_vol, Nv, {base + (int)s[0], base + (int)s[1], base + (int)s[2]}};
  } else {
    src = PatchSource{work + (long long)b * N * N * N, N, {0, 0, 0}};
  }
  const int n = src.n, nn = n * n;
  const float* M = rot + b * 9;
  const float m00 = M[0], m01 = M[1], m02 = M[2], m10 = M[3], m11 = M[4], m12 = M[5], m20 = M[6], m21 = M[7],
              m22 = M[8];
  const int c = N / 2;
  const float fc = (float)c, dz = (float)(z - c);
  float* o = out + (long long)b * ld_out + (long long)z * N * N;
  for (int y = ty; y < N; y += kTy) {
    const float dy = (float)(y - c);
    // sample position c + M (x - c, y - c, z - c) in (x, y, z) index units, stepped along x
    const float bx = fc + m01 * dy + m02 * dz - m00 * fc;
    const float by = fc + m11 * dy + m12 * dz - m10 * fc;
    const float bz = fc + m21 * dy + m22 * dz - m20 * fc;
    for (int x = tx; x < N; x += kTx) {
      const float fxp = (float)x;
      const float px = bx + m00 * fxp, py = by + m10 * fxp, pz = bz + m20 * fxp;
      const float fx0 = floorf(px), fy0 = floorf(py), fz0 = floorf(pz);
      const int x0 = (int)fx0, y0 = (int)fy0, z0 = (int)fz0;
      const float fx = px - fx0, fy = py - fy0, fz = pz - fz0;
      const int sx = x0 + src.off[0], sy = y0 + src.off[1], sz = z0 + src.off[2];
      float res;
      if (far2(x0 - c) + far2(y0 - c) + far2(z0 - c) <= d2max && (unsigned)x0 <= (unsigned)(N - 2) &&
          (unsigned)y0 <= (unsigned)(N - 2) && (unsigned)z0 <= (unsigned)(N - 2) &&
          (unsigned)sx <= (unsigned)(n - 2) && (unsigned)sy <= (unsigned)(n - 2) &&
          (unsigned)sz <= (unsigned)(n - 2)) {
        // the whole cell is inside the patch, the particle and the mask: the
        // weights sum to one, so normalise once after the interpolation
        const float* p = src.v + (sz * nn + sy * n + sx);
        const float c00 = p[0] + fx * (p[1] - p[0]);
        const float c01 = p[n] + fx * (p[n + 1] - p[n]);
        const float c10 = p[nn] + fx * (p[nn + 1] - p[nn]);
        const float c11 = p[nn + n] + fx * (p[nn + n + 1] - p[nn + n]);
        const float c0 = c00 + fy * (c01 - c00), c1 = c10 + fy * (c11 - c10);
        res = ((c0 + fz * (c1 - c0)) - mean) * inv_std;
      } else {
        float acc = 0.f;
#pragma unroll
        for (int kz = 0; kz < 2; ++kz) {
          const int qz = z0 + kz;
          const float wz = kz ? fz : 1.f - fz;
#pragma unroll
          for (int ky = 0; ky < 2; ++ky) {
            const int qy = y0 + ky;
            const float wy = ky ? fy : 1.f - fy;
            const int dyz2 = (qz - c) * (qz - c) + (qy - c) * (qy - c);
#pragma unroll
            for (int kx = 0; kx < 2; ++kx) {
              const int qx = x0 + kx;
              // corners outside the patch (its zero padding included) or the mask are 0
              if ((unsigned)qx >= (unsigned)N || (unsigned)qy >= (unsigned)N || (unsigned)qz >= (unsigned)N)
                continue;
              if (dyz2 + (qx - c) * (qx - c) > d2max) continue;
              const float wgt = wz * wy * (kx ? fx : 1.f - fx);
              acc += wgt * ((fetch(src, qx, qy, qz) - mean) * inv_std);
            }
          }
        }
        res = acc;
      }
      o[y * N + x] = res;
    }
  }
}

// q = exp(a z/2) exp(b y/2) exp(g z/2) as (w, x, y, z)
__device__ void quat_zyz(double a, double b, double g, double q[4]) {
  const double cb = cos(0.5 * b), sb = sin(0.5 * b);
  q[0] = cb * cos(0.5 * (a + g));
  q[1] = -sb * sin(0.5 * (a - g));
  q[2] = sb * cos(0.5 * (a - g));
  q[3] = cb * sin(0.5 * (a + g));
}

__global__ void tracker_update_kernel(const float* __restrict__ al, const float* __restrict__ be,
                                      const float* __restrict__ ga, int B, float* __restrict__ tracker,
                                      float* __restrict__ rot_particle, float* __restrict__ rot_reference) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double t[4];
  for (int i = 0; i < 4; ++i) t[i] = tracker[b * 4 + i];
  if (al) {
    double q[4];
    quat_zyz(al[b], be[b], ga[b], q);
    // exchange x and z, conjugate: r = conj(q~) = (w, -z, -y, -x)
    const double r[4] = {q[0], -q[3], -q[2], -q[1]};
    // Hamilton product t * r
    double u[4];
    u[0] = t[0] * r[0] - t[1] * r[1] - t[2] * r[2] - t[3] * r[3];
    u[1] = t[0] * r[1] + t[1] * r[0] + t[2] * r[3] - t[3] * r[2];
    u[2] = t[0] * r[2] - t[1] * r[3] + t[2] * r[0] + t[3] * r[1];
    u[3] = t[0] * r[3] + t[1] * r[2] - t[2] * r[1] + t[3] * r[0];
    const double inv = 1.0 / sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + u[3] * u[3]);
    for (int i = 0; i < 4; ++i) {
      t[i] = u[i] * inv;
      tracker[b * 4 + i] = (float)t[i];
    }
  }
  const double n2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + t[3] * t[3];
  const double s = n2 > 0.0 ? 2.0 / n2 : 0.0;
  const double w = t[0], x = t[1], y = t[2], z = t[3];
  const double R[9] = {1 - s * (y * y + z * z), s * (x * y - z * w), s * (x * z + y * w),
                       s * (x * y + z * w), 1 - s * (x * x + z * z), s * (y * z - x * w),
                       s * (x * z - y * w), s * (y * z + x * w), 1 - s * (x * x + y * y)};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (rot_particle) rot_particle[b * 9 + i * 3 + j] = (float)R[i * 3 + j];
      if (rot_reference) rot_reference[b * 9 + i * 3 + j] = (float)R[j * 3 + i];
    }
}

}  // namespace

extern "C" {

int mch_extract_rotated(const float* vols, long long ld_vol, int Nv, int B, int N, const float* shift,
                        const float* rot, double radius, float* work, float* partial, float* out,
                        long long ld_out, hipStream_t stream) {
  if (!vols || !rot || !partial || !out || B < 1 || N < 2 || N > 1024 || Nv < N || Nv > 1024 ||
      ld_vol < (long long)Nv * Nv * Nv || ld_out < (long long)N * N * N || !(radius >= 0.0)) {
    mch::set_error("mch_extract_rotated: bad arguments (B=%d N=%d Nv=%d)", B, N, Nv);
    return MCH_ERR_ARG;
  }
  if (!work && shift) {
    mch::set_error("mch_extract_rotated: a work buffer of B*N^3 floats is needed with a shift");
    return MCH_ERR_ARG;
  }
  // mask: sqrt(d2) < radius in FP64 for integer d2, i.e. d2 <= d2max
  const int c = N / 2;
  const long long dmax = 3LL * c * c;
  int d2max = -1;
  for (long long k = (long long)fmin((double)dmax, floor(radius * radius)) + 1; k >= 0; --k)
    if (sqrt((double)k) < radius) { d2max = (int)k; break; }
  hipLaunchKernelGGL(resample_stats_kernel, dim3(N, B), dim3(kThreads), 0, stream, vols, ld_vol, Nv, N, shift,
                     d2max, work, (float4*)partial);
  int rc = mch::hip_check(hipGetLastError(), "resample_stats_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(resample_rotate_kernel, dim3(N * B), dim3(kThreads), 0, stream, vols, ld_vol, Nv, work, N,
                     shift, (const float4*)partial, rot, d2max, out, ld_out);
  return mch::hip_check(hipGetLastError(), "resample_rotate_kernel");
}

int mch_tracker_update(const float* alpha, const float* beta, const float* gamma, int B, float* tracker,
                       float* rot_particle, float* rot_reference, hipStream_t stream) {
  if (!tracker || B < 1 || ((alpha != nullptr) != (beta != nullptr)) || ((alpha != nullptr) != (gamma != nullptr))) {
    mch::set_error("mch_tracker_update: bad arguments");
    return MCH_ERR_ARG;
  }
  hipLaunchKernelGGL(tracker_update_kernel, dim3((B + 255) / 256), dim3(256), 0, stream, alpha, beta, gamma, B,
                     tracker, rot_particle, rot_reference);
  return mch::hip_check(hipGetLastError(), "tracker_update_kernel");
}

}  // extern "C"
