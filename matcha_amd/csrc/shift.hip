// FFT shift match on gfx950.
//
// Reference behaviour (swing-research/Matcha, src/matcha/core/ShiftMatcher.py):
//   rotate_complex_volume :491-537, mask_full :442-453, extract_shifted_patch_batch
//   :558-616, apply_cosine_mask :540-555, rfftn(roll(., -N/2), "forward") :651-656,
//   compute_shift (complex space, no phase normalisation, score normalised,
//   upsample 1) :94-282, estimate_shift_like_relion :619-674.
//
// Per batch (6 launches):
//   bg     : background average of the shifted patch over the outer/cosine shell
//   fwd_x  : shifted + cosine-masked + rolled patch rows, two real rows per
//            complex FFT along x; keeps kx = 0..N/2, written as T[b][kx][z][y]
//   yz     : one workgroup per (b, kx) plane held in LDS: FFT along y and z
//            (1/N^3), ||S||^2, the rotated reference sampled on the fly
//            (trilinear, ifftshift, mask_full, ctf), ||R||^2, P = conj(R) S,
//            inverse FFT along z and y
//   inv_x  : inverse FFT along x of the half axis zero-padded to N, real part,
//            first argmax per tile of lines
//   final  : first argmax per particle, wrap to (-N/2, N/2], flip to the
//            reference's (z,y,x) labels, score / (||S|| ||R||)
// All 1-D transforms are in-LDS mixed-radix (4, 3, 2) Stockham FFTs.
#include "common.hpp"

#include <math.h>
#include <vector>

namespace {

constexpr int kRowThreads = 256;  // x-pass workgroups
constexpr int kRows = 16;         // rows per x-pass workgroup
constexpr int kPlaneThreads = 512;
constexpr int kMaxN = 128;
constexpr int kMaxIt = 4;         // FFT items per thread per stage and chunk
constexpr int kMaxRad = 8;

struct Radices {
  int n;
  int r[kMaxRad];
};

struct ShiftPlan {
  int N, H, max_batch;
  float radius, radius_p, width, wsum;  // cosine mask geometry, sum of weights
  Radices rad;
  float4* refpair;  // (N, N, N): (ref[z][y][x], ref[z][y][x+1]) of the centred reference
  float2* twid;     // (N) e^{-2 pi i j / N}
  float2* T;        // (B, H, N, N) work spectrum
  float* rnorm_part;  // (B, H)
  float* snorm_part;  // (B, H)
  float* bg_part;     // (B, N)
  float* bg_avg;      // (B)
  float* amax_val;    // (B, ntile_x)
  int* amax_idx;      // (B, ntile_x)
  int ntile_x;
  size_t plane_lds;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// (x + i y) * (i c)
__device__ __forceinline__ float2 cmuli(float2 a, float c) { return make_float2(-c * a.y, c * a.x); }

// In-place mixed-radix Stockham FFT of nl lines in LDS: element n of line l is
// buf[l * ls + n * es].  sign = -1 forward, +1 inverse (unscaled).  Work items
// run along the lines (lanes_on_lines, for ls == 1) or along the elements, so
// that consecutive lanes touch consecutive LDS words.  Needs
// nl * N / min(radix) <= NT * kMaxIt.
template <int NT>
__device__ void fft_lds(float2* buf, int nl, int ls, int es, int N, const Radices& rd,
                        const float2* tw, float sign, bool lanes_on_lines) {
  int p = 1;
  for (int s = 0; s < rd.n; ++s) {
    const int R = rd.r[s];
    const int per = N / R;
    const int items = nl * per;
    float2 v[kMaxIt][4];
#pragma unroll
    for (int it = 0; it < kMaxIt; ++it) {
      const int idx = threadIdx.x + it * NT;
      if (idx < items) {
        const int line = lanes_on_lines ? idx % nl : idx / per;
        const int i = lanes_on_lines ? idx / nl : idx % per;
        const float2* src = buf + line * ls + i * es;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (r < R) v[it][r] = src[r * per * es];
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kMaxIt; ++it) {
      const int idx = threadIdx.x + it * NT;
      if (idx >= items) continue;
      const int line = lanes_on_lines ? idx % nl : idx / per;
      const int i = lanes_on_lines ? idx / nl : idx % per;
      const int k = i % p;
      const int step = N / (p * R);
#pragma unroll
      for (int r = 1; r < 4; ++r) {
        if (r < R) {
          float2 w = tw[(r * k * step) % N];
          if (sign > 0.f) w.y = -w.y;
          v[it][r] = cmul(v[it][r], w);
        }
      }
      float2 y[4];
      if (R == 2) {
        y[0] = cadd(v[it][0], v[it][1]);
        y[1] = csub(v[it][0], v[it][1]);
      } else if (R == 3) {
        const float2 t = cadd(v[it][1], v[it][2]);
        const float2 d = csub(v[it][1], v[it][2]);
        const float2 m = make_float2(v[it][0].x - 0.5f * t.x, v[it][0].y - 0.5f * t.y);
        const float2 q = cmuli(d, sign * 0.86602540378443864676f);
        y[0] = cadd(v[it][0], t);
        y[1] = cadd(m, q);
        y[2] = csub(m, q);
      } else {
        const float2 a = cadd(v[it][0], v[it][2]);
        const float2 b = csub(v[it][0], v[it][2]);
        const float2 c = cadd(v[it][1], v[it][3]);
        const float2 d = cmuli(csub(v[it][1], v[it][3]), sign);
        y[0] = cadd(a, c);
        y[1] = cadd(b, d);
        y[2] = csub(a, c);
        y[3] = csub(b, d);
      }
      const int j = (i / p) * p * R + k;
      float2* dst = buf + line * ls + j * es;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r < R) dst[r * p * es] = y[r];
    }
    __syncthreads();
    p *= R;
  }
}

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

__device__ void make_shift_spec(const float* __restrict__ shift_zyx, int b, int N, ShiftSpec& sp) {
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

// cosine-masked value: inner r < radius keeps v, outer r > radius_p is bg,
// the shell blends with 0.5 + 0.5 cos(pi (radius_p - r) / width)
__device__ __forceinline__ float cosine_masked(float v, float bg, int d0, int d1, int d2, float radius,
                                               float radius_p, float width) {
  const float r2 = (float)(d0 * d0 + d1 * d1 + d2 * d2);
  const float r = sqrtf(r2);
  if (r < radius) return v;
  if (r > radius_p) return bg;
  const float rc = 0.5f + 0.5f * cosf(3.14159265358979323846f * (radius_p - r) / width);
  return (1.f - rc) * v + rc * bg;
}

}  // namespace

// ---------------------------------------------------------------------------
// background of the cosine mask
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void shift_bg_kernel(const float* __restrict__ vols, long long ldv,
                                                       int N, const float* __restrict__ shift_zyx,
                                                       float radius, float radius_p, float width,
                                                       float* __restrict__ bg_part) {
  const int z = blockIdx.x, b = blockIdx.y;
  ShiftSpec sp;
  make_shift_spec(shift_zyx, b, N, sp);
  const float* v = vols + (long long)b * ldv;
  float s = 0.f;
  const int dz = z - N / 2;
  for (int t = threadIdx.x; t < N * N; t += blockDim.x) {
    const int y = t / N, x = t - y * N;
    const int dy = y - N / 2, dx = x - N / 2;
    const float r = sqrtf((float)(dx * dx + dy * dy + dz * dz));
    if (r < radius) continue;  // inner
    const float val = patch_value(v, N, sp, z, y, x);
    if (r > radius_p) {
      s += val;
    } else {
      const float rc = 0.5f + 0.5f * cosf(3.14159265358979323846f * (radius_p - r) / width);
      s += rc * val;
    }
  }
  s = mch::wave_sum(s);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) bg_part[b * N + z] = red[0] + red[1] + red[2] + red[3];
}

__global__ void shift_bg_final_kernel(const float* __restrict__ bg_part, int N, int B, float wsum,
                                      float* __restrict__ bg_avg) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float s = 0.f;
  for (int z = 0; z < N; ++z) s += bg_part[b * N + z];
  bg_avg[b] = wsum > 0.f ? s / wsum : 0.f;
}

// ---------------------------------------------------------------------------
// forward FFT along x, two real rows per complex line
// ---------------------------------------------------------------------------
// workgroup: particle b, plane z, rows y0 .. y0+15 of the rolled, shifted,
// cosine-masked patch.  Complex line j carries rows y0+j (real) and y0+8+j (imag).
__global__ __launch_bounds__(kRowThreads) void shift_fwd_x_kernel(
    const float* __restrict__ vols, long long ldv, int N, int H, const float* __restrict__ shift_zyx,
    float radius, float radius_p, float width, const float* __restrict__ bg_avg, Radices rd,
    const float2* __restrict__ twg, float2* __restrict__ T) {
  __shared__ float2 buf[(kRows / 2) * kMaxN];
  __shared__ float2 tw[kMaxN];
  const int ytile = blockIdx.x, z = blockIdx.y, b = blockIdx.z;
  const int y0 = ytile * kRows;
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  ShiftSpec sp;
  make_shift_spec(shift_zyx, b, N, sp);
  const float bg = bg_avg[b];
  const float* v = vols + (long long)b * ldv;
  // roll by -N/2: rolled[i] = patch[(i + N/2) mod N]
  const int a0 = (z + N / 2) % N;
  for (int t = threadIdx.x; t < kRows * N; t += blockDim.x) {
    const int r = t / N, x = t - r * N;
    const int y = y0 + r;
    const int a1 = (y + N / 2) % N, a2 = (x + N / 2) % N;
    const float pv = patch_value(v, N, sp, a0, a1, a2);
    const float val = cosine_masked(pv, bg, a0 - N / 2, a1 - N / 2, a2 - N / 2, radius, radius_p, width);
    const int line = r & (kRows / 2 - 1);
    float* dst = reinterpret_cast<float*>(buf + line * N + x);
    dst[r >= kRows / 2 ? 1 : 0] = val;
  }
  __syncthreads();
  fft_lds<kRowThreads>(buf, kRows / 2, N, 1, N, rd, tw, -1.f, false);
  // unpack A = (Z[k] + conj Z[N-k]) / 2 (row y0+j), B = (Z[k] - conj Z[N-k]) / (2i) (row y0+8+j)
  float2* out = T + (long long)b * H * N * N;
  for (int t = threadIdx.x; t < H * kRows; t += blockDim.x) {
    const int kx = t / kRows, r = t - kx * kRows;
    const int line = r & (kRows / 2 - 1);
    const float2 Zk = buf[line * N + kx];
    const float2 Zn = buf[line * N + (N - kx) % N];
    float2 o;
    if (r < kRows / 2) {
      o = make_float2(0.5f * (Zk.x + Zn.x), 0.5f * (Zk.y - Zn.y));
    } else {
      // (Zk - conj Zn) / (2i)
      const float dx = Zk.x - Zn.x, dy = Zk.y + Zn.y;
      o = make_float2(0.5f * dy, -0.5f * dx);
    }
    out[((long long)kx * N + z) * N + y0 + r] = o;
  }
}

// ---------------------------------------------------------------------------
// one (b, kx) plane: FFT y, z; cross power with the rotated reference; inverse
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kPlaneThreads) void shift_yz_kernel(
    float2* __restrict__ T, const float4* __restrict__ refpair, int N, int H,
    const float* __restrict__ rot, const float* __restrict__ ctf, Radices rd,
    const float2* __restrict__ twg, float* __restrict__ rnorm_part, float* __restrict__ snorm_part) {
  extern __shared__ float4 smem4[];
  float2* plane = reinterpret_cast<float2*>(smem4);  // [z][y]
  float2* tw = plane + N * N;
  __shared__ float red[2][kPlaneThreads / 64];
  const int kx = blockIdx.x, b = blockIdx.y;
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  float2* P = T + ((long long)b * H + kx) * N * N;
  for (int t = threadIdx.x; t < N * N; t += blockDim.x) plane[t] = P[t];
  __syncthreads();
  const int chunk = (kMaxIt * kPlaneThreads * 2) / N;  // lines per FFT call (radix >= 2)
  // forward: along y (lines z, contiguous), then along z (lines y, stride N)
  for (int l0 = 0; l0 < N; l0 += chunk)
    fft_lds<kPlaneThreads>(plane + l0 * N, min(chunk, N - l0), N, 1, N, rd, tw, -1.f, false);
  for (int l0 = 0; l0 < N; l0 += chunk)
    fft_lds<kPlaneThreads>(plane + l0, min(chunk, N - l0), 1, N, N, rd, tw, -1.f, true);
  // cross power with the rotated reference sampled here
  const float* m = rot + b * 9;
  const float c = (float)N * 0.5f;
  const float sc = 2.0f / (float)(N - 1);  // 2 / (size_e - 2) with size_e = N + 1
  const float um = (float)(N - 1);
  const float inv = 1.0f / ((float)N * (float)N * (float)N);
  const int xsrc = kx + N / 2;
  const int jp = kx < N / 2 ? kx : kx - H;
  float ns = 0.f, nr = 0.f;
  for (int t = threadIdx.x; t < N * N; t += blockDim.x) {
    const int kz = t / N, ky = t - kz * N;
    float2 s = plane[t];
    s.x *= inv;
    s.y *= inv;
    ns += s.x * s.x + s.y * s.y;
    float2 rv = make_float2(0.f, 0.f);
    const int kp = kz < N / 2 ? kz : kz - N;
    const int ip = ky < N / 2 ? ky : ky - N;
    if (kp * kp + ip * ip + jp * jp <= (N / 2) * (N / 2)) {
      const int zsrc = (kz + N / 2) % N, ysrc = (ky + N / 2) % N;
      const float bx = (float)xsrc - c, by = (float)ysrc - c, bz = (float)zsrc - c;
      const float rx = m[0] * bx + m[1] * by + m[2] * bz + c;
      const float ry = m[3] * bx + m[4] * by + m[5] * bz + c;
      const float rz = m[6] * bx + m[7] * by + m[8] * bz + c;
      const float ix = ((rx * sc - 1.0f) + 1.0f) * 0.5f * um;
      const float iy = ((ry * sc - 1.0f) + 1.0f) * 0.5f * um;
      const float iz = ((rz * sc - 1.0f) + 1.0f) * 0.5f * um;
      const float fz0 = floorf(iz), fy0 = floorf(iy), fx0 = floorf(ix);
      const int z0 = (int)fz0, y0 = (int)fy0, x0 = (int)fx0;
      const float fz = iz - fz0, fy = iy - fy0, fx = ix - fx0;
      if (x0 >= -1 && x0 < N) {
        const bool left_in = x0 >= 0;
        const int xp = left_in ? x0 : 0;
#pragma unroll
        for (int dz = 0; dz < 2; ++dz) {
          const int zi = z0 + dz;
          if (zi < 0 || zi >= N) continue;
#pragma unroll
          for (int dy = 0; dy < 2; ++dy) {
            const int yi = y0 + dy;
            if (yi < 0 || yi >= N) continue;
            const float w = (dz ? fz : 1.f - fz) * (dy ? fy : 1.f - fy);
            const float4 q = refpair[((long long)zi * N + yi) * N + xp];
            // (value at x0, value at x0 + 1)
            const float2 v0 = left_in ? make_float2(q.x, q.y) : make_float2(0.f, 0.f);
            const float2 v1 = left_in ? make_float2(q.z, q.w) : make_float2(q.x, q.y);
            rv.x += w * ((1.f - fx) * v0.x + fx * v1.x);
            rv.y += w * ((1.f - fx) * v0.y + fx * v1.y);
          }
        }
      }
      if (ctf) {
        const float cv = ctf[(((long long)b * N + kz) * N + ky) * H + kx];
        rv.x *= cv;
        rv.y *= cv;
      }
    }
    nr += rv.x * rv.x + rv.y * rv.y;
    // conj(R) * S
    plane[t] = make_float2(rv.x * s.x + rv.y * s.y, rv.x * s.y - rv.y * s.x);
  }
  ns = mch::wave_sum(ns);
  nr = mch::wave_sum(nr);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = ns;
    red[1][threadIdx.x >> 6] = nr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float a = 0.f, bsum = 0.f;
    for (int w = 0; w < kPlaneThreads / 64; ++w) { a += red[0][w]; bsum += red[1][w]; }
    snorm_part[b * H + kx] = a;
    rnorm_part[b * H + kx] = bsum;
  }
  // inverse: along z, then along y
  for (int l0 = 0; l0 < N; l0 += chunk)
    fft_lds<kPlaneThreads>(plane + l0, min(chunk, N - l0), 1, N, N, rd, tw, 1.f, true);
  for (int l0 = 0; l0 < N; l0 += chunk)
    fft_lds<kPlaneThreads>(plane + l0 * N, min(chunk, N - l0), N, 1, N, rd, tw, 1.f, false);
  for (int t = threadIdx.x; t < N * N; t += blockDim.x) P[t] = plane[t];
}

// ---------------------------------------------------------------------------
// inverse FFT along x of the zero-padded half axis, real part, argmax per tile
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kRowThreads) void shift_inv_x_kernel(
    const float2* __restrict__ T, int N, int H, Radices rd, const float2* __restrict__ twg,
    float* __restrict__ amax_val, int* __restrict__ amax_idx, int ntile_x) {
  __shared__ float2 buf[kRows * kMaxN];
  __shared__ float2 tw[kMaxN];
  __shared__ float rv[4];
  __shared__ int ri[4];
  const int ytile = blockIdx.x, tz = blockIdx.y, b = blockIdx.z;
  const int y0 = ytile * kRows;
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = twg[k];
  const float2* in = T + (long long)b * H * N * N;
  for (int t = threadIdx.x; t < kRows * N; t += blockDim.x) {
    const int x = t / kRows, r = t - x * kRows;
    buf[r * N + x] = (x < H) ? in[((long long)x * N + tz) * N + y0 + r] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<kRowThreads>(buf, kRows, N, 1, N, rd, tw, 1.f, false);
  float bv = -INFINITY;
  int bi = 0x7fffffff;
  for (int t = threadIdx.x; t < kRows * N; t += blockDim.x) {
    const int r = t / N, x = t - r * N;
    const float v = buf[r * N + x].x;
    const int flat = (tz * N + y0 + r) * N + x;
    if (v > bv || (v == bv && flat < bi)) { bv = v; bi = flat; }
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
    const int tile = tz * (N / kRows) + ytile;
    amax_val[(long long)b * ntile_x + tile] = rv[0];
    amax_idx[(long long)b * ntile_x + tile] = ri[0];
  }
}

__global__ __launch_bounds__(256) void shift_final_kernel(const float* __restrict__ amax_val,
                                                          const int* __restrict__ amax_idx,
                                                          int ntile_x, const float* __restrict__ rn,
                                                          const float* __restrict__ sn, int H,
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
  for (int t = threadIdx.x; t < H; t += blockDim.x) {
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
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

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
  if (N < 16 || N > kMaxN || (N % kRows) || max_batch < 1) {
    mch::set_error("mch_shift_plan_create: N=%d unsupported (multiple of %d, <= %d)", N, kRows, kMaxN);
    return MCH_ERR_UNSUPPORTED;
  }
  Radices rd;
  rd.n = 0;
  int n = N;
  while (n % 4 == 0) { rd.r[rd.n++] = 4; n /= 4; }
  while (n % 3 == 0) { rd.r[rd.n++] = 3; n /= 3; }
  while (n % 2 == 0) { rd.r[rd.n++] = 2; n /= 2; }
  if (n != 1 || rd.n > kMaxRad) {
    mch::set_error("mch_shift_plan_create: N=%d is not 2^a 3^b", N);
    return MCH_ERR_UNSUPPORTED;
  }
  // FFT tiling limit of the row passes (kRows lines of N, radix >= 2)
  if (kRows * N / 2 > kRowThreads * kMaxIt) {
    mch::set_error("mch_shift_plan_create: N=%d too large for the row FFT tiling", N);
    return MCH_ERR_UNSUPPORTED;
  }
  const size_t plane_lds = ((size_t)N * N + N) * sizeof(float2);
  int dev = 0, maxlds = 0;
  int rc = mch::hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  rc = mch::hip_check(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
                      "hipDeviceGetAttribute");
  if (rc) return rc;
  if ((size_t)maxlds < plane_lds + 1024) {
    mch::set_error("mch_shift_plan_create: a %dx%d plane needs %zu B of LDS, device allows %d", N, N,
                   plane_lds, maxlds);
    return MCH_ERR_UNSUPPORTED;
  }
  if (plane_lds > 64 * 1024) {
    rc = mch::hip_check(hipFuncSetAttribute((const void*)shift_yz_kernel,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)plane_lds),
                        "hipFuncSetAttribute(shift_yz_kernel)");
    if (rc) return rc;
  }
  ShiftPlan* p = new ShiftPlan();
  memset(p, 0, sizeof(*p));
  p->N = N;
  p->H = N / 2 + 1;
  p->max_batch = max_batch;
  p->rad = rd;
  p->radius = box_radius;
  p->width = 5.0f;
  p->radius_p = box_radius + 5.0f;
  p->plane_lds = plane_lds;
  // sum of background weights: |outer| + sum(raised cosine)
  {
    double wsum = 0.0;
    for (int z = 0; z < N; ++z)
      for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) {
          const double r = sqrt((double)((z - N / 2) * (z - N / 2) + (y - N / 2) * (y - N / 2) +
                                         (x - N / 2) * (x - N / 2)));
          if (r < p->radius) continue;
          if (r > p->radius_p) wsum += 1.0;
          else wsum += (float)(0.5 + 0.5 * cos(M_PI * (p->radius_p - r) / 5.0));
        }
    p->wsum = (float)wsum;
  }
  p->ntile_x = N * (N / kRows);
  std::vector<float2> tw(N);
  for (int k = 0; k < N; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)N;
    tw[k] = make_float2((float)cos(a), (float)sin(a));
  }
  const size_t vol = (size_t)N * N * N;
  const size_t half = (size_t)N * N * p->H;
  std::vector<float4> pair(vol);
  {
    const float2* r = (const float2*)ref_host;
    for (size_t i = 0; i < vol; ++i) {
      const float2 a = r[i];
      const bool last = ((i % N) == (size_t)(N - 1));
      const float2 b = last ? make_float2(0.f, 0.f) : r[i + 1];
      pair[i] = make_float4(a.x, a.y, b.x, b.y);
    }
  }
#define MCH_TRY(expr, what) do { if ((rc = mch::hip_check((expr), what)) != MCH_OK) goto fail; } while (0)
  MCH_TRY(hipMalloc(&p->refpair, vol * sizeof(float4)), "hipMalloc(refpair)");
  MCH_TRY(hipMemcpy(p->refpair, pair.data(), vol * sizeof(float4), hipMemcpyHostToDevice), "copy ref");
  MCH_TRY(hipMalloc(&p->twid, N * sizeof(float2)), "hipMalloc(twid)");
  MCH_TRY(hipMemcpy(p->twid, tw.data(), N * sizeof(float2), hipMemcpyHostToDevice), "copy twid");
  MCH_TRY(hipMalloc(&p->T, (size_t)max_batch * half * sizeof(float2)), "hipMalloc(T)");
  MCH_TRY(hipMalloc(&p->rnorm_part, (size_t)max_batch * p->H * sizeof(float)), "hipMalloc");
  MCH_TRY(hipMalloc(&p->snorm_part, (size_t)max_batch * p->H * sizeof(float)), "hipMalloc");
  MCH_TRY(hipMalloc(&p->bg_part, (size_t)max_batch * N * sizeof(float)), "hipMalloc");
  MCH_TRY(hipMalloc(&p->bg_avg, (size_t)max_batch * sizeof(float)), "hipMalloc");
  MCH_TRY(hipMalloc(&p->amax_val, (size_t)max_batch * p->ntile_x * sizeof(float)), "hipMalloc");
  MCH_TRY(hipMalloc(&p->amax_idx, (size_t)max_batch * p->ntile_x * sizeof(int)), "hipMalloc");
#undef MCH_TRY
  *out_plan = p;
  return MCH_OK;
fail:
  (void)hipFree(p->refpair); (void)hipFree(p->twid); (void)hipFree(p->T);
  (void)hipFree(p->rnorm_part); (void)hipFree(p->snorm_part); (void)hipFree(p->bg_part);
  (void)hipFree(p->bg_avg); (void)hipFree(p->amax_val); (void)hipFree(p->amax_idx);
  delete p;
  return rc;
}

int mch_shift_plan_destroy(void* plan) {
  ShiftPlan* p = (ShiftPlan*)plan;
  if (!p) return MCH_OK;
  (void)hipFree(p->refpair); (void)hipFree(p->twid); (void)hipFree(p->T);
  (void)hipFree(p->rnorm_part); (void)hipFree(p->snorm_part); (void)hipFree(p->bg_part);
  (void)hipFree(p->bg_avg); (void)hipFree(p->amax_val); (void)hipFree(p->amax_idx);
  delete p;
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
  const int N = p->N, H = p->H;
  if (B < 1 || B > p->max_batch || ld_vol < (long long)N * N * N) {
    mch::set_error("mch_shift_search: bad sizes (B=%d max %d)", B, p->max_batch);
    return MCH_ERR_SIZE;
  }
  int rc;
  hipLaunchKernelGGL(shift_bg_kernel, dim3(N, B), dim3(256), 0, stream, subtomos, ld_vol, N,
                     shift_zyx, p->radius, p->radius_p, p->width, p->bg_part);
  if ((rc = mch::hip_check(hipGetLastError(), "shift_bg_kernel"))) return rc;
  hipLaunchKernelGGL(shift_bg_final_kernel, dim3((B + 63) / 64), dim3(64), 0, stream, p->bg_part, N,
                     B, p->wsum, p->bg_avg);
  if ((rc = mch::hip_check(hipGetLastError(), "shift_bg_final_kernel"))) return rc;
  hipLaunchKernelGGL(shift_fwd_x_kernel, dim3(N / kRows, N, B), dim3(kRowThreads), 0, stream,
                     subtomos, ld_vol, N, H, shift_zyx, p->radius, p->radius_p, p->width, p->bg_avg,
                     p->rad, p->twid, p->T);
  if ((rc = mch::hip_check(hipGetLastError(), "shift_fwd_x_kernel"))) return rc;
  hipLaunchKernelGGL(shift_yz_kernel, dim3(H, B), dim3(kPlaneThreads), p->plane_lds, stream, p->T,
                     p->refpair, N, H, rot, ctf, p->rad, p->twid, p->rnorm_part, p->snorm_part);
  if ((rc = mch::hip_check(hipGetLastError(), "shift_yz_kernel"))) return rc;
  hipLaunchKernelGGL(shift_inv_x_kernel, dim3(N / kRows, N, B), dim3(kRowThreads), 0, stream, p->T,
                     N, H, p->rad, p->twid, p->amax_val, p->amax_idx, p->ntile_x);
  if ((rc = mch::hip_check(hipGetLastError(), "shift_inv_x_kernel"))) return rc;
  hipLaunchKernelGGL(shift_final_kernel, dim3(B), dim3(256), 0, stream, p->amax_val, p->amax_idx,
                     p->ntile_x, p->rnorm_part, p->snorm_part, H, N, out_shift, out_score);
  return mch::hip_check(hipGetLastError(), "shift_final_kernel");
}

}  // extern "C"
