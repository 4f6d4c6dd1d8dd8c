// General-size path: grids that are not a power of two (oversized screens
// for a field of view, extruder seed screens of arbitrary width).
//
// The 2-D DFT is two passes of a direct DFT written as a tiled complex GEMM
// that also transposes:  out[b][k][r] = sum_j in[b][r][j] * W^(j*k mod n).
// Two passes give F X F in the original layout.  O(n^3) per grid, FP32
// with exact integer index reduction of the twiddle argument; used for
// one-time seeding and the (rare) field-of-view grids, not the headline
// power-of-two path.  The prologue (spectrum build) and epilogue
// (real/imag split, subharmonics, crop, scale) run as separate kernels.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "fft_core.hpp"
#include "sh_epilogue.hpp"

namespace pt {

constexpr int DT = 32;  // output tile edge
constexpr int DK = 16;  // reduction chunk

template <int S>
__global__ __launch_bounds__(256) void k_dft_T(const float2* __restrict__ in, float2* __restrict__ out, int n) {
  __shared__ float2 xs[DT][DK + 1];
  __shared__ float2 fs[DK][DT + 1];
  const int tid = threadIdx.x;
  const size_t plane = (size_t)n * n;
  const float2* src = in + (size_t)blockIdx.z * plane;
  float2* dst = out + (size_t)blockIdx.z * plane;
  const int r0 = blockIdx.y * DT, k0 = blockIdx.x * DT;
  const int tr = tid % 16, tk = tid / 16;  // lanes walk r (contiguous output)
  float2 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = make_float2(0.f, 0.f);
  for (int j0 = 0; j0 < n; j0 += DK) {
    for (int e = tid; e < DT * DK; e += 256) {
      const int rr = e / DK, jj = e % DK, r = r0 + rr, j = j0 + jj;
      xs[rr][jj] = (r < n && j < n) ? src[(size_t)r * n + j] : make_float2(0.f, 0.f);
    }
    for (int e = tid; e < DK * DT; e += 256) {
      const int jj = e / DT, kk = e % DT, j = j0 + jj, k = k0 + kk;
      float2 w = make_float2(0.f, 0.f);
      if (j < n && k < n) {
        long long t = ((long long)j * k) % n;
        if (2 * t > n) t -= n;  // [-n/2, n/2] -> argument in [-1, 1] half-turns
        float s, c;
        sincospif(2.0f * (float)t / (float)n, &s, &c);
        w = make_float2(c, S * s);
      }
      fs[jj][kk] = w;
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < DK; ++jj) {
      const float2 x0 = xs[tr * 2 + 0][jj], x1 = xs[tr * 2 + 1][jj];
      const float2 f0 = fs[jj][tk * 2 + 0], f1 = fs[jj][tk * 2 + 1];
      acc[0][0] = cadd(acc[0][0], cmul(x0, f0));
      acc[0][1] = cadd(acc[0][1], cmul(x0, f1));
      acc[1][0] = cadd(acc[1][0], cmul(x1, f0));
      acc[1][1] = cadd(acc[1][1], cmul(x1, f1));
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int k = k0 + tk * 2 + c;
    if (k >= n) continue;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int r = r0 + tr * 2 + a;
      if (r < n) dst[(size_t)k * n + r] = acc[a][c];
    }
  }
}

static int dft2_generic(const float2* in, float2* out, float2* work, int batch, int n, int sign, hipStream_t st) {
  if (batch > 65535) return fail(PT_ERR_ARG, "general-size DFT: batch > 65535");
  dim3 g((n + DT - 1) / DT, (n + DT - 1) / DT, batch);
  if (sign > 0) {
    k_dft_T<1><<<g, 256, 0, st>>>(in, work, n);
    k_dft_T<1><<<g, 256, 0, st>>>(work, out, n);
  } else {
    k_dft_T<-1><<<g, 256, 0, st>>>(in, work, n);
    k_dft_T<-1><<<g, 256, 0, st>>>(work, out, n);
  }
  PT_LAUNCH_CHECK();
  return PT_OK;
}

// ---------------------------------------------------------------- prologues
__global__ void k_gen_pro(const float* __restrict__ nre, const float* __restrict__ nim,
                          const float* __restrict__ amp, float2* __restrict__ work, size_t plane, size_t total) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const float a = amp[e % plane];
    work[e] = make_float2(nre[e] * a, nim[e] * a);
  }
}

__global__ void k_frame_pro(const float2* __restrict__ spectra, int L, int n, const double* __restrict__ disp,
                            double inv_period, float2* __restrict__ work, int frames) {
  const size_t plane = (size_t)n * n;
  const size_t total = plane * frames;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int f = (int)(e / plane);
    const size_t o = e % plane;
    const int k = (int)(o / n), j = (int)(o % n);
    const uint32_t kk = (uint32_t)signed_freq(k, n), jj = (uint32_t)signed_freq(j, n);
    float2 acc = make_float2(0.f, 0.f);
    for (int l = 0; l < L; ++l) {
      const uint32_t ux = cycles_fixed(disp[((size_t)f * L + l) * 2 + 0] * inv_period);
      const uint32_t uy = cycles_fixed(disp[((size_t)f * L + l) * 2 + 1] * inv_period);
      const float2 ph = phasor_fixed(kk * ux + jj * uy);
      acc = cadd(acc, cmul(spectra[(size_t)l * plane + o], ph));
    }
    work[e] = acc;
  }
}

// ---------------------------------------------------------------- epilogues
// One thread per output pixel; subharmonic field from the prep tables.
struct GenEpiArgs {
  const float2* work;
  float* out;
  const float* screen_scale;
  SubTables tab;
  int n, count, n_fft, levels, off, n_out, accumulate;
  float scale;
};

__global__ __launch_bounds__(256) void k_gen_epi(GenEpiArgs a) {
  const int n = a.n, P = a.levels;
  const int b = blockIdx.z;
  const int y = blockIdx.x * 32 + (threadIdx.x & 31);
  const int x = blockIdx.y * 8 + (threadIdx.x >> 5);
  if (x >= a.n_out || y >= a.n_out) return;
  const int i = x + a.off, j = y + a.off;
  const int s0 = b, s1 = b + a.n_fft;
  const float2 v = a.work[(size_t)b * n * n + (size_t)i * n + j];
  const size_t oplane = (size_t)a.n_out * a.n_out, o = (size_t)x * a.n_out + y;
  const float l0 = P > 0 ? sub_eval(a.tab, s0, i, j, n, P) : 0.f;
  const float f0 = a.scale * (a.screen_scale ? a.screen_scale[s0] : 1.0f);
  float* d0 = a.out + (size_t)s0 * oplane + o;
  const float v0 = (v.x + l0) * f0;
  *d0 = a.accumulate ? *d0 + v0 : v0;
  if (s1 < a.count) {
    const float l1 = P > 0 ? sub_eval(a.tab, s1, i, j, n, P) : 0.f;
    const float f1 = a.scale * (a.screen_scale ? a.screen_scale[s1] : 1.0f);
    float* d1 = a.out + (size_t)s1 * oplane + o;
    const float v1 = (v.y + l1) * f1;
    *d1 = a.accumulate ? *d1 + v1 : v1;
  }
}

struct FrameEpiArgs {
  const float2* work;
  float* out;
  SubTables tab;
  int n, levels, off, n_out;
  float scale;
};

__global__ __launch_bounds__(256) void k_frame_epi(FrameEpiArgs a) {
  const int n = a.n, P = a.levels, f = blockIdx.z;
  const int y = blockIdx.x * 32 + (threadIdx.x & 31);
  const int x = blockIdx.y * 8 + (threadIdx.x >> 5);
  if (x >= a.n_out || y >= a.n_out) return;
  const int i = x + a.off, j = y + a.off;
  const float2 v = a.work[(size_t)f * n * n + (size_t)i * n + j];
  const float lo = P > 0 ? sub_eval(a.tab, f, i, j, n, P) : 0.f;
  a.out[((size_t)f * a.n_out + x) * a.n_out + y] = (v.x + lo) * a.scale;
}

// Frame-pair epilogue: pair b holds frame 2b in its real part and frame
// 2b + 1 in its imaginary part.
__global__ __launch_bounds__(256) void k_pair_epi(FrameEpiArgs a, int frames) {
  const int n = a.n, P = a.levels, b = blockIdx.z;
  const int y = blockIdx.x * 32 + (threadIdx.x & 31);
  const int x = blockIdx.y * 8 + (threadIdx.x >> 5);
  if (x >= a.n_out || y >= a.n_out) return;
  const int i = x + a.off, j = y + a.off;
  const float2 v = a.work[(size_t)b * n * n + (size_t)i * n + j];
  const int f0 = 2 * b, f1 = f0 + 1;
  const size_t o = (size_t)x * a.n_out + y, oplane = (size_t)a.n_out * a.n_out;
  a.out[(size_t)f0 * oplane + o] = (v.x + (P > 0 ? sub_eval(a.tab, f0, i, j, n, P) : 0.f)) * a.scale;
  if (f1 < frames) a.out[(size_t)f1 * oplane + o] = (v.y + (P > 0 ? sub_eval(a.tab, f1, i, j, n, P) : 0.f)) * a.scale;
}

int launch_pair_epi(const float2* z, float* out, SubTables t, int n, int frames, int levels, int off, int n_out,
                    float scale, hipStream_t st) {
  FrameEpiArgs a;
  a.work = z;
  a.out = out;
  a.tab = t;
  a.n = n;
  a.levels = levels;
  a.off = off;
  a.n_out = n_out;
  a.scale = scale;
  k_pair_epi<<<dim3((n_out + 31) / 32, (n_out + 7) / 8, (frames + 1) / 2), 256, 0, st>>>(a, frames);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

static int grid_stride_blocks(size_t total) {
  size_t b = (total + 255) / 256;
  return (int)(b < 65536 ? b : 65536);
}

}  // namespace pt

using namespace pt;

// Entry points for non-power-of-two grids, called from the dispatchers in
// screens.hip (declared there).
int pt_fft2_generic(const void* in, void* out, void* work, int batch, int n, int sign, void* stream) {
  if (!work) return fail(PT_ERR_ARG, "pt_fft2: n=%d needs a workspace of batch*n*n c64", n);
  return dft2_generic((const float2*)in, (float2*)out, (float2*)work, batch, n, sign, (hipStream_t)stream);
}

int pt_generate_generic(const float* nre, const float* nim, const float* amp, const float* sh_noise,
                        const double* sh_amp, int levels, int count, int n, int off, int n_out, float scale,
                        const float* screen_scale, int accumulate, float* out, void* work, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int n_fft = (count + 1) / 2;
  const size_t plane = (size_t)n * n, total = plane * n_fft;
  float2* w0 = (float2*)work;
  float2* w1 = w0 + total;
  k_gen_pro<<<grid_stride_blocks(total), 256, 0, st>>>(nre, nim, amp, w0, plane, total);
  PT_LAUNCH_CHECK();
  int rc = dft2_generic(w0, w0, w1, n_fft, n, 1, st);
  if (rc) return rc;
  GenEpiArgs a;
  a.work = w0;
  a.out = out;
  a.screen_scale = screen_scale;
  a.tab = sub_tables(w0 + 2 * total, levels, n, count);
  a.n = n;
  a.count = count;
  a.n_fft = n_fft;
  a.levels = levels;
  a.off = off;
  a.n_out = n_out;
  a.accumulate = accumulate;
  a.scale = scale;
  rc = launch_gen_prep(sh_noise, sh_amp, count, levels, n, a.tab, st);
  if (rc) return rc;
  k_gen_epi<<<dim3((n_out + 31) / 32, (n_out + 7) / 8, n_fft), 256, 0, st>>>(a);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

int pt_frames_generic(const void* spectra, int L, int n, double pixel_scale, const double* disp, int frames,
                      const double* coeffs, int levels, int off, int n_out, float scale, float* out, void* work,
                      void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const size_t plane = (size_t)n * n, total = plane * frames;
  float2* w0 = (float2*)work;
  float2* w1 = w0 + total;
  const double inv_period = 1.0 / ((double)n * pixel_scale);
  k_frame_pro<<<grid_stride_blocks(total), 256, 0, st>>>((const float2*)spectra, L, n, disp, inv_period, w0,
                                                         frames);
  PT_LAUNCH_CHECK();
  int rc = dft2_generic(w0, w0, w1, frames, n, 1, st);
  if (rc) return rc;
  FrameEpiArgs a;
  a.work = w0;
  a.out = out;
  a.tab = sub_tables(w0 + 2 * total, levels, n, frames);
  a.n = n;
  a.levels = levels;
  a.off = off;
  a.n_out = n_out;
  a.scale = scale;
  rc = launch_frame_prep(disp, (const double2*)coeffs, frames, L, levels, n, inv_period, a.tab, st);
  if (rc) return rc;
  k_frame_epi<<<dim3((n_out + 31) / 32, (n_out + 7) / 8, frames), 256, 0, st>>>(a);
  PT_LAUNCH_CHECK();
  return PT_OK;
}
