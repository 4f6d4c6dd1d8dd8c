// Fused FFT passes for Monte-Carlo screens (PhaseScreen.generate) and the
// spectral frozen-flow frame engine, gfx950.
//
// A 2-D transform of an N x N grid is two passes over HBM:
//   pass A (rows):    256-thread workgroups (N*N/8 when smaller), M = 8T/N
//                     rows each; the
//                     prologue builds the spectrum on the fly (noise x PSD
//                     amplitude, or the phasor-weighted layer sum) and the
//                     rows are transformed along axis 1 in LDS;
//   pass B (columns): up to 1024-thread workgroups, a tile of M adjacent
//                     columns transformed along axis 0 in LDS; the epilogue
//                     takes the real/imag parts, adds the subharmonic field,
//                     scales, crops and writes float32 screens.
// The intermediate (one c64 grid per transform) is the only extra traffic.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>

#include "common.hpp"
#include "fft_core.hpp"
#include "sh_epilogue.hpp"

namespace pt {

template <int N>
struct RowsCfg {
  static constexpr int T = (N * N / 8 < 256) ? N * N / 8 : 256;
  static constexpr int M = 8 * T / N;
};
template <int N>
struct ColsCfg {
  static constexpr int T = (N * N / 8 < 1024) ? N * N / 8 : 1024;
  static constexpr int M = 8 * T / N;
};

// ----------------------------------------------------------------------------
// plain 2-D transform passes (pt_fft2)
// ----------------------------------------------------------------------------
template <int N, int S>
__global__ __launch_bounds__(256) void k_fft_rows(const float2* __restrict__ in, float2* __restrict__ out, int n1) {
  constexpr int T = RowsCfg<N>::T, M = RowsCfg<N>::M;
  __shared__ float2 lds[M * N];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  init_twiddles<N, S>(tw, tid, T);
  __syncthreads();
  const size_t base = (size_t)blockIdx.y * n1 * N + (size_t)blockIdx.x * M * N;
  auto load = [&](int m, int i) { return in[base + (size_t)m * N + i]; };
  auto store = [&](int m, int i, float2 v) { out[base + (size_t)m * N + i] = v; };
  wg_fft<N, T, M, false, S>(lds, tw, tid, load, store);
}

template <int N, int S>
__global__ __launch_bounds__(ColsCfg<N>::T) void k_fft_cols(const float2* __restrict__ in, float2* __restrict__ out, int n1) {
  constexpr int T = ColsCfg<N>::T, M = ColsCfg<N>::M;
  __shared__ float2 lds[M * N];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  init_twiddles<N, S>(tw, tid, T);
  __syncthreads();
  const size_t base = (size_t)blockIdx.y * N * n1 + (size_t)blockIdx.x * M;
  auto load = [&](int m, int i) { return in[base + (size_t)i * n1 + m]; };
  auto store = [&](int m, int i, float2 v) { out[base + (size_t)i * n1 + m] = v; };
  wg_fft<N, T, M, true, S>(lds, tw, tid, load, store);
}

// ----------------------------------------------------------------------------
// PhaseScreen.generate
// ----------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_gen_rows(const float* __restrict__ nre, const float* __restrict__ nim,
                                                  const float* __restrict__ amp, float2* __restrict__ work) {
  constexpr int T = RowsCfg<N>::T, M = RowsCfg<N>::M;
  __shared__ float2 lds[M * N];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  init_twiddles<N, 1>(tw, tid, T);
  __syncthreads();
  const size_t plane = (size_t)N * N;
  const size_t b = blockIdx.y;
  const size_t r0 = (size_t)blockIdx.x * M * N;
  const float* re = nre + b * plane + r0;
  const float* im = nim + b * plane + r0;
  const float* am = amp + r0;
  float2* out = work + b * plane + r0;
  auto load = [&](int m, int i) {
    const size_t o = (size_t)m * N + i;
    const float a = am[o];
    return make_float2(re[o] * a, im[o] * a);
  };
  auto store = [&](int m, int i, float2 v) { out[(size_t)m * N + i] = v; };
  wg_fft<N, T, M, false, 1>(lds, tw, tid, load, store);
}

struct GenArgs {
  const float2* work;
  float* out;
  const float* sh_noise;
  const double* sh_amp;
  const float* screen_scale;
  int count, n_fft, levels, off, n_out, accumulate, u_in_lds;
  float scale;
};

template <int N>
__global__ __launch_bounds__(ColsCfg<N>::T) void k_gen_cols(GenArgs a) {
  constexpr int T = ColsCfg<N>::T, M = ColsCfg<N>::M;
  __shared__ float2 lds[M * N];
  __shared__ float2 tw[N];
  extern __shared__ __align__(16) unsigned char dyn[];
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const int c0 = blockIdx.x * M;
  if (c0 + M <= a.off || c0 >= a.off + a.n_out) return;  // tile entirely outside the crop
  const int P = a.levels;
  const int s0 = b, s1 = b + a.n_fft;
  const bool has1 = s1 < a.count;
  ShLds sh = sh_carve(dyn, N, P, 2, M, a.u_in_lds != 0);
  init_twiddles<N, 1>(tw, tid, T);
  if (P > 0) {
    auto valid = [&](int t) { return t == 0 || has1; };
    auto block = [&](int t, int p, double2* C) {
      const int s = t == 0 ? s0 : s1;
      const float* re = a.sh_noise + ((size_t)(p * 2 + 0) * a.count + s) * 9;
      const float* im = a.sh_noise + ((size_t)(p * 2 + 1) * a.count + s) * 9;
      for (int k = 0; k < 9; ++k) {
        const double am = a.sh_amp[p * 9 + k];
        C[k] = make_double2((double)re[k] * am, (double)im[k] * am);
      }
    };
    sh_setup(sh, N, P, 2, M, c0, tid, T, block, valid);
  }
  __syncthreads();
  const size_t plane = (size_t)N * N;
  const float2* src = a.work + (size_t)b * plane + c0;
  const size_t oplane = (size_t)a.n_out * a.n_out;
  const float f0 = a.scale * (a.screen_scale ? a.screen_scale[s0] : 1.0f);
  const float f1 = has1 ? a.scale * (a.screen_scale ? a.screen_scale[s1] : 1.0f) : 0.0f;
  auto load = [&](int m, int i) { return src[(size_t)i * N + m]; };
  auto store = [&](int m, int i, float2 v) {
    const int x = i - a.off, y = c0 + m - a.off;
    if ((unsigned)x >= (unsigned)a.n_out || (unsigned)y >= (unsigned)a.n_out) return;
    float l0 = 0.0f, l1 = 0.0f;
    if (P > 0) {
      l0 = sh_eval(sh, 0, m, M, i, N, P);
      if (has1) l1 = sh_eval(sh, 1, m, M, i, N, P);
    }
    const size_t o = (size_t)x * a.n_out + y;
    float* d0 = a.out + (size_t)s0 * oplane + o;
    const float v0 = (v.x + l0) * f0;
    *d0 = a.accumulate ? *d0 + v0 : v0;
    if (has1) {
      float* d1 = a.out + (size_t)s1 * oplane + o;
      const float v1 = (v.y + l1) * f1;
      *d1 = a.accumulate ? *d1 + v1 : v1;
    }
  };
  wg_fft<N, T, M, true, 1>(lds, tw, tid, load, store);
}

// ----------------------------------------------------------------------------
// spectral frozen-flow frames
// ----------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_frame_rows(const float2* __restrict__ spectra, int L,
                                                    const double* __restrict__ disp, double inv_period,
                                                    float2* __restrict__ work) {
  constexpr int T = RowsCfg<N>::T, M = RowsCfg<N>::M;
  __shared__ float2 lds[M * N];
  __shared__ float2 tw[N];
  extern __shared__ __align__(16) unsigned char dyn[];
  float2* py = (float2*)dyn;          // [L][N]
  float2* px = py + (size_t)L * N;    // [L][M]
  uint32_t* UX = (uint32_t*)(px + (size_t)L * M);
  uint32_t* UY = UX + L;
  const int tid = threadIdx.x;
  const int f = blockIdx.y;
  const int row0 = blockIdx.x * M;
  for (int l = tid; l < L; l += T) {
    UX[l] = cycles_fixed(disp[((size_t)f * L + l) * 2 + 0] * inv_period);
    UY[l] = cycles_fixed(disp[((size_t)f * L + l) * 2 + 1] * inv_period);
  }
  init_twiddles<N, 1>(tw, tid, T);
  __syncthreads();
  for (int idx = tid; idx < L * N; idx += T) {
    const int l = idx / N, j = idx % N;
    py[idx] = phasor_fixed((uint32_t)signed_freq(j, N) * UY[l]);
  }
  for (int idx = tid; idx < L * M; idx += T) {
    const int l = idx / M, m = idx % M;
    px[idx] = phasor_fixed((uint32_t)signed_freq(row0 + m, N) * UX[l]);
  }
  __syncthreads();
  const size_t plane = (size_t)N * N;
  const float2* sp = spectra + (size_t)row0 * N;
  auto load = [&](int m, int i) {
    float2 acc = make_float2(0.0f, 0.0f);
    for (int l = 0; l < L; ++l) {
      const float2 s = sp[(size_t)l * plane + (size_t)m * N + i];
      const float2 ph = cmul(px[l * M + m], py[l * N + i]);
      const float2 t = cmul(s, ph);
      acc.x += t.x;
      acc.y += t.y;
    }
    return acc;
  };
  float2* out = work + (size_t)f * plane + (size_t)row0 * N;
  auto store = [&](int m, int i, float2 v) { out[(size_t)m * N + i] = v; };
  wg_fft<N, T, M, false, 1>(lds, tw, tid, load, store);
}

struct FrameArgs {
  const float2* work;
  float* out;
  const double* disp;     // (F, L, 2)
  const double2* coeffs;  // (P, L, 9)
  double inv_period;
  int L, levels, off, n_out, u_in_lds;
  float scale;
};

template <int N>
__global__ __launch_bounds__(ColsCfg<N>::T) void k_frame_cols(FrameArgs a) {
  constexpr int T = ColsCfg<N>::T, M = ColsCfg<N>::M;
  __shared__ float2 lds[M * N];
  __shared__ float2 tw[N];
  extern __shared__ __align__(16) unsigned char dyn[];
  const int tid = threadIdx.x;
  const int f = blockIdx.y;
  const int c0 = blockIdx.x * M;
  if (c0 + M <= a.off || c0 >= a.off + a.n_out) return;
  const int P = a.levels, L = a.L;
  ShLds sh = sh_carve(dyn, N, P, 1, M, a.u_in_lds != 0);
  // shifted 3x3 blocks of this frame: after the epilogue tables in LDS
  double2* Q = (double2*)(dyn + sh_lds_bytes(N, P, 1, M, a.u_in_lds != 0));  // [P][L] (+df_p phasor, x then y)
  double2* SB = Q + 2 * (size_t)P * L;                                         // [P][9]
  init_twiddles<N, 1>(tw, tid, T);
  if (P > 0) {
    frame_sh_blocks(Q, SB, a.disp, a.coeffs, f, L, P, a.inv_period, tid, T);
    auto valid = [&](int t) { return true; };
    auto block = [&](int t, int p, double2* C) {
      for (int k = 0; k < 9; ++k) C[k] = SB[p * 9 + k];
    };
    sh_setup(sh, N, P, 1, M, c0, tid, T, block, valid);
  }
  __syncthreads();
  const size_t plane = (size_t)N * N;
  const float2* src = a.work + (size_t)f * plane + c0;
  float* dst = a.out + (size_t)f * a.n_out * a.n_out;
  auto load = [&](int m, int i) { return src[(size_t)i * N + m]; };
  auto store = [&](int m, int i, float2 v) {
    const int x = i - a.off, y = c0 + m - a.off;
    if ((unsigned)x >= (unsigned)a.n_out || (unsigned)y >= (unsigned)a.n_out) return;
    const float lo = P > 0 ? sh_eval(sh, 0, m, M, i, N, P) : 0.0f;
    dst[(size_t)x * a.n_out + y] = (v.x + lo) * a.scale;
  };
  wg_fft<N, T, M, true, 1>(lds, tw, tid, load, store);
}

// ----------------------------------------------------------------------------
// host dispatch
// ----------------------------------------------------------------------------
static bool is_pow2_fast(int n) { return n >= 32 && n <= 2048 && (n & (n - 1)) == 0; }

// LDS budget per workgroup on gfx950 is 160 KiB.
static const size_t kLdsMax = 160 * 1024;

template <class K>
static int set_dyn_lds(K kernel, size_t bytes) {
  if (bytes > 48 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return fail(PT_ERR_HIP, "hipFuncSetAttribute(%zu B LDS): %s", bytes, hipGetErrorString(e));
  }
  return PT_OK;
}

template <int N>
static int fft2_pow2(const float2* in, float2* out, int batch, int sign, hipStream_t st) {
  dim3 gr(N / RowsCfg<N>::M, batch), gc(N / ColsCfg<N>::M, batch);
  if (sign > 0) {
    k_fft_rows<N, 1><<<gr, RowsCfg<N>::T, 0, st>>>(in, out, N);
    k_fft_cols<N, 1><<<gc, ColsCfg<N>::T, 0, st>>>(out, out, N);
  } else {
    k_fft_rows<N, -1><<<gr, RowsCfg<N>::T, 0, st>>>(in, out, N);
    k_fft_cols<N, -1><<<gc, ColsCfg<N>::T, 0, st>>>(out, out, N);
  }
  PT_LAUNCH_CHECK();
  return PT_OK;
}

static size_t static_lds_cols(int n, int M) { return (size_t)M * n * 8 + (size_t)n * 8; }

template <int N>
static int generate_pow2(const float* nre, const float* nim, const float* amp, GenArgs a, hipStream_t st) {
  constexpr int MC = ColsCfg<N>::M;
  const size_t stat = static_lds_cols(N, MC);
  a.u_in_lds = a.levels > 0 && stat + sh_lds_bytes(N, a.levels, 2, MC, true) <= kLdsMax;
  const size_t dyn = sh_lds_bytes(N, a.levels, 2, MC, a.u_in_lds != 0);
  if (stat + dyn > kLdsMax) return fail(PT_ERR_UNSUPPORTED, "generate: %d subharmonic levels exceed LDS at n=%d", a.levels, N);
  int rc = set_dyn_lds(k_gen_cols<N>, dyn);
  if (rc) return rc;
  k_gen_rows<N><<<dim3(N / RowsCfg<N>::M, a.n_fft), RowsCfg<N>::T, 0, st>>>(nre, nim, amp, (float2*)a.work);
  k_gen_cols<N><<<dim3(N / MC, a.n_fft), ColsCfg<N>::T, dyn, st>>>(a);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

template <int N>
static size_t frames_rows_lds(int L) {
  return (size_t)L * N * 8 + (size_t)L * RowsCfg<N>::M * 8 + (size_t)2 * L * 4;
}

template <int N>
static int frames_rows_pow2(const float2* spectra, int L, const double* disp, int F, double inv_period,
                            float2* work, hipStream_t st) {
  constexpr int MR = RowsCfg<N>::M;
  const size_t dyn_r = frames_rows_lds<N>(L);
  if (dyn_r + (size_t)MR * N * 8 + (size_t)N * 8 > kLdsMax)
    return fail(PT_ERR_UNSUPPORTED, "frames: %d layers exceed the LDS phasor table at n=%d", L, N);
  int rc = set_dyn_lds(k_frame_rows<N>, dyn_r);
  if (rc) return rc;
  k_frame_rows<N><<<dim3(N / MR, F), RowsCfg<N>::T, dyn_r, st>>>(spectra, L, disp, inv_period, work);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

template <int N>
static int frames_cols_pow2(int F, FrameArgs a, hipStream_t st) {
  constexpr int MC = ColsCfg<N>::M;
  const size_t stat = static_lds_cols(N, MC);
  const size_t extra = a.levels > 0 ? (size_t)a.levels * a.L * 16 * 2 + (size_t)a.levels * 9 * 16 : 0;
  a.u_in_lds = a.levels > 0 && stat + sh_lds_bytes(N, a.levels, 1, MC, true) + extra <= kLdsMax;
  const size_t dyn_c = sh_lds_bytes(N, a.levels, 1, MC, a.u_in_lds != 0) + extra;
  if (stat + dyn_c > kLdsMax) return fail(PT_ERR_UNSUPPORTED, "frames: subharmonic tables exceed LDS at n=%d", N);
  int rc = set_dyn_lds(k_frame_cols<N>, dyn_c);
  if (rc) return rc;
  k_frame_cols<N><<<dim3(N / MC, F), ColsCfg<N>::T, dyn_c, st>>>(a);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

template <int N>
static int frames_pow2(const float2* spectra, int L, const double* disp, int F, FrameArgs a, hipStream_t st) {
  int rc = frames_rows_pow2<N>(spectra, L, disp, F, a.inv_period, (float2*)a.work, st);
  if (rc) return rc;
  return frames_cols_pow2<N>(F, a, st);
}

#define PT_POW2_SWITCH(n, CALL) \
  switch (n) {                  \
    case 32: CALL(32);          \
    case 64: CALL(64);          \
    case 128: CALL(128);        \
    case 256: CALL(256);        \
    case 512: CALL(512);        \
    case 1024: CALL(1024);      \
    case 2048: CALL(2048);      \
    default: break;             \
  }

}  // namespace pt

using namespace pt;

// general-size path (generic.hip)
int pt_fft2_generic(const void* in, void* out, void* work, int batch, int n, int sign, void* stream);
int pt_generate_generic(const float* nre, const float* nim, const float* amp, const float* sh_noise,
                        const double* sh_amp, int levels, int count, int n, int off, int n_out, float scale,
                        const float* screen_scale, int accumulate, float* out, void* work, void* stream);
int pt_frames_generic(const void* spectra, int L, int n, double pixel_scale, const double* disp, int frames,
                      const double* coeffs, int levels, int off, int n_out, float scale, float* out, void* work,
                      void* stream);

extern "C" int pt_fft2(const void* in, void* out, void* work, int batch, int n0, int n1, int sign, void* stream) {
  if (!in || !out) return fail(PT_ERR_ARG, "pt_fft2: NULL array");
  if (batch < 0 || n0 < 1 || n1 < 1 || (sign != 1 && sign != -1)) return fail(PT_ERR_ARG, "pt_fft2: bad shape/sign");
  if (batch == 0) return PT_OK;
  hipStream_t st = (hipStream_t)stream;
  if (n0 == n1 && is_pow2_fast(n0)) {
#define CALL(N) return fft2_pow2<N>((const float2*)in, (float2*)out, batch, sign, st);
    PT_POW2_SWITCH(n0, CALL)
#undef CALL
  }
  if (n0 != n1) return fail(PT_ERR_UNSUPPORTED, "pt_fft2: only square grids (got %dx%d)", n0, n1);
  return pt_fft2_generic(in, out, work, batch, n0, sign, stream);
}

extern "C" int pt_generate(const float* noise_re, const float* noise_im, const float* amplitude,
                           const float* sh_noise, const double* sh_amp, int levels, int count, int n,
                           int crop_offset, int n_out, float scale, const float* screen_scale, int accumulate,
                           float* out, void* work, void* stream) {
  if (count < 1) return fail(PT_ERR_ARG, "pt_generate: count must be >= 1");
  if (!noise_re || !noise_im || !amplitude || !out || !work) return fail(PT_ERR_ARG, "pt_generate: NULL array");
  if (levels < 0 || (levels > 0 && (!sh_noise || !sh_amp))) return fail(PT_ERR_ARG, "pt_generate: bad subharmonics");
  if (crop_offset < 0 || n_out < 1 || crop_offset + n_out > n) return fail(PT_ERR_ARG, "pt_generate: crop outside grid");
  GenArgs a;
  a.work = (const float2*)work;
  a.out = out;
  a.sh_noise = sh_noise;
  a.sh_amp = sh_amp;
  a.screen_scale = screen_scale;
  a.count = count;
  a.n_fft = (count + 1) / 2;
  a.levels = levels;
  a.off = crop_offset;
  a.n_out = n_out;
  a.accumulate = accumulate ? 1 : 0;
  a.u_in_lds = 0;
  a.scale = scale;
  if (a.n_fft > 65535) return fail(PT_ERR_ARG, "pt_generate: count > 131070 per call");
  hipStream_t st = (hipStream_t)stream;
  if (is_pow2_fast(n)) {
#define CALL(N) return generate_pow2<N>(noise_re, noise_im, amplitude, a, st);
    PT_POW2_SWITCH(n, CALL)
#undef CALL
  }
  return pt_generate_generic(noise_re, noise_im, amplitude, sh_noise, sh_amp, levels, count, n, crop_offset,
                             n_out, scale, screen_scale, accumulate, out, work, stream);
}

// The two passes of pt_spectral_frames as separate calls (power-of-two n):
// pass A writes the row-transformed layer sum to `work`, pass B finishes the
// column transform and the epilogue.  Same arguments as pt_spectral_frames.
static int frames_check(const void* spectra, int layers, int n, double pixel_scale, const double* disp, int frames,
                        const double* sh_coeffs, int levels, int crop_offset, int n_out, float* out, void* work) {
  if (!disp || !work) return fail(PT_ERR_ARG, "spectral pass: NULL array");
  if (layers < 1 || frames < 0 || levels < 0 || (levels > 0 && !sh_coeffs))
    return fail(PT_ERR_ARG, "spectral pass: bad layers/frames/levels");
  if (!(pixel_scale > 0.0)) return fail(PT_ERR_ARG, "spectral pass: pixel_scale must be > 0");
  if (crop_offset < 0 || n_out < 1 || crop_offset + n_out > n) return fail(PT_ERR_ARG, "spectral pass: crop outside grid");
  if (frames > 65535) return fail(PT_ERR_ARG, "spectral pass: frames > 65535 per call");
  if (!is_pow2_fast(n)) return fail(PT_ERR_UNSUPPORTED, "spectral pass: split passes need a power-of-two n");
  (void)spectra;
  (void)out;
  return PT_OK;
}

extern "C" int pt_spectral_rows(const void* spectra, int layers, int n, double pixel_scale, const double* disp,
                                int frames, void* work, void* stream) {
  if (!spectra) return fail(PT_ERR_ARG, "pt_spectral_rows: NULL spectra");
  int rc = frames_check(spectra, layers, n, pixel_scale, disp, frames, nullptr, 0, 0, n, nullptr, work);
  if (rc || frames == 0) return rc;
  const double inv_period = 1.0 / ((double)n * pixel_scale);
  hipStream_t st = (hipStream_t)stream;
#define CALL(N) return frames_rows_pow2<N>((const float2*)spectra, layers, disp, frames, inv_period, (float2*)work, st);
  PT_POW2_SWITCH(n, CALL)
#undef CALL
  return fail(PT_ERR_UNSUPPORTED, "pt_spectral_rows: n=%d", n);
}

extern "C" int pt_spectral_cols(const void* work, int layers, int n, double pixel_scale, const double* disp,
                                int frames, const double* sh_coeffs, int levels, int crop_offset, int n_out,
                                float scale, float* out, void* stream) {
  if (!out) return fail(PT_ERR_ARG, "pt_spectral_cols: NULL out");
  int rc = frames_check(work, layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out, out,
                        (void*)work);
  if (rc || frames == 0) return rc;
  FrameArgs a;
  a.work = (const float2*)work;
  a.out = out;
  a.disp = disp;
  a.coeffs = (const double2*)sh_coeffs;
  a.inv_period = 1.0 / ((double)n * pixel_scale);
  a.L = layers;
  a.levels = levels;
  a.off = crop_offset;
  a.n_out = n_out;
  a.u_in_lds = 0;
  a.scale = scale;
  hipStream_t st = (hipStream_t)stream;
#define CALL(N) return frames_cols_pow2<N>(frames, a, st);
  PT_POW2_SWITCH(n, CALL)
#undef CALL
  return fail(PT_ERR_UNSUPPORTED, "pt_spectral_cols: n=%d", n);
}

extern "C" int pt_spectral_frames(const void* spectra, int layers, int n, double pixel_scale, const double* disp,
                                  int frames, const double* sh_coeffs, int levels, int crop_offset, int n_out,
                                  float scale, float* out, void* work, void* stream) {
  if (!spectra || !disp || !out || !work) return fail(PT_ERR_ARG, "pt_spectral_frames: NULL array");
  if (layers < 1 || frames < 0 || levels < 0 || (levels > 0 && !sh_coeffs))
    return fail(PT_ERR_ARG, "pt_spectral_frames: bad layers/frames/levels");
  if (!(pixel_scale > 0.0)) return fail(PT_ERR_ARG, "pt_spectral_frames: pixel_scale must be > 0");
  if (crop_offset < 0 || n_out < 1 || crop_offset + n_out > n) return fail(PT_ERR_ARG, "pt_spectral_frames: crop outside grid");
  if (frames == 0) return PT_OK;
  if (frames > 65535) return fail(PT_ERR_ARG, "pt_spectral_frames: frames > 65535 per call");
  FrameArgs a;
  a.work = (const float2*)work;
  a.out = out;
  a.disp = disp;
  a.coeffs = (const double2*)sh_coeffs;
  a.inv_period = 1.0 / ((double)n * pixel_scale);
  a.L = layers;
  a.levels = levels;
  a.off = crop_offset;
  a.n_out = n_out;
  a.u_in_lds = 0;
  a.scale = scale;
  hipStream_t st = (hipStream_t)stream;
  if (is_pow2_fast(n)) {
#define CALL(N) return frames_pow2<N>((const float2*)spectra, layers, disp, frames, a, st);
    PT_POW2_SWITCH(n, CALL)
#undef CALL
  }
  return pt_frames_generic(spectra, layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out,
                           scale, out, work, stream);
}
