// Row and column passes of the fused generate path, shared by the
// Monte-Carlo screens (pt_generate*) and pass B of the frame engine.
//
// The launch shape is a template parameter so that the autotuner can pick
// it per grid size at run time:
//   rows:    TR threads per workgroup, M = 8 TR / N rows per workgroup;
//   columns: TC threads per workgroup, a tile of M = 8 TC / N adjacent
//            columns, workgroups in XCD-contiguous (XCD) or flat order.
// Every shape runs the same per-element arithmetic (the Stockham stages of
// one row or column do not depend on how many rows or columns share a
// workgroup), so all shapes give bit-identical results.  The plan's largest
// radix RX (8 or 4) is a template parameter too; the radix-4 plan rounds
// differently (within 1e-6 relative RMS).  The defaults are the measured
// best on MI355X; tune_variants.hip instantiates the rest.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "fft_core.hpp"
#include "sh_epilogue.hpp"

namespace pt {

constexpr int kRowsT = 256;    // default row-pass workgroup
constexpr int kColsT = 1024;   // default column-pass workgroup

// Threads per workgroup clamped to [N/8, N*N/8]: at least one full row or
// column (eight values per thread), at most the whole grid.
constexpr int clamp_threads(int n, int t) {
  return t < n / 8 ? n / 8 : (t > n * n / 8 ? n * n / 8 : t);
}

template <int N, int TR = kRowsT>
struct RowsCfg {
  static constexpr int T = clamp_threads(N, TR);
  static constexpr int M = 8 * T / N;
};
template <int N, int TC = kColsT>
struct ColsCfg {
  static constexpr int T = clamp_threads(N, TC);
  static constexpr int M = 8 * T / N;
};

// Pass A of generate for G outer-scale groups at once: the transform is
// linear, so the groups' filtered noise is summed before one FFT,
//   spectrum = sum_g (re_g + i im_g) * amp_g,
// where group g's planes start at nre + g * gstride (real) and
// nre + g * gstride + im_off (imaginary), and amp_g = amp + g * N * N.
//
// gen_rows_item transforms row block rb (M rows) of transform b; the
// twiddle table tw holds size N and a barrier separates it from the
// previous use of lds.
template <int N, int T, int M, int RX>
__device__ __forceinline__ void gen_rows_item(const float* __restrict__ nre, ptrdiff_t im_off, size_t gstride, int G,
                                              const float* __restrict__ amp, float2* __restrict__ work, int rb,
                                              int bi, float2* lds, const float2* tw, int tid) {
  const size_t plane = (size_t)N * N;
  const size_t b = bi;
  const size_t r0 = (size_t)rb * M * N;
  const float* re = nre + b * plane + r0;
  const float* am = amp + r0;
  float2* out = work + b * plane + r0;
  auto load = [&](int m, int i) {
    const size_t o = (size_t)m * N + i;
    float2 acc = make_float2(0.f, 0.f);
    for (int g = 0; g < G; ++g) {
      const float* rg = re + (size_t)g * gstride;
      const float* ig = rg + im_off;
      const float a = am[(size_t)g * plane + o];
      // the noise is read once: non-temporal, so it does not evict the
      // amplitude rows that every screen pair re-reads from L2
      acc.x = fmaf(__builtin_nontemporal_load(rg + o), a, acc.x);
      acc.y = fmaf(__builtin_nontemporal_load(ig + o), a, acc.y);
    }
    return acc;
  };
  auto store = [&](int m, int i, float2 v) { store_nt(out + (size_t)m * N + i, v); };
  wg_fft<N, T, M, false, 1, false, RX>(lds, tw, tid, load, store);
}

template <int N, int TR = kRowsT, int RX = 8>
__global__ __launch_bounds__((RowsCfg<N, TR>::T)) void k_gen_rows(const float* __restrict__ nre, ptrdiff_t im_off,
                                                                 size_t gstride, int G, const float* __restrict__ amp,
                                                                 float2* __restrict__ work) {
  constexpr int T = RowsCfg<N, TR>::T, M = RowsCfg<N, TR>::M;
  __shared__ float2 lds[Layout<N, T, M, false>::SIZE];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  init_twiddles<N, 1>(tw, tid, T);
  __syncthreads();
  gen_rows_item<N, T, M, RX>(nre, im_off, gstride, G, amp, work, blockIdx.x, blockIdx.y, lds, tw, tid);
}

// Subharmonic tables of screen s (one 256-thread block per screen); with G
// groups the coefficient blocks of the groups are summed (the low-frequency
// field is linear in them).  Group g's noise is sh_noise + g * P * 2 * count
// * 9, its amplitudes sh_amp + g * P * 9.  With `ubasis` the finished basis
// table t.U is read, else screen 0 also writes it.
__device__ __forceinline__ void gen_prep_set(const float* __restrict__ sh_noise, const double* __restrict__ sh_amp,
                                             int count, int P, int n, int G, SubTables t, int ubasis, int s) {
  __shared__ double2 C[PMAX * 9];
  __shared__ double s_mean;
  const int tid = threadIdx.x;
  for (int q = tid; q < P * 9; q += 256) {
    const int p = q / 9, k = q % 9;
    double2 c = make_double2(0.0, 0.0);
    for (int g = 0; g < G; ++g) {
      const float* sn = sh_noise + (size_t)g * P * 2 * count * 9;
      const double am = sh_amp[(size_t)g * P * 9 + q];
      c.x += (double)sn[((size_t)(p * 2 + 0) * count + s) * 9 + k] * am;
      c.y += (double)sn[((size_t)(p * 2 + 1) * count + s) * 9 + k] * am;
    }
    C[q] = c;
  }
  __syncthreads();
  sub_fill_set(C, P, n, s, t, !ubasis && s == 0, &s_mean, tid, 256, 0, 1, ubasis ? t.U : nullptr);
}

// Column pass with a two-output epilogue.  Transform b yields the sets
// s0 = b * s_mul (real part) and s1 = s0 + s_add (imaginary part, when
// s1 < count): generate pairs screen b with screen b + n_fft (s_mul 1,
// s_add n_fft); the frame engine pairs frames 2b and 2b + 1 (2, 1).
struct GenArgs {
  const float2* work;
  float* out;
  const float* screen_scale;
  SubTables tab;
  int count, n_fft, levels, off, n_out, accumulate, s_mul, s_add;
  float scale;
};

// At n = 1024 the subharmonic staging block (ColSub) lives in the
// transform tile, which is free once the last stage has read it, so the
// basis at the base rows can be staged there too (it used to be evaluated
// per thread and level, a third of that column pass's VALU work).  Smaller
// sizes keep the block beside the tile.
template <int N>
constexpr bool gen_cols_alias() { return N == 1024; }
template <int N, int T, int M>
using GenColSub = ColSub<N, M, 2, T, (N <= 1024)>;
template <int N, int T, int M>
constexpr size_t gen_cols_smem() {
  constexpr size_t tile = (size_t)Layout<N, T, M, true>::SIZE * sizeof(float2);
  constexpr size_t sub = sizeof(GenColSub<N, T, M>);
  return gen_cols_alias<N>() ? (tile > sub ? tile : sub) : tile + sub;
}

// gen_cols_item transforms the column tile at c0 (M columns) of transform
// b and writes its two screens; smem holds the tile and the subharmonic
// staging block (beside the tile, or in it once the transform has read it
// when ALIAS), tw the twiddles of size N, and a barrier separates both from
// their previous use.  A tile outside the crop returns at once
// (workgroup-uniform).
template <int N, bool ACC, int T, int M, int RX, bool ALIAS = gen_cols_alias<N>()>
__device__ __forceinline__ void gen_cols_item(const GenArgs& a, unsigned char* smem, const float2* tw, int c0, int b,
                                              int tid) {
  constexpr size_t TILE_B = (size_t)Layout<N, T, M, true>::SIZE * sizeof(float2);
  using Sub = GenColSub<N, T, M>;
  float2* lds = reinterpret_cast<float2*>(smem);
  Sub& sub = *reinterpret_cast<Sub*>(smem + (ALIAS ? 0 : TILE_B));
  if (c0 + M <= a.off || c0 >= a.off + a.n_out) return;  // tile entirely outside the crop
  const int P = a.levels;
  const int s0 = b * a.s_mul, s1 = s0 + a.s_add;
  const bool has1 = s1 < a.count;
  const int sets[2] = {s0, has1 ? s1 : -1};
  // issued after the barrier: the loads complete under the tile load
  typename Sub::Regs pre;
  if (P > 0) pre = Sub::prefetch(a.tab, sets, c0, P, tid);
  const size_t plane = (size_t)N * N;
  const float2* src = a.work + (size_t)b * plane + c0;
  // The transform leaves each thread column c0 + m and the rows
  // x0 + row_step(c) * N/8, c < 8, in call order.
  float v0[8], v1[8];
  int cnt = 0;
  auto load = [&](int m, int i) { return src[(size_t)i * N + m]; };
  auto store = [&](int m, int i, float2 v) {
    v0[cnt] = v.x;
    v1[cnt] = v.y;
    ++cnt;
  };
  wg_fft<N, T, M, true, 1, false, RX>(lds, tw, tid, load, store);
  if (P > 0) {
    if (ALIAS) __syncthreads();  // every lane has read its last-stage inputs from the tile
    sub.commit(pre, P, tid);
  }
  __syncthreads();
  constexpr int RL = Plan<N, false, RX>::radix(Plan<N, false, RX>::NST - 1);
  const int m = tid % M, x0 = tid / M;
  const int y = c0 + m - a.off;
  if ((unsigned)y >= (unsigned)a.n_out) return;  // no barrier follows
  const size_t oplane = (size_t)a.n_out * a.n_out;
  float* o0 = a.out + (size_t)s0 * oplane + y;
  float* o1 = a.out + (size_t)s1 * oplane + y;
  float prev0[8], prev1[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    prev0[c] = prev1[c] = 0.0f;
    const int x = x0 + Sub::template row_step<RL>(c) * (N / 8) - a.off;
    if (ACC && (unsigned)x < (unsigned)a.n_out) {
      prev0[c] = o0[(size_t)x * a.n_out];
      if (has1) prev1[c] = o1[(size_t)x * a.n_out];
    }
  }
  if (P > 0) sub.template add_low2<RL>(m, x0, P, v0, v1, has1);
  const float f0 = a.scale * (a.screen_scale ? a.screen_scale[s0] : 1.0f);
  const float f1 = has1 ? a.scale * (a.screen_scale ? a.screen_scale[s1] : 1.0f) : 0.0f;
  if (!ACC && a.off == 0 && a.n_out == N) {  // whole grid (uniform): no crop tests
    float* p0 = o0 + (size_t)x0 * N;
#pragma unroll
    for (int c = 0; c < 8; ++c) p0[(size_t)Sub::template row_step<RL>(c) * (N / 8) * N] = v0[c] * f0;
    if (has1) {
      float* p1 = o1 + (size_t)x0 * N;
#pragma unroll
      for (int c = 0; c < 8; ++c) p1[(size_t)Sub::template row_step<RL>(c) * (N / 8) * N] = v1[c] * f1;
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int x = x0 + Sub::template row_step<RL>(c) * (N / 8) - a.off;
    if ((unsigned)x >= (unsigned)a.n_out) continue;
    o0[(size_t)x * a.n_out] = prev0[c] + v0[c] * f0;
    if (has1) o1[(size_t)x * a.n_out] = prev1[c] + v1[c] * f1;
  }
}

template <int N, bool ACC, int TC = kColsT, bool XCD = true, int RX = 8>
__device__ __forceinline__ void gen_cols_body(const GenArgs& a) {
  constexpr int T = ColsCfg<N, TC>::T, M = ColsCfg<N, TC>::M;
  __shared__ __align__(16) unsigned char smem[gen_cols_smem<N, T, M>()];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  const int c0 = (XCD ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x) * M;
  if (c0 + M <= a.off || c0 >= a.off + a.n_out) return;  // tile entirely outside the crop
  init_twiddles<N, 1>(tw, tid, T);
  __syncthreads();
  gen_cols_item<N, ACC, T, M, RX>(a, smem, tw, c0, blockIdx.y, tid);
}

// Writing column pass: the epilogue fits 64 VGPRs and the LDS stays under
// 80 KiB for N <= 1024, so two 1024-thread workgroups share a CU.
template <int N, int TC = kColsT, bool XCD = true, int RX = 8>
__global__ __launch_bounds__((ColsCfg<N, TC>::T)) __attribute__((amdgpu_waves_per_eu(8))) void k_gen_cols_set(GenArgs a) {
  gen_cols_body<N, false, TC, XCD, RX>(a);
}
// Accumulating variant (adds into out); its read-modify-write needs a few
// more registers, so no occupancy target (it would spill).  Default shape
// only.
template <int N>
__global__ __launch_bounds__(ColsCfg<N>::T) void k_gen_cols_acc(GenArgs a) {
  gen_cols_body<N, true>(a);
}

// ----------------------------------------------------------------------------
// Autotuned launch shapes (tune_variants.hip)
// ----------------------------------------------------------------------------
// Variant 0 of each knob is the default shape above.  tune_choice returns
// the variant selected for a grid size (set through pt_tune_set).
int tune_choice(int knob, int n);
// Launch a non-default variant; return PT_ERR_ARG for an unknown one.
int launch_gen_rows_variant(int n, int variant, const float* nre, ptrdiff_t im_off, size_t gstride, int G,
                            const float* amp, float2* work, int n_fft, hipStream_t st);
int launch_gen_cols_variant(int n, int variant, const GenArgs& a, hipStream_t st);
// The column pass of pt_generate* for a power-of-two n (the selected shape).
int launch_gen_cols(int n, const GenArgs& a, hipStream_t st);

}  // namespace pt
