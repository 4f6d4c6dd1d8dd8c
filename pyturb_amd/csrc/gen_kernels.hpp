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
  const size_t plane = (size_t)N * N;
  const size_t b = blockIdx.y;
  const size_t r0 = (size_t)blockIdx.x * M * N;
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

template <int N, bool ACC, int TC = kColsT, bool XCD = true, int RX = 8>
__device__ __forceinline__ void gen_cols_body(const GenArgs& a) {
  constexpr int T = ColsCfg<N, TC>::T, M = ColsCfg<N, TC>::M;
  constexpr bool ALIAS = gen_cols_alias<N>();
  constexpr size_t TILE_B = (size_t)Layout<N, T, M, true>::SIZE * sizeof(float2);
  using Sub = GenColSub<N, T, M>;
  __shared__ __align__(16) unsigned char smem[gen_cols_smem<N, T, M>()];
  __shared__ float2 tw[N];
  float2* lds = reinterpret_cast<float2*>(smem);
  Sub& sub = *reinterpret_cast<Sub*>(smem + (ALIAS ? 0 : TILE_B));
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const int c0 = (XCD ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x) * M;
  if (c0 + M <= a.off || c0 >= a.off + a.n_out) return;  // tile entirely outside the crop
  const int P = a.levels;
  const int s0 = b * a.s_mul, s1 = s0 + a.s_add;
  const bool has1 = s1 < a.count;
  const int sets[2] = {s0, has1 ? s1 : -1};
  init_twiddles<N, 1>(tw, tid, T);
  __syncthreads();
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

}  // namespace pt
