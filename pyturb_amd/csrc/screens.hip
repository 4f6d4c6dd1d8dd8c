// Fused FFT passes for Monte-Carlo screens (PhaseScreen.generate) and the
// spectral frozen-flow frame engine, gfx950.
//
// A 2-D transform of an N x N grid is two passes over HBM:
//   pass A (rows):    256-thread workgroups (N*N/8 when smaller), M = 8T/N
//                     rows each; the prologue builds the spectrum on the fly
//                     (noise x PSD amplitude, or the phasor-weighted layer
//                     sum) and the rows are transformed along axis 1 in LDS;
//   pass B (columns): up to 1024-thread workgroups, a tile of M adjacent
//                     columns transformed along axis 0 in LDS; the epilogue
//                     takes the real/imag parts, adds the subharmonic field
//                     from the prep tables, scales, crops and writes float32.
// A small prep kernel (one block per screen or frame) fills the subharmonic
// tables between the passes.  The intermediate (one c64 grid per transform)
// is the only extra HBM traffic.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <type_traits>

#include "common.hpp"
#include "fft_core.hpp"
#include "boil.hpp"
#include "sh_epilogue.hpp"
#include "gen_kernels.hpp"

namespace pt {

// ----------------------------------------------------------------------------
// prep kernels (subharmonic tables)
// ----------------------------------------------------------------------------
// Subharmonic tables of each screen; with G groups the coefficient blocks
// of the groups are summed (the low-frequency field is linear in them).
// Group g's noise is sh_noise + g * P * 2 * count * 9, its amplitudes
// sh_amp + g * P * 9.
// The basis table U[p][x] = e_p(x), for calls whose sets read it instead of
// each evaluating it (k_gen_prep / k_frame_prep with `ubasis`).
__global__ __launch_bounds__(256) void k_sub_basis(float2* __restrict__ U, int P, int n) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx < P * n) U[idx] = sh_basis_e(idx / n, idx % n, n);
}

__global__ __launch_bounds__(256) void k_gen_prep(const float* __restrict__ sh_noise,
                                                  const double* __restrict__ sh_amp, int count, int P, int n,
                                                  int G, SubTables t, int ubasis) {
  gen_prep_set(sh_noise, sh_amp, count, P, n, G, t, ubasis, blockIdx.x);
}

// One workgroup's share (`part` of `parts`) of frame f's subharmonic tables;
// `dyn` holds (2 P L + 9 P) double2.
__device__ __forceinline__ void frame_prep_body(unsigned char* dyn, const double* __restrict__ disp,
                                                const double2* __restrict__ coeffs, size_t cstride, int L, int P,
                                                int n, double inv_period, SubTables t, int tr, int ubasis, int f,
                                                int part, int parts) {
  __shared__ double s_mean;
  double2* Q = (double2*)dyn;
  double2* SB = Q + 2 * (size_t)P * L;
  const int tid = threadIdx.x;
  frame_sh_blocks(Q, SB, disp, coeffs + (size_t)f * cstride, f, L, P, inv_period, tid, 256, tr != 0);
  sub_fill_set(SB, P, n, f, t, !ubasis && f == 0, &s_mean, tid, 256, part, parts, ubasis ? t.U : nullptr);
}

__global__ __launch_bounds__(256) void k_frame_prep(const double* __restrict__ disp,
                                                    const double2* __restrict__ coeffs, size_t cstride, int L,
                                                    int P, int n, double inv_period, SubTables t, int tr,
                                                    int ubasis) {
  extern __shared__ __align__(16) unsigned char dyn[];
  frame_prep_body(dyn, disp, coeffs, cstride, L, P, n, inv_period, t, tr, ubasis, blockIdx.x, blockIdx.y,
                  gridDim.y);
}

// Workgroups per frame (or pair) of the small per-call table kernels: one
// each for large calls, up to 8 so that a call of a few frames does not run
// on a handful of CUs.
static inline int table_split(int units) { return units >= 256 ? 1 : (256 / units < 8 ? 256 / units : 8); }

// Calls of at least this many sets write the basis table first and let every
// set read it (one small launch instead of P * n evaluations per set).
constexpr int UBASIS_MIN = 16;

static int launch_sub_basis(SubTables t, int P, int n, hipStream_t st) {
  k_sub_basis<<<(P * n + 255) / 256, 256, 0, st>>>(t.U, P, n);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

int launch_gen_prep(const float* sh_noise, const double* sh_amp, int count, int P, int n, SubTables t,
                    hipStream_t st, int groups) {
  if (P == 0) return PT_OK;
  if (P > PMAX) return fail(PT_ERR_UNSUPPORTED, "more than %d subharmonic levels", PMAX);
  const int ub = count >= UBASIS_MIN;
  if (ub) {
    const int rc = launch_sub_basis(t, P, n, st);
    if (rc) return rc;
  }
  k_gen_prep<<<count, 256, 0, st>>>(sh_noise, sh_amp, count, P, n, groups, t, ub);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

int launch_frame_prep(const double* disp, const double2* coeffs, int F, int L, int P, int n, double inv_period,
                      SubTables t, hipStream_t st, size_t cstride, bool transpose) {
  if (P == 0) return PT_OK;
  if (P > PMAX) return fail(PT_ERR_UNSUPPORTED, "more than %d subharmonic levels", PMAX);
  const size_t dyn = (size_t)(2 * P * L + 9 * P) * sizeof(double2);
  if (dyn > 60 * 1024) return fail(PT_ERR_UNSUPPORTED, "frames: %d layers x %d levels exceed LDS", L, P);
  const int ub = F >= UBASIS_MIN;
  if (ub) {
    const int rc = launch_sub_basis(t, P, n, st);
    if (rc) return rc;
  }
  k_frame_prep<<<dim3(F, table_split(F)), 256, dyn, st>>>(disp, coeffs, cstride, L, P, n, inv_period, t,
                                                          transpose ? 1 : 0, ub);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

// ----------------------------------------------------------------------------
// plain 2-D transform passes (pt_fft2)
// ----------------------------------------------------------------------------
template <int N, int S>
__global__ __launch_bounds__(256) void k_fft_rows(const float2* __restrict__ in, float2* __restrict__ out, int n1) {
  constexpr int T = RowsCfg<N>::T, M = RowsCfg<N>::M;
  __shared__ float2 lds[Layout<N, T, M, false>::SIZE];
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
  __shared__ float2 lds[Layout<N, T, M, true>::SIZE];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  init_twiddles<N, S>(tw, tid, T);
  __syncthreads();
  const size_t base = (size_t)blockIdx.y * N * n1 + (size_t)xcd_tile(blockIdx.x, gridDim.x) * M;
  auto load = [&](int m, int i) { return in[base + (size_t)i * n1 + m]; };
  auto store = [&](int m, int i, float2 v) { out[base + (size_t)i * n1 + m] = v; };
  wg_fft<N, T, M, true, S>(lds, tw, tid, load, store);
}


// ----------------------------------------------------------------------------
// PhaseScreen.generate
// ----------------------------------------------------------------------------
// ----------------------------------------------------------------------------
// spectral frozen-flow frames
// ----------------------------------------------------------------------------
// The frame engine transforms two frames at once.  A frame is
// Re(IFFT(X)) with X = sum_l S_l P_l, and Re(IFFT(X)) = IFFT(Herm X) with
// Herm X(k) = (X(k) + conj X(-k)) / 2, which is Hermitian, so
//   IFFT(Herm X_t1 + i Herm X_t2) = frame_t1 + i frame_t2.
// For a real displacement P_l(-k) = conj P_l(k) except on the Nyquist row
// and column, where the signed frequency -N/2 maps to itself; there
//   conj P(-k) = P(k) * E(k),  E = exp(2 pi i ([k0 = N/2] s_0 + [k1 = N/2] s_1) / dx).
// So with A_l = S_l / 2 and B_l(k) = conj(S_l(-k)) / 2 (folded once per
// call by k_frame_fold):
//   Herm X(k) = sum_l P_l(k) (A_l + B_l)            off the Nyquist lines,
//             = sum_l P_l(k) (A_l + B_l E_l(k))       on them.
// fold[l] holds A_l + B_l (A_l on the Nyquist lines); side[l][j] holds B_l
// on the Nyquist row (j = k1) and column (j = N + k0).
//
// Within each row the folded stack is stored chunk-major for pass A: the
// thread for j < n/8 starts from columns j + r * n/8 (r < 8), and column
// c = j + r * n/8 sits at position (r / 2) * (n/4) + 2 j + r % 2.  Chunk
// h = r / 2 of every thread is then one contiguous 16-byte-per-lane run, so
// each of a thread's four 16-byte loads is a fully coalesced wave load
// (with the eight values contiguous per thread, lanes 64 bytes apart, every
// load touched 32 cache lines and the texture-data path saturated).
//
// Uniform step.  The phase of frame t, layer l at k is the fixed-point sum
// s(k0) u_x + s(k1) u_y (mod 2^32, s the signed frequency).  When the second
// frame of every pair is its first moved by one and the same step per layer,
// d_l = u(2p + 1) - u(2p), then P_l,2p+1 = P_l,2p D_l with
// D_l(k) = exp(2 pi i 2^-32 (s(k0) d_x + s(k1) d_y)), and
//   (A_l + B_l)(P_1 + i P_2) = ((A_l + B_l)(1 + i D_l)) P_1,
// so the fold stores (A_l + B_l)(1 + i D_l) (A_l (1 + i D_l) on the Nyquist
// lines) and pass A forms one phasor product per layer and element instead of
// two.  k_frame_step decides per call: the step of pair 0 is d, and every pair
// must match it to within one unit of 2^-32 cycles per axis (host-rounded
// displacements of equally spaced times differ by that much).  Frame 2p + 1
// is then built with phase u(2p) + d, at most N/2 units = N/2 * 2^-32 cycles
// from its own.  Otherwise the fold is plain and pass A forms both products.
//
// The check runs in STEP_BLOCKS workgroups, each writing whether its share of
// the pairs matches: step[b] = 1 (b < STEP_BLOCKS); step[STEP_BLOCKS + 2l] =
// d_x, step[STEP_BLOCKS + 1 + 2l] = d_y.  The step is uniform when every
// step[b] is 1.
constexpr int FOLD_LMAX = 16;  // layers of the fused (paired) frame path
constexpr int STEP_BLOCKS = 16;
constexpr int STEP_WORDS = STEP_BLOCKS + 2 * FOLD_LMAX;
// Word STEP_WORDS of the step region records what the fold in the workspace
// holds: FOLD_PLAIN for the plain Hermitian fold (written by every fold
// without a uniform step, pt_spectral_fold's included), anything else for a
// fold that carries a step factor (1 + i D) and is valid only for the call
// that made it.  The pair tables of pt_spectral_rows_folded check it and
// write NaN tables for a fold that is not plain, so a stepped fold reused by
// mistake gives NaN frames instead of silently wrong ones.
constexpr uint32_t FOLD_PLAIN = 0x504C4E46u;

__global__ __launch_bounds__(1024) void k_frame_step(const double* __restrict__ disp, int L, int F,
                                                     double inv_period, uint32_t* __restrict__ step) {
  __shared__ uint32_t d0[2 * FOLD_LMAX];
  const int tid = threadIdx.x;
  const int E = 2 * L;
  if (F < 2) {
    if (tid == 0) step[blockIdx.x] = 0u;
    return;
  }
  if (tid < E) d0[tid] = cycles_fixed(disp[E + tid] * inv_period) - cycles_fixed(disp[tid] * inv_period);
  __syncthreads();
  int bad = 0;
  const int pairs = F / 2;  // pairs with two frames
  for (int q = blockIdx.x * 1024 + tid; q < (pairs - 1) * E; q += STEP_BLOCKS * 1024) {
    const int p = 1 + q / E, e = q % E;
    const uint32_t u0 = cycles_fixed(disp[(size_t)(2 * p) * E + e] * inv_period);
    const uint32_t u1 = cycles_fixed(disp[(size_t)(2 * p + 1) * E + e] * inv_period);
    const int32_t dd = (int32_t)(u1 - u0 - d0[e]);
    if (dd > 1 || dd < -1) bad = 1;
  }
  bad = __syncthreads_or(bad);
  if (tid == 0) step[blockIdx.x] = bad ? 0u : 1u;
  if (blockIdx.x == 0 && tid < E) step[STEP_BLOCKS + tid] = d0[tid];
}

__device__ __forceinline__ bool step_uniform(const uint32_t* __restrict__ step) {
  if (!step) return false;
  uint32_t all = 1u;
#pragma unroll
  for (int b = 0; b < STEP_BLOCKS; ++b) all &= step[b];
  return all != 0u;
}

// v (1 + i D) with D the uniform step's phasor at (k0, k1), or v itself
// for no uniform step.
__device__ __forceinline__ float2 fold_step(float2 v, const uint32_t* __restrict__ step, int l, int k0, int k1,
                                            int n) {
  if (!step_uniform(step)) return v;
  const uint32_t* d = step + STEP_BLOCKS + 2 * l;
  const uint32_t ph = (uint32_t)signed_freq(k0, n) * d[0] + (uint32_t)signed_freq(k1, n) * d[1];
  const float2 vd = cmul(v, phasor_fixed(ph));
  return make_float2(v.x - vd.y, v.y + vd.x);
}

__device__ __forceinline__ void fold_mark(uint32_t* mark, const uint32_t* step) {
  if (mark && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *mark = step_uniform(step) ? 0u : FOLD_PLAIN;
}

__global__ __launch_bounds__(256) void k_frame_fold(const float2* __restrict__ spectra, int n,
                                                    const uint32_t* __restrict__ step,
                                                    float2* __restrict__ fold, float2* __restrict__ side,
                                                    uint32_t* __restrict__ mark) {
  fold_mark(mark, step);
  const size_t plane = (size_t)n * n;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= plane) return;
  const int l = blockIdx.y;
  const int k0 = (int)(idx / n), k1 = (int)(idx % n);
  const int m0 = (n - k0) & (n - 1), m1 = (n - k1) & (n - 1);
  const float2 sv = spectra[(size_t)l * plane + idx];
  const float2 sm = spectra[(size_t)l * plane + (size_t)m0 * n + m1];
  const float2 A = make_float2(0.5f * sv.x, 0.5f * sv.y), B = make_float2(0.5f * sm.x, -0.5f * sm.y);
  const int h = n / 2, nb = n / 8;
  const int j = k1 % nb, r = k1 / nb;
  const size_t dst = (size_t)l * plane + (size_t)k0 * n + (size_t)(r >> 1) * 2 * nb + 2 * j + (r & 1);
  if (k0 == h || k1 == h) {
    fold[dst] = fold_step(A, step, l, k0, k1, n);
    side[(size_t)l * 2 * n + (k0 == h ? k1 : n + k0)] = B;
  } else {
    fold[dst] = fold_step(make_float2(A.x + B.x, A.y + B.y), step, l, k0, k1, n);
  }
}

// Pass A of a frame pair: the Hermitian layer sums of frames 2p and 2p + 1
// packed as Z = Herm X_t1 + i Herm X_t2, then the row transform.  Shift
// phasors use the exact fixed-point phases; the column factor is split as
// exp(2 pi i (8a + b - N/2) u) = t8[a] * t1[b], (k1 + N/2) mod N = 8a + b,
// so the per-layer LDS tables are N/8 + 8 entries instead of N.
template <int N>
struct PairCfg;
template <int N>
__host__ __device__ constexpr size_t pair_rows_lds(int L) {
  return (size_t)2 * L * (N / 8 + 8 + PairCfg<N>::M) * 8 + (size_t)2 * L * 4;
}

// Per frame pair, the row-independent phasor tables of pass A, computed once
// instead of in each of the N/M row blocks: for frames f in {2p, 2p+1} and
// layers l, t8[a] (a < N/8), t1[b] (b < 8) and the row phase U_x (bit
// pattern in .x).  Layout per pair: [2][L][N/8 + 9] float2.
template <int N>
__host__ __device__ constexpr int pair_tab_stride() { return N / 8 + 9; }

// It also sums the Nyquist-line terms of the pair, corr[j] = sum_l B_l(k) *
// (P_l,1(k) E_l,1(k) + i P_l,2(k) E_l,2(k)), for the Nyquist row k = (N/2, j)
// (j < N) and the Nyquist column k = (j - N, N/2) (N <= j < 2N, k0 != N/2),
// with the phase of P taken exactly as one fixed-point sum.  Pass A adds
// corr to the packed sum of those entries.
// TR (column-first engine): the tables split the row factor instead, t8/t1
// from the row phase U_x, and the per-thread phase entry is U_y.
// With a uniform step (step_uniform) the main term builds the second frame of
// each pair at phase u(2p) + d, so the Nyquist sums use that phase too, and
// the pair stays Hermitian in both parts.  The last pair of an odd frame
// count includes that second frame as well (it is not written).
// Per-layer phasor tables of the Nyquist sums for LB layers at a time,
// T8[a] and T1[b] per frame and varying axis (see pair_tables_body), and one
// term of a sum; shared by k_frame_pair_tables and pass A' with TAB, so the
// two build bit-identical Nyquist terms.
template <int N>
struct NyqTabs {
  static constexpr int LB = N <= 512 ? FOLD_LMAX : 4;
  float2 t8s[LB][2][2][N / 8], t1s[LB][2][2][8];  // [layer][frame][varying axis]
};
// Entry j of frame f's table along axis ax for layer l: T8[j] (j < N/8)
// or T1[j - N/8]; zero for an absent second frame.
template <int N>
__device__ __forceinline__ float2 nyq_tab_value(const uint32_t (&u)[2][FOLD_LMAX][2], int f, int ax, int l, int j,
                                                bool has_b) {
  constexpr int NB = N / 8;
  const uint32_t uv = u[f][l][ax], uo = u[f][l][ax ^ 1];
  // one phasor on either branch (straight-line code, so that entries
  // computed together overlap)
  const float2 v = phasor_fixed(j < NB ? (uint32_t)(8 * j - N / 2 + 1) * uv + (uint32_t)(N / 2) * uo
                                       : (uint32_t)(j - NB) * uv);
  return f != 0 && !has_b ? make_float2(0.f, 0.f) : v;
}
// Both axes' tables (ax_only < 0), or one axis' only.
template <int N>
__device__ __forceinline__ void nyq_build(NyqTabs<N>& tb, const uint32_t (&u)[2][FOLD_LMAX][2], int l0, int nl,
                                          bool has_b, int tid, int nthreads, int ax_only = -1) {
  constexpr int NB = N / 8, RU = 3;  // RU entries per thread at a time (independent chains)
  const int nax = ax_only < 0 ? 2 : 1, total = nl * 2 * nax * (NB + 8);
  for (int e0 = tid; e0 < total; e0 += RU * nthreads) {
    float2 v[RU];
#pragma unroll
    for (int h = 0; h < RU; ++h) {  // an index past the end recomputes the last entry
      const int e = min(e0 + h * nthreads, total - 1);
      const int lb = e / (2 * nax * (NB + 8)), r = e % (2 * nax * (NB + 8));
      const int fa = r / (NB + 8), j = r % (NB + 8);
      const int f = ax_only < 0 ? fa >> 1 : fa, ax = ax_only < 0 ? fa & 1 : ax_only;
      v[h] = nyq_tab_value<N>(u, f, ax, l0 + lb, j, has_b);
    }
#pragma unroll
    for (int h = 0; h < RU; ++h) {
      const int e = e0 + h * nthreads;
      if (e < total) {
        const int lb = e / (2 * nax * (NB + 8)), r = e % (2 * nax * (NB + 8));
        const int fa = r / (NB + 8), j = r % (NB + 8);
        const int f = ax_only < 0 ? fa >> 1 : fa, ax = ax_only < 0 ? fa & 1 : ax_only;
        if (j < NB)
          tb.t8s[lb][f][ax][j] = v[h];
        else
          tb.t1s[lb][f][ax][j - NB] = v[h];
      }
    }
  }
}
// Table position of Nyquist entry k along its varying axis: T8[id >> 3] and
// T1[id & 7].
template <int N>
__device__ __forceinline__ int nyq_index(int k) {
  return k <= N / 2 ? k + N / 2 - 1 : k - N / 2 - 1;
}
// acc += one layer's term from its table values (both frames) and side
// value sq.
__device__ __forceinline__ void nyq_term(float2& acc, float2 t8a, float2 t1a, float2 t8b, float2 t1b, float2 sq) {
  const float2 wa = cmul(t8a, t1a);
  const float2 wb = cmul(t8b, t1b);
  const float2 t = cmul(sq, make_float2(wa.x - wb.y, wa.y + wb.x));
  acc.x += t.x;
  acc.y += t.y;
}
// acc += the term of layer block entry lb for Nyquist entry q (q < 2N, not
// the Nyquist column's own N/2 entry), side value sq
template <int N>
__device__ __forceinline__ void nyq_add(float2& acc, const NyqTabs<N>& tb, int lb, int q, float2 sq) {
  const int ax = q < N ? 1 : 0, k = q < N ? q : q - N;
  const int id = nyq_index<N>(k);
  nyq_term(acc, tb.t8s[lb][0][ax][id >> 3], tb.t1s[lb][0][ax][id & 7], tb.t8s[lb][1][ax][id >> 3],
           tb.t1s[lb][1][ax][id & 7], sq);
}

template <int N, bool TR = false>
__device__ __forceinline__ void pair_tables_body(const double* __restrict__ disp, const float2* __restrict__ side,
                                                 int L, int F, double inv_period, const uint32_t* __restrict__ step,
                                                 float2* __restrict__ tab, float2* __restrict__ corr,
                                                 const uint32_t* __restrict__ guard, int pr, int part, int parts) {
  constexpr int NB = N / 8, S = pair_tab_stride<N>();
  // a reused fold (guard set) must be the plain one, else NaN tables
  const bool poison = guard && *guard != FOLD_PLAIN;
  const float2 nan2 = make_float2(__builtin_nanf(""), __builtin_nanf(""));
  const int first = part * 256 + threadIdx.x, stride = parts * 256;  // entries split over blocks
  const bool uni = step_uniform(step);
  // fixed-point phases of the pair's frames, once per workgroup:
  // u[f][l][axis]; frame 2p + 1 of a uniform step is u(2p) + d, and a frame
  // past the end (odd F, no uniform step) is absent
  __shared__ uint32_t u[2][FOLD_LMAX][2];
  if (threadIdx.x < 4 * L) {
    const int f = threadIdx.x / (2 * L), e = threadIdx.x % (2 * L);
    const bool virt = uni && f == 1;
    const int t = 2 * pr + (virt ? 0 : f);
    uint32_t v = t < F ? cycles_fixed(disp[(size_t)t * 2 * L + e] * inv_period) : 0u;
    if (virt) v += step[STEP_BLOCKS + e];
    u[f][e >> 1][e & 1] = v;
  }
  __syncthreads();
  const bool has_b = uni || 2 * pr + 1 < F;
  float2* tp = tab + (size_t)pr * 2 * L * S;
  // Nyquist sums.  With s'(k) the signed frequency that maps the Nyquist index
  // to +N/2 (the E factor folds into it), a row entry (q < N: k0 = N/2,
  // k1 = q) has phase (N/2) u_x + s'(k1) u_y and a column entry (q = N + k0,
  // k1 = N/2) s'(k0) u_x + (N/2) u_y.  Per layer the phasors along each axis
  // come from two small tables, exp(2 pi i 2^-32 s' u) = T8[a] T1[b] with
  // s' + N/2 - 1 = 8a + b, the other axis' N/2 term folded into T8: one
  // complex product per entry instead of a sine and cosine.
  // Tables of LB layers are built per pair of barriers (at most 37 KiB):
  // every layer at once up to n = 512, blocks of 4 above.
  constexpr int QMAX = (2 * N + 255) / 256;
  constexpr int LB = NyqTabs<N>::LB;
  __shared__ NyqTabs<N> nt;
  float2 acc[QMAX];
#pragma unroll
  for (int i = 0; i < QMAX; ++i) acc[i] = make_float2(0.f, 0.f);
  // n <= 1024: a block's side values are loaded before its tables are built
  // (a load per layer in the sum below would make it a chain of latencies)
  constexpr bool PRE = QMAX * LB <= 48;
  float2 sv[PRE ? LB : 1][PRE ? QMAX : 1];
  for (int l0 = 0; l0 < L; l0 += LB) {
    const int nl = min(LB, L - l0);
    if constexpr (PRE) {
#pragma unroll
      for (int lb = 0; lb < LB; ++lb)
#pragma unroll
        for (int i = 0; i < QMAX; ++i)
          sv[lb][i] = side[(size_t)(l0 + min(lb, nl - 1)) * 2 * N + min(first + stride * i, 2 * N - 1)];
    }
    __syncthreads();  // the previous layers' tables are no longer read
    nyq_build<N>(nt, u, l0, nl, has_b, threadIdx.x, 256);
    __syncthreads();
#pragma unroll
    for (int lb = 0; lb < LB; ++lb) {
      if (lb < nl) {
        const float2* sl = side + (size_t)(l0 + lb) * 2 * N;
#pragma unroll
        for (int i = 0; i < QMAX; ++i) {
          const int q = first + stride * i;
          if (q < 2 * N && !(q >= N && q - N == N / 2)) {
            float2 sq;
            if constexpr (PRE)
              sq = sv[PRE ? lb : 0][PRE ? i : 0];
            else
              sq = sl[q];
            nyq_add<N>(acc[i], nt, lb, q, sq);
          }
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < QMAX; ++i) {
    const int q = first + stride * i;
    if (q < 2 * N) corr[(size_t)pr * 2 * N + q] = poison ? nan2 : acc[i];
  }
  for (int q = first; q < 2 * L * S; q += stride) {
    const int fl = q / S, e = q % S;
    const int fb = fl / L, l = fl % L;
    const uint32_t ux = u[fb][l][0], uy = u[fb][l][1];
    const uint32_t ut = TR ? ux : uy, uo = TR ? uy : ux;  // table axis, per-thread axis
    float2 v;
    if (e < NB)
      v = phasor_fixed((uint32_t)(8 * e - N / 2) * ut);
    else if (e < NB + 8)
      v = phasor_fixed((uint32_t)(e - NB) * ut);
    else
      v = make_float2(__uint_as_float(uo), 0.f);
    tp[q] = poison && e < NB + 8 ? nan2 : v;
  }
}

template <int N, bool TR = false>
__global__ __launch_bounds__(256) void k_frame_pair_tables(const double* __restrict__ disp,
                                                           const float2* __restrict__ side, int L, int F,
                                                           double inv_period, const uint32_t* __restrict__ step,
                                                           float2* __restrict__ tab, float2* __restrict__ corr,
                                                           const uint32_t* __restrict__ guard) {
  pair_tables_body<N, TR>(disp, side, L, F, inv_period, step, tab, corr, guard, blockIdx.x, blockIdx.y, gridDim.y);
}

// Both per-call table kernels of a reused fold in one launch (calls of a few
// frames, where each is a few microseconds of latency on a handful of CUs):
// blocks [0, pairs * ta) build the pair tables (ta per pair), the rest the
// subharmonic tables (tb per frame).  The two are independent.
template <int N, bool TR>
__global__ __launch_bounds__(256) void k_frame_tables_prep(const double* __restrict__ disp,
                                                           const float2* __restrict__ side, int L, int F,
                                                           double inv_period, float2* __restrict__ tab,
                                                           float2* __restrict__ corr,
                                                           const uint32_t* __restrict__ guard, int ta,
                                                           const double2* __restrict__ coeffs, int P,
                                                           SubTables t, int tb) {
  extern __shared__ __align__(16) unsigned char dyn[];
  const int b = blockIdx.x, na = ((F + 1) / 2) * ta;
  if (b < na) {
    pair_tables_body<N, TR>(disp, side, L, F, inv_period, nullptr, tab, corr, guard, b / ta, b % ta, ta);
  } else {
    const int q = b - na;
    frame_prep_body(dyn, disp, coeffs, 0, L, P, N, inv_period, t, TR ? 1 : 0, 0, q / tb, q % tb, tb);
  }
}

// Pass-A workgroups: 256 threads (4 rows at n = 512), four per CU by LDS.
// 512-thread, 8-row workgroups (24 waves per CU instead of 16) measured the
// same: occupancy does not bound this kernel.
template <int N>
struct PairCfg {
  static constexpr int T = (N * N / 8 < 256) ? N * N / 8 : 256;
  static constexpr int M = 8 * T / N;
};

template <int N>
__global__ __launch_bounds__(PairCfg<N>::T) void k_frame_pair_rows(const float2* __restrict__ fold,
                                                                   const float2* __restrict__ corr, int L,
                                                                   const float2* __restrict__ ptab, int F,
                                                                   const uint32_t* __restrict__ step,
                                                                   float2* __restrict__ work) {
  constexpr int T = PairCfg<N>::T, M = PairCfg<N>::M, NB = N / 8;
  static_assert(T == M * NB, "one row position per thread in the first stage");
  __shared__ float2 lds[Layout<N, T, M, false>::SIZE];
  __shared__ float2 tw[N];
  extern __shared__ __align__(16) unsigned char dyn[];
  float2* t8 = (float2*)dyn;              // [2][L][NB]
  float2* t1 = t8 + 2 * L * NB;           // [2][L][8]
  float2* px = t1 + 2 * L * 8;            // [2][L][M]
  uint32_t* U = (uint32_t*)(px + 2 * L * M);  // [2][L] row phases
  const int tid = threadIdx.x;
  const int pr = blockIdx.y;
  const int row0 = blockIdx.x * M;
  const bool two = 2 * pr + 1 < F;
  constexpr int S = pair_tab_stride<N>();
  const float2* tp = ptab + (size_t)pr * 2 * L * S;
  for (int q = tid; q < 2 * L * S; q += T) {
    const int fl = q / S, e = q % S;
    const float2 v = tp[q];
    if (e < NB)
      t8[fl * NB + e] = v;
    else if (e < NB + 8)
      t1[fl * 8 + e - NB] = v;
    else
      U[fl] = __float_as_uint(v.x);
  }
  init_twiddles<N, 1>(tw, tid, T);
  __syncthreads();
  for (int q = tid; q < 2 * L * M; q += T) px[q] = phasor_fixed((uint32_t)signed_freq(row0 + q % M, N) * U[q / M]);
  __syncthreads();
  // first-stage positions of this thread: row m, columns jq + r * NB (r < 8)
  const int m = tid / NB, jq = tid % NB;
  const int k0 = row0 + m;
  const size_t plane = (size_t)N * N;
  const float4* fp = (const float4*)(fold + (size_t)k0 * N) + jq;  // chunk h at float4 h * NB + jq
  // z[r] accumulates the packed sum Herm X_t1 + i Herm X_t2 directly: per layer
  // and element, Q = P_1 + i P_2 (two phasor products folded into 8 FMAs),
  // then one complex multiply-accumulate z += g Q (4 FMAs).  With a uniform
  // step the fold carries (1 + i D), and Q = P_1 (4 FMAs).
  float2 z[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) z[r] = make_float2(0.f, 0.f);
  // For N >= 64 the eight columns share (k1 + N/2) mod 8, so the row phasor
  // times t1 is one product per layer and frame.
  constexpr bool BSAME = NB % 8 == 0;
  const int b0 = (jq + N / 2) & 7;
  auto layer_sum = [&](auto uni_tag) {
    constexpr bool UNI = decltype(uni_tag)::value;
    const bool two2 = !UNI && two;
    for (int l = 0; l < L; ++l) {
      const float2 pa = px[l * M + m], pb = px[(L + l) * M + m];
      const float2 ca0 = cmul(pa, t1[l * 8 + b0]);
      const float2 cb0 = two2 ? cmul(pb, t1[(L + l) * 8 + b0]) : make_float2(0.f, 0.f);
      float2 g[8];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const float4 v = fp[(size_t)l * (plane / 2) + (size_t)h * NB];
        g[2 * h] = make_float2(v.x, v.y);
        g[2 * h + 1] = make_float2(v.z, v.w);
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int qq = (jq + r * NB + N / 2) & (N - 1);
        float2 ca = ca0, cb = cb0;
        if (!BSAME) {
          ca = cmul(pa, t1[l * 8 + (qq & 7)]);
          cb = two2 ? cmul(pb, t1[(L + l) * 8 + (qq & 7)]) : make_float2(0.f, 0.f);
        }
        const float2 ta = t8[l * NB + (qq >> 3)];
        float qx = ca.x * ta.x;
        qx = fmaf(-ca.y, ta.y, qx);
        float qy = ca.x * ta.y;
        qy = fmaf(ca.y, ta.x, qy);
        if (!UNI) {
          const float2 tb = t8[(L + l) * NB + (qq >> 3)];
          qx = fmaf(-cb.x, tb.y, qx);
          qx = fmaf(-cb.y, tb.x, qx);
          qy = fmaf(cb.x, tb.x, qy);
          qy = fmaf(-cb.y, tb.y, qy);
        }
        z[r].x = fmaf(g[r].x, qx, z[r].x);
        z[r].x = fmaf(-g[r].y, qy, z[r].x);
        z[r].y = fmaf(g[r].x, qy, z[r].y);
        z[r].y = fmaf(g[r].y, qx, z[r].y);
      }
    }
  };
  if (step_uniform(step))
    layer_sum(std::true_type{});
  else
    layer_sum(std::false_type{});
  // Nyquist row / column terms of this pair (see k_frame_pair_tables)
  const float2* cp = corr + (size_t)pr * 2 * N;
  if (k0 == N / 2) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const float2 c = cp[jq + r * NB];
      z[r].x += c.x;
      z[r].y += c.y;
    }
  } else if ((N / 2) % NB == jq) {
    const int r = (N / 2) / NB;
    const float2 c = cp[N + k0];
    z[r].x += c.x;
    z[r].y += c.y;
  }
  int cnt = 0;
  auto load = [&](int, int) { return z[cnt++]; };
  float2* out = work + (size_t)pr * plane + (size_t)row0 * N;
  auto store = [&](int mm, int i, float2 v) { store_nt(out + (size_t)mm * N + i, v); };
  wg_fft<N, T, M, false, 1>(lds, tw, tid, load, store);
}

// ----------------------------------------------------------------------------
// Column-first frame engine
// ----------------------------------------------------------------------------
// The same two transforms in the other order: pass A' forms the packed pair
// sum Z and transforms it along columns in tiles of MT adjacent columns;
// pass B' transforms rows and runs the epilogue.  Every HBM access of the
// two passes is then a run of at least MT * 8 bytes (a full 128-byte line
// for MT = 16, n <= 512) or a whole row, where the row-first order's column
// pass writes its float32 frames in MT * 4-byte runs (64 B at n = 512).
// (Narrower tiles for calls of one frame pair, 4x the workgroups, measured
// no faster: a one-pair pass A is bound by its chain of per-layer load
// latencies, not by the CUs it occupies.)
template <int N>
struct PairColsCfg {
  static constexpr int MT = N <= 512 ? 16 : 8192 / N;  // columns per tile (<= 1024 threads)
  static constexpr int T = MT * N / 8;
  // waves per SIMD the register budget must allow: two 1024-thread tiles
  // per CU (64 VGPRs); n = 32 (one wave, 8 VGPR-heavy rows) is left free
  static constexpr int W = N >= 64 ? 8 : 1;
};

// Pass A' for calls of a few frame pairs (FEW_PAIRS at most): 4-column
// tiles (4x the workgroups of PairColsCfg at n <= 512), read from the fold's
// PairColsCfg tiles, with the fold values of PF layers loaded before any is
// summed, so a call is neither one long chain of per-layer load latencies
// nor bound by the L2 bandwidth of a few CUs.  Per element the arithmetic is
// PairColsCfg's (same layer order, same transform), so the intermediates are
// bit-identical.  n >= 128 (one wave of rows per column at least).
constexpr int FEW_PAIRS = 4;
template <int N>
struct FewColsCfg {
  static constexpr int MT = 4;
  static constexpr int T = MT * N / 8;
  // layers per prefetch block, PF x 4 float4 per thread: nine (a 9-layer
  // profile in one block) at one wave per SIMD, four at two or more; the
  // first block loads before the phasor tables are staged
  static constexpr int PF = T <= 256 ? 9 : 4;
  // the same when the launch builds its own tables (k_frame_pair_cols_tab):
  // the first block is requested just before the phasor tables are built,
  // and a wave cannot issue past a long run of loads until most of them
  // have left, so a short block lets the tables start sooner (n = 512, one
  // frame per call: 22.8 us per call with 3 layers, 24.1 us with 9)
  static constexpr int PF_TAB = T <= 256 ? 3 : 4;
};

// Folded stack for pass A' (same values as k_frame_fold): tile t = k1 / mt
// of mt columns; with k0 = j + r * n/8, lane (j, m = k1 % mt) reads its
// values r = 2h, 2h + 1 as float4 number ((h * (n/mt) + t) * n/8 + j) * mt + m
// of the layer, so the lanes of one load read one contiguous run.
__global__ __launch_bounds__(256) void k_frame_fold_tiles(const float2* __restrict__ spectra, int n, int mt,
                                                          const uint32_t* __restrict__ step,
                                                          float2* __restrict__ fold, float2* __restrict__ side,
                                                          uint32_t* __restrict__ mark) {
  fold_mark(mark, step);
  const size_t plane = (size_t)n * n;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= plane) return;
  const int l = blockIdx.y;
  const int k0 = (int)(idx / n), k1 = (int)(idx % n);
  const int m0 = (n - k0) & (n - 1), m1 = (n - k1) & (n - 1);
  const float2 sv = spectra[(size_t)l * plane + idx];
  const float2 sm = spectra[(size_t)l * plane + (size_t)m0 * n + m1];
  const float2 A = make_float2(0.5f * sv.x, 0.5f * sv.y), B = make_float2(0.5f * sm.x, -0.5f * sm.y);
  const int h = n / 2, nb = n / 8;
  const int t = k1 / mt, m = k1 % mt, j = k0 % nb, r = k0 / nb;
  const size_t dst =
      (size_t)l * plane + ((((size_t)(r >> 1) * (n / mt) + t) * nb + j) * mt + m) * 2 + (r & 1);
  if (k0 == h || k1 == h) {
    fold[dst] = fold_step(A, step, l, k0, k1, n);
    side[(size_t)l * 2 * n + (k0 == h ? k1 : n + k0)] = B;
  } else {
    fold[dst] = fold_step(make_float2(A.x + B.x, A.y + B.y), step, l, k0, k1, n);
  }
}

// Bytes of pass A's dynamic LDS: the transform tile, with the phasor tables
// (t8, t1 per layer and frame over the row index, the per-column factors
// and the column phases) aliased at its start; they are dead before the
// first stage writes the tile.
template <int N, int MT = PairColsCfg<N>::MT>
__host__ __device__ constexpr size_t pair_cols_lds(int L) {
  const size_t tile = (size_t)Layout<N, MT * N / 8, MT, true>::SIZE * 8;
  const size_t tabs = (size_t)2 * L * (N / 8 + 8 + MT) * 8 + (size_t)2 * L * 4;
  return tile > tabs ? tile : tabs;
}

// Pass A' of a frame pair over one column tile: per layer the phasor is
// (column factor, per thread) x (row factor t8[a] t1[b], (k0 + N/2) mod N
// = 8a + b); the packed sum in FMA form as in pass A; the Nyquist terms from
// k_frame_pair_tables; then the column transform, written as MT-column
// runs of the c64 intermediate.
// What pass A' needs to build its own per-call tables (TAB): the
// displacements, the fold's side values and marker, and for the subharmonic
// tables of the same frames (P > 0) the coefficients and table set.
struct PairTabArgs {
  const double* disp;
  const float2* side;
  const uint32_t* guard;
  double inv_period;
  const double2* coeffs;
  int P;
  SubTables tab;
  // the Nyquist column's sums, N per pair, written by the launch's producer
  // workgroup of each pair, which then sets nflag[pair] to token; the tile
  // that holds column N/2 waits for it (at most spin_limit polls, NaN
  // frames past that) and clears it
  float2* ncol;
  unsigned long long* nflag;
  unsigned long long token;
  unsigned spin_limit;
};

// The Nyquist column's sums of pair pr (k1 = N/2, every k0 but N/2) with
// k_frame_pair_tables' tables and terms, layers in order, LB per pair of
// barriers: thread tid (of T = N / 2) owns k0 = tid + i T, sc its side
// values.  Every thread of the workgroup calls it.
template <int N, int T>
__device__ __forceinline__ void nyq_col_sums(const uint32_t (&u)[2][FOLD_LMAX][2], const float2 (&sc)[2][FOLD_LMAX],
                                             int L, bool has_b, int tid, float2 (&acc)[2]) {
  static_assert(2 * T == N, "two Nyquist column entries per thread");
  constexpr int LB = NyqTabs<N>::LB;
  __shared__ NyqTabs<N> nt;
#pragma unroll
  for (int i = 0; i < 2; ++i) acc[i] = make_float2(0.f, 0.f);
#pragma unroll
  for (int bl = 0; bl < FOLD_LMAX / LB; ++bl) {
    const int l0 = bl * LB;
    if (l0 >= L) break;  // workgroup-uniform
    const int nl = min(LB, L - l0);
    if (bl > 0) __syncthreads();  // the previous layers' tables are no longer read
    nyq_build<N>(nt, u, l0, nl, has_b, tid, T, 0);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int k0 = tid + i * T;
      if (k0 != N / 2) {
#pragma unroll
        for (int lb = 0; lb < LB; ++lb)
          if (lb < nl) nyq_add<N>(acc[i], nt, lb, N + k0, sc[i][l0 + lb]);
      }
    }
  }
}

// Producer workgroup of pair blockIdx.y: the pair's phases (as the tiles
// work them out), its Nyquist column sums into ta.ncol, then the flag.
template <int N, int T>
__device__ void nyq_col_producer(int L, int F, const PairTabArgs& ta) {
  __shared__ uint32_t u[2][FOLD_LMAX][2];
  const int tid = threadIdx.x, pr = blockIdx.y;
  const uint32_t gv = ta.guard ? *ta.guard : FOLD_PLAIN;
  double dv = 0.0;
  if (tid < 4 * L) {
    const int f = tid / (2 * L), t = 2 * pr + f;
    dv = ta.disp[(size_t)min(t, F - 1) * 2 * L + tid % (2 * L)];
  }
  float2 sc[2][FOLD_LMAX];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int l = 0; l < FOLD_LMAX; ++l)
      sc[i][l] = l < L ? ta.side[(size_t)l * 2 * N + N + tid + i * T] : make_float2(0.f, 0.f);
  if (tid < 4 * L) {
    const int f = tid / (2 * L), e = tid % (2 * L), t = 2 * pr + f;
    u[f][e >> 1][e & 1] = t < F ? cycles_fixed(dv * ta.inv_period) : 0u;
  }
  __syncthreads();
  float2 acc[2];
  nyq_col_sums<N, T>(u, sc, L, 2 * pr + 1 < F, tid, acc);
  const bool poison = gv != FOLD_PLAIN;
  const float2 nan2 = make_float2(__builtin_nanf(""), __builtin_nanf(""));
#pragma unroll
  for (int i = 0; i < 2; ++i) ta.ncol[(size_t)pr * N + tid + i * T] = poison ? nan2 : acc[i];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this wave's sums reach memory before the flag
  __syncthreads();
  if (tid == 0) __hip_atomic_store(ta.nflag + pr, ta.token, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// MT: columns of this launch's tiles; the fold is laid out in tiles of
// PairColsCfg<N>::MT columns (MF).  PF > 1: the fold values of PF layers are
// loaded before the first of them is summed.  TAB: the workgroup builds the
// pair tables of its tile itself (t8, t1, the column phases and its
// Nyquist terms, with k_frame_pair_tables' arithmetic) and its share of the
// pair's subharmonic tables (k_frame_prep's), while its fold loads are in
// flight, instead of reading ptab and corr (calls of a reused fold, no
// uniform step).
template <int N, int MT, int PF, bool TAB = false>
__device__ __forceinline__ void pair_cols_body(const float2* __restrict__ fold, const float2* __restrict__ corr, int L,
                                               const float2* __restrict__ ptab, int F,
                                               const uint32_t* __restrict__ step, float2* __restrict__ work,
                                               const PairTabArgs& ta = PairTabArgs{}) {
  constexpr int T = MT * N / 8, NB = N / 8, MF = PairColsCfg<N>::MT;
  static_assert(MF % MT == 0, "launch tiles inside fold tiles");
  static_assert(T == MT * NB, "one column position per thread in the first stage");
  extern __shared__ __align__(16) unsigned char dyn[];
  __shared__ float2 tw[N];
  float2* lds = (float2*)dyn;             // transform tile (after the tables are consumed)
  float2* t8 = (float2*)dyn;              // [2][L][NB]
  float2* t1 = t8 + 2 * L * NB;           // [2][L][8]
  float2* py = t1 + 2 * L * 8;            // [2][L][MT]
  uint32_t* U = (uint32_t*)(py + 2 * L * MT);  // [2][L] column phases
  const int tid = threadIdx.x;
  const int pr = blockIdx.y;
  const int tile = blockIdx.x, c0 = tile * MT;
  const bool two = 2 * pr + 1 < F;
  constexpr int S = pair_tab_stride<N>();
  // first-stage positions of this thread: column c0 + m, rows j + r * NB (r < 8)
  const int m = tid % MT, j = tid / MT;
  const int k1 = c0 + m;
  const size_t plane = (size_t)N * N;
  // fold tile c0 / MF, column c0 % MF + m of it
  const float4* fp = (const float4*)fold + ((size_t)(c0 / MF) * NB + j) * MF + c0 % MF + m;  // + h * N * NB per chunk
  // PF > 1: the fold values of layers l0 .. l0 + PF - 1 (clamped to the
  // last layer: unconditional loads, the clamped ones never summed)
  constexpr int PB = PF > 1 ? PF : 1;
  float4 gb[PB][4];
  auto load_block = [&](int l0) {
#pragma unroll
    for (int lb = 0; lb < PB; ++lb)
#pragma unroll
      for (int h = 0; h < 4; ++h) gb[lb][h] = fp[(size_t)min(l0 + lb, L - 1) * (plane / 2) + (size_t)h * N * NB];
  };
  // TAB: the phases, the row entries' side values and table values
  // rt[m][l][f] (T8, T1 of entry c0 + m)
  const bool nyq_col = c0 <= N / 2 && N / 2 < c0 + MT;
  __shared__ uint32_t u[TAB ? 2 : 1][FOLD_LMAX][2];
  __shared__ float2 srow[TAB ? MT : 1][FOLD_LMAX];
  __shared__ float2 rt[TAB ? MT : 1][FOLD_LMAX][2][2];
  // TAB: the displacements and the row entries' side values load first,
  // before any tile's fold traffic is queued
  double dv = 0.0;
  float2 sv = make_float2(0.f, 0.f);
  uint32_t gv = FOLD_PLAIN;  // the fold marker (a reused fold must be the plain one)
  if constexpr (TAB) {
    if (ta.guard) gv = *ta.guard;
    if (tid < 4 * L) {
      const int f = tid / (2 * L), t = 2 * pr + f;
      dv = ta.disp[(size_t)min(t, F - 1) * 2 * L + tid % (2 * L)];
    }
    if (tid < MT * L) sv = ta.side[(size_t)(tid / MT) * 2 * N + c0 + tid % MT];
  }
  const bool poison = TAB && gv != FOLD_PLAIN;
  const float2 nan2 = make_float2(__builtin_nanf(""), __builtin_nanf(""));
  // PF > 1: the twiddles and (without TAB) this pair's tables load into
  // registers ahead of the fold (the load counter drains in order, so a
  // load issued after the fold's would wait for all of it)
  constexpr int TWR = PF > 1 ? (N + T - 1) / T : 1, PTR = PF > 1 && !TAB ? (2 * FOLD_LMAX * S + T - 1) / T : 1;
  float2 twv[TWR], ptv[PTR];
  const float2* tp = TAB ? nullptr : ptab + (size_t)pr * 2 * L * S;
  if constexpr (PF > 1) {
#pragma unroll
    for (int r = 0; r < TWR; ++r) {
      const int t = min(tid + r * T, N - 1);
      const TwEntry e = kTw.e[t * (2048 / N)];
      twv[r] = make_float2(e.c, e.s);
    }
    if constexpr (!TAB) {
#pragma unroll
      for (int r = 0; r < PTR; ++r) ptv[r] = tp[min(tid + r * T, 2 * L * S - 1)];
    }
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (!TAB) load_block(0);
  }
  if constexpr (TAB) {
    // fixed-point phases of the pair's frames (a frame past the end is absent)
    if (tid < 4 * L) {
      const int f = tid / (2 * L), e = tid % (2 * L), t = 2 * pr + f;
      u[f][e >> 1][e & 1] = t < F ? cycles_fixed(dv * ta.inv_period) : 0u;
    }
    if (tid < MT * L) srow[tid % MT][tid / MT] = sv;
    __syncthreads();
    // the fold streams in once the displacements and side values are here:
    // issued earlier, it queues those few loads behind every tile's fold
    // traffic (about 4 us of the table phase at n = 512), and the tables
    // below are built while it arrives
    __builtin_amdgcn_sched_barrier(0);
    load_block(0);
    // RU entries per thread at a time, as independent chains (one wave per
    // SIMD hides no latency otherwise); an index past the end recomputes the
    // last entry and is not stored
    constexpr int RU = 3;
    const int ntab = 2 * L * S;
    for (int q0 = tid; q0 < ntab; q0 += RU * T) {
      float2 v[RU];
#pragma unroll
      for (int h = 0; h < RU; ++h) {
        const int q = min(q0 + h * T, ntab - 1);
        const int fl = q / S, e = q % S;
        const int fb = fl / L, l = fl % L;
        const uint32_t ut = u[fb][l][0], uo = u[fb][l][1];  // column-first: the table axis is x
        v[h] = phasor_fixed(e < NB ? (uint32_t)(8 * e - N / 2) * ut : (uint32_t)(e - NB) * ut);
        if (e >= NB + 8) v[h] = make_float2(__uint_as_float(uo), 0.f);
        if (poison && e < NB + 8) v[h] = nan2;
      }
#pragma unroll
      for (int h = 0; h < RU; ++h) {
        const int q = q0 + h * T;
        if (q < ntab) {
          const int fl = q / S, e = q % S;
          if (e < NB)
            t8[fl * NB + e] = v[h];
          else if (e < NB + 8)
            t1[fl * 8 + e - NB] = v[h];
          else
            U[fl] = __float_as_uint(v[h].x);
        }
      }
    }
  } else {
    auto put = [&](int q, float2 v) {
      const int fl = q / S, e = q % S;
      if (e < NB)
        t8[fl * NB + e] = v;
      else if (e < NB + 8)
        t1[fl * 8 + e - NB] = v;
      else
        U[fl] = __float_as_uint(v.x);
    };
    if constexpr (PF > 1) {
#pragma unroll
      for (int r = 0; r < PTR; ++r)
        if (tid + r * T < 2 * L * S) put(tid + r * T, ptv[r]);
    } else {
      for (int q = tid; q < 2 * L * S; q += T) put(q, tp[q]);
    }
  }
  if constexpr (PF > 1) {
#pragma unroll
    for (int r = 0; r < TWR; ++r)
      if (tid + r * T < N) tw[tid + r * T] = twv[r];
  } else {
    init_twiddles<N, 1>(tw, tid, T);
  }
  __syncthreads();
  // and (TAB) the row entries' table values, one phasor each
  for (int q = tid; q < (TAB ? 6 : 2) * L * MT; q += T) {
    if (q < 2 * L * MT) {
      py[q] = phasor_fixed((uint32_t)signed_freq(c0 + q % MT, N) * U[q / MT]);
    } else if constexpr (TAB) {
      const int r = q - 2 * L * MT;
      const int mm = r / (4 * L), l = (r / 4) % L, f = (r >> 1) & 1, w = r & 1;
      const int id = nyq_index<N>(c0 + mm);
      rt[mm][l][f][w] = nyq_tab_value<N>(u, f, 1, l, w == 0 ? id >> 3 : NB + (id & 7), 2 * pr + 1 < F);
    }
  }
  __syncthreads();
  float2 z[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) z[r] = make_float2(0.f, 0.f);
  // For N >= 64 the eight rows share (k0 + N/2) mod 8, so the column factor
  // times t1 is one product per layer and frame; for N = 32 (row step 4)
  // it alternates between two values.
  constexpr bool BSAME = NB % 8 == 0;
  const int b0 = (j + N / 2) & 7;
  // UNI: uniform step, the fold carries (1 + i D) and Q = P_1
  auto layer_sum = [&](auto uni_tag) {
    constexpr bool UNI = decltype(uni_tag)::value;
    const bool two2 = !UNI && two;
    for (int l0 = 0; l0 < L; l0 += PB) {
      if constexpr (PF > 1) {
        if (l0 > 0) load_block(l0);
      }
#pragma unroll
      for (int lb = 0; lb < PB; ++lb) {
      const int l = l0 + lb;
      if (PF > 1 && l >= L) break;  // workgroup-uniform
      const float2 pa = py[l * MT + m], pb = py[(L + l) * MT + m];
      const float2 ca0 = cmul(pa, t1[l * 8 + b0]);
      const float2 cb0 = two2 ? cmul(pb, t1[(L + l) * 8 + b0]) : make_float2(0.f, 0.f);
      float2 g[8];
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const float4 v = PF > 1 ? gb[lb][h] : fp[(size_t)l * (plane / 2) + (size_t)h * N * NB];
        g[2 * h] = make_float2(v.x, v.y);
        g[2 * h + 1] = make_float2(v.z, v.w);
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int qq = (j + r * NB + N / 2) & (N - 1);
        float2 ca = ca0, cb = cb0;
        if (!BSAME) {
          ca = cmul(pa, t1[l * 8 + (qq & 7)]);
          cb = two2 ? cmul(pb, t1[(L + l) * 8 + (qq & 7)]) : make_float2(0.f, 0.f);
        }
        const int a = qq >> 3;
        const float2 ta = t8[l * NB + a];
        float qx = ca.x * ta.x;
        qx = fmaf(-ca.y, ta.y, qx);
        float qy = ca.x * ta.y;
        qy = fmaf(ca.y, ta.x, qy);
        if (!UNI) {
          const float2 tb = t8[(L + l) * NB + a];
          qx = fmaf(-cb.x, tb.y, qx);
          qx = fmaf(-cb.y, tb.x, qx);
          qy = fmaf(cb.x, tb.x, qy);
          qy = fmaf(-cb.y, tb.y, qy);
        }
        z[r].x = fmaf(g[r].x, qx, z[r].x);
        z[r].x = fmaf(-g[r].y, qy, z[r].x);
        z[r].y = fmaf(g[r].x, qy, z[r].y);
        z[r].y = fmaf(g[r].y, qx, z[r].y);
      }
      }
    }
  };
  if (step_uniform(step))
    layer_sum(std::true_type{});
  else
    layer_sum(std::false_type{});
  // Nyquist row / column terms of this pair: corr[k1] on the Nyquist row
  // (k0 = N/2, every k1), corr[N + k0] on the Nyquist column (k1 = N/2,
  // k0 != N/2); TAB: the row entry's sum here, the column's from cl
  const float2* cp = TAB ? nullptr : corr + (size_t)pr * 2 * N;
  bool col_ok = true;
  if constexpr (TAB) {
    if (nyq_col) {  // workgroup-uniform: this pair's producer has written its sums
      __shared__ int s_ok;
      if (tid < 64) {
        int good = 1;
        for (unsigned it = 0;; ++it) {
          const unsigned long long f = __hip_atomic_load(ta.nflag + pr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (f == ta.token) break;
          if (it >= ta.spin_limit) {
            good = 0;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the sums are read fresh after this
        if (tid == 0) s_ok = good;
      }
      __syncthreads();
      col_ok = s_ok != 0;
      if (tid == 0) __hip_atomic_store(ta.nflag + pr, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  constexpr int RN = (N / 2) / NB;  // element r holding row N/2 (for j == (N/2) % NB)
  if (j == (N / 2) % NB) {
    float2 c;
    if constexpr (TAB) {
      float2 acc = make_float2(0.f, 0.f);  // k_frame_pair_tables' terms, layers in order
      for (int l = 0; l < L; ++l)
        nyq_term(acc, rt[m][l][0][0], rt[m][l][0][1], rt[m][l][1][0], rt[m][l][1][1], srow[m][l]);
      c = poison ? nan2 : acc;
    } else {
      c = cp[k1];
    }
    z[RN].x += c.x;
    z[RN].y += c.y;
  }
  if (k1 == N / 2) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int k0 = j + r * NB;
      if (k0 == N / 2) continue;
      const float2 c = TAB ? (col_ok ? ta.ncol[(size_t)pr * N + k0] : nan2) : cp[N + k0];
      z[r].x += c.x;
      z[r].y += c.y;
    }
  }
  __syncthreads();  // every table read precedes the first stage's LDS writes
  int cnt = 0;
  auto load = [&](int, int) { return z[cnt++]; };
  float2* out = work + (size_t)pr * plane + c0;
  auto store = [&](int mm, int i, float2 v) { store_nt(out + (size_t)i * N + mm, v); };
  wg_fft<N, T, MT, true, 1>(lds, tw, tid, load, store);
}

template <int N>
__global__ __launch_bounds__(PairColsCfg<N>::T) __attribute__((amdgpu_waves_per_eu(PairColsCfg<N>::W))) void k_frame_pair_cols(const float2* __restrict__ fold,
                                                                      const float2* __restrict__ corr, int L,
                                                                      const float2* __restrict__ ptab, int F,
                                                                      const uint32_t* __restrict__ step,
                                                                      float2* __restrict__ work) {
  pair_cols_body<N, PairColsCfg<N>::MT, 1>(fold, corr, L, ptab, F, step, work);
}

// Calls of at most FEW_PAIRS pairs (n >= 128).
template <int N>
__global__ __launch_bounds__(FewColsCfg<N>::T) void k_frame_pair_cols_few(const float2* __restrict__ fold,
                                                                        const float2* __restrict__ corr, int L,
                                                                        const float2* __restrict__ ptab, int F,
                                                                        const uint32_t* __restrict__ step,
                                                                        float2* __restrict__ work) {
  pair_cols_body<N, FewColsCfg<N>::MT, FewColsCfg<N>::PF>(fold, corr, L, ptab, F, step, work);
}

// The same for a reused fold with the call's tables built in the same
// launch: blocks x < N / MT are pass A' tiles that build their own pair
// tables (TAB); blocks beyond build the pair's subharmonic tables for pass
// B' (tb per frame, k_frame_prep's arithmetic), and the last one the pair's
// Nyquist column sums for the tile that holds column N/2, on other CUs at
// the same time.  One launch instead of two per call.  Only that tile waits,
// and only for a workgroup of its own launch: every other workgroup runs to
// its end without waiting, so the producer finds a free slot whatever the
// dispatch order and occupancy.
template <int N>
__global__ __launch_bounds__(FewColsCfg<N>::T) void k_frame_pair_cols_tab(const float2* __restrict__ fold, int L, int F,
                                                                        float2* __restrict__ work, PairTabArgs ta,
                                                                        int tb) {
  constexpr int T = FewColsCfg<N>::T, NT = N / FewColsCfg<N>::MT;
  if ((int)blockIdx.x == NT + 2 * tb) {  // workgroup-uniform: the pair's Nyquist column sums
    nyq_col_producer<N, T>(L, F, ta);
    return;
  }
  if ((int)blockIdx.x >= NT) {  // workgroup-uniform
    const int i = blockIdx.x - NT, f = 2 * blockIdx.y + i / tb;
    if (f >= F || ta.P == 0) return;
    extern __shared__ __align__(16) unsigned char dyn[];
    __shared__ double s_mean;
    double2* Q = (double2*)dyn;
    double2* SB = Q + 2 * (size_t)ta.P * L;
    frame_sh_blocks(Q, SB, ta.disp, ta.coeffs, f, L, ta.P, ta.inv_period, threadIdx.x, T, true);
    sub_fill_set(SB, ta.P, N, f, ta.tab, f == 0, &s_mean, threadIdx.x, T, i % tb, tb, nullptr);
    return;
  }
  pair_cols_body<N, FewColsCfg<N>::MT, FewColsCfg<N>::PF_TAB, true>(fold, nullptr, L, nullptr, F, nullptr, work, ta);
}

// Pass B' of the column-first engine: the row transform of pair b (a plan
// whose last stage is radix 8, so a thread's eight outputs lie in one row x
// at columns y0 + d * N/8), then the epilogue: real part to frame 2b,
// imaginary part to frame 2b + 1, plus the subharmonic field
//   low(x, y) = K'[s][x] + sum_p Re(U[p][y] H'[s][p][x])
// from tables of the transposed field (k_frame_prep with tr), i.e. the
// column epilogue's ColSub with rows and columns exchanged; then crop and
// scale, whole rows per wave.
template <int N>
// Seven 256-thread workgroups per CU at n = 512: 22 KiB of LDS each, and at
// most 72 VGPRs (73 uncapped, no spills at 70).
__global__ __launch_bounds__(RowsCfg<N>::T) __attribute__((amdgpu_waves_per_eu(7))) void k_frame_rows_epi(GenArgs a) {
  constexpr int T = RowsCfg<N>::T, M = RowsCfg<N>::M;
  static_assert(Plan<N, true>::radix(Plan<N, true>::NST - 1) == 8, "last stage radix 8");
  using Sub = ColSub<N, M, 2, T>;
  // the subharmonic staging block reuses the transform tile once the last
  // stage has read it: a smaller footprint per workgroup (6 instead of 5 per
  // CU at n = 512) for one more barrier
  constexpr size_t TILE_B = (size_t)Layout<N, T, M, false>::SIZE * sizeof(float2);
  __shared__ __align__(16) unsigned char smem[TILE_B > sizeof(Sub) ? TILE_B : sizeof(Sub)];
  __shared__ float2 tw[N];
  float2* lds = reinterpret_cast<float2*>(smem);
  Sub& sub = *reinterpret_cast<Sub*>(smem);
  const int tid = threadIdx.x;
  const int b = blockIdx.y;
  const int r0 = blockIdx.x * M;
  if (r0 + M <= a.off || r0 >= a.off + a.n_out) return;  // rows entirely outside the crop
  const int P = a.levels;
  const int s0 = b * a.s_mul, s1 = s0 + a.s_add;
  const bool has1 = s1 < a.count;
  const int sets[2] = {s0, has1 ? s1 : -1};
  init_twiddles<N, 1>(tw, tid, T);
  __syncthreads();
  typename Sub::Regs pre;
  if (P > 0) pre = Sub::prefetch(a.tab, sets, r0, P, tid);
  const size_t plane = (size_t)N * N;
  const float2* src = a.work + (size_t)b * plane + (size_t)r0 * N;
  float v0[8], v1[8];
  int cnt = 0;
  // the intermediate is read once: non-temporal, so the tables stay in L2
  auto load = [&](int m, int i) {
    typedef float v2f __attribute__((ext_vector_type(2)));
    const v2f v = __builtin_nontemporal_load((const v2f*)(src + (size_t)m * N + i));
    return make_float2(v.x, v.y);
  };
  auto store = [&](int, int, float2 v) {
    v0[cnt] = v.x;
    v1[cnt] = v.y;
    ++cnt;
  };
  wg_fft<N, T, M, false, 1, true>(lds, tw, tid, load, store);
  if (P > 0) {
    __syncthreads();  // every lane has read its last-stage inputs from the tile
    sub.commit(pre, P, tid);
  }
  __syncthreads();
  const int m = tid / (N / 8), y0 = tid % (N / 8);
  const int x = r0 + m - a.off;
  if ((unsigned)x >= (unsigned)a.n_out) return;  // no barrier follows
  if (P > 0) sub.template add_low2<8>(m, y0, P, v0, v1, has1);
  const size_t oplane = (size_t)a.n_out * a.n_out;
  float* o0 = a.out + (size_t)s0 * oplane + (size_t)x * a.n_out;
  float* o1 = a.out + (size_t)s1 * oplane + (size_t)x * a.n_out;
  const float f = a.scale;
  if (a.off == 0 && a.n_out == N) {  // whole grid (uniform): constant offsets, no tests
    float* p0 = o0 + y0;
#pragma unroll
    for (int c = 0; c < 8; ++c) store_nt(p0 + c * (N / 8), v0[c] * f);
    if (has1) {
      float* p1 = o1 + y0;
#pragma unroll
      for (int c = 0; c < 8; ++c) store_nt(p1 + c * (N / 8), v1[c] * f);
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int y = y0 + c * (N / 8) - a.off;
    if ((unsigned)y >= (unsigned)a.n_out) continue;
    store_nt(o0 + y, v0[c] * f);
    if (has1) store_nt(o1 + y, v1[c] * f);
  }
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

// Fused column epilogues keep the per-column subharmonic terms in registers
// (up to PREG levels), staged through a small LDS block (ColSub) that shares
// the transform tile.
template <int N>
static bool cols_fused_ok(int levels) {
  return levels <= PREG && gen_cols_smem<N, ColsCfg<N>::T, ColsCfg<N>::M>() + (size_t)N * 8 <= kLdsMax;
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

// Column pass of generate and of frame pass B, in the tuned launch shape
// (the accumulating form keeps the default shape).
template <int N>
static int gen_cols_pow2(const GenArgs& a, hipStream_t st) {
  constexpr int MC = ColsCfg<N>::M;
  if (a.accumulate) {
    k_gen_cols_acc<N><<<dim3(N / MC, a.n_fft), ColsCfg<N>::T, 0, st>>>(a);
  } else {
    const int vc = tune_choice(PT_TUNE_GEN_COLS, N);
    if (vc) return launch_gen_cols_variant(N, vc, a, st);
    k_gen_cols_set<N><<<dim3(N / MC, a.n_fft), ColsCfg<N>::T, 0, st>>>(a);
  }
  PT_LAUNCH_CHECK();
  return PT_OK;
}

template <int N>
static int generate_pow2(const float* nre, ptrdiff_t im_off, size_t gstride, int G, const float* amp,
                         const float* sh_noise, const double* sh_amp, GenArgs a, hipStream_t st) {
  const int vr = tune_choice(PT_TUNE_GEN_ROWS, N);
  if (vr) {
    const int rc = launch_gen_rows_variant(N, vr, nre, im_off, gstride, G, amp, (float2*)a.work, a.n_fft, st);
    if (rc) return rc;
  } else {
    k_gen_rows<N><<<dim3(N / RowsCfg<N>::M, a.n_fft), RowsCfg<N>::T, 0, st>>>(nre, im_off, gstride, G, amp,
                                                                              (float2*)a.work);
    PT_LAUNCH_CHECK();
  }
  int rc = launch_gen_prep(sh_noise, sh_amp, a.count, a.levels, N, a.tab, st, G);
  if (rc) return rc;
  return gen_cols_pow2<N>(a, st);
}


// Workspace of the fused frame path (c64 elements): the paired intermediate,
// the folded stack, its Nyquist side table, then the subharmonic tables.
struct FramesLayout {
  size_t fold, side, ptab, corr, step, tab, total;
};

// A value no earlier launch of the process used (the producer/consumer
// flags of k_frame_pair_cols_tab): a clock-seeded odd start, then +1.
static unsigned long long next_launch_token() {
  static std::atomic<unsigned long long> next{
      ((unsigned long long)std::chrono::steady_clock::now().time_since_epoch().count() * 0x9E3779B97F4A7C15ull) | 1ull};
  unsigned long long t = next.fetch_add(1);
  return t ? t : next.fetch_add(1);
}
static inline size_t r32(size_t v) { return (v + 31) & ~(size_t)31; }
static FramesLayout frames_layout(int n, int F, int L, int P) {
  FramesLayout w;
  const size_t plane = (size_t)n * n;
  w.fold = r32((size_t)((F + 1) / 2) * plane);
  w.side = w.fold + r32((size_t)L * plane);
  w.ptab = w.side + r32((size_t)L * 2 * n);
  w.corr = w.ptab + r32((size_t)((F + 1) / 2) * 2 * L * (n / 8 + 9));
  w.step = w.corr + r32((size_t)((F + 1) / 2) * 2 * n);
  w.tab = w.step + r32((size_t)(STEP_WORDS + 2) / 2);  // uint32 words of k_frame_step, then the fold marker
  w.total = w.tab + sub_table_c64(P, n, F);
  return w;
}

// Stages of pass A's call: the fold of the stack (FOLD), the pair tables and
// pass A (PAIRS), or both.
enum { STAGE_FOLD = 1, STAGE_PAIRS = 2, STAGE_ALL = 3 };

// Subharmonic tables built in the same launch as the pair tables
// (k_frame_tables_prep), for pt_spectral_frames_folded's calls of a few
// frames.
struct MergedPrep {
  const double2* coeffs;
  int P;
  SubTables tab;
};

// The pair tables of a reused fold (guard = the fold marker), alone or with
// the subharmonic tables of the same frames in one launch.
template <int N, bool TR>
static int launch_pair_tables(const double* disp, const float2* side, int L, int F, double inv_period,
                              const uint32_t* step, float2* ptab, float2* corr, const uint32_t* guard,
                              const MergedPrep* mp, hipStream_t st) {
  const int pairs = (F + 1) / 2, ta = table_split(pairs);
  if (mp && mp->P > 0) {
    const int tb = table_split(F);
    const size_t dyn = (size_t)(2 * mp->P * L + 9 * mp->P) * sizeof(double2);
    if (dyn > 60 * 1024) return fail(PT_ERR_UNSUPPORTED, "frames: %d layers x %d levels exceed LDS", L, mp->P);
    k_frame_tables_prep<N, TR><<<pairs * ta + F * tb, 256, dyn, st>>>(disp, side, L, F, inv_period, ptab, corr,
                                                                      guard, ta, mp->coeffs, mp->P, mp->tab, tb);
  } else {
    k_frame_pair_tables<N, TR><<<dim3(pairs, ta), 256, 0, st>>>(disp, side, L, F, inv_period, step, ptab, corr,
                                                                guard);
  }
  PT_LAUNCH_CHECK();
  return PT_OK;
}

template <int N>
static int frames_rows_pow2(const float2* spectra, int L, const double* disp, int F, double inv_period,
                            float2* work, hipStream_t st, int stages = STAGE_ALL, const MergedPrep* mp = nullptr) {
  constexpr int MR = PairCfg<N>::M;
  const size_t dyn_r = pair_rows_lds<N>(L);
  if (dyn_r + (size_t)Layout<N, PairCfg<N>::T, MR, false>::SIZE * 8 + (size_t)N * 8 > kLdsMax)
    return fail(PT_ERR_UNSUPPORTED, "frames: %d layers exceed the LDS phasor tables at n=%d", L, N);
  int rc = set_dyn_lds(k_frame_pair_rows<N>, dyn_r);
  if (rc) return rc;
  const FramesLayout w = frames_layout(N, F, L, 0);
  float2* fold = work + w.fold;
  float2* side = work + w.side;
  float2* ptab = work + w.ptab;
  float2* corr = work + w.corr;
  // uniform-step fold only where one call folds and pairs (the step is read
  // from this call's displacements)
  uint32_t* step = stages == STAGE_ALL ? (uint32_t*)(work + w.step) : nullptr;
  uint32_t* mark = (uint32_t*)(work + w.step) + STEP_WORDS;
  if (step) {
    k_frame_step<<<STEP_BLOCKS, 1024, 0, st>>>(disp, L, F, inv_period, step);
    PT_LAUNCH_CHECK();
  }
  if (stages & STAGE_FOLD) {
    k_frame_fold<<<dim3((N * N + 255) / 256, L), 256, 0, st>>>(spectra, N, step, fold, side, mark);
    PT_LAUNCH_CHECK();
  }
  if (!(stages & STAGE_PAIRS)) return PT_OK;
  rc = launch_pair_tables<N, false>(disp, side, L, F, inv_period, step, ptab, corr,
                                    stages == STAGE_PAIRS ? mark : nullptr, mp, st);
  if (rc) return rc;
  k_frame_pair_rows<N><<<dim3(N / MR, (F + 1) / 2), PairCfg<N>::T, dyn_r, st>>>(fold, corr, L, ptab, F, step,
                                                                                 work);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

template <int N>
static int frames_cols_pow2(GenArgs a, hipStream_t st) {
  return gen_cols_pow2<N>(a, st);
}

// Column-first order: fold into column tiles, transposed pair tables, pass A'.
template <int N>
static int frames_cf_rows_pow2(const float2* spectra, int L, const double* disp, int F, double inv_period,
                               float2* work, hipStream_t st, int stages = STAGE_ALL,
                               const MergedPrep* mp = nullptr) {
  constexpr int MT = PairColsCfg<N>::MT;
  const size_t dyn = pair_cols_lds<N>(L);
  if (dyn + (size_t)N * 8 > kLdsMax)
    return fail(PT_ERR_UNSUPPORTED, "frames: %d layers exceed the LDS phasor tables at n=%d", L, N);
  int rc = set_dyn_lds(k_frame_pair_cols<N>, dyn);
  if (rc) return rc;
  const FramesLayout w = frames_layout(N, F, L, 0);
  float2* fold = work + w.fold;
  float2* side = work + w.side;
  float2* ptab = work + w.ptab;
  float2* corr = work + w.corr;
  uint32_t* step = stages == STAGE_ALL ? (uint32_t*)(work + w.step) : nullptr;  // as in frames_rows_pow2
  uint32_t* mark = (uint32_t*)(work + w.step) + STEP_WORDS;
  if (step) {
    k_frame_step<<<STEP_BLOCKS, 1024, 0, st>>>(disp, L, F, inv_period, step);
    PT_LAUNCH_CHECK();
  }
  if (stages & STAGE_FOLD) {
    k_frame_fold_tiles<<<dim3((N * N + 255) / 256, L), 256, 0, st>>>(spectra, N, MT, step, fold, side, mark);
    PT_LAUNCH_CHECK();
  }
  if (!(stages & STAGE_PAIRS)) return PT_OK;
  if constexpr (N >= 128 && N <= 1024) {
    // a reused fold and a few pairs: the tables are built inside pass A'
    // (n = 2048 would spill at 1024 threads)
    if (stages == STAGE_PAIRS && (F + 1) / 2 <= FEW_PAIRS) {
      constexpr int MTF = FewColsCfg<N>::MT;
      PairTabArgs ta;
      ta.disp = disp;
      ta.side = side;
      ta.guard = mark;
      ta.inv_period = inv_period;
      ta.coeffs = mp ? mp->coeffs : nullptr;
      ta.P = mp ? mp->P : 0;
      ta.tab = mp ? mp->tab : SubTables{};
      const size_t prep = (size_t)(2 * ta.P * L + 9 * ta.P) * sizeof(double2);
      const size_t dyn_t = pair_cols_lds<N, MTF>(L) > prep ? pair_cols_lds<N, MTF>(L) : prep;
      rc = set_dyn_lds(k_frame_pair_cols_tab<N>, dyn_t);
      if (rc) return rc;
      const int tb = ta.P > 0 ? 16 : 0;  // subharmonic table blocks per frame
      // the Nyquist column sums go where the unfolded call keeps its
      // correction terms, and the flags after the fold marker
      static_assert((STEP_WORDS + 2) / 2 + FEW_PAIRS <= 32, "flags inside the step words' padding");
      ta.ncol = corr;
      ta.nflag = reinterpret_cast<unsigned long long*>(work + w.step + (STEP_WORDS + 2) / 2);
      ta.token = next_launch_token();
      ta.spin_limit = 1u << 20;
      k_frame_pair_cols_tab<N><<<dim3(N / MTF + 2 * tb + 1, (F + 1) / 2), FewColsCfg<N>::T, dyn_t, st>>>(
          fold, L, F, work, ta, tb);
      PT_LAUNCH_CHECK();
      return PT_OK;
    }
  }
  rc = launch_pair_tables<N, true>(disp, side, L, F, inv_period, step, ptab, corr,
                                   stages == STAGE_PAIRS ? mark : nullptr, mp, st);
  if (rc) return rc;
  if constexpr (N >= 128) {
    if ((F + 1) / 2 <= FEW_PAIRS) {
      constexpr int MTF = FewColsCfg<N>::MT;
      const size_t dyn_f = pair_cols_lds<N, MTF>(L);
      rc = set_dyn_lds(k_frame_pair_cols_few<N>, dyn_f);
      if (rc) return rc;
      k_frame_pair_cols_few<N><<<dim3(N / MTF, (F + 1) / 2), FewColsCfg<N>::T, dyn_f, st>>>(fold, corr, L, ptab, F,
                                                                                          step, work);
      PT_LAUNCH_CHECK();
      return PT_OK;
    }
  }
  k_frame_pair_cols<N><<<dim3(N / MT, (F + 1) / 2), PairColsCfg<N>::T, dyn, st>>>(
      fold, corr, L, ptab, F, step, work);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

// Column-first order, pass B' (the tables must come from a transposed prep).
template <int N>
static int frames_cf_cols_pow2(const GenArgs& a, hipStream_t st) {
  k_frame_rows_epi<N><<<dim3(N / RowsCfg<N>::M, a.n_fft), RowsCfg<N>::T, 0, st>>>(a);
  PT_LAUNCH_CHECK();
  return PT_OK;
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

static bool fused_ok(int n, int levels) {
#define CALL(N) return cols_fused_ok<N>(levels);
  PT_POW2_SWITCH(n, CALL)
#undef CALL
  return false;
}

int launch_gen_cols(int n, const GenArgs& a, hipStream_t st) {
#define CALL(N) return gen_cols_pow2<N>(a, st);
  PT_POW2_SWITCH(n, CALL)
#undef CALL
  return fail(PT_ERR_UNSUPPORTED, "gen_cols: n=%d", n);
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

static bool frames_fused(int n, int layers, int levels) {
  return is_pow2_fast(n) && fused_ok(n, levels) && layers <= FOLD_LMAX;
}

static size_t frames_generic_elems(int n, int F, int P) { return 2 * (size_t)F * n * n + sub_table_c64(P, n, F); }

extern "C" long long pt_frames_work_elems(int n, int frames, int layers, int levels) {
  if (n < 1 || frames < 0 || layers < 1 || levels < 0) return -1;
  return (long long)(frames_fused(n, layers, levels) ? frames_layout(n, frames, layers, levels).total
                                                     : frames_generic_elems(n, frames, levels));
}

extern "C" long long pt_work_elems(int kind, int n, int sets, int levels) {
  // kind 0: generate (sets = screens), 1: spectral frames (sets = frames, any layer count)
  if (n < 1 || sets < 0 || levels < 0) return -1;
  const bool fused = is_pow2_fast(n) && fused_ok(n, levels);
  if (kind == 1) {
    const size_t g = frames_generic_elems(n, sets, levels);
    if (!fused) return (long long)g;
    const size_t f = frames_layout(n, sets, FOLD_LMAX, levels).total;
    return (long long)(f > g ? f : g);
  }
  const size_t nfft = (size_t)(sets + 1) / 2;
  return (long long)((fused ? 1 : 2) * nfft * n * n + sub_table_c64(levels, n, sets));
}

extern "C" int pt_fft2(const void* in, void* out, void* work, int batch, int n0, int n1, int sign, void* stream) {
  if (!in || !out) return fail(PT_ERR_ARG, "pt_fft2: NULL array");
  if (batch < 0 || n0 < 1 || n1 < 1 || (sign != 1 && sign != -1)) return fail(PT_ERR_ARG, "pt_fft2: bad shape/sign");
  if (batch == 0) return PT_OK;
  hipStream_t st = (hipStream_t)stream;
  if (n0 == n1 && is_pow2_fast(n0)) {
    if (batch > 65535) return fail(PT_ERR_ARG, "pt_fft2: batch > 65535");
#define CALL(N) return fft2_pow2<N>((const float2*)in, (float2*)out, batch, sign, st);
    PT_POW2_SWITCH(n0, CALL)
#undef CALL
  }
  if (n0 != n1) return fail(PT_ERR_UNSUPPORTED, "pt_fft2: only square grids (got %dx%d)", n0, n1);
  return pt_fft2_generic(in, out, work, batch, n0, sign, stream);
}

// Shared body of pt_generate (one group, separate real/imaginary planes) and
// pt_generate_groups (G groups, planes at nre + g * gstride [+ im_off]).
static int generate_impl(const char* who, const float* noise_re, ptrdiff_t im_off, size_t gstride, int G,
                         const float* amplitude, const float* sh_noise, const double* sh_amp, int levels,
                         int count, int n, int crop_offset, int n_out, float scale, const float* screen_scale,
                         int accumulate, float* out, void* work, void* stream) {
  if (count < 1) return fail(PT_ERR_ARG, "%s: count must be >= 1", who);
  if (G < 1 || G > 64) return fail(PT_ERR_ARG, "%s: 1..64 groups", who);
  if (!noise_re || !amplitude || !out || !work) return fail(PT_ERR_ARG, "%s: NULL array", who);
  if (levels < 0 || (levels > 0 && (!sh_noise || !sh_amp))) return fail(PT_ERR_ARG, "%s: bad subharmonics", who);
  if (levels > PMAX) return fail(PT_ERR_UNSUPPORTED, "%s: more than %d subharmonic levels", who, PMAX);
  if (crop_offset < 0 || n_out < 1 || crop_offset + n_out > n) return fail(PT_ERR_ARG, "%s: crop outside grid", who);
  const int n_fft = (count + 1) / 2;
  if (n_fft > 65535) return fail(PT_ERR_ARG, "%s: count > 131070 per call", who);
  hipStream_t st = (hipStream_t)stream;
  if (is_pow2_fast(n) && fused_ok(n, levels)) {
    GenArgs a;
    a.work = (const float2*)work;
    a.out = out;
    a.screen_scale = screen_scale;
    a.tab = sub_tables((float2*)work + (size_t)n_fft * n * n, levels, n, count);
    a.count = count;
    a.n_fft = n_fft;
    a.levels = levels;
    a.off = crop_offset;
    a.n_out = n_out;
    a.accumulate = accumulate ? 1 : 0;
    a.s_mul = 1;
    a.s_add = n_fft;
    a.scale = scale;
#define CALL(N) return generate_pow2<N>(noise_re, im_off, gstride, G, amplitude, sh_noise, sh_amp, a, st);
    PT_POW2_SWITCH(n, CALL)
#undef CALL
  }
  // general sizes: one group at a time, accumulating
  const size_t plane = (size_t)n * n;
  for (int g = 0; g < G; ++g) {
    const float* re = noise_re + (size_t)g * gstride;
    int rc = pt_generate_generic(re, re + im_off, amplitude + (size_t)g * plane,
                                 levels ? sh_noise + (size_t)g * levels * 2 * count * 9 : nullptr,
                                 levels ? sh_amp + (size_t)g * levels * 9 : nullptr, levels, count, n, crop_offset,
                                 n_out, scale, screen_scale, (accumulate || g > 0) ? 1 : 0, out, work, stream);
    if (rc) return rc;
  }
  return PT_OK;
}

extern "C" int pt_generate(const float* noise_re, const float* noise_im, const float* amplitude,
                           const float* sh_noise, const double* sh_amp, int levels, int count, int n,
                           int crop_offset, int n_out, float scale, const float* screen_scale, int accumulate,
                           float* out, void* work, void* stream) {
  if (!noise_re || !noise_im) return fail(PT_ERR_ARG, "pt_generate: NULL array");
  return generate_impl("pt_generate", noise_re, noise_im - noise_re, 0, 1, amplitude, sh_noise, sh_amp, levels,
                       count, n, crop_offset, n_out, scale, screen_scale, accumulate, out, work, stream);
}

extern "C" int pt_generate_groups(const float* noise, const float* amplitude, const float* sh_noise,
                                  const double* sh_amp, int groups, int levels, int count, int n, int crop_offset,
                                  int n_out, float scale, const float* screen_scale, int accumulate, float* out,
                                  void* work, void* stream) {
  const size_t n_fft = (size_t)(count + 1) / 2;
  const size_t plane = (size_t)n * n;
  return generate_impl("pt_generate_groups", noise, (ptrdiff_t)(n_fft * plane), 2 * n_fft * plane, groups,
                       amplitude, sh_noise, sh_amp, levels, count, n, crop_offset, n_out, scale, screen_scale,
                       accumulate, out, work, stream);
}

// The two passes of pt_spectral_frames as separate calls (power-of-two n,
// fused epilogue, <= 16 layers): pass A folds the layer stack and writes the
// row-transformed Hermitian sums of frame pairs to `work`; pass B builds the
// subharmonic tables and finishes the column transforms of the pairs.
static int frames_check(int layers, int n, double pixel_scale, const double* disp, int frames, const double* sh_coeffs,
                        int levels, int crop_offset, int n_out, void* work) {
  if (!disp || !work) return fail(PT_ERR_ARG, "spectral frames: NULL array");
  if (layers < 1 || frames < 0 || levels < 0 || (levels > 0 && !sh_coeffs))
    return fail(PT_ERR_ARG, "spectral frames: bad layers/frames/levels");
  if (!(pixel_scale > 0.0)) return fail(PT_ERR_ARG, "spectral frames: pixel_scale must be > 0");
  if (crop_offset < 0 || n_out < 1 || crop_offset + n_out > n) return fail(PT_ERR_ARG, "spectral frames: crop outside grid");
  if (frames > 65535) return fail(PT_ERR_ARG, "spectral frames: frames > 65535 per call");
  if (levels > PMAX) return fail(PT_ERR_UNSUPPORTED, "spectral frames: more than %d subharmonic levels", PMAX);
  return PT_OK;
}

// Transform order of the fused frame engine for size n: 0 column-first
// (pass A' + pass B'), 1 row-first (pass A + pass B); PT_TUNE_FRAMES.
static int frames_order(int n) { return tune_choice(PT_TUNE_FRAMES, n); }

static int spectral_rows_impl(const void* spectra, int layers, int n, double pixel_scale, const double* disp,
                              int frames, void* work, int order, hipStream_t st, int stages = STAGE_ALL,
                              const MergedPrep* mp = nullptr) {
  const double inv_period = 1.0 / ((double)n * pixel_scale);
  if (order == 0) {
#define CALL(N)                                                                                              \
  return frames_cf_rows_pow2<N>((const float2*)spectra, layers, disp, frames, inv_period, (float2*)work, st, \
                                stages, mp);
    PT_POW2_SWITCH(n, CALL)
#undef CALL
  } else {
#define CALL(N)                                                                                              \
  return frames_rows_pow2<N>((const float2*)spectra, layers, disp, frames, inv_period, (float2*)work, st, \
                             stages, mp);
    PT_POW2_SWITCH(n, CALL)
#undef CALL
  }
  return fail(PT_ERR_UNSUPPORTED, "pt_spectral_rows: n=%d", n);
}

extern "C" int pt_spectral_fold(const void* spectra, int layers, int n, int frames, void* work, void* stream) {
  if (!spectra || !work) return fail(PT_ERR_ARG, "pt_spectral_fold: NULL array");
  if (layers < 1 || frames < 0) return fail(PT_ERR_ARG, "pt_spectral_fold: bad layers/frames");
  if (frames > 65535) return fail(PT_ERR_ARG, "spectral frames: frames > 65535 per call");
  if (!is_pow2_fast(n) || layers > FOLD_LMAX)
    return fail(PT_ERR_UNSUPPORTED, "pt_spectral_fold: power-of-two n 32..2048 and <= %d layers only", FOLD_LMAX);
  if (frames == 0) return PT_OK;
  return spectral_rows_impl(spectra, layers, n, 1.0, nullptr, frames, work, frames_order(n), (hipStream_t)stream,
                            STAGE_FOLD);
}

extern "C" int pt_spectral_rows_folded(int layers, int n, double pixel_scale, const double* disp, int frames,
                                       void* work, void* stream) {
  int rc = frames_check(layers, n, pixel_scale, disp, frames, nullptr, 0, 0, n, work);
  if (rc || frames == 0) return rc;
  if (!is_pow2_fast(n) || layers > FOLD_LMAX)
    return fail(PT_ERR_UNSUPPORTED, "pt_spectral_rows_folded: power-of-two n 32..2048 and <= %d layers only",
                FOLD_LMAX);
  return spectral_rows_impl(nullptr, layers, n, pixel_scale, disp, frames, work, frames_order(n), (hipStream_t)stream,
                            STAGE_PAIRS);
}

// Calls of at most this many frames build both table sets in one launch.
constexpr int FOLDED_SMALL_F = 8;

static int spectral_cols_impl(void* work, int layers, int n, double pixel_scale, const double* disp, int frames,
                              const double* sh_coeffs, int levels, int crop_offset, int n_out, float scale,
                              float* out, int order, hipStream_t st, bool prep_done = false);

extern "C" int pt_spectral_frames_folded(int layers, int n, double pixel_scale, const double* disp, int frames,
                                         const double* sh_coeffs, int levels, int crop_offset, int n_out,
                                         float scale, float* out, void* work, void* stream) {
  if (!out) return fail(PT_ERR_ARG, "pt_spectral_frames_folded: NULL out");
  int rc = frames_check(layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out, work);
  if (rc || frames == 0) return rc;
  if (!frames_fused(n, layers, levels))
    return fail(PT_ERR_UNSUPPORTED,
                "pt_spectral_frames_folded: power-of-two n 32..2048, <= %d layers and <= %d levels only",
                FOLD_LMAX, PREG);
  const int order = frames_order(n);  // read once: both passes use the same order
  hipStream_t st = (hipStream_t)stream;
  const bool merge = frames <= FOLDED_SMALL_F && levels > 0;
  MergedPrep mp;
  if (merge) {
    mp.coeffs = (const double2*)sh_coeffs;
    mp.P = levels;
    mp.tab = sub_tables((float2*)work + frames_layout(n, frames, layers, levels).tab, levels, n, frames);
  }
  rc = spectral_rows_impl(nullptr, layers, n, pixel_scale, disp, frames, work, order, st, STAGE_PAIRS,
                          merge ? &mp : nullptr);
  if (rc) return rc;
  return spectral_cols_impl(work, layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out,
                            scale, out, order, st, merge);
}

extern "C" int pt_spectral_rows(const void* spectra, int layers, int n, double pixel_scale, const double* disp,
                                int frames, void* work, void* stream) {
  if (!spectra) return fail(PT_ERR_ARG, "pt_spectral_rows: NULL spectra");
  int rc = frames_check(layers, n, pixel_scale, disp, frames, nullptr, 0, 0, n, work);
  if (rc || frames == 0) return rc;
  if (!is_pow2_fast(n) || layers > FOLD_LMAX)
    return fail(PT_ERR_UNSUPPORTED, "pt_spectral_rows: power-of-two n 32..2048 and <= %d layers only", FOLD_LMAX);
  return spectral_rows_impl(spectra, layers, n, pixel_scale, disp, frames, work, frames_order(n),
                            (hipStream_t)stream);
}
static int spectral_cols_impl(void* work, int layers, int n, double pixel_scale, const double* disp, int frames,
                              const double* sh_coeffs, int levels, int crop_offset, int n_out, float scale,
                              float* out, int order, hipStream_t st, bool prep_done) {
  GenArgs a;
  a.work = (const float2*)work;
  a.out = out;
  a.screen_scale = nullptr;
  a.tab = sub_tables((float2*)work + frames_layout(n, frames, layers, levels).tab, levels, n, frames);
  a.count = frames;
  a.n_fft = (frames + 1) / 2;
  a.levels = levels;
  a.off = crop_offset;
  a.n_out = n_out;
  a.accumulate = 0;
  a.s_mul = 2;
  a.s_add = 1;
  a.scale = scale;
  if (!prep_done) {
    const int rc = launch_frame_prep(disp, (const double2*)sh_coeffs, frames, layers, levels, n,
                                     1.0 / ((double)n * pixel_scale), a.tab, st, 0, order == 0);
    if (rc) return rc;
  }
  if (order == 0) {
#define CALL(N) return frames_cf_cols_pow2<N>(a, st);
    PT_POW2_SWITCH(n, CALL)
#undef CALL
  } else {
#define CALL(N) return frames_cols_pow2<N>(a, st);
    PT_POW2_SWITCH(n, CALL)
#undef CALL
  }
  return fail(PT_ERR_UNSUPPORTED, "pt_spectral_cols: n=%d", n);
}

extern "C" int pt_spectral_cols(void* work, int layers, int n, double pixel_scale, const double* disp, int frames,
                                const double* sh_coeffs, int levels, int crop_offset, int n_out, float scale,
                                float* out, void* stream) {
  if (!out) return fail(PT_ERR_ARG, "pt_spectral_cols: NULL out");
  int rc = frames_check(layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out, work);
  if (rc || frames == 0) return rc;
  if (!frames_fused(n, layers, levels))
    return fail(PT_ERR_UNSUPPORTED, "pt_spectral_cols: power-of-two n, <= %d layers and <= %d levels only",
                FOLD_LMAX, PREG);
  return spectral_cols_impl(work, layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out,
                            scale, out, frames_order(n), (hipStream_t)stream);
}

extern "C" int pt_spectral_frames(const void* spectra, int layers, int n, double pixel_scale, const double* disp,
                                  int frames, const double* sh_coeffs, int levels, int crop_offset, int n_out,
                                  float scale, float* out, void* work, void* stream) {
  if (!spectra || !out) return fail(PT_ERR_ARG, "pt_spectral_frames: NULL array");
  int rc = frames_check(layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out, work);
  if (rc || frames == 0) return rc;
  if (frames_fused(n, layers, levels)) {
    const int order = frames_order(n);  // read once: both passes use the same order
    hipStream_t st = (hipStream_t)stream;
    rc = spectral_rows_impl(spectra, layers, n, pixel_scale, disp, frames, work, order, st);
    if (rc) return rc;
    return spectral_cols_impl(work, layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out,
                              scale, out, order, st);
  }
  return pt_frames_generic(spectra, layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out,
                           scale, out, work, stream);
}

// ----------------------------------------------------------------------------
// boiled spectral frames
// ----------------------------------------------------------------------------
// Workspace (c64 elements): frame-pair spectra, a transform scratch of the
// same size (general n only), the phase table, the per-frame subharmonic
// coefficient sets (c128, two elements each) and the subharmonic tables.
struct BoilLayout {
  size_t z, scratch, phase, cf, tab, total;
};
static BoilLayout boil_layout(int n, int F, int L, int P) {
  BoilLayout w;
  const size_t pairs_plane = (size_t)((F + 1) / 2) * n * n;
  const bool fused = is_pow2_fast(n) && P <= PREG;
  w.z = 0;
  w.scratch = r32(pairs_plane);
  w.phase = w.scratch + (fused ? 0 : r32(pairs_plane));
  w.cf = w.phase + r32((size_t)F * L);
  w.tab = w.cf + r32((size_t)F * P * L * 9 * 2);
  w.total = w.tab + sub_table_c64(P, n, F);
  return w;
}

extern "C" long long pt_boil_work_elems(int n, int frames, int layers, int levels) {
  if (n < 2 || frames < 0 || layers < 1 || levels < 0) return -1;
  return (long long)boil_layout(n, frames, layers, levels).total;
}

extern "C" int pt_boiled_frames(void* spectra, const float* amplitude, const double* inv_tau, const double* f_ref,
                                int layers, int n, double pixel_scale, const double* disp, int frames, int steps,
                                double dt, double* sh_coeffs, const double* sh_amp, int levels, const float* noise,
                                const double* sh_noise, unsigned long long seed, unsigned long long step0,
                                int crop_offset, int n_out, float scale, float* out, void* work, void* stream) {
  if (!spectra || !amplitude || !inv_tau || !f_ref) return fail(PT_ERR_ARG, "pt_boiled_frames: NULL array");
  if (layers < 1 || n < 2 || frames < 0 || steps < frames || levels < 0)
    return fail(PT_ERR_ARG, "pt_boiled_frames: bad sizes (steps >= frames >= 0)");
  if (layers > 16) return fail(PT_ERR_UNSUPPORTED, "pt_boiled_frames: more than 16 layers");
  if (!(dt > 0.0) || !(pixel_scale > 0.0)) return fail(PT_ERR_ARG, "pt_boiled_frames: dt and pixel_scale must be > 0");
  if (levels > 0 && (!sh_coeffs || !sh_amp)) return fail(PT_ERR_ARG, "pt_boiled_frames: NULL subharmonic arrays");
  if (levels > PMAX) return fail(PT_ERR_UNSUPPORTED, "pt_boiled_frames: more than %d levels", PMAX);
  if (frames > 65535) return fail(PT_ERR_ARG, "pt_boiled_frames: frames > 65535 per call");
  if (frames > 0) {
    if (!disp || !out || !work) return fail(PT_ERR_ARG, "pt_boiled_frames: NULL disp/out/work");
    if (crop_offset < 0 || n_out < 1 || crop_offset + n_out > n) return fail(PT_ERR_ARG, "pt_boiled_frames: crop");
  }
  if (steps == 0) return PT_OK;
  hipStream_t st = (hipStream_t)stream;
  const BoilLayout w = boil_layout(n, frames, layers, levels);
  float2* base = (float2*)work;
  BoilArgs b;
  b.spectra = (float2*)spectra;
  b.amp = amplitude;
  b.inv_tau = inv_tau;
  b.f_ref = f_ref;
  b.coeffs = (double2*)sh_coeffs;
  b.sh_amp = sh_amp;
  b.coeff_frames = base ? (double2*)(base + w.cf) : nullptr;
  b.noise = noise;
  b.sh_noise = sh_noise;
  b.phase = base ? (const uint2*)(base + w.phase) : nullptr;
  b.z = base;
  b.n = n;
  b.L = layers;
  b.P = levels;
  b.frames = frames;
  b.steps = steps;
  b.dt = dt;
  b.df = 1.0 / ((double)n * pixel_scale);
  b.step0 = step0;
  b.key0 = (uint32_t)seed;
  b.key1 = (uint32_t)(seed >> 32);
  const double inv_period = b.df;
  int rc = launch_boil(b, disp, inv_period, st);
  if (rc || frames == 0) return rc;
  const SubTables tab = sub_tables(base + w.tab, levels, n, frames);
  rc = launch_frame_prep(disp, (const double2*)(base + w.cf), frames, layers, levels, n, inv_period, tab, st,
                         (size_t)levels * layers * 9);
  if (rc) return rc;
  const int pairs = (frames + 1) / 2;
  if (is_pow2_fast(n) && levels <= PREG) {
    GenArgs a;
    a.work = base;
    a.out = out;
    a.screen_scale = nullptr;
    a.tab = tab;
    a.count = frames;
    a.n_fft = pairs;
    a.levels = levels;
    a.off = crop_offset;
    a.n_out = n_out;
    a.accumulate = 0;
    a.s_mul = 2;
    a.s_add = 1;
    a.scale = scale;
#define CALL(N)                                                                               \
  k_fft_rows<N, 1><<<dim3(N / RowsCfg<N>::M, pairs), RowsCfg<N>::T, 0, st>>>(base, base, N); \
  PT_LAUNCH_CHECK();                                                                          \
  return frames_cols_pow2<N>(a, st);
    PT_POW2_SWITCH(n, CALL)
#undef CALL
  }
  rc = pt_fft2_generic(base, base, base + w.scratch, pairs, n, 1, stream);
  if (rc) return rc;
  return launch_pair_epi(base, out, tab, n, frames, levels, crop_offset, n_out, scale, st);
}
