// Subharmonic (low-frequency) field tables and the shared epilogue pieces.
//
// A prep kernel computes, once per call, for every coefficient set s (a
// Monte-Carlo screen or a frame) and every column y:
//   H[s][p][y] = g_2 + conj(g_0)   and   K[s][y] = sum_p Re(g_1) - mean_s
// (see common.hpp for g and the closed-form grid mean), plus the basis table
// U[p][x] = e_p(x).  The column-pass epilogue then adds
//   low(x, y) = K[s][y] + sum_p Re(U[p][x] * H[s][p][y])
// through ColSub: a thread of a column tile holds one column y and the rows
// x0 + d*N/8 (d < 8, x0 < N/8), and since the basis is exponential in x,
//   e_p(x0 + d*N/8) = e_p(x0) * exp(i*pi*d / (4 * 3^(p+1))),
// so with A_p = e_p(x0) * H_p(y) in registers and the 8-entry step table
// F[p][d] in LDS each pixel costs two FMAs per level.  No per-tile setup, no
// FP64 in the transform kernels.
//
// Tables live in the caller's workspace right after the FFT intermediate,
// in c64 (float2) units:  U: P*n,  H: S*P*n,  K: S*n floats.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "fft_core.hpp"

namespace pt {

constexpr int PMAX = 32;  // subharmonic levels supported
constexpr int PREG = 12;  // levels the fused column epilogues keep in registers

struct SubTables {
  float2* U;
  float2* H;
  float* K;
};

__host__ __device__ inline size_t sub_table_c64(int P, int n, int S) {
  if (P == 0) return 0;
  return (size_t)P * n + (size_t)S * P * n + ((size_t)S * n + 1) / 2;
}

__host__ __device__ inline SubTables sub_tables(float2* base, int P, int n, int S) {
  SubTables t;
  t.U = base;
  t.H = base + (size_t)P * n;
  t.K = (float*)(t.H + (size_t)S * P * n);
  return t;
}

// Fill the tables of one coefficient set (C: P blocks of 9, in LDS) for all
// columns; `writeU` also writes the shared basis table.  256-thread blocks.
// With `parts` > 1, block `part` of the set fills one contiguous share of
// the columns (and a strided share of the basis entries).  The grid-mean
// terms and the per-level column coefficients are computed one level per
// thread; then each thread owns whole columns and runs the level loop, so
// no barrier follows the coefficients (a chunked (level, column) split needed
// two per 32 columns, and a 2000-frame call was bound by that serial chain).
// K sums its level terms in ascending order.
// `basis` (may be NULL) is the finished basis table U: its entries are
// read instead of evaluated per column.
__device__ inline void sub_fill_set(const double2* C, int P, int n, int s, SubTables t, bool writeU,
                                    double* s_mean, int tid, int nthreads, int part = 0, int parts = 1,
                                    const float2* basis = nullptr) {
  __shared__ double terms[PMAX];
  __shared__ double hc[PMAX * 9];  // sh_column_coeffs per level
  for (int p = tid; p < P; p += nthreads) {
    terms[p] = sh_block_mean(C + p * 9, sh_mean_e(p, n));
    sh_column_coeffs(C + p * 9, hc + p * 9);
  }
  if (writeU)
    for (int idx = part * nthreads + tid; idx < P * n; idx += parts * nthreads)
      t.U[idx] = sh_basis_e(idx / n, idx % n, n);
  __syncthreads();
  if (tid == 0) {
    double m = 0.0;
    for (int p = 0; p < P; ++p) m += terms[p];
    *s_mean = m;
  }
  __syncthreads();
  const double mean = *s_mean;
  const int cpp = (n + parts - 1) / parts, y0 = part * cpp, y1 = min(n, y0 + cpp);
  float2* H = t.H + (size_t)s * P * n;
  if ((y1 - y0) * P <= 1024 && y1 - y0 < nthreads) {
    // a narrow share (a call of a few sets split over many workgroups): one
    // (level, column) term per thread instead of one column's levels in
    // sequence, K summed over the levels in the same order afterwards
    __shared__ double g1s[1024];
    const int ny = y1 - y0;
    for (int it = tid; it < ny * P; it += nthreads) {
      const int p = it / ny, y = y0 + it % ny;
      double2 Hd;
      double g1;
      sh_column(hc + p * 9, basis ? basis[(size_t)p * n + y] : sh_basis_e(p, y, n), Hd, g1);
      H[(size_t)p * n + y] = make_float2((float)Hd.x, (float)Hd.y);
      g1s[it] = g1;
    }
    __syncthreads();
    for (int y = tid; y < ny; y += nthreads) {
      double kacc = -mean;
      for (int p = 0; p < P; ++p) kacc += g1s[p * ny + y];
      t.K[(size_t)s * n + y0 + y] = (float)kacc;
    }
    return;
  }
  for (int y = y0 + tid; y < y1; y += nthreads) {
    double kacc = -mean;
    for (int p = 0; p < P; ++p) {
      double2 Hd;
      double g1;
      sh_column(hc + p * 9, basis ? basis[(size_t)p * n + y] : sh_basis_e(p, y, n), Hd, g1);
      H[(size_t)p * n + y] = make_float2((float)Hd.x, (float)Hd.y);
      kacc += g1;
    }
    t.K[(size_t)s * n + y] = (float)kacc;
  }
}

// Shifted 3x3 subharmonic blocks of one frame: for level p,
//   S_p[a][b] = sum_l c[p][l][a][b] * qx_{p,l}^(a-1) * qy_{p,l}^(b-1),
//   q_{p,l} = exp(2*pi*i * s_l / (N dx 3^(p+1)))   (FP64, reduced mod 1).
// With `tr` each block is stored transposed (S_p[b][a]): tables built from
// it describe the transposed field low(y, x), whose per-column terms are
// the per-row terms of low(x, y) (the row-pass epilogue).
// Q needs 2*P*L and SB 9*P double2 of LDS.  Ends with a barrier.
__device__ inline void frame_sh_blocks(double2* Q, double2* SB, const double* disp, const double2* coeffs,
                                       int f, int L, int P, double inv_period, int tid, int nthreads,
                                       bool tr = false) {
  for (int q = tid; q < 2 * P * L; q += nthreads) {
    const int axis = q & 1, pl = q >> 1, p = pl / L, l = pl % L;
    const double u = disp[((size_t)f * L + l) * 2 + axis] * inv_period / pow3(p + 1);
    double s, c;
    sincospi(2.0 * (u - floor(u)), &s, &c);
    Q[q] = make_double2(c, s);
  }
  __syncthreads();
  for (int q = tid; q < 9 * P; q += nthreads) {
    const int p = q / 9, k = q % 9, ia = k / 3, ib = k % 3;
    double2 acc = make_double2(0.0, 0.0);
    // the coefficients of 16 layers are loaded before the first is used (a
    // load per iteration would make the sum a chain of load latencies)
    for (int l0 = 0; l0 < L; l0 += 16) {
      double2 cf[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) cf[i] = coeffs[((size_t)p * L + min(l0 + i, L - 1)) * 9 + k];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int l = l0 + i;
        if (l < L) {
          const double2 qx = Q[2 * (p * L + l) + 0], qy = Q[2 * (p * L + l) + 1];
          const double2 wx = ia == 0 ? dconj(qx) : (ia == 1 ? make_double2(1.0, 0.0) : qx);
          const double2 wy = ib == 0 ? dconj(qy) : (ib == 1 ? make_double2(1.0, 0.0) : qy);
          const double2 t = dcmul(dcmul(cf[i], wx), wy);
          acc.x += t.x;
          acc.y += t.y;
        }
      }
    }
    SB[tr ? p * 9 + ib * 3 + ia : q] = acc;
  }
  __syncthreads();
}

// Subharmonic terms of one column tile, staged through LDS: the row-step
// table F, the basis at the N/8 base rows (N <= 512), and H, K of the
// tile's M columns for up to S coefficient sets (screens / frames).  prefetch() issues the
// global loads into registers before the transform (they complete under
// the tile load and the FFT stages); commit() writes them to LDS after the
// transform, followed by one barrier, so the epilogue never waits on HBM.
template <int N, int M, int S, int T, bool UXL_ = (N <= 512)>
struct ColSub {
  // UXL: the basis at the N/8 base rows is staged in LDS (else e_p(x0) is
  // evaluated per thread and level, a sine and cosine each).  Staging it
  // takes P N/8 entries; callers that place the block in the transform tile
  // (free once the last stage has read it) stage it at any size.
  static constexpr bool UXL = UXL_;
  static constexpr int UXP = UXL ? PREG : 1, UXN = UXL ? N / 8 : 1;
  static constexpr int EMAX = UXP * UXN * (UXL ? 1 : 0) + S * PREG * M + S * M;
  static constexpr int R = (EMAX + T - 1) / T;  // prefetched values per thread
  float2 F[PREG][8];
  float2 Ux[UXP][UXN];
  float2 H[S][PREG][M];
  float K[S][M];

  struct Regs {
    float2 v[R];
  };

  // element q of the packed list [U: P*N/8 (UXL only) | H: P*S*M | K: S*M];
  // H is packed level-major so that every index split divides by a
  // compile-time constant (a split by the run-time P*M costs a full
  // integer division per element).
  __device__ __forceinline__ static float2 fetch(const SubTables& t, const int (&set)[S], int c0, int P, int q) {
    const int eu = UXL ? P * (N / 8) : 0, eh = S * P * M;
    if (q < eu) return t.U[(size_t)(q / (N / 8)) * N + q % (N / 8)];
    q -= eu;
    if (q < eh) {
      const int p = q / (S * M), r = q % (S * M), s = r / M, m = r % M;
      return set[s] >= 0 ? t.H[((size_t)set[s] * P + p) * N + c0 + m] : make_float2(0.f, 0.f);
    }
    q -= eh;
    if (q < S * M) {
      const int s = q / M, m = q % M;
      return make_float2(set[s] >= 0 ? t.K[(size_t)set[s] * N + c0 + m] : 0.f, 0.f);
    }
    return make_float2(0.f, 0.f);
  }

  __device__ __forceinline__ static Regs prefetch(const SubTables& t, const int (&set)[S], int c0, int P, int tid) {
    Regs g;
#pragma unroll
    for (int r = 0; r < R; ++r) g.v[r] = fetch(t, set, c0, P, tid + r * T);
    return g;
  }

  __device__ __forceinline__ void commit(const Regs& g, int P, int tid) {
    const int eu = UXL ? P * (N / 8) : 0, eh = S * P * M;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      int q = tid + r * T;
      if (q < eu) {
        Ux[q / UXN][q % UXN] = g.v[r];
        continue;
      }
      q -= eu;
      if (q < eh) {
        const int p = q / (S * M), rr = q % (S * M);
        H[rr / M][p][rr % M] = g.v[r];
        continue;
      }
      q -= eh;
      if (q < S * M) K[q / M][q % M] = g.v[r].x;
    }
    for (int q = tid; q < P * 8; q += T) {
      const int p = q >> 3, d = q & 7;
      float sn, cs;
      sincospif((float)((double)d / (4.0 * pow3(p + 1))), &sn, &cs);
      F[p][d] = make_float2(cs, sn);
    }
  }

  // Adds low(x0 + d_c * N/8, y) to v[c] for the eight rows of a thread
  // (column m of set s); d_c is the row step of the c-th value the
  // transform hands out (see row_step).  Levels loop at run time, so only
  // the eight sums stay live.
  template <int RL>
  __device__ __forceinline__ void add_low(int s, int m, int x0, int P, float (&v)[8]) const {
    const float k = K[s][m];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] += k;
    for (int p = 0; p < P; ++p) {
      const float2 e = UXL ? Ux[p][UXL ? x0 : 0] : sh_basis_e(p, x0, N);  // e_p(x0)
      const float2 a = cmul(e, H[s][p][m]);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float2 f = F[p][row_step<RL>(c)];
        v[c] = fmaf(-f.y, a.y, fmaf(f.x, a.x, v[c]));  // two FMAs (not mul, fma, add)
      }
    }
  }
  // Both sets of a transform pair (S = 2) at once: the step table and the
  // basis are read once per level for the two outputs.  Same arithmetic
  // per output as add_low.
  template <int RL>
  __device__ __forceinline__ void add_low2(int m, int x0, int P, float (&v0)[8], float (&v1)[8],
                                           bool has1) const {
    static_assert(S == 2, "two coefficient sets");
    if (!has1) {
      add_low<RL>(0, m, x0, P, v0);
      return;
    }
    const float k0 = K[0][m], k1 = K[1][m];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      v0[c] += k0;
      v1[c] += k1;
    }
    for (int p = 0; p < P; ++p) {
      const float2 e = UXL ? Ux[p][UXL ? x0 : 0] : sh_basis_e(p, x0, N);  // e_p(x0)
      const float2 a0 = cmul(e, H[0][p][m]), a1 = cmul(e, H[1][p][m]);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float2 f = F[p][row_step<RL>(c)];
        v0[c] = fmaf(-f.y, a0.y, fmaf(f.x, a0.x, v0[c]));
        v1[c] = fmaf(-f.y, a1.y, fmaf(f.x, a1.x, v1[c]));
      }
    }
  }
  // c-th output of a thread in the last Stockham stage (radix RL): stage
  // block q = c / RL, butterfly output r = c % RL, row x0 + (q + r * 8/RL) * N/8.
  template <int RL>
  __device__ static constexpr int row_step(int c) { return c / RL + (c % RL) * (8 / RL); }
};

// Unbounded-level evaluation straight from the tables (general-size path).
__device__ __forceinline__ float sub_eval(const SubTables& t, int s, int x, int y, int n, int P) {
  float acc = t.K[(size_t)s * n + y];
  for (int p = 0; p < P; ++p) {
    const float2 e = t.U[p * n + x];
    const float2 h = t.H[((size_t)s * P + p) * n + y];
    acc += e.x * h.x - e.y * h.y;
  }
  return acc;
}

// Prep launchers (kernels in screens.hip).  Tables must not alias inputs.
int launch_gen_prep(const float* sh_noise, const double* sh_amp, int count, int P, int n, SubTables t,
                    hipStream_t st, int groups = 1);
// cstride: coefficient sets per frame apart (0: one set for every frame).
// transpose: tables of the transposed field (for the row-pass epilogue of
// the column-first frame engine).
int launch_frame_prep(const double* disp, const double2* coeffs, int F, int L, int P, int n, double inv_period,
                      SubTables t, hipStream_t st, size_t cstride = 0, bool transpose = false);
// Generic-size epilogue of frame-pair transforms: frame 2b = Re, 2b+1 = Im
// of pair b, plus the subharmonic field, cropped and scaled.
int launch_pair_epi(const float2* z, float* out, SubTables t, int n, int frames, int levels, int off, int n_out,
                    float scale, hipStream_t st);

}  // namespace pt
