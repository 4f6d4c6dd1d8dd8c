// Subharmonic (low-frequency) field tables and the shared epilogue pieces.
//
// A prep kernel computes, once per call, for every coefficient set s (a
// Monte-Carlo screen or a frame) and every column y:
//   H[s][p][y] = g_2 + conj(g_0)   and   K[s][y] = sum_p Re(g_1) - mean_s
// (see common.hpp for g and the closed-form grid mean), plus the basis table
// U[p][x] = e_p(x).  The column-pass epilogue then adds
//   low(x, y) = K[s][y] + sum_p Re(U[p][x] * H[s][p][y])
// with H/K in registers (one column per thread) and U in LDS: no per-tile
// setup, no FP64 in the transform kernels.
//
// Tables live in the caller's workspace right after the FFT intermediate,
// in c64 (float2) units:  U: P*n,  H: S*P*n,  K: S*n floats.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

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
__device__ inline void sub_fill_set(const double2* C, int P, int n, int s, SubTables t, bool writeU,
                                    double* s_mean, int tid, int nthreads) {
  if (tid == 0) {
    double m = 0.0;
    for (int p = 0; p < P; ++p) m += sh_block_mean(C + p * 9, sh_mean_e(p, n));
    *s_mean = m;
  }
  if (writeU)
    for (int idx = tid; idx < P * n; idx += nthreads) t.U[idx] = sh_basis_e(idx / n, idx % n, n);
  __syncthreads();
  const double mean = *s_mean;
  for (int y = tid; y < n; y += nthreads) {
    double kacc = -mean;
    for (int p = 0; p < P; ++p) {
      double2 Hd;
      double g1;
      sh_column(C + p * 9, sh_basis_e(p, y, n), Hd, g1);
      t.H[((size_t)s * P + p) * n + y] = make_float2((float)Hd.x, (float)Hd.y);
      kacc += g1;
    }
    t.K[(size_t)s * n + y] = (float)kacc;
  }
}

// Shifted 3x3 subharmonic blocks of one frame: for level p,
//   S_p[a][b] = sum_l c[p][l][a][b] * qx_{p,l}^(a-1) * qy_{p,l}^(b-1),
//   q_{p,l} = exp(2*pi*i * s_l / (N dx 3^(p+1)))   (FP64, reduced mod 1).
// Q needs 2*P*L and SB 9*P double2 of LDS.  Ends with a barrier.
__device__ inline void frame_sh_blocks(double2* Q, double2* SB, const double* disp, const double2* coeffs,
                                       int f, int L, int P, double inv_period, int tid, int nthreads) {
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
    for (int l = 0; l < L; ++l) {
      const double2 qx = Q[2 * (p * L + l) + 0], qy = Q[2 * (p * L + l) + 1];
      const double2 wx = ia == 0 ? dconj(qx) : (ia == 1 ? make_double2(1.0, 0.0) : qx);
      const double2 wy = ib == 0 ? dconj(qy) : (ib == 1 ? make_double2(1.0, 0.0) : qy);
      const double2 t = dcmul(dcmul(coeffs[((size_t)p * L + l) * 9 + k], wx), wy);
      acc.x += t.x;
      acc.y += t.y;
    }
    SB[q] = acc;
  }
  __syncthreads();
}

// One column's K and H in registers.
struct ShRegs {
  int y;
  float k;
  float2 h[PREG];
  __device__ __forceinline__ void load(const SubTables& t, int s, int yy, int n, int P) {
    y = yy;
    k = t.K[(size_t)s * n + yy];
#pragma unroll
    for (int p = 0; p < PREG; ++p)
      if (p < P) h[p] = t.H[((size_t)s * P + p) * n + yy];
  }
  // U: basis table with row stride n (LDS or global)
  __device__ __forceinline__ float eval(const float2* U, int x, int n, int P) const {
    float acc = k;
#pragma unroll
    for (int p = 0; p < PREG; ++p) {
      if (p < P) {
        const float2 e = U[p * n + x];
        acc += e.x * h[p].x - e.y * h[p].y;
      }
    }
    return acc;
  }
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
                    hipStream_t st);
int launch_frame_prep(const double* disp, const double2* coeffs, int F, int L, int P, int n, double inv_period,
                      SubTables t, hipStream_t st);

}  // namespace pt
