// Subharmonic epilogue tables shared by the fused column passes and the
// general-size epilogue kernels (see common.hpp for the algebra).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace pt {

// Dynamic LDS of a column pass: U[P][N] basis values (when it fits), then
// H[NT][M][P] float2, KP[NT][M][P] float, K[NT][M] float, MEAN[NT][P] double.
struct ShLds {
  float2* U;
  float2* H;
  float* KP;
  float* K;
  double* MEAN;
  bool u_in_lds;
};

__host__ __device__ inline size_t sh_lds_bytes(int n, int P, int NT, int M, bool u_in_lds) {
  if (P == 0) return 0;
  size_t b = u_in_lds ? (size_t)n * P * 8 : 0;
  b += (size_t)NT * M * P * 8 + (size_t)NT * M * P * 4 + (size_t)NT * M * 4;
  b = (b + 7) & ~(size_t)7;
  b += (size_t)NT * P * 8;
  return (b + 15) & ~(size_t)15;
}

__device__ inline ShLds sh_carve(unsigned char* dyn, int n, int P, int NT, int M, bool u_in_lds) {
  ShLds s;
  size_t o = 0;
  s.u_in_lds = u_in_lds;
  s.U = (float2*)(dyn + o);
  if (u_in_lds) o += (size_t)n * P * 8;
  s.H = (float2*)(dyn + o);
  o += (size_t)NT * M * P * 8;
  s.KP = (float*)(dyn + o);
  o += (size_t)NT * M * P * 4;
  s.K = (float*)(dyn + o);
  o += (size_t)NT * M * 4;
  o = (o + 7) & ~(size_t)7;
  s.MEAN = (double*)(dyn + o);
  return s;
}

// Fill U (if in LDS), H, KP and MEAN for NT coefficient sets over columns
// c0..c0+M-1.  `block(t, p, C)` writes the 3x3 block of set t, level p.
// `valid(t)` says whether set t is used.  Ends with K filled; callers
// __syncthreads() before reading.
template <class Block, class Valid>
__device__ inline void sh_setup(ShLds& s, int n, int P, int NT, int M, int c0, int tid, int nthreads,
                                Block& block, Valid& valid) {
  if (s.u_in_lds) {
    for (int idx = tid; idx < P * n; idx += nthreads) s.U[idx] = sh_basis_e(idx / n, idx % n, n);
  }
  const int ntask = NT * M * P;
  for (int task = tid; task < ntask + NT * P; task += nthreads) {
    if (task < ntask) {
      const int p = task % P;
      const int m = (task / P) % M;
      const int t = task / (P * M);
      if (!valid(t)) continue;
      double2 C[9];
      block(t, p, C);
      int y = c0 + m;
      if (y >= n) y = n - 1;  // tile overhang (never stored)
      double2 H;
      double g1;
      sh_column(C, sh_basis_e(p, y, n), H, g1);
      s.H[(t * M + m) * P + p] = make_float2((float)H.x, (float)H.y);
      s.KP[(t * M + m) * P + p] = (float)g1;
    } else {
      const int q = task - ntask;
      const int p = q % P, t = q / P;
      if (!valid(t)) continue;
      double2 C[9];
      block(t, p, C);
      s.MEAN[t * P + p] = sh_block_mean(C, sh_mean_e(p, n));
    }
  }
  __syncthreads();
  for (int q = tid; q < NT * M; q += nthreads) {
    const int t = q / M;
    if (!valid(t)) continue;
    double acc = 0.0;
    for (int p = 0; p < P; ++p) acc += (double)s.KP[q * P + p] - s.MEAN[t * P + p];
    s.K[q] = (float)acc;
  }
}

__device__ __forceinline__ float sh_eval(const ShLds& s, int t, int m, int M, int x, int n, int P) {
  float acc = s.K[t * M + m];
  const float2* H = s.H + (size_t)(t * M + m) * P;
  for (int p = 0; p < P; ++p) {
    float2 e = s.u_in_lds ? s.U[p * n + x] : sh_basis_e(p, x, n);
    float2 h = H[p];
    acc += e.x * h.x - e.y * h.y;
  }
  return acc;
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

}  // namespace pt
