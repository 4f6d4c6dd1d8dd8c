// Shared host/device helpers for the pyturb HIP library: status codes,
// error reporting and the subharmonic (low-frequency) epilogue tables.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pyturb_hip.h"

namespace pt {

// Records a formatted message retrievable through pt_last_error() and
// returns `code`.  Thread-local, like errno.
int fail(int code, const char* fmt, ...);

#define PT_LAUNCH_CHECK()                                                   \
  do {                                                                      \
    hipError_t e_ = hipGetLastError();                                      \
    if (e_ != hipSuccess)                                                   \
      return pt::fail(PT_ERR_HIP, "%s: launch failed: %s", __func__,        \
                      hipGetErrorString(e_));                               \
  } while (0)

// ---------------------------------------------------------------------------
// Subharmonic field.  Level p (1-based in the physics, 0-based here) adds
// modes at frequencies (-1, 0, +1) * df / 3^(p+1) per axis.  With
//   e_p(x) = exp(i*pi*(2x - N) / (N * 3^(p+1)))        (the +df_p basis)
//   B_p(x) = [conj(e_p(x)), 1, e_p(x)]
// a 3x3 complex coefficient block C_p contributes Re(B_p(x)^T C_p B_p(y)).
// For a fixed column y this collapses to
//   low(x, y) = K(y) + sum_p Re(e_p(x) * H_p(y)),
//   g_a(y) = sum_b C_p[a][b] B_p(y)[b],  H_p = g_2 + conj(g_0),
//   K(y)  = sum_p Re(g_1) - mean,
// and the grid mean has the closed form sum_p Re(Mb_p^T C_p Mb_p) with
// Mb_p = [conj(m_p), 1, m_p], m_p = mean_x e_p(x)
//      = exp(-i*pi/(N 3^(p+1))) * sin(pi/3^(p+1)) / (N sin(pi/(N 3^(p+1)))).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double pow3(int k) {
  double r = 1.0;
  for (int i = 0; i < k; ++i) r *= 3.0;
  return r;
}

__device__ __forceinline__ float2 sh_basis_e(int p, int x, int n) {
  float s, c;
  float den = (float)((double)n * pow3(p + 1));
  sincospif((float)(2 * x - n) / den, &s, &c);
  return make_float2(c, s);
}

__device__ __forceinline__ double2 sh_mean_e(int p, int n) {
  double a = 1.0 / ((double)n * pow3(p + 1));  // half-cycles of e_p per pixel
  double s1 = sinpi(1.0 / pow3(p + 1));
  double s2 = sinpi(a);
  double mag = s1 / ((double)n * s2);
  double sa, ca;
  sincospi(-a, &sa, &ca);
  return make_double2(mag * ca, mag * sa);
}

__device__ __forceinline__ double2 dcmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 dconj(double2 a) { return make_double2(a.x, -a.y); }

// H_p(y) = g_2 + conj(g_0) and Re(g_1(y)) are affine in the basis value
// e = (c, s): with alpha = C20 + conj(C02), beta = C21 + conj(C01) and
// gamma = C22 + conj(C00),
//   H = beta + alpha conj(e) + gamma e,   Re g_1 = Re C11 + Re((conj(C10) + C12) e).
// sh_column_coeffs folds one block into the nine coefficients
//   H.x = h0 + h2 c + h4 s,  H.y = h1 + h3 c + h5 s,  Re g_1 = h6 + h7 c + h8 s,
// once per block, so each column costs six FP64 FMAs.
__device__ __forceinline__ void sh_column_coeffs(const double2 C[9], double h[9]) {
  const double2 al = make_double2(C[6].x + C[2].x, C[6].y - C[2].y);
  const double2 be = make_double2(C[7].x + C[1].x, C[7].y - C[1].y);
  const double2 ga = make_double2(C[8].x + C[0].x, C[8].y - C[0].y);
  h[0] = be.x;
  h[1] = be.y;
  h[2] = al.x + ga.x;
  h[3] = al.y + ga.y;
  h[4] = al.y - ga.y;
  h[5] = ga.x - al.x;
  h[6] = C[4].x;
  h[7] = C[3].x + C[5].x;
  h[8] = C[3].y - C[5].y;
}

// H_p(y) and Re(g_1(y)) from a block's coefficients and one basis value eb.
__device__ __forceinline__ void sh_column(const double* h, float2 eb, double2& H, double& g1re) {
  const double c = eb.x, s = eb.y;
  H = make_double2(fma(h[4], s, fma(h[2], c, h[0])), fma(h[5], s, fma(h[3], c, h[1])));
  g1re = fma(h[8], s, fma(h[7], c, h[6]));
}

// Re(Mb^T C Mb): the block's contribution to the grid mean.
__device__ __forceinline__ double sh_block_mean(const double2 C[9], double2 mp) {
  double2 mb[3] = {dconj(mp), make_double2(1.0, 0.0), mp};
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) acc += dcmul(dcmul(C[a * 3 + b], mb[a]), mb[b]).x;
  return acc;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Remapping the flat workgroup index with xcd_tile gives every XCD a
// contiguous run of logical tiles, so tiles that share cache lines (column
// tiles of one grid) meet in one L2 instead of being fetched by several.  A
// bijection on [0, nx) for any nx.
__device__ __forceinline__ int xcd_tile(int bx, int nx) {
  const int xcd = bx & 7, k = bx >> 3;
  const int q = nx >> 3, r = nx & 7;
  return (xcd < r ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q) + k;
}

}  // namespace pt
