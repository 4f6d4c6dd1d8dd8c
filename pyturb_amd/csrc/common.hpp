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

// H_p(y) and Re(g_1(y)) for one coefficient block and one basis value eb.
__device__ __forceinline__ void sh_column(const double2 C[9], float2 eb, double2& H, double& g1re) {
  double2 e = make_double2(eb.x, eb.y);
  double2 ec = dconj(e);
  double2 g[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double2 t0 = dcmul(C[a * 3 + 0], ec);
    double2 t2 = dcmul(C[a * 3 + 2], e);
    g[a] = make_double2(t0.x + C[a * 3 + 1].x + t2.x, t0.y + C[a * 3 + 1].y + t2.y);
  }
  H = make_double2(g[2].x + g[0].x, g[2].y - g[0].y);
  g1re = g[1].x;
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
