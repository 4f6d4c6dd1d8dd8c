// Assemat-Wilson extrusion and the fused multi-layer readout, gfx950.
//
// Ring buffers: every layer owns a (cap, W) float32 slab of one
// (layers, cap, W) array.  Local row i of a layer (virtual row base + i in
// the reference's bookkeeping) lives at physical row (start + i) mod cap, so
// "compaction" is an index shift -- no row is ever copied.
//
//  * k_gemm_nt    : innovation terms for every new row of a step at once,
//                   BB = beta * B_unit^T  (K x W), one GEMM reading B once.
//  * k_extrude_rows: one sequential round of the recurrence for up to 16
//                   layers: row = A * z + s_l * BB[k]; one wave per output
//                   element, A streamed once per round for all layers.
//  * k_readout    : rotated, sub-pixel, per-layer-shifted interpolation of
//                   all layers for up to 8 directions (each with its own
//                   per-layer magnification, so NGS and LGS cones share one
//                   launch), summed and scaled into (dirs, n, n).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "fft_core.hpp"

namespace pt {

// ---------------------------------------------------------------------------
// C (K x N) = X (K x J) * Y (N x J)^T, row-major, FP32
// ---------------------------------------------------------------------------
constexpr int GT = 64, GK = 16;

__global__ __launch_bounds__(256) void k_gemm_nt(const float* __restrict__ X, const float* __restrict__ Y,
                                                 float* __restrict__ C, int K, int N, int J) {
  __shared__ float xs[GK][GT + 1];
  __shared__ float ys[GK][GT + 1];
  const int tid = threadIdx.x;
  const int k0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int tn = tid % 16, tk = tid / 16;
  float acc[4][4] = {};
  for (int j0 = 0; j0 < J; j0 += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      const int r = e / GK, jj = e % GK, j = j0 + jj;
      xs[jj][r] = (k0 + r < K && j < J) ? X[(size_t)(k0 + r) * J + j] : 0.f;
      ys[jj][r] = (n0 + r < N && j < J) ? Y[(size_t)(n0 + r) * J + j] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < GK; ++jj) {
      float xv[4], yv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        xv[q] = xs[jj][tk + 16 * q];
        yv[q] = ys[jj][tn + 16 * q];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] += xv[a] * yv[b];
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int k = k0 + tk + 16 * a;
    if (k >= K) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int nn = n0 + tn + 16 * b;
      if (nn < N) C[(size_t)k * N + nn] = acc[a][b];
    }
  }
}

// ---------------------------------------------------------------------------
// one round of the row recurrence
// ---------------------------------------------------------------------------
constexpr int XMAX = 16;  // layers per round
constexpr int MMAX = 4;   // stencil rows

struct RowsArgs {
  const float* A;   // (W, m*W)
  const float* bb;  // (K, W) innovation terms (unscaled)
  float* buf;       // (layers, cap, W)
  long long cap;
  int W, m, nact;
  int layer[XMAX];
  int zrow[XMAX][MMAX];
  int trow[XMAX];
  int bbrow[XMAX];
  float bscale[XMAX];
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_extrude_rows(RowsArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  if (row >= a.W) return;
  const int W = a.W;
  const float* Ar = a.A + (size_t)row * a.m * W;
  float acc[XMAX];
#pragma unroll
  for (int e = 0; e < XMAX; ++e) acc[e] = 0.f;
#pragma unroll
  for (int s = 0; s < MMAX; ++s) {
    if (s >= a.m) break;
    const float* As = Ar + (size_t)s * W;
    const float* zp[XMAX];
#pragma unroll
    for (int e = 0; e < XMAX; ++e)
      zp[e] = a.buf + ((size_t)a.layer[e < a.nact ? e : 0] * a.cap + a.zrow[e < a.nact ? e : 0][s]) * W;
    for (int c = lane; c < W; c += 64) {
      const float av = As[c];
#pragma unroll
      for (int e = 0; e < XMAX; ++e)
        if (e < a.nact) acc[e] = fmaf(av, zp[e][c], acc[e]);
    }
  }
  float mine = 0.f;
#pragma unroll
  for (int e = 0; e < XMAX; ++e) {
    if (e < a.nact) {
      const float t = wave_sum(acc[e]);
      if (lane == e) mine = t;
    }
  }
  if (lane < a.nact) {
    int l = 0, tr = 0, br = 0;
    float sc = 0.f;
#pragma unroll
    for (int e = 0; e < XMAX; ++e)
      if (lane == e) {
        l = a.layer[e];
        tr = a.trow[e];
        br = a.bbrow[e];
        sc = a.bscale[e];
      }
    a.buf[((size_t)l * a.cap + tr) * W + row] = mine + sc * a.bb[(size_t)br * W + row];
  }
}

// ---------------------------------------------------------------------------
// fused readout
// ---------------------------------------------------------------------------
constexpr int RL = 16, RD = 8;

struct ReadoutArgs {
  const float* buf;  // (L, cap, W)
  float* out;        // (D, n, n)
  long long cap;
  int W, n, L, D, accumulate;
  float scale;
  int start[RL];
  int fillm1[RL];
  double cosw[RL], sinw[RL];
  double sa[RD][RL], sp[RD][RL], mag[RD][RL];
};

template <int TAPS>
__device__ __forceinline__ void taps_weights(float t, float* w) {
  if constexpr (TAPS == 2) {
    w[0] = 1.f - t;
    w[1] = t;
  } else if constexpr (TAPS == 4) {  // Catmull-Rom, offsets -1..2
    const float t2 = t * t, t3 = t2 * t;
    w[0] = 0.5f * (-t + 2.f * t2 - t3);
    w[1] = 0.5f * (2.f - 5.f * t2 + 3.f * t3);
    w[2] = 0.5f * (t + 4.f * t2 - 3.f * t3);
    w[3] = 0.5f * (t3 - t2);
  } else {  // Lanczos-3, offsets -2..3, normalised
    float tot = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float x = t - (float)(k - 2);
      float s;
      if (x == 0.f) {
        s = 1.f;
      } else {
        const float x3 = x * (1.f / 3.f);
        s = (sinpif(x) / (3.14159265358979f * x)) * (sinpif(x3) / (3.14159265358979f * x3));
      }
      w[k] = s;
      tot += s;
    }
    const float inv = 1.f / tot;
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] *= inv;
  }
}

template <int TAPS>
__global__ __launch_bounds__(256) void k_readout(ReadoutArgs a) {
  constexpr int OFF = TAPS == 2 ? 0 : (TAPS == 4 ? -1 : -2);
  const int j = blockIdx.x * 32 + (threadIdx.x & 31);
  const int i = blockIdx.y * 8 + (threadIdx.x >> 5);
  if (i >= a.n || j >= a.n) return;
  const double ci = (double)i - 0.5 * (a.n - 1), cj = (double)j - 0.5 * (a.n - 1);
  const double halfw = 0.5 * a.W;
  for (int d = 0; d < a.D; ++d) {
    float acc = 0.f;
    for (int l = 0; l < a.L; ++l) {
      const double mg = a.mag[d][l];
      const double gi = ci * mg, gj = cj * mg;
      const double c = a.cosw[l], s = a.sinw[l];
      const double row = (gi * c + gj * s) + a.sa[d][l];
      const double col = (-gi * s + gj * c + halfw) + a.sp[d][l];
      const double rf = floor(row), cf = floor(col);
      const long long r0 = (long long)rf, c0 = (long long)cf;
      float wr[TAPS], wc[TAPS];
      taps_weights<TAPS>((float)(row - rf), wr);
      taps_weights<TAPS>((float)(col - cf), wc);
      const float* bl = a.buf + (size_t)l * a.cap * a.W;
      const long long fm = a.fillm1[l];
      const long long st = a.start[l];
      float val = 0.f;
#pragma unroll
      for (int u = 0; u < TAPS; ++u) {
        long long rr = r0 + OFF + u;
        rr = rr < 0 ? 0 : (rr > fm ? fm : rr);
        long long ph = st + rr;
        if (ph >= a.cap) ph -= a.cap;
        const float* br = bl + ph * a.W;
        float rs = 0.f;
#pragma unroll
        for (int v = 0; v < TAPS; ++v) {
          long long cc = c0 + OFF + v;
          cc = cc < 0 ? 0 : (cc > a.W - 1 ? a.W - 1 : cc);
          rs = fmaf(wc[v], br[cc], rs);
        }
        val = fmaf(wr[u], rs, val);
      }
      acc += val;
    }
    float* o = a.out + ((size_t)d * a.n + i) * a.n + j;
    const float v = acc * a.scale;
    *o = a.accumulate ? *o + v : v;
  }
}

}  // namespace pt

using namespace pt;

extern "C" int pt_innovation_gemm(const float* beta, const float* B, int rows, int width, float* out,
                                  void* stream) {
  if (!beta || !B || !out) return fail(PT_ERR_ARG, "pt_innovation_gemm: NULL array");
  if (rows < 0 || width < 1) return fail(PT_ERR_ARG, "pt_innovation_gemm: bad sizes");
  if (rows == 0) return PT_OK;
  if ((rows + GT - 1) / GT > 65535) return fail(PT_ERR_ARG, "pt_innovation_gemm: too many rows");
  dim3 g((width + GT - 1) / GT, (rows + GT - 1) / GT);
  k_gemm_nt<<<g, 256, 0, (hipStream_t)stream>>>(beta, B, out, rows, width, width);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

extern "C" int pt_extrude_rows(const float* A, int width, int stencil_rows, const float* bb, float* buf,
                               long long cap, int nact, const int* layer, const int* zrows, const int* trow,
                               const int* bbrow, const float* bscale, void* stream) {
  if (!A || !buf || !layer || !zrows || !trow || !bbrow || !bscale || !bb)
    return fail(PT_ERR_ARG, "pt_extrude_rows: NULL array");
  if (width < 1 || stencil_rows < 1 || cap < stencil_rows + 1 || nact < 0)
    return fail(PT_ERR_ARG, "pt_extrude_rows: bad sizes");
  if (stencil_rows > MMAX) return fail(PT_ERR_UNSUPPORTED, "pt_extrude_rows: stencil_rows > %d", MMAX);
  if (nact == 0) return PT_OK;
  hipStream_t st = (hipStream_t)stream;
  for (int e0 = 0; e0 < nact; e0 += XMAX) {
    RowsArgs a;
    a.A = A;
    a.bb = bb;
    a.buf = buf;
    a.cap = cap;
    a.W = width;
    a.m = stencil_rows;
    a.nact = nact - e0 < XMAX ? nact - e0 : XMAX;
    for (int e = 0; e < XMAX; ++e) {
      const int src = e0 + (e < a.nact ? e : 0);
      a.layer[e] = layer[src];
      for (int s = 0; s < MMAX; ++s) {
        const int r = s < stencil_rows ? zrows[src * stencil_rows + s] : 0;
        if (r < 0 || r >= cap) return fail(PT_ERR_ARG, "pt_extrude_rows: stencil row %d outside ring", r);
        a.zrow[e][s] = r;
      }
      if (trow[src] < 0 || trow[src] >= cap) return fail(PT_ERR_ARG, "pt_extrude_rows: target row outside ring");
      a.trow[e] = trow[src];
      a.bbrow[e] = bbrow[src];
      a.bscale[e] = bscale[src];
    }
    k_extrude_rows<<<(width + 3) / 4, 256, 0, st>>>(a);
    PT_LAUNCH_CHECK();
  }
  return PT_OK;
}

extern "C" int pt_readout(const float* buf, int layers, long long cap, int width, const int* start,
                          const long long* fill, const double* wind_cs, const double* dir_geom, int dirs, int n,
                          int interp, float scale, int accumulate, float* out, void* stream) {
  if (!buf || !start || !fill || !wind_cs || !dir_geom || !out) return fail(PT_ERR_ARG, "pt_readout: NULL array");
  if (layers < 1 || dirs < 1 || n < 1 || width < 1 || cap < 1) return fail(PT_ERR_ARG, "pt_readout: bad sizes");
  if (interp < 0 || interp > 2) return fail(PT_ERR_ARG, "pt_readout: interp must be 0 linear, 1 cubic, 2 lanczos");
  for (int l = 0; l < layers; ++l) {
    if (fill[l] < 1 || fill[l] > cap || start[l] < 0 || start[l] >= cap)
      return fail(PT_ERR_ARG, "pt_readout: layer %d ring state out of range", l);
  }
  hipStream_t st = (hipStream_t)stream;
  dim3 g((n + 31) / 32, (n + 7) / 8);
  for (int d0 = 0; d0 < dirs; d0 += RD) {
    for (int l0 = 0; l0 < layers; l0 += RL) {
      ReadoutArgs a;
      a.buf = buf + (size_t)l0 * cap * width;
      a.out = out + (size_t)d0 * n * n;
      a.cap = cap;
      a.W = width;
      a.n = n;
      a.L = layers - l0 < RL ? layers - l0 : RL;
      a.D = dirs - d0 < RD ? dirs - d0 : RD;
      a.accumulate = (accumulate || l0 > 0) ? 1 : 0;
      a.scale = scale;
      for (int l = 0; l < RL; ++l) {
        const int src = l0 + (l < a.L ? l : 0);
        a.start[l] = start[src];
        a.fillm1[l] = (int)(fill[src] - 1);
        a.cosw[l] = wind_cs[2 * src];
        a.sinw[l] = wind_cs[2 * src + 1];
      }
      for (int d = 0; d < RD; ++d)
        for (int l = 0; l < RL; ++l) {
          const int sd = d0 + (d < a.D ? d : 0), sl = l0 + (l < a.L ? l : 0);
          const double* gp = dir_geom + ((size_t)sd * layers + sl) * 3;
          a.sa[d][l] = gp[0];
          a.sp[d][l] = gp[1];
          a.mag[d][l] = gp[2];
        }
      if (interp == 0)
        k_readout<2><<<g, 256, 0, st>>>(a);
      else if (interp == 1)
        k_readout<4><<<g, 256, 0, st>>>(a);
      else
        k_readout<6><<<g, 256, 0, st>>>(a);
      PT_LAUNCH_CHECK();
    }
  }
  return PT_OK;
}
