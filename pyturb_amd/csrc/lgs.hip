// Spectral LGS cone readout for gfx950: separable cone zoom of every
// layer's screen about its centre, summed over layers.
//
// With a laser guide star at altitude H, layer l is seen through the cone
// scaled by mag_l = 1 - h_l / H, so the reference samples each layer's full
// screen at the pupil grid scaled about the screen centre and sums the
// layers (atmosphere.py:1000-1090).  The per-layer taps are separable and
// isotropic: out(i, j) = sum_l sum_a sum_b w_l,a(i) w_l,b(j) S_l[idx_l,a(i)][idx_l,b(j)],
// with idx, w of shape (L, T, n) built once from the magnifications
// (atmosphere.py:581-616).
//
// k_lgs_zoom: a workgroup owns a 32 x 32 output tile of one frame.  For
// each layer it stages the tile's source window (rows idx[0][i0] ..
// idx[T-1][i1] by the same columns; at most 31 mag + T + 1 <= 38 on a side
// for mag <= 1) and the tile's row taps in LDS, interpolates along rows
// into a 32 x window-width LDS block, then along columns into registers.
// The window and row taps of layer l + 1 are loaded into registers while
// layer l is interpolated and committed to the second LDS buffer during
// its column pass: two barriers per layer, and the loads' latency overlaps
// the arithmetic.  Windows wider than the LDS box (mag > 1) read the screen
// directly.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace pt {

constexpr int LT = 32;     // output tile edge
constexpr int LBOX = 40;   // LDS window edge
constexpr int LNG = 256 / LBOX;                  // row groups of the window lanes (6; lanes 240.. idle)
constexpr int LWR = (LBOX + LNG - 1) / LNG;      // window rows per lane (7)
constexpr int LRR = (LT + LNG - 1) / LNG;        // row-pass rows per lane (6)

struct ZoomWin {
  int r_lo, c_lo, nr, nc;
  bool fits;
};

struct alignas(8) RowTap {
  float w;  // weight
  int off;  // float offset of the source row in the window box
};

// Lanes of the window load, the commit and the row pass own one window
// column cw = tid % LBOX and the rows gw + LNG k (gw = tid / LBOX), so no
// per-value index division remains; a row tap is stored as (weight, LDS
// offset of its source row) and read as one 8-byte value.
template <int T>
__global__ __launch_bounds__(256) void k_lgs_zoom(const float* __restrict__ screens, int L, int F, int ns,
                                                  const int* __restrict__ idx, const float* __restrict__ w, int n,
                                                  float scale, float* __restrict__ out) {
  static_assert(T * LT <= 256, "one row tap per thread");
  __shared__ float box[2][LBOX][LBOX + 1];
  __shared__ float rows[LT][LBOX + 1];
  __shared__ RowTap rt[2][T][LT];  // row taps of the tile's rows
  const int tid = threadIdx.x;
  const int i0 = blockIdx.y * LT, j0 = blockIdx.x * LT, f = blockIdx.z;
  const int i1 = min(i0 + LT, n) - 1, j1 = min(j0 + LT, n) - 1;
  const int jl = tid & 31, il = tid >> 5;  // pixel (i0 + il + 8k, j0 + jl), k < 4
  const int cw = tid % LBOX, gw = tid / LBOX;
  const bool wl_lane = gw < LNG;
  const size_t plane = (size_t)ns * ns;
  const int ta = tid / LT, ti = tid % LT;  // this thread's row tap (tid < T * LT)
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  float pre[LWR];
  float prw = 0.f;
  int pri = 0;
  // window of layer l: idx is non-decreasing in the pixel index and in the tap
  auto window = [&](int l) {
    const int* ix = idx + (size_t)l * T * n;
    ZoomWin g;
    g.r_lo = ix[i0];
    g.c_lo = ix[j0];
    g.nr = ix[(size_t)(T - 1) * n + i1] - g.r_lo + 1;
    g.nc = ix[(size_t)(T - 1) * n + j1] - g.c_lo + 1;
    g.fits = g.nr <= LBOX && g.nc <= LBOX;
    return g;
  };
  auto fetch = [&](int l, const ZoomWin& g) {
    if (!g.fits) return;
    if (wl_lane && cw < g.nc) {
      const float* S = screens + ((size_t)l * F + f) * plane + (size_t)g.r_lo * ns + g.c_lo + cw;
#pragma unroll
      for (int k = 0; k < LWR; ++k) {
        const int r = gw + LNG * k;
        if (r < g.nr) pre[k] = S[(size_t)r * ns];
      }
    }
    if (tid < T * LT) {
      const size_t e = ((size_t)l * T + ta) * n + min(i0 + ti, n - 1);
      prw = w[e];
      pri = (idx[e] - g.r_lo) * (LBOX + 1);
    }
  };
  auto commit = [&](int sel, const ZoomWin& g) {
    if (!g.fits) return;
    if (wl_lane && cw < g.nc) {
#pragma unroll
      for (int k = 0; k < LWR; ++k) {
        const int r = gw + LNG * k;
        if (r < g.nr) box[sel][r][cw] = pre[k];
      }
    }
    if (tid < T * LT) rt[sel][ta][ti] = RowTap{prw, pri};
  };
  ZoomWin cur = window(0);
  fetch(0, cur);
  commit(0, cur);
  __syncthreads();
  for (int l = 0; l < L; ++l) {
    const int sel = l & 1;
    ZoomWin nxt;
    nxt.fits = false;
    if (l + 1 < L) {
      nxt = window(l + 1);
      fetch(l + 1, nxt);  // in flight while layer l is interpolated
    }
    const float* S = screens + ((size_t)l * F + f) * plane;
    const int* ix = idx + (size_t)l * T * n;
    const float* wl = w + (size_t)l * T * n;
    const int j = min(j0 + jl, n - 1);
    float wb[T];
    int cb[T];
#pragma unroll
    for (int b = 0; b < T; ++b) {
      wb[b] = wl[(size_t)b * n + j];
      cb[b] = ix[(size_t)b * n + j];
    }
    if (cur.fits && wl_lane && cw < cur.nc) {
      // rows: rows[i][c] = sum_a w_a(i) box[idx_a(i) - r_lo][c]
      const float* bx = &box[sel][0][0] + cw;
#pragma unroll
      for (int k = 0; k < LRR; ++k) {
        const int i = gw + LNG * k;
        if (i < LT) {
          float v = 0.f;
#pragma unroll
          for (int a = 0; a < T; ++a) {
            const RowTap tp = rt[sel][a][i];
            v = fmaf(tp.w, bx[tp.off], v);
          }
          rows[i][cw] = v;
        }
      }
    }
    __syncthreads();  // rows complete
    if (cur.fits) {
#pragma unroll
      for (int b = 0; b < T; ++b) cb[b] -= cur.c_lo;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float v = 0.f;
#pragma unroll
        for (int b = 0; b < T; ++b) v = fmaf(wb[b], rows[il + 8 * k][cb[b]], v);
        acc[k] += v;
      }
    } else {  // magnification > 1: not unrolled, so it does not set the register budget
#pragma unroll 1
      for (int k = 0; k < 4; ++k) {
        const int i = min(i0 + il + 8 * k, n - 1);
        float v = 0.f;
#pragma unroll
        for (int a = 0; a < T; ++a) {
          const float* row = S + (size_t)ix[(size_t)a * n + i] * ns;
          float r = 0.f;
#pragma unroll
          for (int b = 0; b < T; ++b) r = fmaf(wb[b], row[cb[b]], r);
          v = fmaf(wl[(size_t)a * n + i], r, v);
        }
        acc[k] += v;
      }
    }
    if (l + 1 < L) {
      // box[sel ^ 1] and its taps were last read in layer l - 1's row pass
      commit(sel ^ 1, nxt);
      cur = nxt;
    }
    __syncthreads();  // rows free for the next row pass; the committed window visible
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = i0 + il + 8 * k, j = j0 + jl;
    if (i < n && j < n) out[((size_t)f * n + i) * n + j] = acc[k] * scale;
  }
}

}  // namespace pt

using namespace pt;

extern "C" int pt_lgs_zoom(const float* screens, int layers, int frames, int ns, const int* idx, const float* w,
                           int taps, int n, float scale, float* out, void* stream) {
  if (!screens || !idx || !w || !out) return fail(PT_ERR_ARG, "pt_lgs_zoom: NULL array");
  if (layers < 1 || frames < 0 || ns < 1 || n < 1) return fail(PT_ERR_ARG, "pt_lgs_zoom: bad sizes");
  if (frames > 65535) return fail(PT_ERR_ARG, "pt_lgs_zoom: frames > 65535 per call");
  if (frames == 0) return PT_OK;
  hipStream_t st = (hipStream_t)stream;
  const dim3 g((n + LT - 1) / LT, (n + LT - 1) / LT, frames);
  if (taps == 2)
    k_lgs_zoom<2><<<g, 256, 0, st>>>(screens, layers, frames, ns, idx, w, n, scale, out);
  else if (taps == 4)
    k_lgs_zoom<4><<<g, 256, 0, st>>>(screens, layers, frames, ns, idx, w, n, scale, out);
  else if (taps == 6)
    k_lgs_zoom<6><<<g, 256, 0, st>>>(screens, layers, frames, ns, idx, w, n, scale, out);
  else
    return fail(PT_ERR_ARG, "pt_lgs_zoom: taps must be 2 (linear), 4 (cubic) or 6 (lanczos)");
  PT_LAUNCH_CHECK();
  return PT_OK;
}
