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
// for mag <= 1) into LDS with coalesced row reads, interpolates along rows
// into a 32 x window-width LDS block, then along columns into registers.
// Windows wider than the LDS box (mag > 1) read the screen directly.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace pt {

constexpr int LT = 32;     // output tile edge
constexpr int LBOX = 40;   // LDS window edge

template <int T>
__global__ __launch_bounds__(256) void k_lgs_zoom(const float* __restrict__ screens, int L, int F, int ns,
                                                  const int* __restrict__ idx, const float* __restrict__ w, int n,
                                                  float scale, float* __restrict__ out) {
  __shared__ float box[LBOX][LBOX + 1];
  __shared__ float rows[LT][LBOX + 1];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.y * LT, j0 = blockIdx.x * LT, f = blockIdx.z;
  const int i1 = min(i0 + LT, n) - 1, j1 = min(j0 + LT, n) - 1;
  const int jl = tid & 31, il = tid >> 5;  // pixel (i0 + il + 8k, j0 + jl), k < 4
  const size_t plane = (size_t)ns * ns;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int l = 0; l < L; ++l) {
    const float* S = screens + ((size_t)l * F + f) * plane;
    const int* ix = idx + (size_t)l * T * n;
    const float* wl = w + (size_t)l * T * n;
    // window: idx is non-decreasing in the pixel index and in the tap
    const int r_lo = ix[i0], r_hi = ix[(size_t)(T - 1) * n + i1];
    const int c_lo = ix[j0], c_hi = ix[(size_t)(T - 1) * n + j1];
    const int nr = r_hi - r_lo + 1, nc = c_hi - c_lo + 1;
    if (nr <= LBOX && nc <= LBOX) {
      for (int q = tid; q < nr * LBOX; q += 256) {
        const int r = q / LBOX, c = q % LBOX;
        if (c < nc) box[r][c] = S[(size_t)(r_lo + r) * ns + c_lo + c];
      }
      __syncthreads();
      // rows: rows[i][c] = sum_a w_a(i) box[idx_a(i) - r_lo][c]
      for (int q = tid; q < LT * nc; q += 256) {
        const int i = q / nc, c = q % nc;
        const int gi = min(i0 + i, n - 1);
        float v = 0.f;
#pragma unroll
        for (int a = 0; a < T; ++a) v = fmaf(wl[(size_t)a * n + gi], box[ix[(size_t)a * n + gi] - r_lo][c], v);
        rows[i][c] = v;
      }
      __syncthreads();
      const int j = min(j0 + jl, n - 1);
      float wb[T];
      int cb[T];
#pragma unroll
      for (int b = 0; b < T; ++b) {
        wb[b] = wl[(size_t)b * n + j];
        cb[b] = ix[(size_t)b * n + j] - c_lo;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        float v = 0.f;
#pragma unroll
        for (int b = 0; b < T; ++b) v = fmaf(wb[b], rows[il + 8 * k][cb[b]], v);
        acc[k] += v;
      }
      __syncthreads();  // box and rows are refilled for the next layer
    } else {
      const int j = min(j0 + jl, n - 1);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int i = min(i0 + il + 8 * k, n - 1);
        float v = 0.f;
#pragma unroll
        for (int a = 0; a < T; ++a) {
          const float* row = S + (size_t)ix[(size_t)a * n + i] * ns;
          float r = 0.f;
#pragma unroll
          for (int b = 0; b < T; ++b) r = fmaf(wl[(size_t)b * n + j], row[ix[(size_t)b * n + j]], r);
          v = fmaf(wl[(size_t)a * n + i], r, v);
        }
        acc[k] += v;
      }
    }
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
