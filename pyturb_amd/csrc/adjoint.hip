// Adjoint of the spectral frozen-flow engine (the frame model of
// src/pyturb/atmosphere.py:882-938 with the layer spectra of
// src/pyturb/atmosphere.py:532-579 written as gain_l * amp_{L0_l}(k) * w_l(k)).
//
// Forward, per frame t and layer l (s = displacement in metres):
//   OPD_t = scale * ( Re(sum_k X_t(k) e^{+2 pi i k.x / n}) + low_t ),
//   X_t(k) = sum_l gain_l * amp_l(k) * w_l(k) * exp(2 pi i (fx s_x + fy s_y)),
//   low_t  = Re(sum_{p,a,b} Q_pab B_pab(x, y)) - mean,  Q_pab = sum_l c_plab q_pla q_plb.
// For a real cotangent G_t = dLoss/dOPD_t, with Ghat_t = sum_x G_t(x) e^{+2 pi i k.x/n}
// (the same unnormalised +i transform) and Gamma_t,pab = sum_x (G_t - mean G_t) B_pab:
//   dLoss/dgain_l  = scale * Re sum_k U_l P Ghat          + sub terms,
//   dLoss/ds_x     = scale * gain_l * Re sum_k 2 pi i fx U_l P Ghat + ...,
//   dLoss/dL0_l    = scale * gain_l * Re sum_k rho_l U_l P Ghat     + ...,
// where U_l = amp_l * w_l is the unit-gain spectrum, recomputed here from
// the white noise (never stored), and rho_l = d ln amp_l / d L0.
//
// Steps (one stream):
//   1. k_adj_pack   : frames 2m, 2m+1 packed as Z_m = G_2m + i G_2m+1, plus the
//                     per-row subharmonic projections h of both frames;
//   2. pt_fft2(+1)  : Zhat_m in place; Ghat of each frame is the Hermitian /
//                     anti-Hermitian part of Zhat_m;
//   3. k_adj_phasor : per (frame, layer) shift phasor tables (exact fixed-point
//                     phase, identical to the forward pass);
//   4. k_adj_reduce : per (frame pair, pixel chunk, layer) the four sums
//                     U P Ghat * {1, fx, fy, rho}, block-reduced to partials;
//   5. k_adj_gamma  : Gamma_t from h (FP64);
//   6. k_adj_final  : FP64 sum of partials + subharmonic terms -> (T, L, 4).

#include "common.hpp"
#include "fft_core.hpp"
#include "sh_epilogue.hpp"

namespace pt {

constexpr int ADJ_LMAX = 16;   // layers per call
constexpr int ADJ_PPT = 8;     // pixels per thread in k_adj_reduce
constexpr int ADJ_CH = 256 * ADJ_PPT;

struct LayerMap {
  float gain[ADJ_LMAX];
  int group[ADJ_LMAX];
};

// ---------------------------------------------------------------------------
// forward helper: spectra[l] = gain_l * amp[group_l] * (w_re + i w_im)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_layer_spectra(const float* __restrict__ noise,
                                                       const float* __restrict__ amp, int n, int L,
                                                       LayerMap map, float2* __restrict__ out) {
  const size_t plane = (size_t)n * n;
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int l = blockIdx.y;
  if (idx >= plane || l >= L) return;
  const float a = map.gain[l] * amp[(size_t)map.group[l] * plane + idx];
  const float wr = noise[(size_t)(2 * l) * plane + idx];
  const float wi = noise[(size_t)(2 * l + 1) * plane + idx];
  out[(size_t)l * plane + idx] = make_float2(a * wr, a * wi);
}

// ---------------------------------------------------------------------------
// wave / block sums (64-lane wavefronts, 256-thread blocks)
// ---------------------------------------------------------------------------
template <class V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Sums NV values over the block; the totals land in red[0..NV).  red holds
// 4 * NV entries.  Ends with a barrier.
template <int NV, class V>
__device__ __forceinline__ void block_sum(V (&v)[NV], V* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const V s = wave_sum(v[q]);
    if (lane == 0) red[wave * NV + q] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int q = threadIdx.x;
    red[q] = red[q] + red[NV + q] + red[2 * NV + q] + red[3 * NV + q];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// 1. packing + per-row subharmonic projections
//    h[m][x] = [sum_y G_0, sum_y G_1, (sum_y G_0 e_p(y), sum_y G_1 e_p(y)) for p < P]
//    (floats: 2 + 4P).  One wave per row.
// ---------------------------------------------------------------------------
struct PackArgs {
  const float* grad;   // (T, n_out, n_out)
  float2* Z;           // (pairs, n, n)
  float* h;            // (pairs, n, 2 + 4P)
  const float2* E;     // (P, n) e_p(y)
  int n, T, P, off, n_out;
};

__global__ __launch_bounds__(256) void k_adj_pack(PackArgs a) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int x = blockIdx.x * 4 + wave;
  const int m = blockIdx.y;
  if (x >= a.n) return;
  const int t0 = 2 * m, t1 = 2 * m + 1;
  const int xr = x - a.off;
  const bool rowin = xr >= 0 && xr < a.n_out;
  const size_t img = (size_t)a.n_out * a.n_out;
  float r0 = 0.f, r1 = 0.f;
  float hr0[PREG], hi0[PREG], hr1[PREG], hi1[PREG];
#pragma unroll
  for (int p = 0; p < PREG; ++p) hr0[p] = hi0[p] = hr1[p] = hi1[p] = 0.f;
  float2* zrow = a.Z + ((size_t)m * a.n + x) * a.n;
  for (int y = lane; y < a.n; y += 64) {
    const int yr = y - a.off;
    float g0 = 0.f, g1 = 0.f;
    if (rowin && yr >= 0 && yr < a.n_out) {
      const size_t o = (size_t)xr * a.n_out + yr;
      g0 = a.grad[(size_t)t0 * img + o];
      if (t1 < a.T) g1 = a.grad[(size_t)t1 * img + o];
    }
    zrow[y] = make_float2(g0, g1);
    r0 += g0;
    r1 += g1;
#pragma unroll
    for (int p = 0; p < PREG; ++p) {
      if (p >= a.P) break;
      const float2 e = a.E[(size_t)p * a.n + y];
      hr0[p] = fmaf(g0, e.x, hr0[p]);
      hi0[p] = fmaf(g0, e.y, hi0[p]);
      hr1[p] = fmaf(g1, e.x, hr1[p]);
      hi1[p] = fmaf(g1, e.y, hi1[p]);
    }
  }
  float* hrow = a.h + ((size_t)m * a.n + x) * (2 + 4 * a.P);
  r0 = wave_sum(r0);
  r1 = wave_sum(r1);
  if (lane == 0) {
    hrow[0] = r0;
    hrow[1] = r1;
  }
#pragma unroll
  for (int p = 0; p < PREG; ++p) {
    if (p >= a.P) break;
    const float v0 = wave_sum(hr0[p]), v1 = wave_sum(hi0[p]);
    const float v2 = wave_sum(hr1[p]), v3 = wave_sum(hi1[p]);
    if (lane == 0) {
      hrow[2 + 4 * p + 0] = v0;
      hrow[2 + 4 * p + 1] = v1;
      hrow[2 + 4 * p + 2] = v2;
      hrow[2 + 4 * p + 3] = v3;
    }
  }
}

// basis table E[p][y] = e_p(y)
__global__ __launch_bounds__(256) void k_adj_basis(float2* E, int P, int n) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx < P * n) E[idx] = sh_basis_e(idx / n, idx % n, n);
}

// ---------------------------------------------------------------------------
// 3. phasor tables px[t][l][i], py[t][l][j]: the forward's exact phases
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_phasor(const double* __restrict__ disp, int T, int L, int n,
                                                    double inv_period, float2* px, float2* py) {
  const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (size_t)T * L * n) return;
  const int i = (int)(idx % n);
  const size_t tl = idx / n;
  const uint32_t ux = cycles_fixed(disp[tl * 2 + 0] * inv_period);
  const uint32_t uy = cycles_fixed(disp[tl * 2 + 1] * inv_period);
  const uint32_t k = (uint32_t)signed_freq(i, n);
  px[idx] = phasor_fixed(k * ux);
  py[idx] = phasor_fixed(k * uy);
}

// ---------------------------------------------------------------------------
// 4. spectral sums.  Block (chunk, pair m); thread pixels k = chunk*CH + e*256 + tid.
//    partial[((m * chunks + chunk) * L + l) * 16 + f * 8 + q], q: re/im of
//    {U P Ghat, fx U P Ghat, fy U P Ghat, rho U P Ghat} for frame 2m + f.
// ---------------------------------------------------------------------------
struct ReduceArgs {
  const float* noise;   // (L, 2, n, n)
  const float* amp;     // (G, n, n)
  const float* dlnamp;  // (G, n, n)
  const float2* Zhat;   // (pairs, n, n)
  const float2* px;     // (T, L, n)
  const float2* py;
  float* partial;
  LayerMap map;
  int n, log2n, L, T, chunks;
  float inv_period;
};

__global__ __launch_bounds__(256) void k_adj_reduce(ReduceArgs a) {
  __shared__ float red[4 * 16];
  const int tid = threadIdx.x;
  const int chunk = blockIdx.x, m = blockIdx.y;
  const int n = a.n;
  const size_t plane = (size_t)n * n;
  const int t0 = 2 * m;
  const bool has1 = t0 + 1 < a.T;
  float2 g0[ADJ_PPT], g1[ADJ_PPT];
  int ii[ADJ_PPT], jj[ADJ_PPT];
  bool ok[ADJ_PPT];
  const float2* Z = a.Zhat + (size_t)m * plane;
#pragma unroll
  for (int e = 0; e < ADJ_PPT; ++e) {
    const size_t k = (size_t)chunk * ADJ_CH + e * 256 + tid;
    ok[e] = k < plane;
    const int i = ok[e] ? (int)(k >> a.log2n) : 0;
    const int j = ok[e] ? (int)(k & (n - 1)) : 0;
    ii[e] = i;
    jj[e] = j;
    const float2 z = Z[(size_t)i * n + j];
    const float2 zm = Z[(size_t)((n - i) & (n - 1)) * n + ((n - j) & (n - 1))];
    // Ghat_0 = (Z + conj(Zm)) / 2,  Ghat_1 = (Z - conj(Zm)) / (2i)
    g0[e] = ok[e] ? make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y)) : make_float2(0.f, 0.f);
    g1[e] = ok[e] ? make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x)) : make_float2(0.f, 0.f);
  }
  for (int l = 0; l < a.L; ++l) {
    const int g = a.map.group[l];
    const float* wr = a.noise + (size_t)(2 * l) * plane;
    const float* wi = wr + plane;
    const float* am = a.amp + (size_t)g * plane;
    const float* rh = a.dlnamp + (size_t)g * plane;
    const float2* px0 = a.px + ((size_t)t0 * a.L + l) * n;
    const float2* py0 = a.py + ((size_t)t0 * a.L + l) * n;
    const float2* px1 = px0 + (size_t)a.L * n;
    const float2* py1 = py0 + (size_t)a.L * n;
    float acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.f;
#pragma unroll
    for (int e = 0; e < ADJ_PPT; ++e) {
      if (!ok[e]) continue;
      const size_t k = (size_t)ii[e] * n + jj[e];
      const float A = am[k], rho = rh[k];
      const float2 U = make_float2(A * wr[k], A * wi[k]);
      const float fx = (float)signed_freq(ii[e], n) * a.inv_period;
      const float fy = (float)signed_freq(jj[e], n) * a.inv_period;
      {
        const float2 P = cmul(px0[ii[e]], py0[jj[e]]);
        const float2 V = cmul(cmul(U, P), g0[e]);
        acc[0] += V.x;
        acc[1] += V.y;
        acc[2] = fmaf(fx, V.x, acc[2]);
        acc[3] = fmaf(fx, V.y, acc[3]);
        acc[4] = fmaf(fy, V.x, acc[4]);
        acc[5] = fmaf(fy, V.y, acc[5]);
        acc[6] = fmaf(rho, V.x, acc[6]);
        acc[7] = fmaf(rho, V.y, acc[7]);
      }
      if (has1) {
        const float2 P = cmul(px1[ii[e]], py1[jj[e]]);
        const float2 V = cmul(cmul(U, P), g1[e]);
        acc[8] += V.x;
        acc[9] += V.y;
        acc[10] = fmaf(fx, V.x, acc[10]);
        acc[11] = fmaf(fx, V.y, acc[11]);
        acc[12] = fmaf(fy, V.x, acc[12]);
        acc[13] = fmaf(fy, V.y, acc[13]);
        acc[14] = fmaf(rho, V.x, acc[14]);
        acc[15] = fmaf(rho, V.y, acc[15]);
      }
    }
    block_sum<16>(acc, red);
    if (tid < 16) a.partial[(((size_t)m * a.chunks + chunk) * a.L + l) * 16 + tid] = red[tid];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 5. Gamma[t][p][a][b] = sum_x e_pa(x) h_pb(x) - (sum G) m_pa m_pb   (FP64)
//    with e_p0 = conj(e_p), e_p1 = 1, e_p2 = e_p and likewise h.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_gamma(const float* __restrict__ h, const float2* __restrict__ E, int n,
                                                   int T, int P, double2* __restrict__ gam) {
  __shared__ double red[4 * 18];
  __shared__ double total;
  const int t = blockIdx.x;
  const int m = t >> 1, f = t & 1;
  const int stride = 2 + 4 * P;
  const float* hb = h + (size_t)m * n * stride;
  {
    double v[1] = {0.0};
    for (int x = threadIdx.x; x < n; x += 256) v[0] += (double)hb[(size_t)x * stride + f];
    double one[1] = {v[0]};
    block_sum<1>(one, red);
    if (threadIdx.x == 0) total = red[0];
    __syncthreads();
  }
  for (int p = 0; p < P; ++p) {
    double acc[18];
#pragma unroll
    for (int q = 0; q < 18; ++q) acc[q] = 0.0;
    for (int x = threadIdx.x; x < n; x += 256) {
      const float* hr = hb + (size_t)x * stride;
      const double2 H = make_double2(hr[2 + 4 * p + 2 * f], hr[2 + 4 * p + 2 * f + 1]);
      const double2 hv[3] = {dconj(H), make_double2(hr[f], 0.0), H};
      const float2 ef = E[(size_t)p * n + x];
      const double2 ev[3] = {make_double2(ef.x, -ef.y), make_double2(1.0, 0.0), make_double2(ef.x, ef.y)};
#pragma unroll
      for (int ai = 0; ai < 3; ++ai)
#pragma unroll
        for (int bi = 0; bi < 3; ++bi) {
          const double2 w = dcmul(ev[ai], hv[bi]);
          acc[2 * (ai * 3 + bi)] += w.x;
          acc[2 * (ai * 3 + bi) + 1] += w.y;
        }
    }
    block_sum<18>(acc, red);
    if (threadIdx.x < 9) {
      const int ai = threadIdx.x / 3, bi = threadIdx.x % 3;
      const double2 mp = sh_mean_e(p, n);
      const double2 mv[3] = {dconj(mp), make_double2(1.0, 0.0), mp};
      const double2 mm = dcmul(mv[ai], mv[bi]);
      gam[((size_t)t * P + p) * 9 + threadIdx.x] =
          make_double2(red[2 * threadIdx.x] - total * mm.x, red[2 * threadIdx.x + 1] - total * mm.y);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 6. per (frame, layer): FP64 sum of the partials, subharmonic terms, output
//    out[(t * L + l) * 4 + {0: dgain, 1: ds_x, 2: ds_y, 3: dL0}]
// ---------------------------------------------------------------------------
struct FinalArgs {
  const float* partial;
  const double2* gam;      // (T, P, 9)
  const double2* sh_unit;  // (P, L, 9) unit-gain coefficients
  const double* sh_rho;    // (P, L, 9) d ln amp / d L0 of the coefficients
  const double* disp;      // (T, L, 2)
  double* out;
  double gain[ADJ_LMAX];
  int T, L, P, n, chunks;
  double inv_period, scale;
};

__global__ __launch_bounds__(64) void k_adj_final(FinalArgs a) {
  const int idx = blockIdx.x * 64 + threadIdx.x;
  if (idx >= a.T * a.L) return;
  const int t = idx / a.L, l = idx % a.L;
  const int m = t >> 1, f = t & 1;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int c = 0; c < a.chunks; ++c) {
    const float* pp = a.partial + (((size_t)m * a.chunks + c) * a.L + l) * 16 + f * 8;
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] += (double)pp[q];
  }
  const double g = a.gain[l];
  const double two_pi = 6.283185307179586;
  double dg = s[0], dsx = -two_pi * g * s[3], dsy = -two_pi * g * s[5], dl0 = g * s[6];
  if (a.P > 0) {
    const double sx = a.disp[((size_t)t * a.L + l) * 2 + 0];
    const double sy = a.disp[((size_t)t * a.L + l) * 2 + 1];
    for (int p = 0; p < a.P; ++p) {
      const double dfp = a.inv_period / pow3(p + 1);
      double2 qx[3], qy[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double sn, cs;
        sincospi(2.0 * sx * dfp * (k - 1), &sn, &cs);
        qx[k] = make_double2(cs, sn);
        sincospi(2.0 * sy * dfp * (k - 1), &sn, &cs);
        qy[k] = make_double2(cs, sn);
      }
      for (int ab = 0; ab < 9; ++ab) {
        const int ai = ab / 3, bi = ab % 3;
        const size_t ci = ((size_t)p * a.L + l) * 9 + ab;
        const double2 w = dcmul(dcmul(dcmul(a.sh_unit[ci], qx[ai]), qy[bi]), a.gam[((size_t)t * a.P + p) * 9 + ab]);
        dg += w.x;
        dsx += -two_pi * g * dfp * (ai - 1) * w.y;
        dsy += -two_pi * g * dfp * (bi - 1) * w.y;
        dl0 += g * a.sh_rho[ci] * w.x;
      }
    }
  }
  double* o = a.out + (size_t)idx * 4;
  o[0] = a.scale * dg;
  o[1] = a.scale * dsx;
  o[2] = a.scale * dsy;
  o[3] = a.scale * dl0;
}

// ---------------------------------------------------------------------------
// workspace layout (c64 units, each section rounded to 32 elements)
// ---------------------------------------------------------------------------
static inline size_t r32(size_t v) { return (v + 31) & ~(size_t)31; }

struct AdjLayout {
  size_t Z, px, py, E, h, gam, partial, total;
};

static AdjLayout adj_layout(int n, int T, int P) {
  const size_t pairs = (size_t)(T + 1) / 2;
  const size_t chunks = ((size_t)n * n + ADJ_CH - 1) / ADJ_CH;
  AdjLayout w;
  size_t o = 0;
  w.Z = o;
  o += r32(pairs * n * n);
  w.px = o;
  o += r32((size_t)T * ADJ_LMAX * n);
  w.py = o;
  o += r32((size_t)T * ADJ_LMAX * n);
  w.E = o;
  o += r32((size_t)P * n);
  w.h = o;
  o += r32((pairs * n * (2 + 4 * (size_t)P) + 1) / 2);
  w.gam = o;
  o += r32((size_t)T * P * 9 * 2);
  w.partial = o;
  o += r32(pairs * chunks * ADJ_LMAX * 8);
  w.total = o;
  return w;
}

}  // namespace pt

using namespace pt;

extern "C" long long pt_adjoint_work_elems(int n, int frames, int levels) {
  if (n < 1 || frames < 0 || levels < 0) return -1;
  return (long long)adj_layout(n, frames, levels).total;
}

static bool adj_pow2(int n) { return n >= 32 && n <= 2048 && (n & (n - 1)) == 0; }

extern "C" int pt_layer_spectra(const float* noise, const float* amplitude, const int* group, const double* gain,
                                int layers, int n, void* spectra, void* stream) {
  if (!noise || !amplitude || !group || !gain || !spectra) return fail(PT_ERR_ARG, "pt_layer_spectra: NULL array");
  if (layers < 1 || layers > ADJ_LMAX || n < 1) return fail(PT_ERR_ARG, "pt_layer_spectra: 1..%d layers", ADJ_LMAX);
  LayerMap map;
  for (int l = 0; l < ADJ_LMAX; ++l) {
    const int src = l < layers ? l : 0;
    if (group[src] < 0) return fail(PT_ERR_ARG, "pt_layer_spectra: negative group index");
    map.gain[l] = (float)gain[src];
    map.group[l] = group[src];
  }
  const size_t plane = (size_t)n * n;
  dim3 g((unsigned)((plane + 255) / 256), layers);
  k_layer_spectra<<<g, 256, 0, (hipStream_t)stream>>>(noise, amplitude, n, layers, map, (float2*)spectra);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

extern "C" int pt_spectral_frames_adjoint(const float* noise, const float* amplitude, const float* dlnamp,
                                          const int* group, const double* gain, int layers, int n,
                                          double pixel_scale, const double* disp, int frames, const float* grad,
                                          int crop_offset, int n_out, const double* sh_unit, const double* sh_rho,
                                          int levels, double scale, double* out, void* work, void* stream) {
  if (!noise || !amplitude || !dlnamp || !group || !gain || !disp || !grad || !out || !work)
    return fail(PT_ERR_ARG, "pt_spectral_frames_adjoint: NULL array");
  if (layers < 1 || layers > ADJ_LMAX) return fail(PT_ERR_ARG, "pt_spectral_frames_adjoint: 1..%d layers", ADJ_LMAX);
  if (frames < 0 || frames > 65535) return fail(PT_ERR_ARG, "pt_spectral_frames_adjoint: 0..65535 frames");
  if (levels < 0 || (levels > 0 && (!sh_unit || !sh_rho))) return fail(PT_ERR_ARG, "pt_spectral_frames_adjoint: bad subharmonics");
  if (levels > PREG) return fail(PT_ERR_UNSUPPORTED, "pt_spectral_frames_adjoint: more than %d levels", PREG);
  if (!(pixel_scale > 0.0)) return fail(PT_ERR_ARG, "pt_spectral_frames_adjoint: pixel_scale must be > 0");
  if (crop_offset < 0 || n_out < 1 || crop_offset + n_out > n)
    return fail(PT_ERR_ARG, "pt_spectral_frames_adjoint: crop outside grid");
  if (!adj_pow2(n)) return fail(PT_ERR_UNSUPPORTED, "pt_spectral_frames_adjoint: power-of-two n 32..2048 only");
  if (frames == 0) return PT_OK;
  hipStream_t st = (hipStream_t)stream;
  const AdjLayout w = adj_layout(n, frames, levels);
  float2* base = (float2*)work;
  float2* Z = base + w.Z;
  float2* px = base + w.px;
  float2* py = base + w.py;
  float2* E = base + w.E;
  float* h = (float*)(base + w.h);
  double2* gam = (double2*)(base + w.gam);
  float* partial = (float*)(base + w.partial);
  const int pairs = (frames + 1) / 2;
  const int chunks = (int)(((size_t)n * n + ADJ_CH - 1) / ADJ_CH);
  const double inv_period = 1.0 / ((double)n * pixel_scale);

  LayerMap map;
  for (int l = 0; l < ADJ_LMAX; ++l) {
    const int src = l < layers ? l : 0;
    if (group[src] < 0) return fail(PT_ERR_ARG, "pt_spectral_frames_adjoint: negative group index");
    map.gain[l] = 1.0f;  // the sums use unit-gain spectra
    map.group[l] = group[src];
  }
  if (levels > 0) {
    k_adj_basis<<<(levels * n + 255) / 256, 256, 0, st>>>(E, levels, n);
    PT_LAUNCH_CHECK();
  }
  PackArgs pa;
  pa.grad = grad;
  pa.Z = Z;
  pa.h = h;
  pa.E = E;
  pa.n = n;
  pa.T = frames;
  pa.P = levels;
  pa.off = crop_offset;
  pa.n_out = n_out;
  k_adj_pack<<<dim3((n + 3) / 4, pairs), 256, 0, st>>>(pa);
  PT_LAUNCH_CHECK();
  int rc = pt_fft2(Z, Z, nullptr, pairs, n, n, 1, stream);
  if (rc) return rc;
  const size_t ntab = (size_t)frames * layers * n;
  k_adj_phasor<<<(unsigned)((ntab + 255) / 256), 256, 0, st>>>(disp, frames, layers, n, inv_period, px, py);
  PT_LAUNCH_CHECK();
  ReduceArgs ra;
  ra.noise = noise;
  ra.amp = amplitude;
  ra.dlnamp = dlnamp;
  ra.Zhat = Z;
  ra.px = px;
  ra.py = py;
  ra.partial = partial;
  ra.map = map;
  ra.n = n;
  ra.log2n = 0;
  while ((1 << ra.log2n) < n) ++ra.log2n;
  ra.L = layers;
  ra.T = frames;
  ra.chunks = chunks;
  ra.inv_period = (float)inv_period;
  k_adj_reduce<<<dim3(chunks, pairs), 256, 0, st>>>(ra);
  PT_LAUNCH_CHECK();
  if (levels > 0) {
    k_adj_gamma<<<frames, 256, 0, st>>>(h, E, n, frames, levels, gam);
    PT_LAUNCH_CHECK();
  }
  FinalArgs fa;
  fa.partial = partial;
  fa.gam = gam;
  fa.sh_unit = (const double2*)sh_unit;
  fa.sh_rho = sh_rho;
  fa.disp = disp;
  fa.out = out;
  for (int l = 0; l < ADJ_LMAX; ++l) fa.gain[l] = gain[l < layers ? l : 0];
  fa.T = frames;
  fa.L = layers;
  fa.P = levels;
  fa.n = n;
  fa.chunks = chunks;
  fa.inv_period = inv_period;
  fa.scale = scale;
  k_adj_final<<<(frames * layers + 63) / 64, 64, 0, st>>>(fa);
  PT_LAUNCH_CHECK();
  return PT_OK;
}
