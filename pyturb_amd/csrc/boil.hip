// Spectral boiling for gfx950: per-mode AR(1) evolution of the stored layer
// spectra and subharmonic coefficients, fused with the frame layer sum.
//
// The reference boils after every frame (atmosphere.py:1092-1154): each
// Fourier mode of each layer relaxes toward a fresh draw of its stationary
// distribution,
//   S <- a S + sqrt(1 - a^2) (n0 + i n1) A,   a = exp(-dt / tau(f)),
//   tau(f) = tau_boil * max(|f| / f_ref, 1)^(-2/3)   (atmosphere.py:618-658),
// so frame k is built from the state after k steps and frames are no
// longer independent of each other.  k_boil_pairs walks the frames in
// order with the state in registers: a thread owns a mirror pair of modes
// (k, -k) of every layer, forms each frame's layer sum at both, takes the
// Hermitian part (frame = Re IFFT(X) = IFFT(Herm X)), packs frames 2p and
// 2p+1 into one complex grid Z_p = Herm X_2p + i Herm X_2p+1 and then
// applies the AR(1) step.  The stack is read and written once per call,
// not once per frame; the transforms of Z_p are the frame engine's own.
//
// Noise: either the caller's normals in the reference's order
// ((steps, 2, L, n, n) f32 and (steps, 2, P, L, 3, 3) f64, the draws of
// atmosphere.py:1114-1124), or a counter-based Philox4x32-10 stream
// (Salmon, Moraes, Dror & Shaw, SC 2011) keyed by the boiling seed and
// counted by (mode, layer, step), with the Box-Muller transform.  The
// counter makes the stream independent of how frames are chunked into
// calls.
#include <hip/hip_runtime.h>

#include "boil.hpp"
#include "common.hpp"
#include "fft_core.hpp"

namespace pt {

// ---------------------------------------------------------------------------
// Philox4x32-10 and Box-Muller
// ---------------------------------------------------------------------------
struct Ctr4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ Ctr4 philox4x32_10(Ctr4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;  // round multipliers
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;  // Weyl key increments
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32 x 32 -> 64-bit multiply each (v_mad_u64_u32) instead of a
    // separate high and low product
    const uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c = Ctr4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// The hardware sine and cosine take their argument in revolutions: on
// [0, 1) they are within 1.3e-7 of sin(2 pi x), cos(2 pi x) (measured on
// gfx950 against FP64 over 2^22 points), as are log (1.9e-7 relative) and
// square root (9.3e-8).  No range reduction is needed for either use here.
__device__ __forceinline__ float2 sincos_rev(float x) {
  return make_float2(__builtin_amdgcn_cosf(x), __builtin_amdgcn_sinf(x));
}

// exp(2 pi i phase 2^-32), phase in fixed-point cycles (24 significant bits)
__device__ __forceinline__ float2 phasor_rev(uint32_t phase) {
  return sincos_rev((float)(phase >> 8) * 5.9604644775390625e-08f);  // 2^-24
}

// Two standard normals from two 32-bit words: u1 in (0, 1], u2 in [0, 1).
__device__ __forceinline__ float2 box_muller(uint32_t a, uint32_t b) {
  const float u1 = ((float)a + 1.0f) * 2.3283064365386963e-10f;  // 2^-32
  const float u2 = (float)(b >> 8) * 5.9604644775390625e-08f;    // 2^-24
  const float r = __builtin_amdgcn_sqrtf(-2.0f * __logf(u1));
  const float2 cs = sincos_rev(u2);
  return make_float2(r * cs.x, r * cs.y);
}

// The same two normals in FP64 (subharmonic coefficients).
__device__ __forceinline__ double2 box_muller_d(uint32_t a, uint32_t b) {
  const double u1 = ((double)a + 1.0) * 2.3283064365386963e-10;
  const double u2 = (double)(b >> 8) * 5.9604644775390625e-08;
  const double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  return make_double2(r * c, r * s);
}

// Counter words: x = mode (main) or coefficient index (subharmonic), y =
// layer (main) or SH_TAG (subharmonic), (z, w) = the 64-bit step index.
constexpr uint32_t SH_TAG = 0x80000000u;

// ---------------------------------------------------------------------------
// per-frame phase table: fixed-point cycles of every layer's displacement
// ---------------------------------------------------------------------------
__global__ void k_boil_phase(const double* __restrict__ disp, int count, double inv_period, uint2* __restrict__ tab) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  tab[q] = make_uint2(cycles_fixed(disp[2 * (size_t)q] * inv_period), cycles_fixed(disp[2 * (size_t)q + 1] * inv_period));
}

// ---------------------------------------------------------------------------
// the fused boil + layer-sum kernel
// ---------------------------------------------------------------------------
// A quad of lanes (LP = 4 for 3+ layers) owns mode e = (k0, k1), k0 <= n/2,
// and its mirror m = (-k0, -k1) mod n when e < m, or e alone when e = m
// (the four self-conjugate modes); quads with e > m (half of rows 0 and
// n/2) idle.  Lane g of the quad holds layers g, g + LP, ...: each frame's
// layer sums are reduced across the quad with lane shuffles, so per-lane
// state is a quarter of the stack and four times as many waves hide the
// latency of the per-layer transcendental chain.  |f| and hence
// a = exp(-dt/tau) are equal at e and m.  Instantiated per layer count, so
// every per-layer array stays in registers.
template <int L>
__global__ __launch_bounds__(256) void k_boil_pairs(BoilArgs b) {
  constexpr int LP = L >= 3 ? 4 : L;         // lanes per mode pair
  constexpr int JB = (L + LP - 1) / LP;      // layers per lane
  const int n = b.n;
  const size_t plane = (size_t)n * n;
  const size_t tidx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tidx >= (size_t)(n / 2 + 1) * n * LP) return;  // whole quads
  const int g = (int)(tidx % LP);
  const size_t pidx = tidx / LP;
  const int k0 = (int)(pidx / n), k1 = (int)(pidx % n);
  const int mr0 = (n - k0) % n, m1 = (n - k1) % n;
  const size_t e = (size_t)k0 * n + k1, m = (size_t)mr0 * n + m1;
  if (e > m) return;  // the whole quad
  const bool self = e == m;
  const uint32_t k0s = (uint32_t)signed_freq(k0, n), k1s = (uint32_t)signed_freq(k1, n);
  const uint32_t m0s = (uint32_t)signed_freq(mr0, n), m1s = (uint32_t)signed_freq(m1, n);
  // off the Nyquist row and column the mirror's signed frequencies are the
  // negatives, so its phasor is the conjugate
  const bool nyq = 2 * k0 == n || 2 * k1 == n;
  const double fx = (double)(int32_t)k0s * b.df, fy = (double)(int32_t)k1s * b.df;
  const double fr = sqrt(fx * fx + fy * fy);
  float2 se[JB], sm[JB];
  float a[JB], ce[JB], cm[JB];
#pragma unroll
  for (int j = 0; j < JB; ++j) {
    const int l = g + LP * j;
    se[j] = sm[j] = make_float2(0.f, 0.f);
    a[j] = 1.f;
    ce[j] = cm[j] = 0.f;
    if (l < L) {
      se[j] = b.spectra[(size_t)l * plane + e];
      sm[j] = b.spectra[(size_t)l * plane + m];
      // a = exp(-dt / tau), 1/tau = inv_tau * max(fr / f_ref, 1)^(2/3); FP64 setup once per call
      const double ratio = fmax(fr / b.f_ref[l], 1.0);
      const double al = exp(-b.dt * b.inv_tau[l] * cbrt(ratio * ratio));
      const double bl = sqrt(fmax(1.0 - al * al, 0.0));
      a[j] = (float)al;
      ce[j] = (float)(bl * b.amp[(size_t)l * plane + e]);
      cm[j] = (float)(bl * b.amp[(size_t)l * plane + m]);
    }
  }
  float2 h1 = make_float2(0.f, 0.f);
  const uint32_t mode = (uint32_t)e;
  for (int s = 0; s < b.steps; ++s) {
    if (s < b.frames) {
      const uint2* ph = b.phase + (size_t)s * L;
      float2 xe = make_float2(0.f, 0.f), xm = make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 0; j < JB; ++j) {
        const int l = g + LP * j;
        if (l < L) {
          const uint2 u = ph[l];
          const float2 pe = phasor_rev(k0s * u.x + k1s * u.y);
          xe = cadd(xe, cmul(se[j], pe));
          const float2 pm = nyq ? phasor_rev(m0s * u.x + m1s * u.y) : make_float2(pe.x, -pe.y);
          xm = cadd(xm, cmul(sm[j], pm));
        }
      }
#pragma unroll
      for (int o = 1; o < LP; o <<= 1) {
        xe.x += __shfl_xor(xe.x, o, LP);
        xe.y += __shfl_xor(xe.y, o, LP);
        xm.x += __shfl_xor(xm.x, o, LP);
        xm.y += __shfl_xor(xm.y, o, LP);
      }
      if (self) xm = xe;
      // Herm X(e) = (X(e) + conj X(m)) / 2, Herm X(m) = conj Herm X(e)
      const float2 h = make_float2(0.5f * (xe.x + xm.x), 0.5f * (xe.y - xm.y));
      if ((s & 1) == 0) h1 = h;
      if (g == 0 && ((s & 1) == 1 || s == b.frames - 1)) {
        const float2 h2 = (s & 1) ? h : make_float2(0.f, 0.f);
        float2* z = b.z + (size_t)(s >> 1) * plane;
        z[e] = make_float2(h1.x - h2.y, h1.y + h2.x);  // h1 + i h2
        if (!self) z[m] = make_float2(h1.x + h2.y, h2.x - h1.y);  // conj h1 + i conj h2
      }
    }
    // AR(1) step of both modes of this lane's layers
    const unsigned long long gs = b.step0 + (unsigned long long)s;
#pragma unroll
    for (int j = 0; j < JB; ++j) {
      const int l = g + LP * j;
      if (l >= L || (ce[j] == 0.f && cm[j] == 0.f && a[j] == 1.f)) continue;  // frozen layer
      float2 ne, nm;
      if (b.noise) {
        const float* nz = b.noise + (size_t)s * 2 * L * plane + (size_t)l * plane;
        ne = make_float2(nz[e], nz[(size_t)L * plane + e]);
        nm = make_float2(nz[m], nz[(size_t)L * plane + m]);
      } else {
        const Ctr4 r = philox4x32_10(Ctr4{mode, (uint32_t)l, (uint32_t)gs, (uint32_t)(gs >> 32)}, b.key0, b.key1);
        ne = box_muller(r.x, r.y);
        nm = box_muller(r.z, r.w);
      }
      se[j] = make_float2(fmaf(a[j], se[j].x, ce[j] * ne.x), fmaf(a[j], se[j].y, ce[j] * ne.y));
      if (self)
        sm[j] = se[j];
      else
        sm[j] = make_float2(fmaf(a[j], sm[j].x, cm[j] * nm.x), fmaf(a[j], sm[j].y, cm[j] * nm.y));
    }
  }
#pragma unroll
  for (int j = 0; j < JB; ++j) {
    const int l = g + LP * j;
    if (l >= L) continue;
    b.spectra[(size_t)l * plane + e] = se[j];
    if (!self) b.spectra[(size_t)l * plane + m] = sm[j];
  }
}

// ---------------------------------------------------------------------------
// subharmonic coefficients: the same AR(1) per 3x3 cell, in FP64
// ---------------------------------------------------------------------------
// Level p (0-based) has cell frequencies (i - 1, j - 1) * df / 3^(p+1)
// (fourier.py:203-205); the coefficients of frame s (< frames) are written
// to cf[s] before step s, the state after `steps` steps back to coeffs.
__global__ void k_boil_coeffs(BoilArgs b) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int count = b.P * b.L * 9;
  if (q >= count) return;
  const int p = q / (b.L * 9), l = (q / 9) % b.L, cell = q % 9;
  const double fp = b.df / pow3(p + 1);
  const double fx = (double)(cell / 3 - 1) * fp, fy = (double)(cell % 3 - 1) * fp;
  const double ratio = fmax(sqrt(fx * fx + fy * fy) / b.f_ref[l], 1.0);
  const double a = exp(-b.dt * b.inv_tau[l] * cbrt(ratio * ratio));
  const double c = sqrt(fmax(1.0 - a * a, 0.0)) * b.sh_amp[q];
  double2 v = b.coeffs[q];
  for (int s = 0; s < b.steps; ++s) {
    if (s < b.frames) b.coeff_frames[(size_t)s * count + q] = v;
    if (c == 0.0 && a == 1.0) continue;
    double2 nz;
    if (b.sh_noise) {
      const double* w = b.sh_noise + (size_t)s * 2 * count;
      nz = make_double2(w[q], w[count + q]);
    } else {
      const unsigned long long g = b.step0 + (unsigned long long)s;
      const Ctr4 r = philox4x32_10(Ctr4{(uint32_t)q, SH_TAG, (uint32_t)g, (uint32_t)(g >> 32)}, b.key0, b.key1);
      nz = box_muller_d(r.x, r.y);
    }
    v = make_double2(a * v.x + c * nz.x, a * v.y + c * nz.y);
  }
  b.coeffs[q] = v;
}

// ---------------------------------------------------------------------------
// plain normals (extruder boiling): element i is word pair (i % 4) / 2,
// component i % 2, of Philox call (i / 4, stream)
// ---------------------------------------------------------------------------
__global__ void k_normals(float* __restrict__ out, long long count, uint32_t key0, uint32_t key1,
                          unsigned long long stream, float scale) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (4 * c >= count) return;
  const Ctr4 r = philox4x32_10(Ctr4{(uint32_t)c, (uint32_t)((unsigned long long)c >> 32), (uint32_t)stream,
                                    (uint32_t)(stream >> 32)},
                               key0, key1);
  const float2 p = box_muller(r.x, r.y), q = box_muller(r.z, r.w);
  const float v[4] = {p.x, p.y, q.x, q.y};
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (4 * c + k < count) out[4 * c + k] = scale * v[k];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int launch_boil(const BoilArgs& b, const double* disp, double inv_period, hipStream_t st) {
  if (b.frames > 0) {
    const int count = b.frames * b.L;
    k_boil_phase<<<(count + 255) / 256, 256, 0, st>>>(disp, count, inv_period, (uint2*)b.phase);
    PT_LAUNCH_CHECK();
  }
  const size_t threads = (size_t)(b.n / 2 + 1) * b.n * (b.L >= 3 ? 4 : b.L);  // lanes per mode pair
  const int blocks = (int)((threads + 255) / 256);
  switch (b.L) {
#define PT_BOIL_CASE(L) \
  case L:               \
    k_boil_pairs<L><<<blocks, 256, 0, st>>>(b); \
    break;
    PT_BOIL_CASE(1) PT_BOIL_CASE(2) PT_BOIL_CASE(3) PT_BOIL_CASE(4) PT_BOIL_CASE(5) PT_BOIL_CASE(6)
    PT_BOIL_CASE(7) PT_BOIL_CASE(8) PT_BOIL_CASE(9) PT_BOIL_CASE(10) PT_BOIL_CASE(11) PT_BOIL_CASE(12)
    PT_BOIL_CASE(13) PT_BOIL_CASE(14) PT_BOIL_CASE(15) PT_BOIL_CASE(16)
#undef PT_BOIL_CASE
    default:
      return fail(PT_ERR_UNSUPPORTED, "boiling: %d layers (at most 16)", b.L);
  }
  PT_LAUNCH_CHECK();
  if (b.P > 0) {
    const int count = b.P * b.L * 9;
    k_boil_coeffs<<<(count + 63) / 64, 64, 0, st>>>(b);
    PT_LAUNCH_CHECK();
  }
  return PT_OK;
}

}  // namespace pt

using namespace pt;

extern "C" int pt_normals(float* out, long long count, unsigned long long seed, unsigned long long stream_id,
                          float scale, void* stream) {
  if (!out || count < 0) return fail(PT_ERR_ARG, "pt_normals: bad arguments");
  if (count == 0) return PT_OK;
  const long long calls = (count + 3) / 4;
  k_normals<<<(unsigned)((calls + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      out, count, (uint32_t)seed, (uint32_t)(seed >> 32), stream_id, scale);
  PT_LAUNCH_CHECK();
  return PT_OK;
}
