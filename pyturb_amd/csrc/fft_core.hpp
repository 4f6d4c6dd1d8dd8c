// LDS-resident Stockham FFT engine for gfx950 (CDNA4, wave64).
//
// One workgroup transforms M independent complex sequences of length N
// (N a power of two, 32 <= N <= 2048).  Every thread owns exactly eight
// complex values per stage: one radix-8 butterfly, two radix-4 or four
// radix-2 butterflies, so M * N == 8 * T for a T-thread workgroup.  The
// plan's largest radix RX is 8 (the default: log2(N)/3 LDS round trips) or
// 4 (log2(N)/2 round trips of simpler butterflies), a tuning choice.
//
// The first stage reads its inputs through a caller-supplied prologue
// functor straight from HBM (so layer sums, PSD filtering and noise
// colouring fuse into the load), and the last stage hands its outputs to an
// epilogue functor that writes HBM (so real/imag split, subharmonic add,
// scaling and cropping fuse into the store).  Intermediate stages exchange
// through LDS in place: read, barrier, write, barrier.
//
// Two data layouts are supported:
//   ROWS  (COLS=false): sequence m occupies LDS row m (padded); butterfly b of a
//         stage maps to (m = b / (N/R), j = b % (N/R)), so consecutive lanes
//         walk along a sequence -> coalesced row access in HBM.
//   COLS  (COLS=true):  element i of sequence m sits at LDS[row(i)*M + m];
//         butterfly b maps to (m = b % M, j = b / M), so consecutive lanes
//         walk across M adjacent columns -> M*8-byte segments in HBM.
//
// Stockham autosort step (radix R, stride Ns = product of earlier radices):
//   v[r] = x[j + r*N/R] * W_{Ns*R}^{(j mod Ns)*r},  v = DFT_R(v),
//   y[(j/Ns)*Ns*R + (j mod Ns) + r*Ns] = v[r].
// Twiddles come from an N-entry table tw[t] = exp(S*2*pi*i*t/N) in LDS.
// S = +1 is the unnormalised inverse DFT (the reference's ifft2 * n^2).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// Streaming stores for data that no kernel of the same pass reads again (a
// pass's output, re-read by the next pass from HBM anyway): non-temporal, so
// they do not evict the lines the pass itself re-reads from L2.
__device__ __forceinline__ void store_nt(float2* p, float2 v) {
  typedef float v2f __attribute__((ext_vector_type(2)));
  v2f w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, (v2f*)p);
}
__device__ __forceinline__ void store_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by S*i
template <int S>
__device__ __forceinline__ float2 mul_si(float2 a) {
  return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int S>
__device__ __forceinline__ void dft2(float2& a0, float2& a1) {
  float2 t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

template <int S>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_si<S>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <int S>
__device__ __forceinline__ void dft8(float2* v) {
  // even/odd split into two 4-point DFTs, recombined with W_8^k
  dft4<S>(v[0], v[2], v[4], v[6]);
  dft4<S>(v[1], v[3], v[5], v[7]);
  const float c = 0.70710678118654752440f;
  float2 o1 = make_float2(c * (v[3].x - S * v[3].y), c * (S * v[3].x + v[3].y));
  float2 o2 = mul_si<S>(v[5]);
  float2 o3 = make_float2(c * (-v[7].x - S * v[7].y), c * (S * v[7].x - v[7].y));
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

template <int R, int S>
__device__ __forceinline__ void dftR(float2* v) {
  if constexpr (R == 8) {
    dft8<S>(v);
  } else if constexpr (R == 4) {
    dft4<S>(v[0], v[1], v[2], v[3]);
  } else {
    dft2<S>(v[0], v[1]);
  }
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n >> 1); }

// Stage plan: radix-RX stages plus one smaller stage when log2 N is not a
// multiple of log2 RX (RX = 8: one radix-4 or radix-2 stage; RX = 4: one
// radix-2 stage).  The small stage is last by default; SF (small first) puts
// it first, so that with RX = 8 the last stage is radix 8 for every N and a
// thread's eight outputs lie in one sequence (what a per-row epilogue needs).
template <int N, bool SF = false, int RX = 8>
struct Plan {
  static_assert((N & (N - 1)) == 0 && N >= 32 && N <= 2048, "power-of-two 32..2048");
  static_assert(RX == 8 || RX == 4, "largest radix 8 or 4");
  static constexpr int LOG = ilog2(N);
  static constexpr int LR = ilog2(RX);
  static constexpr int NFULL = LOG / LR;
  static constexpr int REM = LOG % LR;
  static constexpr int NST = NFULL + (REM ? 1 : 0);
  static constexpr int radix(int s) {
    return SF ? ((REM && s == 0) ? (1 << REM) : RX) : (s < NFULL ? RX : (1 << REM));
  }
  static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * radix(s - 1); }
};

// Twiddles exp(2*pi*i*t/2048), t < 2048, evaluated at compile time: the
// angle is reduced to an octant (|x| <= pi/4, from the exact 1/8-turn
// symmetries), then sine and cosine are double-precision Taylor sums
// (error below 1e-17) rounded once to float.  Every power-of-two N <= 2048
// reads entry t * 2048/N, the same angle exactly.  Filling the LDS table
// from this read-only array costs one load per entry where evaluating
// sincospif in every workgroup cost about 40 VALU instructions per entry.
struct TwEntry {
  float c, s;
};
struct TwTable {
  TwEntry e[2048];
};
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double tw_sin_series(double x) {  // |x| <= pi/4
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double tw_cos_series(double x) {  // |x| <= pi/4
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
constexpr TwTable make_tw_table() {
  TwTable t{};
  for (int i = 0; i < 2048; ++i) {
    // i/2048 turn = octant o (256 entries each) + remainder r, |r| <= 128
    const int o = (i + 128) / 256, r = i - 256 * o;
    const double x = kTwoPi * r / 2048.0;
    const double sr = tw_sin_series(x), cr = tw_cos_series(x);
    // rotate (cos r, sin r) by o eighths of a turn
    const double h = 0.70710678118654752440084436210485;
    double c = 0.0, s = 0.0;
    switch (o & 7) {
      case 0: c = cr; s = sr; break;
      case 1: c = h * (cr - sr); s = h * (cr + sr); break;
      case 2: c = -sr; s = cr; break;
      case 3: c = -h * (cr + sr); s = h * (cr - sr); break;
      case 4: c = -cr; s = -sr; break;
      case 5: c = -h * (cr - sr); s = -h * (cr + sr); break;
      case 6: c = sr; s = -cr; break;
      default: c = h * (cr + sr); s = -h * (cr - sr); break;
    }
    t.e[i].c = (float)c;
    t.e[i].s = (float)s;
  }
  return t;
}
__device__ constexpr TwTable kTw = make_tw_table();

// Twiddle table exp(S*2*pi*i*t/N), t < N, copied cooperatively into LDS.
template <int N, int S>
__device__ __forceinline__ void init_twiddles(float2* tw, int tid, int nthreads) {
  for (int t = tid; t < N; t += nthreads) {
    const TwEntry e = kTw.e[t * (2048 / N)];
    tw[t] = make_float2(e.c, S * e.s);
  }
}

// Element i of a sequence is stored at padded row i + i/8: the radix-8
// Stockham write pattern (index 8j + r across lanes) then spreads over all
// LDS banks instead of hitting one bank group 4-16 times (64 banks x 4 B).
// (An XOR swizzle, i ^ ((i/8) mod 32), and a padded twiddle table remove
// the remaining read conflicts of row transforms, 76 -> 22 extra cycles per
// wave, but cost more VALU address work than they save in these VALU-bound
// kernels: measured slower.)
template <int N>
struct Pad {
  static constexpr int ROWS = N + N / 8;
  __device__ static __forceinline__ int at(int i) { return i + (i >> 3); }
};

// Column tiles (COLS): a half-wave (32 lanes) covers 32/M rows of M*8 bytes;
// the radix-8 write pattern puts those rows 8 apart, i.e. on the same
// banks.  Rows are stored permuted inside each aligned group of eight,
// row ^ ((row / 8) mod (32/M)), which sends them to distinct bank slices
// while consecutive rows (the read pattern) stay distinct too.  No padding.
template <int N, int M>
struct Swz {
  static constexpr int MASK = M >= 32 ? 0 : 32 / M - 1;
  __device__ static __forceinline__ int row(int i) { return i ^ ((i >> 3) & MASK); }
};

template <int N, int T, int M, bool COLS>
struct Layout {
  static_assert(M * N == 8 * T, "eight values per thread");
  static constexpr int SIZE = COLS ? M * N : M * Pad<N>::ROWS;  // float2 elements of LDS
  __device__ static __forceinline__ int addr(int m, int i) {
    return COLS ? Swz<N, M>::row(i) * M + m : m * Pad<N>::ROWS + Pad<N>::at(i);
  }
};

// One Stockham stage ST.  For ST == 0 inputs come from `load(m, i)`; for the
// last stage outputs go to `store(m, i, v)`; otherwise LDS.
template <int N, int T, int M, bool COLS, int S, int ST, bool SF, int RX, class Load, class Store>
__device__ __forceinline__ void fft_stage(float2* lds, const float2* tw, int tid, Load& load, Store& store) {
  using P = Plan<N, SF, RX>;
  using Lo = Layout<N, T, M, COLS>;
  constexpr int R = P::radix(ST);
  constexpr int NS = P::ns(ST);
  constexpr int Q = 8 / R;
  constexpr int NR = N / R;
  constexpr bool LAST = (ST == P::NST - 1);
  float2 v[8];
  int mq[Q], jq[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    int b = q * T + tid;
    mq[q] = COLS ? (b % M) : (b / NR);
    jq[q] = COLS ? (b / M) : (b % NR);
  }
  if constexpr (ST == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int r = 0; r < R; ++r) v[q * R + r] = load(mq[q], jq[q] + r * NR);
  } else {
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int r = 0; r < R; ++r) v[q * R + r] = lds[Lo::addr(mq[q], jq[q] + r * NR)];
    __syncthreads();  // every read of this stage precedes any write
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if constexpr (NS > 1) {
      const int k = jq[q] % NS;
      constexpr int STEP = N / (NS * R);
#pragma unroll
      for (int r = 1; r < R; ++r) v[q * R + r] = cmul(v[q * R + r], tw[k * r * STEP]);
    }
    dftR<R, S>(&v[q * R]);
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int k = jq[q] % NS;
    const int d = (jq[q] / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if constexpr (LAST) {
        store(mq[q], d + r * NS, v[q * R + r]);
      } else {
        lds[Lo::addr(mq[q], d + r * NS)] = v[q * R + r];
      }
    }
  }
  if constexpr (!LAST) {
    __syncthreads();
    fft_stage<N, T, M, COLS, S, ST + 1, SF, RX>(lds, tw, tid, load, store);
  }
}

template <int N, int T, int M, bool COLS, int S, bool SF = false, int RX = 8, class Load, class Store>
__device__ __forceinline__ void wg_fft(float2* lds, const float2* tw, int tid, Load& load, Store& store) {
  fft_stage<N, T, M, COLS, S, 0, SF, RX>(lds, tw, tid, load, store);
}

// ---------------------------------------------------------------------------
// Shift-theorem phasors with exact argument reduction.  A displacement s
// [m] on an N-point grid of pitch dx gives, for signed frequency index k,
// the phase 2*pi*k*u with u = s/(N*dx) cycles.  The host-independent device
// code keeps u as a 32-bit fixed-point fraction of a cycle, so k*U wraps
// modulo 2^32 exactly and only the final conversion rounds.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t cycles_fixed(double u) {
  double f = u - floor(u);                   // [0, 1)
  double s = f * 4294967296.0;               // 2^32
  uint64_t q = (uint64_t)llrint(s);
  return (uint32_t)q;                        // 2^32 wraps to 0
}

__device__ __forceinline__ float2 phasor_fixed(uint32_t phase) {
  // phase * 2^-32 cycles -> signed half-cycles in [-1, 1)
  float h = (float)(int32_t)phase * 4.656612873077393e-10f;  // 2^-31
  float s, c;
  sincospif(h, &s, &c);
  return make_float2(c, s);
}

__device__ __forceinline__ int signed_freq(int k, int n) { return k < (n + 1) / 2 ? k : k - n; }

}  // namespace pt
