// Assemat-Wilson extrusion and the fused multi-layer readout, gfx950.
//
// Ring buffers: every layer owns a (cap, W) float32 slab of one
// (layers, cap, W) array.  Local row i of a layer (virtual row base + i in
// the reference's bookkeeping) lives at physical row (start + i) mod cap, so
// "compaction" is an index shift -- no row is ever copied.
//
//  * innovation terms for every new row of a step at once,
//    BB = beta * B_unit^T (K x W), B streamed once per step;
//  * one sequential round of the recurrence for up to 32 layers,
//    row = A * z + s_l * BB[k], A streamed once per round for all layers.
//    Both are k_mfma_gemv (a W x m*ld matrix against <= 32 vectors on the
//    FP32 matrix cores).
// Rows have a pitch `ld` (multiple of 4 floats, zero padding) so every
// matrix and ring row is 16-byte aligned.
//  * k_readout    : rotated, sub-pixel, per-layer-shifted interpolation of
//                   all layers for up to 8 directions (each with its own
//                   per-layer magnification, so NGS and LGS cones share one
//                   launch), summed and scaled into (dirs, n, n).
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "fft_core.hpp"

namespace pt {

// ---------------------------------------------------------------------------
// multi-vector GEMV: y_e[w] = sum_s sum_c M[w][s*ld + c] * v_{e,s}[c]
//                             (+ scale_e * add_e[w])
// Both the innovation terms (M = B_unit, one segment, v = beta rows) and a
// recurrence round (M = A, m segments, v = the m stencil rows of each layer)
// are this shape: a W x (m*ld) matrix streamed once against up to 32 short
// vectors.  float4 loads (ld % 4 == 0, 16-byte aligned rows, zero padding
// beyond the row width).
// ---------------------------------------------------------------------------
constexpr int VT = 16;         // vectors per MFMA column tile
constexpr int VMAX = 2 * VT;   // vectors per launch
constexpr int MMAX = 4;        // stencil rows (segments)
constexpr int FAR_RMAX = 4;    // far stencil rows
constexpr int FAR_COLS = 8;    // far points per far row

struct GemvArgs {
  const float* M;
  long long ldm;  // matrix row pitch (m * ld, + the far segment)
  int W, ld, m, nv;
  int zw;    // ZV4 == false: valid floats per vector (4-byte aligned, no padding)
  int nfar;  // far points: matrix columns [m*ld, m*ld + nfar) (zero padded to 16)
  const float* v[VMAX][MMAX];
  const float* far[VMAX][FAR_RMAX];  // far row r of vector e (ring rows)
  float* out[VMAX];
  const float* add[VMAX];
  float scale[VMAX];
};
static_assert(sizeof(GemvArgs) <= 4096, "kernel argument block");

// C (W x nv) = M (W x m*ld) * Z (m*ld x nv) on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32: A lane l = M[l%16][l/16], B lane l = Z[l/16][l%16],
// D lane l register r = C[4(l/16) + r][l%16]).  A workgroup owns 16 matrix
// rows and NT column tiles of 16 vectors (nv <= 16 NT); its MF_WAVES waves
// split the K = m*ld columns in chunks of 16.  In a chunk, lane l loads the
// float4 of row l%16 at k = 16c + 4(l/16) and the same float4 of vector
// l%16 of each tile; the four MFMAs per tile (one per float4 component) then
// cover the chunk with k permuted inside it, identically for both operands,
// so the products pair up exactly.  Each matrix value is read once per
// launch for up to 16 NT vectors, each vector value is reused for 16 matrix
// rows.  Partial tiles are summed across waves through LDS, then
// out[e][row] = C + scale[e] * add[e][row].
// ZV4: vectors are 16-byte aligned and zero-padded to ld (ring rows), else
// plain rows of zw floats read element by element (caller's innovations).
constexpr int MF_WAVES = 16;
typedef float mfma_v4f __attribute__((ext_vector_type(4)));

// s + a * b as one explicit fused multiply-add (whatever the contraction
// flags), so every kernel that finishes a row this way gives the same bits
__device__ __forceinline__ float add_product(float s, float a, float b) { return __builtin_fmaf(a, b, s); }

template <bool ZV4, int NT>
__global__ __launch_bounds__(64 * MF_WAVES) void k_mfma_gemv(GemvArgs a) {
  __shared__ float red[NT][MF_WAVES][16][17];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = blockIdx.x * 16;
  const int ri = lane & 15, kq = lane >> 4;
  const float* arow = a.M + (size_t)min(row0 + ri, a.W - 1) * a.ldm;  // tail rows recompute the last row
  const int cps = (a.ld + 15) >> 4;  // 16-wide chunks per segment (ld % 4 == 0; k >= ld reads as zero)
  mfma_v4f acc[NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) acc[j] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < MMAX; ++s) {
    if (s >= a.m) break;
    const float* zs[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      zs[j] = nullptr;
#pragma unroll
      for (int e = 0; e < VT; ++e)
        if (e == ri) zs[j] = a.v[j * VT + e][s];
    }
    const float* as = arow + (size_t)s * a.ld;
    for (int c = wave; c < cps; c += MF_WAVES) {
      const int k = 16 * c + 4 * kq;
      const bool in = k < a.ld;
      const float4 x = in ? *(const float4*)(as + k) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const bool live = j * VT + ri < a.nv;  // this lane's vector (B operand column)
        float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ZV4) {
          if (in && live) z = *(const float4*)(zs[j] + k);
        } else if (live) {
          if (k < a.zw) z.x = zs[j][k];
          if (k + 1 < a.zw) z.y = zs[j][k + 1];
          if (k + 2 < a.zw) z.z = zs[j][k + 2];
          if (k + 3 < a.zw) z.w = zs[j][k + 3];
        }
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, z.x, acc[j], 0, 0, 0);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, z.y, acc[j], 0, 0, 0);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.z, z.z, acc[j], 0, 0, 0);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w, z.w, acc[j], 0, 0, 0);
      }
    }
  }
  // Far points (ring rows only): point q of the segment is column
  // (2 (q % 8) + 1) W / 16 of far row q / 8.  In chunk c, lane quarter kq
  // covers q = 16 c + 4 kq + {0..3}: far row 2 c + kq / 2, columns
  // 4 (kq % 2) + {0..3}.
  if (ZV4 && a.nfar > 0) {
    const float* fz[NT][FAR_RMAX];
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int r = 0; r < FAR_RMAX; ++r) {
        fz[j][r] = nullptr;
#pragma unroll
        for (int e = 0; e < VT; ++e)
          if (e == ri) fz[j][r] = a.far[j * VT + e][r];
      }
    const float* as = arow + (size_t)a.m * a.ld;
    const int fch = (a.nfar + 15) >> 4;
    for (int c = wave; c < fch; c += MF_WAVES) {
      const int k = 16 * c + 4 * kq;
      const float4 x = *(const float4*)(as + k);
      const int r = 2 * c + (kq >> 1);
      const int i0 = 4 * (kq & 1);
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const bool live = j * VT + ri < a.nv && k < a.nfar;  // nfar % 8 == 0: whole quarters
        float z[4] = {0.f, 0.f, 0.f, 0.f};
        if (live) {
          const float* fr = r == 0 ? fz[j][0] : r == 1 ? fz[j][1] : r == 2 ? fz[j][2] : fz[j][3];
#pragma unroll
          for (int u = 0; u < 4; ++u) z[u] = fr[((2 * (i0 + u) + 1) * a.W) / (2 * FAR_COLS)];
        }
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, z[0], acc[j], 0, 0, 0);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, z[1], acc[j], 0, 0, 0);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.z, z[2], acc[j], 0, 0, 0);
        acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w, z[3], acc[j], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) red[j][wave][4 * kq + r][ri] = acc[j][r];
  __syncthreads();
  for (int q = threadIdx.x; q < NT * 256; q += 64 * MF_WAVES) {
    const int j = q >> 8, rr = (q >> 4) & 15, e = j * VT + (q & 15);
    if (e < a.nv && row0 + rr < a.W) {
      float sum = 0.f;
#pragma unroll
      for (int w = 0; w < MF_WAVES; ++w) sum += red[j][w][rr][q & 15];
      float* out = nullptr;
      const float* add = nullptr;
      float sc = 0.f;
#pragma unroll
      for (int k = 0; k < VMAX; ++k)
        if (e == k) {
          out = a.out[k];
          add = a.add[k];
          sc = a.scale[k];
        }
      out[row0 + rr] = add ? add_product(sum, sc, add[row0 + rr]) : sum;
    }
  }
}

static int launch_gemv(GemvArgs& a, bool zv4, hipStream_t st) {
  const int g = (a.W + 15) / 16, b = 64 * MF_WAVES;
  if (a.nv > VT) {
    if (zv4)
      k_mfma_gemv<true, 2><<<g, b, 0, st>>>(a);
    else
      k_mfma_gemv<false, 2><<<g, b, 0, st>>>(a);
  } else {
    if (zv4)
      k_mfma_gemv<true, 1><<<g, b, 0, st>>>(a);
    else
      k_mfma_gemv<false, 1><<<g, b, 0, st>>>(a);
  }
  PT_LAUNCH_CHECK();
  return PT_OK;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

// ---------------------------------------------------------------------------
// One extrusion step in one launch (k_extrude_step)
// ---------------------------------------------------------------------------
// The step's innovation terms and every recurrence round of k_mfma_gemv in a
// single launch: workgroup g owns matrix rows [16 g, 16 g + 16) of A and B
// (one workgroup per CU), keeps its slice of A in registers for the whole
// step (A is read from HBM once per step instead of once per round), keeps
// its share of the innovation terms B beta in LDS, and the rounds are
// separated by a grid barrier.  Within a workgroup the arithmetic is
// k_mfma_gemv's (same chunk order, same MFMA tiles, same cross-wave sum, the
// same fused multiply-add at the end), so the rows are bit-identical to
// pt_innovation_gemm + pt_extrude_rows_far.
//
// Grid barrier: every workgroup owns a 64-bit flag, the number of barriers
// it has passed since the sync words were zeroed; at a barrier it writes its
// flag (after its rows are written back) and one wave polls all flags of the
// grid until each has reached the barrier's number -- no read-modify-write
// atomics, so nothing serialises on one address.  sync[0] holds the barrier
// count at the start of the next call, sync[1] the abort word.  A wait is
// bounded by spin_limit polls: on expiry the workgroup sets the abort word
// and *status and leaves, and a workgroup that sees the abort word leaves
// too (the word is checked before the flags), so every workgroup ends.
constexpr int XS_LMAX = 16;              // layers per launch
constexpr int XS_VMAX = 256;             // innovation rows per step
constexpr int XS_RMAX = 32;              // rounds per step (rows of one layer)
constexpr int XS_CWMAX = 9;              // A chunks per segment and virtual wave in registers (row pitch <= 2304)
constexpr int XI_GRP = 3;                // innovation: chunks per prefetch group
constexpr int XI_NG = (XS_CWMAX + XI_GRP - 1) / XI_GRP;  // innovation: groups per virtual wave
constexpr int XS_GRP = 3;                // rounds: chunks per prefetch group
constexpr int XS_NG = (XS_CWMAX + XS_GRP - 1) / XS_GRP;  // rounds: groups per virtual wave and segment
constexpr int XS_NBUF = 4;               // rounds: prefetch ring (loads run XS_NBUF - 1 groups ahead)
constexpr int XS_NT = 2;                 // innovation vector tiles per pass (32 vectors)
constexpr int XS_GRIDMAX = 144;          // workgroups (row pitch <= 2304)
static_assert(XS_GRIDMAX <= 3 * 64, "a barrier poll reads three flags per lane");
constexpr int XS_FLAGS = 16;             // sync word of workgroup g's flag: XS_FLAGS + g
static_assert((XS_FLAGS + XS_GRIDMAX) * 8 <= PT_EXTRUDE_SYNC_BYTES, "sync words");

struct StepArgs {
  const float* A;  // (W, ldm): m stencil segments of ld, then the far segment
  const float* B;  // (W, ld)
  float* buf;      // (L, cap, ld) rings
  long long ldm, cap;
  int W, ld, m, nfar, L, rounds, nvec, cw;  // cw: A chunks per wave and segment
  int far_d[FAR_RMAX];                      // far rows behind the newest stencil row
  int t0[XS_LMAX], cnt[XS_LMAX], boff[XS_LMAX];
  long long bpitch[XS_LMAX];                // floats between a layer's innovation rows
  const float* beta[XS_LMAX];
  float scale[XS_LMAX];
  unsigned long long* sync;
  int* status;
  unsigned spin_limit;
};
static_assert(sizeof(StepArgs) <= 4096, "kernel argument block");

// physical ring row of virtual row r (rings smaller than 2^31 floats: 32 bits)
__device__ __forceinline__ int ring_row32(int r, int cap) {
  r %= cap;
  return r < 0 ? r + cap : r;
}

// Index of the (e+1)-th set bit of `live` (the layer in column e), or -1.
__device__ __forceinline__ int nth_layer(unsigned live, int e) {
  unsigned m = live;
#pragma unroll
  for (int j = 0; j < XS_LMAX - 1; ++j)
    if (j < e) m &= m - 1;
  return m ? __ffs(m) - 1 : -1;
}

// Grid barrier number `target`, in two halves.  step_arrive (wave 0 only,
// after it stored the workgroup's rows: only wave 0 stores rows, so its
// write-back covers them) publishes the workgroup's flag; step_wait (every
// thread; the last wave polls, so polling starts while wave 0 still writes)
// returns false when the barrier timed out here or anywhere in the grid.
__device__ __forceinline__ void step_arrive(const StepArgs& a, unsigned long long target) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this wave's row stores reach memory before the flag
  if (threadIdx.x == 0)
    __hip_atomic_store(a.sync + XS_FLAGS + blockIdx.x, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ bool step_wait(const StepArgs& a, unsigned long long target, int poll_wave) {
  __shared__ int ok;
  const int lane = threadIdx.x & 63;
  if ((int)(threadIdx.x >> 6) == poll_wave) {
    // lane q polls the flags of workgroups q, q + 64 and q + 128 (clamped to
    // the grid: a repeated flag changes nothing); the four loads of a poll
    // are independent, so a poll costs one memory round trip
    const unsigned long long* flags = a.sync + XS_FLAGS;
    const int g1 = (int)gridDim.x - 1;
    const unsigned long long *f0 = flags + min(lane, g1), *f1 = flags + min(lane + 64, g1),
                             *f2 = flags + min(lane + 128, g1);
    int good = 1;
    unsigned it = 0;
    for (;;) {
      const unsigned long long abort = __hip_atomic_load(a.sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long v0 = __hip_atomic_load(f0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long v1 = __hip_atomic_load(f1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned long long v2 = __hip_atomic_load(f2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const bool done = v0 >= target && v1 >= target && v2 >= target;
      if (__any(abort != 0) || ++it > a.spin_limit) {  // the abort word is checked first
        good = 0;
        break;
      }
      if (__all(done)) break;
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other workgroups' rows are read fresh after this
    if (lane == 0) {
      if (!good) {
        __hip_atomic_store(a.sync + 1, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(a.status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      ok = good;
    }
  }
  __syncthreads();
  return ok != 0;
}

// 512 threads: each wave plays two of k_mfma_gemv's 16 waves (virtual waves
// wave and wave + 8), with separate accumulators, so the A slices of both
// fit its registers (two waves per SIMD) and the sums keep k_mfma_gemv's
// order exactly.  Loads run ahead of the MFMAs: one group of XI_GRP chunks
// in the innovation phase, XS_NBUF - 1 groups of XS_GRP chunks in a round.
constexpr int XS_WAVES = MF_WAVES / 2;

// The innovation terms B beta of every new row of the step into bbl, with
// k_mfma_gemv<false, NT>'s arithmetic (same chunks per virtual wave, same
// MFMA tiles, same cross-wave sum).  Every load is unconditional, from an
// address clamped into its row, and out-of-range values are selected away
// afterwards: a load under a per-lane branch would make the compiler drain
// all loads in flight where the branches join.  Innovation rows are 16-byte
// aligned with a pitch of a multiple of 4 floats (float4 loads; a row is
// read up to its width rounded up to 4).

__device__ __forceinline__ void step_innovations(const StepArgs& a, float (&red)[XS_NT][MF_WAVES][16][17],
                                                 float (&bbl)[XS_VMAX][16], int wave, int ri, int kq, int row0,
                                                 int cps) {
  const float* brow = a.B + (size_t)min(row0 + ri, a.W - 1) * a.ld;
  for (int v0 = 0; v0 < a.nvec; v0 += XS_NT * VT) {
    const int nv = a.nvec - v0 < XS_NT * VT ? a.nvec - v0 : XS_NT * VT;
    const float* zs[XS_NT];
    bool zl[XS_NT];
#pragma unroll
    for (int j = 0; j < XS_NT; ++j) {
      const int v = v0 + j * VT + ri;
      zs[j] = brow;  // any readable row; its values are selected away
      zl[j] = false;
      if (j * VT + ri < nv) {
        // the layer of innovation row v (kernel arguments read with uniform indices only)
        for (int q = 0; q < a.L; ++q)
          if (a.cnt[q] > 0 && v >= a.boff[q] && v < a.boff[q] + a.cnt[q]) {
            zs[j] = a.beta[q] + (size_t)(v - a.boff[q]) * a.bpitch[q];
            zl[j] = true;
          }
      }
    }
    mfma_v4f acc[2][XS_NT];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int j = 0; j < XS_NT; ++j) acc[h][j] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
    float4 xb[2][XI_GRP], zb[2][XS_NT][XI_GRP];
    // group g: virtual wave wave + 8 (g / XI_NG), chunks XI_GRP (g % XI_NG) + [0, XI_GRP)
    auto load_group = [&](int g, float4 (&x)[XI_GRP], float4 (&z)[XS_NT][XI_GRP]) {
      const int vw = wave + XS_WAVES * (g / XI_NG);
#pragma unroll
      for (int u = 0; u < XI_GRP; ++u) {
        const int i = XI_GRP * (g % XI_NG) + u, c = vw + MF_WAVES * i;
        const bool ex = i < XS_CWMAX && i < a.cw && c < cps;  // wave-uniform; MFMAs skipped when false
        const int k = ex ? 16 * c + 4 * kq : 4 * kq;             // ld % 16 == 0: k + 3 < ld
        x[u] = *(const float4*)(brow + k);
#pragma unroll
        for (int j = 0; j < XS_NT; ++j) {
          // floats k..k+3 of the row: the 4-float window [kb, kb + 4) with
          // kb = min(k, W - 4) lies inside the row; the last chunk shifts it by
          // d = k - kb, and values past the row's end are +0
          // floats past the row's end (k + q >= W) are masked to +0 bitwise, so
          // whatever the row's padding holds never reaches the MFMA
          const float4 zz = *(const float4*)(zs[j] + (k < a.W ? k : 0));
          const int n = zl[j] ? a.W - k : 0;  // valid floats of this float4 (<= 0: none)
          z[j][u] = make_float4(__uint_as_float(__float_as_uint(zz.x) & (n > 0 ? ~0u : 0u)),
                                __uint_as_float(__float_as_uint(zz.y) & (n > 1 ? ~0u : 0u)),
                                __uint_as_float(__float_as_uint(zz.z) & (n > 2 ? ~0u : 0u)),
                                __uint_as_float(__float_as_uint(zz.w) & (n > 3 ? ~0u : 0u)));
        }
      }
    };
    auto mfma_group = [&](int g, const float4 (&x)[XI_GRP], const float4 (&z)[XS_NT][XI_GRP]) {
      const int h = g / XI_NG, vw = wave + XS_WAVES * h;
#pragma unroll
      for (int u = 0; u < XI_GRP; ++u) {
        const int i = XI_GRP * (g % XI_NG) + u, c = vw + MF_WAVES * i;
        if (i < XS_CWMAX && i < a.cw && c < cps) {  // wave-uniform
#pragma unroll
          for (int j = 0; j < XS_NT; ++j) {
            if (j * VT >= nv) break;  // tiles in use (as k_mfma_gemv<false, NT>)
            acc[h][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[u].x, z[j][u].x, acc[h][j], 0, 0, 0);
            acc[h][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[u].y, z[j][u].y, acc[h][j], 0, 0, 0);
            acc[h][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[u].z, z[j][u].z, acc[h][j], 0, 0, 0);
            acc[h][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[u].w, z[j][u].w, acc[h][j], 0, 0, 0);
          }
        }
      }
    };
    load_group(0, xb[0], zb[0]);
#pragma unroll
    for (int g = 0; g < 2 * XI_NG; ++g) {  // (virtual wave, group)
      if (g + 1 < 2 * XI_NG) load_group(g + 1, xb[(g + 1) & 1], zb[(g + 1) & 1]);
      mfma_group(g, xb[g & 1], zb[g & 1]);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int j = 0; j < XS_NT; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[j][wave + XS_WAVES * h][4 * kq + r][ri] = acc[h][j][r];
    __syncthreads();
    for (int q = threadIdx.x; q < XS_NT * 256; q += 64 * XS_WAVES) {
      const int j = q >> 8, rr = (q >> 4) & 15, e = j * VT + (q & 15);
      if (e < nv) {
        float sum = 0.f;
#pragma unroll
        for (int w = 0; w < MF_WAVES; ++w) sum += red[j][w][rr][q & 15];
        bbl[v0 + e][rr] = sum;
      }
    }
    __syncthreads();
  }
}

template <int M>
__global__ __launch_bounds__(64 * XS_WAVES) void k_extrude_step(StepArgs a) {
  __shared__ float red[XS_NT][MF_WAVES][16][17];
  __shared__ float bbl[XS_VMAX][16];
  __shared__ float rowv[16][17];
  __shared__ float4 afar[2][64];  // the far segment of A (virtual waves 0 and 1 at most)
  __shared__ float4 alast[2][M][XS_WAVES][64];  // chunk XS_CWMAX - 1 of every segment (registers are full)
  __shared__ unsigned long long base_s;
  __shared__ int lay_cnt[XS_LMAX], lay_t0[XS_LMAX], lay_boff[XS_LMAX];
  __shared__ float lay_sc[XS_LMAX];
  // per round r and column e (the e-th layer with cnt > r, or none): the
  // float offsets in buf of its stencil rows (RT_Z + s), far rows
  // (RT_F + j) and target row (RT_T), and the layer (RT_L, -1 for none)
  constexpr int RT_Z = 0, RT_F = 2, RT_T = RT_F + FAR_RMAX, RT_L = RT_T + 1;
  static_assert(M <= RT_F && RT_L < 8, "round table entry");
  __shared__ int rtab[XS_RMAX][16][8];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = blockIdx.x * 16;
  const int ri = lane & 15, kq = lane >> 4;
  if (threadIdx.x == 0) {
    base_s = __hip_atomic_load(a.sync, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the per-layer tables, for the rounds' per-lane lookups (kernel arguments
    // are read with uniform indices only)
#pragma unroll
    for (int q = 0; q < XS_LMAX; ++q) {
      const bool in = q < a.L;
      lay_cnt[q] = in ? a.cnt[q] : 0;
      lay_t0[q] = in ? a.t0[q] : 0;
      lay_boff[q] = in ? a.boff[q] : 0;
      lay_sc[q] = in ? a.scale[q] : 0.f;
    }
  }
  const int cps = (a.ld + 15) >> 4;
  // --- innovation terms B beta of every new row, into LDS
  step_innovations(a, red, bbl, wave, ri, kq, row0, cps);
  // --- this workgroup's rows of A, in registers for the whole step: for
  // virtual wave vw, chunk i of segment s is chunk vw + 16 i of the segment;
  // the far segment's chunk is vw (static register indices throughout)
  const float* arow = a.A + (size_t)min(row0 + ri, a.W - 1) * a.ldm;
  const int fch = (a.nfar + 15) >> 4;
  float4 areg[2][M][XS_CWMAX - 1];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int vw = wave + XS_WAVES * h;
#pragma unroll
    for (int s = 0; s < M; ++s)
#pragma unroll
      for (int i = 0; i < XS_CWMAX; ++i) {
        const int c = vw + MF_WAVES * i;
        const bool ex = i < a.cw && c < cps;  // chunks that do not exist are never multiplied
        const float4 v = *(const float4*)(arow + (size_t)s * a.ld + (ex ? 16 * c + 4 * kq : 4 * kq));
        if (i < XS_CWMAX - 1)
          areg[h][s][i] = v;
        else
          alast[h][s][wave][lane] = v;
      }
  }
  if (wave < fch) afar[wave][lane] = *(const float4*)(arow + (size_t)M * a.ld + 16 * wave + 4 * kq);
  // the round table (its integer work runs while the A loads are in flight;
  // ring rows as 32-bit float offsets: the rings hold fewer than 2^31 floats)
  for (int q = threadIdx.x; q < a.rounds * 16; q += 64 * XS_WAVES) {
    const int r = q >> 4, e = q & 15;
    unsigned live = 0;
#pragma unroll
    for (int l = 0; l < XS_LMAX; ++l) live |= (lay_cnt[l] > r ? 1u : 0u) << l;
    const int le = nth_layer(live, e);
    int* t = rtab[r][e];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = 0;
    t[RT_L] = le;
    if (le >= 0) {
      const int cap = (int)a.cap, slab = le * cap;
      const int target = ring_row32(lay_t0[le] + r, cap);
#pragma unroll
      for (int s = 0; s < M; ++s) t[RT_Z + s] = (slab + ring_row32(target - M + s, cap)) * a.ld;
#pragma unroll
      for (int j = 0; j < FAR_RMAX; ++j)
        if (8 * j < a.nfar) t[RT_F + j] = (slab + ring_row32(target - 1 - a.far_d[j], cap)) * a.ld;
      t[RT_T] = (slab + target) * a.ld;
    }
  }
  __syncthreads();
  const unsigned long long base = base_s;
  // --- the rounds: round r extrudes row r of every layer that has one.
  // Virtual wave vw sums its chunks of stencil segment 0, ..., M - 1 and
  // then its far points (k_mfma_gemv's order) in acc[h].  Segments
  // 0 .. M - 2 and the far rows lie two or more rows behind round r's
  // target row, so every workgroup wrote them before the barrier that ends
  // round r - 1: their loads and MFMAs run before that barrier is waited
  // for, and only segment M - 1 (the rows of round r - 1) follows it.
  int zo[M], fo[FAR_RMAX];
  float zf[2][4];
  // entries of round r (column ri of the round table): a column without a
  // layer reads row offset 0 and feeds only its own unused output column
  auto setup = [&](int r) {
    const int* t = rtab[r][ri];
#pragma unroll
    for (int s = 0; s < M; ++s) zo[s] = t[RT_Z + s];
#pragma unroll
    for (int j = 0; j < FAR_RMAX; ++j) fo[j] = t[RT_F + j];
    // far points of both virtual waves (k_mfma_gemv's far segment)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int vw = wave + XS_WAVES * h;
      const int fr = 2 * vw + (kq >> 1), i0 = 4 * (kq & 1);
#pragma unroll
      for (int u = 0; u < 4; ++u) zf[h][u] = 0.f;
      if (vw < fch) {  // wave-uniform
        const int fb = fr == 0 ? fo[0] : fr == 1 ? fo[1] : fr == 2 ? fo[2] : fo[3];
#pragma unroll
        for (int u = 0; u < 4; ++u) zf[h][u] = a.buf[(unsigned)(fb + ((2 * (i0 + u) + 1) * a.W) / (2 * FAR_COLS))];
      }
    }
  };
  // chunks in groups of XS_GRP, loaded XS_NBUF - 1 groups ahead of their
  // MFMAs: GO groups (h, s, gi) of the old segments s < M - 1, then GN
  // groups (h, gi) of segment M - 1
  constexpr int GO = 2 * (M - 1) * XS_NG, GN = 2 * XS_NG;
  float4 zb[XS_NBUF][XS_GRP];
  mfma_v4f acc[2];
  auto load_chunks = [&](int h, int s, int gi, float4 (&z)[XS_GRP]) {
    const int vw = wave + XS_WAVES * h;
#pragma unroll
    for (int u = 0; u < XS_GRP; ++u) {
      const int i = XS_GRP * gi + u, c = vw + MF_WAVES * i;
      // unconditional loads (the compiler drains loads in flight where
      // branches around loads join): a chunk that does not exist is never
      // multiplied, and a column without a layer (le < 0, zo = 0) only
      // feeds its own unused output column
      const int k = (i < XS_CWMAX && i < a.cw && c < cps) ? 16 * c + 4 * kq : 4 * kq;
      z[u] = *(const float4*)(a.buf + (unsigned)(zo[s] + k));
    }
  };
  auto mfma_chunks = [&](int h, int s, int gi, const float4 (&z)[XS_GRP]) {
    const int vw = wave + XS_WAVES * h;
#pragma unroll
    for (int u = 0; u < XS_GRP; ++u) {
      const int i = XS_GRP * gi + u, c = vw + MF_WAVES * i;
      if (i < XS_CWMAX && i < a.cw && c < cps) {  // wave-uniform
        const float4 x = i < XS_CWMAX - 1 ? areg[h][s][i] : alast[h][s][wave][lane];
        acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, z[u].x, acc[h], 0, 0, 0);
        acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, z[u].y, acc[h], 0, 0, 0);
        acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.z, z[u].z, acc[h], 0, 0, 0);
        acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w, z[u].w, acc[h], 0, 0, 0);
      }
    }
  };
  // round r's old segments (before the barrier that ends round r - 1)
  auto old_part = [&]() {
    acc[0] = acc[1] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
    if constexpr (GO > 0) {
      constexpr int PS = (M > 1 ? M - 1 : 1) * XS_NG;  // groups per virtual wave
#pragma unroll
      for (int g = 0; g < XS_NBUF - 1; ++g)
        if (g < GO) load_chunks(g / PS, (g / XS_NG) % (M - 1), g % XS_NG, zb[g]);
#pragma unroll
      for (int g = 0; g < GO; ++g) {
        const int gp = g + XS_NBUF - 1;
        if (gp < GO) load_chunks(gp / PS, (gp / XS_NG) % (M - 1), gp % XS_NG, zb[gp % XS_NBUF]);
        mfma_chunks(g / PS, (g / XS_NG) % (M - 1), g % XS_NG, zb[g % XS_NBUF]);
      }
    }
  };
  // segment M - 1 and the far points, then each virtual wave's partial tile into LDS
  auto new_part = [&]() {
#pragma unroll
    for (int g = 0; g < XS_NBUF - 1; ++g)
      if (g < GN) load_chunks(g / XS_NG, M - 1, g % XS_NG, zb[g]);
#pragma unroll
    for (int g = 0; g < GN; ++g) {
      const int gp = g + XS_NBUF - 1;
      if (gp < GN) load_chunks(gp / XS_NG, M - 1, gp % XS_NG, zb[gp % XS_NBUF]);
      const int h = g / XS_NG, vw = wave + XS_WAVES * h;
      mfma_chunks(h, M - 1, g % XS_NG, zb[g % XS_NBUF]);
      if (g % XS_NG == XS_NG - 1) {  // the last group of virtual wave vw: far points, then its partial tile
        if (vw < fch) {
          // points past nfar meet A's zero padding: zero them so they stay finite
          const bool fin = 16 * vw + 4 * kq < a.nfar;
          const float4 x = afar[vw][lane];
          acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, fin ? zf[h][0] : 0.f, acc[h], 0, 0, 0);
          acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, fin ? zf[h][1] : 0.f, acc[h], 0, 0, 0);
          acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.z, fin ? zf[h][2] : 0.f, acc[h], 0, 0, 0);
          acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w, fin ? zf[h][3] : 0.f, acc[h], 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) red[0][vw][4 * kq + q][ri] = acc[h][q];
      }
    }
  };
  setup(0);
  for (int r = 0; r < a.rounds; ++r) {
    old_part();
    if (r > 0 && !step_wait(a, base + r, XS_WAVES - 1)) return;
    new_part();
    __syncthreads();
    if (threadIdx.x < 256) {
      const int rr = threadIdx.x >> 4, e = threadIdx.x & 15;
      const int l = rtab[r][e][RT_L];
      if (l >= 0) {
        float sum = 0.f;
#pragma unroll
        for (int w = 0; w < MF_WAVES; ++w) sum += red[0][w][rr][e];
        rowv[rr][e] = add_product(sum, lay_sc[l], bbl[lay_boff[l] + r][rr]);
      }
    }
    __syncthreads();
    const bool more = r + 1 < a.rounds;
    if (wave == 0) {  // one wave stores the rows, so one write-back publishes them
      const int e = lane & 15;
      if (rtab[r][e][RT_L] >= 0) {
        float* dst = a.buf + (unsigned)(rtab[r][e][RT_T] + row0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int rr = (lane >> 4) + 4 * q;
          if (row0 + rr < a.W) dst[rr] = rowv[rr][e];
        }
      }
      if (more) step_arrive(a, base + r + 1);
    }
    if (more) setup(r + 1);  // round r + 1's rows, far-row loads issued before its barrier
  }
  // the barrier count for the next call (every workgroup read it before barrier 0)
  if (blockIdx.x == 0 && threadIdx.x == 0)
    __hip_atomic_store(a.sync, base + (a.rounds - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// fused readout
// ---------------------------------------------------------------------------
// Instruction-lean gather: the host splits each (direction, layer) shift
// into an integer part and a fraction (float64 -> int + float), so the
// per-pixel coordinate math is float32 on magnitudes <= ~n (abs. error
// ~1e-4 px) and all indexing is 32-bit.
constexpr int RL = 16, RD = 8;

struct ReadoutArgs {
  const float* buf;  // (L, cap, ld)
  float* out;        // (D, n, n)
  int cap, W, ld, n, L, D, accumulate;
  float scale;
  int start[RL];
  int fillm1[RL];
  float cosw[RL], sinw[RL];
  int r_int[RD][RL], c_int[RD][RL];
  float r_frac[RD][RL], c_frac[RD][RL], mag[RD][RL];
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

// Tap weights and their derivatives with respect to the fraction t.
template <int TAPS>
__device__ __forceinline__ void taps_dweights(float t, float* w, float* dw) {
  if constexpr (TAPS == 2) {
    w[0] = 1.f - t;
    w[1] = t;
    dw[0] = -1.f;
    dw[1] = 1.f;
  } else if constexpr (TAPS == 4) {
    const float t2 = t * t, t3 = t2 * t;
    w[0] = 0.5f * (-t + 2.f * t2 - t3);
    w[1] = 0.5f * (2.f - 5.f * t2 + 3.f * t3);
    w[2] = 0.5f * (t + 4.f * t2 - 3.f * t3);
    w[3] = 0.5f * (t3 - t2);
    dw[0] = 0.5f * (-1.f + 4.f * t - 3.f * t2);
    dw[1] = 0.5f * (-10.f * t + 9.f * t2);
    dw[2] = 0.5f * (1.f + 8.f * t - 9.f * t2);
    dw[3] = 0.5f * (3.f * t2 - 2.f * t);
  } else {  // normalised Lanczos-3: w_k = s_k / sum s, s_k = sinc(x) sinc(x/3), x = t - (k - 2)
    float sv[6], ds[6], tot = 0.f, dtot = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float x = t - (float)(k - 2);
      if (x == 0.f) {
        sv[k] = 1.f;
        ds[k] = 0.f;
      } else {
        const float x3 = x * (1.f / 3.f);
        float s1, c1, s3, c3;
        sincospif(x, &s1, &c1);
        sincospif(x3, &s3, &c3);
        const float pi = 3.14159265358979f;
        const float a = s1 / (pi * x), b = s3 / (pi * x3);
        const float da = (c1 - a) / x, db = (c3 - b) / x3;  // d sinc(y)/dy = (cos(pi y) - sinc(y)) / y
        sv[k] = a * b;
        ds[k] = da * b + a * db * (1.f / 3.f);
      }
      tot += sv[k];
      dtot += ds[k];
    }
    const float inv = 1.f / tot;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      w[k] = sv[k] * inv;
      dw[k] = (ds[k] - w[k] * dtot) * inv;
    }
  }
}

// Adjoint of the readout with respect to each (direction, layer)'s geometry
// (row shift sa, column shift sp, magnification mag) for a cotangent grad:
//   d/dsa  = scale * sum_px grad * d interp / d row,
//   d/dsp  = scale * sum_px grad * d interp / d col,
//   d/dmag = scale * sum_px grad * (d/drow (ci c + cj s) + d/dcol (-ci s + cj c)),
// ci = i - (n-1)/2, cj = j - (n-1)/2.  Taps are the forward's (same clamping,
// same FP32 coordinates); the per-block sums go to partial[((d * nblk + blk)
// * L + l) * 3 + q] and are added in FP64 by k_readout_adj_final.
template <int TAPS>
__global__ __launch_bounds__(256) void k_readout_adj(ReadoutArgs a, const float* __restrict__ grad,
                                                     float* __restrict__ partial) {
  constexpr int OFF = TAPS == 2 ? 0 : (TAPS == 4 ? -1 : -2);
  __shared__ float red[4 * 3];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = blockIdx.x * 32 + (threadIdx.x & 31);
  const int i = blockIdx.y * 8 + (threadIdx.x >> 5);
  const bool in = i < a.n && j < a.n;
  const int d = blockIdx.z;
  const int nblk = gridDim.x * gridDim.y, blk = blockIdx.y * gridDim.x + blockIdx.x;
  const float h = 0.5f * (float)(a.n - 1);
  const float ci = in ? (float)i - h : 0.f, cj = in ? (float)j - h : 0.f;
  const float G = in ? grad[((size_t)d * a.n + i) * a.n + j] * a.scale : 0.f;
  for (int l = 0; l < a.L; ++l) {
    float q3[3] = {0.f, 0.f, 0.f};
    if (in) {
      const float mg = a.mag[d][l];
      const float gi = ci * mg, gj = cj * mg;
      const float c = a.cosw[l], s = a.sinw[l];
      const float row = fmaf(gi, c, fmaf(gj, s, a.r_frac[d][l]));
      const float col = fmaf(-gi, s, fmaf(gj, c, a.c_frac[d][l]));
      const float rf = floorf(row), cf = floorf(col);
      const int r0 = (int)rf + a.r_int[d][l], c0 = (int)cf + a.c_int[d][l];
      float wr[TAPS], dwr[TAPS], wc[TAPS], dwc[TAPS];
      taps_dweights<TAPS>(row - rf, wr, dwr);
      taps_dweights<TAPS>(col - cf, wc, dwc);
      const float* bl = a.buf + (size_t)l * a.cap * a.ld;
      const int fm = a.fillm1[l], st = a.start[l];
      float vr = 0.f, vc = 0.f;
#pragma unroll
      for (int u = 0; u < TAPS; ++u) {
        int ph = st + min(max(r0 + OFF + u, 0), fm);
        ph = ph >= a.cap ? ph - a.cap : ph;
        const float* br = bl + (size_t)ph * a.ld;
        float rw = 0.f, rdw = 0.f;
#pragma unroll
        for (int v = 0; v < TAPS; ++v) {
          const float b = br[min(max(c0 + OFF + v, 0), a.W - 1)];
          rw = fmaf(wc[v], b, rw);
          rdw = fmaf(dwc[v], b, rdw);
        }
        vr = fmaf(dwr[u], rw, vr);
        vc = fmaf(wr[u], rdw, vc);
      }
      const float gr = G * vr, gc = G * vc;
      q3[0] = gr;
      q3[1] = gc;
      q3[2] = gr * (ci * c + cj * s) + gc * (cj * c - ci * s);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      float v = q3[q];
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
      if (lane == 0) red[wave * 3 + q] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      const int q = threadIdx.x;
      partial[(((size_t)d * nblk + blk) * a.L + l) * 3 + q] = red[q] + red[3 + q] + red[6 + q] + red[9 + q];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(64) void k_readout_adj_final(const float* __restrict__ partial, int D, int L, int nblk,
                                                          int ldo, double* __restrict__ out) {
  const int idx = blockIdx.x * 64 + threadIdx.x;
  if (idx >= D * L * 3) return;
  const int q = idx % 3, l = (idx / 3) % L, d = idx / (3 * L);
  double acc = 0.0;
  for (int b = 0; b < nblk; ++b) acc += (double)partial[(((size_t)d * nblk + b) * L + l) * 3 + q];
  out[((size_t)d * ldo + l) * 3 + q] = acc;
}

template <int TAPS>
__global__ __launch_bounds__(256) void k_readout(ReadoutArgs a) {
  constexpr int OFF = TAPS == 2 ? 0 : (TAPS == 4 ? -1 : -2);
  const int j = blockIdx.x * 32 + (threadIdx.x & 31);
  const int i = blockIdx.y * 8 + (threadIdx.x >> 5);
  if (i >= a.n || j >= a.n) return;
  const float h = 0.5f * (float)(a.n - 1);
  const float ci = (float)i - h, cj = (float)j - h;  // exact half-integers
  const int d = blockIdx.z;  // one direction per block: short dependency chains
  {
    float acc = 0.f;
#pragma unroll 3
    for (int l = 0; l < a.L; ++l) {
      const float mg = a.mag[d][l];
      const float gi = ci * mg, gj = cj * mg;
      const float c = a.cosw[l], s = a.sinw[l];
      const float row = fmaf(gi, c, fmaf(gj, s, a.r_frac[d][l]));
      const float col = fmaf(-gi, s, fmaf(gj, c, a.c_frac[d][l]));
      const float rf = floorf(row), cf = floorf(col);
      const int r0 = (int)rf + a.r_int[d][l], c0 = (int)cf + a.c_int[d][l];
      float wr[TAPS], wc[TAPS];
      taps_weights<TAPS>(row - rf, wr);
      taps_weights<TAPS>(col - cf, wc);
      const float* bl = a.buf + (size_t)l * a.cap * a.ld;
      const int fm = a.fillm1[l], st = a.start[l];
      int cc[TAPS];
#pragma unroll
      for (int v = 0; v < TAPS; ++v) cc[v] = min(max(c0 + OFF + v, 0), a.W - 1);
      float val = 0.f;
#pragma unroll
      for (int u = 0; u < TAPS; ++u) {
        int ph = st + min(max(r0 + OFF + u, 0), fm);
        ph = ph >= a.cap ? ph - a.cap : ph;
        const float* br = bl + ph * a.ld;
        float rs = 0.f;
#pragma unroll
        for (int v = 0; v < TAPS; ++v) rs = fmaf(wc[v], br[cc[v]], rs);
        val = fmaf(wr[u], rs, val);
      }
      acc += val;
    }
    float* o = a.out + ((size_t)d * a.n + i) * a.n + j;
    const float v = acc * a.scale;
    *o = a.accumulate ? *o + v : v;
  }
}

constexpr int TB = 32;  // tile edge

// Geometry of the tiled readout, per (direction, layer), in FP64:
//   row(i, j) = R0 + i P + j Q,   col(i, j) = C0 - i Q + j P,
// P = mag cos, Q = mag sin, R0 = sa - h (P + Q), C0 = W/2 + sp + h (Q - P),
// h = (n - 1) / 2 -- the reference's along/perp grids plus the shifts
// (extrude.py:376-383, 807-822).  Each workgroup evaluates its tile origin in
// FP64 and the pixels relative to it in FP32 (offsets < 32 px), so a
// coordinate is exact to ~1e-6 px at any grid size.
constexpr int TRL = 12, TRD = 8;  // layers, directions per launch
struct TileArgs {
  const float* buf;  // (L, cap, ld)
  float* out;        // (D, n, n)
  int cap, W, ld, n, L, D, accumulate;
  float scale;
  int start[TRL];
  int fillm1[TRL];
  double R0[TRD][TRL], C0[TRD][TRL], P[TRD][TRL], Q[TRD][TRL];
};
static_assert(sizeof(TileArgs) <= 4096, "kernel argument block");

// Box of a tile's footprint in one layer: local ring rows r0.. and buffer
// columns c0.. (before clamping), extents nr x nc; (fr, fc) is the tile
// origin pixel relative to the box origin.
struct TileGeo {
  int r0, c0, nr, nc;
  float fr, fc, p, q;
  bool fits;
};

// Box edge per tap count: for magnification <= 1 (every NGS direction and
// LGS cone) a tile's footprint spans at most floor(31 sqrt 2) = 43 rows and
// columns between its extreme pixels, plus the taps and a spare on each
// side: 48 / 50 / 52 rows, one lane per column.  A larger footprint takes
// the direct gather.
template <int TAPS>
struct TileCfg {
  static constexpr int OFF = TAPS == 2 ? 0 : (TAPS == 4 ? -1 : -2);
  static constexpr int BH = 46 + TAPS;       // box rows
  static constexpr int BW = 64;              // box columns (lanes)
  static constexpr int BP = BW + 1;          // LDS pitch
  static constexpr int KR = (BH + 3) / 4;    // box rows per wave
};

template <int TAPS>
__device__ __forceinline__ TileGeo tile_geo(const TileArgs& a, int d, int l, int i0, int j0) {
  constexpr int OFF = TileCfg<TAPS>::OFF;
  const double P = a.P[d][l], Q = a.Q[d][l];
  const double rb = fma((double)i0, P, fma((double)j0, Q, a.R0[d][l]));
  const double cb = fma((double)i0, -Q, fma((double)j0, P, a.C0[d][l]));
  TileGeo g;
  g.p = (float)P;
  g.q = (float)Q;
  // the map is affine: a full tile's extremes are at its corners, offsets
  // 31 (min / max of 0 and 31 p, 31 q); an edge tile uses the same box
  constexpr float E = (float)(TB - 1);
  const float erl = E * (fminf(g.p, 0.f) + fminf(g.q, 0.f)), erh = E * (fmaxf(g.p, 0.f) + fmaxf(g.q, 0.f));
  const float ecl = E * (fminf(-g.q, 0.f) + fminf(g.p, 0.f)), ech = E * (fmaxf(-g.q, 0.f) + fmaxf(g.p, 0.f));
  // one spare row/column each side absorbs the FP32 rounding of interior pixels
  g.r0 = (int)floor(rb + (double)erl) + OFF - 1;
  g.c0 = (int)floor(cb + (double)ecl) + OFF - 1;
  g.nr = (int)floor(rb + (double)erh) + OFF + TAPS + 1 - g.r0;
  g.nc = (int)floor(cb + (double)ech) + OFF + TAPS + 1 - g.c0;
  g.fits = g.nr <= TileCfg<TAPS>::BH && g.nc <= TileCfg<TAPS>::BW;
  g.fr = (float)(rb - (double)g.r0);
  g.fc = (float)(cb - (double)g.c0);
  return g;
}

// Tiled readout for n >= 32: a workgroup owns a 32 x 32 pixel tile of one
// direction.  For each layer the tile's footprint in the ring buffer plus
// the tap margins is staged in LDS (one lane per column, coalesced rows;
// rows clamped to the valid ring rows and columns to [0, W - 1] exactly as
// the per-tap clamp does), and the taps are read from LDS.  The box of layer
// l + 1 is loaded into registers while layer l is interpolated (two LDS
// buffers, one barrier per layer).  A footprint larger than the box
// (magnification > 1) takes the direct gather for that layer.
//  * Workgroups run in XCD-contiguous order over (tile, direction),
//    direction fastest: one XCD takes a band of neighbouring tiles with all
//    their directions, whose footprints (overlapping box margins, the low
//    layers' near-identical boxes across directions) meet in its L2.
//  * Staging: lane k of each wave works out the ring offset of the wave's
//    box row k once per layer; the loads take it with v_readlane, so a box
//    row costs one address add.  Every lane loads (columns past the box are
//    clamped reads of the row), no lane masking.
//  * The per-layer box geometry (FP64 tile origin, box extents) is worked
//    out once per wave, lane l for layer l, and broadcast to scalar
//    registers with v_readlane when the layer comes up: ten instructions per
//    layer instead of the FP64 arithmetic in every lane.
template <int TAPS>
__global__ __launch_bounds__(256) void k_readout_tile(TileArgs a, int tiles_x, int ntiles) {
  using C = TileCfg<TAPS>;
  constexpr int OFF = C::OFF, BP = C::BP, KR = C::KR;
  static_assert(KR <= 64, "one lane per box row of a wave");
  __shared__ float box[2][C::BH * BP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int t = xcd_tile((int)blockIdx.x, ntiles);
  const int d = t % a.D, tile = t / a.D;
  const int i0 = (tile / tiles_x) * TB, j0 = (tile % tiles_x) * TB;
  const int jl = tid & 31, il = tid >> 5;  // pixels (i0 + il + 8k, j0 + jl), k < 4
  const float fj = (float)jl;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  float pre[KR];
  const int wv = __builtin_amdgcn_readfirstlane(wave);  // wave-uniform box-row tests
  auto fetch = [&](int l, const TileGeo& g) {
    if (!g.fits) return;
    // the slab is < 2^30 floats: 32-bit byte offsets from its base
    const char* bl = (const char*)(a.buf + (size_t)l * a.cap * a.ld);
    const int fm = a.fillm1[l];
    const unsigned ph = (unsigned)(a.start[l] + min(max(g.r0 + wave + 4 * min(lane, KR - 1), 0), fm));
    const unsigned rowoff = __umul24(min(ph, ph - (unsigned)a.cap), (unsigned)a.ld) << 2;  // ph mod cap
    const unsigned coloff = (unsigned)min(max(g.c0 + lane, 0), a.W - 1) << 2;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      if (wv + 4 * k < g.nr) {
        const unsigned ro = (unsigned)__builtin_amdgcn_readlane((int)rowoff, k);
        pre[k] = *(const float*)(bl + (ro + coloff));
      }
    }
  };
  auto commit = [&](int sel, const TileGeo& g) {
    if (!g.fits) return;
#pragma unroll
    for (int k = 0; k < KR; ++k) {
      if (wv + 4 * k < g.nr) box[sel][(wave + 4 * k) * BP + lane] = pre[k];
    }
  };
  // lane l < L: layer l's box; geo(l) reads it back into scalar registers
  const TileGeo mine = tile_geo<TAPS>(a, d, min(lane, a.L - 1), i0, j0);
  auto geo = [&](int l) {
    TileGeo g;
    g.r0 = __builtin_amdgcn_readlane(mine.r0, l);
    g.c0 = __builtin_amdgcn_readlane(mine.c0, l);
    g.nr = __builtin_amdgcn_readlane(mine.nr, l);
    g.nc = __builtin_amdgcn_readlane(mine.nc, l);
    g.fr = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.fr), l));
    g.fc = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.fc), l));
    g.p = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.p), l));
    g.q = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.q), l));
    g.fits = __builtin_amdgcn_readlane((int)mine.fits, l) != 0;
    return g;
  };
  TileGeo cur = geo(0);
  fetch(0, cur);
  commit(0, cur);
  __syncthreads();
  for (int l = 0; l < a.L; ++l) {
    const int sel = l & 1;
    TileGeo nxt;
    nxt.fits = false;
    if (l + 1 < a.L) {
      nxt = geo(l + 1);
      fetch(l + 1, nxt);  // in flight while layer l is interpolated
    }
    // per-thread row terms: row = rj + di p, col = cj - di q
    const float rj = fmaf(fj, cur.q, cur.fr), cj = fmaf(fj, cur.p, cur.fc);
    if (cur.fits) {
      const float* bx = box[sel];
      // coordinates kept inside the box (a guard: interior pixels are at
      // least one spare row/column from its edges)
      const float rhi = (float)(cur.nr - TAPS - OFF), chi = (float)(cur.nc - TAPS - OFF);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float di = (float)(il + 8 * k);
        const float row = __builtin_amdgcn_fmed3f(fmaf(di, cur.p, rj), (float)-OFF, rhi);
        const float col = __builtin_amdgcn_fmed3f(fmaf(-di, cur.q, cj), (float)-OFF, chi);
        const float rf = floorf(row), cf = floorf(col);
        const float* bp = bx + ((int)fmaf(rf, (float)BP, cf) + OFF * (BP + 1));  // tap (0, 0)
        float val;
        if constexpr (TAPS == 2) {
          const float tr = row - rf, tc = col - cf;
          const float top = fmaf(tc, bp[1] - bp[0], bp[0]);
          const float bot = fmaf(tc, bp[BP + 1] - bp[BP], bp[BP]);
          val = fmaf(tr, bot - top, top);
        } else {
          float wr[TAPS], wc[TAPS];
          taps_weights<TAPS>(row - rf, wr);
          taps_weights<TAPS>(col - cf, wc);
          val = 0.f;
#pragma unroll
          for (int u = 0; u < TAPS; ++u) {
            float rs = 0.f;
#pragma unroll
            for (int v = 0; v < TAPS; ++v) rs = fmaf(wc[v], bp[u * BP + v], rs);
            val = fmaf(wr[u], rs, val);
          }
        }
        acc[k] += val;
      }
    } else {  // direct gather with the per-tap clamps
      const float* bl = a.buf + (size_t)l * a.cap * a.ld;
      const int fm = a.fillm1[l], st = a.start[l];
#pragma unroll 1
      for (int k = 0; k < 4; ++k) {
        const float di = (float)(il + 8 * k);
        const float row = fmaf(di, cur.p, rj), col = fmaf(-di, cur.q, cj);
        const float rf = floorf(row), cf = floorf(col);
        const int r0 = (int)rf + cur.r0, c0 = (int)cf + cur.c0;
        float wr[TAPS], wc[TAPS];
        taps_weights<TAPS>(row - rf, wr);
        taps_weights<TAPS>(col - cf, wc);
        float val = 0.f;
#pragma unroll
        for (int u = 0; u < TAPS; ++u) {
          int ph = st + min(max(r0 + OFF + u, 0), fm);
          ph = ph >= a.cap ? ph - a.cap : ph;
          const float* brow = bl + (unsigned)ph * (unsigned)a.ld;
          float rs = 0.f;
#pragma unroll
          for (int v = 0; v < TAPS; ++v) rs = fmaf(wc[v], brow[min(max(c0 + OFF + v, 0), a.W - 1)], rs);
          val = fmaf(wr[u], rs, val);
        }
        acc[k] += val;
      }
    }
    if (l + 1 < a.L) {
      // buffer sel ^ 1 was last read in iteration l - 1, before the previous barrier
      commit(sel ^ 1, nxt);
      __syncthreads();
      cur = nxt;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = i0 + il + 8 * k, j = j0 + jl;
    if (i >= a.n || j >= a.n) continue;
    float* o = a.out + ((size_t)d * a.n + i) * a.n + j;
    const float v = acc[k] * a.scale;
    *o = a.accumulate ? *o + v : v;
  }
}

}  // namespace pt

using namespace pt;

extern "C" int pt_innovation_gemm(const float* const* beta_rows, const float* B, int rows, int width, int ld,
                                  float* out, void* stream) {
  if (!beta_rows || !B || !out) return fail(PT_ERR_ARG, "pt_innovation_gemm: NULL array");
  if (rows < 0 || width < 1 || ld < width || (ld & 3)) return fail(PT_ERR_ARG, "pt_innovation_gemm: bad sizes (ld %% 4)");
  if (!aligned16(B) || !aligned16(out)) return fail(PT_ERR_ARG, "pt_innovation_gemm: B and out must be 16-byte aligned");
  for (int k = 0; k < rows; ++k)
    if (!beta_rows[k] || ((uintptr_t)beta_rows[k] & 3))
      return fail(PT_ERR_ARG, "pt_innovation_gemm: innovation row %d NULL or not 4-byte aligned", k);
  hipStream_t st = (hipStream_t)stream;
  for (int k0 = 0; k0 < rows; k0 += VMAX) {
    GemvArgs a = {};
    a.M = B;
    a.ldm = ld;
    a.W = width;
    a.ld = ld;
    a.m = 1;
    a.zw = width;
    a.nv = rows - k0 < VMAX ? rows - k0 : VMAX;
    for (int e = 0; e < VMAX; ++e) {
      const int k = k0 + (e < a.nv ? e : 0);
      a.v[e][0] = beta_rows[k];
      a.out[e] = out + (size_t)k * ld;
      a.add[e] = nullptr;
      a.scale[e] = 0.f;
    }
    int rc = launch_gemv(a, false, st);
    if (rc) return rc;
  }
  return PT_OK;
}

static int extrude_rows_impl(const char* who, const float* A, int width, int ld, int stencil_rows, int far_points,
                             const float* bb, float* buf, long long cap, int nact, const int* layer,
                             const int* zrows, const int* far_rows, const int* trow, const int* bbrow,
                             const float* bscale, void* stream) {
  if (!A || !buf || !layer || !zrows || !trow || !bbrow || !bscale || !bb || (far_points && !far_rows))
    return fail(PT_ERR_ARG, "%s: NULL array", who);
  if (width < 1 || ld < width || (ld & 3) || stencil_rows < 1 || cap < stencil_rows + 1 || nact < 0)
    return fail(PT_ERR_ARG, "%s: bad sizes (ld %% 4 == 0, ld >= width)", who);
  if (stencil_rows > MMAX) return fail(PT_ERR_UNSUPPORTED, "%s: stencil_rows > %d", who, MMAX);
  if (far_points < 0 || far_points % FAR_COLS || far_points > FAR_COLS * FAR_RMAX)
    return fail(PT_ERR_ARG, "%s: far_points must be a multiple of %d in [0, %d]", who, FAR_COLS,
                FAR_COLS * FAR_RMAX);
  if (!aligned16(A) || !aligned16(buf) || !aligned16(bb)) return fail(PT_ERR_ARG, "%s: arrays must be 16-byte aligned", who);
  const int R = far_points / FAR_COLS;
  hipStream_t st = (hipStream_t)stream;
  for (int e0 = 0; e0 < nact; e0 += VMAX) {
    GemvArgs a = {};
    a.M = A;
    a.ldm = (long long)stencil_rows * ld + ((far_points + 15) / 16) * 16;
    a.W = width;
    a.ld = ld;
    a.m = stencil_rows;
    a.nfar = far_points;
    a.nv = nact - e0 < VMAX ? nact - e0 : VMAX;
    for (int e = 0; e < VMAX; ++e) {
      const int src = e0 + (e < a.nv ? e : 0);
      const size_t slab = (size_t)layer[src] * cap;
      for (int s = 0; s < MMAX; ++s) {
        const int r = zrows[src * stencil_rows + (s < stencil_rows ? s : 0)];
        if (r < 0 || r >= cap) return fail(PT_ERR_ARG, "%s: stencil row %d outside ring", who, r);
        a.v[e][s] = buf + (slab + r) * ld;
      }
      for (int j = 0; j < FAR_RMAX; ++j) {
        const int r = R ? far_rows[src * R + (j < R ? j : 0)] : 0;
        if (r < 0 || r >= cap) return fail(PT_ERR_ARG, "%s: far row %d outside ring", who, r);
        a.far[e][j] = buf + (slab + r) * ld;
      }
      if (trow[src] < 0 || trow[src] >= cap) return fail(PT_ERR_ARG, "%s: target row outside ring", who);
      a.out[e] = buf + (slab + trow[src]) * ld;
      a.add[e] = bb + (size_t)bbrow[src] * ld;
      a.scale[e] = bscale[src];
    }
    int rc = launch_gemv(a, true, st);
    if (rc) return rc;
  }
  return PT_OK;
}

extern "C" int pt_extrude_rows(const float* A, int width, int ld, int stencil_rows, const float* bb, float* buf,
                               long long cap, int nact, const int* layer, const int* zrows, const int* trow,
                               const int* bbrow, const float* bscale, void* stream) {
  return extrude_rows_impl("pt_extrude_rows", A, width, ld, stencil_rows, 0, bb, buf, cap, nact, layer, zrows,
                           nullptr, trow, bbrow, bscale, stream);
}

extern "C" int pt_extrude_rows_far(const float* A, int width, int ld, int stencil_rows, int far_points,
                                   const float* bb, float* buf, long long cap, int nact, const int* layer,
                                   const int* zrows, const int* far_rows, const int* trow, const int* bbrow,
                                   const float* bscale, void* stream) {
  return extrude_rows_impl("pt_extrude_rows_far", A, width, ld, stencil_rows, far_points, bb, buf, cap, nact,
                           layer, zrows, far_rows, trow, bbrow, bscale, stream);
}

// Geometry of one chunk of up to RD directions x RL layers: host-split
// integer and fractional shifts, FP32 rotation and magnification.
static int pack_readout(ReadoutArgs& a, const float* buf, int layers, long long cap, int width, int ld,
                        const int* start, const long long* fill, const double* wind_cs, const double* dir_geom,
                        int dirs, int n, int d0, int l0) {
  a.buf = buf + (size_t)l0 * cap * ld;
  a.out = nullptr;
  a.cap = (int)cap;
  a.W = width;
  a.ld = ld;
  a.n = n;
  a.L = layers - l0 < RL ? layers - l0 : RL;
  a.D = dirs - d0 < RD ? dirs - d0 : RD;
  a.accumulate = 0;
  a.scale = 1.f;
  for (int l = 0; l < RL; ++l) {
    const int src = l0 + (l < a.L ? l : 0);
    a.start[l] = start[src];
    a.fillm1[l] = (int)(fill[src] - 1);
    a.cosw[l] = (float)wind_cs[2 * src];
    a.sinw[l] = (float)wind_cs[2 * src + 1];
  }
  for (int d = 0; d < RD; ++d)
    for (int l = 0; l < RL; ++l) {
      const int sd = d0 + (d < a.D ? d : 0), sl = l0 + (l < a.L ? l : 0);
      const double* gp = dir_geom + ((size_t)sd * layers + sl) * 3;
      const double rs = gp[0], cs = 0.5 * width + gp[1];  // row / column shifts
      const double rfl = floor(rs), cfl = floor(cs);
      if (!(fabs(rfl) <= 1e9) || !(fabs(cfl) <= 1e9)) return fail(PT_ERR_ARG, "readout: shift out of range");
      a.r_int[d][l] = (int)rfl;
      a.c_int[d][l] = (int)cfl;
      a.r_frac[d][l] = (float)(rs - rfl);
      a.c_frac[d][l] = (float)(cs - cfl);
      a.mag[d][l] = (float)gp[2];
    }
  return PT_OK;
}

static int readout_check(const char* who, const float* buf, int layers, long long cap, int width, int ld,
                         const int* start, const long long* fill, const double* wind_cs, const double* dir_geom,
                         int dirs, int n, int interp) {
  if (!buf || !start || !fill || !wind_cs || !dir_geom) return fail(PT_ERR_ARG, "%s: NULL array", who);
  if (layers < 1 || dirs < 1 || n < 1 || width < 1 || cap < 1 || ld < width) return fail(PT_ERR_ARG, "%s: bad sizes", who);
  if (cap * (long long)ld > (1LL << 30)) return fail(PT_ERR_UNSUPPORTED, "%s: ring slab over 2^30 floats", who);
  if (interp < 0 || interp > 2) return fail(PT_ERR_ARG, "%s: interp must be 0 linear, 1 cubic, 2 lanczos", who);
  for (int l = 0; l < layers; ++l) {
    if (fill[l] < 1 || fill[l] > cap || start[l] < 0 || start[l] >= cap)
      return fail(PT_ERR_ARG, "%s: layer %d ring state out of range", who, l);
  }
  return PT_OK;
}

// Geometry of one chunk of up to TRD directions x TRL layers for the tiled
// readout (FP64 affine coefficients, see TileArgs).
static int pack_tile(TileArgs& a, const float* buf, int layers, long long cap, int width, int ld, const int* start,
                     const long long* fill, const double* wind_cs, const double* dir_geom, int dirs, int n, int d0,
                     int l0) {
  a.buf = buf + (size_t)l0 * cap * ld;
  a.out = nullptr;
  a.cap = (int)cap;
  a.W = width;
  a.ld = ld;
  a.n = n;
  a.L = layers - l0 < TRL ? layers - l0 : TRL;
  a.D = dirs - d0 < TRD ? dirs - d0 : TRD;
  a.accumulate = 0;
  a.scale = 1.f;
  const double h = 0.5 * (double)(n - 1);
  for (int l = 0; l < TRL; ++l) {
    const int src = l0 + (l < a.L ? l : 0);
    a.start[l] = start[src];
    a.fillm1[l] = (int)(fill[src] - 1);
  }
  for (int d = 0; d < TRD; ++d)
    for (int l = 0; l < TRL; ++l) {
      const int sd = d0 + (d < a.D ? d : 0), sl = l0 + (l < a.L ? l : 0);
      const double* gp = dir_geom + ((size_t)sd * layers + sl) * 3;
      const double c = wind_cs[2 * sl], s = wind_cs[2 * sl + 1], mg = gp[2];
      const double P = mg * c, Q = mg * s;
      const double R0 = gp[0] - h * (P + Q), C0 = 0.5 * width + gp[1] + h * (Q - P);
      if (!(fabs(R0) <= 1e9) || !(fabs(C0) <= 1e9) || !(fabs(P) <= 1e3) || !(fabs(Q) <= 1e3))
        return fail(PT_ERR_ARG, "readout: geometry out of range");
      a.R0[d][l] = R0;
      a.C0[d][l] = C0;
      a.P[d][l] = P;
      a.Q[d][l] = Q;
    }
  return PT_OK;
}

extern "C" int pt_readout(const float* buf, int layers, long long cap, int width, int ld, const int* start,
                          const long long* fill, const double* wind_cs, const double* dir_geom, int dirs, int n,
                          int interp, float scale, int accumulate, float* out, void* stream) {
  if (!out) return fail(PT_ERR_ARG, "pt_readout: NULL array");
  int rc = readout_check("pt_readout", buf, layers, cap, width, ld, start, fill, wind_cs, dir_geom, dirs, n, interp);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (n >= TB) {
    for (int d0 = 0; d0 < dirs; d0 += TRD) {
      for (int l0 = 0; l0 < layers; l0 += TRL) {
        TileArgs a;
        rc = pack_tile(a, buf, layers, cap, width, ld, start, fill, wind_cs, dir_geom, dirs, n, d0, l0);
        if (rc) return rc;
        a.out = out + (size_t)d0 * n * n;
        a.accumulate = (accumulate || l0 > 0) ? 1 : 0;
        a.scale = scale;
        const int tx = (n + TB - 1) / TB, nt = tx * tx * a.D;
        if (interp == 0)
          k_readout_tile<2><<<nt, 256, 0, st>>>(a, tx, nt);
        else if (interp == 1)
          k_readout_tile<4><<<nt, 256, 0, st>>>(a, tx, nt);
        else
          k_readout_tile<6><<<nt, 256, 0, st>>>(a, tx, nt);
        PT_LAUNCH_CHECK();
      }
    }
    return PT_OK;
  }
  for (int d0 = 0; d0 < dirs; d0 += RD) {
    for (int l0 = 0; l0 < layers; l0 += RL) {
      ReadoutArgs a;
      rc = pack_readout(a, buf, layers, cap, width, ld, start, fill, wind_cs, dir_geom, dirs, n, d0, l0);
      if (rc) return rc;
      a.out = out + (size_t)d0 * n * n;
      a.accumulate = (accumulate || l0 > 0) ? 1 : 0;
      a.scale = scale;
      const dim3 g((n + 31) / 32, (n + 7) / 8, a.D);
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

static size_t readout_adj_blocks(int n) { return (size_t)((n + 31) / 32) * ((n + 7) / 8); }

extern "C" long long pt_readout_adjoint_work_elems(int n) {
  if (n < 1) return -1;
  return (long long)((size_t)RD * readout_adj_blocks(n) * RL * 3);
}

extern "C" int pt_readout_adjoint(const float* buf, int layers, long long cap, int width, int ld, const int* start,
                                  const long long* fill, const double* wind_cs, const double* dir_geom, int dirs,
                                  int n, int interp, float scale, const float* grad, double* out, float* work,
                                  void* stream) {
  if (!grad || !out || !work) return fail(PT_ERR_ARG, "pt_readout_adjoint: NULL array");
  int rc = readout_check("pt_readout_adjoint", buf, layers, cap, width, ld, start, fill, wind_cs, dir_geom, dirs,
                         n, interp);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int nblk = (int)readout_adj_blocks(n);
  for (int d0 = 0; d0 < dirs; d0 += RD) {
    for (int l0 = 0; l0 < layers; l0 += RL) {
      ReadoutArgs a;
      rc = pack_readout(a, buf, layers, cap, width, ld, start, fill, wind_cs, dir_geom, dirs, n, d0, l0);
      if (rc) return rc;
      a.scale = scale;
      const float* g = grad + (size_t)d0 * n * n;
      const dim3 grid((n + 31) / 32, (n + 7) / 8, a.D);
      if (interp == 0)
        k_readout_adj<2><<<grid, 256, 0, st>>>(a, g, work);
      else if (interp == 1)
        k_readout_adj<4><<<grid, 256, 0, st>>>(a, g, work);
      else
        k_readout_adj<6><<<grid, 256, 0, st>>>(a, g, work);
      PT_LAUNCH_CHECK();
      const int cnt = a.D * a.L * 3;
      k_readout_adj_final<<<(cnt + 63) / 64, 64, 0, st>>>(work, a.D, a.L, nblk, layers,
                                                           out + ((size_t)d0 * layers + l0) * 3);
      PT_LAUNCH_CHECK();
    }
  }
  return PT_OK;
}

// Row blend into a ring: dst row (dst_start + r) mod dst_cap =
// a * itself + b * src row r, for r < rows, over `width` floats (a = 0
// writes b * src without reading dst).  The extruder's boiling blend
// (src/pyturb/extrude.py:740-752) and the seeding of fresh windows.
__global__ void k_rows_blend(float* __restrict__ dst, long long dst_pitch, int dst_start, int dst_cap,
                             const float* __restrict__ src, long long src_pitch, int width, float a, float b) {
  const int r = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= width) return;
  int pr = dst_start + r;
  pr %= dst_cap;
  float* d = dst + (size_t)pr * dst_pitch + c;
  const float s = b * src[(size_t)r * src_pitch + c];
  *d = a == 0.f ? s : fmaf(a, *d, s);
}

extern "C" int pt_rows_blend(float* dst, long long dst_pitch, int dst_start, int dst_cap, const float* src,
                             long long src_pitch, int rows, int width, float a, float b, void* stream) {
  if (!dst || !src) return fail(PT_ERR_ARG, "pt_rows_blend: NULL array");
  if (rows < 0 || width < 1 || dst_cap < 1 || dst_start < 0 || dst_start >= dst_cap || rows > dst_cap ||
      dst_pitch < width || src_pitch < width)
    return fail(PT_ERR_ARG, "pt_rows_blend: bad sizes");
  if (rows == 0) return PT_OK;
  if (rows > 65535) return fail(PT_ERR_ARG, "pt_rows_blend: more than 65535 rows");
  k_rows_blend<<<dim3((width + 255) / 256, rows), 256, 0, (hipStream_t)stream>>>(dst, dst_pitch, dst_start, dst_cap,
                                                                                  src, src_pitch, width, a, b);
  PT_LAUNCH_CHECK();
  return PT_OK;
}

// Workgroups of k_extrude_step<m> that the current device holds at once
// (occupancy per CU x CUs), cached per device and stencil; -1 on error.
//
// The step is an ordinary launch on the caller's stream, not a cooperative
// one: a cooperative launch goes to a separate hardware queue, and each hop
// between it and the stream's queue cost about 9 us per step (the readout
// runs on the stream).  The grid barrier only needs every workgroup of the
// launch to become resident eventually, and with grid <= this capacity it
// does: the stream is in order, so the previous kernel of the stream has
// finished; no workgroup of this launch holds a slot that another of its
// workgroups needs; and a CU held by another stream's kernel frees up when
// that kernel's workgroups end, which does not depend on this launch.  The
// bounded wait (spin_limit polls, then the abort word) still ends every
// workgroup if that takes too long.
static int step_capacity(const void* kern, int m) {
  static int cache[64][3];  // [device][m]: capacity + 1, 0 = not yet known
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
  if (cache[dev][m] > 0) return cache[dev][m] - 1;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * XS_WAVES, 0) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  cache[dev][m] = per_cu * cus + 1;
  return per_cu * cus;
}

extern "C" int pt_extrude_step(const float* A, const float* B, int width, int ld, int stencil_rows, int far_points,
                               const int* far_rows_back, float* buf, long long cap, int layers, const int* first_row,
                               const int* rows, const float* const* beta, const long long* beta_pitch,
                               const float* bscale, void* sync, int* status, unsigned spin_limit,
                               void* stream) {
  if (!A || !B || !buf || !first_row || !rows || !beta || !beta_pitch || !bscale || !sync || !status ||
      (far_points && !far_rows_back))
    return fail(PT_ERR_ARG, "pt_extrude_step: NULL array");
  if (width < 4 || ld < width || (ld & 3) || stencil_rows < 1 || stencil_rows > MMAX || layers < 1 || cap < 1)
    return fail(PT_ERR_ARG, "pt_extrude_step: bad sizes");
  if (ld & 15) return fail(PT_ERR_UNSUPPORTED, "pt_extrude_step: row pitch not a multiple of 16 floats");
  if (far_points < 0 || far_points % FAR_COLS || far_points > FAR_COLS * FAR_RMAX)
    return fail(PT_ERR_ARG, "pt_extrude_step: far_points must be a multiple of %d in [0, %d]", FAR_COLS,
                FAR_COLS * FAR_RMAX);
  if (!aligned16(A) || !aligned16(B) || !aligned16(buf) || ((uintptr_t)sync & 7))
    return fail(PT_ERR_ARG, "pt_extrude_step: A, B and buf must be 16-byte aligned, sync 8-byte aligned");
  StepArgs a = {};
  a.A = A;
  a.B = B;
  a.buf = buf;
  a.ld = ld;
  a.W = width;
  a.m = stencil_rows;
  a.nfar = far_points;
  a.ldm = (long long)stencil_rows * ld + ((far_points + 15) / 16) * 16;
  a.cap = cap;
  a.L = layers;
  a.cw = ((ld + 15) / 16 + MF_WAVES - 1) / MF_WAVES;
  a.sync = (unsigned long long*)sync;
  a.status = status;
  a.spin_limit = spin_limit;
  if ((long long)layers * cap * ld >= (1LL << 31))
    return fail(PT_ERR_UNSUPPORTED, "pt_extrude_step: rings of 2^31 floats or more");
  if (layers > XS_LMAX || a.cw > XS_CWMAX || stencil_rows > 2)
    return fail(PT_ERR_UNSUPPORTED,
                "pt_extrude_step: more than %d layers, more than 2 stencil rows or a row too wide for the register slice",
                XS_LMAX);
  for (int j = 0; j < FAR_RMAX; ++j) a.far_d[j] = far_points && j < far_points / FAR_COLS ? far_rows_back[j] : 0;
  // round r reads its far rows before the barrier that ends round r - 1, so
  // they must lie behind the newest stencil row (written in round r - 1)
  for (int j = 0; j < far_points / FAR_COLS; ++j)
    if (far_rows_back[j] < 1)
      return fail(PT_ERR_UNSUPPORTED, "pt_extrude_step: far row %d is not behind the newest stencil row", j);
  int nvec = 0, rounds = 0;
  for (int l = 0; l < layers; ++l) {
    if (rows[l] < 0 || first_row[l] < 0 || first_row[l] >= cap) return fail(PT_ERR_ARG, "pt_extrude_step: bad rows");
    if (rows[l] > 0 && (!beta[l] || ((uintptr_t)beta[l] & 15) || (beta_pitch[l] & 3) || beta_pitch[l] < width))
      return fail(PT_ERR_ARG,
                  "pt_extrude_step: layer %d innovations NULL, not 16-byte aligned, pitch not a multiple of 4 or "
                  "below width", l);
    a.t0[l] = first_row[l];
    a.cnt[l] = rows[l];
    a.boff[l] = nvec;
    a.beta[l] = beta[l];
    a.bpitch[l] = beta_pitch[l];
    a.scale[l] = bscale[l];
    nvec += rows[l];
    rounds = rows[l] > rounds ? rows[l] : rounds;
  }
  if (nvec > XS_VMAX) return fail(PT_ERR_UNSUPPORTED, "pt_extrude_step: more than %d new rows per step", XS_VMAX);
  if (rounds > XS_RMAX)
    return fail(PT_ERR_UNSUPPORTED, "pt_extrude_step: more than %d new rows of one layer per step", XS_RMAX);
  if (rounds == 0) return PT_OK;
  a.nvec = nvec;
  a.rounds = rounds;
  const int grid = (width + 15) / 16;
  if (grid > XS_GRIDMAX) return fail(PT_ERR_UNSUPPORTED, "pt_extrude_step: width over %d", 16 * XS_GRIDMAX);
  void* args[] = {&a};
  const void* kern = stencil_rows == 1 ? (const void*)k_extrude_step<1> : (const void*)k_extrude_step<2>;
  const int resident = step_capacity(kern, stencil_rows);
  if (resident < 0) return fail(PT_ERR_HIP, "pt_extrude_step: occupancy query failed");
  if (grid > resident)
    return fail(PT_ERR_UNSUPPORTED, "pt_extrude_step: %d workgroups, at most %d resident at once", grid, resident);
  hipError_t e = hipLaunchKernel(kern, dim3(grid), dim3(64 * XS_WAVES), args, 0, (hipStream_t)stream);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(PT_ERR_HIP, "pt_extrude_step: launch of %d workgroups: %s", grid, hipGetErrorString(e));
  }
  return PT_OK;
}
