// Launch-shape variants of the generate row and column passes, and the
// tuning state that selects them (pt_tune_*).
//
// Knob PT_TUNE_GEN_ROWS (k_gen_rows): threads per workgroup x largest radix
//   variant 0: 256 radix 8 (default), 1: 128 radix 8, 2: 512 radix 8,
//           3: 256 radix 4, 4: 512 radix 4.
// Knob PT_TUNE_GEN_COLS (k_gen_cols_set, Monte-Carlo screens and frame
// pass B): threads per workgroup x workgroup order x largest radix
//   variant 0: 1024 XCD-contiguous radix 8 (default), 1: 512 XCD, 2: 256 XCD,
//           3: 1024 flat, 4: 512 flat, 5: 256 flat (radix 8),
//           6: 1024 XCD radix 4, 7: 1024 flat radix 4.
// The radix-4 plans take log2(N)/2 LDS round trips of two radix-4
// butterflies per thread instead of log2(N)/3 of one radix-8 butterfly; they
// round differently from the radix-8 plan (within 1e-6 relative RMS).  The
// column selection applies to the row-first frame engine's pass B too.
// Knob PT_TUNE_FRAMES (fused frame engine): transform order
//   variant 0: column-first (pass A' column tiles, pass B' rows + epilogue),
//           1: row-first (pass A rows, pass B column tiles + epilogue).
// Thread counts are clamped to [N/8, N*N/8]; pt_tune_info reports the
// effective shape (for PT_TUNE_FRAMES: pass A's threads and rows or
// columns per workgroup), so variants that coincide at a size can be skipped.
//
// The selection is process-wide host state, read when a call launches: a
// change applies to later calls, not to work already enqueued, and is not
// synchronised against calls made from other threads at the same time.
#include <hip/hip_runtime.h>

#include "gen_kernels.hpp"

namespace pt {

namespace {

constexpr int kKnobs = 3;
constexpr int kLogMin = 5, kLogMax = 11;  // N = 32 .. 2048
constexpr int kRowsThreads[] = {256, 128, 512, 256, 512};
constexpr int kRowsRadix[] = {8, 8, 8, 4, 4};
constexpr int kColsThreads[] = {1024, 512, 256, 1024, 512, 256, 1024, 1024};
constexpr bool kColsXcd[] = {true, true, true, false, false, false, true, false};
constexpr int kColsRadix[] = {8, 8, 8, 8, 8, 8, 4, 4};
constexpr int kVariants[kKnobs] = {5, 8, 2};

int g_choice[kKnobs][kLogMax + 1] = {};

int log2_pow2(int n) {
  if (n < (1 << kLogMin) || n > (1 << kLogMax) || (n & (n - 1))) return -1;
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

template <int N>
int rows_variant(int v, const float* nre, ptrdiff_t im_off, size_t gstride, int G, const float* amp, float2* work,
                 int n_fft, hipStream_t st) {
#define PT_ROWS(TR, RX)                                                                                 \
  k_gen_rows<N, TR, RX><<<dim3(N / RowsCfg<N, TR>::M, n_fft), RowsCfg<N, TR>::T, 0, st>>>(nre, im_off, \
                                                                                          gstride, G, amp, work)
  switch (v) {
    case 1: PT_ROWS(128, 8); break;
    case 2: PT_ROWS(512, 8); break;
    case 3: PT_ROWS(256, 4); break;
    case 4: PT_ROWS(512, 4); break;
    default: return fail(PT_ERR_ARG, "gen_rows: no variant %d", v);
  }
#undef PT_ROWS
  PT_LAUNCH_CHECK();
  return PT_OK;
}

template <int N>
int cols_variant(int v, const GenArgs& a, hipStream_t st) {
#define PT_COLS(TC, X, RX) \
  k_gen_cols_set<N, TC, X, RX><<<dim3(N / ColsCfg<N, TC>::M, a.n_fft), ColsCfg<N, TC>::T, 0, st>>>(a)
  switch (v) {
    case 1: PT_COLS(512, true, 8); break;
    case 2: PT_COLS(256, true, 8); break;
    case 3: PT_COLS(1024, false, 8); break;
    case 4: PT_COLS(512, false, 8); break;
    case 5: PT_COLS(256, false, 8); break;
    case 6: PT_COLS(1024, true, 4); break;
    case 7: PT_COLS(1024, false, 4); break;
    default: return fail(PT_ERR_ARG, "gen_cols: no variant %d", v);
  }
#undef PT_COLS
  PT_LAUNCH_CHECK();
  return PT_OK;
}

#define PT_TUNE_SWITCH(n, CALL) \
  switch (n) {                  \
    case 32: CALL(32);          \
    case 64: CALL(64);          \
    case 128: CALL(128);        \
    case 256: CALL(256);        \
    case 512: CALL(512);        \
    case 1024: CALL(1024);      \
    case 2048: CALL(2048);      \
    default: break;             \
  }

}  // namespace

int tune_choice(int knob, int n) {
  const int l = log2_pow2(n);
  return (knob >= 0 && knob < kKnobs && l >= 0) ? g_choice[knob][l] : 0;
}

int launch_gen_rows_variant(int n, int variant, const float* nre, ptrdiff_t im_off, size_t gstride, int G,
                            const float* amp, float2* work, int n_fft, hipStream_t st) {
#define CALL(N) return rows_variant<N>(variant, nre, im_off, gstride, G, amp, work, n_fft, st);
  PT_TUNE_SWITCH(n, CALL)
#undef CALL
  return fail(PT_ERR_UNSUPPORTED, "gen_rows: n=%d", n);
}

int launch_gen_cols_variant(int n, int variant, const GenArgs& a, hipStream_t st) {
#define CALL(N) return cols_variant<N>(variant, a, st);
  PT_TUNE_SWITCH(n, CALL)
#undef CALL
  return fail(PT_ERR_UNSUPPORTED, "gen_cols: n=%d", n);
}

}  // namespace pt

using namespace pt;

extern "C" int pt_tune_variants(int knob) {
  if (knob < 0 || knob >= kKnobs) return -fail(PT_ERR_ARG, "pt_tune_variants: unknown knob %d", knob);
  return kVariants[knob];
}

extern "C" int pt_tune_info(int knob, int n, int variant, int* threads, int* tile, int* xcd_order, int* radix) {
  if (knob < 0 || knob >= kKnobs) return fail(PT_ERR_ARG, "pt_tune_info: unknown knob %d", knob);
  if (log2_pow2(n) < 0) return fail(PT_ERR_ARG, "pt_tune_info: n=%d is not a power of two in 32..2048", n);
  if (variant < 0 || variant >= kVariants[knob])
    return fail(PT_ERR_ARG, "pt_tune_info: knob %d has no variant %d", knob, variant);
  int t;
  if (knob == PT_TUNE_FRAMES)  // pass A: 16-column tiles up to n = 512, else 1024 threads; 256-thread rows
    t = variant == 0 ? (n <= 512 ? 2 * n : 1024) : clamp_threads(n, 256);
  else
    t = clamp_threads(n, knob == PT_TUNE_GEN_ROWS ? kRowsThreads[variant] : kColsThreads[variant]);
  if (threads) *threads = t;
  if (tile) *tile = 8 * t / n;
  if (xcd_order) *xcd_order = knob == PT_TUNE_GEN_COLS ? (kColsXcd[variant] ? 1 : 0) : 0;
  if (radix)
    *radix = knob == PT_TUNE_GEN_ROWS ? kRowsRadix[variant] : knob == PT_TUNE_GEN_COLS ? kColsRadix[variant] : 8;
  return PT_OK;
}

extern "C" int pt_tune_set(int knob, int n, int variant) {
  if (knob < 0 || knob >= kKnobs) return fail(PT_ERR_ARG, "pt_tune_set: unknown knob %d", knob);
  const int l = log2_pow2(n);
  if (l < 0) return fail(PT_ERR_ARG, "pt_tune_set: n=%d is not a power of two in 32..2048", n);
  if (variant < 0 || variant >= kVariants[knob])
    return fail(PT_ERR_ARG, "pt_tune_set: knob %d has no variant %d", knob, variant);
  g_choice[knob][l] = variant;
  return PT_OK;
}

extern "C" int pt_tune_get(int knob, int n) {
  if (knob < 0 || knob >= kKnobs) return -fail(PT_ERR_ARG, "pt_tune_get: unknown knob %d", knob);
  if (log2_pow2(n) < 0) return -fail(PT_ERR_ARG, "pt_tune_get: n=%d is not a power of two in 32..2048", n);
  return g_choice[knob][log2_pow2(n)];
}
