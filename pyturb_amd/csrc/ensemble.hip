// Monte-Carlo ensembles of several grid sizes in three persistent launches
// (pt_generate_ensemble).
//
// pt_generate_groups runs one size class per call: a row pass, the
// subharmonic basis and prep, and a column pass, four launches per class.
// Here every class of the ensemble shares each launch:
//   * k_ens_rows walks a work list of (class, transform, row block) items,
//     then the subharmonic basis items of every class;
//   * k_ens_prep builds every screen's subharmonic tables (one block per
//     screen of every class);
//   * the column pass: per class, the tuned k_gen_cols_set launch of
//     pt_generate_groups (the default), or k_ens_cols, a persistent kernel
//     over every class's (transform, column tile) items.
// The persistent row pass runs 2 % faster than the per-class row passes of
// the mixed 256^2 / 512^2 / 1024^2 ensemble; the persistent column pass moves
// the same HBM bytes as the per-class ones but ran 42 % slower (2.82 ms
// against 1.99 ms), so the per-class column launches are the default.
// The row and column kernels are persistent: their grids are what the device
// holds at once (occupancy x CUs, a multiple of the 8 XCDs), and workgroup g
// takes items g, g + grid, g + 2 grid, ...  Each class's item range starts at
// a multiple of 8 and the column tiles of a transform are remapped with
// xcd_tile, so (with workgroup g on XCD g mod 8) the neighbouring column tiles
// of a transform, which share 128-byte lines of its intermediate, meet in one
// L2, as in the per-class column kernel.  The classes are walked largest
// first: the n = 1024 column tiles (8 columns, 64 bytes of each line) rely on
// that sharing, which holds while the workgroups still run in step; after
// the ~100 items of the smaller classes they drift apart (measured: 1024^2
// after 512^2 ran 39 % slower than its own launches, first 10 %).
// (Taking items from per-XCD queues with atomic counters instead was
// measured slower still, 40-50 % over the per-class launches.)
//
// An item runs the per-class kernels' device code for its size
// (gen_rows_item / gen_cols_item with the default shapes and the radix-8
// plan), so every screen is bit-identical to pt_generate_groups'.  The LDS
// of the persistent kernels is dynamic, sized for the largest class, and
// the twiddle table is rebuilt when a workgroup moves to another size.  The
// row kernel is capped at the per-class kernel's occupancy (6 waves per
// SIMD), so that the loop does not keep hoisted per-thread addresses in
// registers at the cost of resident workgroups.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "fft_core.hpp"
#include "gen_kernels.hpp"
#include "sh_epilogue.hpp"

namespace pt {

namespace {

constexpr int kEnsMax = 8;        // size classes per call
constexpr int kEnsRowsT = kRowsT;  // row workgroup: 256 threads, M = 2048 / n rows
constexpr int kEnsColsT = kColsT;  // column workgroup: 1024 threads, M = 8192 / n columns
constexpr int kEnsNMin = 128, kEnsNMax = 2048;

struct EnsClass {
  // row pass (pt_generate_groups' noise layout: (G, 2, n_fft, n, n))
  const float* nre;
  ptrdiff_t im_off;
  size_t gstride;
  const float* amp;
  float2* work;
  // subharmonic prep
  const float* sh_noise;
  const double* sh_amp;
  int ubasis;
  // column pass
  GenArgs g;
  int n;
  int row_first, row_items, basis_items;  // row-kernel items: transforms x row blocks, then basis
  int prep_first;                         // prep blocks: one per screen
  int col_first, col_items;               // column-kernel items: transforms x column tiles
};

struct EnsArgs {
  EnsClass c[kEnsMax];
  int classes, G, P;
  int row_total, prep_total, col_total;
};
static_assert(sizeof(EnsArgs) <= 4096, "kernel argument block");

// The thread index as a value the compiler cannot see through: an item's
// per-thread addresses are then computed inside the item, as in the
// per-class kernels, instead of being hoisted out of the item loop for
// every size path at once and kept live in registers (or spilled).
__device__ __forceinline__ int fresh(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// class of item `it` given each class's first item (ascending)
template <class First>
__device__ __forceinline__ int ens_class(const EnsArgs& a, int it, First first) {
  int k = 0;
  while (k + 1 < a.classes && it >= first(a.c[k + 1])) ++k;
  return k;
}

__device__ __forceinline__ void ens_twiddles(float2* tw, int n, int tid, int T) {
  switch (n) {
    case 128: init_twiddles<128, 1>(tw, tid, T); break;
    case 256: init_twiddles<256, 1>(tw, tid, T); break;
    case 512: init_twiddles<512, 1>(tw, tid, T); break;
    case 1024: init_twiddles<1024, 1>(tw, tid, T); break;
    default: init_twiddles<2048, 1>(tw, tid, T); break;
  }
}

template <int N>
__device__ __forceinline__ void ens_rows_item(const EnsClass& c, int G, int local, float2* lds, const float2* tw,
                                              int tid) {
  constexpr int T = kEnsRowsT, M = RowsCfg<N, T>::M, NB = N / M;
  gen_rows_item<N, T, M, 8>(c.nre, c.im_off, c.gstride, G, c.amp, c.work, local % NB, local / NB, lds, tw, tid);
}

__global__ __launch_bounds__(kEnsRowsT) __attribute__((amdgpu_waves_per_eu(6))) void k_ens_rows(EnsArgs a) {
  constexpr int T = kEnsRowsT;
  extern __shared__ __align__(16) unsigned char dyn[];
  float2* lds = reinterpret_cast<float2*>(dyn);  // M (n + n/8) = 9 T float2 at every n
  float2* tw = lds + 9 * T;
  const int tid = threadIdx.x;
  int tw_n = 0;
  for (int it = blockIdx.x; it < a.row_total; it += gridDim.x) {
    const int k = ens_class(a, it, [](const EnsClass& c) { return c.row_first; });
    const EnsClass& c = a.c[k];
    const int local = it - c.row_first;
    if (local >= c.row_items) {  // the class's subharmonic basis U[p][x] = e_p(x)
      const int idx = (local - c.row_items) * T + tid;
      if (idx < a.P * c.n) c.g.tab.U[idx] = sh_basis_e(idx / c.n, idx % c.n, c.n);
      continue;
    }
    __syncthreads();  // the previous item's last reads of lds and tw are done
    const int t = fresh(tid);
    if (c.n != tw_n) {
      ens_twiddles(tw, c.n, t, T);
      tw_n = c.n;
      __syncthreads();
    }
    switch (c.n) {
      case 128: ens_rows_item<128>(c, a.G, local, lds, tw, t); break;
      case 256: ens_rows_item<256>(c, a.G, local, lds, tw, t); break;
      case 512: ens_rows_item<512>(c, a.G, local, lds, tw, t); break;
      case 1024: ens_rows_item<1024>(c, a.G, local, lds, tw, t); break;
      default: ens_rows_item<2048>(c, a.G, local, lds, tw, t); break;
    }
  }
}

__global__ __launch_bounds__(256) void k_ens_prep(EnsArgs a) {
  const int k = ens_class(a, (int)blockIdx.x, [](const EnsClass& c) { return c.prep_first; });
  const EnsClass& c = a.c[k];
  gen_prep_set(c.sh_noise, c.sh_amp, c.g.count, a.P, c.n, a.G, c.g.tab, c.ubasis, (int)blockIdx.x - c.prep_first);
}

// A class's column LDS: the per-class kernel's tile and staging block (the
// block beside the tile up to n = 512, in it at n = 1024 and above), then
// that class's twiddles, so the largest need (78 KiB for the 512^2 class)
// keeps two workgroups per CU.
template <int N>
__host__ __device__ constexpr size_t ens_cols_smem_n() {
  return gen_cols_smem<N, kEnsColsT, ColsCfg<N, kEnsColsT>::M>() + (size_t)N * sizeof(float2);
}
__host__ __device__ constexpr size_t cmax(size_t x, size_t y) { return x > y ? x : y; }
// tile, staging block and twiddles of a size class
__host__ __device__ constexpr size_t ens_cols_smem(int n) {
  return n == 128 ? ens_cols_smem_n<128>()
         : n == 256 ? ens_cols_smem_n<256>()
         : n == 512 ? ens_cols_smem_n<512>()
         : n == 1024 ? ens_cols_smem_n<1024>()
                     : ens_cols_smem_n<2048>();
}
template <int N>
__device__ __forceinline__ float2* ens_cols_tw(unsigned char* smem) {
  return reinterpret_cast<float2*>(smem + gen_cols_smem<N, kEnsColsT, ColsCfg<N, kEnsColsT>::M>());
}

template <int N>
__device__ __forceinline__ void ens_cols_item(const EnsClass& c, int local, unsigned char* smem, int tid,
                                              bool fresh_tw) {
  constexpr int T = kEnsColsT, M = ColsCfg<N, T>::M, NT = N / M;
  float2* tw = ens_cols_tw<N>(smem);
  if (fresh_tw) {
    init_twiddles<N, 1>(tw, tid, T);
    __syncthreads();
  }
  gen_cols_item<N, false, T, M, 8>(c.g, smem, tw, xcd_tile(local % NT, NT) * M, local / NT, tid);
}

__global__ __launch_bounds__(kEnsColsT) __attribute__((amdgpu_waves_per_eu(8))) void k_ens_cols(EnsArgs a) {
  constexpr int T = kEnsColsT;
  extern __shared__ __align__(16) unsigned char dyn[];
  // the tile and staging block at the start of LDS (compile-time addresses,
  // as in the per-class kernel), the class's twiddles after them
  unsigned char* smem = dyn;
  const int tid = threadIdx.x;
  int tw_n = 0;
  for (int it = blockIdx.x; it < a.col_total; it += gridDim.x) {
    const int k = ens_class(a, it, [](const EnsClass& c) { return c.col_first; });
    const EnsClass& c = a.c[k];
    const int local = it - c.col_first;
    if (local >= c.col_items) continue;  // padding to the next multiple of 8
    __syncthreads();  // the previous item's last reads of smem are done
    const int t = fresh(tid);
    const bool new_n = c.n != tw_n;
    tw_n = c.n;
    switch (c.n) {
      case 128: ens_cols_item<128>(c, local, smem, t, new_n); break;
      case 256: ens_cols_item<256>(c, local, smem, t, new_n); break;
      case 512: ens_cols_item<512>(c, local, smem, t, new_n); break;
      case 1024: ens_cols_item<1024>(c, local, smem, t, new_n); break;
      default: ens_cols_item<2048>(c, local, smem, t, new_n); break;
    }
  }
}

// Persistent grid of `kern`: workgroups resident at once, a multiple of 8;
// -1 on error.  Raises the kernel's dynamic-LDS limit first when needed.
int ens_grid(const void* kern, int threads, size_t dyn) {
  int dev = 0, per_cu = 0, cus = 0;
  if (dyn > 64 * 1024 &&
      hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess)
    return -1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, dyn) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return -1;
  const int g = per_cu * cus / 8 * 8;
  return g > 0 ? g : (per_cu * cus > 0 ? per_cu * cus : -1);
}


}  // namespace

}  // namespace pt

using namespace pt;

extern "C" int pt_generate_ensemble(int classes, const int* n, const int* count, const float* const* noise,
                                    const float* const* amplitude, const float* const* sh_noise,
                                    const double* const* sh_amp, int groups, int levels,
                                    const float* const* screen_scale, float scale, float* const* out,
                                    void* const* work, int persistent_cols, void* stream) {
  if (classes < 1 || classes > kEnsMax)
    return fail(PT_ERR_ARG, "pt_generate_ensemble: 1..%d size classes", kEnsMax);
  if (!n || !count || !noise || !amplitude || !out || !work)
    return fail(PT_ERR_ARG, "pt_generate_ensemble: NULL array");
  if (groups < 1 || groups > 64) return fail(PT_ERR_ARG, "pt_generate_ensemble: 1..64 groups");
  if (levels < 0 || (levels > 0 && (!sh_noise || !sh_amp)))
    return fail(PT_ERR_ARG, "pt_generate_ensemble: bad subharmonics");
  if (levels > PREG)
    return fail(PT_ERR_UNSUPPORTED, "pt_generate_ensemble: more than %d subharmonic levels", PREG);
  EnsArgs a = {};
  a.classes = classes;
  a.G = groups;
  a.P = levels;
  int nmax = 0;
  // walk order: largest n first (stable)
  int ord[kEnsMax];
  for (int k = 0; k < classes; ++k) ord[k] = k;
  for (int i = 1; i < classes; ++i)
    for (int j = i; j > 0 && n[ord[j]] > n[ord[j - 1]]; --j) {
      const int tmp = ord[j];
      ord[j] = ord[j - 1];
      ord[j - 1] = tmp;
    }
  long long rows = 0, prep = 0, cols = 0;
  for (int w = 0; w < classes; ++w) {
    const int k = ord[w];
    const int nk = n[k];
    if (nk < kEnsNMin || nk > kEnsNMax || (nk & (nk - 1)))
      return fail(PT_ERR_UNSUPPORTED, "pt_generate_ensemble: class %d: n=%d (power of two %d..%d only)", k, nk,
                  kEnsNMin, kEnsNMax);
    if (count[k] < 1 || (count[k] + 1) / 2 > 65535)
      return fail(PT_ERR_ARG, "pt_generate_ensemble: class %d: count must be 1..131070", k);
    if (!noise[k] || !amplitude[k] || !out[k] || !work[k] || (levels > 0 && (!sh_noise[k] || !sh_amp[k])))
      return fail(PT_ERR_ARG, "pt_generate_ensemble: class %d: NULL array", k);
    EnsClass& c = a.c[w];
    const int n_fft = (count[k] + 1) / 2;
    const size_t plane = (size_t)nk * nk;
    c.nre = noise[k];
    c.im_off = (ptrdiff_t)(n_fft * plane);
    c.gstride = 2 * (size_t)n_fft * plane;
    c.amp = amplitude[k];
    c.work = (float2*)work[k];
    c.sh_noise = levels ? sh_noise[k] : nullptr;
    c.sh_amp = levels ? sh_amp[k] : nullptr;
    c.ubasis = count[k] >= 16;  // as pt_generate_groups: the basis table first for 16+ screens
    c.n = nk;
    GenArgs& g = c.g;
    g.work = c.work;
    g.out = out[k];
    g.screen_scale = screen_scale ? screen_scale[k] : nullptr;
    g.tab = sub_tables(c.work + (size_t)n_fft * plane, levels, nk, count[k]);
    g.count = count[k];
    g.n_fft = n_fft;
    g.levels = levels;
    g.off = 0;
    g.n_out = nk;
    g.accumulate = 0;
    g.s_mul = 1;
    g.s_add = n_fft;
    g.scale = scale;
    const int rm = 8 * kEnsRowsT / nk;    // rows per row item
    const int cm = 8 * kEnsColsT / nk;    // columns per column item
    c.row_first = (int)rows;
    c.row_items = n_fft * (nk / rm);  // < 2^31: n_fft <= 65535, nk / rm <= 1024
    c.basis_items = (levels && c.ubasis) ? (levels * nk + kEnsRowsT - 1) / kEnsRowsT : 0;
    rows += (c.row_items + c.basis_items + 7) / 8 * 8;
    c.prep_first = (int)prep;
    prep += levels ? count[k] : 0;
    c.col_first = (int)cols;
    c.col_items = n_fft * (nk / cm);
    cols += (c.col_items + 7) / 8 * 8;
    nmax = nk > nmax ? nk : nmax;
  }
  if (rows >= (1LL << 31) || prep >= (1LL << 31) || cols >= (1LL << 31))
    return fail(PT_ERR_ARG, "pt_generate_ensemble: too many items");
  a.row_total = (int)rows;
  a.prep_total = (int)prep;
  a.col_total = (int)cols;
  size_t cols_dyn = 0;
  for (int k = 0; k < classes; ++k) cols_dyn = cmax(cols_dyn, ens_cols_smem(n[k]));
  const size_t rows_dyn = (size_t)(9 * kEnsRowsT + nmax) * sizeof(float2);
  hipStream_t st = (hipStream_t)stream;
  const int gr = ens_grid((const void*)k_ens_rows, kEnsRowsT, rows_dyn);
  const int gc = persistent_cols ? ens_grid((const void*)k_ens_cols, kEnsColsT, cols_dyn) : 0;
  if (gr < 0 || gc < 0) {
    (void)hipGetLastError();
    return fail(PT_ERR_HIP, "pt_generate_ensemble: occupancy query failed (LDS %zu / %zu B)", rows_dyn, cols_dyn);
  }
  k_ens_rows<<<gr, kEnsRowsT, rows_dyn, st>>>(a);
  PT_LAUNCH_CHECK();
  if (a.prep_total > 0) {
    k_ens_prep<<<a.prep_total, 256, 0, st>>>(a);
    PT_LAUNCH_CHECK();
  }
  if (persistent_cols) {
    k_ens_cols<<<gc, kEnsColsT, cols_dyn, st>>>(a);
    PT_LAUNCH_CHECK();
    return PT_OK;
  }
  for (int w = 0; w < classes; ++w) {  // pt_generate_groups' column launch of each class
    const int rc = launch_gen_cols(a.c[w].n, a.c[w].g, st);
    if (rc) return rc;
  }
  return PT_OK;
}
