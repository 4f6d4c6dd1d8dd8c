// pt_spectral_frames without a caller workspace, for callers that cannot
// allocate device memory (code under HIP-graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// The workspace of the frame engine (pair intermediate, folded stack, tables)
// lives in `scratch`, a file-scope device array of PT_NOALLOC_SCRATCH_BYTES
// that every GPU the library runs on holds until the process ends.  In host
// code `scratch` names a host-side stand-in, not device memory, so the host
// never passes it to a kernel: on the first call on each device (with that
// device current) hipGetSymbolAddress gives that device's own copy of the
// array, and that address is kept per device and passed to the kernels as the
// workspace.  The entry point calls no hipMalloc, hipFree or allocating
// library routine.  Every call on a GPU shares that GPU's array; calls that
// overlap on one GPU (concurrent streams or threads) race on it.  Nothing
// in the frame engine reads workspace contents left by an earlier call (every
// section is written before it is read within the call), so no clearing is
// needed between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"

namespace {

constexpr size_t kScratchElems = PT_NOALLOC_SCRATCH_BYTES / 8;  // c64 elements
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float2 scratch[kScratchElems];

std::mutex g_mutex;
void* g_ptr[kMaxDevices] = {};

// That device's address of `scratch` (looked up once per device) into *out;
// a status code.
int scratch_ptr(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return pt::fail(PT_ERR_HIP, "pt_spectral_frames_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return pt::fail(PT_ERR_UNSUPPORTED, "pt_spectral_frames_noalloc: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return pt::fail(PT_ERR_HIP, "pt_spectral_frames_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return PT_OK;
}

}  // namespace

extern "C" int pt_spectral_frames_noalloc(const void* spectra, int layers, int n, double pixel_scale,
                                          const double* disp, int frames, const double* sh_coeffs, int levels,
                                          int crop_offset, int n_out, float scale, float* out, void* stream) {
  const long long need = pt_frames_work_elems(n, frames, layers, levels);
  if (need < 0) return pt::fail(PT_ERR_ARG, "pt_spectral_frames_noalloc: bad n/frames/layers/levels");
  if ((unsigned long long)need > kScratchElems)
    return pt::fail(PT_ERR_ARG,
                    "pt_spectral_frames_noalloc: the call needs %lld bytes of workspace, over the %lld-byte "
                    "static scratch (fewer frames per call, or pt_spectral_frames with a workspace)",
                    need * 8, (long long)PT_NOALLOC_SCRATCH_BYTES);
  void* work = nullptr;
  const int rc = scratch_ptr(&work);
  if (rc) return rc;
  return pt_spectral_frames(spectra, layers, n, pixel_scale, disp, frames, sh_coeffs, levels, crop_offset, n_out,
                            scale, out, work, stream);
}
