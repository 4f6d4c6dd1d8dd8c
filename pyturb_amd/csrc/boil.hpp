// Spectral boiling kernels (boil.hip): arguments and launcher.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pt {

struct BoilArgs {
  float2* spectra;            // (L, n, n) c64 state, updated in place
  const float* amp;           // (L, n, n) f32 PSD amplitude per layer
  const double* inv_tau;      // (L) 1 / tau_boil [1/s], 0 for a frozen layer
  const double* f_ref;        // (L) reference frequency [1/m]
  double2* coeffs;            // (P, L, 3, 3) c128 subharmonic state, in place
  const double* sh_amp;       // (P, L, 3, 3) f64 subharmonic amplitudes
  double2* coeff_frames;      // (frames, P, L, 3, 3) c128: coefficients of each frame
  const float* noise;         // (steps, 2, L, n, n) f32 caller normals, or null (Philox)
  const double* sh_noise;     // (steps, 2, P, L, 3, 3) f64 caller normals, or null
  const uint2* phase;         // (frames, L) fixed-point cycles (filled by launch_boil)
  float2* z;                  // (ceil(frames / 2), n, n) c64 frame-pair spectra
  int n, L, P, frames, steps;
  double dt, df;              // step [s]; grid frequency spacing 1 / (n dx) [1/m]
  unsigned long long step0;   // global index of the first step (Philox counter)
  uint32_t key0, key1;        // Philox key (the boiling seed)
};

// Phase table, the fused boil + layer-sum pass and the coefficient pass.
// n even (mirror pairs), L <= 16.
int launch_boil(const BoilArgs& b, const double* disp, double inv_period, hipStream_t st);

}  // namespace pt
