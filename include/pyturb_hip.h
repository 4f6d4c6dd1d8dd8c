/*
 * pyturb_hip.h -- C ABI of the MI355X (gfx950) phase-screen library.
 *
 * Every array argument is a DEVICE pointer (HBM) owned by the caller; sizes
 * are plain integers; `stream` is a hipStream_t passed as void* (NULL = the
 * null stream).  Calls are asynchronous on `stream` and return a status
 * code; on failure pt_last_error() returns a message.  No entry point
 * allocates device memory: workspaces are caller-owned and their sizes are
 * stated below.  Calls sharing a workspace on one stream serialise
 * correctly; concurrent calls must not share a workspace.
 *
 * Complex arrays are interleaved (re, im) float32 ("c64") or float64
 * ("c128").  Transforms are the reference's unnormalised inverse DFT,
 * ifft2(x) * n0 * n1 (sign +1); sign -1 is the forward DFT.
 *
 * Each entry point replaces the array-module call chain the reference runs
 * for the same step; the replaced code is cited (paths are relative to the
 * reference checkout).
 */
#ifndef PYTURB_HIP_H
#define PYTURB_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define PT_ABI_VERSION 6

#define PT_OK 0
#define PT_ERR_ARG 1         /* invalid argument (shape, count, NULL)      */
#define PT_ERR_HIP 2         /* HIP runtime error (launch, copy)           */
#define PT_ERR_UNSUPPORTED 3 /* size outside what the kernels implement    */

/* Launch-shape knobs of the autotuner (pt_tune_*). */
#define PT_TUNE_GEN_ROWS 0 /* row pass of pt_generate*: threads per workgroup     */
#define PT_TUNE_GEN_COLS 1 /* column pass of pt_generate* and frame pass B:       */
                           /* threads per workgroup x workgroup order             */
#define PT_TUNE_FRAMES 2   /* transform order of the fused frame engine:          */
                           /* 0 column-first, 1 row-first                         */

/* Last error message of the calling thread ("" if none). */
const char* pt_last_error(void);

/* PT_ABI_VERSION of the loaded library. */
int pt_abi_version(void);

/*
 * Workspace (in c64 elements) that pt_generate (kind 0, sets = count) or
 * pt_spectral_frames / pt_spectral_rows+cols (kind 1, sets = frames, any
 * layer count) needs for an n x n grid with `levels` subharmonic levels:
 * the FFT intermediate (doubled on the general-size path) followed by the
 * subharmonic tables.  -1 on bad arguments.  Host-only, no device work.
 */
long long pt_work_elems(int kind, int n, int sets, int levels);

/*
 * Exact workspace (c64 elements) of pt_spectral_frames, or of
 * pt_spectral_rows + pt_spectral_cols, for `layers` layers.  pt_work_elems
 * kind 1 is an upper bound of this for any layer count.  -1 on bad input.
 */
long long pt_frames_work_elems(int n, int frames, int layers, int levels);

/*
 * Unnormalised 2-D DFT over the last two axes of `in` (batch, n0, n1) c64,
 * written to `out` (same shape; may equal `in`).  Powers of two from 32 to
 * 2048 per axis run the LDS Stockham passes; other sizes run the direct
 * DFT path and then need `work` of batch*n0*n1 c64 (else may be NULL).
 * Replaces: src/pyturb/fourier.py:259, src/pyturb/atmosphere.py:918,978,
 *           src/pyturb/flow.py:110 (xp.fft.ifft2 * n^2).
 */
int pt_fft2(const void* in, void* out, void* work, int batch, int n0, int n1,
            int sign, void* stream);

/*
 * Fused Monte-Carlo screen generation, one PhaseScreen.generate(count) call
 * on injected noise (src/pyturb/fourier.py:233-282, and the crop/scale of
 * src/pyturb/atmosphere.py:1213-1219 when summing L0 groups).
 *
 *   noise_re, noise_im : (n_fft, n, n) f32, n_fft = (count+1)/2, the real and
 *                        imaginary white-noise planes (fourier.py:257-258)
 *   amplitude          : (n, n) f32 PSD filter (fourier.py:174-191)
 *   sh_noise           : (levels, 2, count, 3, 3) f32 subharmonic noise
 *                        (fourier.py:271-273), NULL when levels == 0
 *   sh_amp             : (levels, 3, 3) f64 subharmonic amplitudes
 *   screen_scale       : (count,) f32 per-screen factor or NULL; output is
 *                        (screen + low) * scale * screen_scale[s]
 *   out                : (count, n_out, n_out) f32; screen s pixel (x, y) is
 *                        full-grid pixel (x + crop_offset, y + crop_offset);
 *                        accumulate != 0 adds into out
 *   work               : pt_work_elems(0, n, count, levels) c64: the FFT
 *                        intermediate, then the subharmonic tables
 * Screen k < n_fft is the real part of FFT k, screen n_fft + k its
 * imaginary part (fourier.py:260).
 */
int pt_generate(const float* noise_re, const float* noise_im,
                const float* amplitude, const float* sh_noise,
                const double* sh_amp, int levels, int count, int n,
                int crop_offset, int n_out, float scale,
                const float* screen_scale, int accumulate, float* out,
                void* work, void* stream);

/*
 * Atmosphere.sample for several outer-scale groups in one pass
 * (src/pyturb/atmosphere.py:439-470,1191-1219: one generator per distinct
 * L0, their screens summed).  The transform and the subharmonic field are
 * linear, so the groups' filtered noise and subharmonic coefficients are
 * summed before a single transform per screen pair:
 *   out = (sum_g screen_g + low_g) * scale * screen_scale[s]
 *   noise     : (groups, 2, n_fft, n, n) f32, each group's real then
 *               imaginary planes (as pt_generate's noise_re / noise_im)
 *   amplitude : (groups, n, n) f32
 *   sh_noise  : (groups, levels, 2, count, 3, 3) f32, NULL when levels == 0
 *   sh_amp    : (groups, levels, 3, 3) f64
 *   work      : pt_work_elems(0, n, count, levels) c64
 * groups == 1 is pt_generate.  1 <= groups <= 64.
 */
int pt_generate_groups(const float* noise, const float* amplitude,
                       const float* sh_noise, const double* sh_amp, int groups,
                       int levels, int count, int n, int crop_offset,
                       int n_out, float scale, const float* screen_scale,
                       int accumulate, float* out, void* work, void* stream);

/*
 * A Monte-Carlo ensemble of several grid sizes (Atmosphere.sample or
 * PhaseScreen.generate once per size, src/pyturb/atmosphere.py:1191-1219,
 * fourier.py:233-282) with launches shared by every size class: a
 * persistent row-pass kernel over all classes' (transform, row block) items
 * and one subharmonic prep; then, with persistent_cols, a persistent
 * column-pass kernel over all classes' (transform, column tile) items, else
 * each class's column launch of pt_generate_groups.  Class k is
 * pt_generate_groups with crop_offset = 0, n_out = n[k], accumulate = 0,
 * and its screens are bit-identical to that call's under the default
 * launch shapes (the persistent passes run the default shapes and radix-8
 * plans; a tuned radix-4 plan of pt_generate_groups rounds differently):
 *   n[k]            : power of two, 128..2048 (else PT_ERR_UNSUPPORTED)
 *   count[k]        : screens of class k
 *   noise[k], amplitude[k], sh_noise[k], sh_amp[k], out[k]: as
 *                     pt_generate_groups' arrays of class k (HOST arrays of
 *                     device pointers); sh_noise / sh_amp may be NULL when
 *                     levels == 0
 *   screen_scale    : NULL, or a host array of per-class device pointers
 *                     (each NULL or (count[k]) f32)
 *   work[k]         : pt_work_elems(0, n[k], count[k], levels) c64 each
 * 1 <= classes <= 8, 1 <= groups <= 64, levels <= 12.
 */
int pt_generate_ensemble(int classes, const int* n, const int* count,
                         const float* const* noise,
                         const float* const* amplitude,
                         const float* const* sh_noise,
                         const double* const* sh_amp, int groups, int levels,
                         const float* const* screen_scale, float scale,
                         float* const* out, void* const* work,
                         int persistent_cols, void* stream);

/*
 * Spectral frozen-flow frames (src/pyturb/atmosphere.py:882-938 on-axis,
 * :940-998 for directions; src/pyturb/flow.py:96-129 for one layer).
 *
 *   spectra   : (layers, n, n) c64 fixed layer spectra (atmosphere.py:545)
 *   disp      : (frames, layers, 2) f64 displacement (s_x, s_y) in metres of
 *               each layer for each output frame (v*t + h*tan(theta))
 *   sh_coeffs : (levels, layers, 3, 3) c128 subharmonic coefficients, NULL
 *               when levels == 0
 *   out       : (frames, n_out, n_out) f32,
 *               out = scale * Re(ifft2(sum_l S_l P_x P_y) * n^2 + low)
 *               cropped at crop_offset
 *   work      : pt_frames_work_elems(n, frames, layers, levels) c64 (or the
 *               layer-independent bound pt_work_elems(1, n, frames, levels))
 * Shift-theorem phases are reduced exactly (32-bit fixed-point cycles), so
 * large travel keeps full FP32 phase accuracy.  When every frame pair moves
 * by the same fixed-point step per layer (to within one unit of 2^-32
 * cycles; equally spaced times), the call folds that step into the stack and
 * builds each pair's second frame at the first frame's phase plus the step:
 * half the phasor work, results within 1e-6 of the general path.  The call
 * decides this on the device; the decision lives in `work`.
 */
int pt_spectral_frames(const void* spectra, int layers, int n,
                       double pixel_scale, const double* disp, int frames,
                       const double* sh_coeffs, int levels, int crop_offset,
                       int n_out, float scale, float* out, void* work,
                       void* stream);

/*
 * pt_spectral_frames with no workspace argument, for callers that cannot
 * allocate device memory (code under HIP-graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * workspace is a static device array of PT_NOALLOC_SCRATCH_BYTES per GPU,
 * held by the library until the process ends; calls whose
 * pt_frames_work_elems(n, frames, layers, levels) * 8 bytes exceed it return
 * PT_ERR_ARG (512^2 with 9 layers and 8 levels: up to 44 frames per call).
 * The call allocates nothing and runs the kernels of pt_spectral_frames, so
 * its frames are bit-identical to pt_spectral_frames'.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (other streams
 * or threads) race and must be serialised by the caller.
 */
#define PT_NOALLOC_SCRATCH_BYTES (64LL << 20)
int pt_spectral_frames_noalloc(const void* spectra, int layers, int n,
                               double pixel_scale, const double* disp,
                               int frames, const double* sh_coeffs, int levels,
                               int crop_offset, int n_out, float scale,
                               float* out, void* stream);

/*
 * pt_spectral_frames split into its two passes (power-of-two n, <= 16
 * layers, <= 12 subharmonic levels), for callers that schedule or time them
 * separately.  Frames are transformed in pairs: Re(ifft2(X)) equals the
 * transform of the Hermitian part of X, so frames 2p and 2p+1 share one
 * complex transform of Herm(X_2p) + i Herm(X_2p+1).
 *   pass A: fold the layer stack into its Hermitian part (plus an exact
 *           Nyquist-line correction table), then work[p] = row-transformed
 *           packed sum of frame pair p
 *   pass B: subharmonic tables (one block per frame), then the column
 *           transform of each pair; real part -> frame 2p, imaginary part ->
 *           frame 2p+1, + subharmonics + crop + scale into out
 * pt_spectral_rows then pt_spectral_cols on one stream == pt_spectral_frames;
 * the workspace is pt_frames_work_elems(n, frames, layers, levels).
 */
int pt_spectral_rows(const void* spectra, int layers, int n,
                     double pixel_scale, const double* disp, int frames,
                     void* work, void* stream);
int pt_spectral_cols(void* work, int layers, int n, double pixel_scale,
                     const double* disp, int frames, const double* sh_coeffs,
                     int levels, int crop_offset, int n_out, float scale,
                     float* out, void* stream);

/*
 * The fold of pt_spectral_rows on its own, and pt_spectral_rows without it.
 * The fold written by pt_spectral_fold (the Hermitian-folded layer stack
 * and its Nyquist side table, in `work`) depends only on the spectra, the
 * layer count, n, the frame count (which places it in the workspace) and
 * the transform order (PT_TUNE_FRAMES).  A caller whose spectra do not
 * change between calls of the same frame count, such as a loop replaying a
 * captured HIP graph, folds once with pt_spectral_fold and then calls
 * pt_spectral_rows_folded + pt_spectral_cols per step.
 * Only pt_spectral_fold's output may be reused this way: the fold that
 * pt_spectral_rows or pt_spectral_frames leave in `work` carries that
 * call's uniform step when it had one.  The workspace records which fold
 * it holds, and pt_spectral_rows_folded on any other fold (a stepped one,
 * or a workspace never folded) produces NaN frames.
 * No counterpart in the reference (its per-frame layer sum is
 * atmosphere.py:910-917).  Power-of-two n only.
 */
int pt_spectral_fold(const void* spectra, int layers, int n, int frames,
                     void* work, void* stream);
int pt_spectral_rows_folded(int layers, int n, double pixel_scale,
                            const double* disp, int frames, void* work,
                            void* stream);

/*
 * pt_spectral_rows_folded + pt_spectral_cols in one call, on the fold that
 * pt_spectral_fold left in `work` (same rules).  For calls of at most 8
 * frames the per-call tables are built with pass A (n = 128..1024: pass A
 * builds its own pair tables and further workgroups of the same launch the
 * subharmonic tables and the Nyquist column's sums; two kernels per call),
 * which is what a loop of a few frames per call waits on.  One workgroup
 * per frame pair waits (bounded) for a producer workgroup of its own launch
 * through a flag in `work`: calls on one `work` must not overlap (as for
 * every call that writes it).  `disp` may point to pinned host memory
 * (hipHostMalloc): the kernels read it during the call, so it must not be
 * rewritten before the call has finished.  Same results as the two calls.  No counterpart in the reference (its
 * per-frame layer sum is atmosphere.py:910-917).  Power-of-two n only.
 */
int pt_spectral_frames_folded(int layers, int n, double pixel_scale,
                              const double* disp, int frames,
                              const double* sh_coeffs, int levels,
                              int crop_offset, int n_out, float scale,
                              float* out, void* work, void* stream);

/*
 * Ring-buffer and matrix rows below have a pitch `ld` >= width, a multiple
 * of 4 floats, with zero padding past `width`; arrays are 16-byte aligned.
 *
 * Innovation terms of the Assemat-Wilson recurrence for a batch of new rows:
 *   out[k][w] = sum_j B_unit[w][j] * beta_k[j]      (out = beta * B^T)
 *   beta_rows: HOST array of `rows` device pointers, each to `width`
 *   contiguous f32 white-noise values (4-byte aligned, no padding needed),
 *   so every layer's innovations are read where the caller drew them;
 *   B_unit (width, ld) f32, out (rows, ld) f32.
 *   The per-layer r0^(-5/6) factor is applied by pt_extrude_rows.
 * Replaces the B @ beta half of src/pyturb/extrude.py:427-429 and
 * src/pyturb/infinite.py:278-280, batched over every row of a step (as
 * extrude.py:686-689 does for boiling).
 */
int pt_innovation_gemm(const float* const* beta_rows, const float* B_unit,
                       int rows, int width, int ld, float* out, void* stream);

/*
 * One round of the row recurrence for `nact` ring buffers:
 *   buf[layer[e]][trow[e]][:] = A @ concat(buf[layer[e]][zrows[e][0..m-1]])
 *                              + bscale[e] * bb[bbrow[e]][:]
 *   A (width, m*ld) f32, stencil row s in columns [s*ld, s*ld + width)
 *   (shared by the layers of one outer scale); buf (layers, cap, ld) f32
 *   ring buffers (physical row indices); bb from pt_innovation_gemm.
 *   HOST arrays: layer, trow, bbrow (nact) int, zrows (nact, m) int,
 *   bscale (nact) f32.  m <= 4.
 * Replaces src/pyturb/extrude.py:424-430 (and infinite.py:274-281) for all
 * layers that extrude a row in this round.
 */
int pt_extrude_rows(const float* A, int width, int ld, int stencil_rows,
                    const float* bb, float* buf, long long cap, int nact,
                    const int* layer, const int* zrows, const int* trow,
                    const int* bbrow, const float* bscale, void* stream);

/*
 * pt_extrude_rows with a stencil of m full rows plus `far_points` single
 * pixels further back (far_points = 8 R, R <= 4 far rows of 8 points).
 * No counterpart in the reference, whose stencil is full rows only
 * (src/pyturb/extrude.py:272-275); with far points the new row is
 *   A @ concat(stencil rows, far points) + bscale * bb[bbrow]
 * where far point q is column (2 (q % 8) + 1) * width / 16 (integer
 * division) of far row q / 8.  A is (width, m*ld + round_up(far_points, 16))
 * with the far points in columns [m*ld, m*ld + far_points) (zero padding
 * after them).  HOST array far_rows (nact, R) int: the physical ring row of
 * each far row of each entry.  Other arguments as pt_extrude_rows;
 * far_points = 0 is pt_extrude_rows.
 */
int pt_extrude_rows_far(const float* A, int width, int ld, int stencil_rows,
                        int far_points, const float* bb, float* buf,
                        long long cap, int nact, const int* layer,
                        const int* zrows, const int* far_rows, const int* trow,
                        const int* bbrow, const float* bscale, void* stream);

/*
 * One extrusion step in one launch: the innovation terms of every new row
 * and every round of the recurrence (pt_innovation_gemm followed by the
 * pt_extrude_rows_far rounds, bit-identical results), for layers that share
 * one A and one B.  Layer l extrudes rows[l] rows at physical ring rows
 * (first_row[l] + i) mod cap, i < rows[l], from its innovations
 * beta[l] + i * beta_pitch[l] (HOST arrays of device pointers, 16-byte
 * aligned, and pitches in floats, multiples of 4; a row is read up to its
 * width rounded up to 4, the floats past the width are ignored), each row's
 * stencil being the m ring rows before it (and the
 * far rows far_rows_back[j] + 1 behind it).  One launch on `stream` whose
 * workgroups (one per 16 rows of A) are all resident at once: each
 * workgroup keeps its 16 rows of A in registers for the whole step, and
 * the rounds are separated by a grid barrier on `sync`, PT_EXTRUDE_SYNC_BYTES
 * of device memory (8-byte aligned) that is zero before the first call and
 * is passed unchanged to every later call on the same stream (it counts the
 * barriers passed).  A barrier wait is bounded by spin_limit polls; on
 * expiry, or when the abort word (the second 8-byte word of sync) is
 * nonzero, every workgroup leaves the kernel, *status (host-visible, e.g.
 * pinned) is set to 1 and the rows of the step are incomplete: zero sync
 * again before the next call.  PT_ERR_UNSUPPORTED (nothing launched) for
 * more than 16 layers, more than 256 new rows or 32 of one layer, stencil_rows > 2, a row
 * pitch ld above 2304 or not a multiple of 16, a far_rows_back entry below
 * 1 (a round reads its far rows before the previous round's barrier), or a
 * GPU that cannot hold the grid; the caller then uses the
 * round-by-round entry points.  Replaces src/pyturb/extrude.py:424-430 for
 * every layer of a set_time step.  Other arguments as pt_extrude_rows_far.
 */
#define PT_EXTRUDE_SYNC_BYTES 2048
int pt_extrude_step(const float* A, const float* B, int width, int ld,
                    int stencil_rows, int far_points, const int* far_rows_back,
                    float* buf, long long cap, int layers,
                    const int* first_row, const int* rows,
                    const float* const* beta, const long long* beta_pitch,
                    const float* bscale, void* sync, int* status,
                    unsigned spin_limit, void* stream);

/*
 * Fused multi-layer, multi-direction readout of the ring buffers
 * (src/pyturb/extrude.py:754-915 and the CUDA readout kernels
 * extrude.py:91-198; one launch instead of one per direction):
 *   out[d][i][j] = scale * sum_l interp(buf[l], row, col) (+ out if accumulate)
 *   g = (i - (n-1)/2) * mag[d][l]  (likewise for j),
 *   row = g_i cos_l + g_j sin_l + sa[d][l],
 *   col = -g_i sin_l + g_j cos_l + width/2 + sp[d][l],
 *   rows clamped to [0, fill_l - 1] (local ring rows, physical row
 *   (start_l + r) mod cap), columns to [0, width - 1].
 *   HOST arrays: start (layers) int, fill (layers) int64, wind_cs
 *   (layers, 2) f64 = (cos, sin), dir_geom (dirs, layers, 3) f64 =
 *   (sa, sp, mag).  interp: 0 linear, 1 Catmull-Rom cubic, 2 Lanczos-3.
 *   buf (layers, cap, ld), out (dirs, n, n) f32 device.
 */
int pt_readout(const float* buf, int layers, long long cap, int width, int ld,
               const int* start, const long long* fill, const double* wind_cs,
               const double* dir_geom, int dirs, int n, int interp, float scale,
               int accumulate, float* out, void* stream);

/*
 * Adjoint of pt_readout with respect to each (direction, layer)'s readout
 * geometry, for a cotangent grad = dLoss/d(out) (dirs, n, n) f32 device:
 *   out[(d * layers + l) * 3 + {0, 1, 2}] = dLoss/d sa, dLoss/d sp, dLoss/d mag
 * (f64 device; sa, sp, mag as in dir_geom).  Same taps and clamping as the
 * forward; the interpolation weights are differentiated analytically
 * (linear, Catmull-Rom, normalised Lanczos-3).  The reference computes no
 * derivatives; this serves fitting wind travel and viewing directions
 * through the extruder readout.  work: pt_readout_adjoint_work_elems(n)
 * float32 elements.  Other arguments as pt_readout.
 */
long long pt_readout_adjoint_work_elems(int n);
int pt_readout_adjoint(const float* buf, int layers, long long cap, int width,
                       int ld, const int* start, const long long* fill,
                       const double* wind_cs, const double* dir_geom, int dirs,
                       int n, int interp, float scale, const float* grad,
                       double* out, float* work, void* stream);

/*
 * Differentiable frame model (profile and wind fitting).  The reference has
 * no derivatives; these entry points serve a fit of per-layer Cn2 weights,
 * r0, L0 and winds to OPD frames.  The layer spectra are written as
 *   S_l(k) = gain_l * amplitude[group_l](k) * (w_l,re(k) + i w_l,im(k)),
 * gain_l = r0_l^(-5/6) for unit-r0 amplitudes (src/pyturb/atmosphere.py:
 * 401-402 and 532-579, src/pyturb/fourier.py:143,174-191), so the forward
 * model is pt_layer_spectra followed by pt_spectral_frames.
 *
 * pt_layer_spectra: noise (layers, 2, n, n) f32 white-noise planes,
 *   amplitude (groups, n, n) f32; HOST arrays group (layers) int and gain
 *   (layers) f64; spectra (layers, n, n) c64 out.  layers <= 16.
 */
int pt_layer_spectra(const float* noise, const float* amplitude,
                     const int* group, const double* gain, int layers, int n,
                     void* spectra, void* stream);

/* Workspace (c64 elements) of pt_spectral_frames_adjoint; -1 on bad input. */
long long pt_adjoint_work_elems(int n, int frames, int levels);

/*
 * Adjoint of pt_layer_spectra + pt_spectral_frames for a real cotangent
 * grad = dLoss/d(out) (frames, n_out, n_out) f32 (the output crop of the
 * forward call).  Layer spectra are recomputed from the noise, never read.
 *   dlnamp  : (groups, n, n) f32, d ln(amplitude) / d L0
 *   sh_unit : (levels, layers, 3, 3) c128 unit-gain subharmonic coefficients
 *             (the forward's sh_coeffs are gain_l * sh_unit)
 *   sh_rho  : (levels, layers, 3, 3) f64, d ln|sh_unit| / d L0
 *   disp    : (frames, layers, 2) f64 device, as in pt_spectral_frames
 *   out     : (frames, layers, 4) f64 device:
 *             dLoss/dgain_l, dLoss/ds_x, dLoss/ds_y, dLoss/dL0_l per frame
 *             (sum over frames for the per-layer parameters)
 *   group, gain: HOST arrays as for pt_layer_spectra.
 * Power-of-two n from 32 to 2048, layers <= 16, levels <= 12.
 */
int pt_spectral_frames_adjoint(const float* noise, const float* amplitude,
                               const float* dlnamp, const int* group,
                               const double* gain, int layers, int n,
                               double pixel_scale, const double* disp,
                               int frames, const float* grad, int crop_offset,
                               int n_out, const double* sh_unit,
                               const double* sh_rho, int levels, double scale,
                               double* out, void* work, void* stream);

/* Workspace (c64 elements) of pt_boiled_frames; -1 on bad input. */
long long pt_boil_work_elems(int n, int frames, int layers, int levels);

/*
 * Spectral frames with boiling: `frames` frames at the displacements disp,
 * each followed by one boiling step of dt seconds, then steps - frames more
 * steps (frames = 0: boil only).  Frame k is built from the state after k
 * steps, as Atmosphere.frames does (src/pyturb/atmosphere.py:1128-1154),
 * and the state is left after `steps` steps.  One boiling step is the
 * per-mode AR(1) update of src/pyturb/atmosphere.py:1092-1126 with the time
 * constants of :618-658:
 *   S <- a S + sqrt(1 - a^2) (n0 + i n1) A,   a = exp(-dt / tau(f)),
 *   tau(f) = tau_boil * max(|f| / f_ref, 1)^(-2/3),
 * on the layer spectra and on the subharmonic coefficients.
 *   spectra   : (layers, n, n) c64 state, in place
 *   amplitude : (layers, n, n) f32 PSD amplitude of each layer
 *   inv_tau   : (layers) f64 device, 1 / tau_boil (0 for a frozen layer)
 *   f_ref     : (layers) f64 device, 1 / L0 (or 1 / (n dx) for L0 = inf)
 *   disp      : (frames, layers, 2) f64 device, displacements [m]
 *   sh_coeffs : (levels, layers, 3, 3) c128 state, in place
 *   sh_amp    : (levels, layers, 3, 3) f64 subharmonic amplitudes
 *   noise     : (steps, 2, layers, n, n) f32 standard normals in the
 *               reference's draw order, or NULL for the device stream
 *               (Philox4x32-10 keyed by `seed`, counted from step `step0`)
 *   sh_noise  : (steps, 2, levels, layers, 3, 3) f64, or NULL likewise
 *   out       : (frames, n_out, n_out) f32, crop and scale as in
 *               pt_spectral_frames.
 * Replaces Atmosphere._boil_step + _integrate per frame
 * (src/pyturb/atmosphere.py:882-938, 1092-1154).  layers <= 16.
 */
int pt_boiled_frames(void* spectra, const float* amplitude,
                     const double* inv_tau, const double* f_ref, int layers,
                     int n, double pixel_scale, const double* disp, int frames,
                     int steps, double dt, double* sh_coeffs,
                     const double* sh_amp, int levels, const float* noise,
                     const double* sh_noise, unsigned long long seed,
                     unsigned long long step0, int crop_offset, int n_out,
                     float scale, float* out, void* work, void* stream);

/*
 * Spectral LGS cone readout: every layer's screen sampled at the pupil grid
 * scaled about the screen centre by the layer's cone magnification, with
 * separable taps, and the layers summed:
 *   out[f][i][j] = scale * sum_l sum_a sum_b w[l][a][i] w[l][b][j]
 *                  * screens[l][f][idx[l][a][i]][idx[l][b][j]]
 *   screens : (layers, frames, ns, ns) f32 per-layer screens (each layer
 *             transformed on its own, subharmonics included)
 *   idx, w  : (layers, taps, n) int32 / f32 device tap tables, idx clipped
 *             to [0, ns) and non-decreasing along both the pixel and the tap
 *             axis (as src/pyturb/atmosphere.py:581-616 builds them)
 *   taps    : 2 linear, 4 cubic, 6 lanczos;  out : (frames, n, n) f32.
 * Replaces Atmosphere._lgs_zoom (src/pyturb/atmosphere.py:1047-1090).
 */
int pt_lgs_zoom(const float* screens, int layers, int frames, int ns,
                const int* idx, const float* w, int taps, int n, float scale,
                float* out, void* stream);

/*
 * `count` standard normals times `scale` into out (f32): element i is
 * component i % 2 of Box-Muller pair (i % 4) / 2 of the Philox4x32-10 call
 * with counter (i / 4, stream_id) and key `seed`.  The extruder's boiling
 * noise (the reference draws it from one generator, extrude.py:648-650).
 */
int pt_normals(float* out, long long count, unsigned long long seed,
               unsigned long long stream_id, float scale, void* stream);

/*
 * dst row (dst_start + r) mod dst_cap = a * itself + b * src row r, for
 * r < rows and `width` floats (a = 0: dst is not read).  Pitches in floats.
 * The boiling blend of src/pyturb/extrude.py:740-752 on a ring buffer, and
 * the seeding of fresh windows.
 */
int pt_rows_blend(float* dst, long long dst_pitch, int dst_start, int dst_cap,
                  const float* src, long long src_pitch, int rows, int width,
                  float a, float b, void* stream);

/*
 * Autotuner state: the launch shape and FFT plan of the generate row pass
 * (PT_TUNE_GEN_ROWS) and of the column pass shared by pt_generate* and the
 * frame engine's pass B (PT_TUNE_GEN_COLS), and the frame engine's transform
 * order (PT_TUNE_FRAMES), selected per power-of-two grid size n in 32..2048.
 * No counterpart in the reference, whose transforms are library ifft2 calls
 * (fourier.py:259, atmosphere.py:918).  Variant 0 is the built-in default.
 * Variants that differ only in launch shape give bit-identical results;
 * the radix-4 plans and the two transform orders round differently (within
 * 1e-6 relative RMS).  The selection is process-wide host state read when a
 * call launches (not synchronised with calls from other threads).
 *
 * pt_tune_variants: number of variants of a knob (negative error code on a
 * bad knob).  pt_tune_info: effective threads per workgroup, rows or
 * columns per workgroup (tile), XCD-contiguous order (0/1) and the plan's
 * largest radix (8 or 4) of a variant at size n; any output pointer may be
 * NULL.  pt_tune_set / pt_tune_get:
 * select / read the variant for size n (pt_tune_get returns it, or a
 * negative error code).
 */
int pt_tune_variants(int knob);
int pt_tune_info(int knob, int n, int variant, int* threads, int* tile,
                 int* xcd_order, int* radix);
int pt_tune_set(int knob, int n, int variant);
int pt_tune_get(int knob, int n);

#ifdef __cplusplus
}
#endif

#endif /* PYTURB_HIP_H */
