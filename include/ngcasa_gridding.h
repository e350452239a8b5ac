/*
 * ngcasa_gridding.h — C ABI of the MI355X gridding library (libngcasa_hip.so).
 *
 * This is the lower drop-in boundary of ngCASA's imaging hot path.  Each entry
 * point replaces one numba kernel (or one numpy/dask step) of the reference
 * and keeps its argument meaning:
 *
 *   ngc_standard_grid          _standard_grid_jit
 *                              ngcasa/imaging/_imaging_utils/_standard_grid.py:243-371
 *                              (called by _standard_grid_numpy_wrap :123-177 and
 *                              _standard_grid_psf_numpy_wrap :180-236)
 *   ngc_imaging_weight_degrid  _standard_imaging_weight_degrid_jit
 *                              _standard_grid.py:467-518 (wrapper :443-464)
 *   ngc_briggs_factors         calculate_briggs_parms
 *                              ngcasa/imaging/make_imaging_weight.py:198-216
 *   ngc_aperture_grid          _aperture_grid_jit  _aperture_grid.py:377-513
 *                              (wrappers :294-371) and, with NGC_GRID_WEIGHT,
 *                              _aperture_weight_grid_jit  _aperture_grid.py:181-291
 *                              (wrapper :146-178)
 *   ngc_grid_to_image          ifftshift/ifft2/fftshift + _remove_padding +
 *                              correct_image: make_psf.py:117-131,
 *                              make_image.py:116-130, make_mosaic_pb.py:119-129,
 *                              make_image_with_gcf.py:143-147
 *   ngc_normalize_mosaic       _normalize 'forward'  _imaging_utils/_normalize.py:17-89
 *   ngc_gcf_patterns           make_baseline_patterns + FFT of the a_term branch of
 *                              make_gridding_convolution_function.py:181-191,245-246,394-412
 *   ngc_gcf_resize             resize_and_calc_support  :361-391 (crop/normalise)
 *   ngc_phase_gradient         make_phase_gradient  :331-359 (the table)
 *   ngc_direction_rotate       apply_rotation_matrix + apply_phasor
 *                              ngcasa/imaging/direction_rotate.py:213-268
 *
 * Conventions
 *   - Large arrays (uvw, visibilities, weights, flags, grids, images) are
 *     DEVICE pointers on the context's device.  Small parameter arrays
 *     (freq_chan, chan_map, pol_map, cgk_1d, cf maps, field ids, supports,
 *     briggs factors) are HOST pointers; the library copies them.
 *   - Array layouts are the reference's, C order: visibilities and weights
 *     [n_time, n_baseline, n_chan, n_pol], uvw [n_time, n_baseline, 3] in
 *     metres, grids [n_imag_chan, n_imag_pol, n_u, n_v] (the gridder's own
 *     orientation, before the reference's moveaxis), complex values
 *     interleaved (re, im).
 *   - Flags: the reference marks flagged samples by NaN (cngi/vis/apply_flags.py:49-53).
 *     NaN is honoured exactly as the reference does; an optional u8 mask
 *     [n_time, n_baseline, n_chan, n_pol] (nonzero = flagged) is treated the
 *     same as a NaN weight.  Pass NULL for none.
 *   - `accumulate` = 0 overwrites grid and sum_weight; 1 adds to them (the
 *     reference's in-place convention with caller-zeroed outputs).
 *   - Calls are ordered on `stream` (a hipStream_t; NULL = default stream) and
 *     return before the work completes, except where a result must come back
 *     to the host (ngc_briggs_factors) or a host-visible error must be checked.
 *   - Every function returns 0 on success and a negative NGC_ERR_* code on
 *     failure; ngc_last_error() gives the message (thread-local).  No C++
 *     exception crosses this boundary.
 *   - A context owns the device workspace, the cached kernel tables and the
 *     rocFFT plans.  Calls on one context are serialised internally; separate
 *     contexts may be used from separate threads.
 */
#ifndef NGCASA_GRIDDING_H
#define NGCASA_GRIDDING_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define NGC_OK 0
#define NGC_ERR_ARG (-1)         /* invalid argument (shape, dtype, range) */
#define NGC_ERR_HIP (-2)         /* a HIP runtime call failed */
#define NGC_ERR_UNSUPPORTED (-3) /* valid input that this build does not handle */
#define NGC_ERR_FIELD (-4)       /* a row's field id is absent from field_id */
#define NGC_ERR_FFT (-5)         /* rocFFT failed */

/* gridding modes */
#define NGC_GRID_IMAGE 0   /* complex grid of vis*weight (make_image, make_image_with_gcf) */
#define NGC_GRID_PSF 1     /* grid of weight (make_psf, make_psf_with_gcf) */
#define NGC_GRID_DENSITY 2 /* weight density: nearest cell and its conjugate (make_imaging_weight) */
#define NGC_GRID_WEIGHT 3  /* A-term weight image at the grid centre (make_mosaic_pb) */

/* precision forms */
#define NGC_PREC_FP64 0       /* plain FP64 drop-in: reference loop, FP64 atomics */
#define NGC_PREC_FP64_TILED 1 /* uv-tiled binning, FP64 LDS accumulation */
#define NGC_PREC_FP32_KAHAN 2 /* uv-tiled binning, FP32 taps and register runs
                                 flushed into an LDS subgrid whose cells are
                                 FP64 (the 8 bytes of a (sum, compensation)
                                 pair of floats, with at least its accuracy);
                                 FP64 for coordinates and sum_weight */
#define NGC_PREC_FP32 3       /* as FP32_KAHAN with plain FP32 LDS cells */

/* element types */
#define NGC_C64 0
#define NGC_C128 1
#define NGC_F32 2
#define NGC_F64 3
#define NGC_I32 4
#define NGC_I64 5
#define NGC_BOOL 6            /* one byte per element, 0 or 1 */

/* weighting for ngc_briggs_factors (make_imaging_weight.py:198-213) */
#define NGC_WEIGHT_UNIFORM 1
#define NGC_WEIGHT_BRIGGS 2

typedef struct ngc_context ngc_context;

/* Create a context on HIP device `device`. */
int ngc_context_create(int device, ngc_context **ctx);
int ngc_context_destroy(ngc_context *ctx);
/* Message of the last failure on this thread ("" if none). */
const char *ngc_last_error(void);
/* Library version string. */
const char *ngc_version(void);

/*
 * Standard (prolate-spheroidal) gridder: replaces _standard_grid_jit.
 *   mode       NGC_GRID_IMAGE, NGC_GRID_PSF or NGC_GRID_DENSITY
 *   precision  NGC_PREC_*; grid element type follows it:
 *              FP64 forms: complex128 (IMAGE) / float64 (PSF, DENSITY);
 *              FP32 forms: complex64 (IMAGE) / float32 (PSF, DENSITY)
 *   vis        device [T,B,C,P] of vis_dtype (NGC_C64|NGC_C128); NULL unless IMAGE
 *   weight     device [T,B,C,P] of weight_dtype (NGC_F32|NGC_F64)
 *   chan_map   host int32 [n_chan] -> image channel; pol_map host int32 [n_pol]
 *   cgk_1d     host float64 [cgk_len] oversampled half kernel
 *              (_gridding_convolutional_kernels.py:151-157); density mode uses {1.0}
 *   delta_l/m  cell size in radians with the reference's sign (cell[0] negated)
 *   grid       device out [n_imag_chan, n_imag_pol, n_u, n_v]
 *   sum_weight device float64 out [n_imag_chan, n_imag_pol]
 */
int ngc_standard_grid(ngc_context *ctx, void *stream, int mode, int precision,
                      const double *uvw, const void *vis, int vis_dtype,
                      const void *weight, int weight_dtype, const uint8_t *flag,
                      const double *freq_chan, const int32_t *chan_map,
                      const int32_t *pol_map, const double *cgk_1d,
                      int64_t cgk_len, int64_t n_time, int64_t n_baseline,
                      int64_t n_chan, int64_t n_pol, int64_t n_imag_chan,
                      int64_t n_imag_pol, int64_t n_u, int64_t n_v,
                      double delta_l, double delta_m, int64_t support,
                      int64_t oversampling, int accumulate, void *grid,
                      double *sum_weight);

/*
 * Allocation-free form of ngc_standard_grid (no counterpart in the reference:
 * it is for callers that cannot allocate device memory during a call, such
 * as code under HIP graph capture, tight serving loops and harnesses that
 * forbid hipMalloc).  Same arguments, results and error behaviour as
 * ngc_standard_grid, in every mode and precision.  It calls no hipMalloc,
 * hipFree or allocating library routine: every device buffer beyond the
 * inputs and outputs (parameter tables, tap-sum table, bin counts, offsets,
 * order and queues, bin records) lives in one static device array of
 * NGC_NOALLOC_SCRATCH_BYTES per GPU, held from the library's load until the
 * process ends and addressed once per device at ngc_context_create.  A call
 * whose tables and bin records do not fit is refused with NGC_ERR_ARG before
 * grid or sum_weight are touched (64 MiB holds about 1.6M bin records of
 * 40 B, i.e. 4 pols of complex64, or 2.8M of 24 B; a record is one (row,
 * channel, uv tile touched)).  All calls on one GPU share the array: calls
 * that overlap on one GPU (different streams or contexts with no ordering
 * between them) race.  Each call clears what must start at zero itself.
 */
#define NGC_NOALLOC_SCRATCH_BYTES (64ull << 20)
int ngc_standard_grid_noalloc(ngc_context *ctx, void *stream, int mode, int precision,
                              const double *uvw, const void *vis, int vis_dtype,
                              const void *weight, int weight_dtype, const uint8_t *flag,
                              const double *freq_chan, const int32_t *chan_map,
                              const int32_t *pol_map, const double *cgk_1d,
                              int64_t cgk_len, int64_t n_time, int64_t n_baseline,
                              int64_t n_chan, int64_t n_pol, int64_t n_imag_chan,
                              int64_t n_imag_pol, int64_t n_u, int64_t n_v,
                              double delta_l, double delta_m, int64_t support,
                              int64_t oversampling, int accumulate, void *grid,
                              double *sum_weight);

/*
 * Briggs/uniform factors from a density grid: replaces calculate_briggs_parms.
 *   density_grid  device float64 [n_imag_chan, n_imag_pol, n_u, n_v]
 *   sum_weight    device float64 [n_imag_chan, n_imag_pol]
 *   factors       HOST float64 out [2, n_imag_chan, n_imag_pol]
 * Synchronises `stream`.
 */
int ngc_briggs_factors(ngc_context *ctx, void *stream, const double *density_grid,
                       const double *sum_weight, int64_t n_imag_chan,
                       int64_t n_imag_pol, int64_t n_u, int64_t n_v,
                       int weighting, double robust, double *factors);

/*
 * Imaging-weight degridder: replaces _standard_imaging_weight_degrid_jit.
 *   density_grid    device float64 [n_imag_chan, n_imag_pol, n_u, n_v]
 *   briggs_factors  host float64 [2, n_imag_chan, n_imag_pol]
 *   natural_weight  device [T,B,C,P] of weight_dtype
 *   imaging_weight  device out [T,B,C,P] of out_dtype (NGC_F32|NGC_F64)
 */
int ngc_imaging_weight_degrid(ngc_context *ctx, void *stream, const double *uvw,
                              const void *natural_weight, int weight_dtype,
                              const uint8_t *flag, const double *freq_chan,
                              const int32_t *chan_map, const int32_t *pol_map,
                              const double *density_grid,
                              const double *briggs_factors, int64_t n_time,
                              int64_t n_baseline, int64_t n_chan, int64_t n_pol,
                              int64_t n_imag_chan, int64_t n_imag_pol,
                              int64_t n_u, int64_t n_v, double delta_l,
                              double delta_m, void *imaging_weight,
                              int out_dtype);

/*
 * Aperture (mosaic A-term) gridder: replaces _aperture_grid_jit (modes IMAGE
 * and PSF) and _aperture_weight_grid_jit (mode NGC_GRID_WEIGHT).
 *   field            device int32 [T,B] field id per row (<= -1 skips the row)
 *   field_id         host int32 [n_field]: the gcf's field ids; position = index
 *                    into phase_gradient
 *   cf_*_map         host int32 maps [n_baseline], [n_chan], [n_pol]
 *   conv_kernel      device float64 [n_cfb, n_cfc, n_cfp, k_u, k_v]
 *                    (CONV_KERNEL, or WEIGHT_CONV_KERNEL for NGC_GRID_WEIGHT)
 *   support          host int32 [n_cfb, n_cfc, n_cfp, 2]
 *   phase_gradient   device complex128 [n_field, k_u, k_v]
 *   grid             device out complex128 [n_imag_chan, n_imag_pol, n_u, n_v]
 *                    (complex64 for NGC_PREC_FP32_KAHAN)
 * Returns NGC_ERR_FIELD if a gridded row's field id is not in field_id (the
 * reference raises IndexError); this check synchronises `stream`.
 */
int ngc_aperture_grid(ngc_context *ctx, void *stream, int mode, int precision,
                      const double *uvw, const void *vis, int vis_dtype,
                      const void *weight, int weight_dtype, const uint8_t *flag,
                      const int32_t *field, const int32_t *field_id,
                      int64_t n_field, const double *freq_chan,
                      const int32_t *chan_map, const int32_t *pol_map,
                      const int32_t *cf_baseline_map, const int32_t *cf_chan_map,
                      const int32_t *cf_pol_map, const double *conv_kernel,
                      int64_t n_cfb, int64_t n_cfc, int64_t n_cfp, int64_t k_u,
                      int64_t k_v, const int32_t *support,
                      const double *phase_gradient, int64_t n_time,
                      int64_t n_baseline, int64_t n_chan, int64_t n_pol,
                      int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
                      int64_t n_v, double delta_l, double delta_m,
                      int64_t os_u, int64_t os_v, int accumulate, void *grid,
                      double *sum_weight);

/*
 * Grid -> image: ifftshift, unnormalised inverse 2-D FFT, fftshift, crop of
 * the centre n_l x n_m, real part x (n_u*n_v), then division by sum_weight
 * (0 -> 1) and by corr_l[l]*corr_m[m] when those are given.  The transform is
 * rocFFT, or for single-precision grids whose axis lengths are products of 2,
 * 3, 5, 7, 11 and 13 up to 8192 the library's own LDS row FFTs with a tiled
 * transpose (see ngc_set_fft_mode).
 *   grid        device [n_planes, n_u, n_v] of grid_dtype (C64, C128, F32, F64);
 *               single-precision grids are transformed in FP32, double in FP64
 *   sum_weight  device float64 [n_planes] or NULL
 *   corr_l/m    host float64 [n_l] / [n_m] or NULL (grdsf correction factors)
 *   image       device float64 out [n_planes, n_l, n_m]
 */
int ngc_grid_to_image(ngc_context *ctx, void *stream, const void *grid,
                      int grid_dtype, int64_t n_planes, int64_t n_u,
                      int64_t n_v, int64_t n_l, int64_t n_m,
                      const double *sum_weight, const double *corr_l,
                      const double *corr_m, double *image);

/*
 * Mosaic normalisation, 'forward' direction of _normalize
 * (_imaging_utils/_normalize.py:39-89):
 *   image = (image / sum_weight') / (sinc_l*sinc_m * ps_corr * norm_pb),
 *   zero where pb < pb_limit (pb_limit > 0), optional float32 round trip.
 *   image      device float64 in/out [n_planes, n_l, n_m]
 *   sum_weight device float64 [n_planes] (0 -> 1)
 *   ps_corr    device float64 [n_l, n_m] (PS_CORR_IMAGE) or NULL (ones)
 *   norm_pb    device float64 [n_planes, n_l, n_m]: PB for flat_noise,
 *              WEIGHT_PB for flat_sky, NULL for 'none'
 *   pb         device float64 [n_planes, n_l, n_m] for the pb_limit mask, or NULL
 *   os_u/os_v  GCF oversampling for the sinc correction
 */
int ngc_normalize_mosaic(ngc_context *ctx, void *stream, double *image,
                         const double *sum_weight, const double *ps_corr,
                         const double *norm_pb, const double *pb,
                         int64_t n_planes, int64_t n_l, int64_t n_m,
                         int64_t os_u, int64_t os_v, double pb_limit,
                         int single_precision);

/*
 * Phase-centre rotation: replaces apply_rotation_matrix and apply_phasor of
 * direction_rotate (ngcasa/imaging/direction_rotate.py:213-268); the 3 x 3
 * matrices and phase-rotation vectors of calc_rotation_mats (:148-196) come
 * from the host.
 *   field            device int32 [T,B]; one valid id (>= 0) per integration,
 *                    rows with INT32_MIN ignored, else NGC_ERR_FIELD
 *   field_ids        host int32 [n_ids]: the ids the tables are listed for
 *   uvw_rotmat       host float64 [n_ids][3][3]: uvw_rot = uvw (row) @ M
 *   phase_rotation   host float64 [n_ids][3]
 *   uvw_out          device float64 [T,B,3]
 *   vis_out          device [T,B,C,P]: complex64 when single_precision, else
 *                    complex128; vis * exp(2 pi i (uvw_rot . p) f / c), with
 *                    the w term left out under common-tangent reprojection
 * Synchronises `stream` once (the field check).
 */
int ngc_direction_rotate(ngc_context *ctx, void *stream, const double *uvw, const void *vis,
                         int vis_dtype, const int32_t *field, const int32_t *field_ids,
                         int64_t n_ids, const double *uvw_rotmat, const double *phase_rotation,
                         const double *freq_chan, int64_t n_time, int64_t n_baseline,
                         int64_t n_chan, int64_t n_pol, int common_tangent, int single_precision,
                         double *uvw_out, void *vis_out);

/*
 * A-term convolution-function patterns: replaces make_baseline_patterns and
 * the real part of fftshift(fft2(ifftshift(.))) in
 * make_gridding_convolution_function (make_gridding_convolution_function.py:
 * 181-191, 245-246, 394-412; patterns _make_pb_symmetric.py:135-235).
 *   pb_freq      host [n_freq] pattern frequencies (Hz)
 *   dish, blockage host [n_dish] diameters (m)
 *   pairs        host int32 [n_pairs][2] dish-type pairs
 *   casa         1: CASA Airy (casa_airy), 0: obscured Airy (airy)
 *   ipower       1: voltage-pattern products (CONV_KERNEL), 2: power (WEIGHT_CONV_KERNEL)
 *   cell_l/m     pattern cell = image cell (radians, cell[0] negated) x GCF oversampling
 *   out          device float64 [n_pairs, n_freq, n_u, n_v]
 *   abs_max/min  host [n_pairs*n_freq] max and min of |out| per plane, or NULL
 *   centre_row   host [n_pairs*n_freq][n_u]: out[..][:, n_v//2], or NULL
 *   centre_col   host [n_pairs*n_freq][n_v]: out[..][n_u//2, :], or NULL
 * Synchronises `stream` (the support cut is decided on the host from these).
 */
int ngc_gcf_patterns(ngc_context *ctx, void *stream, const double *pb_freq, int64_t n_freq,
                     const double *dish, const double *blockage, int64_t n_dish,
                     const int32_t *pairs, int64_t n_pairs, int casa, int ipower,
                     double cell_l, double cell_m, int64_t n_u, int64_t n_v, double *out,
                     double *abs_max, double *abs_min, double *centre_row,
                     double *centre_col);

/*
 * Crop and normalisation of resize_and_calc_support
 * (make_gridding_convolution_function.py:361-391): the centre k_u x k_v window
 * of each pattern plane divided by Re(sum over the centred (support+1)*os
 * window) / (os_u*os_v).
 *   full     device float64 [n_items, n_u, n_v]
 *   support  host int32 [n_items][2]
 *   out      device float64 [n_items, k_u, k_v]
 */
int ngc_gcf_resize(ngc_context *ctx, void *stream, const double *full, int64_t n_items,
                   int64_t n_u, int64_t n_v, const int32_t *support, int64_t os_u,
                   int64_t os_v, int64_t k_u, int64_t k_v, double *out);

/*
 * Phase-gradient table of make_phase_gradient
 * (make_gridding_convolution_function.py:349-359):
 *   out[f][x][y] = exp(i (x*pix[f][0] + y*pix[f][1])), x in [-k_u//2, k_u - k_u//2)
 *   pix  host float64 [n_field][2] (radians per oversampled kernel pixel)
 *   out  device complex128 [n_field, k_u, k_v]
 */
int ngc_phase_gradient(ngc_context *ctx, void *stream, const double *pix, int64_t n_field,
                       int64_t k_u, int64_t k_v, double *out);

/*
 * Flags applied to one data variable: replaces the per-variable body of
 * apply_flags (cngi/vis/apply_flags.py:49-55), dv.where(flag == 0).astype(dv.dtype).
 * out[i] = data[i] where flag[i] == 0, else NaN (NGC_F32/F64, both parts for
 * NGC_C64/C128), the integer minimum (NGC_I32/I64) or 1 (NGC_BOOL).
 *   data, out  device arrays of n elements of `dtype` (out may equal data)
 *   flag       device uint8 [n]
 */
int ngc_apply_flags(ngc_context *ctx, void *stream, const void *data, int dtype,
                    const uint8_t *flag, int64_t n, void *out);

/*
 * Primary beam image: replaces the map_blocks call of make_pb (make_pb.py:100-104)
 * over _casa_airy_disk (casa != 0) / _airy_disk (_imaging_utils/_make_pb_symmetric.py:
 * 24-130).  Pixel (l, m) of channel c and dish d is the Airy power (ipower 2) or
 * voltage (ipower 1) pattern at radius argument
 *   r = sqrt(((l - center_l) cell_l)^2 + ((m - center_m) cell_m)^2) * (2 pi f_c / c) * D_d / 2,
 * and exactly 1 at the centre pixel; the value is repeated over the n_pol pols.
 *   freq_chan  host float64 [n_chan];  dish, blockage  host float64 [n_dish] (metres)
 *   cell_l, cell_m  radians
 *   out  device float64 [n_l, n_m, n_chan, n_pol, n_dish]
 */
int ngc_make_pb(ngc_context *ctx, void *stream, const double *freq_chan, int64_t n_chan,
                const double *dish, const double *blockage, int64_t n_dish, int casa,
                int ipower, int64_t n_l, int64_t n_m, int64_t center_l, int64_t center_m,
                double cell_l, double cell_m, int64_t n_pol, double *out);

/*
 * Visibility prediction.  The reference has no implementation: its degridding
 * graph prints that degridding is still to be implemented
 * (_imaging_utils/_standard_grid.py:429-430) and predict_modelvis_image.py:20
 * is a stub naming the steps (correction, Fourier transform, degrid).
 *
 * ngc_model_to_grid: model image [n_planes, n_l, n_m] (device float64,
 * Jy/pixel), divided by corr_l[l] corr_m[m] (host; both NULL: none), padded to
 * the centre of an n_u x n_v grid (start n_u//2 - n_l//2) and transformed as
 * fftshift(fft2(ifftshift(.))) (unnormalised forward): the inverse placement
 * of ngc_grid_to_image.  grid: device [n_planes, n_u, n_v] of grid_dtype
 * (NGC_C64|NGC_C128).  Complex64 grids whose lengths are products of 2, 3, 5,
 * 7, 11 and 13 up to 8192 use the LDS row FFTs, others rocFFT (see
 * ngc_set_fft_mode).
 *
 * ngc_standard_degrid: for every (time, baseline, chan, pol) sample placed on
 * the grid as the standard gridder places it (same uv scale, centre cell,
 * sub-cell offset, bounds and taps, _standard_grid.py:273-360),
 *   V = sum_taps conv * grid[chan_map[c], pol_map[p], u0+i, v0+j]
 * divided by sum_taps conv when `normalize`; 0 for samples off the grid or
 * with NaN uvw.  The adjoint of ngc_standard_grid with unit weights.
 *   grid  device [n_imag_chan, n_imag_pol, n_u, n_v] of grid_dtype (C64|C128)
 *   vis   device out [T, B, C, P] of vis_dtype (C64|C128)
 *   uvw device; freq_chan, chan_map, pol_map, cgk_1d host
 * Complex64 grids with support <= 8 and 1, 2 or 4 pols are degridded from
 * LDS tiles (samples binned by the tile of their footprint, workspace of
 * 8 bytes per sample); the others, and every call after
 * ngc_set_degrid_mode(ctx, 0), by a per-sample gather.  Both give the same
 * bits.
 */
int ngc_model_to_grid(ngc_context *ctx, void *stream, const double *image, int64_t n_planes,
                      int64_t n_l, int64_t n_m, int64_t n_u, int64_t n_v, const double *corr_l,
                      const double *corr_m, void *grid, int grid_dtype);
int ngc_standard_degrid(ngc_context *ctx, void *stream, const double *uvw,
                        const double *freq_chan, const int32_t *chan_map,
                        const int32_t *pol_map, const double *cgk_1d, int64_t cgk_len,
                        int64_t n_time, int64_t n_baseline, int64_t n_chan, int64_t n_pol,
                        int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u, int64_t n_v,
                        double delta_l, double delta_m, int64_t support, int64_t oversampling,
                        const void *grid, int grid_dtype, void *vis, int vis_dtype,
                        int normalize);

/*
 * ngc_aperture_degrid: visibility prediction with the mosaic A-term kernels,
 * the adjoint of ngc_aperture_grid (image mode, _aperture_grid.py:377-513)
 * with unit weights.  For every (time, baseline, chan, pol) sample that the
 * aperture gridder places (row field id found in field_id, same uv scale,
 * centre cell, sub-cell offsets in os_u / os_v, bounds by the largest
 * support, CK * PG[field] taps),
 *   V = sum_taps conj(CK * PG) * grid[chan_map[c], pol_map[p], cell]
 * divided by the kernel's own tap sum, sum_taps CK (without the phase
 * gradient), when `normalize`: the beam is 1 at the field's pointing centre,
 * so a point source there predicts its flux and one elsewhere its flux times
 * the beam at that offset; 0 for rows on field <= -1 and samples off the grid.
 * Arrays as for ngc_aperture_grid: field device int32 [T,B]; field_id,
 * freq_chan, chan_map, pol_map, cf_*_map and support host; conv_kernel device
 * float64 [n_cfb, n_cfc, n_cfp, k_u, k_v]; phase_gradient device complex128
 * [n_field, k_u, k_v]; grid device [n_imag_chan, n_imag_pol, n_u, n_v] of
 * grid_dtype (C64|C128); vis device out [T, B, C, P] of vis_dtype (C64|C128).
 * Sums are FP64 for either grid precision.  A phase gradient that factors as
 * PG[x][y] = PG[x][K_v/2] PG[K_u/2][y] (to 1e-11 absolute; the gcf builder's
 * exp(i(x px + y py)) does) takes the tap-parallel kernel; any other takes a
 * one-thread-per-sample kernel with the same result.
 * Returns NGC_ERR_FIELD if a row's field id is not in field_id (synchronises
 * `stream`).
 */
int ngc_aperture_degrid(ngc_context *ctx, void *stream, const double *uvw, const int32_t *field,
                        const int32_t *field_id, int64_t n_field, const double *freq_chan,
                        const int32_t *chan_map, const int32_t *pol_map,
                        const int32_t *cf_baseline_map, const int32_t *cf_chan_map,
                        const int32_t *cf_pol_map, const double *conv_kernel, int64_t n_cfb,
                        int64_t n_cfc, int64_t n_cfp, int64_t k_u, int64_t k_v,
                        const int32_t *support, const void *phase_gradient, int64_t n_time,
                        int64_t n_baseline, int64_t n_chan, int64_t n_pol, int64_t n_imag_chan,
                        int64_t n_imag_pol, int64_t n_u, int64_t n_v, double delta_l,
                        double delta_m, int64_t os_u, int64_t os_v, const void *grid,
                        int grid_dtype, void *vis, int vis_dtype, int normalize);

/*
 * Gridding plans (no counterpart in the reference, whose graph layer rebuilds
 * its dask graph on every call).  A plan fixes every argument of one
 * ngc_standard_grid call in a tiled precision (IMAGE or PSF mode): the input,
 * grid and sum_weight pointers, the shapes and the tables.  Creation runs the
 * autotuner when `autotune` is nonzero (tile rows 8/16/32 x 1/2/4/8 channels per
 * binning wavefront, each timed on these inputs; the choice is cached per
 * shape in the context), sizes the bins with one counting pass, and captures
 * the whole pipeline (count, scan, fill, largest-first bin order, persistent
 * gridding kernel) in a hipGraph.  ngc_grid_plan_execute replays it on
 * `stream` with one launch, asynchronously; the inputs may be rewritten in
 * place between replays.  ngc_grid_plan_status synchronises `stream` and
 * fails if the inputs now bin more records than the plan holds (1/16 spare).
 */
typedef struct ngc_grid_plan ngc_grid_plan;
int ngc_grid_plan_create(ngc_context *ctx, void *stream, int mode, int precision,
                         const double *uvw, const void *vis, int vis_dtype, const void *weight,
                         int weight_dtype, const uint8_t *flag, const double *freq_chan,
                         const int32_t *chan_map, const int32_t *pol_map, const double *cgk_1d,
                         int64_t cgk_len, int64_t n_time, int64_t n_baseline, int64_t n_chan,
                         int64_t n_pol, int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
                         int64_t n_v, double delta_l, double delta_m, int64_t support,
                         int64_t oversampling, int accumulate, void *grid, double *sum_weight,
                         int autotune, ngc_grid_plan **plan);
/*
 * ngc_aperture_grid_plan_create: the same for one ngc_aperture_grid call in a
 * tiled precision (IMAGE or PSF mode), so that a mosaic's pointings x
 * channels x pols replay as one captured graph (the reference builds one
 * dask task per chunk, ngcasa/imaging/_imaging_utils/_aperture_grid.py:25-142).
 * Arguments as ngc_aperture_grid, plus `autotune` (tile edge 16/24/32 timed
 * on these inputs; cached per shape) and the returned plan.  The convolution
 * kernels and phase gradients are read at creation (their re-laid tables are
 * built once); uvw, vis, weight, flag and field may be rewritten in place
 * between replays.  ngc_grid_plan_status also fails with NGC_ERR_FIELD when a
 * replay met a row whose field id is not in field_id.  Execute, status,
 * describe and destroy are the ngc_grid_plan_* calls.
 */
int ngc_aperture_grid_plan_create(
    ngc_context *ctx, void *stream, int mode, int precision, const double *uvw,
    const void *vis, int vis_dtype, const void *weight, int weight_dtype, const uint8_t *flag,
    const int32_t *field, const int32_t *field_id, int64_t n_field, const double *freq_chan,
    const int32_t *chan_map, const int32_t *pol_map, const int32_t *cf_baseline_map,
    const int32_t *cf_chan_map, const int32_t *cf_pol_map, const double *conv_kernel,
    int64_t n_cfb, int64_t n_cfc, int64_t n_cfp, int64_t k_u, int64_t k_v,
    const int32_t *support, const double *phase_gradient, int64_t n_time, int64_t n_baseline,
    int64_t n_chan, int64_t n_pol, int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
    int64_t n_v, double delta_l, double delta_m, int64_t os_u, int64_t os_v, int accumulate,
    void *grid, double *sum_weight, int autotune, ngc_grid_plan **plan);
int ngc_grid_plan_execute(ngc_grid_plan *plan, void *stream);
int ngc_grid_plan_status(ngc_grid_plan *plan, void *stream);
int ngc_grid_plan_describe(ngc_grid_plan *plan, char *buf, int64_t len);
int ngc_grid_plan_destroy(ngc_grid_plan *plan);

/*
 * Dense layers of the learned A-term kernel surrogate (no counterpart in the
 * reference).  All arrays are device FP32; GEMM operands are rounded to bf16
 * as they are loaded and multiplied on the matrix cores with FP32
 * accumulation.
 *   ngc_bf16_gemm      C[m][n] = act(sum_k A(m,k) B(k,n) + bias[n]) (+ C[m][n] when
 *                      accumulate), A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn],
 *                      act 0 = identity, 1 = tanh, 2 = times the tanh derivative
 *                      1 - aux[m][n]^2 (aux: activations, row stride ldc; the
 *                      backward step of a tanh layer fused); Z (or NULL)
 *                      receives the pre-activation values (row stride ldc).
 *                      Output blocks that cannot fill the device split K; the
 *                      partial products are summed in split order.
 *   ngc_tanh_backward  dZ = dA * (1 - A^2) for A = tanh(Z)
 *   ngc_col_sum        out[n] = sum_m X[m][n]   (X is M x N, row major)
 *   ngc_mse_grad       g = 2 (y - t) / n, *loss = mean (y - t)^2 (device double)
 *   ngc_adam_step      Adam (Kingma & Ba 2015) on FP32 master weights, step >= 1
 */
int ngc_bf16_gemm(ngc_context *ctx, void *stream, int64_t M, int64_t N, int64_t K,
                  const float *A, int64_t sam, int64_t sak, const float *B, int64_t sbk,
                  int64_t sbn, const float *bias, int act, int accumulate, float *C, int64_t ldc,
                  float *Z, const float *aux);
int ngc_tanh_backward(ngc_context *ctx, void *stream, int64_t n, const float *dA, const float *A,
                      float *dZ);
int ngc_col_sum(ngc_context *ctx, void *stream, int64_t M, int64_t N, const float *X,
                float *out);
int ngc_mse_grad(ngc_context *ctx, void *stream, int64_t n, const float *y, const float *t,
                 float *g, double *loss);
int ngc_adam_step(ngc_context *ctx, void *stream, int64_t n, float *w, const float *g, float *m,
                  float *v, double lr, double beta1, double beta2, double eps, int64_t step);

/*
 * Training with a device-side step counter (no counterpart in the reference),
 * so that a whole training step can be captured once in a graph and
 * replayed: nothing step-dependent is passed from the host.
 *   ngc_mlp_sample_rows   minibatch of step k = *counter: rows
 *                         splitmix64(seed ^ splitmix64(k*batch + j)) mod n of
 *                         x [n, dx] and t [n, dt] into xb [batch, dx], tb [batch, dt]
 *   ngc_adam_schedule     sched = {cosine-decayed rate lr -> lr_min over `steps`
 *                         at step k, 1 - beta1^s, 1 - beta2^s} with
 *                         s = step0 + k + 1; then *counter = k + 1
 *   ngc_adam_step_sched   the Adam update with the rate and bias corrections
 *                         read from sched (device float [3])
 * counter: device int64; all arrays device FP32.
 */
int ngc_mlp_sample_rows(ngc_context *ctx, void *stream, const float *x, const float *t,
                        int64_t n, int64_t dx, int64_t dt, int64_t batch, uint64_t seed,
                        const int64_t *counter, float *xb, float *tb);
int ngc_adam_schedule(ngc_context *ctx, void *stream, int64_t *counter, float *sched,
                      double lr, double lr_min, int64_t steps, int64_t step0, double beta1,
                      double beta2);
int ngc_adam_step_sched(ngc_context *ctx, void *stream, int64_t n, float *w, const float *g,
                        float *m, float *v, const float *sched, double beta1, double beta2,
                        double eps);

/*
 * Tile shape of the uv-tiled gridders: 8, 16 or 32 grid rows per tile (the
 * tile is always 32 columns wide), or 0 (the default) for 32 rows on calls
 * with at least 65,536 (image channel, 32 x 32 tile) bins and 16 rows below
 * that.  Set by the autotuner (no counterpart in the reference, whose gridder
 * has no tiles).
 */
int ngc_set_tile_rows(ngc_context *ctx, int tile_rows);

/*
 * Wavefronts per tile of the uv-tiled standard gridder (no counterpart in the
 * reference): 1, 2 or 4 wavefronts of one workgroup share the tile's LDS
 * subgrid, each gridding a contiguous share of the tile's records into every
 * pol.  0 (the default) = 4 for the single-precision forms, 2 for FP64 tiled.
 * Results agree to the rounding of the register runs, whose grouping of the
 * records follows the shares.
 */
int ngc_set_tile_waves(ngc_context *ctx, int waves);

/*
 * Channels per binning wavefront of the uv-tiled standard gridder (no
 * counterpart in the reference): a wavefront bins 64/lanes integrations x
 * `lanes` adjacent channels of one baseline (1, 2, 4 or 8; default 2).  The
 * plan autotuner chooses it per shape.
 */
int ngc_set_bin_lanes(ngc_context *ctx, int lanes);

/*
 * Phase timing (instrumentation; no counterpart in the reference, which has
 * no timing inside its kernels).  When enabled, each gridding call records
 * HIP events on its stream around its phases; ngc_profile_read synchronises
 * that stream and returns up to `max` (name, milliseconds) pairs of the last
 * call: names are written as a ';'-separated list into `names` (capacity
 * `names_len`), durations into `ms`.  Returns the number of phases.
 */
/*
 * FFTs of single-precision grids in ngc_grid_to_image and ngc_model_to_grid
 * (no counterpart in the reference, which uses numpy/dask FFTs): mode 1
 * (default) uses the LDS kernels where the lengths allow, mode 0 always rocFFT.
 */
int ngc_set_fft_mode(ngc_context *ctx, int mode);

/*
 * Standard degridding path (no counterpart in the reference, whose degridder
 * is a stub): mode 1 (default) LDS tiles where ngc_standard_degrid allows
 * them, mode 0 always the per-sample gather.
 */
int ngc_set_degrid_mode(ngc_context *ctx, int mode);

int ngc_profile_enable(ngc_context *ctx, int enable);
int ngc_profile_read(ngc_context *ctx, char *names, int64_t names_len, double *ms,
                     int64_t max);

#ifdef __cplusplus
}
#endif

#endif /* NGCASA_GRIDDING_H */
