/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C, FP64, single-threaded restatement of the four numba kernels on
 * ngCASA's gridding path.  It is the checker for the HIP library: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  Nothing in the product links or calls it.
 *
 * Every loop keeps the reference's loop order, skip rules, rounding rules
 * and bounds rules:
 *   ngcasa/imaging/_imaging_utils/_standard_grid.py:243-371  (_standard_grid_jit)
 *   ngcasa/imaging/_imaging_utils/_standard_grid.py:467-518  (_standard_imaging_weight_degrid_jit)
 *   ngcasa/imaging/_imaging_utils/_aperture_grid.py:377-513  (_aperture_grid_jit)
 *   ngcasa/imaging/_imaging_utils/_aperture_grid.py:181-291  (_aperture_weight_grid_jit)
 *
 * Semantics restated from the numba source text:
 *   - `~np.isnan(x)` on a numba boolean is logical not, so NaN u/v and NaN
 *     weighted data are skipped.
 *   - `int(x)` truncates toward zero (LLVM fptosi); `math.floor` floors.
 *   - numba does no bounds checking: an index equal to the axis length lands
 *     at the flat C-order offset.  The density gridder's conjugate cell is not
 *     bounds-checked by the reference (_standard_grid.py:317-320, 363-364);
 *     this restatement writes such a tap at its flat offset when that offset
 *     lies inside the grid allocation and drops it otherwise.
 *   - complex*real products are done as numba does them: the real operand is
 *     promoted to complex with zero imaginary part.  For finite values this
 *     equals scaling both components, which is what is written here.
 *
 * Complex arrays are interleaved (re, im) doubles.  Build with
 * -ffp-contract=off so that no multiply-add is fused.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static const double C_LIGHT = 299792458.0;

static inline int is_nan(double x) { return x != x; }

/* ----------------------------------------------------------------------- */
/* _standard_grid_jit (_standard_grid.py:243-371)                            */
/* ----------------------------------------------------------------------- */
/*
 * grid        : [n_imag_chan, n_imag_pol, n_u, n_v], complex (interleaved) if
 *               complex_grid else real
 * sum_weight  : [n_imag_chan, n_imag_pol]
 * vis         : [n_time, n_baseline, n_chan, n_pol] complex (ignored if do_psf)
 * uvw         : [n_time, n_baseline, 3]
 * weight      : [n_time, n_baseline, n_chan, n_pol]
 */
void oracle_standard_grid(double *grid, double *sum_weight, int do_psf,
                          int do_imaging_weight, int complex_grid,
                          const double *vis, const double *uvw,
                          const double *freq_chan, const int64_t *chan_map,
                          const int64_t *pol_map, const double *weight,
                          const double *cgk_1d, int64_t n_time,
                          int64_t n_baseline, int64_t n_chan, int64_t n_pol,
                          int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
                          int64_t n_v, double delta_l, double delta_m,
                          int64_t support, int64_t oversampling)
{
    double *uv_scale0 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    double *uv_scale1 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    /* _standard_grid.py:274-276: -(freq*delta*n)/c, evaluated left to right */
    for (int64_t c = 0; c < n_chan; c++) {
        uv_scale0[c] = -((freq_chan[c] * delta_l) * (double)n_u) / C_LIGHT;
        uv_scale1[c] = -((freq_chan[c] * delta_m) * (double)n_v) / C_LIGHT;
    }
    const int64_t support_center = support / 2;
    const int64_t uv_center0 = n_u / 2, uv_center1 = n_v / 2;
    const int64_t start_support = -support_center;
    const int64_t end_support = support - support_center;
    const int64_t plane = n_u * n_v;
    const int64_t grid_cells = n_imag_chan * n_imag_pol * plane;

    for (int64_t t = 0; t < n_time; t++) {
        for (int64_t b = 0; b < n_baseline; b++) {
            for (int64_t ch = 0; ch < n_chan; ch++) {
                const int64_t a_chan = chan_map[ch];
                const double *uvw_tb = uvw + (t * n_baseline + b) * 3;
                const double u = uvw_tb[0] * uv_scale0[ch];
                const double v = uvw_tb[1] * uv_scale1[ch];
                if (is_nan(u) || is_nan(v))
                    continue;
                const double u_pos = u + (double)uv_center0;
                const double v_pos = v + (double)uv_center1;
                const double u_pos_conj = -u + (double)uv_center0;
                const double v_pos_conj = -v + (double)uv_center1;
                const int64_t u_center_indx = (int64_t)(u_pos + 0.5);
                const int64_t v_center_indx = (int64_t)(v_pos + 0.5);
                const int64_t u_center_indx_conj = (int64_t)(u_pos_conj + 0.5);
                const int64_t v_center_indx_conj = (int64_t)(v_pos_conj + 0.5);
                if (!((u_center_indx + support_center < n_u) &&
                      (v_center_indx + support_center < n_v) &&
                      (u_center_indx - support_center >= 0) &&
                      (v_center_indx - support_center >= 0)))
                    continue;
                const double u_offset = (double)u_center_indx - u_pos;
                const int64_t u_center_offset_indx =
                    (int64_t)floor(u_offset * (double)oversampling + 0.5);
                const double v_offset = (double)v_center_indx - v_pos;
                const int64_t v_center_offset_indx =
                    (int64_t)floor(v_offset * (double)oversampling + 0.5);

                const int64_t vis_base = ((t * n_baseline + b) * n_chan + ch) * n_pol;
                for (int64_t p = 0; p < n_pol; p++) {
                    double wd_re, wd_im, sel_weight;
                    if (do_psf) {
                        if (n_pol >= 2 && do_imaging_weight) {
                            wd_re = (weight[vis_base + 0] + weight[vis_base + 1]) / 2.0;
                            sel_weight = wd_re;
                        } else {
                            sel_weight = weight[vis_base + p];
                            wd_re = weight[vis_base + p];
                        }
                        wd_im = 0.0;
                    } else {
                        sel_weight = weight[vis_base + p];
                        wd_re = vis[2 * (vis_base + p)] * weight[vis_base + p];
                        wd_im = vis[2 * (vis_base + p) + 1] * weight[vis_base + p];
                    }
                    if (is_nan(wd_re) || is_nan(wd_im))
                        continue;
                    if (wd_re == 0.0 && wd_im == 0.0)
                        continue;
                    const int64_t a_pol = pol_map[p];
                    const int64_t pbase = (a_chan * n_imag_pol + a_pol) * plane;
                    double norm = 0.0;
                    for (int64_t iv = start_support; iv < end_support; iv++) {
                        const int64_t v_indx = v_center_indx + iv;
                        const int64_t v_off = llabs(oversampling * iv + v_center_offset_indx);
                        const double conv_v = cgk_1d[v_off];
                        const int64_t v_indx_conj = v_center_indx_conj + iv;
                        for (int64_t iu = start_support; iu < end_support; iu++) {
                            const int64_t u_indx = u_center_indx + iu;
                            const int64_t u_off = llabs(oversampling * iu + u_center_offset_indx);
                            const double conv_u = cgk_1d[u_off];
                            const double conv = conv_u * conv_v;
                            const int64_t cell = pbase + u_indx * n_v + v_indx;
                            if (complex_grid) {
                                grid[2 * cell] = grid[2 * cell] + conv * wd_re;
                                grid[2 * cell + 1] = grid[2 * cell + 1] + conv * wd_im;
                            } else {
                                grid[cell] = grid[cell] + conv * wd_re;
                            }
                            norm = norm + conv;
                            if (do_imaging_weight) {
                                const int64_t u_indx_conj = u_center_indx_conj + iu;
                                const int64_t ccell = pbase + u_indx_conj * n_v + v_indx_conj;
                                if (ccell >= 0 && ccell < grid_cells) {
                                    if (complex_grid) {
                                        grid[2 * ccell] = grid[2 * ccell] + conv * wd_re;
                                        grid[2 * ccell + 1] = grid[2 * ccell + 1] + conv * wd_im;
                                    } else {
                                        grid[ccell] = grid[ccell] + conv * wd_re;
                                    }
                                }
                            }
                        }
                    }
                    const int64_t sw = a_chan * n_imag_pol + a_pol;
                    sum_weight[sw] = sum_weight[sw] + sel_weight * norm;
                    if (do_imaging_weight)
                        sum_weight[sw] = sum_weight[sw] + sel_weight * norm;
                }
            }
        }
    }
    free(uv_scale0);
    free(uv_scale1);
}

/* ----------------------------------------------------------------------- */
/* Visibility degridding (prediction): the gather counterpart of the loop    */
/* above.  The reference has no implementation (_standard_grid.py:429-430    */
/* prints that it is missing; predict_modelvis_image.py:20 is a stub), so    */
/* this restatement is defined as the adjoint of oracle_standard_grid with   */
/* unit weights: the same uv_scale, centre cell, sub-cell offset, bounds and */
/* tap index, gathering instead of scattering.                               */
/* ----------------------------------------------------------------------- */
/*
 * vis   : out [n_time, n_baseline, n_chan, n_pol] complex (interleaved);
 *         V = sum_taps conv * G[cell] (divided by sum_taps conv when
 *         normalize), 0 for samples that are not placed on the grid
 * grid  : [n_imag_chan, n_imag_pol, n_u, n_v] complex (interleaved)
 */
void oracle_standard_degrid(double *vis, const double *grid, const double *uvw,
                            const double *freq_chan, const int64_t *chan_map,
                            const int64_t *pol_map, const double *cgk_1d,
                            int64_t n_time, int64_t n_baseline, int64_t n_chan,
                            int64_t n_pol, int64_t n_imag_chan, int64_t n_imag_pol,
                            int64_t n_u, int64_t n_v, double delta_l, double delta_m,
                            int64_t support, int64_t oversampling, int normalize)
{
    (void)n_imag_chan;
    double *uv_scale0 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    double *uv_scale1 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    for (int64_t c = 0; c < n_chan; c++) {
        uv_scale0[c] = -((freq_chan[c] * delta_l) * (double)n_u) / C_LIGHT;
        uv_scale1[c] = -((freq_chan[c] * delta_m) * (double)n_v) / C_LIGHT;
    }
    const int64_t support_center = support / 2;
    const int64_t uv_center0 = n_u / 2, uv_center1 = n_v / 2;
    const int64_t start_support = -support_center;
    const int64_t end_support = support - support_center;
    const int64_t plane = n_u * n_v;
    for (int64_t t = 0; t < n_time; t++) {
        for (int64_t b = 0; b < n_baseline; b++) {
            for (int64_t ch = 0; ch < n_chan; ch++) {
                const int64_t vis_base = ((t * n_baseline + b) * n_chan + ch) * n_pol;
                for (int64_t p = 0; p < n_pol; p++) {
                    vis[2 * (vis_base + p)] = 0.0;
                    vis[2 * (vis_base + p) + 1] = 0.0;
                }
                const int64_t a_chan = chan_map[ch];
                const double *uvw_tb = uvw + (t * n_baseline + b) * 3;
                const double u = uvw_tb[0] * uv_scale0[ch];
                const double v = uvw_tb[1] * uv_scale1[ch];
                if (is_nan(u) || is_nan(v))
                    continue;
                const double u_pos = u + (double)uv_center0;
                const double v_pos = v + (double)uv_center1;
                const int64_t u_center_indx = (int64_t)(u_pos + 0.5);
                const int64_t v_center_indx = (int64_t)(v_pos + 0.5);
                if (!((u_center_indx + support_center < n_u) &&
                      (v_center_indx + support_center < n_v) &&
                      (u_center_indx - support_center >= 0) &&
                      (v_center_indx - support_center >= 0)))
                    continue;
                const int64_t u_center_offset_indx = (int64_t)floor(
                    ((double)u_center_indx - u_pos) * (double)oversampling + 0.5);
                const int64_t v_center_offset_indx = (int64_t)floor(
                    ((double)v_center_indx - v_pos) * (double)oversampling + 0.5);
                for (int64_t p = 0; p < n_pol; p++) {
                    const int64_t pbase = (a_chan * n_imag_pol + pol_map[p]) * plane;
                    double re = 0.0, im = 0.0, norm = 0.0;
                    for (int64_t iv = start_support; iv < end_support; iv++) {
                        const int64_t v_indx = v_center_indx + iv;
                        const double conv_v =
                            cgk_1d[llabs(oversampling * iv + v_center_offset_indx)];
                        for (int64_t iu = start_support; iu < end_support; iu++) {
                            const int64_t u_indx = u_center_indx + iu;
                            const double conv_u =
                                cgk_1d[llabs(oversampling * iu + u_center_offset_indx)];
                            const double conv = conv_u * conv_v;
                            const int64_t cell = pbase + u_indx * n_v + v_indx;
                            re = re + conv * grid[2 * cell];
                            im = im + conv * grid[2 * cell + 1];
                            norm = norm + conv;
                        }
                    }
                    if (normalize && norm != 0.0) {
                        re = re / norm;
                        im = im / norm;
                    }
                    vis[2 * (vis_base + p)] = re;
                    vis[2 * (vis_base + p) + 1] = im;
                }
            }
        }
    }
    free(uv_scale0);
    free(uv_scale1);
}

/* ----------------------------------------------------------------------- */
/* _standard_imaging_weight_degrid_jit (_standard_grid.py:467-518)          */
/* ----------------------------------------------------------------------- */
/*
 * imaging_weight      : out [n_time, n_baseline, n_chan, n_pol], zero-filled
 * grid_imaging_weight : [n_u, n_v, n_imag_chan, n_imag_pol] (image orientation,
 *                       as the reference indexes it after its moveaxis)
 * briggs_factors      : [2, n_imag_chan, n_imag_pol]
 */
void oracle_imaging_weight_degrid(double *imaging_weight,
                                  const double *grid_imaging_weight,
                                  const double *briggs_factors, const double *uvw,
                                  const double *freq_chan, const int64_t *chan_map,
                                  const int64_t *pol_map,
                                  const double *natural_imaging_weight,
                                  int64_t n_time, int64_t n_baseline,
                                  int64_t n_chan, int64_t n_pol,
                                  int64_t n_imag_chan, int64_t n_imag_pol,
                                  int64_t n_u, int64_t n_v, double delta_l,
                                  double delta_m)
{
    double *uv_scale0 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    double *uv_scale1 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    for (int64_t c = 0; c < n_chan; c++) {
        uv_scale0[c] = -((freq_chan[c] * delta_l) * (double)n_u) / C_LIGHT;
        uv_scale1[c] = -((freq_chan[c] * delta_m) * (double)n_v) / C_LIGHT;
    }
    const int64_t uv_center0 = n_u / 2, uv_center1 = n_v / 2;
    for (int64_t t = 0; t < n_time; t++) {
        for (int64_t b = 0; b < n_baseline; b++) {
            for (int64_t ch = 0; ch < n_chan; ch++) {
                const int64_t a_chan = chan_map[ch];
                const double *uvw_tb = uvw + (t * n_baseline + b) * 3;
                const double u = uvw_tb[0] * uv_scale0[ch];
                const double v = uvw_tb[1] * uv_scale1[ch];
                if (is_nan(u) || is_nan(v))
                    continue;
                const double u_pos = u + (double)uv_center0;
                const double v_pos = v + (double)uv_center1;
                const int64_t u_center_indx = (int64_t)(u_pos + 0.5);
                const int64_t v_center_indx = (int64_t)(v_pos + 0.5);
                if (!((u_center_indx < n_u) && (v_center_indx < n_v) &&
                      (u_center_indx >= 0) && (v_center_indx >= 0)))
                    continue;
                const int64_t base = ((t * n_baseline + b) * n_chan + ch) * n_pol;
                for (int64_t p = 0; p < n_pol; p++) {
                    const int64_t a_pol = pol_map[p];
                    if (n_pol == 2)
                        imaging_weight[base + p] =
                            (natural_imaging_weight[base + 0] + natural_imaging_weight[base + 1]) / 2.0;
                    else
                        imaging_weight[base + p] = natural_imaging_weight[base + p];
                    const double nw = natural_imaging_weight[base + p];
                    if (!is_nan(nw) && nw != 0.0) {
                        const double g = grid_imaging_weight
                            [((u_center_indx * n_v + v_center_indx) * n_imag_chan + a_chan) * n_imag_pol + a_pol];
                        if (!is_nan(g) && g != 0.0) {
                            const double bg =
                                briggs_factors[(0 * n_imag_chan + a_chan) * n_imag_pol + a_pol] * g +
                                briggs_factors[(1 * n_imag_chan + a_chan) * n_imag_pol + a_pol];
                            imaging_weight[base + p] = imaging_weight[base + p] / bg;
                        }
                    }
                }
            }
        }
    }
    free(uv_scale0);
    free(uv_scale1);
}

/* ----------------------------------------------------------------------- */
/* _aperture_grid_jit (_aperture_grid.py:377-513)                           */
/* ----------------------------------------------------------------------- */
/*
 * grid          : out [n_imag_chan, n_imag_pol, n_u, n_v] complex
 * vis           : [T, B, C, P] complex (ignored if do_psf)
 * field         : [T, B] int64 field id per row (<= -1 skips the row)
 * field_id      : [n_field] the gcf's field ids (field index = position)
 * conv_kernel   : [n_cfb, n_cfc, n_cfp, K_u, K_v] real
 * weight_support: [n_cfb, n_cfc, n_cfp, 2] int64
 * phase_gradient: [n_field, K_u, K_v] complex
 * Returns 0, or -1 when a row's field id is not in field_id (the reference
 * raises IndexError from np.where(...)[0][0]).
 */
int oracle_aperture_grid(double *grid, double *sum_weight, int do_psf,
                         const double *vis, const double *uvw,
                         const double *freq_chan, const int64_t *chan_map,
                         const int64_t *pol_map, const int64_t *cf_baseline_map,
                         const int64_t *cf_chan_map, const int64_t *cf_pol_map,
                         const double *imaging_weight, const double *conv_kernel,
                         int64_t n_cfb, int64_t n_cfc, int64_t n_cfp,
                         int64_t k_u, int64_t k_v, int64_t n_time,
                         int64_t n_baseline, int64_t n_chan, int64_t n_pol,
                         int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
                         int64_t n_v, double delta_l, double delta_m,
                         const int64_t *weight_support, int64_t os_u,
                         int64_t os_v, const int64_t *field,
                         const int64_t *field_id, int64_t n_field,
                         const double *phase_gradient)
{
    double *uv_scale0 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    double *uv_scale1 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    for (int64_t c = 0; c < n_chan; c++) {
        uv_scale0[c] = -((freq_chan[c] * delta_l) * (double)n_u) / C_LIGHT;
        uv_scale1[c] = -((freq_chan[c] * delta_m) * (double)n_v) / C_LIGHT;
    }
    const int64_t uv_center0 = n_u / 2, uv_center1 = n_v / 2;
    int64_t max_support_center = weight_support[0];
    for (int64_t i = 1; i < n_cfb * n_cfc * n_cfp * 2; i++)
        if (weight_support[i] > max_support_center)
            max_support_center = weight_support[i];
    const int64_t conv_v_center = k_v / 2;
    const int64_t conv_u_center = k_u / 2;
    const int64_t kplane = k_u * k_v;
    const int64_t plane = n_u * n_v;

    for (int64_t t = 0; t < n_time; t++) {
        for (int64_t b = 0; b < n_baseline; b++) {
            const int64_t fval = field[t * n_baseline + b];
            if (!(fval > -1))
                continue;
            int64_t field_indx = -1;
            for (int64_t f = 0; f < n_field; f++)
                if (field_id[f] == fval) { field_indx = f; break; }
            if (field_indx < 0) {
                free(uv_scale0);
                free(uv_scale1);
                return -1;
            }
            const double *pg = phase_gradient + 2 * field_indx * kplane;
            const int64_t cf_baseline = cf_baseline_map[b];
            for (int64_t ch = 0; ch < n_chan; ch++) {
                const int64_t cf_chan = cf_chan_map[ch];
                const int64_t a_chan = chan_map[ch];
                const double *uvw_tb = uvw + (t * n_baseline + b) * 3;
                const double u = uvw_tb[0] * uv_scale0[ch];
                const double v = uvw_tb[1] * uv_scale1[ch];
                if (is_nan(u) || is_nan(v))
                    continue;
                const double u_pos = u + (double)uv_center0;
                const double v_pos = v + (double)uv_center1;
                const int64_t u_center_indx = (int64_t)(u_pos + 0.5);
                const int64_t v_center_indx = (int64_t)(v_pos + 0.5);
                if (!((u_center_indx + max_support_center < n_u) &&
                      (v_center_indx + max_support_center < n_v) &&
                      (u_center_indx - max_support_center >= 0) &&
                      (v_center_indx - max_support_center >= 0)))
                    continue;
                const double u_offset = (double)u_center_indx - u_pos;
                const int64_t u_center_offset_indx =
                    (int64_t)floor(u_offset * (double)os_u + 0.5) + conv_u_center;
                const double v_offset = (double)v_center_indx - v_pos;
                const int64_t v_center_offset_indx =
                    (int64_t)floor(v_offset * (double)os_v + 0.5) + conv_v_center;
                const int64_t base = ((t * n_baseline + b) * n_chan + ch) * n_pol;
                for (int64_t p = 0; p < n_pol; p++) {
                    double wd_re, wd_im;
                    const double iw = imaging_weight[base + p];
                    if (do_psf) {
                        wd_re = iw;
                        wd_im = 0.0;
                    } else {
                        wd_re = vis[2 * (base + p)] * iw;
                        wd_im = vis[2 * (base + p) + 1] * iw;
                    }
                    if (is_nan(wd_re) || is_nan(wd_im))
                        continue;
                    if (wd_re == 0.0 && wd_im == 0.0)
                        continue;
                    const int64_t cf_pol = cf_pol_map[p];
                    const int64_t a_pol = pol_map[p];
                    const int64_t sidx = ((cf_baseline * n_cfc + cf_chan) * n_cfp + cf_pol) * 2;
                    const int64_t support_u = weight_support[sidx + 0];
                    const int64_t support_v = weight_support[sidx + 1];
                    const int64_t scu = support_u / 2, scv = support_v / 2;
                    const double *ck = conv_kernel + ((cf_baseline * n_cfc + cf_chan) * n_cfp + cf_pol) * kplane;
                    const int64_t pbase = (a_chan * n_imag_pol + a_pol) * plane;
                    double norm_re = 0.0, norm_im = 0.0;
                    for (int64_t iv = -scv; iv < support_v - scv; iv++) {
                        const int64_t v_indx = v_center_indx + iv;
                        const int64_t cf_v_indx = os_v * iv + v_center_offset_indx;
                        for (int64_t iu = -scu; iu < support_u - scu; iu++) {
                            const int64_t u_indx = u_center_indx + iu;
                            const int64_t cf_u_indx = os_u * iu + u_center_offset_indx;
                            const int64_t kidx = cf_u_indx * k_v + cf_v_indx;
                            /* CKPG = CK * PG[field] (_aperture_grid.py:429) */
                            const double c_re = ck[kidx] * pg[2 * kidx];
                            const double c_im = ck[kidx] * pg[2 * kidx + 1];
                            const int64_t cell = pbase + u_indx * n_v + v_indx;
                            /* complex product conv * weighted_data */
                            grid[2 * cell] = grid[2 * cell] + (c_re * wd_re - c_im * wd_im);
                            grid[2 * cell + 1] = grid[2 * cell + 1] + (c_re * wd_im + c_im * wd_re);
                            norm_re = norm_re + c_re;
                            norm_im = norm_im + c_im;
                        }
                    }
                    const int64_t sw = a_chan * n_imag_pol + a_pol;
                    if (do_psf)
                        sum_weight[sw] = sum_weight[sw] + iw * norm_re;
                    else /* Re(norm**2) = re^2 - im^2 (_aperture_grid.py:511) */
                        sum_weight[sw] = sum_weight[sw] + iw * (norm_re * norm_re - norm_im * norm_im);
                }
            }
        }
    }
    free(uv_scale0);
    free(uv_scale1);
    return 0;
}

/* ----------------------------------------------------------------------- */
/* Aperture degridding (no counterpart in the reference, whose prediction    */
/* is a stub: ngcasa/imaging/predict_modelvis_image.py:20).  Defined as the  */
/* adjoint of oracle_aperture_grid's loop with unit weights: the same rows,  */
/* fields, cells, offsets and CK*PG taps, each sample reading               */
/*   V = sum_taps conj(CK*PG) G[cell]                                        */
/* and, when normalize, V / sum_taps CK: the kernel's own tap sum without    */
/* the phase gradient, so a point source at a field's pointing centre        */
/* predicts its flux and one elsewhere its flux times the beam there.        */
/* Samples that are not placed on the grid get 0.                            */
/* ----------------------------------------------------------------------- */
int oracle_aperture_degrid(double *vis, const double *grid, const double *uvw,
                           const double *freq_chan, const int64_t *chan_map,
                           const int64_t *pol_map, const int64_t *cf_baseline_map,
                           const int64_t *cf_chan_map, const int64_t *cf_pol_map,
                           const double *conv_kernel, int64_t n_cfb, int64_t n_cfc,
                           int64_t n_cfp, int64_t k_u, int64_t k_v, int64_t n_time,
                           int64_t n_baseline, int64_t n_chan, int64_t n_pol,
                           int64_t n_imag_pol, int64_t n_u, int64_t n_v,
                           double delta_l, double delta_m, const int64_t *weight_support,
                           int64_t os_u, int64_t os_v, const int64_t *field,
                           const int64_t *field_id, int64_t n_field,
                           const double *phase_gradient, int normalize)
{
    double *uv_scale0 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    double *uv_scale1 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    for (int64_t c = 0; c < n_chan; c++) {
        uv_scale0[c] = -((freq_chan[c] * delta_l) * (double)n_u) / C_LIGHT;
        uv_scale1[c] = -((freq_chan[c] * delta_m) * (double)n_v) / C_LIGHT;
    }
    const int64_t uv_center0 = n_u / 2, uv_center1 = n_v / 2;
    int64_t max_support_center = weight_support[0];
    for (int64_t i = 1; i < n_cfb * n_cfc * n_cfp * 2; i++)
        if (weight_support[i] > max_support_center)
            max_support_center = weight_support[i];
    const int64_t conv_v_center = k_v / 2;
    const int64_t conv_u_center = k_u / 2;
    const int64_t kplane = k_u * k_v;
    const int64_t plane = n_u * n_v;
    for (int64_t i = 0; i < 2 * n_time * n_baseline * n_chan * n_pol; i++)
        vis[i] = 0.0;

    for (int64_t t = 0; t < n_time; t++) {
        for (int64_t b = 0; b < n_baseline; b++) {
            const int64_t fval = field[t * n_baseline + b];
            if (!(fval > -1))
                continue;
            int64_t field_indx = -1;
            for (int64_t f = 0; f < n_field; f++)
                if (field_id[f] == fval) { field_indx = f; break; }
            if (field_indx < 0) {
                free(uv_scale0);
                free(uv_scale1);
                return -1;
            }
            const double *pg = phase_gradient + 2 * field_indx * kplane;
            const int64_t cf_baseline = cf_baseline_map[b];
            for (int64_t ch = 0; ch < n_chan; ch++) {
                const int64_t cf_chan = cf_chan_map[ch];
                const int64_t a_chan = chan_map[ch];
                const double *uvw_tb = uvw + (t * n_baseline + b) * 3;
                const double u = uvw_tb[0] * uv_scale0[ch];
                const double v = uvw_tb[1] * uv_scale1[ch];
                if (is_nan(u) || is_nan(v))
                    continue;
                const double u_pos = u + (double)uv_center0;
                const double v_pos = v + (double)uv_center1;
                const int64_t u_center_indx = (int64_t)(u_pos + 0.5);
                const int64_t v_center_indx = (int64_t)(v_pos + 0.5);
                if (!((u_center_indx + max_support_center < n_u) &&
                      (v_center_indx + max_support_center < n_v) &&
                      (u_center_indx - max_support_center >= 0) &&
                      (v_center_indx - max_support_center >= 0)))
                    continue;
                const double u_offset = (double)u_center_indx - u_pos;
                const int64_t u_center_offset_indx =
                    (int64_t)floor(u_offset * (double)os_u + 0.5) + conv_u_center;
                const double v_offset = (double)v_center_indx - v_pos;
                const int64_t v_center_offset_indx =
                    (int64_t)floor(v_offset * (double)os_v + 0.5) + conv_v_center;
                const int64_t base = ((t * n_baseline + b) * n_chan + ch) * n_pol;
                for (int64_t p = 0; p < n_pol; p++) {
                    const int64_t cf_pol = cf_pol_map[p];
                    const int64_t a_pol = pol_map[p];
                    const int64_t sidx = ((cf_baseline * n_cfc + cf_chan) * n_cfp + cf_pol) * 2;
                    const int64_t support_u = weight_support[sidx + 0];
                    const int64_t support_v = weight_support[sidx + 1];
                    const int64_t scu = support_u / 2, scv = support_v / 2;
                    const double *ck = conv_kernel + ((cf_baseline * n_cfc + cf_chan) * n_cfp + cf_pol) * kplane;
                    const int64_t pbase = (a_chan * n_imag_pol + a_pol) * plane;
                    double re = 0.0, im = 0.0, norm = 0.0;
                    for (int64_t iv = -scv; iv < support_v - scv; iv++) {
                        const int64_t v_indx = v_center_indx + iv;
                        const int64_t cf_v_indx = os_v * iv + v_center_offset_indx;
                        for (int64_t iu = -scu; iu < support_u - scu; iu++) {
                            const int64_t u_indx = u_center_indx + iu;
                            const int64_t cf_u_indx = os_u * iu + u_center_offset_indx;
                            const int64_t kidx = cf_u_indx * k_v + cf_v_indx;
                            const double c_re = ck[kidx] * pg[2 * kidx];
                            const double c_im = ck[kidx] * pg[2 * kidx + 1];
                            const int64_t cell = pbase + u_indx * n_v + v_indx;
                            const double g_re = grid[2 * cell], g_im = grid[2 * cell + 1];
                            /* conj(c) * g */
                            re = re + (c_re * g_re + c_im * g_im);
                            im = im + (c_re * g_im - c_im * g_re);
                            norm = norm + ck[kidx];
                        }
                    }
                    if (normalize && norm != 0.0) {
                        /* divide by the kernel's own tap sum, without the
                           phase gradient: the beam is 1 at the pointing centre */
                        re = re / norm;
                        im = im / norm;
                    }
                    vis[2 * (base + p)] = re;
                    vis[2 * (base + p) + 1] = im;
                }
            }
        }
    }
    free(uv_scale0);
    free(uv_scale1);
    return 0;
}

/* ----------------------------------------------------------------------- */
/* _aperture_weight_grid_jit (_aperture_grid.py:181-291)                    */
/* ----------------------------------------------------------------------- */
int oracle_aperture_weight_grid(double *grid, double *sum_weight,
                                const double *uvw, const double *freq_chan,
                                const int64_t *chan_map, const int64_t *pol_map,
                                const int64_t *cf_baseline_map,
                                const int64_t *cf_chan_map,
                                const int64_t *cf_pol_map,
                                const double *imaging_weight,
                                const double *weight_conv_kernel, int64_t n_cfb,
                                int64_t n_cfc, int64_t n_cfp, int64_t k_u,
                                int64_t k_v, int64_t n_time, int64_t n_baseline,
                                int64_t n_chan, int64_t n_pol,
                                int64_t n_imag_chan, int64_t n_imag_pol,
                                int64_t n_u, int64_t n_v, double delta_l,
                                double delta_m, const int64_t *weight_support,
                                int64_t os_u, int64_t os_v, const int64_t *field,
                                const int64_t *field_id, int64_t n_field,
                                const double *phase_gradient)
{
    double *uv_scale0 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    double *uv_scale1 = (double *)malloc(sizeof(double) * (size_t)n_chan);
    for (int64_t c = 0; c < n_chan; c++) {
        uv_scale0[c] = -((freq_chan[c] * delta_l) * (double)n_u) / C_LIGHT;
        uv_scale1[c] = -((freq_chan[c] * delta_m) * (double)n_v) / C_LIGHT;
    }
    const int64_t uv_center0 = n_u / 2, uv_center1 = n_v / 2;
    const int64_t u_center = uv_center0, v_center = uv_center1;
    int64_t max_support_center = weight_support[0];
    for (int64_t i = 1; i < n_cfb * n_cfc * n_cfp * 2; i++)
        if (weight_support[i] > max_support_center)
            max_support_center = weight_support[i];
    const int64_t conv_v_center = k_v / 2;
    const int64_t conv_u_center = k_u / 2;
    const int64_t kplane = k_u * k_v;
    const int64_t plane = n_u * n_v;

    for (int64_t t = 0; t < n_time; t++) {
        for (int64_t b = 0; b < n_baseline; b++) {
            const int64_t fval = field[t * n_baseline + b];
            if (!(fval > -1))
                continue;
            int64_t field_indx = -1;
            for (int64_t f = 0; f < n_field; f++)
                if (field_id[f] == fval) { field_indx = f; break; }
            if (field_indx < 0) {
                free(uv_scale0);
                free(uv_scale1);
                return -1;
            }
            const double *pg = phase_gradient + 2 * field_indx * kplane;
            const int64_t cf_baseline = cf_baseline_map[b];
            for (int64_t ch = 0; ch < n_chan; ch++) {
                const int64_t cf_chan = cf_chan_map[ch];
                const int64_t a_chan = chan_map[ch];
                const double *uvw_tb = uvw + (t * n_baseline + b) * 3;
                const double u = uvw_tb[0] * uv_scale0[ch];
                const double v = uvw_tb[1] * uv_scale1[ch];
                if (is_nan(u) || is_nan(v))
                    continue;
                const double u_pos = u + (double)uv_center0;
                const double v_pos = v + (double)uv_center1;
                const int64_t u_center_indx = (int64_t)(u_pos + 0.5);
                const int64_t v_center_indx = (int64_t)(v_pos + 0.5);
                if (!((u_center_indx + max_support_center < n_u) &&
                      (v_center_indx + max_support_center < n_v) &&
                      (u_center_indx - max_support_center >= 0) &&
                      (v_center_indx - max_support_center >= 0)))
                    continue;
                const int64_t base = ((t * n_baseline + b) * n_chan + ch) * n_pol;
                for (int64_t p = 0; p < n_pol; p++) {
                    const double wd = imaging_weight[base + p];
                    if (is_nan(wd) || wd == 0.0)
                        continue;
                    const int64_t cf_pol = cf_pol_map[p];
                    const int64_t a_pol = pol_map[p];
                    const int64_t sidx = ((cf_baseline * n_cfc + cf_chan) * n_cfp + cf_pol) * 2;
                    const int64_t support_u = weight_support[sidx + 0];
                    const int64_t support_v = weight_support[sidx + 1];
                    const int64_t scu = support_u / 2, scv = support_v / 2;
                    const double *ck = weight_conv_kernel + ((cf_baseline * n_cfc + cf_chan) * n_cfp + cf_pol) * kplane;
                    const int64_t pbase = (a_chan * n_imag_pol + a_pol) * plane;
                    double norm_re = 0.0;
                    for (int64_t iv = -scv; iv < support_v - scv; iv++) {
                        const int64_t v_indx = v_center + iv;
                        const int64_t cf_v_indx = os_v * iv + conv_v_center;
                        for (int64_t iu = -scu; iu < support_u - scu; iu++) {
                            const int64_t u_indx = u_center + iu;
                            const int64_t cf_u_indx = os_u * iu + conv_u_center;
                            const int64_t kidx = cf_u_indx * k_v + cf_v_indx;
                            const double c_re = ck[kidx] * pg[2 * kidx];
                            const double c_im = ck[kidx] * pg[2 * kidx + 1];
                            const int64_t cell = pbase + u_indx * n_v + v_indx;
                            grid[2 * cell] = grid[2 * cell] + c_re * wd;
                            grid[2 * cell + 1] = grid[2 * cell + 1] + c_im * wd;
                            norm_re = norm_re + c_re;
                        }
                    }
                    const int64_t sw = a_chan * n_imag_pol + a_pol;
                    sum_weight[sw] = sum_weight[sw] + wd * norm_re;
                }
            }
        }
    }
    free(uv_scale0);
    free(uv_scale1);
    return 0;
}

/* ----------------------------------------------------------------------- */
/* Chunked, multi-threaded driver of oracle_standard_grid, the way the       */
/* reference's graph layer runs it (_standard_grid.py:23-120): one task per  */
/* (time, baseline) chunk with a private grid, then a pairwise tree-sum.     */
/* Used only as the timed CPU baseline.  Returns the thread count used.      */
/* ----------------------------------------------------------------------- */
int oracle_standard_grid_chunked(double *grid, double *sum_weight, int do_psf,
                                 int complex_grid, const double *vis,
                                 const double *uvw, const double *freq_chan,
                                 const int64_t *chan_map, const int64_t *pol_map,
                                 const double *weight, const double *cgk_1d,
                                 int64_t n_time, int64_t n_baseline,
                                 int64_t n_chan, int64_t n_pol,
                                 int64_t n_imag_chan, int64_t n_imag_pol,
                                 int64_t n_u, int64_t n_v, double delta_l,
                                 double delta_m, int64_t support,
                                 int64_t oversampling, int64_t time_chunk,
                                 int n_threads)
{
    const int64_t n_chunks = (n_time + time_chunk - 1) / time_chunk;
    const int64_t gsize = n_imag_chan * n_imag_pol * n_u * n_v * (complex_grid ? 2 : 1);
    const int64_t swsize = n_imag_chan * n_imag_pol;
    double **grids = (double **)calloc((size_t)n_chunks, sizeof(double *));
    double **sws = (double **)calloc((size_t)n_chunks, sizeof(double *));
    int used = 1;
#ifdef _OPENMP
    if (n_threads > 0)
        omp_set_num_threads(n_threads);
#pragma omp parallel
    {
#pragma omp single
        used = omp_get_num_threads();
    }
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t k = 0; k < n_chunks; k++) {
        const int64_t t0 = k * time_chunk;
        const int64_t nt = (t0 + time_chunk <= n_time) ? time_chunk : n_time - t0;
        grids[k] = (double *)calloc((size_t)gsize, sizeof(double));
        sws[k] = (double *)calloc((size_t)swsize, sizeof(double));
        const int64_t rows = n_baseline * n_chan * n_pol;
        oracle_standard_grid(grids[k], sws[k], do_psf, 0, complex_grid,
                             vis ? vis + 2 * t0 * rows : NULL,
                             uvw + 3 * t0 * n_baseline, freq_chan, chan_map,
                             pol_map, weight + t0 * rows, cgk_1d, nt,
                             n_baseline, n_chan, n_pol, n_imag_chan, n_imag_pol,
                             n_u, n_v, delta_l, delta_m, support, oversampling);
    }
    /* pairwise tree-sum (_standard_grid.py:109-120) */
    int64_t n = n_chunks;
    while (n > 1) {
        int64_t m = 0;
        for (int64_t i = 0; i < n; i += 2) {
            if (i < n - 1) {
                double *a = grids[i], *bb = grids[i + 1];
#ifdef _OPENMP
#pragma omp parallel for
#endif
                for (int64_t j = 0; j < gsize; j++)
                    a[j] = a[j] + bb[j];
                for (int64_t j = 0; j < swsize; j++)
                    sws[i][j] = sws[i][j] + sws[i + 1][j];
                free(grids[i + 1]);
                free(sws[i + 1]);
            }
            grids[m] = grids[i];
            sws[m] = sws[i];
            m++;
        }
        n = m;
    }
    memcpy(grid, grids[0], sizeof(double) * (size_t)gsize);
    memcpy(sum_weight, sws[0], sizeof(double) * (size_t)swsize);
    free(grids[0]);
    free(sws[0]);
    free(grids);
    free(sws);
    return used;
}
