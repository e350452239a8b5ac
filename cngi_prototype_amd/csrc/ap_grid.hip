// Aperture (mosaic A-term) gridder for gfx950.
//
// Replaces _aperture_grid_jit (ngcasa/imaging/_imaging_utils/_aperture_grid.py:377-513)
// and _aperture_weight_grid_jit (:181-291) behind ngc_aperture_grid().
//
// Plain FP64 drop-in: one thread per (time, baseline, channel) sample walks
// the pols and the support_u x support_v taps in the reference's loop order.
// The reference multiplies the whole kernel cube by the field's phase gradient
// on every field change (:428-430); here each tap forms CK*PG on the fly from
// the two tables, which gives the same FP64 products.
#include "ngc_internal.hpp"

namespace ngc {

struct ApArgs {
    const double *uvw;
    const void *vis;
    const void *weight;
    const uint8_t *flag;
    const int *field;          // [T,B]
    const int *field_id;       // [n_field]
    const double *uv_scale;    // [2][n_chan]
    const int *chan_map, *pol_map, *cfb, *cfc, *cfp;
    const double *ck;          // [n_cfb, n_cfc, n_cfp, k_u, k_v]
    const int *support;        // [n_cfb, n_cfc, n_cfp, 2]
    const double2 *pg;         // [n_field, k_u, k_v]
    int64_t n_field, n_cfb, n_cfc, n_cfp, k_u, k_v;
    int64_t n_time, n_baseline, n_chan, n_pol, n_imag_chan, n_imag_pol, n_u, n_v;
    int64_t os_u, os_v, max_support;
    int vis_dtype, weight_dtype;
    int *error;
};

__device__ inline double ap_weight(const ApArgs &a, int64_t i) {
    double w = a.weight_dtype == NGC_F32 ? (double)((const float *)a.weight)[i]
                                         : ((const double *)a.weight)[i];
    if (a.flag && a.flag[i]) w = __longlong_as_double(0x7ff8000000000000LL);
    return w;
}

__device__ inline double2 ap_vis(const ApArgs &a, int64_t i) {
    if (a.vis_dtype == NGC_C64) {
        const float2 v = ((const float2 *)a.vis)[i];
        return make_double2(v.x, v.y);
    }
    return ((const double2 *)a.vis)[i];
}

template <int MODE>
__global__ void __launch_bounds__(256)
ap_grid_plain_kernel(ApArgs a, double *grid, double *sum_weight) {
#pragma clang fp contract(off)
    const int64_t n_rc = a.n_time * a.n_baseline * a.n_chan;
    const int64_t plane = a.n_u * a.n_v;
    const int64_t kplane = a.k_u * a.k_v;
    const int64_t cu = a.n_u / 2, cv = a.n_v / 2;
    const int64_t conv_u_center = a.k_u / 2, conv_v_center = a.k_v / 2;
    const int64_t ms = a.max_support;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rc;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ch = i % a.n_chan, row = i / a.n_chan;
        const int64_t b = row % a.n_baseline;
        const int fval = a.field[row];
        if (!(fval > -1)) continue;
        int64_t fidx = -1;
        for (int64_t f = 0; f < a.n_field; f++)
            if (a.field_id[f] == fval) {
                fidx = f;
                break;
            }
        if (fidx < 0) {
            atomicOr(a.error, 1);
            continue;
        }
        // pixel position with the u-axis oversampling, then the v offset index
        // again with the v-axis oversampling (_aperture_grid.py:436-451)
        Pix px, pxv;
        if (!pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                            a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os_u, px))
            continue;
        pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                       a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os_v, pxv);
        if (!((px.uc + ms < a.n_u) && (px.vc + ms < a.n_v) && (px.uc - ms >= 0) &&
              (px.vc - ms >= 0)))
            continue;
        const int64_t u_off = px.du + conv_u_center;
        const int64_t v_off = pxv.dv + conv_v_center;
        const int64_t cf_b = a.cfb[b], cf_c = a.cfc[ch];
        const int64_t a_chan = a.chan_map[ch];
        const int64_t base = i * a.n_pol;
        const double2 *pg = a.pg + fidx * kplane;
        for (int p = 0; p < a.n_pol; p++) {
            const double w = ap_weight(a, base + p);
            double wr = w, wi = 0.0;
            if (MODE == NGC_GRID_IMAGE) {
                const double2 v = ap_vis(a, base + p);
                wr = v.x * w;
                wi = v.y * w;
            }
            if (isnan(wr) || isnan(wi)) continue;
            if (wr == 0.0 && wi == 0.0) continue;
            const int64_t cf_p = a.cfp[p];
            const int64_t a_pol = a.pol_map[p];
            const int64_t sidx = ((cf_b * a.n_cfc + cf_c) * a.n_cfp + cf_p) * 2;
            const int64_t su = a.support[sidx], sv = a.support[sidx + 1];
            const int64_t scu = su / 2, scv = sv / 2;
            const double *ck = a.ck + ((cf_b * a.n_cfc + cf_c) * a.n_cfp + cf_p) * kplane;
            const int64_t pbase = (a_chan * a.n_imag_pol + a_pol) * plane;
            double nr = 0.0, ni = 0.0;
            for (int64_t iv = -scv; iv < sv - scv; iv++) {
                const int64_t v_indx = (MODE == NGC_GRID_WEIGHT ? cv : px.vc) + iv;
                const int64_t cf_v = a.os_v * iv + (MODE == NGC_GRID_WEIGHT ? conv_v_center : v_off);
                for (int64_t iu = -scu; iu < su - scu; iu++) {
                    const int64_t u_indx = (MODE == NGC_GRID_WEIGHT ? cu : px.uc) + iu;
                    const int64_t cf_u =
                        a.os_u * iu + (MODE == NGC_GRID_WEIGHT ? conv_u_center : u_off);
                    const int64_t k = cf_u * a.k_v + cf_v;
                    const double kv = ck[k];
                    const double2 g = pg[k];
                    const double c_re = kv * g.x, c_im = kv * g.y;
                    const int64_t cell = pbase + u_indx * a.n_v + v_indx;
                    atomicAdd(&grid[2 * cell], c_re * wr - c_im * wi);
                    atomicAdd(&grid[2 * cell + 1], c_re * wi + c_im * wr);
                    nr = nr + c_re;
                    ni = ni + c_im;
                }
            }
            const int64_t sw = a_chan * a.n_imag_pol + a_pol;
            if (MODE == NGC_GRID_IMAGE)
                atomicAdd(&sum_weight[sw], w * (nr * nr - ni * ni));  // Re(norm**2)
            else
                atomicAdd(&sum_weight[sw], w * nr);
        }
    }
}

}  // namespace ngc

using namespace ngc;

extern "C" int ngc_aperture_grid(
    ngc_context *ctx, void *stream, int mode, int precision, const double *uvw,
    const void *vis, int vis_dtype, const void *weight, int weight_dtype, const uint8_t *flag,
    const int32_t *field, const int32_t *field_id, int64_t n_field, const double *freq_chan,
    const int32_t *chan_map, const int32_t *pol_map, const int32_t *cf_baseline_map,
    const int32_t *cf_chan_map, const int32_t *cf_pol_map, const double *conv_kernel,
    int64_t n_cfb, int64_t n_cfc, int64_t n_cfp, int64_t k_u, int64_t k_v,
    const int32_t *support, const double *phase_gradient, int64_t n_time, int64_t n_baseline,
    int64_t n_chan, int64_t n_pol, int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
    int64_t n_v, double delta_l, double delta_m, int64_t os_u, int64_t os_v, int accumulate,
    void *grid, double *sum_weight) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(mode == NGC_GRID_IMAGE || mode == NGC_GRID_PSF || mode == NGC_GRID_WEIGHT,
              NGC_ERR_ARG, "mode must be IMAGE, PSF or WEIGHT");
    NGC_CHECK(precision >= NGC_PREC_FP64 && precision <= NGC_PREC_FP32, NGC_ERR_ARG,
              "unknown precision");
    const int64_t n_samples = n_time * n_baseline * n_chan;
    NGC_CHECK(field_id && freq_chan && chan_map && pol_map && cf_baseline_map && cf_chan_map &&
                  cf_pol_map && conv_kernel && support && phase_gradient && grid && sum_weight,
              NGC_ERR_ARG, "NULL array");
    NGC_CHECK((uvw && weight && field) || n_samples == 0, NGC_ERR_ARG,
              "NULL uvw, weight or field");
    NGC_CHECK(mode != NGC_GRID_IMAGE || vis || n_samples == 0, NGC_ERR_ARG, "image mode needs vis");
    NGC_CHECK(n_field >= 1 && n_cfb >= 1 && n_cfc >= 1 && n_cfp >= 1 && k_u >= 1 && k_v >= 1,
              NGC_ERR_ARG, "bad kernel dimensions");
    NGC_CHECK(os_u >= 1 && os_v >= 1, NGC_ERR_ARG, "oversampling must be >= 1");
    for (int64_t b = 0; b < n_baseline; b++)
        NGC_CHECK(cf_baseline_map[b] >= 0 && cf_baseline_map[b] < n_cfb, NGC_ERR_ARG,
                  "cf_baseline_map range");
    for (int64_t c = 0; c < n_chan; c++) {
        NGC_CHECK(cf_chan_map[c] >= 0 && cf_chan_map[c] < n_cfc, NGC_ERR_ARG, "cf_chan_map range");
        NGC_CHECK(chan_map[c] >= 0 && chan_map[c] < n_imag_chan, NGC_ERR_ARG, "chan_map range");
    }
    for (int64_t p = 0; p < n_pol; p++) {
        NGC_CHECK(cf_pol_map[p] >= 0 && cf_pol_map[p] < n_cfp, NGC_ERR_ARG, "cf_pol_map range");
        NGC_CHECK(pol_map[p] >= 0 && pol_map[p] < n_imag_pol, NGC_ERR_ARG, "pol_map range");
    }
    // max(SUPPORT) is the half-width of the reference's bounds check
    // (_aperture_grid.py:397,447); every tap must then index inside the kernel
    int64_t ms = support[0];
    const int64_t n_sup = n_cfb * n_cfc * n_cfp * 2;
    for (int64_t i = 1; i < n_sup; i++) ms = support[i] > ms ? support[i] : ms;
    for (int64_t i = 0; i < n_sup; i += 2) {
        const int64_t su = support[i], sv = support[i + 1];
        NGC_CHECK(su >= 1 && sv >= 1, NGC_ERR_ARG, "support must be >= 1");
        const int64_t reach_u = os_u * (su - su / 2 - 1) + os_u / 2 + 1;
        const int64_t reach_v = os_v * (sv - sv / 2 - 1) + os_v / 2 + 1;
        const int64_t low_u = os_u * (su / 2) + os_u / 2 + 1;
        const int64_t low_v = os_v * (sv / 2) + os_v / 2 + 1;
        NGC_CHECK(k_u / 2 + reach_u < k_u && k_u / 2 - low_u >= 0 && k_v / 2 + reach_v < k_v &&
                      k_v / 2 - low_v >= 0,
                  NGC_ERR_ARG, "support x oversampling exceeds the kernel size");
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    std::vector<double> scale;
    uv_scale_host(freq_chan, n_chan, delta_l, delta_m, n_u, n_v, scale);
    void *d_scale, *d_cmap, *d_pmap, *d_cfb, *d_cfc, *d_cfp, *d_sup, *d_fid;
    NGC_TRY(upload(ctx, s, "uv_scale", scale.data(), scale.size() * sizeof(double), &d_scale));
    NGC_TRY(upload(ctx, s, "chan_map", padded_map(chan_map, n_chan).data(), n_chan * sizeof(int32_t) + 4, &d_cmap));
    NGC_TRY(upload(ctx, s, "pol_map", pol_map, n_pol * sizeof(int32_t), &d_pmap));
    NGC_TRY(upload(ctx, s, "cf_bl_map", padded_map(cf_baseline_map, n_baseline).data(),
                   n_baseline * sizeof(int32_t) + 4, &d_cfb));
    NGC_TRY(upload(ctx, s, "cf_chan_map", padded_map(cf_chan_map, n_chan).data(),
                   n_chan * sizeof(int32_t) + 4, &d_cfc));
    NGC_TRY(upload(ctx, s, "cf_pol_map", cf_pol_map, n_pol * sizeof(int32_t), &d_cfp));
    NGC_TRY(upload(ctx, s, "support", support, n_sup * sizeof(int32_t), &d_sup));
    NGC_TRY(upload(ctx, s, "field_id", field_id, n_field * sizeof(int32_t), &d_fid));
    const int64_t cells = n_imag_chan * n_imag_pol * n_u * n_v;
    NGC_HIP(hipMemsetAsync(ctx->d_error, 0, sizeof(int), s));
    if (precision != NGC_PREC_FP64) {
        bool pol_identity = n_imag_pol == n_pol, cfp_uniform = true;
        for (int64_t p = 0; p < n_pol; p++) {
            pol_identity &= pol_map[p] == p;
            cfp_uniform &= cf_pol_map[p] == cf_pol_map[0];
        }
        ApTiledCall c{};
        c.mode = mode;
        c.precision = precision;
        c.uvw = uvw;
        c.vis = vis;
        c.vis_dtype = vis_dtype;
        c.weight = weight;
        c.weight_dtype = weight_dtype;
        c.flag = flag;
        c.field = field;
        c.d_field_id = (const int *)d_fid;
        c.n_field = n_field;
        c.d_uv_scale = (const double *)d_scale;
        c.d_chan_map = (const int *)d_cmap;
        c.d_cfb = (const int *)d_cfb;
        c.d_cfc = (const int *)d_cfc;
        c.d_support = (const int *)d_sup;
        c.n_cfc = n_cfc;
        c.n_cfp = n_cfp;
        c.cfp0 = cf_pol_map[0];
        c.n_cf = n_cfb * n_cfc * n_cfp;
        c.conv_kernel = conv_kernel;
        c.phase_gradient = phase_gradient;
        c.k_u = k_u;
        c.k_v = k_v;
        c.n_time = n_time;
        c.n_baseline = n_baseline;
        c.n_chan = n_chan;
        c.n_pol = n_pol;
        c.n_imag_chan = n_imag_chan;
        c.n_imag_pol = n_imag_pol;
        c.n_u = n_u;
        c.n_v = n_v;
        c.os_u = os_u;
        c.os_v = os_v;
        c.max_support = ms;
        c.pol_identity = pol_identity;
        c.cfp_uniform = cfp_uniform;
        c.accumulate = accumulate;
        c.grid = grid;
        c.sum_weight = sum_weight;
        profile_new_call(ctx, s);
        NGC_TRY(ap_tiled_run(ctx, s, c));
        int err = 0;
        NGC_HIP(hipMemcpyAsync(&err, ctx->d_error, sizeof(int), hipMemcpyDeviceToHost, s));
        NGC_HIP(hipStreamSynchronize(s));
        NGC_CHECK(err == 0, NGC_ERR_FIELD, "a row's field id is not in field_id");
        return NGC_OK;
    }
    if (!accumulate) {
        NGC_HIP(hipMemsetAsync(grid, 0, cells * 2 * sizeof(double), s));
        NGC_HIP(hipMemsetAsync(sum_weight, 0, n_imag_chan * n_imag_pol * sizeof(double), s));
    }
    ApArgs a{};
    a.uvw = uvw;
    a.vis = vis;
    a.weight = weight;
    a.flag = flag;
    a.field = field;
    a.field_id = (const int *)d_fid;
    a.uv_scale = (const double *)d_scale;
    a.chan_map = (const int *)d_cmap;
    a.pol_map = (const int *)d_pmap;
    a.cfb = (const int *)d_cfb;
    a.cfc = (const int *)d_cfc;
    a.cfp = (const int *)d_cfp;
    a.ck = conv_kernel;
    a.support = (const int *)d_sup;
    a.pg = (const double2 *)phase_gradient;
    a.n_field = n_field;
    a.n_cfb = n_cfb;
    a.n_cfc = n_cfc;
    a.n_cfp = n_cfp;
    a.k_u = k_u;
    a.k_v = k_v;
    a.n_time = n_time;
    a.n_baseline = n_baseline;
    a.n_chan = n_chan;
    a.n_pol = n_pol;
    a.n_imag_chan = n_imag_chan;
    a.n_imag_pol = n_imag_pol;
    a.n_u = n_u;
    a.n_v = n_v;
    a.os_u = os_u;
    a.os_v = os_v;
    a.max_support = ms;
    a.vis_dtype = vis_dtype;
    a.weight_dtype = weight_dtype;
    a.error = ctx->d_error;
    const int64_t n_rc = n_time * n_baseline * n_chan;
    if (n_rc > 0) {
        int64_t nb = ceil_div(n_rc, 256);
        if (nb > (1 << 20)) nb = 1 << 20;
        double *g = (double *)grid;
        if (mode == NGC_GRID_IMAGE)
            ap_grid_plain_kernel<NGC_GRID_IMAGE><<<(int)nb, 256, 0, s>>>(a, g, sum_weight);
        else if (mode == NGC_GRID_PSF)
            ap_grid_plain_kernel<NGC_GRID_PSF><<<(int)nb, 256, 0, s>>>(a, g, sum_weight);
        else
            ap_grid_plain_kernel<NGC_GRID_WEIGHT><<<(int)nb, 256, 0, s>>>(a, g, sum_weight);
        NGC_HIP(hipGetLastError());
    }
    int err = 0;
    NGC_HIP(hipMemcpyAsync(&err, ctx->d_error, sizeof(int), hipMemcpyDeviceToHost, s));
    NGC_HIP(hipStreamSynchronize(s));
    NGC_CHECK(err == 0, NGC_ERR_FIELD, "a row's field id is not in field_id");
    return NGC_OK;
}
