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
#include <algorithm>
#include <functional>
#include <vector>

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

namespace {

// The arguments of ngc_aperture_grid.
struct ApHost {
    int mode, precision;
    const double *uvw;
    const void *vis;
    int vis_dtype;
    const void *weight;
    int weight_dtype;
    const uint8_t *flag;
    const int32_t *field, *field_id;
    int64_t n_field;
    const double *freq_chan;
    const int32_t *chan_map, *pol_map, *cf_baseline_map, *cf_chan_map, *cf_pol_map;
    const double *conv_kernel;
    int64_t n_cfb, n_cfc, n_cfp, k_u, k_v;
    const int32_t *support;
    const double *phase_gradient;
    int64_t n_time, n_baseline, n_chan, n_pol, n_imag_chan, n_imag_pol, n_u, n_v;
    double delta_l, delta_m;
    int64_t os_u, os_v;
    int accumulate;
    void *grid;
    double *sum_weight;
};

// Device copies of the small host tables.
struct ApDevTables {
    void *scale, *cmap, *pmap, *cfb, *cfc, *cfp, *sup, *fid, *fmap;
    int64_t fmap_len;
};

// Argument checks; *ms = max(SUPPORT), the half-width of the reference's
// bounds check (_aperture_grid.py:397,447), with which every tap must index
// inside the kernel.
int ap_validate(const ApHost &h, int64_t *ms_out) {
    NGC_CHECK(h.mode == NGC_GRID_IMAGE || h.mode == NGC_GRID_PSF || h.mode == NGC_GRID_WEIGHT,
              NGC_ERR_ARG, "mode must be IMAGE, PSF or WEIGHT");
    NGC_CHECK(h.precision >= NGC_PREC_FP64 && h.precision <= NGC_PREC_FP32, NGC_ERR_ARG,
              "unknown precision");
    const int64_t n_samples = h.n_time * h.n_baseline * h.n_chan;
    NGC_CHECK(h.field_id && h.freq_chan && h.chan_map && h.pol_map && h.cf_baseline_map &&
                  h.cf_chan_map && h.cf_pol_map && h.conv_kernel && h.support &&
                  h.phase_gradient && h.grid && h.sum_weight,
              NGC_ERR_ARG, "NULL array");
    NGC_CHECK((h.uvw && h.weight && h.field) || n_samples == 0, NGC_ERR_ARG,
              "NULL uvw, weight or field");
    NGC_CHECK(h.mode != NGC_GRID_IMAGE || h.vis || n_samples == 0, NGC_ERR_ARG,
              "image mode needs vis");
    NGC_CHECK(h.n_field >= 1 && h.n_cfb >= 1 && h.n_cfc >= 1 && h.n_cfp >= 1 && h.k_u >= 1 &&
                  h.k_v >= 1,
              NGC_ERR_ARG, "bad kernel dimensions");
    NGC_CHECK(h.os_u >= 1 && h.os_v >= 1, NGC_ERR_ARG, "oversampling must be >= 1");
    for (int64_t b = 0; b < h.n_baseline; b++)
        NGC_CHECK(h.cf_baseline_map[b] >= 0 && h.cf_baseline_map[b] < h.n_cfb, NGC_ERR_ARG,
                  "cf_baseline_map range");
    for (int64_t c = 0; c < h.n_chan; c++) {
        NGC_CHECK(h.cf_chan_map[c] >= 0 && h.cf_chan_map[c] < h.n_cfc, NGC_ERR_ARG,
                  "cf_chan_map range");
        NGC_CHECK(h.chan_map[c] >= 0 && h.chan_map[c] < h.n_imag_chan, NGC_ERR_ARG,
                  "chan_map range");
    }
    for (int64_t p = 0; p < h.n_pol; p++) {
        NGC_CHECK(h.cf_pol_map[p] >= 0 && h.cf_pol_map[p] < h.n_cfp, NGC_ERR_ARG,
                  "cf_pol_map range");
        NGC_CHECK(h.pol_map[p] >= 0 && h.pol_map[p] < h.n_imag_pol, NGC_ERR_ARG,
                  "pol_map range");
    }
    int64_t ms = h.support[0];
    const int64_t n_sup = h.n_cfb * h.n_cfc * h.n_cfp * 2;
    for (int64_t i = 1; i < n_sup; i++) ms = h.support[i] > ms ? h.support[i] : ms;
    for (int64_t i = 0; i < n_sup; i += 2) {
        const int64_t su = h.support[i], sv = h.support[i + 1];
        NGC_CHECK(su >= 1 && sv >= 1, NGC_ERR_ARG, "support must be >= 1");
        const int64_t reach_u = h.os_u * (su - su / 2 - 1) + h.os_u / 2 + 1;
        const int64_t reach_v = h.os_v * (sv - sv / 2 - 1) + h.os_v / 2 + 1;
        const int64_t low_u = h.os_u * (su / 2) + h.os_u / 2 + 1;
        const int64_t low_v = h.os_v * (sv / 2) + h.os_v / 2 + 1;
        NGC_CHECK(h.k_u / 2 + reach_u < h.k_u && h.k_u / 2 - low_u >= 0 &&
                      h.k_v / 2 + reach_v < h.k_v && h.k_v / 2 - low_v >= 0,
                  NGC_ERR_ARG, "support x oversampling exceeds the kernel size");
    }
    *ms_out = ms;
    return NGC_OK;
}

using Upload = std::function<int(const char *, const void *, size_t, void **)>;

int ap_upload_tables(const ApHost &h, const Upload &up, ApDevTables &d) {
    std::vector<double> scale;
    uv_scale_host(h.freq_chan, h.n_chan, h.delta_l, h.delta_m, h.n_u, h.n_v, scale);
    const int64_t n_sup = h.n_cfb * h.n_cfc * h.n_cfp * 2;
    NGC_TRY(up("uv_scale", scale.data(), scale.size() * sizeof(double), &d.scale));
    NGC_TRY(up("chan_map", padded_map(h.chan_map, h.n_chan).data(),
               h.n_chan * sizeof(int32_t) + 4, &d.cmap));
    NGC_TRY(up("pol_map", h.pol_map, h.n_pol * sizeof(int32_t), &d.pmap));
    NGC_TRY(up("cf_bl_map", padded_map(h.cf_baseline_map, h.n_baseline).data(),
               h.n_baseline * sizeof(int32_t) + 4, &d.cfb));
    NGC_TRY(up("cf_chan_map", padded_map(h.cf_chan_map, h.n_chan).data(),
               h.n_chan * sizeof(int32_t) + 4, &d.cfc));
    NGC_TRY(up("cf_pol_map", h.cf_pol_map, h.n_pol * sizeof(int32_t), &d.cfp));
    NGC_TRY(up("support", h.support, n_sup * sizeof(int32_t), &d.sup));
    NGC_TRY(up("field_id", h.field_id, h.n_field * sizeof(int32_t), &d.fid));
    // field id -> its first index in field_id (-1 if absent), when the ids are
    // small non-negative integers; otherwise the kernels search field_id
    int64_t id_max = -1;
    bool small = true;
    for (int64_t f = 0; f < h.n_field; f++) {
        small &= h.field_id[f] >= 0 && h.field_id[f] < (1 << 20);
        id_max = std::max<int64_t>(id_max, h.field_id[f]);
    }
    d.fmap = nullptr;
    d.fmap_len = 0;
    if (small && id_max >= 0) {
        std::vector<int32_t> map((size_t)id_max + 1, -1);
        for (int64_t f = h.n_field - 1; f >= 0; f--) map[(size_t)h.field_id[f]] = (int32_t)f;
        d.fmap_len = id_max + 1;
        NGC_TRY(up("field_map", map.data(), map.size() * sizeof(int32_t), &d.fmap));
    }
    return NGC_OK;
}

ApTiledCall ap_tiled_call(const ApHost &h, const ApDevTables &d, int64_t ms) {
    bool pol_identity = h.n_imag_pol == h.n_pol, cfp_uniform = true;
    for (int64_t p = 0; p < h.n_pol; p++) {
        pol_identity &= h.pol_map[p] == p;
        cfp_uniform &= h.cf_pol_map[p] == h.cf_pol_map[0];
    }
    ApTiledCall c{};
    c.mode = h.mode;
    c.precision = h.precision;
    c.uvw = h.uvw;
    c.vis = h.vis;
    c.vis_dtype = h.vis_dtype;
    c.weight = h.weight;
    c.weight_dtype = h.weight_dtype;
    c.flag = h.flag;
    c.field = h.field;
    c.d_field_id = (const int *)d.fid;
    c.d_field_map = (const int *)d.fmap;
    c.field_map_len = d.fmap_len;
    c.n_field = h.n_field;
    c.d_uv_scale = (const double *)d.scale;
    c.d_chan_map = (const int *)d.cmap;
    c.d_cfb = (const int *)d.cfb;
    c.d_cfc = (const int *)d.cfc;
    c.d_support = (const int *)d.sup;
    c.support_host = h.support;
    c.n_cfc = h.n_cfc;
    c.n_cfp = h.n_cfp;
    c.cfp0 = h.cf_pol_map[0];
    c.n_cf = h.n_cfb * h.n_cfc * h.n_cfp;
    c.conv_kernel = h.conv_kernel;
    c.phase_gradient = h.phase_gradient;
    c.k_u = h.k_u;
    c.k_v = h.k_v;
    c.n_time = h.n_time;
    c.n_baseline = h.n_baseline;
    c.n_chan = h.n_chan;
    c.n_pol = h.n_pol;
    c.n_imag_chan = h.n_imag_chan;
    c.n_imag_pol = h.n_imag_pol;
    c.n_u = h.n_u;
    c.n_v = h.n_v;
    c.os_u = h.os_u;
    c.os_v = h.os_v;
    c.max_support = ms;
    c.pol_identity = pol_identity;
    c.cfp_uniform = cfp_uniform;
    c.accumulate = h.accumulate;
    c.grid = h.grid;
    c.sum_weight = h.sum_weight;
    return c;
}

}  // namespace

#define NGC_AP_HOST_ARGS                                                                      \
    ApHost h{mode,        precision,   uvw,          vis,         vis_dtype,    weight,        \
             weight_dtype, flag,       field,        field_id,    n_field,      freq_chan,     \
             chan_map,    pol_map,     cf_baseline_map, cf_chan_map, cf_pol_map, conv_kernel, \
             n_cfb,       n_cfc,       n_cfp,        k_u,         k_v,          support,       \
             phase_gradient, n_time,   n_baseline,   n_chan,      n_pol,        n_imag_chan,   \
             n_imag_pol,  n_u,         n_v,          delta_l,     delta_m,      os_u,          \
             os_v,        accumulate,  grid,         sum_weight}

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
    const NGC_AP_HOST_ARGS;
    int64_t ms = 0;
    NGC_TRY(ap_validate(h, &ms));
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    ApDevTables d{};
    NGC_TRY(ap_upload_tables(h, [&](const char *name, const void *host, size_t bytes, void **dev) {
        return upload(ctx, s, name, host, bytes, dev);
    }, d));
    const int64_t cells = n_imag_chan * n_imag_pol * n_u * n_v;
    NGC_HIP(hipMemsetAsync(ctx->d_error, 0, sizeof(int), s));
    if (precision != NGC_PREC_FP64) {
        const ApTiledCall c = ap_tiled_call(h, d, ms);
        profile_new_call(ctx, s);
        NGC_TRY(ap_tiled_run(ctx, s, c));
        int err = 0;
        NGC_HIP(hipMemcpyAsync(&err, ctx->d_error, sizeof(int), hipMemcpyDeviceToHost, s));
        NGC_HIP(hipStreamSynchronize(s));
        NGC_CHECK(err == 0, NGC_ERR_FIELD, "a row's field id is not in field_id");
        return NGC_OK;
    }
    void *d_scale = d.scale, *d_cmap = d.cmap, *d_pmap = d.pmap, *d_cfb = d.cfb, *d_cfc = d.cfc,
         *d_cfp = d.cfp, *d_sup = d.sup, *d_fid = d.fid;
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

extern "C" int ngc_aperture_grid_plan_create(
    ngc_context *ctx, void *stream, int mode, int precision, const double *uvw,
    const void *vis, int vis_dtype, const void *weight, int weight_dtype, const uint8_t *flag,
    const int32_t *field, const int32_t *field_id, int64_t n_field, const double *freq_chan,
    const int32_t *chan_map, const int32_t *pol_map, const int32_t *cf_baseline_map,
    const int32_t *cf_chan_map, const int32_t *cf_pol_map, const double *conv_kernel,
    int64_t n_cfb, int64_t n_cfc, int64_t n_cfp, int64_t k_u, int64_t k_v,
    const int32_t *support, const double *phase_gradient, int64_t n_time, int64_t n_baseline,
    int64_t n_chan, int64_t n_pol, int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
    int64_t n_v, double delta_l, double delta_m, int64_t os_u, int64_t os_v, int accumulate,
    void *grid, double *sum_weight, int autotune, ngc_grid_plan **plan) {
    NGC_CHECK(ctx && plan, NGC_ERR_ARG, "ctx or plan is NULL");
    const NGC_AP_HOST_ARGS;
    int64_t ms = 0;
    NGC_TRY(ap_validate(h, &ms));
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    ngc_grid_plan *p = new ngc_grid_plan();
    p->ctx = ctx;
    ApDevTables d{};
    int rc = ap_upload_tables(h, [&](const char *, const void *host, size_t bytes, void **dev) {
        return plan_upload(p, s, host, bytes, dev);
    }, d);
    if (rc == NGC_OK) {
        const ApTiledCall c = ap_tiled_call(h, d, ms);
        // the autotuner's eager runs report field errors in the context's flag
        (void)hipMemsetAsync(ctx->d_error, 0, sizeof(int), s);
        rc = ap_plan_create(ctx, s, c, autotune, p);
    }
    if (rc != NGC_OK) {
        (void)hipStreamSynchronize(s);
        plan_free(p);
        return rc;
    }
    *plan = p;
    return NGC_OK;
}
