// Grid -> image (rocFFT), Briggs factors, imaging-weight degridder and mosaic
// normalisation for gfx950.
//
//   ngc_grid_to_image        make_psf.py:117-131, make_image.py:116-130,
//                            make_mosaic_pb.py:119-129, _remove_padding.py:20-30
//   ngc_briggs_factors       make_imaging_weight.py:198-216
//   ngc_imaging_weight_degrid _standard_grid.py:467-518
//   ngc_normalize_mosaic     _normalize.py:39-89
#include <rocfft/rocfft.h>

#include "ngc_internal.hpp"

namespace ngc {

struct FftPlan {
    rocfft_plan plan = nullptr;
    rocfft_execution_info info = nullptr;
    size_t work_bytes = 0;
};

static std::once_flag g_rocfft_once;
static rocfft_status g_rocfft_setup = rocfft_status_success;

int fft_destroy_plans(ngc_context *ctx) {
    for (auto &kv : ctx->plans) {
        if (kv.second->info) rocfft_execution_info_destroy(kv.second->info);
        if (kv.second->plan) rocfft_plan_destroy(kv.second->plan);
        delete kv.second;
    }
    ctx->plans.clear();
    return NGC_OK;
}

static int get_plan(ngc_context *ctx, int64_t n_u, int64_t n_v, int64_t batch, bool dbl,
                    FftPlan **out, bool forward = false) {
    std::call_once(g_rocfft_once, [] { g_rocfft_setup = rocfft_setup(); });
    NGC_CHECK(g_rocfft_setup == rocfft_status_success, NGC_ERR_FFT, "rocfft_setup failed");
    auto key = std::make_tuple(n_u, n_v, batch, (dbl ? 1 : 0) + (forward ? 2 : 0));
    auto it = ctx->plans.find(key);
    if (it != ctx->plans.end()) {
        *out = it->second;
        return NGC_OK;
    }
    FftPlan *p = new FftPlan();
    size_t lengths[2] = {(size_t)n_v, (size_t)n_u};  // fastest axis first
    rocfft_status st = rocfft_plan_create(
        &p->plan, rocfft_placement_inplace,
        forward ? rocfft_transform_type_complex_forward : rocfft_transform_type_complex_inverse,
        dbl ? rocfft_precision_double : rocfft_precision_single, 2, lengths, (size_t)batch,
        nullptr);
    if (st != rocfft_status_success) {
        delete p;
        set_error("rocfft_plan_create failed (status " + std::to_string((int)st) + ")");
        return NGC_ERR_FFT;
    }
    rocfft_plan_get_work_buffer_size(p->plan, &p->work_bytes);
    rocfft_execution_info_create(&p->info);
    ctx->plans[key] = p;
    *out = p;
    return NGC_OK;
}

// A[ku][kv] = grid[(ku + n_u//2) % n_u][(kv + n_v//2) % n_v]  (ifftshift)
template <typename IN, typename OUT, int NC>
__global__ void ifftshift_kernel(const IN *grid, OUT *buf, int64_t n_planes, int64_t n_u,
                                 int64_t n_v) {
    const int64_t plane = n_u * n_v;
    const int64_t n = n_planes * plane;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pl = i / plane;
        const int64_t r = i - pl * plane;
        const int64_t ku = r / n_v, kv = r - ku * n_v;
        int64_t su = ku + n_u / 2;
        if (su >= n_u) su -= n_u;
        int64_t sv = kv + n_v / 2;
        if (sv >= n_v) sv -= n_v;
        const int64_t src = pl * plane + su * n_v + sv;
        OUT re = (OUT)grid[NC * src];
        OUT im = NC == 2 ? (OUT)grid[NC * src + NC - 1] : (OUT)0;
        buf[2 * i] = re;
        buf[2 * i + 1] = im;
    }
}

// image[pl][l][m] = ((Re B[(l - n_l//2) mod n_u][(m - n_m//2) mod n_v]) / sw') /
//                   (corr_l[l]*corr_m[m])   with B the unnormalised inverse FFT
template <typename T>
__global__ void crop_correct_kernel(const T *buf, int64_t n_planes, int64_t n_u,
                                    int64_t n_v, int64_t n_l, int64_t n_m,
                                    const double *sum_weight, const double *corr,
                                    double *image) {
#pragma clang fp contract(off)
    const int64_t n = n_planes * n_l * n_m;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pl = i / (n_l * n_m);
        const int64_t r = i - pl * n_l * n_m;
        const int64_t l = r / n_m, m = r - l * n_m;
        int64_t ku = l - n_l / 2;
        if (ku < 0) ku += n_u;
        int64_t kv = m - n_m / 2;
        if (kv < 0) kv += n_v;
        double x = (double)buf[2 * (pl * n_u * n_v + ku * n_v + kv)];
        if (sum_weight) {
            double sw = sum_weight[pl];
            if (sw == 0.0) sw = 1.0;
            x = x / sw;
        }
        if (corr) x = x / (corr[l] * corr[n_l + m]);
        image[i] = x;
    }
}

// In-place batched 2-D complex FFT of `buf` ([batch, n_u, n_v], interleaved),
// unnormalised, forward or inverse.
int fft_execute(ngc_context *ctx, hipStream_t s, void *buf, int64_t n_u, int64_t n_v,
                int64_t batch, bool dbl, bool forward) {
    FftPlan *plan;
    NGC_TRY(get_plan(ctx, n_u, n_v, batch, dbl, &plan, forward));
    void *work = nullptr;
    if (plan->work_bytes) NGC_TRY(workspace(ctx, "fft_work", plan->work_bytes, &work));
    rocfft_execution_info_set_stream(plan->info, s);
    if (plan->work_bytes) rocfft_execution_info_set_work_buffer(plan->info, work, plan->work_bytes);
    void *bufs[1] = {buf};
    rocfft_status st = rocfft_execute(plan->plan, bufs, nullptr, plan->info);
    NGC_CHECK(st == rocfft_status_success, NGC_ERR_FFT, "rocfft_execute failed");
    return NGC_OK;
}

static int blocks_for(int64_t n) {
    int64_t b = ceil_div(n, 256);
    if (b > (1 << 20)) b = 1 << 20;
    return (int)(b < 1 ? 1 : b);
}

static int grid_to_image_impl(ngc_context *ctx, hipStream_t s, const void *grid, int dt,
                              int64_t n_planes, int64_t n_u, int64_t n_v, int64_t n_l,
                              int64_t n_m, const double *sum_weight, const double *corr_l,
                              const double *corr_m, double *image) {
    NGC_CHECK(grid && image, NGC_ERR_ARG, "NULL grid or image");
    NGC_CHECK(n_planes >= 1 && n_u >= 1 && n_v >= 1 && n_l >= 1 && n_m >= 1 && n_l <= n_u &&
                  n_m <= n_v,
              NGC_ERR_ARG, "bad image dimensions");
    NGC_CHECK(dt == NGC_C64 || dt == NGC_C128 || dt == NGC_F32 || dt == NGC_F64, NGC_ERR_ARG,
              "grid dtype");
    NGC_CHECK((corr_l == nullptr) == (corr_m == nullptr), NGC_ERR_ARG,
              "corr_l and corr_m must both be given or both NULL");
    const bool dbl = (dt == NGC_C128 || dt == NGC_F64);
    const int64_t cells = n_planes * n_u * n_v;
    const size_t elem = dbl ? 2 * sizeof(double) : 2 * sizeof(float);
    double *d_corr = nullptr;
    if (corr_l) {
        std::vector<double> c(n_l + n_m);
        for (int64_t i = 0; i < n_l; i++) c[i] = corr_l[i];
        for (int64_t i = 0; i < n_m; i++) c[n_l + i] = corr_m[i];
        void *p;
        NGC_TRY(upload(ctx, s, "corr", c.data(), c.size() * sizeof(double), &p));
        d_corr = (double *)p;
    }
    profile_new_call(ctx, s);
    // single-precision grids of smooth lengths: the LDS transform (shift,
    // rows, transpose, rows, crop); otherwise rocFFT
    if (!dbl && ctx->fft_lds && lds_fft_supported(n_u) && lds_fft_supported(n_v))
        return lds_grid_to_image(ctx, s, grid, dt, n_planes, n_u, n_v, n_l, n_m, sum_weight,
                                 d_corr, image);
    void *buf;
    NGC_TRY(workspace(ctx, "fft_buffer", cells * elem, &buf));
    const int nb = blocks_for(cells);
    phase_begin(ctx, s, "ifftshift");
    switch (dt) {
    case NGC_C64:
        ifftshift_kernel<float, float, 2><<<nb, 256, 0, s>>>((const float *)grid, (float *)buf,
                                                            n_planes, n_u, n_v);
        break;
    case NGC_F32:
        ifftshift_kernel<float, float, 1><<<nb, 256, 0, s>>>((const float *)grid, (float *)buf,
                                                            n_planes, n_u, n_v);
        break;
    case NGC_C128:
        ifftshift_kernel<double, double, 2><<<nb, 256, 0, s>>>(
            (const double *)grid, (double *)buf, n_planes, n_u, n_v);
        break;
    default:
        ifftshift_kernel<double, double, 1><<<nb, 256, 0, s>>>(
            (const double *)grid, (double *)buf, n_planes, n_u, n_v);
        break;
    }
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    FftPlan *plan;
    NGC_TRY(get_plan(ctx, n_u, n_v, n_planes, dbl, &plan));
    void *work = nullptr;
    if (plan->work_bytes) NGC_TRY(workspace(ctx, "fft_work", plan->work_bytes, &work));
    rocfft_execution_info_set_stream(plan->info, s);
    if (plan->work_bytes) rocfft_execution_info_set_work_buffer(plan->info, work, plan->work_bytes);
    void *bufs[1] = {buf};
    phase_begin(ctx, s, "rocfft_inverse");
    rocfft_status st = rocfft_execute(plan->plan, bufs, nullptr, plan->info);
    NGC_CHECK(st == rocfft_status_success, NGC_ERR_FFT, "rocfft_execute failed");
    phase_end(ctx, s);
    const int nb2 = blocks_for(n_planes * n_l * n_m);
    phase_begin(ctx, s, "crop_correct");
    if (dbl)
        crop_correct_kernel<double><<<nb2, 256, 0, s>>>((const double *)buf, n_planes, n_u, n_v,
                                                        n_l, n_m, sum_weight, d_corr, image);
    else
        crop_correct_kernel<float><<<nb2, 256, 0, s>>>((const float *)buf, n_planes, n_u, n_v,
                                                       n_l, n_m, sum_weight, d_corr, image);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}

// ---------------------------------------------------------------------------
// Briggs factors: sum over (u, v) of density^2 per plane, FP64
// ---------------------------------------------------------------------------
__global__ void sumsq_kernel(const double *g, int64_t plane, double *out) {
    const int64_t pl = blockIdx.y;
    const double *p = g + pl * plane;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < plane;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double x = p[i];
        acc += x * x;
    }
    acc = wave_sum(acc);
    __shared__ double s[4];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&out[pl], s[0] + s[1] + s[2] + s[3]);
}

// ---------------------------------------------------------------------------
// imaging-weight degridder (_standard_grid.py:467-518)
// ---------------------------------------------------------------------------
struct DegridArgs {
    const double *uvw;
    const void *nat;
    const uint8_t *flag;
    const double *uv_scale;
    const int *chan_map, *pol_map;
    const double *grid;    // [n_imag_chan, n_imag_pol, n_u, n_v]
    const double *factors; // [2, n_imag_chan, n_imag_pol]
    int64_t n_time, n_baseline, n_chan, n_pol, n_imag_chan, n_imag_pol, n_u, n_v;
    int weight_dtype, out_dtype;
    void *out;
};

__device__ inline double nat_w(const DegridArgs &a, int64_t i) {
    double w = a.weight_dtype == NGC_F32 ? (double)((const float *)a.nat)[i]
                                         : ((const double *)a.nat)[i];
    if (a.flag && a.flag[i]) w = __longlong_as_double(0x7ff8000000000000LL);
    return w;
}

// A wavefront takes 8 consecutive integrations x 8 consecutive channels of
// one baseline: the weights it reads and the imaging weights it writes are
// runs of 8 channels, and the density cells of neighbouring integrations in
// one channel's plane share cache lines (with the channel fastest, every lane
// read a different plane).
__global__ void degrid_kernel(DegridArgs a) {
#pragma clang fp contract(off)
    const int64_t n_cb = (a.n_chan + 7) / 8, n_tb = (a.n_time + 7) / 8;
    const int64_t n_thr = a.n_baseline * n_tb * n_cb * 64;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n_thr;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int lane = (int)(j & 63);
        const int64_t w = j >> 6;
        const int64_t cb = w % n_cb, r = w / n_cb;
        const int64_t tb = r % n_tb, bl = r / n_tb;
        const int64_t ch = cb * 8 + (lane & 7), t = tb * 8 + (lane >> 3);
        if (ch >= a.n_chan || t >= a.n_time) continue;
        const int64_t row = t * a.n_baseline + bl;
        const int64_t i = row * a.n_chan + ch;
        const int64_t base = i * a.n_pol;
        Pix px;
        bool inb = pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                                  a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, 0, px);
        inb = inb && (px.uc < a.n_u) && (px.vc < a.n_v) && (px.uc >= 0) && (px.vc >= 0);
        const int64_t a_chan = a.chan_map[ch];
        for (int p = 0; p < a.n_pol; p++) {
            double out = 0.0;  // imaging_weight starts at zero (_standard_grid.py:460)
            if (inb) {
                const int64_t a_pol = a.pol_map[p];
                if (a.n_pol == 2)
                    out = (nat_w(a, base) + nat_w(a, base + 1)) / 2.0;
                else
                    out = nat_w(a, base + p);
                const double nw = nat_w(a, base + p);
                if (!isnan(nw) && nw != 0.0) {
                    const double g =
                        a.grid[((a_chan * a.n_imag_pol + a_pol) * a.n_u + px.uc) * a.n_v + px.vc];
                    if (!isnan(g) && g != 0.0) {
                        const int64_t fi = a_chan * a.n_imag_pol + a_pol;
                        const double bg =
                            a.factors[fi] * g + a.factors[a.n_imag_chan * a.n_imag_pol + fi];
                        out = out / bg;
                    }
                }
            }
            if (a.out_dtype == NGC_F32)
                ((float *)a.out)[base + p] = (float)out;
            else
                ((double *)a.out)[base + p] = out;
        }
    }
}

// ---------------------------------------------------------------------------
// mosaic normalisation (_normalize.py:39-89)
// ---------------------------------------------------------------------------
__device__ inline double sinc_np(double x) {
    // numpy.sinc: sin(pi x)/(pi x), 1 at 0
    if (x == 0.0) return 1.0;
    const double y = M_PI * x;
    return sin(y) / y;
}

__global__ void normalize_kernel(double *image, const double *sum_weight, const double *ps_corr,
                                 const double *norm_pb, const double *pb, int64_t n_planes,
                                 int64_t n_l, int64_t n_m, int64_t os_u, int64_t os_v,
                                 double pb_limit, int single_precision) {
#pragma clang fp contract(off)
    const int64_t n = n_planes * n_l * n_m;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pl = i / (n_l * n_m);
        const int64_t r = i - pl * n_l * n_m;
        const int64_t l = r / n_m, m = r - l * n_m;
        double sw = sum_weight[pl];
        if (sw == 0.0) sw = 1.0;
        const double sx = sinc_np((double)(l - n_l / 2) / (double)(n_l * os_u));
        const double sy = sinc_np((double)(m - n_m / 2) / (double)(n_m * os_v));
        const double osc = sx * sy;
        const double ps = ps_corr ? ps_corr[l * n_m + m] : 1.0;
        const double nimg = norm_pb ? ps * norm_pb[i] : ps;
        double x = (image[i] / sw) / (osc * nimg);
        if (pb_limit > 0 && pb && pb[i] < pb_limit) x = 0.0;
        if (single_precision) x = (double)(float)x;
        image[i] = x;
    }
}

}  // namespace ngc

using namespace ngc;

extern "C" {

int ngc_grid_to_image(ngc_context *ctx, void *stream, const void *grid, int grid_dtype,
                      int64_t n_planes, int64_t n_u, int64_t n_v, int64_t n_l, int64_t n_m,
                      const double *sum_weight, const double *corr_l, const double *corr_m,
                      double *image) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    return grid_to_image_impl(ctx, s, grid, grid_dtype, n_planes, n_u, n_v, n_l, n_m,
                              sum_weight, corr_l, corr_m, image);
}

int ngc_set_fft_mode(ngc_context *ctx, int mode) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(mode == 0 || mode == 1, NGC_ERR_ARG, "mode must be 0 (rocFFT) or 1 (LDS)");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->fft_lds = mode;
    return NGC_OK;
}

int ngc_briggs_factors(ngc_context *ctx, void *stream, const double *density_grid,
                       const double *sum_weight, int64_t n_imag_chan, int64_t n_imag_pol,
                       int64_t n_u, int64_t n_v, int weighting, double robust,
                       double *factors) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(density_grid && sum_weight && factors, NGC_ERR_ARG, "NULL array");
    NGC_CHECK(weighting == NGC_WEIGHT_UNIFORM || weighting == NGC_WEIGHT_BRIGGS, NGC_ERR_ARG,
              "weighting must be NGC_WEIGHT_UNIFORM or NGC_WEIGHT_BRIGGS");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    const int64_t np = n_imag_chan * n_imag_pol;
    std::vector<double> sq(np, 0.0), sw(np, 0.0);
    if (weighting == NGC_WEIGHT_BRIGGS) {
        void *d_sq;
        NGC_TRY(workspace(ctx, "sumsq", np * sizeof(double), &d_sq));
        NGC_HIP(hipMemsetAsync(d_sq, 0, np * sizeof(double), s));
        const int64_t plane = n_u * n_v;
        int bx = (int)ceil_div(plane, 256 * 16);
        if (bx < 1) bx = 1;
        if (bx > 1024) bx = 1024;
        dim3 grid(bx, (unsigned)np);
        profile_new_call(ctx, s);
        phase_begin(ctx, s, "briggs_sumsq");
        sumsq_kernel<<<grid, 256, 0, s>>>(density_grid, plane, (double *)d_sq);
        NGC_HIP(hipGetLastError());
        phase_end(ctx, s);
        NGC_HIP(hipMemcpyAsync(sq.data(), d_sq, np * sizeof(double), hipMemcpyDeviceToHost, s));
        NGC_HIP(hipMemcpyAsync(sw.data(), sum_weight, np * sizeof(double),
                               hipMemcpyDeviceToHost, s));
        NGC_HIP(hipStreamSynchronize(s));
        // f0 = (5*10^-R)^2 / (sum g^2 / sum w), f1 = 1
        const double r = 5.0 * std::pow(10.0, -robust);
        for (int64_t i = 0; i < np; i++) {
            factors[i] = (r * r) / (sq[i] / sw[i]);
            factors[np + i] = 1.0;
        }
    } else {
        for (int64_t i = 0; i < np; i++) {
            factors[i] = 1.0;
            factors[np + i] = 0.0;
        }
    }
    return NGC_OK;
}

int ngc_imaging_weight_degrid(ngc_context *ctx, void *stream, const double *uvw,
                              const void *natural_weight, int weight_dtype, const uint8_t *flag,
                              const double *freq_chan, const int32_t *chan_map,
                              const int32_t *pol_map, const double *density_grid,
                              const double *briggs_factors, int64_t n_time,
                              int64_t n_baseline, int64_t n_chan, int64_t n_pol,
                              int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
                              int64_t n_v, double delta_l, double delta_m,
                              void *imaging_weight, int out_dtype) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(density_grid && briggs_factors && freq_chan && chan_map && pol_map, NGC_ERR_ARG,
              "NULL array");
    NGC_CHECK((uvw && natural_weight && imaging_weight) || n_time * n_baseline * n_chan == 0,
              NGC_ERR_ARG, "NULL uvw, weight or output");
    NGC_CHECK(weight_dtype == NGC_F32 || weight_dtype == NGC_F64, NGC_ERR_ARG, "weight dtype");
    NGC_CHECK(out_dtype == NGC_F32 || out_dtype == NGC_F64, NGC_ERR_ARG, "out dtype");
    for (int64_t c = 0; c < n_chan; c++)
        NGC_CHECK(chan_map[c] >= 0 && chan_map[c] < n_imag_chan, NGC_ERR_ARG, "chan_map range");
    for (int64_t p = 0; p < n_pol; p++)
        NGC_CHECK(pol_map[p] >= 0 && pol_map[p] < n_imag_pol, NGC_ERR_ARG, "pol_map range");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    std::vector<double> scale;
    uv_scale_host(freq_chan, n_chan, delta_l, delta_m, n_u, n_v, scale);
    void *d_scale, *d_cmap, *d_pmap, *d_fac;
    NGC_TRY(upload(ctx, s, "uv_scale", scale.data(), scale.size() * sizeof(double), &d_scale));
    NGC_TRY(upload(ctx, s, "chan_map", padded_map(chan_map, n_chan).data(), n_chan * sizeof(int32_t) + 4,
                   &d_cmap));
    NGC_TRY(upload(ctx, s, "pol_map", pol_map, n_pol * sizeof(int32_t), &d_pmap));
    NGC_TRY(upload(ctx, s, "briggs", briggs_factors, 2 * n_imag_chan * n_imag_pol * sizeof(double),
                   &d_fac));
    DegridArgs a{};
    a.uvw = uvw;
    a.nat = natural_weight;
    a.flag = flag;
    a.uv_scale = (const double *)d_scale;
    a.chan_map = (const int *)d_cmap;
    a.pol_map = (const int *)d_pmap;
    a.grid = density_grid;
    a.factors = (const double *)d_fac;
    a.n_time = n_time;
    a.n_baseline = n_baseline;
    a.n_chan = n_chan;
    a.n_pol = n_pol;
    a.n_imag_chan = n_imag_chan;
    a.n_imag_pol = n_imag_pol;
    a.n_u = n_u;
    a.n_v = n_v;
    a.weight_dtype = weight_dtype;
    a.out_dtype = out_dtype;
    a.out = imaging_weight;
    const int64_t n_rc = n_time * n_baseline * n_chan;
    if (n_rc == 0) return NGC_OK;
    profile_new_call(ctx, s);
    phase_begin(ctx, s, "imaging_weight_degrid");
    degrid_kernel<<<blocks_for(n_baseline * ceil_div(n_time, (int64_t)8) *
                               ceil_div(n_chan, (int64_t)8) * 64),
                    256, 0, s>>>(a);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}

int ngc_normalize_mosaic(ngc_context *ctx, void *stream, double *image,
                         const double *sum_weight, const double *ps_corr,
                         const double *norm_pb, const double *pb, int64_t n_planes,
                         int64_t n_l, int64_t n_m, int64_t os_u, int64_t os_v,
                         double pb_limit, int single_precision) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(image && sum_weight, NGC_ERR_ARG, "NULL array");
    NGC_CHECK(os_u >= 1 && os_v >= 1, NGC_ERR_ARG, "oversampling must be >= 1");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    const int64_t n = n_planes * n_l * n_m;
    if (n == 0) return NGC_OK;
    normalize_kernel<<<blocks_for(n), 256, 0, s>>>(image, sum_weight, ps_corr, norm_pb, pb,
                                                    n_planes, n_l, n_m, os_u, os_v, pb_limit,
                                                    single_precision);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

}  // extern "C"
