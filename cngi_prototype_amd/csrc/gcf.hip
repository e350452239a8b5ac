// A-term gridding convolution functions and phase gradients for gfx950.
//
//   ngc_gcf_patterns     make_baseline_patterns + the real part of
//                        fftshift(fft2(ifftshift(.))) of
//                        make_gridding_convolution_function.py:245-246, 394-412,
//                        with _casa_airy_disk_rorder / _airy_disk_rorder
//                        (_imaging_utils/_make_pb_symmetric.py:135-235)
//   ngc_gcf_resize       crop and normalisation of resize_and_calc_support
//                        (make_gridding_convolution_function.py:361-391)
//   ngc_phase_gradient   the table of make_phase_gradient (:349-359)
//   ngc_make_pb          the primary beam image of make_pb (make_pb.py:28-123) with
//                        _casa_airy_disk / _airy_disk (_make_pb_symmetric.py:24-130)
#include "ngc_internal.hpp"

namespace ngc {

struct PatternArgs {
    const double *freq;      // [n_freq]
    const double *dish;      // [n_dish]
    const double *blockage;  // [n_dish]
    const int *pairs;        // [n_pairs][2]
    int n_freq, n_pairs, casa, ipower;
    double cell_l, cell_m;   // pattern-grid cell (image cell x oversampling), radians
    int64_t n_u, n_v;
};

// Voltage (ipower = 1) or power (ipower = 2) pattern of dish d at radius
// argument r = |l| * k * D/2, in the reference's operation order.
__device__ inline double airy_pattern(double r, double dish, double blockage, int casa,
                                      int ipower) {
#pragma clang fp contract(off)
    double v;
    if (blockage == 0.0) {
        v = (2.0 * j1(r)) / r;
    } else if (casa) {
        const double area = (dish / blockage) * (dish / blockage);
        const double len = dish / blockage;
        v = (((area * 2.0) * j1(r)) / r - (2.0 * j1(r * len)) / (r * len)) / (area - 1.0);
    } else {
        const double e = blockage / dish;
        v = ((2.0 * j1(r)) / r - ((2.0 * e) * j1(r * e)) / r) / (1.0 - e * e);
    }
    return ipower == 2 ? v * v : v;
}

// Baseline pattern P = VP_i * VP_j at pattern pixel (x, y), written
// ifftshift-ed as a complex value into the FFT buffer.
__global__ void pattern_kernel(PatternArgs a, double2 *buf) {
#pragma clang fp contract(off)
    const int64_t plane = a.n_u * a.n_v;
    const int64_t n = (int64_t)a.n_pairs * a.n_freq * plane;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t item = i / plane;
        const int64_t r0 = i - item * plane;
        const int64_t ku = r0 / a.n_v, kv = r0 - ku * a.n_v;
        // buffer index k holds pattern pixel (k + n//2) mod n
        int64_t x = ku + a.n_u / 2;
        if (x >= a.n_u) x -= a.n_u;
        int64_t y = kv + a.n_v / 2;
        if (y >= a.n_v) y -= a.n_v;
        const int pair = (int)(item / a.n_freq);
        const int f = (int)(item - (int64_t)pair * a.n_freq);
        const int64_t cx = a.n_u / 2, cy = a.n_v / 2;
        double p = 1.0;
        if (!(x == cx && y == cy)) {
            const double xl = (double)(x - cx) * a.cell_l;
            const double ym = (double)(y - cy) * a.cell_m;
            const double k = (2.0 * M_PI * a.freq[f]) / kCLight;
            for (int s = 0; s < 2; s++) {
                const int d = a.pairs[2 * pair + s];
                const double aperture = a.dish[d] / 2.0;
                const double r = (sqrt(xl * xl + ym * ym) * k) * aperture;
                p = (s == 0) ? airy_pattern(r, a.dish[d], a.blockage[d], a.casa, a.ipower)
                             : p * airy_pattern(r, a.dish[d], a.blockage[d], a.casa, a.ipower);
            }
        }
        buf[i] = make_double2(p, 0.0);
    }
}

struct PbArgs {
    const double *freq;      // [n_chan]
    const double *dish;      // [n_dish]
    const double *blockage;  // [n_dish]
    int64_t n_l, n_m, n_chan, n_pol, n_dish;
    int64_t c_l, c_m;        // image centre pixel
    double cell_l, cell_m;   // radians
    int casa, ipower;
};

// One thread per (l, m, chan, dish): the dish's pattern at image pixel (l, m),
// written to every pol of out[l][m][chan][pol][dish].
__global__ void pb_kernel(PbArgs a, double *out) {
#pragma clang fp contract(off)
    const int64_t n = a.n_l * a.n_m * a.n_chan * a.n_dish;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = i % a.n_dish;
        const int64_t rest = i / a.n_dish;
        const int64_t c = rest % a.n_chan;
        const int64_t lm = rest / a.n_chan;
        const int64_t l = lm / a.n_m, m = lm - l * a.n_m;
        double p = 1.0;
        if (!(l == a.c_l && m == a.c_m)) {
            const double xl = (double)(l - a.c_l) * a.cell_l;
            const double ym = (double)(m - a.c_m) * a.cell_m;
            const double k = (2.0 * M_PI * a.freq[c]) / kCLight;
            const double r = (sqrt(xl * xl + ym * ym) * k) * (a.dish[d] / 2.0);
            p = airy_pattern(r, a.dish[d], a.blockage[d], a.casa, a.ipower);
        }
        double *o = out + (lm * a.n_chan + c) * a.n_pol * a.n_dish + d;
        for (int64_t q = 0; q < a.n_pol; q++) o[q * a.n_dish] = p;
    }
}

// out[item][u][v] = Re(B[(u - n_u//2) mod n_u][(v - n_v//2) mod n_v]) (fftshift)
__global__ void fftshift_real_kernel(const double2 *buf, int64_t n_items, int64_t n_u,
                                     int64_t n_v, double *out) {
    const int64_t plane = n_u * n_v;
    const int64_t n = n_items * plane;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t item = i / plane;
        const int64_t r0 = i - item * plane;
        const int64_t u = r0 / n_v, v = r0 - u * n_v;
        int64_t ku = u - n_u / 2;
        if (ku < 0) ku += n_u;
        int64_t kv = v - n_v / 2;
        if (kv < 0) kv += n_v;
        out[i] = buf[item * plane + ku * n_v + kv].x;
    }
}

// max |x| and min |x| per item
__global__ void absminmax_kernel(const double *a, int64_t plane, double *mx, double *mn) {
    const int64_t item = blockIdx.y;
    const double *p = a + item * plane;
    double hi = 0.0, lo = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < plane;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double x = fabs(p[i]);
        hi = fmax(hi, x);
        lo = fmin(lo, x);
    }
    for (int off = 32; off > 0; off >>= 1) {
        hi = fmax(hi, __shfl_xor(hi, off, 64));
        lo = fmin(lo, __shfl_xor(lo, off, 64));
    }
    __shared__ double sh[4], sl[4];
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6] = hi;
        sl[threadIdx.x >> 6] = lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            sh[0] = fmax(sh[0], sh[w]);
            sl[0] = fmin(sl[0], sl[w]);
        }
        // non-negative doubles order like their bit patterns
        atomicMax((unsigned long long *)&mx[item], (unsigned long long)__double_as_longlong(sh[0]));
        atomicMin((unsigned long long *)&mn[item], (unsigned long long)__double_as_longlong(sl[0]));
    }
}

// crop the centre K window and divide by Re(sum over the (support+1)*os
// window) / (os_u*os_v)
__global__ void resize_kernel(const double *full, int64_t n_u, int64_t n_v, const int *support,
                              int os_u, int os_v, int64_t k_u, int64_t k_v, double *out) {
    const int64_t item = blockIdx.x;
    const double *src = full + item * n_u * n_v;
    const int64_t su = n_u / 2 - k_u / 2, sv = n_v / 2 - k_v / 2;
    const int64_t eu_n = (int64_t)(support[2 * item] + 1) * os_u;
    const int64_t ev_n = (int64_t)(support[2 * item + 1] + 1) * os_v;
    const int64_t eu = k_u / 2 - eu_n / 2, ev = k_v / 2 - ev_n / 2;
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < eu_n * ev_n; i += blockDim.x) {
        const int64_t a = i / ev_n, b = i - a * ev_n;
        acc += src[(su + eu + a) * n_v + (sv + ev + b)];
    }
    acc = wave_sum(acc);
    __shared__ double sred[16];
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = acc;
    __syncthreads();
    double total = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) total += sred[w];
    const double nf = total / (double)(os_u * os_v);
    for (int64_t i = threadIdx.x; i < k_u * k_v; i += blockDim.x) {
        const int64_t a = i / k_v, b = i - a * k_v;
        out[item * k_u * k_v + i] = src[(su + a) * n_v + (sv + b)] / nf;
    }
}

// PG[f][x][y] = exp(i (x*px_f + y*py_f)), x in [-K_u//2, K_u - K_u//2)
__global__ void phase_gradient_kernel(const double *pix, int64_t n_field, int64_t k_u,
                                      int64_t k_v, double2 *out) {
#pragma clang fp contract(off)
    const int64_t plane = k_u * k_v;
    const int64_t n = n_field * plane;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = i / plane;
        const int64_t r0 = i - f * plane;
        const int64_t a = r0 / k_v, b = r0 - a * k_v;
        const double x = (double)(a - k_u / 2), y = (double)(b - k_v / 2);
        const double ph = x * pix[2 * f] + y * pix[2 * f + 1];
        double sn, cs;
        sincos(ph, &sn, &cs);
        out[i] = make_double2(cs, sn);
    }
}

static int nblocks(int64_t n) {
    int64_t b = ceil_div(n, 256);
    if (b > (1 << 20)) b = 1 << 20;
    return (int)(b < 1 ? 1 : b);
}

}  // namespace ngc

using namespace ngc;

extern "C" int ngc_gcf_patterns(ngc_context *ctx, void *stream, const double *pb_freq,
                                int64_t n_freq, const double *dish, const double *blockage,
                                int64_t n_dish, const int32_t *pairs, int64_t n_pairs,
                                int casa, int ipower, double cell_l, double cell_m,
                                int64_t n_u, int64_t n_v, double *out, double *abs_max,
                                double *abs_min, double *centre_row, double *centre_col) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(pb_freq && dish && blockage && pairs && out, NGC_ERR_ARG, "NULL array");
    NGC_CHECK(n_freq >= 1 && n_dish >= 1 && n_pairs >= 1 && n_u >= 2 && n_v >= 2, NGC_ERR_ARG,
              "bad dimensions");
    NGC_CHECK(ipower == 1 || ipower == 2, NGC_ERR_ARG, "ipower must be 1 or 2");
    for (int64_t i = 0; i < 2 * n_pairs; i++)
        NGC_CHECK(pairs[i] >= 0 && pairs[i] < n_dish, NGC_ERR_ARG, "pair index out of range");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    void *d_freq, *d_dish, *d_blk, *d_pairs, *buf, *d_mx, *d_mn;
    NGC_TRY(upload(ctx, s, "gcf_freq", pb_freq, n_freq * sizeof(double), &d_freq));
    NGC_TRY(upload(ctx, s, "gcf_dish", dish, n_dish * sizeof(double), &d_dish));
    NGC_TRY(upload(ctx, s, "gcf_blk", blockage, n_dish * sizeof(double), &d_blk));
    NGC_TRY(upload(ctx, s, "gcf_pairs", pairs, 2 * n_pairs * sizeof(int32_t), &d_pairs));
    const int64_t items = n_pairs * n_freq;
    const int64_t plane = n_u * n_v;
    NGC_TRY(workspace(ctx, "gcf_fft", items * plane * sizeof(double2), &buf));
    PatternArgs a;
    a.freq = (const double *)d_freq;
    a.dish = (const double *)d_dish;
    a.blockage = (const double *)d_blk;
    a.pairs = (const int *)d_pairs;
    a.n_freq = (int)n_freq;
    a.n_pairs = (int)n_pairs;
    a.casa = casa;
    a.ipower = ipower;
    a.cell_l = cell_l;
    a.cell_m = cell_m;
    a.n_u = n_u;
    a.n_v = n_v;
    pattern_kernel<<<nblocks(items * plane), 256, 0, s>>>(a, (double2 *)buf);
    NGC_HIP(hipGetLastError());
    NGC_TRY(fft_execute(ctx, s, buf, n_u, n_v, items, true, true));
    fftshift_real_kernel<<<nblocks(items * plane), 256, 0, s>>>((const double2 *)buf, items, n_u,
                                                                 n_v, out);
    NGC_HIP(hipGetLastError());
    if (abs_max || abs_min) {
        NGC_TRY(workspace(ctx, "gcf_mx", items * sizeof(double), &d_mx));
        NGC_TRY(workspace(ctx, "gcf_mn", items * sizeof(double), &d_mn));
        NGC_HIP(hipMemsetAsync(d_mx, 0, items * sizeof(double), s));
        NGC_HIP(hipMemsetAsync(d_mn, 0x7f, items * sizeof(double), s));
        dim3 g((unsigned)std::min<int64_t>(ceil_div(plane, 256 * 8), 1024), (unsigned)items);
        absminmax_kernel<<<g, 256, 0, s>>>(out, plane, (double *)d_mx, (double *)d_mn);
        NGC_HIP(hipGetLastError());
        if (abs_max)
            NGC_HIP(hipMemcpyAsync(abs_max, d_mx, items * sizeof(double), hipMemcpyDeviceToHost, s));
        if (abs_min)
            NGC_HIP(hipMemcpyAsync(abs_min, d_mn, items * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    for (int64_t it = 0; it < items; it++) {
        const double *pl = out + it * plane;
        if (centre_row)  // pattern[n_u//2 .. n_u-1][n_v//2]: the +u walk of calc_conv_size
            NGC_HIP(hipMemcpy2DAsync(centre_row + it * n_u, sizeof(double),
                                     pl + (n_v / 2), n_v * sizeof(double), sizeof(double), n_u,
                                     hipMemcpyDeviceToHost, s));
        if (centre_col)  // pattern[n_u//2][0 .. n_v-1]: the +v walk
            NGC_HIP(hipMemcpyAsync(centre_col + it * n_v, pl + (n_u / 2) * n_v,
                                   n_v * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    NGC_HIP(hipStreamSynchronize(s));
    return NGC_OK;
}

extern "C" int ngc_gcf_resize(ngc_context *ctx, void *stream, const double *full,
                              int64_t n_items, int64_t n_u, int64_t n_v, const int32_t *support,
                              int64_t os_u, int64_t os_v, int64_t k_u, int64_t k_v, double *out) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(full && support && out, NGC_ERR_ARG, "NULL array");
    NGC_CHECK(k_u <= n_u && k_v <= n_v && k_u >= 1 && k_v >= 1, NGC_ERR_ARG,
              "kernel window larger than the pattern grid");
    for (int64_t i = 0; i < n_items; i++) {
        NGC_CHECK((support[2 * i] + 1) * os_u <= k_u && (support[2 * i + 1] + 1) * os_v <= k_v &&
                      support[2 * i] >= 0 && support[2 * i + 1] >= 0,
                  NGC_ERR_ARG, "support window larger than the kernel window");
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    void *d_sup;
    NGC_TRY(upload(ctx, s, "gcf_support", support, 2 * n_items * sizeof(int32_t), &d_sup));
    if (n_items == 0) return NGC_OK;
    resize_kernel<<<(unsigned)n_items, 256, 0, s>>>(full, n_u, n_v, (const int *)d_sup, (int)os_u,
                                                    (int)os_v, k_u, k_v, out);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

extern "C" int ngc_phase_gradient(ngc_context *ctx, void *stream, const double *pix,
                                  int64_t n_field, int64_t k_u, int64_t k_v, double *out) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(pix && out, NGC_ERR_ARG, "NULL array");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    void *d_pix;
    NGC_TRY(upload(ctx, s, "pg_pix", pix, 2 * n_field * sizeof(double), &d_pix));
    const int64_t n = n_field * k_u * k_v;
    if (n == 0) return NGC_OK;
    phase_gradient_kernel<<<nblocks(n), 256, 0, s>>>((const double *)d_pix, n_field, k_u, k_v,
                                                     (double2 *)out);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

extern "C" int ngc_make_pb(ngc_context *ctx, void *stream, const double *freq_chan,
                           int64_t n_chan, const double *dish, const double *blockage,
                           int64_t n_dish, int casa, int ipower, int64_t n_l, int64_t n_m,
                           int64_t center_l, int64_t center_m, double cell_l, double cell_m,
                           int64_t n_pol, double *out) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(freq_chan && dish && blockage && out, NGC_ERR_ARG, "NULL array");
    NGC_CHECK(n_chan >= 1 && n_dish >= 1 && n_pol >= 1 && n_l >= 1 && n_m >= 1, NGC_ERR_ARG,
              "bad dimensions");
    NGC_CHECK(center_l >= 0 && center_l < n_l && center_m >= 0 && center_m < n_m, NGC_ERR_ARG,
              "image centre outside the image");
    NGC_CHECK(ipower == 1 || ipower == 2, NGC_ERR_ARG, "ipower must be 1 or 2");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    void *d_freq, *d_dish, *d_blk;
    NGC_TRY(upload(ctx, s, "pb_freq", freq_chan, n_chan * sizeof(double), &d_freq));
    NGC_TRY(upload(ctx, s, "pb_dish", dish, n_dish * sizeof(double), &d_dish));
    NGC_TRY(upload(ctx, s, "pb_blk", blockage, n_dish * sizeof(double), &d_blk));
    PbArgs a;
    a.freq = (const double *)d_freq;
    a.dish = (const double *)d_dish;
    a.blockage = (const double *)d_blk;
    a.n_l = n_l;
    a.n_m = n_m;
    a.n_chan = n_chan;
    a.n_pol = n_pol;
    a.n_dish = n_dish;
    a.c_l = center_l;
    a.c_m = center_m;
    a.cell_l = cell_l;
    a.cell_m = cell_m;
    a.casa = casa;
    a.ipower = ipower;
    pb_kernel<<<nblocks(n_l * n_m * n_chan * n_dish), 256, 0, s>>>(a, out);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}
