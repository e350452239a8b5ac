// Visibility prediction (the major-cycle inverse of the gridding path).
//
// The reference has no implementation: its degridding graph prints that
// "Degridding of visibilities and psf still needs to be implemented"
// (ngcasa/imaging/_imaging_utils/_standard_grid.py:429-430) and
// predict_modelvis_image (ngcasa/imaging/predict_modelvis_image.py:20) is a
// stub whose docstring names the steps: model image -> gridding correction ->
// Fourier transform -> degrid.  Here:
//
//   ngc_model_to_grid     model image [planes, n_l, n_m] (Jy/pixel) divided by
//                         the cropped correction image, padded to the centre of
//                         the n_u x n_v grid, fftshift(fft2(ifftshift(.)))
//                         (unnormalised forward): the inverse placement of
//                         ngc_grid_to_image; complex64 grids of lengths with
//                         prime factors up to 13 by the LDS row FFTs, others
//                         by rocFFT
//   ngc_standard_degrid   V = sum_taps conv * G[cell] (/ sum_taps conv), the
//                         adjoint of the standard gridder's loop with unit
//                         weights: same uv scale, centre cell, sub-cell offset,
//                         bounds and tap index (_standard_grid.py:273-360)
//   ngc_aperture_degrid   V = sum_taps conj(CK PG) G[cell] (/ sum_taps CK),
//                         the adjoint of the aperture gridder's loop with unit
//                         weights (_aperture_grid.py:377-513), per field,
//                         baseline, channel and pol kernel
//
// The standard degrid kernel is a gather: one thread per sample; the 64 lanes
// of a wavefront are 64 consecutive integrations of one (baseline, channel),
// whose footprints in one grid plane overlap, so their tap reads share cache
// lines.  Its inner 7-tap sums run in the grid's precision, the outer sums and
// the normalisation in FP64.  The aperture kernels are described where they
// are defined.
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "ngc_internal.hpp"

namespace ngc {

// buf[pl][ku][kv] = X[l][m] / (corr_l[l] corr_m[m]) for the model pixel
// (l, m) = ((ku + n_l//2) mod n_u, (kv + n_m//2) mod n_v) when inside the
// model, else 0 (the padded image, ifftshift-ed)
template <typename T>
__global__ void model_pad_kernel(const double *image, int64_t n_planes, int64_t n_l,
                                 int64_t n_m, int64_t n_u, int64_t n_v, const double *corr,
                                 T *buf) {
#pragma clang fp contract(off)
    const int64_t plane = n_u * n_v;
    const int64_t n = n_planes * plane;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pl = i / plane;
        const int64_t r = i - pl * plane;
        const int64_t ku = r / n_v, kv = r - ku * n_v;
        int64_t l = ku + n_l / 2;
        if (l >= n_u) l -= n_u;
        int64_t m = kv + n_m / 2;
        if (m >= n_v) m -= n_v;
        double x = 0.0;
        if (l < n_l && m < n_m) {
            x = image[(pl * n_l + l) * n_m + m];
            if (corr) x = x / (corr[l] * corr[n_l + m]);
        }
        buf[2 * i] = (T)x;
        buf[2 * i + 1] = (T)0;
    }
}

// grid[pl][u][v] = buf[pl][(u - n_u//2) mod n_u][(v - n_v//2) mod n_v]  (fftshift)
template <typename T, typename G>
__global__ void fftshift_complex_kernel(const T *buf, int64_t n_planes, int64_t n_u,
                                        int64_t n_v, G *grid) {
    const int64_t plane = n_u * n_v;
    const int64_t n = n_planes * plane;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pl = i / plane;
        const int64_t r = i - pl * plane;
        const int64_t u = r / n_v, v = r - u * n_v;
        int64_t ku = u - n_u / 2;
        if (ku < 0) ku += n_u;
        int64_t kv = v - n_v / 2;
        if (kv < 0) kv += n_v;
        const int64_t src = pl * plane + ku * n_v + kv;
        grid[2 * i] = (G)buf[2 * src];
        grid[2 * i + 1] = (G)buf[2 * src + 1];
    }
}

struct DegridStdArgs {
    const double *uvw;
    const double *uv_scale;  // [2][n_chan]
    const int *chan_map, *pol_map;
    const double *cgk;
    int64_t cgk_len;
    int64_t n_time, n_baseline, n_chan, n_pol, n_imag_pol, n_u, n_v, support, os;
    int normalize;
};

// Tap weights of one axis: c[i] = cgk[|os (i - S//2) + d|] for i < S.
template <int SMAX>
__device__ inline void axis_taps(const double *s_k, int64_t S, int64_t os, int64_t d,
                                 double (&c)[SMAX]) {
#pragma unroll
    for (int i = 0; i < SMAX; i++) {
        const int64_t k = os * (i - S / 2) + d;
        c[i] = i < S ? s_k[min(k < 0 ? -k : k, (int64_t)511)] : 0.0;
    }
}

// One thread per sample (all pols); the 64 lanes of a wavefront are 64
// consecutive integrations of one (baseline, channel), whose footprints in the
// same grid plane overlap, so their tap reads share cache lines.  SMAX is the
// compile-time bound of the support (taps beyond the support weigh 0 and are
// skipped); NP the number of pols (0: n_pol at run time).
//
// The kernel is separable, so each pol's sum is taken as
// sum_u cu[u] (sum_v cv[v] g[u][v]): the inner sums along v in the grid's
// precision (FP32 taps for a complex64 grid, 7 terms each), the outer sums
// and the normalisation by the tap sum (sum cu)(sum cv) in FP64.
template <typename G, typename V, int SMAX, int NP>
__global__ void __launch_bounds__(256) standard_degrid_kernel(DegridStdArgs a, const G *grid,
                                                              V *vis) {
    __shared__ double s_k[512];
    for (int i = threadIdx.x; i < 512; i += blockDim.x)
        s_k[i] = i < a.cgk_len ? a.cgk[i] : 0.0;
    __syncthreads();
    const int64_t n_tb = (a.n_time + 63) / 64;
    const int64_t n_thr = a.n_baseline * a.n_chan * n_tb * 64;
    const int64_t sc = a.support / 2;
    const int64_t plane = a.n_u * a.n_v;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_thr;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t w = i >> 6;
        const int64_t t = (w % n_tb) * 64 + (i & 63);
        const int64_t r = w / n_tb;
        const int64_t ch = r % a.n_chan, b = r / a.n_chan;
        if (t >= a.n_time) continue;
        const int64_t row = t * a.n_baseline + b;
        V *out = vis + 2 * ((row * a.n_chan + ch) * a.n_pol);
        Pix px;
        bool on = pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                                 a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os, px);
        on = on && (px.uc + sc < a.n_u) && (px.vc + sc < a.n_v) && (px.uc - sc >= 0) &&
             (px.vc - sc >= 0);
        if (!on) {
            for (int64_t p = 0; p < a.n_pol; p++) {
                out[2 * p] = (V)0;
                out[2 * p + 1] = (V)0;
            }
            continue;
        }
        double cu[SMAX], cv[SMAX];
        axis_taps<SMAX>(s_k, a.support, a.os, px.du, cu);
        axis_taps<SMAX>(s_k, a.support, a.os, px.dv, cv);
        G cvg[SMAX];
        double su = 0.0, sv = 0.0;
#pragma unroll
        for (int i = 0; i < SMAX; i++) {
            cvg[i] = (G)cv[i];
            su += cu[i];
            sv += cv[i];
        }
        const double norm = su * sv;   // the sum of all taps
        const int64_t a_chan = a.chan_map[ch];
        const int64_t u0 = px.uc - sc, v0 = px.vc - sc;
        const int n_pol = NP > 0 ? NP : (int)a.n_pol;
        const int S = (int)a.support;
#pragma unroll
        for (int p = 0; p < n_pol; p++) {
            const G *g = grid + 2 * ((a_chan * a.n_imag_pol + a.pol_map[p]) * plane +
                                     u0 * a.n_v + v0);
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int iu = 0; iu < SMAX; iu++) {
                if (iu >= S) break;
                const G *gr = g + 2 * (int64_t)iu * a.n_v;
                G ir = (G)0, ii = (G)0;
#pragma unroll
                for (int iv = 0; iv < SMAX; iv++) {
                    if (iv < S) {
                        ir = fma(cvg[iv], gr[2 * iv], ir);
                        ii = fma(cvg[iv], gr[2 * iv + 1], ii);
                    }
                }
                re = fma(cu[iu], (double)ir, re);
                im = fma(cu[iu], (double)ii, im);
            }
            if (a.normalize && norm != 0.0) {
                re = re / norm;
                im = im / norm;
            }
            out[2 * p] = (V)re;
            out[2 * p + 1] = (V)im;
        }
    }
}

// ---------------------------------------------------------------------------
// Tiled standard degridding (complex64 grids, support <= 8, 1, 2 or 4 pols).
// The gather above reads 49 taps per pol from L1 for every sample; here each
// sample is binned by the 16 x 32 tile of its footprint's first cell
// (u0 = uc - S//2, v0 = vc - S//2), and one workgroup per non-empty bin
// stages the tile with a 7-cell halo for every pol in LDS (23 x 39 cells, the
// footprint of every sample of the bin) and degrids its samples from there,
// one lane per sample with the gather's own arithmetic (same taps, same sums
// in the same order), so both paths give the same bits.
//   degrid_bin_kernel<false>  count: one wavefront per 64 consecutive
//                             integrations of one (baseline, channel); lanes
//                             of equal bin form runs, one atomic per run
//   hipcub exclusive scan     bin offsets
//   degrid_bin_kernel<true>   fill: the same runs take their ranges from the
//                             bin cursors and write (sample, record) entries;
//                             samples off the grid get zero visibilities here
//   standard_degrid_tiled_kernel  one workgroup per bin (empty bins return)
// Record: bits 0-3 u0 - tile row, 4-8 v0 - tile column, 9-16 du + 128,
// 17-24 dv + 128.
// ---------------------------------------------------------------------------
constexpr int kDgTU = 16, kDgTV = 32, kDgHalo = 7;
constexpr int kDgRows = kDgTU + kDgHalo, kDgCols = kDgTV + kDgHalo;   // 23 x 39

struct DgBins {
    int64_t n_tu, n_tv, n_bins;
    unsigned int *count;    // [n_bins + 1]
    unsigned int *offset;   // [n_bins + 1], exclusive scan of count
    unsigned int *cursor;   // [n_bins], fill cursors
    uint2 *entries;         // (sample index row * n_chan + chan, record)
};

template <bool FILL, typename V>
__global__ void __launch_bounds__(256) degrid_bin_kernel(DegridStdArgs a, DgBins bn, V *vis) {
    const int64_t n_tb = (a.n_time + 63) / 64;
    const int64_t n_thr = a.n_baseline * a.n_chan * n_tb * 64;
    const int64_t sc = a.support / 2;
    // n_thr is a multiple of 64 and the stride of 256, so every wavefront
    // runs the loop a uniform number of times (the ballots need all lanes)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_thr;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int lane = (int)(i & 63);
        const int64_t w = i >> 6;
        const int64_t t = (w % n_tb) * 64 + lane;
        const int64_t r = w / n_tb;
        const int64_t ch = r % a.n_chan, b = r / a.n_chan;
        int64_t key = -1, rc = 0;
        uint32_t rec = 0;
        if (t < a.n_time) {
            const int64_t row = t * a.n_baseline + b;
            rc = row * a.n_chan + ch;
            Pix px;
            bool on = pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                                     a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os, px);
            on = on && (px.uc + sc < a.n_u) && (px.vc + sc < a.n_v) && (px.uc - sc >= 0) &&
                 (px.vc - sc >= 0);
            if (on) {
                const int64_t u0 = px.uc - sc, v0 = px.vc - sc;
                const int64_t tu = u0 / kDgTU, tv = v0 / kDgTV;
                key = ((int64_t)a.chan_map[ch] * bn.n_tu + tu) * bn.n_tv + tv;
                rec = (uint32_t)(u0 - tu * kDgTU) | (uint32_t)(v0 - tv * kDgTV) << 4 |
                      (uint32_t)(px.du + 128) << 9 | (uint32_t)(px.dv + 128) << 17;
            } else if (FILL) {
                V *out = vis + 2 * rc * a.n_pol;
                for (int64_t p = 0; p < 2 * a.n_pol; p++) out[p] = (V)0;
            }
        }
        // runs of equal keys among consecutive lanes
        const int64_t prev = __shfl_up(key, 1, 64);
        const bool boundary = key < 0 || lane == 0 || prev != key;
        const unsigned long long bounds = __ballot(boundary);
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
        const int start = 63 - __clzll(bounds & upto);          // this lane's run head
        const unsigned long long after = bounds & ~upto;
        const int next = after ? __ffsll(after) - 1 : 64;       // first lane of the next run
        unsigned int base = 0;
        if (key >= 0 && lane == start) {
            if (FILL)
                base = atomicAdd(&bn.cursor[key], (unsigned int)(next - start));
            else
                atomicAdd(&bn.count[key], (unsigned int)(next - start));
        }
        if (FILL) {
            base = __shfl(base, start, 64);
            if (key >= 0) bn.entries[base + (lane - start)] = make_uint2((uint32_t)rc, rec);
        }
    }
}

// SC: the support when known at compile time (the tap loops then have no
// branches and the two samples' reads interleave), 0 for any support <= 8
template <int NP, typename V, int SC>
__global__ void __launch_bounds__(256)
standard_degrid_tiled_kernel(DegridStdArgs a, const float *grid, V *vis, DgBins bn) {
    __shared__ double s_k[512];
    __shared__ float2 s_g[NP][kDgRows][kDgCols];
    const int64_t bin = blockIdx.x;
    const unsigned int beg = bn.offset[bin], end = bn.offset[bin + 1];
    if (beg == end) return;   // uniform: the workgroup's bin is empty
    for (int i = threadIdx.x; i < 512; i += blockDim.x) s_k[i] = i < a.cgk_len ? a.cgk[i] : 0.0;
    const int64_t tv = bin % bn.n_tv, rest = bin / bn.n_tv;
    const int64_t tu = rest % bn.n_tu, a_chan = rest / bn.n_tu;
    const int64_t r0 = tu * kDgTU, c0 = tv * kDgTV;
    const int64_t plane = a.n_u * a.n_v;
    const float2 *g2 = reinterpret_cast<const float2 *>(grid);
    // stage the tile and its halo for every pol: all loads first, then LDS
    constexpr int kCells = NP * kDgRows * kDgCols;
    constexpr int kPer = (kCells + 255) / 256;
    float2 x[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int idx = threadIdx.x + k * 256;
        const int p = idx / (kDgRows * kDgCols), rem = idx - p * (kDgRows * kDgCols);
        const int r = rem / kDgCols, c = rem - r * kDgCols;
        const int64_t gu = r0 + r, gv = c0 + c;
        x[k] = make_float2(0.f, 0.f);
        if (idx < kCells && gu < a.n_u && gv < a.n_v)
            x[k] = g2[(a_chan * a.n_imag_pol + a.pol_map[p]) * plane + gu * a.n_v + gv];
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) {
        const int idx = threadIdx.x + k * 256;
        if (idx < kCells) (&s_g[0][0][0])[idx] = x[k];
    }
    __syncthreads();
    const int S = SC > 0 ? SC : (int)a.support;
    constexpr int SL = SC > 0 ? SC : 8;   // tap loop bound
    // one sample per lane at a time, two in flight: the second sample's
    // entry, taps and LDS reads overlap the first's (the arithmetic of each
    // is the gather's, so the bits are the same)
    auto degrid_one = [&](const uint2 en, double (&re)[NP], double (&im)[NP]) {
        const uint32_t rec = en.y;
        const int ru = rec & 15, rv = (rec >> 4) & 31;
        const int64_t du = (int64_t)((rec >> 9) & 255) - 128, dv = (int64_t)((rec >> 17) & 255) - 128;
        double cu[8], cv[8];
        axis_taps<8>(s_k, a.support, a.os, du, cu);
        axis_taps<8>(s_k, a.support, a.os, dv, cv);
        float cvg[8];
        double su = 0.0, sv = 0.0;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            cvg[i] = (float)cv[i];
            su += cu[i];
            sv += cv[i];
        }
        const double norm = su * sv;
#pragma unroll
        for (int p = 0; p < NP; p++) {
            double r = 0.0, m = 0.0;
#pragma unroll
            for (int iu = 0; iu < SL; iu++) {
                if (SC == 0 && iu >= S) break;
                const float2 *gr = &s_g[p][ru + iu][rv];
                float ir = 0.f, ii = 0.f;
#pragma unroll
                for (int iv = 0; iv < SL; iv++) {
                    if (SC > 0 || iv < S) {
                        const float2 gval = gr[iv];
                        ir = fmaf(cvg[iv], gval.x, ir);
                        ii = fmaf(cvg[iv], gval.y, ii);
                    }
                }
                r = fma(cu[iu], (double)ir, r);
                m = fma(cu[iu], (double)ii, m);
            }
            if (a.normalize && norm != 0.0) {
                r = r / norm;
                m = m / norm;
            }
            re[p] = r;
            im[p] = m;
        }
    };
    auto store = [&](const uint2 en, const double (&re)[NP], const double (&im)[NP]) {
        V *out = vis + 2 * ((int64_t)en.x * NP);
#pragma unroll
        for (int p = 0; p < NP; p++) {
            out[2 * p] = (V)re[p];
            out[2 * p + 1] = (V)im[p];
        }
    };
    const unsigned int stride = blockDim.x;
    for (unsigned int e = beg + threadIdx.x; e < end; e += 2 * stride) {
        const bool two = e + stride < end;
        const uint2 en0 = bn.entries[e];
        const uint2 en1 = bn.entries[two ? e + stride : e];
        double re0[NP], im0[NP], re1[NP], im1[NP];
        degrid_one(en0, re0, im0);
        degrid_one(en1, re1, im1);
        store(en0, re0, im0);
        if (two) store(en1, re1, im1);
    }
}

static int pblocks(int64_t n) {
    int64_t b = ceil_div(n, 256);
    if (b > (1 << 20)) b = 1 << 20;
    return (int)(b < 1 ? 1 : b);
}

}  // namespace ngc

using namespace ngc;

extern "C" int ngc_model_to_grid(ngc_context *ctx, void *stream, const double *image,
                                 int64_t n_planes, int64_t n_l, int64_t n_m, int64_t n_u,
                                 int64_t n_v, const double *corr_l, const double *corr_m,
                                 void *grid, int grid_dtype) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(image && grid, NGC_ERR_ARG, "NULL image or grid");
    NGC_CHECK(n_planes >= 1 && n_l >= 1 && n_m >= 1 && n_l <= n_u && n_m <= n_v, NGC_ERR_ARG,
              "bad dimensions (the model must fit the grid)");
    NGC_CHECK(grid_dtype == NGC_C64 || grid_dtype == NGC_C128, NGC_ERR_ARG,
              "grid dtype must be C64 or C128");
    NGC_CHECK((corr_l == nullptr) == (corr_m == nullptr), NGC_ERR_ARG,
              "corr_l and corr_m must both be given or both NULL");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    const bool dbl = grid_dtype == NGC_C128;
    const int64_t cells = n_planes * n_u * n_v;
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
    // complex64 grids of smooth lengths: the LDS transform (padded rows,
    // transpose, shifted rows, transpose); otherwise pad, rocFFT, fftshift
    if (!dbl && ctx->fft_lds && lds_fft_supported(n_u) && lds_fft_supported(n_v))
        return lds_model_to_grid(ctx, s, image, n_planes, n_l, n_m, n_u, n_v, d_corr, grid);
    void *buf;
    NGC_TRY(workspace(ctx, "fft_buffer", cells * (dbl ? 16 : 8), &buf));
    const int nb = pblocks(cells);
    phase_begin(ctx, s, "model_pad");
    if (dbl)
        model_pad_kernel<double><<<nb, 256, 0, s>>>(image, n_planes, n_l, n_m, n_u, n_v, d_corr,
                                                    (double *)buf);
    else
        model_pad_kernel<float><<<nb, 256, 0, s>>>(image, n_planes, n_l, n_m, n_u, n_v, d_corr,
                                                   (float *)buf);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    phase_begin(ctx, s, "rocfft_forward");
    NGC_TRY(fft_execute(ctx, s, buf, n_u, n_v, n_planes, dbl, true));
    phase_end(ctx, s);
    phase_begin(ctx, s, "fftshift");
    if (dbl)
        fftshift_complex_kernel<double, double><<<nb, 256, 0, s>>>(
            (const double *)buf, n_planes, n_u, n_v, (double *)grid);
    else
        fftshift_complex_kernel<float, float><<<nb, 256, 0, s>>>(
            (const float *)buf, n_planes, n_u, n_v, (float *)grid);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}

extern "C" int ngc_standard_degrid(ngc_context *ctx, void *stream, const double *uvw,
                                   const double *freq_chan, const int32_t *chan_map,
                                   const int32_t *pol_map, const double *cgk_1d, int64_t cgk_len,
                                   int64_t n_time, int64_t n_baseline, int64_t n_chan,
                                   int64_t n_pol, int64_t n_imag_chan, int64_t n_imag_pol,
                                   int64_t n_u, int64_t n_v, double delta_l, double delta_m,
                                   int64_t support, int64_t oversampling, const void *grid,
                                   int grid_dtype, void *vis, int vis_dtype, int normalize) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(n_time >= 0 && n_baseline >= 0 && n_chan >= 1 && n_pol >= 1 &&
                  n_imag_chan >= 1 && n_imag_pol >= 1 && n_u >= 1 && n_v >= 1 && support >= 1 &&
                  oversampling >= 0,
              NGC_ERR_ARG, "bad dimensions");
    NGC_CHECK(freq_chan && chan_map && pol_map && cgk_1d && cgk_len >= 1, NGC_ERR_ARG,
              "NULL parameter array");
    NGC_CHECK(cgk_len <= 512, NGC_ERR_UNSUPPORTED, "cgk_1d longer than 512");
    NGC_CHECK(support <= 16, NGC_ERR_UNSUPPORTED, "support above 16");
    NGC_CHECK(grid_dtype == NGC_C64 || grid_dtype == NGC_C128, NGC_ERR_ARG,
              "grid dtype must be C64 or C128");
    NGC_CHECK(vis_dtype == NGC_C64 || vis_dtype == NGC_C128, NGC_ERR_ARG,
              "vis dtype must be C64 or C128");
    for (int64_t c = 0; c < n_chan; c++)
        NGC_CHECK(chan_map[c] >= 0 && chan_map[c] < n_imag_chan, NGC_ERR_ARG,
                  "chan_map entry out of range");
    for (int64_t p = 0; p < n_pol; p++)
        NGC_CHECK(pol_map[p] >= 0 && pol_map[p] < n_imag_pol, NGC_ERR_ARG,
                  "pol_map entry out of range");
    // every tap index must fall inside the table: |os*i + d| <= os*(S//2 + 1/2)
    NGC_CHECK(oversampling * (support / 2) + oversampling / 2 < cgk_len || support == 1,
              NGC_ERR_ARG, "cgk_1d shorter than the support needs");
    if (n_time == 0 || n_baseline == 0) return NGC_OK;
    NGC_CHECK(uvw && grid && vis, NGC_ERR_ARG, "NULL uvw, grid or vis");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    std::vector<double> scale;
    uv_scale_host(freq_chan, n_chan, delta_l, delta_m, n_u, n_v, scale);
    void *d_scale, *d_cm, *d_pm, *d_cgk;
    NGC_TRY(upload(ctx, s, "dg_scale", scale.data(), scale.size() * sizeof(double), &d_scale));
    NGC_TRY(upload(ctx, s, "dg_chan_map", chan_map, n_chan * sizeof(int32_t), &d_cm));
    NGC_TRY(upload(ctx, s, "dg_pol_map", pol_map, n_pol * sizeof(int32_t), &d_pm));
    NGC_TRY(upload(ctx, s, "dg_cgk", cgk_1d, cgk_len * sizeof(double), &d_cgk));
    DegridStdArgs a;
    a.uvw = uvw;
    a.uv_scale = (const double *)d_scale;
    a.chan_map = (const int *)d_cm;
    a.pol_map = (const int *)d_pm;
    a.cgk = (const double *)d_cgk;
    a.cgk_len = cgk_len;
    a.n_time = n_time;
    a.n_baseline = n_baseline;
    a.n_chan = n_chan;
    a.n_pol = n_pol;
    a.n_imag_pol = n_imag_pol;
    a.n_u = n_u;
    a.n_v = n_v;
    a.support = support;
    a.os = oversampling;
    a.normalize = normalize ? 1 : 0;
    const int64_t n_thr = n_baseline * n_chan * ceil_div(n_time, 64) * 64;
    const int nb = pblocks(n_thr);
    profile_new_call(ctx, s);
    const int64_t n_tu = ceil_div(n_u, (int64_t)kDgTU), n_tv = ceil_div(n_v, (int64_t)kDgTV);
    const int64_t n_bins = n_imag_chan * n_tu * n_tv;
    if (ctx->degrid_tiled && grid_dtype == NGC_C64 && support <= 8 && oversampling <= 200 &&
        (n_pol == 1 || n_pol == 2 || n_pol == 4) && n_time * n_baseline * n_chan < (int64_t(1) << 32) &&
        n_bins < (int64_t(1) << 31)) {
        // the tiled path: bin, scan, fill, one workgroup per bin
        const int64_t n_rc = n_time * n_baseline * n_chan;
        DgBins bn;
        bn.n_tu = n_tu;
        bn.n_tv = n_tv;
        bn.n_bins = n_bins;
        void *p;
        NGC_TRY(workspace(ctx, "dg_count", (n_bins + 1) * sizeof(unsigned int), &p));
        bn.count = (unsigned int *)p;
        NGC_TRY(workspace(ctx, "dg_offset", (n_bins + 1) * sizeof(unsigned int), &p));
        bn.offset = (unsigned int *)p;
        NGC_TRY(workspace(ctx, "dg_cursor", n_bins * sizeof(unsigned int), &p));
        bn.cursor = (unsigned int *)p;
        NGC_TRY(workspace(ctx, "dg_entries", n_rc * sizeof(uint2), &p));
        bn.entries = (uint2 *)p;
        size_t scan_bytes = 0;
        NGC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, bn.count, bn.offset,
                                                 (int)(n_bins + 1), s));
        void *scan_tmp;
        NGC_TRY(workspace(ctx, "dg_scan", std::max<size_t>(scan_bytes, 8), &scan_tmp));
        const bool v64 = vis_dtype == NGC_C64;
        phase_begin(ctx, s, "degrid_bin_count");
        NGC_HIP(hipMemsetAsync(bn.count, 0, (n_bins + 1) * sizeof(unsigned int), s));
        if (v64)
            degrid_bin_kernel<false, float><<<nb, 256, 0, s>>>(a, bn, (float *)vis);
        else
            degrid_bin_kernel<false, double><<<nb, 256, 0, s>>>(a, bn, (double *)vis);
        NGC_HIP(hipGetLastError());
        phase_end(ctx, s);
        phase_begin(ctx, s, "degrid_bin_scan");
        NGC_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, bn.count, bn.offset,
                                                 (int)(n_bins + 1), s));
        NGC_HIP(hipMemcpyAsync(bn.cursor, bn.offset, n_bins * sizeof(unsigned int),
                               hipMemcpyDeviceToDevice, s));
        phase_end(ctx, s);
        phase_begin(ctx, s, "degrid_bin_fill");
        if (v64)
            degrid_bin_kernel<true, float><<<nb, 256, 0, s>>>(a, bn, (float *)vis);
        else
            degrid_bin_kernel<true, double><<<nb, 256, 0, s>>>(a, bn, (double *)vis);
        NGC_HIP(hipGetLastError());
        phase_end(ctx, s);
        phase_begin(ctx, s, "standard_degrid");
        auto tiled = [&](auto np_c) {
            constexpr int NP = decltype(np_c)::value;
            if (support == 7) {
                if (v64)
                    standard_degrid_tiled_kernel<NP, float, 7><<<(unsigned)n_bins, 256, 0, s>>>(
                        a, (const float *)grid, (float *)vis, bn);
                else
                    standard_degrid_tiled_kernel<NP, double, 7><<<(unsigned)n_bins, 256, 0, s>>>(
                        a, (const float *)grid, (double *)vis, bn);
            } else if (v64) {
                standard_degrid_tiled_kernel<NP, float, 0><<<(unsigned)n_bins, 256, 0, s>>>(
                    a, (const float *)grid, (float *)vis, bn);
            } else {
                standard_degrid_tiled_kernel<NP, double, 0><<<(unsigned)n_bins, 256, 0, s>>>(
                    a, (const float *)grid, (double *)vis, bn);
            }
        };
        if (n_pol == 1)
            tiled(std::integral_constant<int, 1>{});
        else if (n_pol == 2)
            tiled(std::integral_constant<int, 2>{});
        else
            tiled(std::integral_constant<int, 4>{});
        NGC_HIP(hipGetLastError());
        phase_end(ctx, s);
        return NGC_OK;
    }
    phase_begin(ctx, s, "standard_degrid");
    auto launch = [&](auto smax, auto np_c) {
        constexpr int SM = decltype(smax)::value;
        constexpr int NP = decltype(np_c)::value;
        if (grid_dtype == NGC_C64 && vis_dtype == NGC_C64)
            standard_degrid_kernel<float, float, SM, NP><<<nb, 256, 0, s>>>(
                a, (const float *)grid, (float *)vis);
        else if (grid_dtype == NGC_C64)
            standard_degrid_kernel<float, double, SM, NP><<<nb, 256, 0, s>>>(
                a, (const float *)grid, (double *)vis);
        else if (vis_dtype == NGC_C64)
            standard_degrid_kernel<double, float, SM, NP><<<nb, 256, 0, s>>>(
                a, (const double *)grid, (float *)vis);
        else
            standard_degrid_kernel<double, double, SM, NP><<<nb, 256, 0, s>>>(
                a, (const double *)grid, (double *)vis);
    };
    auto by_pol = [&](auto smax) {
        if (n_pol == 1)
            launch(smax, std::integral_constant<int, 1>{});
        else if (n_pol == 2)
            launch(smax, std::integral_constant<int, 2>{});
        else if (n_pol == 4)
            launch(smax, std::integral_constant<int, 4>{});
        else
            launch(smax, std::integral_constant<int, 0>{});
    };
    if (support <= 8)
        by_pol(std::integral_constant<int, 8>{});
    else
        by_pol(std::integral_constant<int, 16>{});
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}

// ---------------------------------------------------------------------------
// Aperture (A-term) degridding: the adjoint of the aperture gridder's loop
// (_aperture_grid.py:377-513) with unit weights.  One thread per (row,
// channel); the 64 lanes of a wavefront are 64 consecutive integrations of
// one (baseline, channel).  Per pol, V = sum conj(CK*PG) G[cell] over the
// gridder's cells, offsets and taps (v outer, u inner, FP64), divided by the
// kernel's tap sum sum CK (no phase gradient) when normalize.  Taps whose kernel index falls outside the
// K_u x K_v table weigh 0, as in the tiled gridder.  This general form runs
// when the phase gradient does not factor into u and v parts (or the tables
// of the tap-parallel form below would be too large).
// ---------------------------------------------------------------------------
namespace ngc {

struct DegridApArgs {
    const double *uvw;
    const int *field;       // device [T, B]
    const int *field_id;    // [n_field]
    int n_field;
    const double *uv_scale;  // [2][n_chan]
    const int *chan_map, *pol_map, *cfb, *cfc, *cfp, *support;
    const double *ck;        // [n_cfb, n_cfc, n_cfp, k_u, k_v]
    const double2 *pg;       // [n_field, k_u, k_v]
    int64_t n_cfc, n_cfp, k_u, k_v;
    int64_t n_time, n_baseline, n_chan, n_pol, n_imag_pol, n_u, n_v, os_u, os_v;
    int64_t max_sc;          // the largest support (the gridder's bounds margin)
    int normalize;
    int *error;
};

template <typename G, typename V>
__global__ void __launch_bounds__(256) aperture_degrid_kernel(DegridApArgs a, const G *grid,
                                                              V *vis) {
    const int64_t n_tb = (a.n_time + 63) / 64;
    const int64_t n_thr = a.n_baseline * a.n_chan * n_tb * 64;
    const int64_t plane = a.n_u * a.n_v;
    const int64_t kplane = a.k_u * a.k_v;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_thr;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t w = i >> 6;
        const int64_t t = (w % n_tb) * 64 + (i & 63);
        const int64_t r = w / n_tb;
        const int64_t ch = r % a.n_chan, b = r / a.n_chan;
        if (t >= a.n_time) continue;
        const int64_t row = t * a.n_baseline + b;
        V *out = vis + 2 * ((row * a.n_chan + ch) * a.n_pol);
        for (int64_t p = 0; p < a.n_pol; p++) {
            out[2 * p] = (V)0;
            out[2 * p + 1] = (V)0;
        }
        const int fval = a.field[row];
        if (!(fval > -1)) continue;
        int fidx = -1;
        for (int f = 0; f < a.n_field; f++)
            if (a.field_id[f] == fval) {
                fidx = f;
                break;
            }
        if (fidx < 0) {
            atomicOr(a.error, 1);
            continue;
        }
        Pix px, pxv;
        if (!pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                            a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os_u, px))
            continue;
        pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                       a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os_v, pxv);
        if (!((px.uc + a.max_sc < a.n_u) && (px.vc + a.max_sc < a.n_v) &&
              (px.uc - a.max_sc >= 0) && (px.vc - a.max_sc >= 0)))
            continue;
        const int64_t u_off = px.du + a.k_u / 2, v_off = pxv.dv + a.k_v / 2;
        const int64_t a_chan = a.chan_map[ch];
        const double2 *pg = a.pg + fidx * kplane;
        for (int64_t p = 0; p < a.n_pol; p++) {
            const int64_t cf = (a.cfb[b] * a.n_cfc + a.cfc[ch]) * a.n_cfp + a.cfp[p];
            const int64_t su = a.support[2 * cf], sv = a.support[2 * cf + 1];
            const int64_t scu = su / 2, scv = sv / 2;
            const double *ck = a.ck + cf * kplane;
            const G *g = grid + 2 * ((a_chan * a.n_imag_pol + a.pol_map[p]) * plane);
            double re = 0.0, im = 0.0, nrm = 0.0;
            for (int64_t iv = -scv; iv < sv - scv; iv++) {
                const int64_t kv = a.os_v * iv + v_off;
                const int64_t gv = px.vc + iv;
                for (int64_t iu = -scu; iu < su - scu; iu++) {
                    const int64_t ku = a.os_u * iu + u_off;
                    if (ku < 0 || ku >= a.k_u || kv < 0 || kv >= a.k_v) continue;
                    const int64_t kidx = ku * a.k_v + kv;
                    const double c = ck[kidx];
                    const double cr = c * pg[kidx].x, ci = c * pg[kidx].y;
                    const int64_t cell = (px.uc + iu) * a.n_v + gv;
                    const double gr = (double)g[2 * cell], gi = (double)g[2 * cell + 1];
                    re += cr * gr + ci * gi;   // conj(c) g
                    im += cr * gi - ci * gr;
                    nrm += c;
                }
            }
            if (a.normalize && nrm != 0.0) {   // the beam is 1 at the pointing centre
                re = re / nrm;
                im = im / nrm;
            }
            out[2 * p] = (V)re;
            out[2 * p + 1] = (V)im;
        }
    }
}

// ---------------------------------------------------------------------------
// Tap-parallel form of the same sum, for phase gradients of the form
// exp(i(x px + y py)) (the gcf builder's): PG[x][y] = PGu[x] PGv[y] with
// PGu[x] = PG[x][cv], PGv[y] = PG[cu][y] (so PG[cu][cv] = 1), so a tap's
// weight is CK * PGu * PGv from two short factor tables instead of a
// [field][K_u][K_v] table.
//
// A wavefront takes 64 integrations of one (baseline, channel), so the
// convolution function and its support are wave-uniform.  Each lane first
// places one integration (cell, sub-cell offset, field); then, in 16 rounds,
// 16-lane groups each take one integration and run over its su * sv taps in
// order, lane l taking taps l, l + 16, ...  The taps of one integration are
// one contiguous block of the re-laid kernel, so a group's kernel reads are
// 128 contiguous bytes and its grid reads follow rows of the footprint.  (One
// lane per integration had every lane read its own block: 64 cache lines per
// load, and the kernel ran at the L2's bandwidth.)  The group's partial sums
// are added across its 16 lanes.  Pols that share a convolution function
// (cf_pol_map constant) share the tap weights.
//
// All sums and tables are FP64, whatever the grid's precision: towards an
// offset pointing CK PG turns many times across the support, the sums cancel
// to a small fraction of sum |CK PG|, and FP32 sums lose that fraction's
// digits (1.5e-4 relative error on a 25-pointing mosaic).  FP64 multiply-adds
// issue at the rate of unpacked FP32 ones on gfx950.
//
// Tables (rebuilt per call):
//  - adg_taps: CK re-laid per (convolution function, sub-cell offset pair)
//    as a block of su * sv taps, row-major in (ru, rv):
//    ckr[cf_base[cf] + ((ou + hu) nwv + ov + hv) su sv + ru sv + rv] =
//    CK[os_u (ru - su/2) + ou + K_u/2][os_v (rv - sv/2) + ov + K_v/2], zero
//    where that index leaves the K_u x K_v table (a zero weight, as in the
//    general form);
//  - adg_pgu / adg_pgv: the factors re-laid the same way, per (field,
//    sub-cell offset) as W = R + 1 entries (R the largest support),
//    pur[(f nwu + ou + hu) W + iu + R/2] = PGu[os_u iu + ou + K_u/2] for the
//    tap offset iu = ru - su/2, zero where the index leaves the table, so a
//    group's factor reads are contiguous as well.
// ---------------------------------------------------------------------------
struct AdgTab {
    int hu, hv;               // sub-cell offset half-ranges
    int R, W;                 // largest support; factor entries per offset (R + 1)
    int64_t n_cf;
    int max_block;            // largest su * sv
    const int64_t *cf_base;   // [n_cf] first element of each function's blocks
};

__global__ void adg_taps_kernel(DegridApArgs a, AdgTab t, double *ckr) {
    const int nwu = 2 * t.hu + 1, nwv = 2 * t.hv + 1;
    const int64_t nwo = (int64_t)nwu * nwv;
    const int64_t n = t.n_cf * nwo * t.max_block;
    const int64_t kplane = a.k_u * a.k_v;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(i % t.max_block);
        const int64_t o = (i / t.max_block) % nwo;
        const int64_t cf = i / (nwo * t.max_block);
        const int su = a.support[cf * 2], sv = a.support[cf * 2 + 1];
        if (j >= su * sv) continue;
        const int ru = j / sv, rv = j - ru * sv;
        const int64_t ov = o % nwv - t.hv, ou = o / nwv - t.hu;
        const int64_t x = a.os_u * (ru - su / 2) + ou + a.k_u / 2;
        const int64_t y = a.os_v * (rv - sv / 2) + ov + a.k_v / 2;
        double v = 0.0;
        if (x >= 0 && x < a.k_u && y >= 0 && y < a.k_v) v = a.ck[cf * kplane + x * a.k_v + y];
        ckr[t.cf_base[cf] + o * (int64_t)(su * sv) + j] = v;
    }
}

// sets *bad when some PG[x][y] differs from PG[x][cv] PG[cu][y] by more than
// 1e-11 (the table has unit modulus; the gcf builder's exp(i(x px + y py))
// factors to ~1e-13).  An accepted table then moves a normalised prediction
// by at most 1e-11 of sum |CK G| / sum CK.
__global__ void adg_pg_check_kernel(DegridApArgs a, int *bad) {
#pragma clang fp contract(off)
    const int64_t cu = a.k_u / 2, cv = a.k_v / 2;
    const int64_t kplane = a.k_u * a.k_v;
    const int64_t n = (int64_t)a.n_field * kplane;
    int b = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = i / kplane, r = i - f * kplane;
        const int64_t x = r / a.k_v, y = r - x * a.k_v;
        const double2 gx = a.pg[f * kplane + x * a.k_v + cv];
        const double2 gy = a.pg[f * kplane + cu * a.k_v + y];
        const double2 g = a.pg[i];
        const double pr = gx.x * gy.x - gx.y * gy.y, pi = gx.x * gy.y + gx.y * gy.x;
        b |= !(fabs(pr - g.x) <= 1e-11 && fabs(pi - g.y) <= 1e-11);
    }
    if (b) atomicOr(bad, 1);
}

// pur[f][ou][iu] = PG[f][os_u (iu - R/2) + ou + K_u/2][cv], 0 off the table;
// pvr alike with PG[f][cu][.]
__global__ void adg_pg_factor_kernel(DegridApArgs a, AdgTab t, double *pur, double *pvr) {
    const int64_t cu = a.k_u / 2, cv = a.k_v / 2;
    const int64_t kplane = a.k_u * a.k_v;
    const int64_t nwu = 2 * t.hu + 1, nwv = 2 * t.hv + 1;
    const int64_t nu = (int64_t)a.n_field * nwu * t.W, n = nu + (int64_t)a.n_field * nwv * t.W;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const bool is_u = i < nu;
        const int64_t j = is_u ? i : i - nu;
        const int64_t nw = is_u ? nwu : nwv;
        const int64_t e = j % t.W - t.R / 2;
        const int64_t o = (j / t.W) % nw - (is_u ? t.hu : t.hv);
        const int64_t f = j / (t.W * nw);
        const int64_t x = (is_u ? a.os_u : a.os_v) * e + o + (is_u ? a.k_u : a.k_v) / 2;
        double2 g = make_double2(0.0, 0.0);
        if (is_u && x >= 0 && x < a.k_u) g = a.pg[f * kplane + x * a.k_v + cv];
        if (!is_u && x >= 0 && x < a.k_v) g = a.pg[f * kplane + cu * a.k_v + x];
        double *dst = is_u ? pur : pvr;
        dst[2 * j] = g.x;
        dst[2 * j + 1] = g.y;
    }
}

__device__ inline double group16_sum(double v) {
    v += __shfl_xor(v, 8);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 1);
    return v;
}

// NP pols per pass sharing one convolution function (NP = 1 when the pols'
// functions differ); G: the grid's element type.
template <typename G, typename V, int NP>
__global__ void __launch_bounds__(256) aperture_degrid_taps_kernel(DegridApArgs a, AdgTab t,
                                                                   const double *__restrict__ ckr,
                                                                   const double *__restrict__ pgu,
                                                                   const double *__restrict__ pgv,
                                                                   const G *__restrict__ grid,
                                                                   V *vis) {
    const int64_t n_tb = (a.n_time + 63) / 64;
    const int64_t n_w = a.n_baseline * a.n_chan * n_tb;
    const int lane = (int)(threadIdx.x & 63);
    const int grp = lane >> 4, l = lane & 15;
    const int64_t plane = a.n_u * a.n_v;
    const int nwu = 2 * t.hu + 1, nwv = 2 * t.hv + 1;
    const int64_t w0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t ws = ((int64_t)gridDim.x * blockDim.x) >> 6;
    // this wave's tap offsets for the support it last used: {2 ru, 2 rv,
    // 2 (ru n_v + rv)} per tap, so the tap loop reads them instead of
    // dividing its tap index by sv
    // Then each group's sums go to its integration's slot (2 NP + 2 doubles),
    // and each lane normalises and writes its own integration.
    extern __shared__ int4 s_taps[];
    int4 *taps = s_taps + (threadIdx.x >> 6) * t.max_block;
    double *res = reinterpret_cast<double *>(s_taps + 4 * t.max_block) +
                  (threadIdx.x >> 6) * 64 * (2 * NP + 2);
    int tab_su = 0, tab_sv = 0;
    for (int64_t w = w0; w < n_w; w += ws) {   // wave-uniform
        const int64_t t0 = (w % n_tb) * 64;
        const int64_t r = w / n_tb;
        const int64_t ch = r % a.n_chan, b = r / a.n_chan;
        // placement: lane = integration t0 + lane; fidx < 0 marks nothing to predict
        int uc = 0, vc = 0, du = 0, dv = 0, fidx = -1;
        const int64_t tm = t0 + lane;
        if (tm < a.n_time) {
            const int64_t row = tm * a.n_baseline + b;
            const int fval = a.field[row];
            int f_i = -1;
            if (fval > -1) {
                for (int f = 0; f < a.n_field; f++)
                    if (a.field_id[f] == fval) {
                        f_i = f;
                        break;
                    }
                if (f_i < 0) atomicOr(a.error, 1);
            }
            Pix px, pxv;
            if (f_i >= 0 &&
                pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                               a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os_u, px)) {
                pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                               a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os_v, pxv);
                if ((px.uc + a.max_sc < a.n_u) && (px.vc + a.max_sc < a.n_v) &&
                    (px.uc - a.max_sc >= 0) && (px.vc - a.max_sc >= 0)) {
                    // |du| <= os_u/2 + 1 by construction of the offset; kept as a guard
                    if (px.du < -t.hu || px.du > t.hu || pxv.dv < -t.hv || pxv.dv > t.hv) {
                        atomicOr(a.error, 2);
                    } else {
                        uc = (int)px.uc;
                        vc = (int)px.vc;
                        du = (int)px.du;
                        dv = (int)pxv.dv;
                        fidx = f_i;
                    }
                }
            }
            if (fidx < 0) {
                V *out = vis + 2 * ((row * a.n_chan + ch) * a.n_pol);
                for (int64_t p = 0; p < a.n_pol; p++) {
                    out[2 * p] = (V)0;
                    out[2 * p + 1] = (V)0;
                }
            }
        }
        const int64_t a_chan = a.chan_map[ch];
        const int64_t cfbc = (a.cfb[b] * a.n_cfc + a.cfc[ch]) * a.n_cfp;
        for (int64_t p0 = 0; p0 < a.n_pol; p0 += NP) {
            const int64_t cf = cfbc + a.cfp[p0];
            const int su = a.support[2 * cf], sv = a.support[2 * cf + 1];
            const int scu = su / 2, scv = sv / 2;
            const int n_tap = su * sv;
            if (su != tab_su || sv != tab_sv) {   // wave-uniform
                __builtin_amdgcn_wave_barrier();
                for (int j = lane; j < n_tap; j += 64) {
                    const int ru = j / sv, rv = j - ru * sv;
                    taps[j] = make_int4(2 * ru, 2 * rv, 2 * (ru * (int)a.n_v + rv), 0);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                tab_su = su;
                tab_sv = sv;
            }
            // this lane's integration: its taps' block, factor rows and
            // footprint corner (32-bit element offsets; the host checks the
            // table and plane sizes)
            int koff = 0, puo = 0, pvo = 0, goff = 0;
            if (fidx >= 0) {
                koff = (int)(t.cf_base[cf] + ((du + t.hu) * nwv + dv + t.hv) * n_tap);
                puo = 2 * ((fidx * nwu + du + t.hu) * t.W + t.R / 2 - scu);
                pvo = 2 * ((fidx * nwv + dv + t.hv) * t.W + t.R / 2 - scv);
                goff = 2 * ((uc - scu) * (int)a.n_v + vc - scv);
            }
            const G *gpl[NP];
#pragma unroll
            for (int q = 0; q < NP; q++)
                gpl[q] = grid + 2 * ((a_chan * a.n_imag_pol + a.pol_map[p0 + q]) * plane);
            for (int rd = 0; rd < 16; rd++) {
                const int src = 4 * rd + grp;
                const int s_f = __shfl(fidx, src);
                const int s_k = __shfl(koff, src), s_g = __shfl(goff, src);
                const int s_pu = __shfl(puo, src), s_pv = __shfl(pvo, src);
                if (s_f < 0) continue;   // the whole group skips
                const double *kb = ckr + s_k;
                const double *pu = pgu + s_pu;
                const double *pv = pgv + s_pv;
                const G *gs[NP];
#pragma unroll
                for (int q = 0; q < NP; q++) gs[q] = gpl[q] + s_g;
                double sr[NP], si[NP], nr = 0.0;
#pragma unroll
                for (int q = 0; q < NP; q++) sr[q] = si[q] = 0.0;
                auto tap = [&](int j) {
                    const int4 o = taps[j];
                    const double k = kb[j];
                    const double ur = pu[o.x], ui = pu[o.x + 1];
                    const double vr = pv[o.y], vi = pv[o.y + 1];
                    const double cr = k * fma(ur, vr, -ui * vi);
                    const double ci = k * fma(ur, vi, ui * vr);
                    nr += k;
#pragma unroll
                    for (int q = 0; q < NP; q++) {   // conj(c) g
                        const double gr = (double)gs[q][o.z], gi = (double)gs[q][o.z + 1];
                        sr[q] = fma(cr, gr, sr[q]);
                        sr[q] = fma(ci, gi, sr[q]);
                        si[q] = fma(cr, gi, si[q]);
                        si[q] = fma(-ci, gr, si[q]);
                    }
                };
                // full passes of 16 taps (a wave-uniform count), then the rest
                const int n_full = n_tap >> 4;
#pragma unroll 2
                for (int it = 0; it < n_full; it++) tap(l + 16 * it);
                if (l + 16 * n_full < n_tap) tap(l + 16 * n_full);
                nr = group16_sum(nr);
#pragma unroll
                for (int q = 0; q < NP; q++) {
                    sr[q] = group16_sum(sr[q]);
                    si[q] = group16_sum(si[q]);
                }
                if (l == 0) {
                    double *o = res + src * (2 * NP + 2);
#pragma unroll
                    for (int q = 0; q < NP; q++) {
                        o[2 * q] = sr[q];
                        o[2 * q + 1] = si[q];
                    }
                    o[2 * NP] = nr;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (fidx >= 0) {
                const double *o = res + lane * (2 * NP + 2);
                const double nr = o[2 * NP];
                const int64_t row = tm * a.n_baseline + b;
                V *out = vis + 2 * ((row * a.n_chan + ch) * a.n_pol + p0);
#pragma unroll
                for (int q = 0; q < NP; q++) {
                    double re = o[2 * q], im = o[2 * q + 1];
                    if (a.normalize && nr != 0.0) {   // the kernel's own tap sum
                        re = re / nr;
                        im = im / nr;
                    }
                    out[2 * q] = (V)re;
                    out[2 * q + 1] = (V)im;
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

}  // namespace ngc

extern "C" int ngc_aperture_degrid(ngc_context *ctx, void *stream, const double *uvw,
                                   const int32_t *field, const int32_t *field_id,
                                   int64_t n_field, const double *freq_chan,
                                   const int32_t *chan_map, const int32_t *pol_map,
                                   const int32_t *cf_baseline_map, const int32_t *cf_chan_map,
                                   const int32_t *cf_pol_map, const double *conv_kernel,
                                   int64_t n_cfb, int64_t n_cfc, int64_t n_cfp, int64_t k_u,
                                   int64_t k_v, const int32_t *support,
                                   const void *phase_gradient, int64_t n_time,
                                   int64_t n_baseline, int64_t n_chan, int64_t n_pol,
                                   int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
                                   int64_t n_v, double delta_l, double delta_m, int64_t os_u,
                                   int64_t os_v, const void *grid, int grid_dtype, void *vis,
                                   int vis_dtype, int normalize) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(n_time >= 0 && n_baseline >= 0 && n_chan >= 1 && n_pol >= 1 &&
                  n_imag_chan >= 1 && n_imag_pol >= 1 && n_u >= 1 && n_v >= 1 && n_field >= 1 &&
                  n_cfb >= 1 && n_cfc >= 1 && n_cfp >= 1 && k_u >= 1 && k_v >= 1 &&
                  os_u >= 1 && os_v >= 1,
              NGC_ERR_ARG, "bad dimensions");
    NGC_CHECK(freq_chan && chan_map && pol_map && field_id && cf_baseline_map && cf_chan_map &&
                  cf_pol_map && support,
              NGC_ERR_ARG, "NULL parameter array");
    NGC_CHECK(grid_dtype == NGC_C64 || grid_dtype == NGC_C128, NGC_ERR_ARG,
              "grid dtype must be C64 or C128");
    NGC_CHECK(vis_dtype == NGC_C64 || vis_dtype == NGC_C128, NGC_ERR_ARG,
              "vis dtype must be C64 or C128");
    for (int64_t c = 0; c < n_chan; c++) {
        NGC_CHECK(chan_map[c] >= 0 && chan_map[c] < n_imag_chan, NGC_ERR_ARG,
                  "chan_map entry out of range");
        NGC_CHECK(cf_chan_map[c] >= 0 && cf_chan_map[c] < n_cfc, NGC_ERR_ARG,
                  "cf_chan_map entry out of range");
    }
    for (int64_t p = 0; p < n_pol; p++) {
        NGC_CHECK(pol_map[p] >= 0 && pol_map[p] < n_imag_pol, NGC_ERR_ARG,
                  "pol_map entry out of range");
        NGC_CHECK(cf_pol_map[p] >= 0 && cf_pol_map[p] < n_cfp, NGC_ERR_ARG,
                  "cf_pol_map entry out of range");
    }
    for (int64_t b = 0; b < n_baseline; b++)
        NGC_CHECK(cf_baseline_map[b] >= 0 && cf_baseline_map[b] < n_cfb, NGC_ERR_ARG,
                  "cf_baseline_map entry out of range");
    int64_t max_sc = 0;
    for (int64_t i = 0; i < n_cfb * n_cfc * n_cfp * 2; i++) {
        NGC_CHECK(support[i] >= 1, NGC_ERR_ARG, "support entries must be >= 1");
        max_sc = std::max<int64_t>(max_sc, support[i]);
    }
    if (n_time == 0 || n_baseline == 0) return NGC_OK;
    NGC_CHECK(uvw && field && conv_kernel && phase_gradient && grid && vis, NGC_ERR_ARG,
              "NULL device array");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    std::vector<double> scale;
    uv_scale_host(freq_chan, n_chan, delta_l, delta_m, n_u, n_v, scale);
    void *d_scale, *d_fid, *d_cm, *d_pm, *d_cfb, *d_cfc, *d_cfp, *d_sup;
    NGC_TRY(upload(ctx, s, "adg_scale", scale.data(), scale.size() * sizeof(double), &d_scale));
    NGC_TRY(upload(ctx, s, "adg_field_id", field_id, n_field * sizeof(int32_t), &d_fid));
    NGC_TRY(upload(ctx, s, "adg_chan_map", chan_map, n_chan * sizeof(int32_t), &d_cm));
    NGC_TRY(upload(ctx, s, "adg_pol_map", pol_map, n_pol * sizeof(int32_t), &d_pm));
    NGC_TRY(upload(ctx, s, "adg_cfb", cf_baseline_map, n_baseline * sizeof(int32_t), &d_cfb));
    NGC_TRY(upload(ctx, s, "adg_cfc", cf_chan_map, n_chan * sizeof(int32_t), &d_cfc));
    NGC_TRY(upload(ctx, s, "adg_cfp", cf_pol_map, n_pol * sizeof(int32_t), &d_cfp));
    NGC_TRY(upload(ctx, s, "adg_support", support, n_cfb * n_cfc * n_cfp * 2 * sizeof(int32_t),
                   &d_sup));
    DegridApArgs a;
    a.uvw = uvw;
    a.field = field;
    a.field_id = (const int *)d_fid;
    a.n_field = (int)n_field;
    a.uv_scale = (const double *)d_scale;
    a.chan_map = (const int *)d_cm;
    a.pol_map = (const int *)d_pm;
    a.cfb = (const int *)d_cfb;
    a.cfc = (const int *)d_cfc;
    a.cfp = (const int *)d_cfp;
    a.support = (const int *)d_sup;
    a.ck = conv_kernel;
    a.pg = (const double2 *)phase_gradient;
    a.n_cfc = n_cfc;
    a.n_cfp = n_cfp;
    a.k_u = k_u;
    a.k_v = k_v;
    a.n_time = n_time;
    a.n_baseline = n_baseline;
    a.n_chan = n_chan;
    a.n_pol = n_pol;
    a.n_imag_pol = n_imag_pol;
    a.n_u = n_u;
    a.n_v = n_v;
    a.os_u = os_u;
    a.os_v = os_v;
    a.max_sc = max_sc;
    a.normalize = normalize ? 1 : 0;
    a.error = ctx->d_error;
    const int64_t n_thr = n_baseline * n_chan * ceil_div(n_time, 64) * 64;
    const int nb = pblocks(n_thr);
    profile_new_call(ctx, s);

    // the separable form needs PG = PGu PGv and re-laid taps of modest size
    AdgTab t;
    t.hu = (int)(os_u / 2 + 1);
    t.hv = (int)(os_v / 2 + 1);
    t.R = (int)max_sc;
    t.W = t.R + 1;
    t.n_cf = n_cfb * n_cfc * n_cfp;
    const int64_t elem = 8;
    const int64_t nwo = (int64_t)(2 * t.hu + 1) * (2 * t.hv + 1);
    std::vector<int64_t> cf_base(t.n_cf);
    int64_t n_ckr = 0;
    t.max_block = 1;
    for (int64_t cf = 0; cf < t.n_cf; cf++) {
        const int64_t blk = (int64_t)support[2 * cf] * support[2 * cf + 1];
        cf_base[cf] = n_ckr;
        n_ckr += nwo * blk;
        t.max_block = std::max<int>(t.max_block, (int)blk);
    }
    // the tap-offset tables take max_block int4s per wavefront in LDS, and the
    // offsets are 32-bit
    bool sep = n_ckr * elem <= (int64_t(1) << 30) && t.max_block <= 1024 &&
               2 * n_u * n_v < (int64_t(1) << 31) &&
               2 * n_field * (2 * t.hu + 3) * (int64_t)(max_sc + 1) < (int64_t(1) << 31) &&
               2 * n_field * (2 * t.hv + 3) * (int64_t)(max_sc + 1) < (int64_t(1) << 31);
    if (sep) {
        NGC_HIP(hipMemsetAsync(ctx->d_error + 1, 0, sizeof(int), s));
        adg_pg_check_kernel<<<pblocks(n_field * k_u * k_v), 256, 0, s>>>(a, ctx->d_error + 1);
        NGC_HIP(hipGetLastError());
        int bad = 0;
        NGC_HIP(hipMemcpyAsync(&bad, ctx->d_error + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        NGC_HIP(hipStreamSynchronize(s));
        sep = bad == 0;
    }
    NGC_HIP(hipMemsetAsync(ctx->d_error, 0, sizeof(int), s));
    if (sep) {
        const int64_t lu = (2 * t.hu + 1) * t.W, lv = (2 * t.hv + 1) * t.W;
        void *d_ckr, *d_pgu, *d_pgv;
        NGC_TRY(workspace(ctx, "adg_ckr", n_ckr * elem, &d_ckr));
        NGC_TRY(workspace(ctx, "adg_pgu", n_field * lu * 2 * elem, &d_pgu));
        NGC_TRY(workspace(ctx, "adg_pgv", n_field * lv * 2 * elem, &d_pgv));
        void *d_cfbase;
        NGC_TRY(upload(ctx, s, "adg_cf_base", cf_base.data(), t.n_cf * sizeof(int64_t),
                       &d_cfbase));
        t.cf_base = (const int64_t *)d_cfbase;
        bool shared = true;
        for (int64_t p = 1; p < n_pol; p++) shared = shared && cf_pol_map[p] == cf_pol_map[0];
        const int np = shared && (n_pol == 1 || n_pol == 2 || n_pol == 4) ? (int)n_pol : 1;
        phase_begin(ctx, s, "aperture_degrid_tables");
        adg_taps_kernel<<<pblocks(t.n_cf * nwo * t.max_block), 256, 0, s>>>(a, t,
                                                                           (double *)d_ckr);
        adg_pg_factor_kernel<<<pblocks(n_field * (lu + lv)), 256, 0, s>>>(a, t, (double *)d_pgu,
                                                                           (double *)d_pgv);
        NGC_HIP(hipGetLastError());
        phase_end(ctx, s);
        phase_begin(ctx, s, "aperture_degrid");
        auto launch = [&](auto g, auto v, auto npc) {
            using G = decltype(g);
            using V = decltype(v);
            constexpr int NPv = decltype(npc)::value;
            aperture_degrid_taps_kernel<G, V, NPv>
                <<<nb, 256, 4 * (t.max_block * sizeof(int4) + 64 * (2 * NPv + 2) * sizeof(double)),
                   s>>>(
                a, t, (const double *)d_ckr, (const double *)d_pgu, (const double *)d_pgv,
                (const G *)grid, (V *)vis);
        };
        auto by_np = [&](auto g, auto v) {
            if (np == 4)
                launch(g, v, std::integral_constant<int, 4>{});
            else if (np == 2)
                launch(g, v, std::integral_constant<int, 2>{});
            else
                launch(g, v, std::integral_constant<int, 1>{});
        };
        if (grid_dtype == NGC_C64 && vis_dtype == NGC_C64)
            by_np(float{}, float{});
        else if (grid_dtype == NGC_C64)
            by_np(float{}, double{});
        else if (vis_dtype == NGC_C64)
            by_np(double{}, float{});
        else
            by_np(double{}, double{});
    } else {
        phase_begin(ctx, s, "aperture_degrid");
        if (grid_dtype == NGC_C64 && vis_dtype == NGC_C64)
            aperture_degrid_kernel<float, float><<<nb, 256, 0, s>>>(a, (const float *)grid,
                                                                    (float *)vis);
        else if (grid_dtype == NGC_C64)
            aperture_degrid_kernel<float, double><<<nb, 256, 0, s>>>(a, (const float *)grid,
                                                                     (double *)vis);
        else if (vis_dtype == NGC_C64)
            aperture_degrid_kernel<double, float><<<nb, 256, 0, s>>>(a, (const double *)grid,
                                                                     (float *)vis);
        else
            aperture_degrid_kernel<double, double><<<nb, 256, 0, s>>>(a, (const double *)grid,
                                                                      (double *)vis);
    }
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    int err = 0;
    NGC_HIP(hipMemcpyAsync(&err, ctx->d_error, sizeof(int), hipMemcpyDeviceToHost, s));
    NGC_HIP(hipStreamSynchronize(s));
    NGC_CHECK(!(err & 2), NGC_ERR_UNSUPPORTED, "sub-cell offset outside the re-laid taps");
    NGC_CHECK(err == 0, NGC_ERR_FIELD, "a row's field id is not in field_id");
    return NGC_OK;
}

extern "C" int ngc_set_degrid_mode(ngc_context *ctx, int mode) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(mode == 0 || mode == 1, NGC_ERR_ARG, "mode must be 0 (gather) or 1 (tiled)");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->degrid_tiled = mode;
    return NGC_OK;
}
