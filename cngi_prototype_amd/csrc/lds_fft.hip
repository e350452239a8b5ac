// Single-precision 2-D inverse FFT of gridded planes for the image transform
// (make_image.py:116-130: ifftshift, ifft2, fftshift, crop and correction),
// for axis lengths whose prime factors are at most 13 (C2: 2457 = 3^3 7 13).
//
//   fft_rows_kernel      one workgroup per row: the row is loaded (through the
//                        ifftshift, for the first pass) into LDS and
//                        transformed there by a mixed-radix Stockham FFT
//                        (radices 8, 4, 2, 13, 11, 7, 5, 3): stage s with
//                        radix R reads R values N/R apart, twiddles them by
//                        the powers of exp(i 2 pi k / (Ns R)) (Ns = product of
//                        the earlier radices, k = j mod Ns), takes their
//                        R-point DFT against the R-th roots of unity held in
//                        registers, and writes the results Ns apart from
//                        (j - k) R + k; two LDS buffers alternate.  The
//                        twiddles come from one table exp(i 2 pi t / N), t < N,
//                        computed in FP64 on the host.
//   fft_transpose_kernel the planes transposed through 32 x 32 LDS tiles, so
//                        that the second axis is again a row pass
//   fft_crop_kernel      crop to the image, fftshift, real part, sum_weight and
//                        correction (as crop_correct_kernel), reading the
//                        transposed planes through LDS tiles
//
// Passes over the planes: rows (read + write), transpose (read + write), rows
// (read + write), crop (reads the image region); the library transform it
// replaces took 66 ms for 256 planes of 2457^2, about 9 % of HBM.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ngc_internal.hpp"

namespace ngc {

constexpr int kFftMaxN = 8192;
constexpr int kFftMaxStages = 16;
constexpr int kFftBlock = 256;

struct FftRadices {
    int n, count;
    int r[kFftMaxStages];
};

static bool fft_factor(int64_t n, FftRadices &f) {
    if (n < 2 || n > kFftMaxN) return false;
    f.n = (int)n;
    f.count = 0;
    int64_t m = n;
    for (int r : {8, 4, 2, 13, 11, 7, 5, 3}) {
        while (m % r == 0) {
            if (f.count == kFftMaxStages) return false;
            f.r[f.count++] = r;
            m /= r;
        }
    }
    return m == 1;
}

bool lds_fft_supported(int64_t n) {
    FftRadices f;
    return fft_factor(n, f);
}

__device__ inline float2 fft_cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One Stockham stage of radix R over the N values in A, written to B.
template <int R>
__device__ void fft_stage(const float2 *__restrict__ A, float2 *__restrict__ B, int N, int Ns,
                          const float2 *__restrict__ tw) {
    const int NR = N / R;
    const int step = NR / Ns;  // N / (Ns R): twiddle index per unit of r k
    float2 w[R];               // w[t] = exp(i 2 pi t / R) (sign of the table)
#pragma unroll
    for (int t = 0; t < R; t++) w[t] = tw[t * NR];
    for (int j = threadIdx.x; j < NR; j += blockDim.x) {
        const int k = j % Ns;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = A[j + r * NR];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = fft_cmul(v[r], tw[r * k * step]);
        }
        const int o = (j - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) {
            float2 acc = v[0];
#pragma unroll
            for (int m = 1; m < R; m++) {
                const float2 c = w[(m * q) % R];
                acc.x += v[m].x * c.x - v[m].y * c.y;
                acc.y += v[m].x * c.y + v[m].y * c.x;
            }
            B[o + q * Ns] = acc;
        }
    }
}

// The whole transform of the N values in buffer A (scratch B); returns the
// buffer that holds the result.
__device__ float2 *fft_run(float2 *A, float2 *B, const FftRadices &f,
                           const float2 *__restrict__ tw) {
    int Ns = 1;
    for (int s = 0; s < f.count; s++) {
        const int R = f.r[s];
        switch (R) {
        case 2: fft_stage<2>(A, B, f.n, Ns, tw); break;
        case 3: fft_stage<3>(A, B, f.n, Ns, tw); break;
        case 4: fft_stage<4>(A, B, f.n, Ns, tw); break;
        case 5: fft_stage<5>(A, B, f.n, Ns, tw); break;
        case 7: fft_stage<7>(A, B, f.n, Ns, tw); break;
        case 8: fft_stage<8>(A, B, f.n, Ns, tw); break;
        case 11: fft_stage<11>(A, B, f.n, Ns, tw); break;
        default: fft_stage<13>(A, B, f.n, Ns, tw); break;
        }
        __syncthreads();
        float2 *t = A;
        A = B;
        B = t;
        Ns *= R;
    }
    return A;
}

// IN: 0 complex rows of `in` (transformed in place is allowed), 1 complex64
// grid, 2 float32 grid; for 1 and 2 the row is read through the ifftshift:
// A[ku][kv] = grid[(ku + n_u/2) mod n_u][(kv + N/2) mod N].
template <int IN>
__global__ void __launch_bounds__(kFftBlock)
fft_rows_kernel(const float *in, float2 *out, int64_t n_rows, int64_t n_u, FftRadices f,
                const float2 *__restrict__ tw) {
    extern __shared__ float2 s_fft[];
    const int N = f.n;
    for (int64_t row = blockIdx.x; row < n_rows; row += gridDim.x) {
        float2 *A = s_fft, *B = s_fft + N;
        if (IN == 0) {
            const float2 *src = reinterpret_cast<const float2 *>(in) + row * N;
            for (int v = threadIdx.x; v < N; v += blockDim.x) A[v] = src[v];
        } else {
            const int64_t pl = row / n_u, ku = row - pl * n_u;
            int64_t su = ku + n_u / 2;
            if (su >= n_u) su -= n_u;
            const int64_t base = (pl * n_u + su) * N;
            const int half = N / 2;
            for (int v = threadIdx.x; v < N; v += blockDim.x) {
                int sv = v + half;
                if (sv >= N) sv -= N;
                if (IN == 1)
                    A[v] = reinterpret_cast<const float2 *>(in)[base + sv];
                else
                    A[v] = make_float2(in[base + sv], 0.f);
            }
        }
        __syncthreads();
        const float2 *res = fft_run(A, B, f, tw);
        float2 *dst = out + row * N;
        for (int v = threadIdx.x; v < N; v += blockDim.x) dst[v] = res[v];
        __syncthreads();  // the row's LDS is free for the next one
    }
}

// out[pl][c][r] = in[pl][r][c], [rows x cols] planes, 32 x 32 tiles
__global__ void __launch_bounds__(256)
fft_transpose_kernel(const float2 *__restrict__ in, float2 *__restrict__ out, int64_t n_planes,
                     int64_t rows, int64_t cols) {
    __shared__ float2 tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    const int64_t c0 = (int64_t)blockIdx.x * 32, r0 = (int64_t)blockIdx.y * 32;
    for (int64_t pl = blockIdx.z; pl < n_planes; pl += gridDim.z) {
        const float2 *ip = in + pl * rows * cols;
        float2 *op = out + pl * rows * cols;
        for (int k = ty; k < 32; k += 8) {
            const int64_t r = r0 + k, c = c0 + tx;
            if (r < rows && c < cols) tile[k][tx] = ip[r * cols + c];
        }
        __syncthreads();
        for (int k = ty; k < 32; k += 8) {
            const int64_t c = c0 + k, r = r0 + tx;
            if (c < cols && r < rows) op[c * rows + r] = tile[tx][k];
        }
        __syncthreads();
    }
}

// image[pl][l][m] = ((Re T[pl][kv(m)][ku(l)]) / sw') / (corr_l[l] corr_m[m]) with
// ku(l) = (l - n_l/2) mod n_u, kv(m) = (m - n_m/2) mod n_v and T the transposed
// (v-major) transform, as crop_correct_kernel computes it from the u-major one
__global__ void __launch_bounds__(256)
fft_crop_kernel(const float2 *__restrict__ T, int64_t n_planes, int64_t n_u, int64_t n_v,
                int64_t n_l, int64_t n_m, const double *sum_weight, const double *corr,
                double *image) {
#pragma clang fp contract(off)
    __shared__ float tile[32][33];   // [m - m0][l - l0]
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int64_t l0 = (int64_t)blockIdx.x * 32, m0 = (int64_t)blockIdx.y * 32;
    for (int64_t pl = blockIdx.z; pl < n_planes; pl += gridDim.z) {
        const float2 *tp = T + pl * n_u * n_v;
        for (int k = ty; k < 32; k += 8) {
            const int64_t m = m0 + k, l = l0 + tx;
            if (m < n_m && l < n_l) {
                int64_t kv = m - n_m / 2;
                if (kv < 0) kv += n_v;
                int64_t ku = l - n_l / 2;
                if (ku < 0) ku += n_u;
                tile[k][tx] = tp[kv * n_u + ku].x;
            }
        }
        __syncthreads();
        double sw = 1.0;
        if (sum_weight) {
            sw = sum_weight[pl];
            if (sw == 0.0) sw = 1.0;
        }
        for (int k = ty; k < 32; k += 8) {
            const int64_t l = l0 + k, m = m0 + tx;
            if (l < n_l && m < n_m) {
                double x = (double)tile[tx][k];
                if (sum_weight) x = x / sw;
                if (corr) x = x / (corr[l] * corr[n_l + m]);
                image[(pl * n_l + l) * n_m + m] = x;
            }
        }
        __syncthreads();
    }
}

static int fft_twiddles(ngc_context *ctx, hipStream_t s, int64_t n, int sign, const char *name,
                        const float2 **out) {
    std::vector<float> t(2 * (size_t)n);
    for (int64_t k = 0; k < n; k++) {
        const double a = 2.0 * M_PI * (double)k / (double)n;
        t[2 * k] = (float)std::cos(a);
        t[2 * k + 1] = (float)(sign * std::sin(a));
    }
    void *p;
    NGC_TRY(upload(ctx, s, name, t.data(), t.size() * sizeof(float), &p));
    *out = (const float2 *)p;
    return NGC_OK;
}

template <int IN>
static int fft_rows(hipStream_t s, const float *in, float2 *out, int64_t n_rows, int64_t n_u,
                    const FftRadices &f, const float2 *tw) {
    const size_t lds = 2 * (size_t)f.n * sizeof(float2);
    auto kern = fft_rows_kernel<IN>;
    NGC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    const int64_t blocks = std::min<int64_t>(n_rows, int64_t(1) << 20);
    kern<<<(unsigned)blocks, kFftBlock, lds, s>>>(in, out, n_rows, n_u, f, tw);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

int lds_grid_to_image(ngc_context *ctx, hipStream_t s, const void *grid, int dt,
                      int64_t n_planes, int64_t n_u, int64_t n_v, int64_t n_l, int64_t n_m,
                      const double *sum_weight, const double *d_corr, double *image) {
    FftRadices fu, fv;
    NGC_CHECK(fft_factor(n_u, fu) && fft_factor(n_v, fv), NGC_ERR_UNSUPPORTED,
              "axis length not a product of 2, 3, 5, 7, 11, 13 up to 8192");
    NGC_CHECK(dt == NGC_C64 || dt == NGC_F32, NGC_ERR_ARG, "single-precision grids only");
    const size_t cells = (size_t)n_planes * n_u * n_v;
    void *b0, *b1;
    NGC_TRY(workspace(ctx, "fft_buffer", cells * sizeof(float2), &b0));
    NGC_TRY(workspace(ctx, "fft_transposed", cells * sizeof(float2), &b1));
    const float2 *tw_v, *tw_u;
    NGC_TRY(fft_twiddles(ctx, s, n_v, +1, "fft_tw_rows", &tw_v));
    NGC_TRY(fft_twiddles(ctx, s, n_u, +1, "fft_tw_cols", &tw_u));
    phase_begin(ctx, s, "fft_rows");
    if (dt == NGC_C64)
        NGC_TRY(fft_rows<1>(s, (const float *)grid, (float2 *)b0, n_planes * n_u, n_u, fv, tw_v));
    else
        NGC_TRY(fft_rows<2>(s, (const float *)grid, (float2 *)b0, n_planes * n_u, n_u, fv, tw_v));
    phase_end(ctx, s);
    phase_begin(ctx, s, "fft_transpose");
    const dim3 tg((unsigned)ceil_div(n_v, 32), (unsigned)ceil_div(n_u, 32),
                  (unsigned)std::min<int64_t>(n_planes, 65535));
    fft_transpose_kernel<<<tg, 256, 0, s>>>((const float2 *)b0, (float2 *)b1, n_planes, n_u, n_v);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    phase_begin(ctx, s, "fft_cols");
    NGC_TRY(fft_rows<0>(s, (const float *)b1, (float2 *)b1, n_planes * n_v, n_v, fu, tw_u));
    phase_end(ctx, s);
    phase_begin(ctx, s, "crop_correct");
    const dim3 cg((unsigned)ceil_div(n_l, 32), (unsigned)ceil_div(n_m, 32),
                  (unsigned)std::min<int64_t>(n_planes, 65535));
    fft_crop_kernel<<<cg, 256, 0, s>>>((const float2 *)b1, n_planes, n_u, n_v, n_l, n_m,
                                       sum_weight, d_corr, image);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}

}  // namespace ngc
