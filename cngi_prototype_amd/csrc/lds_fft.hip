// Single-precision 2-D FFTs of gridded planes for axis lengths whose prime
// factors are at most 13 (a 2048^2 image padded by 1.2: 2457 = 3^3 7 13):
// the inverse image transform (make_image.py:116-130: ifftshift, ifft2,
// fftshift, crop and correction) and the forward placement of a model image
// for prediction.
//
//   fft_rows_kernel      one workgroup per row: the row is loaded (through the
//                        ifftshift, for the first pass) into LDS and
//                        transformed there by a mixed-radix Stockham FFT
//                        (radices 8, 4, 2, 13, 11, 9, 7, 5, 3): stage s with
//                        radix R reads R values N/R apart, twiddles them by
//                        the powers of exp(i 2 pi k / (Ns R)) (Ns = product of
//                        the earlier radices, k = j mod Ns by a multiply-high
//                        with a host-computed reciprocal), takes their R-point
//                        DFT with the R-th roots of unity as compile-time
//                        constants (odd R in the paired form, (R-1)^2 real
//                        multiply-adds instead of 4 (R-1)^2; R = 2, 4, 8 as
//                        radix-2 butterflies) and writes the results Ns apart
//                        from (j - k) R + k; two LDS buffers alternate.  The
//                        twiddles come from one table exp(i 2 pi t / N), t < N,
//                        computed in FP64 on the host.
//   fft_cols_a_kernel,   the columns without a transpose (four-step, N = N1 N2):
//   fft_cols_b_kernel    step A transforms the N1 rows r1 N2 + r2 of a
//                        32-column block in LDS and twiddles them in place;
//                        step B transforms N2 consecutive rows and writes the
//                        result rows k1 + N1 k2 straight into the cropped,
//                        corrected image (or, forward, the fftshifted grid)
//   fft_transpose_kernel for column lengths with no split N1 <= N2 <= 128: the
//                        planes transposed through 64 x 64 LDS tiles, so that
//                        the second axis is again a row pass, then
//   fft_crop_kernel      crop to the image, fftshift, real part, sum_weight and
//                        correction (as crop_correct_kernel), reading the
//                        transposed planes through LDS tiles
//
// Passes over the planes: rows (read + write), step A (read + write), step B
// (read, and the image written); rocFFT took 66 ms for the transform of 256
// planes of 2457^2, about 9 % of HBM.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ngc_internal.hpp"

namespace ngc {

constexpr int kFftMaxN = 8192;
constexpr int kFftMaxStages = 16;
constexpr int kFftBlock = 256;

// Stage s: radix r, Ns = product of the earlier radices, step = N / (Ns r)
// (twiddle table stride), magic = floor(2^32 / Ns) + 1, so that
// j / Ns = umulhi(j, magic) for every j < N <= 8192 (j Ns < 2^32).
struct FftStage {
    int r, ns, step;
    uint32_t magic;
};

struct FftRadices {
    int n, count;
    FftStage st[kFftMaxStages];
};

static bool fft_factor(int64_t n, FftRadices &f) {
    if (n < 2 || n > kFftMaxN) return false;
    f.n = (int)n;
    f.count = 0;
    int64_t m = n, ns = 1;
    for (int r : {8, 4, 2, 13, 11, 9, 7, 5, 3}) {
        while (m % r == 0) {
            if (f.count == kFftMaxStages) return false;
            FftStage &s = f.st[f.count++];
            s.r = r;
            s.ns = (int)ns;
            s.step = (int)(n / (ns * r));
            s.magic = ns > 1 ? (uint32_t)((uint64_t(1) << 32) / (uint64_t)ns + 1) : 0u;
            ns *= r;
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

// cos and sin of 2 pi t / R in double, by Taylor series on [-pi, pi], so that
// the R-th roots of unity are compile-time constants of the butterflies
constexpr double kPi = 3.14159265358979323846;

constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 30; k++) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}

constexpr double taylor_sin(double x) {
    double term = x, sum = x;
    for (int k = 1; k < 30; k++) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}

// c[t] + i s[t] = exp(SIGN i 2 pi t / R)
template <int R, int SIGN>
struct DftRoots {
    float c[R], s[R];
    constexpr DftRoots() : c(), s() {
        for (int t = 0; t < R; t++) {
            double x = 2.0 * kPi * t / R;
            if (x > kPi) x -= 2.0 * kPi;
            c[t] = (float)taylor_cos(x);
            s[t] = (float)(SIGN * taylor_sin(x));
        }
    }
};

// y_q = sum_m v_m exp(SIGN i 2 pi m q / R), in place.  Odd R pairs m with
// R - m: with a_m = v_m + v_{R-m}, b_m = v_m - v_{R-m}, P = v_0 + sum a_m c_{mq}
// and Q = sum s_{mq} (i b_m), y_q = P + Q and y_{R-q} = P - Q.  R = 2, 4, 8
// are radix-2 butterflies whose multiplications by +-i are swaps.
template <int SIGN>
__device__ inline float2 fft_mul_i(float2 z) {   // z exp(SIGN i pi / 2)
    return SIGN > 0 ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
}

template <int R, int SIGN>
__device__ inline void fft_dft(float2 (&v)[R]) {
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = make_float2(a.x + b.x, a.y + b.y);
        v[1] = make_float2(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 4) {
        const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
        const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
        const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
        const float2 d13 = fft_mul_i<SIGN>(make_float2(v[1].x - v[3].x, v[1].y - v[3].y));
        v[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
        v[2] = make_float2(s02.x - s13.x, s02.y - s13.y);
        v[1] = make_float2(d02.x + d13.x, d02.y + d13.y);
        v[3] = make_float2(d02.x - d13.x, d02.y - d13.y);
    } else if constexpr (R == 8) {
        float2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        fft_dft<4, SIGN>(e);
        fft_dft<4, SIGN>(o);
        constexpr float h = 0.70710678118654752f;
        // o_q exp(SIGN i 2 pi q / 8)
        o[1] = make_float2(h * (o[1].x - SIGN * o[1].y), h * (o[1].y + SIGN * o[1].x));
        o[2] = fft_mul_i<SIGN>(o[2]);
        o[3] = make_float2(h * (-o[3].x - SIGN * o[3].y), h * (-o[3].y + SIGN * o[3].x));
#pragma unroll
        for (int q = 0; q < 4; q++) {
            v[q] = make_float2(e[q].x + o[q].x, e[q].y + o[q].y);
            v[q + 4] = make_float2(e[q].x - o[q].x, e[q].y - o[q].y);
        }
    } else {
        static_assert(R % 2 == 1, "odd radix");
        constexpr int H = (R - 1) / 2;
        constexpr DftRoots<R, SIGN> W{};
        float2 a[H], b[H];
#pragma unroll
        for (int m = 1; m <= H; m++) {
            a[m - 1] = make_float2(v[m].x + v[R - m].x, v[m].y + v[R - m].y);
            b[m - 1] = make_float2(v[m].x - v[R - m].x, v[m].y - v[R - m].y);
        }
        const float2 v0 = v[0];
        float2 y0 = v0;
#pragma unroll
        for (int m = 0; m < H; m++) y0 = make_float2(y0.x + a[m].x, y0.y + a[m].y);
        v[0] = y0;
#pragma unroll
        for (int q = 1; q <= H; q++) {
            float2 P = v0, Q = make_float2(0.f, 0.f);
#pragma unroll
            for (int m = 1; m <= H; m++) {
                const float c = W.c[(m * q) % R], s = W.s[(m * q) % R];
                P.x += a[m - 1].x * c;
                P.y += a[m - 1].y * c;
                Q.x -= b[m - 1].y * s;
                Q.y += b[m - 1].x * s;
            }
            v[q] = make_float2(P.x + Q.x, P.y + Q.y);
            v[R - q] = make_float2(P.x - Q.x, P.y - Q.y);
        }
    }
}

// One Stockham stage of radix R over the N values of one transform, written
// from A to B: the butterfly j (k = j mod Ns) reads A[j + r N/R], twiddles
// input r by tw[r k step TS] = exp(SIGN i 2 pi r k / (Ns R)) and writes output
// q to B[(j - k) R + k + q Ns].  Value e of the transform sits at LDS index
// e STR + c, and the transform's L threads take butterflies lt, lt + L, ...
// (rows: L = 256, STR = 1, c = 0; column blocks: L = 8 threads per column of
// STR = 32 columns).  TS is the stride of the twiddle table: a table for N TS
// values serves a transform of N.
template <int R, int SIGN, int L, int STR>
__device__ void fft_stage(const float2 *__restrict__ A, float2 *__restrict__ B, int N,
                          const FftStage st, const float2 *__restrict__ tw, int TS, int lt,
                          int c) {
    const int NR = N / R;
    const int Ns = st.ns;
    for (int j = lt; j < NR; j += L) {
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = A[(j + r * NR) * STR + c];
        int jd = j, k = 0;
        if (Ns > 1) {
            jd = (int)__umulhi((uint32_t)j, st.magic);
            k = j - jd * Ns;
            const int ks = k * st.step * TS;
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = fft_cmul(v[r], tw[r * ks]);
        }
        fft_dft<R, SIGN>(v);
        float2 *o = B + (jd * Ns * R + k) * STR + c;
#pragma unroll
        for (int q = 0; q < R; q++) o[q * Ns * STR] = v[q];
    }
}

// The whole transform in buffer A (scratch B); returns the buffer that holds
// the result.
template <int SIGN, int L, int STR>
__device__ float2 *fft_run(float2 *A, float2 *B, const FftRadices &f,
                           const float2 *__restrict__ tw, int TS, int lt, int c) {
    for (int s = 0; s < f.count; s++) {
        const FftStage st = f.st[s];
        switch (st.r) {
        case 2: fft_stage<2, SIGN, L, STR>(A, B, f.n, st, tw, TS, lt, c); break;
        case 3: fft_stage<3, SIGN, L, STR>(A, B, f.n, st, tw, TS, lt, c); break;
        case 4: fft_stage<4, SIGN, L, STR>(A, B, f.n, st, tw, TS, lt, c); break;
        case 5: fft_stage<5, SIGN, L, STR>(A, B, f.n, st, tw, TS, lt, c); break;
        case 7: fft_stage<7, SIGN, L, STR>(A, B, f.n, st, tw, TS, lt, c); break;
        case 8: fft_stage<8, SIGN, L, STR>(A, B, f.n, st, tw, TS, lt, c); break;
        case 9: fft_stage<9, SIGN, L, STR>(A, B, f.n, st, tw, TS, lt, c); break;
        case 11: fft_stage<11, SIGN, L, STR>(A, B, f.n, st, tw, TS, lt, c); break;
        default: fft_stage<13, SIGN, L, STR>(A, B, f.n, st, tw, TS, lt, c); break;
        }
        __syncthreads();
        float2 *t = A;
        A = B;
        B = t;
    }
    return A;
}

// Where a row pass reads and writes.  IN: 0 complex rows of `in` (in place
// is allowed); 1 complex64 grid, 2 float32 grid, both read through the
// ifftshift: A[ku][kv] = grid[(ku + n_u/2) mod n_u][(kv + N/2) mod N], i.e.
// grid value w of the row goes to A[(w - N/2) mod N]; 3 the FP64 model image
// [planes, n_l, n_m] padded to the rows_per_plane x N plane and ifftshift-ed,
// divided by the correction image:
// A[ku][kv] = X[l][m] / (corr[l] corr[n_l + m]) with l = (ku + n_l/2) mod
// rows_per_plane, m = (kv + n_m/2) mod N, zero outside the model.  OUT: 0 the
// row to the same row of `out`; 1 through the fftshift, row r of a plane to
// row (r + rows_per_plane/2) mod rows_per_plane and value v to (v + N/2) mod N.
struct FftRowArgs {
    const void *in;
    float2 *out;
    int64_t n_rows, rows_per_plane;
    int64_t n_l, n_m;        // IN 3
    const double *corr;      // IN 3: [n_l + n_m] or NULL
};

// Each thread issues its PER loads of a row (contiguous from the row's
// uniform base, but for the model's wrap) before it writes any of them to LDS.
template <int IN, int OUT, int SIGN, int PER>
__global__ void __launch_bounds__(kFftBlock) __attribute__((amdgpu_waves_per_eu(PER <= 16 ? 4 : 2)))
fft_rows_kernel(FftRowArgs g, FftRadices f, const float2 *__restrict__ tw) {
    static_assert(IN != 3 || OUT == 0, "model rows are written unshifted");
    extern __shared__ float2 s_fft[];
    const int N = f.n;
    const int half = (IN == 1 || IN == 2) ? N / 2 : 0;
    const int64_t rpp = g.rows_per_plane;
    for (int64_t row = blockIdx.x; row < g.n_rows; row += gridDim.x) {
        float2 *A = s_fft, *B = s_fft + N;
        const int64_t pl = row / rpp, ku = row - pl * rpp;
        float2 x[PER];
        if (IN == 3) {
            int64_t l = ku + g.n_l / 2;
            if (l >= rpp) l -= rpp;
            const bool in_model = l < g.n_l;
            if (!in_model) {   // a padding row: its transform is zero
                float2 *dst = g.out + row * N;
                for (int w = threadIdx.x; w < N; w += kFftBlock) dst[w] = make_float2(0.f, 0.f);
                continue;
            }
            const double *src = reinterpret_cast<const double *>(g.in) + (pl * g.n_l + l) * g.n_m;
            const int hm = (int)(g.n_m / 2);
            // all loads first (model values and correction factors), then the
            // divisions; branch-free, indices outside the model read element 0
            // and are masked afterwards (guarded loads with their divisions
            // had waited for one another)
            double v[PER], c[PER];
            const double *cm = g.corr ? g.corr + g.n_l : src;
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int a = threadIdx.x + i * kFftBlock;
                int m = a + hm;
                if (m >= N) m -= N;
                const int mc = (a < N && m < g.n_m) ? m : 0;
                v[i] = src[mc];
                c[i] = cm[mc];
            }
            const double cl = g.corr ? g.corr[l] : 1.0;
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int a = threadIdx.x + i * kFftBlock;
                int m = a + hm;
                if (m >= N) m -= N;
                double y = v[i];
                if (g.corr) y = y / (cl * c[i]);
                x[i] = make_float2((a < N && m < g.n_m) ? (float)y : 0.f, 0.f);
            }
        } else {
            int64_t src_row = row;
            if (IN != 0) {
                int64_t su = ku + rpp / 2;
                if (su >= rpp) su -= rpp;
                src_row = pl * rpp + su;
            }
#pragma unroll
            for (int i = 0; i < PER; i++) {
                const int w = threadIdx.x + i * kFftBlock;
                if (w < N) {
                    if (IN == 2)
                        x[i] = make_float2((reinterpret_cast<const float *>(g.in) + src_row * N)[w],
                                           0.f);
                    else
                        x[i] = (reinterpret_cast<const float2 *>(g.in) + src_row * N)[w];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int w = threadIdx.x + i * kFftBlock;
            int a = w - half;
            if (a < 0) a += N;
            if (w < N) A[a] = x[i];
        }
        __syncthreads();
        const float2 *res = fft_run<SIGN, kFftBlock, 1>(A, B, f, tw, 1, threadIdx.x, 0);
        int64_t dst_row = row;
        if (OUT == 1) {
            int64_t r = ku + rpp / 2;
            if (r >= rpp) r -= rpp;
            dst_row = pl * rpp + r;
        }
        float2 *dst = g.out + dst_row * N;
        const int oh = OUT == 1 ? N / 2 : 0;   // dst[w] = res[(w - N/2) mod N]
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int w = threadIdx.x + i * kFftBlock;
            int v = w - oh;
            if (v < 0) v += N;
            if (w < N) dst[w] = res[v];
        }
        __syncthreads();  // the row's LDS is free for the next one
    }
}

// out[pl][c][r] = in[pl][r][c], [rows x cols] planes, 64 x 64 tiles: each
// thread issues its 16 loads of a tile before it writes any of them to LDS.
// Workgroups go round-robin to the 8 XCDs, so the tile index is remapped to
// give each XCD a contiguous run of tiles: neighbouring tiles share the
// cache lines at their edges (rows are not 128-byte aligned for odd lengths),
// which are then fetched and written back through one L2.
constexpr int kTrTile = 64;

__device__ inline void xcd_tile(unsigned &tx, unsigned &ty) {
    const unsigned nx = gridDim.x, total = gridDim.x * gridDim.y;
    const unsigned id = blockIdx.y * nx + blockIdx.x;
    const unsigned per = total / 8;
    const unsigned t = id < per * 8 ? (id % 8) * per + id / 8 : id;
    tx = t % nx;
    ty = t / nx;
}

__global__ void __launch_bounds__(256)
fft_transpose_kernel(const float2 *__restrict__ in, float2 *__restrict__ out, int64_t n_planes,
                     int64_t rows, int64_t cols) {
    __shared__ float2 tile[kTrTile][kTrTile + 1];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // 64 x 4
    unsigned bx, by;
    xcd_tile(bx, by);
    const int64_t c0 = (int64_t)bx * kTrTile, r0 = (int64_t)by * kTrTile;
    constexpr int K = kTrTile / 4;
    for (int64_t pl = blockIdx.z; pl < n_planes; pl += gridDim.z) {
        const float2 *ip = in + pl * rows * cols;
        float2 *op = out + pl * rows * cols;
        float2 x[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
            const int64_t r = r0 + ty + 4 * i, c = c0 + tx;
            if (r < rows && c < cols) x[i] = ip[r * cols + c];
        }
#pragma unroll
        for (int i = 0; i < K; i++) tile[ty + 4 * i][tx] = x[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < K; i++) {
            const int64_t c = c0 + ty + 4 * i, r = r0 + tx;
            if (c < cols && r < rows) op[c * rows + r] = tile[tx][ty + 4 * i];
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
    unsigned bx, by;
    xcd_tile(bx, by);
    const int64_t l0 = (int64_t)bx * 32, m0 = (int64_t)by * 32;
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

// ---------------------------------------------------------------------------
// Column transforms without a transpose (four-step): the length-N transform
// along u of the [N][n_cols] planes, N = N1 N2, input row n = r1 N2 + r2,
// output row k = k1 + N1 k2:
//   X[k1 + N1 k2] = sum_r2 w_N^(r2 k1) w_N2^(r2 k2) sum_r1 x[r1 N2 + r2] w_N1^(r1 k1)
// Step A, per (plane, r2, 32-column block): the N1 rows r1 N2 + r2 through LDS,
// their length-N1 transforms, the twiddle w_N^(r2 k1), written back in place
// to rows k1 N2 + r2.  Step B, per (plane, k1, block): the N2 consecutive rows
// k1 N2 + r2, their length-N2 transforms, and output k2 written as row
// k1 + N1 k2 of the result: cropped, corrected and real into the image
// (inverse), or through the fftshift into the grid (forward).  Each row
// segment a workgroup reads or writes is 32 values (256 bytes) wide; 8
// threads share a column.  The length-N1 and length-N2 twiddles are the
// length-N table read at strides N2 and N1.
// ---------------------------------------------------------------------------
constexpr int kColW = 32;
constexpr int kColL = kFftBlock / kColW;   // threads per column

__device__ inline int64_t xcd_block() {
    const int64_t total = gridDim.x, id = blockIdx.x;
    const int64_t per = total / 8;
    return id < per * 8 ? (id % 8) * per + id / 8 : id;
}

struct ColArgs {
    float2 *buf;             // [n_planes][N][n_cols], after the row pass
    int64_t n_planes;
    int N, n_cols, N1, N2;
    // step B outputs
    double *image;           // inverse: [n_planes][n_l][n_m]
    int64_t n_l, n_m;
    const double *sum_weight, *corr;
    float2 *grid;            // forward: [n_planes][N][n_cols]
};

template <int SIGN, int PER>
__global__ void __launch_bounds__(kFftBlock)
fft_cols_a_kernel(ColArgs a, FftRadices f1, const float2 *__restrict__ tw) {
    extern __shared__ float2 s_fft[];
    const int N1 = a.N1, N2 = a.N2, nc = a.n_cols;
    const int n_cg = (nc + kColW - 1) / kColW;
    const int64_t b = xcd_block();
    const int cg = (int)(b % n_cg);
    const int64_t rest = b / n_cg;
    const int r2 = (int)(rest % N2);
    const int64_t pl = rest / N2;
    const int c = threadIdx.x % kColW, lt = threadIdx.x / kColW;
    const int col = cg * kColW + c;
    const bool cok = col < nc;
    float2 *base = a.buf + (pl * a.N + r2) * nc + col;   // row r1 N2 + r2 at r1 N2 nc
    const int64_t rs = (int64_t)N2 * nc;
    float2 *A = s_fft, *B = s_fft + N1 * kColW;
    float2 x[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int r1 = lt + i * kColL;
        if (r1 < N1 && cok) x[i] = base[r1 * rs];
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int r1 = lt + i * kColL;
        if (r1 < N1) A[r1 * kColW + c] = cok ? x[i] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const float2 *res = fft_run<SIGN, kColL, kColW>(A, B, f1, tw, N2, lt, c);
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int k1 = lt + i * kColL;
        if (k1 < N1 && cok) base[k1 * rs] = fft_cmul(res[k1 * kColW + c], tw[r2 * k1]);
    }
}

// MODE 0: image[pl][l][m] = (Re X[ku][kv] / sw') / (corr[l] corr[n_l + m]) for
// l = (ku + n_l/2) mod N < n_l, m = (kv + n_m/2) mod n_cols < n_m, as
// crop_correct_kernel computes it.  MODE 1: grid[pl][u][v] = X[ku][kv] for
// u = (ku + N/2) mod N, v = (kv + n_cols/2) mod n_cols (the fftshift).
template <int SIGN, int PER, int MODE>
__global__ void __launch_bounds__(kFftBlock)
fft_cols_b_kernel(ColArgs a, FftRadices f2, const float2 *__restrict__ tw) {
#pragma clang fp contract(off)
    extern __shared__ float2 s_fft[];
    const int N1 = a.N1, N2 = a.N2, nc = a.n_cols, N = a.N;
    const int n_cg = (nc + kColW - 1) / kColW;
    const int64_t b = xcd_block();
    const int cg = (int)(b % n_cg);
    const int64_t rest = b / n_cg;
    const int k1 = (int)(rest % N1);
    const int64_t pl = rest / N1;
    const int c = threadIdx.x % kColW, lt = threadIdx.x / kColW;
    const int col = cg * kColW + c;
    const bool cok = col < nc;
    const float2 *base = a.buf + (pl * N + (int64_t)k1 * N2) * nc + col;
    float2 *A = s_fft, *B = s_fft + N2 * kColW;
    float2 x[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int r2 = lt + i * kColL;
        if (r2 < N2 && cok) x[i] = base[(int64_t)r2 * nc];
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int r2 = lt + i * kColL;
        if (r2 < N2) A[r2 * kColW + c] = cok ? x[i] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const float2 *res = fft_run<SIGN, kColL, kColW>(A, B, f2, tw, N1, lt, c);
    if (MODE == 0) {
        int m = col + (int)(a.n_m / 2);
        if (m >= nc) m -= nc;
        const bool mok = cok && m < a.n_m;
        double sw = 1.0;
        if (a.sum_weight) {
            sw = a.sum_weight[pl];
            if (sw == 0.0) sw = 1.0;
        }
        const double cm = (a.corr && mok) ? a.corr[a.n_l + m] : 1.0;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int k2 = lt + i * kColL;
            if (k2 >= N2) continue;
            int l = k1 + N1 * k2 + (int)(a.n_l / 2);
            if (l >= N) l -= N;
            if (mok && l < a.n_l) {
                double v = (double)res[k2 * kColW + c].x;
                if (a.sum_weight) v = v / sw;
                if (a.corr) v = v / (a.corr[l] * cm);
                a.image[(pl * a.n_l + l) * a.n_m + m] = v;
            }
        }
    } else {
        int v = col + nc / 2;
        if (v >= nc) v -= nc;
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const int k2 = lt + i * kColL;
            if (k2 >= N2) continue;
            int u = k1 + N1 * k2 + N / 2;
            if (u >= N) u -= N;
            if (cok) a.grid[(pl * N + u) * nc + v] = res[k2 * kColW + c];
        }
    }
}

// N = N1 N2 with N1 <= N2 <= 128 and N2 as small as possible, or N1 = 1 for
// N <= 128 (step A is then skipped); false when no such split exists.
static bool col_split(int64_t N, int &N1, int &N2) {
    if (N <= 128) {
        N1 = 1;
        N2 = (int)N;
        return true;
    }
    for (int n2 = 2; n2 <= 128; n2++) {
        if (N % n2 == 0 && N / n2 <= n2) {
            N1 = (int)(N / n2);
            N2 = n2;
            return true;
        }
    }
    return false;
}

template <int SIGN, int PER>
static int cols_a_launch(hipStream_t s, const ColArgs &a, const FftRadices &f1,
                         const float2 *tw, int64_t blocks) {
    const size_t lds = 2 * (size_t)a.N1 * kColW * sizeof(float2);
    auto kern = fft_cols_a_kernel<SIGN, PER>;
    NGC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    kern<<<(unsigned)blocks, kFftBlock, lds, s>>>(a, f1, tw);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

template <int SIGN, int PER, int MODE>
static int cols_b_launch(hipStream_t s, const ColArgs &a, const FftRadices &f2,
                         const float2 *tw, int64_t blocks) {
    const size_t lds = 2 * (size_t)a.N2 * kColW * sizeof(float2);
    auto kern = fft_cols_b_kernel<SIGN, PER, MODE>;
    NGC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    kern<<<(unsigned)blocks, kFftBlock, lds, s>>>(a, f2, tw);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

// PER = rows per thread, ceil(N1 or N2 / 8) rounded up to 2, 4, 8, 16
template <int SIGN>
static int cols_a(hipStream_t s, const ColArgs &a, const FftRadices &f1, const float2 *tw,
                  int64_t blocks) {
    if (a.N1 <= 2 * kColL) return cols_a_launch<SIGN, 2>(s, a, f1, tw, blocks);
    if (a.N1 <= 4 * kColL) return cols_a_launch<SIGN, 4>(s, a, f1, tw, blocks);
    if (a.N1 <= 8 * kColL) return cols_a_launch<SIGN, 8>(s, a, f1, tw, blocks);
    return cols_a_launch<SIGN, 16>(s, a, f1, tw, blocks);
}

template <int SIGN, int MODE>
static int cols_b(hipStream_t s, const ColArgs &a, const FftRadices &f2, const float2 *tw,
                  int64_t blocks) {
    if (a.N2 <= 2 * kColL) return cols_b_launch<SIGN, 2, MODE>(s, a, f2, tw, blocks);
    if (a.N2 <= 4 * kColL) return cols_b_launch<SIGN, 4, MODE>(s, a, f2, tw, blocks);
    if (a.N2 <= 8 * kColL) return cols_b_launch<SIGN, 8, MODE>(s, a, f2, tw, blocks);
    return cols_b_launch<SIGN, 16, MODE>(s, a, f2, tw, blocks);
}

// Steps A and B over the row-pass output in `buf` (phases "fft_cols_a",
// "fft_cols_b"); false in *done when the column length has no split or the
// block counts exceed one launch (the caller then transposes).
template <int SIGN, int MODE>
static int cols_four_step(ngc_context *ctx, hipStream_t s, ColArgs a, const float2 *tw,
                          bool *done) {
    *done = false;
    int N1, N2;
    if (!col_split(a.N, N1, N2)) return NGC_OK;
    a.N1 = N1;
    a.N2 = N2;
    const int64_t n_cg = ceil_div((int64_t)a.n_cols, (int64_t)kColW);
    const int64_t blocks_a = a.n_planes * N2 * n_cg, blocks_b = a.n_planes * N1 * n_cg;
    if (blocks_a >= (int64_t(1) << 31) || blocks_b >= (int64_t(1) << 31)) return NGC_OK;
    FftRadices f1, f2;
    NGC_CHECK(fft_factor(N2, f2) && (N1 == 1 || fft_factor(N1, f1)), NGC_ERR_UNSUPPORTED,
              "column split not factorable");
    if (N1 > 1) {
        phase_begin(ctx, s, "fft_cols_a");
        NGC_TRY((cols_a<SIGN>(s, a, f1, tw, blocks_a)));
        phase_end(ctx, s);
    }
    phase_begin(ctx, s, "fft_cols_b");
    NGC_TRY((cols_b<SIGN, MODE>(s, a, f2, tw, blocks_b)));
    phase_end(ctx, s);
    *done = true;
    return NGC_OK;
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

template <int IN, int OUT, int SIGN, int PER>
static int fft_rows_launch(hipStream_t s, const FftRowArgs &g, const FftRadices &f,
                           const float2 *tw) {
    const size_t lds = 2 * (size_t)f.n * sizeof(float2);
    auto kern = fft_rows_kernel<IN, OUT, SIGN, PER>;
    NGC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    const int64_t blocks = std::min<int64_t>(g.n_rows, int64_t(1) << 20);
    kern<<<(unsigned)blocks, kFftBlock, lds, s>>>(g, f, tw);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

// PER = loads per thread of one row, ceil(N / 256) rounded up to 4, 8, 12, 16, 32
template <int IN, int OUT, int SIGN>
static int fft_rows(hipStream_t s, const FftRowArgs &g, const FftRadices &f, const float2 *tw) {
    if (f.n <= 4 * kFftBlock) return fft_rows_launch<IN, OUT, SIGN, 4>(s, g, f, tw);
    if (f.n <= 8 * kFftBlock) return fft_rows_launch<IN, OUT, SIGN, 8>(s, g, f, tw);
    if (f.n <= 12 * kFftBlock) return fft_rows_launch<IN, OUT, SIGN, 12>(s, g, f, tw);
    if (f.n <= 16 * kFftBlock) return fft_rows_launch<IN, OUT, SIGN, 16>(s, g, f, tw);
    return fft_rows_launch<IN, OUT, SIGN, 32>(s, g, f, tw);
}

static FftRowArgs row_args(const void *in, void *out, int64_t n_rows, int64_t rows_per_plane) {
    FftRowArgs g{};
    g.in = in;
    g.out = (float2 *)out;
    g.n_rows = n_rows;
    g.rows_per_plane = rows_per_plane;
    return g;
}

static int fft_transpose(hipStream_t s, const void *in, void *out, int64_t n_planes,
                         int64_t rows, int64_t cols) {
    const dim3 tg((unsigned)ceil_div(cols, kTrTile), (unsigned)ceil_div(rows, kTrTile),
                  (unsigned)std::min<int64_t>(n_planes, 65535));
    fft_transpose_kernel<<<tg, 256, 0, s>>>((const float2 *)in, (float2 *)out, n_planes, rows,
                                            cols);
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
    const float2 *tw_v, *tw_u;
    NGC_TRY(fft_twiddles(ctx, s, n_v, +1, "fft_tw_rows", &tw_v));
    NGC_TRY(fft_twiddles(ctx, s, n_u, +1, "fft_tw_cols", &tw_u));
    phase_begin(ctx, s, "fft_rows");
    const FftRowArgs g1 = row_args(grid, b0, n_planes * n_u, n_u);
    if (dt == NGC_C64)
        NGC_TRY((fft_rows<1, 0, +1>(s, g1, fv, tw_v)));
    else
        NGC_TRY((fft_rows<2, 0, +1>(s, g1, fv, tw_v)));
    phase_end(ctx, s);
    // columns in place, the crop and correction fused into their last step
    ColArgs ca{};
    ca.buf = (float2 *)b0;
    ca.n_planes = n_planes;
    ca.N = (int)n_u;
    ca.n_cols = (int)n_v;
    ca.image = image;
    ca.n_l = n_l;
    ca.n_m = n_m;
    ca.sum_weight = sum_weight;
    ca.corr = d_corr;
    bool done = false;
    NGC_TRY((cols_four_step<+1, 0>(ctx, s, ca, tw_u, &done)));
    if (done) return NGC_OK;
    // otherwise transpose, rows again, crop
    NGC_TRY(workspace(ctx, "fft_transposed", cells * sizeof(float2), &b1));
    phase_begin(ctx, s, "fft_transpose");
    NGC_TRY(fft_transpose(s, b0, b1, n_planes, n_u, n_v));
    phase_end(ctx, s);
    phase_begin(ctx, s, "fft_cols");
    NGC_TRY((fft_rows<0, 0, +1>(s, row_args(b1, b1, n_planes * n_v, n_v), fu, tw_u)));
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

// The forward placement of the model (the inverse of lds_grid_to_image):
// grid = fftshift(fft2(ifftshift(padded model / correction))) for complex64
// grids: rows of the padded model (IN 3), then the four-step columns with the
// fftshift in their last step; for column lengths without a split, transpose,
// rows through the fftshift (OUT 1) and transpose into the grid.
int lds_model_to_grid(ngc_context *ctx, hipStream_t s, const double *image, int64_t n_planes,
                      int64_t n_l, int64_t n_m, int64_t n_u, int64_t n_v, const double *d_corr,
                      void *grid) {
    FftRadices fu, fv;
    NGC_CHECK(fft_factor(n_u, fu) && fft_factor(n_v, fv), NGC_ERR_UNSUPPORTED,
              "axis length not a product of 2, 3, 5, 7, 11, 13 up to 8192");
    const size_t cells = (size_t)n_planes * n_u * n_v;
    void *b0, *b1;
    NGC_TRY(workspace(ctx, "fft_buffer", cells * sizeof(float2), &b0));
    const float2 *tw_v, *tw_u;
    NGC_TRY(fft_twiddles(ctx, s, n_v, -1, "fft_tw_rows_forward", &tw_v));
    NGC_TRY(fft_twiddles(ctx, s, n_u, -1, "fft_tw_cols_forward", &tw_u));
    phase_begin(ctx, s, "fft_rows");
    FftRowArgs g1 = row_args(image, b0, n_planes * n_u, n_u);
    g1.n_l = n_l;
    g1.n_m = n_m;
    g1.corr = d_corr;
    NGC_TRY((fft_rows<3, 0, -1>(s, g1, fv, tw_v)));
    phase_end(ctx, s);
    // columns in place, the fftshift fused into their last step
    ColArgs ca{};
    ca.buf = (float2 *)b0;
    ca.n_planes = n_planes;
    ca.N = (int)n_u;
    ca.n_cols = (int)n_v;
    ca.grid = (float2 *)grid;
    bool done = false;
    NGC_TRY((cols_four_step<-1, 1>(ctx, s, ca, tw_u, &done)));
    if (done) return NGC_OK;
    // otherwise transpose, shifted rows, transpose into the grid
    NGC_TRY(workspace(ctx, "fft_transposed", cells * sizeof(float2), &b1));
    phase_begin(ctx, s, "fft_transpose");
    NGC_TRY(fft_transpose(s, b0, b1, n_planes, n_u, n_v));
    phase_end(ctx, s);
    phase_begin(ctx, s, "fft_cols");
    NGC_TRY((fft_rows<0, 1, -1>(s, row_args(b1, b0, n_planes * n_v, n_v), fu, tw_u)));
    phase_end(ctx, s);
    phase_begin(ctx, s, "fft_transpose_out");
    NGC_TRY(fft_transpose(s, b0, grid, n_planes, n_v, n_u));
    phase_end(ctx, s);
    return NGC_OK;
}

}  // namespace ngc
