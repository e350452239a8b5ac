// Dense layers of the learned A-term kernel surrogate, on gfx950 matrix cores.
//
// No counterpart in the reference: the surrogate replaces per-channel
// evaluations of the exact FFT-of-aperture convolution function
// (make_gridding_convolution_function.py:161-280).  The network is a plain
// MLP; every product runs as bf16 x bf16 -> FP32 MFMA (v_mfma_f32_16x16x16_bf16):
// operands are held in FP32 in memory (the optimiser's master copy) and rounded
// to bf16 (round to nearest even) as they are loaded, so the weights the
// network computes with are bf16 and every accumulation is FP32.
//
//   ngc_bf16_gemm     C = act(op(A) op(B) + bias), strided A and B (so X^T dY and
//                     dY W^T need no transposed copies), optional pre-activation
//                     output, optional accumulate into C; act 2 multiplies by
//                     the tanh derivative 1 - aux^2 (the backward step fused)
//   ngc_tanh_backward dZ = dA * (1 - A^2)
//   ngc_col_sum       column sums (bias gradients)
//   ngc_mse_grad      d/dy of mean((y - t)^2) and the loss
//   ngc_adam_step     Adam update of FP32 master weights
#include <algorithm>

#include "ngc_internal.hpp"

namespace ngc {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (NaN kept quiet)
__device__ inline short to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (short)0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

struct GemmArgs {
    const float *A, *B, *bias;
    float *C, *Z;
    const float *aux;  // act 2: activations whose tanh derivative scales the result
    int64_t M, N, K;
    int64_t sam, sak, sbk, sbn, ldc;
    int act, accumulate;
    int64_t k_chunk;   // K range of one split (a multiple of kGemmBK)
    float *partial;    // [splits][M][N] FP32 partial products when split
};

constexpr int kGemmBM = 64, kGemmBN = 64, kGemmBK = 32;
constexpr int kGemmLd = kGemmBK + 8;   // LDS row pitch in bf16 (80 B, 8 B aligned)

__device__ inline void gemm_store(const GemmArgs &g, int64_t row, int64_t col, float z) {
    if (g.bias) z += g.bias[col];
    if (g.Z) g.Z[row * g.ldc + col] = z;
    float y = z;
    if (g.act == 1) {
        y = tanhf(z);
    } else if (g.act == 2) {
        const float a = g.aux[row * g.ldc + col];
        y = z * (1.f - a * a);
    }
    float *c = &g.C[row * g.ldc + col];
    *c = g.accumulate ? *c + y : y;
}

// One workgroup computes a 64 x 64 block of C over the K range of its split
// (blockIdx.z).  Per 32-wide K step the four wavefronts load the A (64 x 32)
// and B (32 x 64) tiles with the contiguous dimension of each operand across
// consecutive lanes (so X^T and W^T operands are read coalesced too), round to
// bf16 and store both k-contiguous in LDS; each wavefront then runs a 32 x 32
// block as 2 x 2 tiles of v_mfma_f32_16x16x16_bf16 over two K sub-steps.
// Operand layout of the instruction (lane l, element j = 0..3):
//   A[l % 16][4 (l / 16) + j],  B[4 (l / 16) + j][l % 16],
//   D: row 4 (l / 16) + j, column l % 16.
// Unsplit launches apply bias, activation and accumulate directly; split
// launches write FP32 partial blocks that gemm_reduce_kernel sums in split
// order.
__global__ void __launch_bounds__(256) bf16_gemm_kernel(GemmArgs g) {
    __shared__ __align__(16) short As[kGemmBM * kGemmLd];
    __shared__ __align__(16) short Bs[kGemmBN * kGemmLd];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t mt = (int64_t)blockIdx.y * kGemmBM, nt = (int64_t)blockIdx.x * kGemmBN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int64_t kb = (int64_t)blockIdx.z * g.k_chunk;
    const int64_t ke = g.K < kb + g.k_chunk ? g.K : kb + g.k_chunk;
    const bool a_mc = g.sam == 1 && g.sak != 1;   // A contiguous along m
    const bool b_kc = g.sbk == 1 && g.sbn != 1;   // B contiguous along k
    const int lr = lane & 15, lk = (lane >> 4) * 4;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int64_t k0 = kb; k0 < ke; k0 += kGemmBK) {
        if (a_mc) {
            const int kk = tid >> 3, mm = (tid & 7) * 8;
            const int64_t k = k0 + kk;
            for (int j = 0; j < 8; j++) {
                const int64_t m = mt + mm + j;
                As[(mm + j) * kGemmLd + kk] =
                    (m < g.M && k < ke) ? to_bf16(g.A[m * g.sam + k * g.sak]) : (short)0;
            }
        } else {
            const int mm = tid >> 2, kk = (tid & 3) * 8;
            const int64_t m = mt + mm;
            for (int j = 0; j < 8; j++) {
                const int64_t k = k0 + kk + j;
                As[mm * kGemmLd + kk + j] =
                    (m < g.M && k < ke) ? to_bf16(g.A[m * g.sam + k * g.sak]) : (short)0;
            }
        }
        if (b_kc) {
            const int nn = tid >> 2, kk = (tid & 3) * 8;
            const int64_t n = nt + nn;
            for (int j = 0; j < 8; j++) {
                const int64_t k = k0 + kk + j;
                Bs[nn * kGemmLd + kk + j] =
                    (n < g.N && k < ke) ? to_bf16(g.B[k * g.sbk + n * g.sbn]) : (short)0;
            }
        } else {
            const int kk = tid >> 3, nn = (tid & 7) * 8;
            const int64_t k = k0 + kk;
            for (int j = 0; j < 8; j++) {
                const int64_t n = nt + nn + j;
                Bs[(nn + j) * kGemmLd + kk] =
                    (n < g.N && k < ke) ? to_bf16(g.B[k * g.sbk + n * g.sbn]) : (short)0;
            }
        }
        __syncthreads();
        for (int ks = 0; ks < kGemmBK; ks += 16) {
            bf16x4 a[2], b[2];
            for (int mi = 0; mi < 2; mi++)
                a[mi] = *(const bf16x4 *)&As[(wm + 16 * mi + lr) * kGemmLd + ks + lk];
            for (int ni = 0; ni < 2; ni++)
                b[ni] = *(const bf16x4 *)&Bs[(wn + 16 * ni + lr) * kGemmLd + ks + lk];
            for (int mi = 0; mi < 2; mi++)
                for (int ni = 0; ni < 2; ni++)
                    acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(
                        a[mi], b[ni], acc[mi][ni], 0, 0, 0);
        }
        __syncthreads();
    }
    for (int mi = 0; mi < 2; mi++)
        for (int ni = 0; ni < 2; ni++)
            for (int j = 0; j < 4; j++) {
                const int64_t row = mt + wm + 16 * mi + lk + j;
                const int64_t col = nt + wn + 16 * ni + lr;
                if (row >= g.M || col >= g.N) continue;
                if (gridDim.z > 1)
                    g.partial[((int64_t)blockIdx.z * g.M + row) * g.N + col] = acc[mi][ni][j];
                else
                    gemm_store(g, row, col, acc[mi][ni][j]);
            }
}

// Sum of the split-K partial blocks in split order, then the epilogue.
__global__ void __launch_bounds__(256) gemm_reduce_kernel(GemmArgs g, int splits) {
    const int64_t mn = g.M * g.N;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < mn;
         i += (int64_t)gridDim.x * blockDim.x) {
        float z = 0.f;
        for (int s = 0; s < splits; s++) z += g.partial[s * mn + i];
        gemm_store(g, i / g.N, i % g.N, z);
    }
}

__global__ void tanh_backward_kernel(int64_t n, const float *dA, const float *A, float *dZ) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dZ[i] = dA[i] * (1.f - A[i] * A[i]);
}

// one workgroup per column block: out[c] = sum_r X[r][c]
__global__ void __launch_bounds__(256) col_sum_kernel(int64_t M, int64_t N, const float *X,
                                                      float *out) {
    __shared__ float s[256];
    const int64_t c = (int64_t)blockIdx.x;
    float acc = 0.f;
    for (int64_t r = threadIdx.x; r < M; r += blockDim.x) acc += X[r * N + c];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = s[0];
}

__global__ void __launch_bounds__(256) mse_grad_kernel(int64_t n, const float *y, const float *t,
                                                       float *g, double *loss) {
    __shared__ double s[256];
    double acc = 0.0;
    const float scale = 2.f / (float)n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float d = y[i] - t[i];
        g[i] = scale * d;
        acc += (double)d * (double)d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, s[0] / (double)n);
}

__global__ void adam_kernel(int64_t n, float *w, const float *g, float *m, float *v, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void zero_f64_kernel(double *p) { *p = 0.0; }

// Device-side training schedule, so that one training step can be captured in
// a graph and replayed: `counter` holds the number of steps taken in this run.
__device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// rows r_i = hash(seed, step, i) mod n of x [n, dx] and t [n, dt] into the
// minibatch xb [batch, dx], tb [batch, dt]
__global__ void sample_rows_kernel(const float *x, const float *t, int64_t n, int64_t dx,
                                   int64_t dt, int64_t batch, uint64_t seed,
                                   const int64_t *counter, float *xb, float *tb) {
    const uint64_t k = (uint64_t)*counter;
    const int64_t width = dx + dt;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < batch * width;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i / width, c = i - j * width;
        const int64_t r = (int64_t)(splitmix64(seed ^ splitmix64(k * (uint64_t)batch +
                                                                  (uint64_t)j)) %
                                    (uint64_t)n);
        if (c < dx)
            xb[j * dx + c] = x[r * dx + c];
        else
            tb[j * dt + c - dx] = t[r * dt + c - dx];
    }
}

// sched = {learning rate (cosine decay from lr to lr_min over `steps`),
// 1 - b1^s, 1 - b2^s} for Adam step s = step0 + k + 1, k = *counter; then ++k.
__global__ void adam_schedule_kernel(int64_t *counter, float *sched, double lr, double lr_min,
                                     int64_t steps, int64_t step0, double b1, double b2) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int64_t k = *counter;
        const double span = steps > 1 ? (double)(steps - 1) : 1.0;
        const double rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + cos(M_PI * (double)k / span));
        const double s = (double)(step0 + k + 1);
        sched[0] = (float)rate;
        sched[1] = (float)(1.0 - pow(b1, s));
        sched[2] = (float)(1.0 - pow(b2, s));
        *counter = k + 1;
    }
}

__global__ void adam_sched_kernel(int64_t n, float *w, const float *g, float *m, float *v,
                                  const float *sched, float b1, float b2, float eps) {
    const float lr = sched[0], c1 = sched[1], c2 = sched[2];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

static int blocks_of(int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 8192));
}

}  // namespace ngc

using namespace ngc;

extern "C" {

int ngc_bf16_gemm(ngc_context *ctx, void *stream, int64_t M, int64_t N, int64_t K,
                  const float *A, int64_t sam, int64_t sak, const float *B, int64_t sbk,
                  int64_t sbn, const float *bias, int act, int accumulate, float *C, int64_t ldc,
                  float *Z, const float *aux) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(M >= 0 && N >= 0 && K >= 0 && ldc >= N, NGC_ERR_ARG, "bad GEMM dimensions");
    NGC_CHECK(act == 0 || act == 1 || act == 2, NGC_ERR_ARG,
              "act must be 0 (none), 1 (tanh) or 2 (times 1 - aux^2)");
    NGC_CHECK(act != 2 || aux, NGC_ERR_ARG, "act 2 needs aux");
    if (M == 0 || N == 0) return NGC_OK;
    NGC_CHECK(A && B && C, NGC_ERR_ARG, "NULL GEMM operand");
    NGC_CHECK(ceil_div(M, kGemmBM) < 65536, NGC_ERR_UNSUPPORTED, "M too large");
    NGC_HIP(hipSetDevice(ctx->device));
    NGC_CHECK(ceil_div(N, kGemmBN) < (int64_t(1) << 31), NGC_ERR_UNSUPPORTED, "N too large");
    // split K when the output blocks alone cannot fill the device, keeping at
    // least 4 K steps per split
    const int64_t tiles = ceil_div(M, kGemmBM) * ceil_div(N, kGemmBN);
    const int64_t ksteps = std::max<int64_t>(1, ceil_div(K, kGemmBK));
    int64_t splits = 1;
    if (tiles < ctx->n_cu && ksteps >= 8)
        splits = std::min<int64_t>(ksteps / 4, ceil_div(2 * (int64_t)ctx->n_cu, tiles));
    const int64_t k_chunk = ceil_div(ksteps, splits) * kGemmBK;
    splits = std::max<int64_t>(1, ceil_div(K, k_chunk));
    GemmArgs g{A, B, bias, C, Z, aux, M, N, K, sam, sak, sbk, sbn, ldc, act, accumulate,
               K > 0 ? k_chunk : kGemmBK, nullptr};
    hipStream_t s = (hipStream_t)stream;
    if (splits > 1) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        NGC_TRY(enter_stream(ctx, s));
        void *part = nullptr;
        NGC_TRY(workspace(ctx, "gemm_partial", (size_t)(splits * M * N) * sizeof(float), &part));
        g.partial = (float *)part;
        dim3 grid((unsigned)ceil_div(N, kGemmBN), (unsigned)ceil_div(M, kGemmBM),
                  (unsigned)splits);
        bf16_gemm_kernel<<<grid, 256, 0, s>>>(g);
        NGC_HIP(hipGetLastError());
        gemm_reduce_kernel<<<blocks_of(M * N), 256, 0, s>>>(g, (int)splits);
        NGC_HIP(hipGetLastError());
        return NGC_OK;
    }
    dim3 grid((unsigned)ceil_div(N, kGemmBN), (unsigned)ceil_div(M, kGemmBM), 1);
    bf16_gemm_kernel<<<grid, 256, 0, s>>>(g);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

int ngc_tanh_backward(ngc_context *ctx, void *stream, int64_t n, const float *dA, const float *A,
                      float *dZ) {
    NGC_CHECK(ctx && (n == 0 || (dA && A && dZ)), NGC_ERR_ARG, "bad arguments");
    if (n == 0) return NGC_OK;
    NGC_HIP(hipSetDevice(ctx->device));
    tanh_backward_kernel<<<blocks_of(n), 256, 0, (hipStream_t)stream>>>(n, dA, A, dZ);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

int ngc_col_sum(ngc_context *ctx, void *stream, int64_t M, int64_t N, const float *X,
                float *out) {
    NGC_CHECK(ctx && X && out && M >= 0 && N >= 0, NGC_ERR_ARG, "bad arguments");
    NGC_CHECK(N < (int64_t(1) << 31), NGC_ERR_UNSUPPORTED, "too many columns");
    if (N == 0) return NGC_OK;
    NGC_HIP(hipSetDevice(ctx->device));
    col_sum_kernel<<<(unsigned)N, 256, 0, (hipStream_t)stream>>>(M, N, X, out);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

int ngc_mse_grad(ngc_context *ctx, void *stream, int64_t n, const float *y, const float *t,
                 float *g, double *loss) {
    NGC_CHECK(ctx && y && t && g && loss && n > 0, NGC_ERR_ARG, "bad arguments");
    NGC_HIP(hipSetDevice(ctx->device));
    zero_f64_kernel<<<1, 1, 0, (hipStream_t)stream>>>(loss);
    mse_grad_kernel<<<std::min(blocks_of(n), 1024), 256, 0, (hipStream_t)stream>>>(n, y, t, g,
                                                                                   loss);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

int ngc_adam_step(ngc_context *ctx, void *stream, int64_t n, float *w, const float *g, float *m,
                  float *v, double lr, double beta1, double beta2, double eps, int64_t step) {
    NGC_CHECK(ctx && w && g && m && v && n >= 0 && step >= 1, NGC_ERR_ARG, "bad arguments");
    if (n == 0) return NGC_OK;
    NGC_HIP(hipSetDevice(ctx->device));
    const float c1 = (float)(1.0 - std::pow(beta1, (double)step));
    const float c2 = (float)(1.0 - std::pow(beta2, (double)step));
    adam_kernel<<<blocks_of(n), 256, 0, (hipStream_t)stream>>>(
        n, w, g, m, v, (float)lr, (float)beta1, (float)beta2, (float)eps, c1, c2);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

int ngc_mlp_sample_rows(ngc_context *ctx, void *stream, const float *x, const float *t,
                        int64_t n, int64_t dx, int64_t dt, int64_t batch, uint64_t seed,
                        const int64_t *counter, float *xb, float *tb) {
    NGC_CHECK(ctx && x && t && counter && xb && tb, NGC_ERR_ARG, "bad arguments");
    NGC_CHECK(n >= 1 && dx >= 1 && dt >= 1 && batch >= 1, NGC_ERR_ARG, "bad dimensions");
    NGC_HIP(hipSetDevice(ctx->device));
    sample_rows_kernel<<<blocks_of(batch * (dx + dt)), 256, 0, (hipStream_t)stream>>>(
        x, t, n, dx, dt, batch, seed, counter, xb, tb);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

int ngc_adam_schedule(ngc_context *ctx, void *stream, int64_t *counter, float *sched,
                      double lr, double lr_min, int64_t steps, int64_t step0, double beta1,
                      double beta2) {
    NGC_CHECK(ctx && counter && sched && steps >= 1 && step0 >= 0, NGC_ERR_ARG,
              "bad arguments");
    NGC_HIP(hipSetDevice(ctx->device));
    adam_schedule_kernel<<<1, 64, 0, (hipStream_t)stream>>>(counter, sched, lr, lr_min, steps,
                                                            step0, beta1, beta2);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

int ngc_adam_step_sched(ngc_context *ctx, void *stream, int64_t n, float *w, const float *g,
                        float *m, float *v, const float *sched, double beta1, double beta2,
                        double eps) {
    NGC_CHECK(ctx && w && g && m && v && sched && n >= 0, NGC_ERR_ARG, "bad arguments");
    if (n == 0) return NGC_OK;
    NGC_HIP(hipSetDevice(ctx->device));
    adam_sched_kernel<<<blocks_of(n), 256, 0, (hipStream_t)stream>>>(
        n, w, g, m, v, sched, (float)beta1, (float)beta2, (float)eps);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

}  // extern "C"
