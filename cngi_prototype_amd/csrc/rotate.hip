// Phase-centre rotation of uvw and visibilities for gfx950: replaces
// apply_rotation_matrix and apply_phasor of direction_rotate
// (ngcasa/imaging/direction_rotate.py:213-268) behind ngc_direction_rotate().
//
// One pre-pass workgroup per integration finds that integration's field (the
// reference requires one valid field id per integration, :221-224 and
// :248-251); then one thread per (row, channel) rotates the row's uvw with
// the field's 3 x 3 matrix, forms the phase (the rotated u, v (and w without
// common-tangent reprojection) dotted with the field's phase-rotation
// vector) and multiplies every pol by exp(2 pi i phase f / c), all in FP64.
// With single_precision the result is stored as complex64 (the reference
// rounds through complex64).  Streaming: 8-16 B of visibility in and out per
// sample, 24 B of uvw per row.
#include "ngc_internal.hpp"

namespace ngc {

constexpr int32_t kIntNan = INT32_MIN;

struct RotArgs {
    const double *uvw;
    const void *vis;
    int vis_dtype;
    const int *field;
    const int *field_ids;   // [n_ids]
    int n_ids;
    const double *rotmat;   // [n_ids][3][3]
    const double *phrot;    // [n_ids][3]
    const double *freq;     // [n_chan]
    int64_t n_time, n_baseline, n_chan, n_pol;
    int common_tangent, single_precision;
    double *uvw_out;
    void *vis_out;
    int *time_field;        // [n_time] index into field_ids
    int *error;
};

// per integration: the one field id of its valid rows (>= 0); rows with the
// integer NaN are ignored; any other negative id, no valid row, two different
// ids or an id absent from field_ids is an error
__global__ void __launch_bounds__(256) time_field_kernel(RotArgs a) {
    __shared__ int s_min[256], s_max[256], s_bad[256];
    const int64_t t = blockIdx.x;
    int mn = INT32_MAX, mx = -1, bad = 0;
    for (int64_t b = threadIdx.x; b < a.n_baseline; b += blockDim.x) {
        const int f = a.field[t * a.n_baseline + b];
        if (f >= 0) {
            mn = min(mn, f);
            mx = max(mx, f);
        } else if (f != kIntNan) {
            bad = 1;
        }
    }
    s_min[threadIdx.x] = mn;
    s_max[threadIdx.x] = mx;
    s_bad[threadIdx.x] = bad;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_min[threadIdx.x] = min(s_min[threadIdx.x], s_min[threadIdx.x + w]);
            s_max[threadIdx.x] = max(s_max[threadIdx.x], s_max[threadIdx.x + w]);
            s_bad[threadIdx.x] |= s_bad[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int idx = -1;
        for (int k = 0; k < a.n_ids; k++)
            if (a.field_ids[k] == s_min[0]) {
                idx = k;
                break;
            }
        if (s_bad[0] || s_max[0] < 0 || s_min[0] != s_max[0] || idx < 0) {
            atomicOr(a.error, 1);
            idx = 0;
        }
        a.time_field[t] = idx;
    }
}

__global__ void __launch_bounds__(256) rotate_kernel(RotArgs a) {
#pragma clang fp contract(off)
    const int64_t n = a.n_time * a.n_baseline * a.n_chan;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ch = i % a.n_chan, row = i / a.n_chan;
        const int64_t t = row / a.n_baseline;
        const int k = a.time_field[t];
        const double *M = a.rotmat + 9 * k;
        const double *P = a.phrot + 3 * k;
        const double u = a.uvw[row * 3], v = a.uvw[row * 3 + 1], w = a.uvw[row * 3 + 2];
        // row vector times matrix: uvw_rot[j] = sum_i uvw[i] M[i][j]
        double r[3];
        for (int j = 0; j < 3; j++) r[j] = u * M[j] + v * M[3 + j] + w * M[6 + j];
        if (ch == 0) {
            a.uvw_out[row * 3] = r[0];
            a.uvw_out[row * 3 + 1] = r[1];
            a.uvw_out[row * 3 + 2] = r[2];
        }
        double pd = r[0] * P[0] + r[1] * P[1];
        if (!a.common_tangent) pd = pd + r[2] * P[2];
        const double ph = 2.0 * M_PI * pd * a.freq[ch] / kCLight;
        double s, c;
        sincos(ph, &s, &c);
        for (int64_t p = 0; p < a.n_pol; p++) {
            const int64_t e = i * a.n_pol + p;
            double vr, vi;
            if (a.vis_dtype == NGC_C64) {
                const float2 x = ((const float2 *)a.vis)[e];
                vr = x.x;
                vi = x.y;
            } else {
                const double2 x = ((const double2 *)a.vis)[e];
                vr = x.x;
                vi = x.y;
            }
            const double orr = vr * c - vi * s, oi = vr * s + vi * c;
            if (a.single_precision)
                ((float2 *)a.vis_out)[e] = make_float2((float)orr, (float)oi);
            else
                ((double2 *)a.vis_out)[e] = make_double2(orr, oi);
        }
    }
}

}  // namespace ngc

using namespace ngc;

extern "C" int ngc_direction_rotate(ngc_context *ctx, void *stream, const double *uvw,
                                    const void *vis, int vis_dtype, const int32_t *field,
                                    const int32_t *field_ids, int64_t n_ids,
                                    const double *uvw_rotmat, const double *phase_rotation,
                                    const double *freq_chan, int64_t n_time, int64_t n_baseline,
                                    int64_t n_chan, int64_t n_pol, int common_tangent,
                                    int single_precision, double *uvw_out, void *vis_out) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(n_time >= 0 && n_baseline >= 0 && n_chan >= 1 && n_pol >= 1 && n_ids >= 1 &&
                  n_ids < (1 << 30),
              NGC_ERR_ARG, "bad dimensions");
    NGC_CHECK(vis_dtype == NGC_C64 || vis_dtype == NGC_C128, NGC_ERR_ARG,
              "vis dtype must be C64 or C128");
    NGC_CHECK(field_ids && uvw_rotmat && phase_rotation && freq_chan, NGC_ERR_ARG,
              "NULL table");
    const int64_t n_rows = n_time * n_baseline;
    if (n_rows == 0) return NGC_OK;
    NGC_CHECK(uvw && vis && field && uvw_out && vis_out, NGC_ERR_ARG, "NULL array");
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(enter_stream(ctx, s));
    void *d_ids, *d_rot, *d_ph, *d_fr, *d_tf;
    NGC_TRY(upload(ctx, s, "rot_field_ids", field_ids, n_ids * sizeof(int32_t), &d_ids));
    NGC_TRY(upload(ctx, s, "rot_mats", uvw_rotmat, n_ids * 9 * sizeof(double), &d_rot));
    NGC_TRY(upload(ctx, s, "rot_phase", phase_rotation, n_ids * 3 * sizeof(double), &d_ph));
    NGC_TRY(upload(ctx, s, "rot_freq", freq_chan, n_chan * sizeof(double), &d_fr));
    NGC_TRY(workspace(ctx, "rot_time_field", n_time * sizeof(int), &d_tf));
    NGC_HIP(hipMemsetAsync(ctx->d_error, 0, sizeof(int), s));
    RotArgs a{};
    a.uvw = uvw;
    a.vis = vis;
    a.vis_dtype = vis_dtype;
    a.field = field;
    a.field_ids = (const int *)d_ids;
    a.n_ids = (int)n_ids;
    a.rotmat = (const double *)d_rot;
    a.phrot = (const double *)d_ph;
    a.freq = (const double *)d_fr;
    a.n_time = n_time;
    a.n_baseline = n_baseline;
    a.n_chan = n_chan;
    a.n_pol = n_pol;
    a.common_tangent = common_tangent ? 1 : 0;
    a.single_precision = single_precision ? 1 : 0;
    a.uvw_out = uvw_out;
    a.vis_out = vis_out;
    a.time_field = (int *)d_tf;
    a.error = ctx->d_error;
    NGC_CHECK(n_time < (int64_t(1) << 31), NGC_ERR_UNSUPPORTED, "too many integrations");
    profile_new_call(ctx, s);
    phase_begin(ctx, s, "rotate_fields");
    time_field_kernel<<<(unsigned)n_time, 256, 0, s>>>(a);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    int err = 0;
    NGC_HIP(hipMemcpyAsync(&err, ctx->d_error, sizeof(int), hipMemcpyDeviceToHost, s));
    NGC_HIP(hipStreamSynchronize(s));
    NGC_CHECK(err == 0, NGC_ERR_FIELD,
              "direction_rotate needs one field id per integration, present in field_ids");
    const int64_t n = n_rows * n_chan;
    int64_t nb = ceil_div(n, 256);
    if (nb > (1 << 20)) nb = 1 << 20;
    phase_begin(ctx, s, "rotate");
    rotate_kernel<<<(int)nb, 256, 0, s>>>(a);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}
