// Visibility feed: flags applied to data variables on the GPU.
//
//   ngc_apply_flags  the per-variable body of apply_flags (cngi/vis/apply_flags.py:49-55):
//                    out = data where flag == 0, else the dtype's value of NaN
//                    (xarray's where() gives NaN, then astype() back to the
//                    variable's dtype: NaN for floating and complex types,
//                    the integer minimum for integers, True for bool).
//
// A streaming, HBM-bound pass (read data + flag, write data): one thread per
// 16 bytes of output where the element size allows, grid-stride.
#include "ngc_internal.hpp"

namespace ngc {

template <typename T>
__device__ inline T flagged_value();
template <>
__device__ inline float flagged_value<float>() { return __int_as_float(0x7fc00000); }
template <>
__device__ inline double flagged_value<double>() {
    return __longlong_as_double(0x7ff8000000000000LL);
}
template <>
__device__ inline float2 flagged_value<float2>() {
    const float q = __int_as_float(0x7fc00000);
    return make_float2(q, q);
}
template <>
__device__ inline double2 flagged_value<double2>() {
    const double q = __longlong_as_double(0x7ff8000000000000LL);
    return make_double2(q, q);
}
template <>
__device__ inline int32_t flagged_value<int32_t>() { return INT32_MIN; }
template <>
__device__ inline int64_t flagged_value<int64_t>() { return INT64_MIN; }
template <>
__device__ inline uint8_t flagged_value<uint8_t>() { return 1; }

template <typename T>
__global__ void apply_flags_kernel(const T *data, const uint8_t *flag, int64_t n, T *out) {
    const T q = flagged_value<T>();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = flag[i] ? q : data[i];
}

static int feed_blocks(int64_t n) {
    const int64_t b = ceil_div(n, 256);
    return (int)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

template <typename T>
static int launch_apply(hipStream_t s, const void *data, const uint8_t *flag, int64_t n,
                        void *out) {
    apply_flags_kernel<T><<<feed_blocks(n), 256, 0, s>>>((const T *)data, flag, n, (T *)out);
    return hipGetLastError() == hipSuccess ? NGC_OK : NGC_ERR_HIP;
}

}  // namespace ngc

using namespace ngc;

extern "C" int ngc_apply_flags(ngc_context *ctx, void *stream, const void *data, int dtype,
                               const uint8_t *flag, int64_t n, void *out) {
    NGC_CHECK(ctx, NGC_ERR_ARG, "ctx is NULL");
    NGC_CHECK(n >= 0, NGC_ERR_ARG, "negative element count");
    if (n == 0) return NGC_OK;
    NGC_CHECK(data && flag && out, NGC_ERR_ARG, "NULL array");
    NGC_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    int rc;
    switch (dtype) {
        case NGC_C64: rc = launch_apply<float2>(s, data, flag, n, out); break;
        case NGC_C128: rc = launch_apply<double2>(s, data, flag, n, out); break;
        case NGC_F32: rc = launch_apply<float>(s, data, flag, n, out); break;
        case NGC_F64: rc = launch_apply<double>(s, data, flag, n, out); break;
        case NGC_I32: rc = launch_apply<int32_t>(s, data, flag, n, out); break;
        case NGC_I64: rc = launch_apply<int64_t>(s, data, flag, n, out); break;
        case NGC_BOOL: rc = launch_apply<uint8_t>(s, data, flag, n, out); break;
        default: NGC_CHECK(false, NGC_ERR_ARG, "unsupported dtype");
    }
    NGC_CHECK(rc == NGC_OK, NGC_ERR_HIP, "apply_flags launch failed");
    return NGC_OK;
}
