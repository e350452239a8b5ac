// Internal declarations of libngcasa_hip: context, error reporting, device
// helpers shared by the gridding kernels.  Written for gfx950 (CDNA4):
// 64-lane wavefronts, 160 KB LDS per CU, FP64 atomics in LDS and L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "ngcasa_gridding.h"

namespace ngc {

constexpr double kCLight = 299792458.0;

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);

struct Status {
    int code = NGC_OK;
    explicit operator bool() const { return code == NGC_OK; }
};

#define NGC_HIP(call)                                                           \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) {                                                 \
            ::ngc::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return NGC_ERR_HIP;                                                 \
        }                                                                       \
    } while (0)

#define NGC_CHECK(cond, code, msg)     \
    do {                               \
        if (!(cond)) {                 \
            ::ngc::set_error(msg);     \
            return (code);             \
        }                              \
    } while (0)

#define NGC_TRY(expr)            \
    do {                         \
        int rc_ = (expr);        \
        if (rc_ != NGC_OK)       \
            return rc_;          \
    } while (0)

// ---------------------------------------------------------------------------
// context: device, workspace arena, cached FFT plans
// ---------------------------------------------------------------------------
struct FftPlan;

struct Buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace ngc

struct ngc_profile_mark {
    std::string name;
    hipEvent_t start, stop;
};

struct ngc_context {
    int device = 0;
    bool profile = false;
    hipStream_t profile_stream = nullptr;
    std::vector<ngc_profile_mark> marks;   // phases of the last call
    std::vector<hipEvent_t> event_pool;
    std::mutex mu;
    hipStream_t last_stream = nullptr;
    bool has_last_stream = false;
    // named workspace buffers, grown on demand, freed with the context
    std::map<std::string, ngc::Buffer> ws;
    // rocFFT plans keyed by (n_u, n_v, batch, double_precision)
    std::map<std::tuple<int64_t, int64_t, int64_t, int>, ngc::FftPlan *> plans;
    int *d_error = nullptr;  // device error flags [4]
    int tile_rows = 0;       // uv-tile rows of the tiled gridder (8, 16 or 32; 0: by size)
    int bin_chan_lanes = 2;  // channels per binning wavefront (1, 2, 4, 8)
    int tile_waves = 0;      // wavefronts per tile of the tiled gridder (0: default)
    int n_cu = 256;          // compute units of the device
    int fft_lds = 1;         // single-precision image transforms: LDS kernels where supported
    int degrid_tiled = 1;    // standard degridding: LDS tiles where supported
    // autotuned (tile rows, binning channel lanes) per gridding shape
    std::map<std::string, std::pair<int, int>> tuned;
};

// A captured gridding pipeline: its graph, the device buffers it owns and its
// device flags ([1]: the inputs bin more records than the buffers hold;
// [3]: a row's field id is not in field_id).
struct ngc_grid_plan {
    ngc_context *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t cap_stream = nullptr;
    std::vector<void *> owned;
    int *flags = nullptr;
    std::string info;
};

namespace ngc {

void plan_free(ngc_grid_plan *p);
int plan_alloc(ngc_grid_plan *p, size_t bytes, void **out);
int plan_upload(ngc_grid_plan *p, hipStream_t s, const void *host, size_t bytes, void **out);

// Returns a device buffer of at least `bytes` under `name`; contents are
// undefined.  The context must be locked.
int workspace(ngc_context *ctx, const char *name, size_t bytes, void **out);

// Switch the context to `stream`, synchronising the previous stream if it
// differs (workspace buffers are reused across calls).
int enter_stream(ngc_context *ctx, hipStream_t stream);

// Copy a small host array to a named device workspace buffer (synchronous
// w.r.t. the host buffer).
int upload(ngc_context *ctx, hipStream_t stream, const char *name,
           const void *host, size_t bytes, void **dev);

// An int32 map of n entries followed by one zero entry (kernels that look one
// channel or baseline ahead read the pad); the caller's array itself is read
// for exactly n entries.
inline std::vector<int32_t> padded_map(const int32_t *map, int64_t n) {
    std::vector<int32_t> v((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; i++) v[(size_t)i] = map[i];
    return v;
}

// Device address of the allocation-free gridder's scratch array on `device`
// (which must be current), looked up once per device; called by
// ngc_context_create.
int noalloc_scratch_setup(int device);

int fft_destroy_plans(ngc_context *ctx);
// The LDS image transform (lds_fft.hip): axis lengths with prime factors
// <= 13 up to 8192; single-precision grids.  d_corr: device [n_l + n_m] or NULL.
bool lds_fft_supported(int64_t n);
int lds_grid_to_image(ngc_context *ctx, hipStream_t s, const void *grid, int dt,
                      int64_t n_planes, int64_t n_u, int64_t n_v, int64_t n_l, int64_t n_m,
                      const double *sum_weight, const double *d_corr, double *image);
int lds_model_to_grid(ngc_context *ctx, hipStream_t s, const double *image, int64_t n_planes,
                      int64_t n_l, int64_t n_m, int64_t n_u, int64_t n_v, const double *d_corr,
                      void *grid);
int fft_execute(ngc_context *ctx, hipStream_t s, void *buf, int64_t n_u, int64_t n_v,
                int64_t batch, bool dbl, bool forward);

// Phase timing: begin a new call (clears the previous marks), then bracket
// phases with phase_begin/phase_end.  No-ops unless profiling is enabled.
void profile_new_call(ngc_context *ctx, hipStream_t s);
void phase_begin(ngc_context *ctx, hipStream_t s, const char *name);
void phase_end(ngc_context *ctx, hipStream_t s);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// One validated ngc_aperture_grid call in a tiled precision form (device
// copies of the small tables already uploaded).
struct ApTiledCall {
    int mode, precision;
    const double *uvw;
    const void *vis;
    int vis_dtype;
    const void *weight;
    int weight_dtype;
    const uint8_t *flag;
    const int *field, *d_field_id;
    const int *d_field_map;   // index in field_id of each field id (first match), or -1
    int64_t field_map_len;    // 0 when there is no map (ids too large)
    int64_t n_field;
    const double *d_uv_scale;
    const int *d_chan_map, *d_cfb, *d_cfc, *d_support;
    const int32_t *support_host;  // [n_cf][2], the caller's host copy
    int64_t n_cfc, n_cfp, cfp0, n_cf;
    const double *conv_kernel, *phase_gradient;
    int64_t k_u, k_v;
    int64_t n_time, n_baseline, n_chan, n_pol, n_imag_chan, n_imag_pol, n_u, n_v;
    int64_t os_u, os_v, max_support;
    bool pol_identity, cfp_uniform;
    int accumulate;
    void *grid;
    double *sum_weight;
};
int ap_tiled_run(ngc_context *ctx, hipStream_t s, const ApTiledCall &c);
// The tiled pipeline of one call captured into plan p (image and PSF modes;
// the tables of c must be owned by p).
int ap_plan_create(ngc_context *ctx, hipStream_t s, const ApTiledCall &c, int autotune,
                   ngc_grid_plan *p);

// uv scale per channel, in the reference's operation order:
// -(freq*delta*n_uv)/c  (_standard_grid.py:274-276)
inline void uv_scale_host(const double *freq, int64_t n_chan, double delta_l,
                          double delta_m, int64_t n_u, int64_t n_v,
                          std::vector<double> &out) {
    out.resize(2 * n_chan);
    for (int64_t c = 0; c < n_chan; c++) {
        out[c] = -((freq[c] * delta_l) * (double)n_u) / kCLight;
        out[n_chan + c] = -((freq[c] * delta_m) * (double)n_v) / kCLight;
    }
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
template <typename T> struct VisLoad;
template <> struct VisLoad<float2> {
    __device__ static inline double2 ld(const void *p, int64_t i) {
        float2 v = reinterpret_cast<const float2 *>(p)[i];
        return make_double2(v.x, v.y);
    }
};
template <> struct VisLoad<double2> {
    __device__ static inline double2 ld(const void *p, int64_t i) {
        return reinterpret_cast<const double2 *>(p)[i];
    }
};
template <typename T> __device__ inline double ld_weight(const void *p, int64_t i) {
    return (double)reinterpret_cast<const T *>(p)[i];
}

// Unsigned 32-bit division by a divisor fixed at launch, as a multiply-high,
// a subtract and two shifts (T. Granlund and P. L. Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, the round-up method of
// their figure 4.1): with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0) where t = mulhi(m, n),
// exact for every n < 2^32.
struct FastDiv {
    uint32_t d = 1, m = 1;
    int s1 = 0, s2 = 0;
    // d = 0 (an empty axis, where nothing is divided) is taken as 1
    static FastDiv of(uint32_t d) {
        FastDiv f;
        if (d == 0) d = 1;
        f.d = d;
        int l = 0;
        while ((uint64_t(1) << l) < d) l++;
        f.m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
        f.s1 = l < 1 ? l : 1;
        f.s2 = l > 1 ? l - 1 : 0;
        return f;
    }
    __host__ __device__ uint32_t div(uint32_t n) const {
        const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
        return (t + ((n - t) >> s1)) >> s2;
    }
};

// Pixel position of one (row, channel) sample, with the reference's exact
// FP64 operation order and rounding (_standard_grid.py:299-324):
//   u = uvw_u*scale; u_pos = u + n_u//2; centre = int(u_pos+0.5) (truncation);
//   offset index = floor((centre - u_pos)*os + 0.5).
// No multiply-add is fused here, so the integer results match the reference
// bit for bit.  Returns false when u or v is NaN.
struct Pix {
    int64_t uc, vc;        // centre cell
    int64_t ucj, vcj;      // centre cell of the conjugate position
    int64_t du, dv;        // oversampled offset index
};

__device__ inline bool pixel_position(double uvw_u, double uvw_v, double s0,
                                      double s1, int64_t n_u, int64_t n_v,
                                      int64_t os, Pix &p) {
#pragma clang fp contract(off)
    const double u = uvw_u * s0;
    const double v = uvw_v * s1;
    if (isnan(u) || isnan(v)) return false;
    const double cu = (double)(n_u / 2), cv = (double)(n_v / 2);
    const double u_pos = u + cu;
    const double v_pos = v + cv;
    const double u_pos_conj = -u + cu;
    const double v_pos_conj = -v + cv;
    p.uc = (int64_t)(u_pos + 0.5);
    p.vc = (int64_t)(v_pos + 0.5);
    p.ucj = (int64_t)(u_pos_conj + 0.5);
    p.vcj = (int64_t)(v_pos_conj + 0.5);
    const double uo = (double)p.uc - u_pos;
    const double vo = (double)p.vc - v_pos;
    p.du = (int64_t)floor(uo * (double)os + 0.5);
    p.dv = (int64_t)floor(vo * (double)os + 0.5);
    return true;
}

__device__ inline int64_t iabs64(int64_t x) { return x < 0 ? -x : x; }

// wave-level reduction of a double (64 lanes)
__device__ inline double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

}  // namespace ngc
