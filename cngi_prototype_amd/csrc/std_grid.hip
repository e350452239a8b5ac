// Standard (prolate-spheroidal) gridder for gfx950.
//
// Replaces _standard_grid_jit (ngcasa/imaging/_imaging_utils/_standard_grid.py:243-371)
// behind ngc_standard_grid().  Two implementations:
//
//  * plain FP64 drop-in (NGC_PREC_FP64, and every density-mode call): one
//    thread per (time, baseline, channel) sample walks the pols and the
//    support^2 taps in the reference's loop order and scatters with FP64
//    atomics into the grid in HBM.
//
//  * uv-tiled (NGC_PREC_FP64_TILED, NGC_PREC_FP32_KAHAN): samples are binned
//    by (image channel, 32x32 uv tile) with a deterministic, order-preserving
//    counting sort (wave-aggregated atomics keep runs of consecutive
//    integrations of one baseline together).  One workgroup owns one tile: it
//    streams the tile's bin, accumulates every tap into an LDS subgrid (one
//    lane per tap of an 8x8 tap window) and writes each grid cell to HBM once.
//    In FP32 the LDS subgrid is folded into per-thread Kahan-Neumaier
//    accumulators after every batch of 256 samples.  Pixel positions and
//    sum_weight stay FP64.
#include <hipcub/hipcub.hpp>

#include "ngc_internal.hpp"

namespace ngc {

constexpr int kTile = 32;          // tile edge in cells (both axes)
constexpr int kBlock = 256;        // threads per workgroup
constexpr int kMaxTiledSupport = 8;

struct StdArgs {
    const double *uvw;
    const void *vis;
    const void *weight;
    const uint8_t *flag;
    const double *uv_scale;  // [2][n_chan]
    const int *chan_map;     // [n_chan]
    const int *pol_map;      // [n_pol]
    int64_t n_time, n_baseline, n_chan, n_pol;
    int64_t n_imag_chan, n_imag_pol, n_u, n_v;
    int64_t support, os;
    int vis_dtype, weight_dtype;
};

__device__ inline double2 load_vis(const void *p, int64_t i, int dt) {
    if (dt == NGC_C64) {
        float2 v = reinterpret_cast<const float2 *>(p)[i];
        return make_double2((double)v.x, (double)v.y);
    }
    return reinterpret_cast<const double2 *>(p)[i];
}

__device__ inline double load_w(const void *p, int64_t i, int dt) {
    if (dt == NGC_F32) return (double)reinterpret_cast<const float *>(p)[i];
    return reinterpret_cast<const double *>(p)[i];
}

// weight of pol p of sample `base` (= rc*n_pol), NaN when flagged
__device__ inline double weight_of(const StdArgs &a, int64_t base, int p) {
    double w = load_w(a.weight, base + p, a.weight_dtype);
    if (a.flag && a.flag[base + p]) w = __longlong_as_double(0x7ff8000000000000LL);
    return w;
}

// ---------------------------------------------------------------------------
// plain FP64 kernel: the reference loop, one thread per (row, channel)
// ---------------------------------------------------------------------------
template <int MODE>
__global__ void __launch_bounds__(kBlock)
std_grid_plain_kernel(StdArgs a, const double *cgk, int64_t cgk_len, double *grid,
                      double *sum_weight, int sw_in_lds) {
#pragma clang fp contract(off)
    extern __shared__ double s_sw[];
    const int64_t n_sw = a.n_imag_chan * a.n_imag_pol;
    if (sw_in_lds) {
        for (int64_t i = threadIdx.x; i < n_sw; i += blockDim.x) s_sw[i] = 0.0;
        __syncthreads();
    }
    const int64_t sc = a.support / 2;
    const int64_t start_s = -sc, end_s = a.support - sc;
    const int64_t plane = a.n_u * a.n_v;
    const int64_t cells = a.n_imag_chan * a.n_imag_pol * plane;
    const int64_t n_rc = a.n_time * a.n_baseline * a.n_chan;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rc;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ch = i % a.n_chan;
        const int64_t row = i / a.n_chan;
        Pix px;
        if (!pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                            a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os, px))
            continue;
        if (!((px.uc + sc < a.n_u) && (px.vc + sc < a.n_v) && (px.uc - sc >= 0) &&
              (px.vc - sc >= 0)))
            continue;
        const int64_t a_chan = a.chan_map[ch];
        const int64_t base = i * a.n_pol;
        for (int p = 0; p < a.n_pol; p++) {
            double wr, wi = 0.0, sel;
            if (MODE == NGC_GRID_IMAGE) {
                const double w = weight_of(a, base, p);
                const double2 v = load_vis(a.vis, base + p, a.vis_dtype);
                sel = w;
                wr = v.x * w;
                wi = v.y * w;
            } else if (MODE == NGC_GRID_DENSITY && a.n_pol >= 2) {
                wr = (weight_of(a, base, 0) + weight_of(a, base, 1)) / 2.0;
                sel = wr;
            } else {
                wr = weight_of(a, base, p);
                sel = wr;
            }
            if (isnan(wr) || isnan(wi)) continue;
            if (wr == 0.0 && wi == 0.0) continue;
            const int64_t a_pol = a.pol_map[p];
            const int64_t pbase = (a_chan * a.n_imag_pol + a_pol) * plane;
            double norm = 0.0;
            for (int64_t iv = start_s; iv < end_s; iv++) {
                const int64_t v_indx = px.vc + iv;
                const int64_t kv = iabs64(a.os * iv + px.dv);
                const double conv_v = kv < cgk_len ? cgk[kv] : 0.0;
                const int64_t v_indx_conj = px.vcj + iv;
                for (int64_t iu = start_s; iu < end_s; iu++) {
                    const int64_t u_indx = px.uc + iu;
                    const int64_t ku = iabs64(a.os * iu + px.du);
                    const double conv_u = ku < cgk_len ? cgk[ku] : 0.0;
                    const double conv = conv_u * conv_v;
                    const int64_t cell = pbase + u_indx * a.n_v + v_indx;
                    if (MODE == NGC_GRID_IMAGE) {
                        atomicAdd(&grid[2 * cell], conv * wr);
                        atomicAdd(&grid[2 * cell + 1], conv * wi);
                    } else {
                        atomicAdd(&grid[cell], conv * wr);
                    }
                    norm = norm + conv;
                    if (MODE == NGC_GRID_DENSITY) {
                        // the reference does not bounds-check the conjugate
                        // cell; numba lands it at its flat C-order offset.  Keep
                        // that when the offset lies inside the grid, drop it
                        // otherwise.
                        const int64_t u_indx_conj = px.ucj + iu;
                        const int64_t ccell = pbase + u_indx_conj * a.n_v + v_indx_conj;
                        if (ccell >= 0 && ccell < cells) atomicAdd(&grid[ccell], conv * wr);
                    }
                }
            }
            const int64_t sw = a_chan * a.n_imag_pol + a_pol;
            double add = sel * norm;
            if (sw_in_lds) {
                atomicAdd(&s_sw[sw], add);
                if (MODE == NGC_GRID_DENSITY) atomicAdd(&s_sw[sw], add);
            } else {
                atomicAdd(&sum_weight[sw], add);
                if (MODE == NGC_GRID_DENSITY) atomicAdd(&sum_weight[sw], add);
            }
        }
    }
    if (sw_in_lds) {
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n_sw; i += blockDim.x)
            if (s_sw[i] != 0.0) atomicAdd(&sum_weight[i], s_sw[i]);
    }
}

// ---------------------------------------------------------------------------
// binning: deterministic counting sort of samples into (image chan, tile) bins
// ---------------------------------------------------------------------------
// Samples are enumerated channel-major, then baseline, then time, so one
// wavefront usually covers 64 consecutive integrations of one baseline track.
struct BinGeom {
    int64_t ntu, ntv;  // tiles per axis
};

__device__ inline bool sample_of(const StdArgs &a, int64_t i, int64_t &rc, Pix &px,
                                 int64_t &a_chan) {
    const int64_t bt = a.n_baseline * a.n_time;
    const int64_t ch = i / bt;
    const int64_t rem = i - ch * bt;
    const int64_t b = rem / a.n_time;
    const int64_t t = rem - b * a.n_time;
    const int64_t row = t * a.n_baseline + b;
    rc = row * a.n_chan + ch;
    if (!pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                        a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os, px))
        return false;
    const int64_t sc = a.support / 2;
    if (!((px.uc + sc < a.n_u) && (px.vc + sc < a.n_v) && (px.uc - sc >= 0) &&
          (px.vc - sc >= 0)))
        return false;
    if (a.flag) {
        bool any = false;
        for (int p = 0; p < a.n_pol; p++) any |= (a.flag[rc * a.n_pol + p] == 0);
        if (!any) return false;
    }
    a_chan = a.chan_map[ch];
    return true;
}

// k-th overlapped tile (k = 0..3) of a valid sample, or -1
__device__ inline int64_t tile_key(const StdArgs &a, const BinGeom &g, const Pix &px,
                                   int64_t a_chan, int k) {
    const int64_t sc = a.support / 2;
    const int64_t tu0 = (px.uc - sc) / kTile, tu1 = (px.uc + sc) / kTile;
    const int64_t tv0 = (px.vc - sc) / kTile, tv1 = (px.vc + sc) / kTile;
    const int64_t tu = tu0 + (k & 1), tv = tv0 + (k >> 1);
    if (tu > tu1 || tv > tv1) return -1;
    return (a_chan * g.ntu + tu) * g.ntv + tv;
}

template <bool FILL>
__global__ void __launch_bounds__(kBlock)
std_bin_kernel(StdArgs a, BinGeom g, unsigned long long *counts, int4 *entries) {
    const int64_t n = a.n_time * a.n_baseline * a.n_chan;
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n;
         i0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        int64_t rc = 0, a_chan = 0;
        Pix px{};
        const bool ok = (i < n) && sample_of(a, i, rc, px, a_chan);
        for (int k = 0; k < 4; k++) {
            int64_t key = ok ? tile_key(a, g, px, a_chan, k) : -1;
            uint64_t pending = __ballot(key >= 0);
            while (pending) {
                const int leader = __ffsll((long long)pending) - 1;
                const int64_t lkey = __shfl(key, leader, 64);
                const uint64_t same = __ballot(key == lkey);
                unsigned long long basev = 0;
                if (lane == leader)
                    basev = atomicAdd(&counts[lkey], (unsigned long long)__popcll(same));
                if (FILL) {
                    basev = __shfl(basev, leader, 64);
                    if (key == lkey) {
                        const unsigned long long pos = basev + __popcll(same & below);
                        const int64_t hu = px.uc / kTile, hv = px.vc / kTile;
                        const int64_t hkey = (a_chan * g.ntu + hu) * g.ntv + hv;
                        int4 e;
                        e.x = (int)rc;
                        e.y = (int)((px.uc & 0xffff) | (px.vc << 16));
                        e.z = (int)((px.du & 0xffff) | ((px.dv & 0xffff) << 16));
                        e.w = (hkey == lkey) ? 1 : 0;
                        entries[pos] = e;
                    }
                }
                if (key == lkey) key = -1;
                pending &= ~same;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// tiled gridding kernel: one workgroup per (image channel, tile) bin
// ---------------------------------------------------------------------------
template <typename ACC> struct AccTraits;
template <> struct AccTraits<float> {
    static constexpr bool kKahan = true;
    static constexpr int kPad = 1;
};
template <> struct AccTraits<double> {
    static constexpr bool kKahan = false;
    static constexpr int kPad = 1;
};

template <int NP, int NC, typename ACC> struct Stage {
    int u0, v0, du, dv;
    ACC wd[NP][NC];
};

template <int MODE, int NP, typename ACC>
__global__ void __launch_bounds__(kBlock)
std_grid_tiled_kernel(StdArgs a, BinGeom g, const unsigned long long *offsets,
                      const int4 *entries, const double *cgk, int cgk_len,
                      const double *norm_tab, int norm_half, void *grid_out,
                      double *sum_weight, int accumulate) {
    constexpr int NC = (MODE == NGC_GRID_IMAGE) ? 2 : 1;
    constexpr int RS = kTile * NC + AccTraits<ACC>::kPad;  // LDS row stride
    constexpr int F = NP * kTile * kTile * NC;              // accumulators per tile
    constexpr int PER = F / kBlock;                          // owned per thread
    static_assert(F % kBlock == 0, "tile must split evenly over the workgroup");

    __shared__ ACC s_acc[NP * kTile * RS];
    __shared__ Stage<NP, NC, ACC> s_stage[kBlock / 64][64];
    __shared__ ACC s_k[512];
    __shared__ double s_sw[kBlock / 64][NP];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t bin = blockIdx.x;
    const int64_t tiles = g.ntu * g.ntv;
    const int64_t a_chan = bin / tiles;
    const int64_t trem = bin - a_chan * tiles;
    const int64_t u_lo = (trem / g.ntv) * kTile;
    const int64_t v_lo = (trem % g.ntv) * kTile;

    for (int i = tid; i < NP * kTile * RS; i += kBlock) s_acc[i] = (ACC)0;
    for (int i = tid; i < cgk_len && i < 512; i += kBlock) s_k[i] = (ACC)cgk[i];

    ACC ksum[PER], kcomp[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) { ksum[j] = 0; kcomp[j] = 0; }
    double swp[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) swp[p] = 0.0;
    __syncthreads();

    const int sc = (int)(a.support / 2);
    const int S = (int)a.support;
    const int os = (int)a.os;
    const int ta = lane & 7, tb = lane >> 3;
    const bool tap_on = (ta < S) && (tb < S);
    const int iu = ta - sc, iv = tb - sc;
    const int nw = 2 * norm_half + 1;

    const int64_t start = (int64_t)offsets[bin], end = (int64_t)offsets[bin + 1];
    for (int64_t b0 = start; b0 < end; b0 += kBlock) {
        const int64_t e = b0 + wave * 64 + lane;
        if (e < end) {
            const int4 en = entries[e];
            const int64_t rc = (int64_t)(unsigned)en.x;
            Stage<NP, NC, ACC> st;
            st.u0 = en.y & 0xffff;
            st.v0 = (en.y >> 16) & 0xffff;
            st.du = (int)(short)(en.z & 0xffff);
            st.dv = (int)(short)((en.z >> 16) & 0xffff);
            const bool home = en.w & 1;
            const int64_t base = rc * NP;
            const int ndu = min(max(st.du + norm_half, 0), nw - 1);
            const int ndv = min(max(st.dv + norm_half, 0), nw - 1);
            const double norm = home ? norm_tab[ndu * nw + ndv] : 0.0;
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const double w = weight_of(a, base, p);
                double wr = w, wi = 0.0;
                if (MODE == NGC_GRID_IMAGE) {
                    const double2 v = load_vis(a.vis, base + p, a.vis_dtype);
                    wr = v.x * w;
                    wi = v.y * w;
                }
                const bool use = !(isnan(wr) || isnan(wi)) && !(wr == 0.0 && wi == 0.0);
                st.wd[p][0] = use ? (ACC)wr : (ACC)0;
                if (NC == 2) st.wd[p][NC - 1] = use ? (ACC)wi : (ACC)0;
                if (use && home) swp[p] += w * norm;
            }
            s_stage[wave][lane] = st;
        }
        __syncthreads();
        const int64_t nh64 = end - (b0 + wave * 64);
        const int nh = nh64 > 64 ? 64 : (nh64 < 0 ? 0 : (int)nh64);
        for (int j = 0; j < nh; j++) {
            const Stage<NP, NC, ACC> st = s_stage[wave][j];
            if (tap_on) {
                const int lu = st.u0 + iu - (int)u_lo;
                const int lv = st.v0 + iv - (int)v_lo;
                if ((unsigned)lu < (unsigned)kTile && (unsigned)lv < (unsigned)kTile) {
                    const int ku = abs(os * iu + st.du);
                    const int kv = abs(os * iv + st.dv);
                    const ACC kw = (ku < cgk_len ? s_k[ku] : (ACC)0) * (kv < cgk_len ? s_k[kv] : (ACC)0);
                    ACC *cell = &s_acc[lu * RS + lv * NC];
#pragma unroll
                    for (int p = 0; p < NP; p++)
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            atomicAdd(cell + p * kTile * RS + c, kw * st.wd[p][c]);
                }
            }
        }
        __syncthreads();
        if (AccTraits<ACC>::kKahan) {
            // fold the batch's LDS partial sums into compensated registers
#pragma unroll
            for (int j = 0; j < PER; j++) {
                const int k = tid + j * kBlock;
                const int c = k % NC;
                const int lv = (k / NC) % kTile;
                const int lu = (k / (NC * kTile)) % kTile;
                const int p = k / (NC * kTile * kTile);
                ACC *src = &s_acc[(p * kTile + lu) * RS + lv * NC + c];
                const ACC x = *src;
                *src = (ACC)0;
                const ACC s = ksum[j];
                const ACC t = s + x;
                if (fabs(s) >= fabs(x))
                    kcomp[j] += (s - t) + x;
                else
                    kcomp[j] += (x - t) + s;
                ksum[j] = t;
            }
            __syncthreads();
        }
    }

    // write the tile: every cell of the grid is written exactly once
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int k = tid + j * kBlock;
        const int c = k % NC;
        const int lv = (k / NC) % kTile;
        const int lu = (k / (NC * kTile)) % kTile;
        const int p = k / (NC * kTile * kTile);
        const int64_t gu = u_lo + lu, gv = v_lo + lv;
        if (gu >= a.n_u || gv >= a.n_v) continue;
        ACC val;
        if (AccTraits<ACC>::kKahan)
            val = ksum[j] + kcomp[j];
        else
            val = s_acc[(p * kTile + lu) * RS + lv * NC + c];
        const int64_t gi = (((a_chan * a.n_imag_pol + p) * a.n_u + gu) * a.n_v + gv) * NC + c;
        ACC *gp = reinterpret_cast<ACC *>(grid_out) + gi;
        if (accumulate)
            *gp = *gp + val;
        else
            *gp = val;
    }

    // sum_weight: FP64 partials of the samples whose home tile this is
#pragma unroll
    for (int p = 0; p < NP; p++) {
        const double s = wave_sum(swp[p]);
        if (lane == 0) s_sw[wave][p] = s;
    }
    __syncthreads();
    if (tid < NP) {
        double s = 0.0;
        for (int w = 0; w < kBlock / 64; w++) s += s_sw[w][tid];
        if (s != 0.0) atomicAdd(&sum_weight[a_chan * a.n_imag_pol + tid], s);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int launch_blocks(int64_t n) {
    int64_t b = ceil_div(n, kBlock);
    if (b > 1 << 20) b = 1 << 20;
    if (b < 1) b = 1;
    return (int)b;
}

// Sum of the support^2 kernel taps for every oversampled offset pair, in the
// reference's loop order (_standard_grid.py:342-360).  The per-sample
// sum_weight term sel_weight*norm is then bit-identical to the reference's.
static void norm_table(const double *cgk, int64_t cgk_len, int64_t support, int64_t os, int half,
                       std::vector<double> &tab) {
    const int nw = 2 * half + 1;
    tab.assign((size_t)nw * nw, 0.0);
    const int64_t sc = support / 2;
    for (int du = -half; du <= half; du++)
        for (int dv = -half; dv <= half; dv++) {
            double norm = 0.0;
            for (int64_t iv = -sc; iv < support - sc; iv++) {
                const int64_t kv = std::llabs(os * iv + dv);
                const double cv = kv < cgk_len ? cgk[kv] : 0.0;
                for (int64_t iu = -sc; iu < support - sc; iu++) {
                    const int64_t ku = std::llabs(os * iu + du);
                    const double cu = ku < cgk_len ? cgk[ku] : 0.0;
                    volatile double conv = cu * cv;
                    norm = norm + conv;
                }
            }
            tab[(size_t)(du + half) * nw + (dv + half)] = norm;
        }
}

template <int MODE, typename ACC>
static int launch_tiled(int64_t np, dim3 grid, hipStream_t s, const StdArgs &a,
                        const BinGeom &g, const unsigned long long *off,
                        const int4 *ent, const double *cgk, int cgk_len,
                        const double *ntab, int half, void *out, double *sw,
                        int accumulate) {
    switch (np) {
    case 1:
        std_grid_tiled_kernel<MODE, 1, ACC><<<grid, kBlock, 0, s>>>(
            a, g, off, ent, cgk, cgk_len, ntab, half, out, sw, accumulate);
        break;
    case 2:
        std_grid_tiled_kernel<MODE, 2, ACC><<<grid, kBlock, 0, s>>>(
            a, g, off, ent, cgk, cgk_len, ntab, half, out, sw, accumulate);
        break;
    case 3:
        std_grid_tiled_kernel<MODE, 3, ACC><<<grid, kBlock, 0, s>>>(
            a, g, off, ent, cgk, cgk_len, ntab, half, out, sw, accumulate);
        break;
    case 4:
        std_grid_tiled_kernel<MODE, 4, ACC><<<grid, kBlock, 0, s>>>(
            a, g, off, ent, cgk, cgk_len, ntab, half, out, sw, accumulate);
        break;
    default:
        set_error("tiled gridder supports 1..4 polarisations");
        return NGC_ERR_UNSUPPORTED;
    }
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

static int standard_grid_impl(ngc_context *ctx, hipStream_t s, int mode, int precision,
                              StdArgs a, const double *freq, const int32_t *chan_map,
                              const int32_t *pol_map, const double *cgk, int64_t cgk_len,
                              int accumulate, void *grid, double *sum_weight,
                              double delta_l, double delta_m) {
    NGC_CHECK(mode == NGC_GRID_IMAGE || mode == NGC_GRID_PSF || mode == NGC_GRID_DENSITY,
              NGC_ERR_ARG, "mode must be IMAGE, PSF or DENSITY");
    NGC_CHECK(precision >= NGC_PREC_FP64 && precision <= NGC_PREC_FP32_KAHAN, NGC_ERR_ARG,
              "unknown precision");
    const int64_t n_samples = a.n_time * a.n_baseline * a.n_chan;
    NGC_CHECK(grid && sum_weight, NGC_ERR_ARG, "NULL grid or sum_weight");
    NGC_CHECK((a.uvw && a.weight) || n_samples == 0, NGC_ERR_ARG, "NULL uvw or weight");
    NGC_CHECK(mode != NGC_GRID_IMAGE || a.vis || n_samples == 0, NGC_ERR_ARG,
              "image mode needs vis");
    NGC_CHECK(a.vis_dtype == NGC_C64 || a.vis_dtype == NGC_C128 || mode != NGC_GRID_IMAGE,
              NGC_ERR_ARG, "vis dtype must be C64 or C128");
    NGC_CHECK(a.weight_dtype == NGC_F32 || a.weight_dtype == NGC_F64, NGC_ERR_ARG,
              "weight dtype must be F32 or F64");
    NGC_CHECK(a.n_time >= 0 && a.n_baseline >= 0 && a.n_chan >= 0 && a.n_pol >= 1 &&
                  a.n_imag_chan >= 1 && a.n_imag_pol >= 1 && a.n_u >= 1 && a.n_v >= 1 &&
                  a.support >= 1 && a.os >= 0,
              NGC_ERR_ARG, "bad dimensions");
    NGC_CHECK(freq && chan_map && pol_map && cgk && cgk_len >= 1, NGC_ERR_ARG,
              "NULL parameter array");
    for (int64_t c = 0; c < a.n_chan; c++)
        NGC_CHECK(chan_map[c] >= 0 && chan_map[c] < a.n_imag_chan, NGC_ERR_ARG,
                  "chan_map entry out of range");
    bool pol_identity = (a.n_imag_pol == a.n_pol);
    for (int64_t p = 0; p < a.n_pol; p++) {
        NGC_CHECK(pol_map[p] >= 0 && pol_map[p] < a.n_imag_pol, NGC_ERR_ARG,
                  "pol_map entry out of range");
        pol_identity = pol_identity && (pol_map[p] == p);
    }
    // every tap index |os*i + offset| must fall inside cgk_1d
    const int64_t sc = a.support / 2;
    const int64_t far = sc > a.support - 1 - sc ? sc : a.support - 1 - sc;
    const int64_t max_tap = a.os * far + (a.os + 1) / 2;
    NGC_CHECK(max_tap < cgk_len, NGC_ERR_ARG, "cgk_1d too short for support and oversampling");

    std::vector<double> scale;
    uv_scale_host(freq, a.n_chan, delta_l, delta_m, a.n_u, a.n_v, scale);
    void *d_scale, *d_cmap, *d_pmap, *d_cgk;
    NGC_TRY(upload(ctx, s, "uv_scale", scale.data(), scale.size() * sizeof(double), &d_scale));
    NGC_TRY(upload(ctx, s, "chan_map", chan_map, a.n_chan * sizeof(int32_t) + 4, &d_cmap));
    NGC_TRY(upload(ctx, s, "pol_map", pol_map, a.n_pol * sizeof(int32_t), &d_pmap));
    NGC_TRY(upload(ctx, s, "cgk", cgk, cgk_len * sizeof(double), &d_cgk));
    a.uv_scale = (const double *)d_scale;
    a.chan_map = (const int *)d_cmap;
    a.pol_map = (const int *)d_pmap;

    const bool single = (precision == NGC_PREC_FP32_KAHAN);
    const int64_t cells = a.n_imag_chan * a.n_imag_pol * a.n_u * a.n_v;
    const int64_t nc = (mode == NGC_GRID_IMAGE) ? 2 : 1;
    const int64_t n_sw = a.n_imag_chan * a.n_imag_pol;
    const int64_t n_rc = a.n_time * a.n_baseline * a.n_chan;

    const bool tiled = precision != NGC_PREC_FP64 && mode != NGC_GRID_DENSITY;
    profile_new_call(ctx, s);
    if (!tiled) {
        if (single) {
            // FP32 grid requested for a density grid: the density gridder is
            // FP64 only (it feeds FP64 Briggs factors)
            set_error("density gridding is FP64 only (precision NGC_PREC_FP64 or FP64_TILED)");
            return NGC_ERR_UNSUPPORTED;
        }
        if (!accumulate) {
            NGC_HIP(hipMemsetAsync(grid, 0, cells * nc * sizeof(double), s));
            NGC_HIP(hipMemsetAsync(sum_weight, 0, n_sw * sizeof(double), s));
        }
        if (n_rc == 0) return NGC_OK;
        const int sw_lds = n_sw <= 4096 ? 1 : 0;
        const size_t shmem = sw_lds ? n_sw * sizeof(double) : 0;
        const int blocks = launch_blocks(n_rc);
        double *g = (double *)grid;
        phase_begin(ctx, s, "std_grid_plain");
        if (mode == NGC_GRID_IMAGE)
            std_grid_plain_kernel<NGC_GRID_IMAGE><<<blocks, kBlock, shmem, s>>>(
                a, (const double *)d_cgk, cgk_len, g, sum_weight, sw_lds);
        else if (mode == NGC_GRID_PSF)
            std_grid_plain_kernel<NGC_GRID_PSF><<<blocks, kBlock, shmem, s>>>(
                a, (const double *)d_cgk, cgk_len, g, sum_weight, sw_lds);
        else
            std_grid_plain_kernel<NGC_GRID_DENSITY><<<blocks, kBlock, shmem, s>>>(
                a, (const double *)d_cgk, cgk_len, g, sum_weight, sw_lds);
        NGC_HIP(hipGetLastError());
        phase_end(ctx, s);
        return NGC_OK;
    }

    // ---------------- tiled path ----------------
    NGC_CHECK(a.support <= kMaxTiledSupport && a.support >= 2, NGC_ERR_UNSUPPORTED,
              "tiled gridder supports 2 <= support <= 8");
    NGC_CHECK(pol_identity, NGC_ERR_UNSUPPORTED, "tiled gridder needs pol_map = arange(n_pol)");
    NGC_CHECK(a.n_pol <= 4, NGC_ERR_UNSUPPORTED, "tiled gridder supports up to 4 pols");
    NGC_CHECK(a.n_u < 65536 && a.n_v < 65536, NGC_ERR_UNSUPPORTED, "grid axis >= 65536");
    NGC_CHECK(n_rc < (int64_t(1) << 31), NGC_ERR_UNSUPPORTED, "more than 2^31 samples per call");
    NGC_CHECK(a.os < 30000, NGC_ERR_UNSUPPORTED, "oversampling too large");
    NGC_CHECK(cgk_len <= 512, NGC_ERR_UNSUPPORTED, "cgk_1d longer than 512");

    const int half = (int)(a.os / 2 + 1);
    std::vector<double> ntab;
    norm_table(cgk, cgk_len, a.support, a.os, half, ntab);
    void *d_ntab;
    NGC_TRY(upload(ctx, s, "norm_tab", ntab.data(), ntab.size() * sizeof(double), &d_ntab));

    BinGeom g;
    g.ntu = ceil_div(a.n_u, kTile);
    g.ntv = ceil_div(a.n_v, kTile);
    const int64_t n_bins = a.n_imag_chan * g.ntu * g.ntv;
    NGC_CHECK(n_bins < (int64_t(1) << 31), NGC_ERR_UNSUPPORTED, "too many tiles");

    void *d_counts, *d_off, *d_scan_tmp, *d_ent;
    NGC_TRY(workspace(ctx, "bin_counts", (n_bins + 1) * sizeof(unsigned long long), &d_counts));
    NGC_TRY(workspace(ctx, "bin_offsets", (n_bins + 1) * sizeof(unsigned long long), &d_off));
    NGC_HIP(hipMemsetAsync(d_counts, 0, (n_bins + 1) * sizeof(unsigned long long), s));

    const int blocks = launch_blocks(n_rc);
    phase_begin(ctx, s, "bin_count");
    if (n_rc > 0) {
        std_bin_kernel<false><<<blocks, kBlock, 0, s>>>(a, g, (unsigned long long *)d_counts,
                                                          nullptr);
        NGC_HIP(hipGetLastError());
    }
    phase_end(ctx, s);
    phase_begin(ctx, s, "bin_scan");
    size_t tmp_bytes = 0;
    NGC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes,
                                             (unsigned long long *)d_counts,
                                             (unsigned long long *)d_off, n_bins + 1, s));
    NGC_TRY(workspace(ctx, "scan_tmp", tmp_bytes, &d_scan_tmp));
    NGC_HIP(hipcub::DeviceScan::ExclusiveSum(d_scan_tmp, tmp_bytes,
                                             (unsigned long long *)d_counts,
                                             (unsigned long long *)d_off, n_bins + 1, s));
    phase_end(ctx, s);
    unsigned long long total = 0;
    NGC_HIP(hipMemcpyAsync(&total, (unsigned long long *)d_off + n_bins, sizeof(total),
                           hipMemcpyDeviceToHost, s));
    NGC_HIP(hipStreamSynchronize(s));
    NGC_TRY(workspace(ctx, "bin_entries", (size_t)(total + 1) * sizeof(int4), &d_ent));
    // cursors start at the offsets
    NGC_HIP(hipMemcpyAsync(d_counts, d_off, (n_bins + 1) * sizeof(unsigned long long),
                           hipMemcpyDeviceToDevice, s));
    phase_begin(ctx, s, "bin_fill");
    if (n_rc > 0) {
        std_bin_kernel<true><<<blocks, kBlock, 0, s>>>(a, g, (unsigned long long *)d_counts,
                                                         (int4 *)d_ent);
        NGC_HIP(hipGetLastError());
    }
    phase_end(ctx, s);
    if (!accumulate) NGC_HIP(hipMemsetAsync(sum_weight, 0, n_sw * sizeof(double), s));

    dim3 gridDim((unsigned)n_bins);
    phase_begin(ctx, s, "std_grid_tiled");
    struct PhaseGuard {
        ngc_context *c;
        hipStream_t st;
        ~PhaseGuard() { phase_end(c, st); }
    } guard{ctx, s};
    const unsigned long long *off = (const unsigned long long *)d_off;
    const int4 *ent = (const int4 *)d_ent;
    if (single) {
        if (mode == NGC_GRID_IMAGE)
            return launch_tiled<NGC_GRID_IMAGE, float>(a.n_pol, gridDim, s, a, g, off, ent,
                                                       (const double *)d_cgk, (int)cgk_len,
                                                       (const double *)d_ntab, half, grid,
                                                       sum_weight, accumulate);
        return launch_tiled<NGC_GRID_PSF, float>(a.n_pol, gridDim, s, a, g, off, ent,
                                                 (const double *)d_cgk, (int)cgk_len,
                                                 (const double *)d_ntab, half, grid,
                                                 sum_weight, accumulate);
    }
    if (mode == NGC_GRID_IMAGE)
        return launch_tiled<NGC_GRID_IMAGE, double>(a.n_pol, gridDim, s, a, g, off, ent,
                                                    (const double *)d_cgk, (int)cgk_len,
                                                    (const double *)d_ntab, half, grid,
                                                    sum_weight, accumulate);
    return launch_tiled<NGC_GRID_PSF, double>(a.n_pol, gridDim, s, a, g, off, ent,
                                              (const double *)d_cgk, (int)cgk_len,
                                              (const double *)d_ntab, half, grid, sum_weight,
                                              accumulate);
}

}  // namespace ngc

extern "C" int ngc_standard_grid(ngc_context *ctx, void *stream, int mode, int precision,
                                 const double *uvw, const void *vis, int vis_dtype,
                                 const void *weight, int weight_dtype, const uint8_t *flag,
                                 const double *freq_chan, const int32_t *chan_map,
                                 const int32_t *pol_map, const double *cgk_1d,
                                 int64_t cgk_len, int64_t n_time, int64_t n_baseline,
                                 int64_t n_chan, int64_t n_pol, int64_t n_imag_chan,
                                 int64_t n_imag_pol, int64_t n_u, int64_t n_v,
                                 double delta_l, double delta_m, int64_t support,
                                 int64_t oversampling, int accumulate, void *grid,
                                 double *sum_weight) {
    if (!ctx) {
        ngc::set_error("ctx is NULL");
        return NGC_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    hipStream_t s = (hipStream_t)stream;
    NGC_TRY(ngc::enter_stream(ctx, s));
    ngc::StdArgs a{};
    a.uvw = uvw;
    a.vis = vis;
    a.weight = weight;
    a.flag = flag;
    a.n_time = n_time;
    a.n_baseline = n_baseline;
    a.n_chan = n_chan;
    a.n_pol = n_pol;
    a.n_imag_chan = n_imag_chan;
    a.n_imag_pol = n_imag_pol;
    a.n_u = n_u;
    a.n_v = n_v;
    a.support = support;
    a.os = oversampling;
    a.vis_dtype = vis_dtype;
    a.weight_dtype = weight_dtype;
    return ngc::standard_grid_impl(ctx, s, mode, precision, a, freq_chan, chan_map, pol_map,
                                   cgk_1d, cgk_len, accumulate, grid, sum_weight, delta_l,
                                   delta_m);
}
