// Standard (prolate-spheroidal) gridder for gfx950.
//
// Replaces _standard_grid_jit (ngcasa/imaging/_imaging_utils/_standard_grid.py:243-371)
// behind ngc_standard_grid().  Two implementations:
//
//  * plain FP64 drop-in (NGC_PREC_FP64, and density-mode calls with supports
//    above 1): one thread per (time, baseline, channel) sample walks the pols
//    and the support^2 taps in the reference's loop order and scatters with
//    FP64 atomics into the grid in HBM.  Density mode with the one-tap kernel
//    make_imaging_weight uses goes through std_density_runs_kernel instead:
//    the same values, one atomic per run of samples sharing a cell.
//
//  * uv-tiled (NGC_PREC_FP64_TILED, NGC_PREC_FP32_KAHAN, NGC_PREC_FP32): a
//    count pass and a fill pass bin every (row, channel) sample into the
//    (image channel, TU x 32 cell) tiles its footprint touches with an
//    order-preserving counting sort (wave-aggregated atomics keep runs of
//    consecutive integrations of one baseline together).  The fill pass
//    stores per sample the cell, the sub-cell offsets and the weighted
//    visibilities of all pols in the accumulation type, and adds the FP64
//    sum_weight term.  One workgroup of 1, 2 or 4 wavefronts (each taking a
//    contiguous share of the tile's records) then grids one tile into an LDS
//    subgrid (lane-owned residue classes, register runs; see the kernel) and
//    writes each grid cell to HBM once.  Pixel positions and sum_weight stay FP64; FP32_KAHAN accumulates
//    the FP32 register runs into FP64 LDS cells.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <type_traits>

#include "ngc_internal.hpp"

namespace ngc {

constexpr int kTile = 32;          // tile edge in cells (both axes)
constexpr int kTileLog2 = 5;
static_assert(kTile == 1 << kTileLog2, "kTileLog2");
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
// density grid for support 1 (make_imaging_weight grids the weights with a
// one-tap kernel, make_imaging_weight.py:154-161): one wavefront per 64
// consecutive integrations of one (baseline, channel), so that lanes whose
// cell (or conjugate cell) is the same form runs, and each run adds its sum
// with one FP64 atomic per pol instead of one per lane.  The values, the
// bounds, the conjugate rule and the sum_weight terms are those of
// std_grid_plain_kernel<NGC_GRID_DENSITY>; only the order of the FP64
// additions differs (atomics have none).
// ---------------------------------------------------------------------------
__device__ inline double run_sum(double v, int lane, int start) {
    // inclusive sum over lanes start .. lane
    for (int off = 1; off < 64; off <<= 1) {
        const double y = __shfl_up(v, off, 64);
        if (lane - off >= start) v += y;
    }
    return v;
}

__global__ void __launch_bounds__(kBlock)
std_density_runs_kernel(StdArgs a, const double *cgk, int64_t cgk_len, double *grid,
                        double *sum_weight) {
#pragma clang fp contract(off)
    const int64_t n_tb = (a.n_time + 63) / 64;
    const int64_t n_thr = a.n_baseline * a.n_chan * n_tb * 64;
    const int64_t plane = a.n_u * a.n_v;
    const int64_t cells = a.n_imag_chan * a.n_imag_pol * plane;
    // n_thr is a multiple of 64 and the stride of kBlock, so every wavefront
    // runs the loop a uniform number of times (the ballots need all lanes)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_thr;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int lane = (int)(i & 63);
        const int64_t w = i >> 6;
        const int64_t t = (w % n_tb) * 64 + lane;
        const int64_t r = w / n_tb;
        const int64_t ch = r % a.n_chan, b = r / a.n_chan;
        int64_t key = -1, ckey = -1;   // cell and conjugate cell within a plane
        double val = 0.0;              // conv * weight of this sample
        if (t < a.n_time) {
            const int64_t row = t * a.n_baseline + b;
            Pix px;
            if (pixel_position(a.uvw[row * 3], a.uvw[row * 3 + 1], a.uv_scale[ch],
                               a.uv_scale[a.n_chan + ch], a.n_u, a.n_v, a.os, px) &&
                px.uc < a.n_u && px.vc < a.n_v && px.uc >= 0 && px.vc >= 0) {
                const int64_t base = (row * a.n_chan + ch) * a.n_pol;
                const double wr = a.n_pol >= 2
                                      ? (weight_of(a, base, 0) + weight_of(a, base, 1)) / 2.0
                                      : weight_of(a, base, 0);
                if (!isnan(wr) && wr != 0.0) {
                    const int64_t ku = iabs64(px.du), kv = iabs64(px.dv);
                    const double conv_v = kv < cgk_len ? cgk[kv] : 0.0;
                    const double conv_u = ku < cgk_len ? cgk[ku] : 0.0;
                    val = (conv_u * conv_v) * wr;
                    key = px.uc * a.n_v + px.vc;
                    ckey = px.ucj * a.n_v + px.vcj;
                }
            }
        }
        const int64_t a_chan = a.chan_map[ch];   // one channel per wavefront
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
        for (int pass = 0; pass < 2; pass++) {   // the cell, then the conjugate cell
            const int64_t k = pass ? ckey : key;
            const int64_t prev = __shfl_up(k, 1, 64);
            const unsigned long long bounds = __ballot(lane == 0 || prev != k);
            const int start = 63 - __clzll(bounds & upto);
            const unsigned long long after = bounds & ~upto;
            const int next = after ? __ffsll(after) - 1 : 64;
            const double sum = run_sum(val, lane, start);
            if (key >= 0 && lane == next - 1) {
                for (int64_t p = 0; p < a.n_pol; p++) {
                    const int64_t pbase = (a_chan * a.n_imag_pol + a.pol_map[p]) * plane;
                    const int64_t cell = pbase + k;
                    // the reference does not bounds-check the conjugate cell:
                    // kept at its flat C-order offset when inside the grid
                    if (pass == 0 || (cell >= 0 && cell < cells)) atomicAdd(&grid[cell], sum);
                }
            }
        }
        // sum_weight: sel * norm = conv * weight per pol, added twice (once per
        // cell) as the reference does; one atomic pair per pol per wavefront
        double sw = val;
        for (int off = 32; off > 0; off >>= 1) sw += __shfl_xor(sw, off, 64);
        if (lane == 0 && sw != 0.0) {
            for (int64_t p = 0; p < a.n_pol; p++) {
                double *d = &sum_weight[a_chan * a.n_imag_pol + a.pol_map[p]];
                atomicAdd(d, sw);
                atomicAdd(d, sw);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// binning: deterministic counting sort of samples into (image chan, tile) bins
// ---------------------------------------------------------------------------
// A wavefront covers a block of (t_lanes integrations) x (ch_lanes channels)
// of one baseline: for each integration the ch_lanes channels are adjacent in
// memory, so the visibility and weight reads are coalesced, and the
// integrations of one channel land in the same tile, so the wave's entries
// reach each bin as a contiguous run in time order.
struct BinGeom {
    int64_t ntu, ntv;      // tiles per axis
    int tu;                // tile rows (u extent); the v extent is kTile
    int tu_log2;           // log2(tu)
    int ch_log2;           // log2(ch_lanes)
    int64_t n_tb, n_cb;    // time blocks, channel blocks
    FastDiv div_cb, div_tb;  // by n_cb and n_tb (n_baseline n_tb n_cb < 2^32)
};

struct Sample {
    int64_t t, b, ch, rc;  // rc = (t*n_baseline + b)*n_chan + ch
};

__device__ inline bool sample_at(const StdArgs &a, const BinGeom &g, int64_t wave_id, int lane,
                                 Sample &sm) {
    const int chl = 1 << g.ch_log2;
    const int tl = 64 >> g.ch_log2;
    // wave_id is the same in every lane: the divisions are scalar work
    const uint32_t wid = (uint32_t)wave_id;
    const uint32_t r = g.div_cb.div(wid);
    const uint32_t cb = wid - r * g.div_cb.d;
    const uint32_t b = g.div_tb.div(r);
    const uint32_t tb = r - b * g.div_tb.d;
    sm.t = (int64_t)tb * tl + (lane >> g.ch_log2);
    sm.ch = cb * chl + (lane & (chl - 1));
    sm.b = b;
    if (b >= a.n_baseline || sm.t >= a.n_time || sm.ch >= a.n_chan) return false;
    sm.rc = (sm.t * a.n_baseline + b) * a.n_chan + sm.ch;
    return true;
}

// centre cell, offsets and on-grid test of a sample at (u, v) metres, with
// `any_unflagged` from its flags; false when it is not gridded
// (s0, s1 = the channel's uv scales, loaded with the sample's data)
__device__ inline bool place(const StdArgs &a, double u, double v, double s0, double s1,
                             bool any_unflagged, Pix &px) {
    if (!any_unflagged) return false;
    if (!pixel_position(u, v, s0, s1, a.n_u, a.n_v, a.os, px)) return false;
    const int64_t sc = a.support / 2;
    return (px.uc + sc < a.n_u) && (px.vc + sc < a.n_v) && (px.uc - sc >= 0) &&
           (px.vc - sc >= 0);
}

// k-th overlapped tile (k = 0..3) of a placed sample, or -1
__device__ inline int64_t tile_key(const StdArgs &a, const BinGeom &g, const Pix &px,
                                   int64_t a_chan, int k) {
    // placed samples have 0 <= uc - sc and uc + sc < n_u < 65536 (same for v)
    const int sc = (int)(a.support / 2);
    const int uc = (int)px.uc, vc = (int)px.vc;
    const int tu0 = (uc - sc) >> g.tu_log2, tu1 = (uc + sc) >> g.tu_log2;
    const int tv0 = (vc - sc) >> kTileLog2, tv1 = (vc + sc) >> kTileLog2;
    const int tu = tu0 + (k & 1), tv = tv0 + (k >> 1);
    if (tu > tu1 || tv > tv1) return -1;
    return (a_chan * g.ntu + tu) * g.ntv + tv;
}

// Entry record of one sample in one bin: packed centre cell and offsets, then
// the weighted visibilities of every pol (ACC precision), padded to 8 B (the
// header is read as one 8-byte word).  The fill pass's entry stores are the
// largest part of its time, so records carry no further padding: 40 B for
// 4 pols and 24 B for 2 pols of complex FP32.
template <int NP, int NC, typename ACC> struct Entry {
    int uv;      // u0 | v0 << 16
    int duv;     // du & 0xffff | dv << 16
    ACC wd[NP][NC];
};
template <int NP, int NC, typename ACC> struct EntrySize {
    static constexpr int raw = (int)sizeof(Entry<NP, NC, ACC>);
    static constexpr int bytes = (raw + 7) / 8 * 8;
};

// The fast load path of the binning passes (FAST = true): uint8 flags, FP32
// weights and, in image mode, complex64 visibilities, each array aligned to
// one sample's NP elements (NP = 1, 2 or 4), so that a sample's flags,
// weights and visibilities are one vector load each, issued together.  The
// general path branches on the dtypes and alignment, and the compiler then
// waits for each load before issuing the next.
template <int NP> struct FastSample {
    static_assert(NP == 1 || NP == 2 || NP == 4, "fast loads for 1, 2 or 4 pols");
    using FT = typename std::conditional<
        NP == 4, uint32_t, typename std::conditional<NP == 2, uint16_t, uint8_t>::type>::type;
    FT f;
    float w[NP];
    float2 v[NP];
    template <bool FILLV, bool IMAGE> __device__ void load(const StdArgs &a, int64_t base) {
        f = *reinterpret_cast<const FT *>(a.flag + base);
        if constexpr (FILLV) {
            const float *wp = reinterpret_cast<const float *>(a.weight) + base;
            if constexpr (NP == 4) {
                const float4 t = *reinterpret_cast<const float4 *>(wp);
                w[0] = t.x; w[NP > 1 ? 1 : 0] = t.y; w[NP > 2 ? 2 : 0] = t.z; w[NP - 1] = t.w;
            } else if constexpr (NP == 2) {
                const float2 t = *reinterpret_cast<const float2 *>(wp);
                w[0] = t.x; w[NP - 1] = t.y;
            } else {
                w[0] = *wp;
            }
            if constexpr (IMAGE) {
                const float2 *vp = reinterpret_cast<const float2 *>(a.vis) + base;
                if constexpr (NP == 1) {
                    v[0] = *vp;
                } else {
#pragma unroll
                    for (int h = 0; h < NP / 2; h++) {
                        const float4 t = reinterpret_cast<const float4 *>(vp)[h];
                        v[2 * h] = make_float2(t.x, t.y);
                        v[2 * h + 1] = make_float2(t.z, t.w);
                    }
                }
            }
        }
    }
    __device__ void unpack(bool (&fl)[NP], double (&wts)[NP], double2 (&vis)[NP]) const {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            fl[p] = ((f >> (8 * p)) & 0xffu) != 0;
            wts[p] = w[p];
            vis[p] = make_double2(v[p].x, v[p].y);
        }
    }
};

// whether the fast load path applies to a call's arrays
static bool bin_fast_loads(const StdArgs &a, int mode, int64_t np) {
    if (!(np == 1 || np == 2 || np == 4) || !a.flag) return false;
    if ((uintptr_t)a.flag % np != 0) return false;
    if (a.weight_dtype != NGC_F32 || (uintptr_t)a.weight % (4 * np) != 0) return false;
    if (mode == NGC_GRID_IMAGE &&
        (a.vis_dtype != NGC_C64 || (uintptr_t)a.vis % (np == 1 ? 8 : 16) != 0))
        return false;
    return true;
}

// Count pass (FILL = false) and fill pass (FILL = true).  The fill pass also
// accumulates sum_weight in FP64: sel_weight * norm(du, dv) per valid pol,
// the reference's per-sample term (_standard_grid.py:366).
template <bool FILL, int MODE, int NP, typename ACC, bool FAST>
__global__ void __launch_bounds__(kBlock)
std_bin_kernel(StdArgs a, BinGeom g, unsigned long long *counts, char *entries,
               unsigned long long cap, int *overflow, const double *norm_tab, int norm_half,
               double *sum_weight, int sw_in_lds) {
    constexpr int NC = (MODE == NGC_GRID_IMAGE) ? 2 : 1;
    using E = Entry<NP, NC, ACC>;
    constexpr int EB = EntrySize<NP, NC, ACC>::bytes;
    extern __shared__ double s_sw[];
    const int64_t n_sw = a.n_imag_chan * a.n_imag_pol;
    if (FILL && sw_in_lds) {
        for (int64_t i = threadIdx.x; i < n_sw; i += blockDim.x) s_sw[i] = 0.0;
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t n_waves = a.n_baseline * g.n_tb * g.n_cb;
    const int waves_per_block = blockDim.x >> 6;
    const int nw = 2 * norm_half + 1;
    for (int64_t w0 = (int64_t)blockIdx.x * waves_per_block; w0 < n_waves;
         w0 += (int64_t)gridDim.x * waves_per_block) {
        const int64_t wave_id = w0 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        Sample sm{};
        Pix px{};
        const bool have = (wave_id < n_waves) && sample_at(a, g, wave_id, lane, sm);
        // all of the sample's loads are issued before the placement arithmetic
        const int64_t base = sm.rc * NP;
        double uu = 0.0, vv = 0.0, s0 = 0.0, s1 = 0.0;
        int64_t a_chan_ld = 0;
        bool fl[NP];
        double wts[NP];
        double2 vis[NP];
        bool any_unflagged = false;
        if (have) {
            const int64_t row = sm.t * a.n_baseline + sm.b;
            uu = a.uvw[row * 3];
            vv = a.uvw[row * 3 + 1];
            s0 = a.uv_scale[sm.ch];
            s1 = a.uv_scale[a.n_chan + sm.ch];
            a_chan_ld = a.chan_map[sm.ch];
            if constexpr (FAST) {
                // one vector load per array, all issued before any is used
                FastSample<NP> fs;
                fs.template load<FILL, MODE == NGC_GRID_IMAGE>(a, base);
                fs.unpack(fl, wts, vis);
#pragma unroll
                for (int p = 0; p < NP; p++) any_unflagged |= !fl[p];
            } else {
                // 4 pols: one word of flags, one 16-byte load of FP32 weights and two
                // of complex64 visibilities when the arrays are aligned for it
                if (NP == 4 && a.flag && ((uintptr_t)a.flag & 3) == 0) {
                    const uint32_t f4 = *reinterpret_cast<const uint32_t *>(a.flag + base);
#pragma unroll
                    for (int p = 0; p < NP; p++) fl[p] = ((f4 >> (8 * p)) & 0xffu) != 0;
                } else {
#pragma unroll
                    for (int p = 0; p < NP; p++) fl[p] = a.flag ? a.flag[base + p] != 0 : false;
                }
#pragma unroll
                for (int p = 0; p < NP; p++) any_unflagged |= !fl[p];
                if (FILL) {
                    if (NP == 4 && a.weight_dtype == NGC_F32 && ((uintptr_t)a.weight & 15) == 0) {
                        const float4 w4 = *reinterpret_cast<const float4 *>(
                            reinterpret_cast<const float *>(a.weight) + base);
                        wts[0] = w4.x;
                        wts[NP > 1 ? 1 : 0] = w4.y;
                        wts[NP > 2 ? 2 : 0] = w4.z;
                        wts[NP > 3 ? 3 : 0] = w4.w;
                    } else {
#pragma unroll
                        for (int p = 0; p < NP; p++) wts[p] = load_w(a.weight, base + p, a.weight_dtype);
                    }
                    if (MODE == NGC_GRID_IMAGE) {
                        if (NP == 4 && a.vis_dtype == NGC_C64 && ((uintptr_t)a.vis & 15) == 0) {
                            const float4 *vp = reinterpret_cast<const float4 *>(
                                reinterpret_cast<const float2 *>(a.vis) + base);
                            const float4 v01 = vp[0], v23 = vp[1];
                            vis[0] = make_double2(v01.x, v01.y);
                            vis[NP > 1 ? 1 : 0] = make_double2(v01.z, v01.w);
                            vis[NP > 2 ? 2 : 0] = make_double2(v23.x, v23.y);
                            vis[NP > 3 ? 3 : 0] = make_double2(v23.z, v23.w);
                        } else {
#pragma unroll
                            for (int p = 0; p < NP; p++) vis[p] = load_vis(a.vis, base + p, a.vis_dtype);
                        }
                    }
                }
            }
        }
        const bool ok = have && place(a, uu, vv, s0, s1, any_unflagged, px);
        const int64_t a_chan = ok ? a_chan_ld : 0;
        E rec;
        double norm = 0.0;
        bool use[NP];
        double wvs[NP];
        if (FILL && ok) {
            rec.uv = (int)((px.uc & 0xffff) | (px.vc << 16));
            rec.duv = (int)((px.du & 0xffff) | ((px.dv & 0xffff) << 16));
            const int ndu = min(max((int)px.du + norm_half, 0), nw - 1);
            const int ndv = min(max((int)px.dv + norm_half, 0), nw - 1);
            // (used after the counter atomics are issued, which hides its latency)
            norm = norm_tab[ndu * nw + ndv];
#pragma unroll
            for (int p = 0; p < NP; p++) {
                // the weight is NaN when flagged
                const double wv = fl[p] ? __longlong_as_double(0x7ff8000000000000LL) : wts[p];
                double wr = wv, wi = 0.0;
                if (MODE == NGC_GRID_IMAGE) {
                    wr = vis[p].x * wv;
                    wi = vis[p].y * wv;
                }
                use[p] = !(isnan(wr) || isnan(wi)) && !(wr == 0.0 && wi == 0.0);
                wvs[p] = wv;
                rec.wd[p][0] = use[p] ? (ACC)wr : (ACC)0;
                if (NC == 2) rec.wd[p][NC - 1] = use[p] ? (ACC)wi : (ACC)0;
            }
        }
        // group the lanes by tile for each of the (up to 4) tiles a sample
        // overlaps: the group's lowest lane leads, every lane knows its rank
        int64_t key[4];
        int leader[4], rank[4], cnt[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            key[k] = ok ? tile_key(a, g, px, a_chan, k) : -1;
            leader[k] = lane;
            rank[k] = 0;
            cnt[k] = 0;
            uint64_t pending = __ballot(key[k] >= 0);
            while (pending) {
                const int l = __ffsll((long long)pending) - 1;
                const int64_t lkey = __shfl(key[k], l, 64);
                const uint64_t same = __ballot(key[k] == lkey);
                if (key[k] == lkey) {
                    leader[k] = l;
                    rank[k] = __popcll(same & below);
                    cnt[k] = __popcll(same);
                }
                pending &= ~same;
            }
        }
        // one counter atomic per group, all issued before any result is used
        unsigned long long basev[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            basev[k] = 0;
            if (key[k] >= 0 && rank[k] == 0)
                basev[k] = atomicAdd(&counts[key[k]], (unsigned long long)cnt[k]);
        }
        if (FILL && ok) {
            // the FP64 sum_weight terms, while the counter atomics are in flight
#pragma unroll
            for (int p = 0; p < NP; p++) {
                if (use[p]) {
                    const int64_t swi = a_chan * a.n_imag_pol + p;
                    if (sw_in_lds)
                        atomicAdd(&s_sw[swi], wvs[p] * norm);
                    else
                        atomicAdd(&sum_weight[swi], wvs[p] * norm);
                }
            }
        }
        if (FILL) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const unsigned long long b0 = __shfl(basev[k], leader[k], 64);
                if (key[k] >= 0) {
                    const unsigned long long pos = b0 + (unsigned long long)rank[k];
                    if (pos < cap)
                        *reinterpret_cast<E *>(entries + pos * EB) = rec;
                    else
                        atomicOr(overflow, 1);
                }
            }
        }
    }
    if (FILL && sw_in_lds) {
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n_sw; i += blockDim.x)
            if (s_sw[i] != 0.0) atomicAdd(&sum_weight[i], s_sw[i]);
    }
}

// ---------------------------------------------------------------------------
// tiled gridding kernel: one workgroup per (image channel, TU x 32 tile) bin
// ---------------------------------------------------------------------------
// Work distribution after J. W. Romein, "An efficient work-distribution
// strategy for gridding radio-telescope data on GPUs" (ICS 2012), applied to
// an LDS subgrid: lane (a, b) of a wavefront owns every tile cell with
// u = a (mod 8) and v = b (mod 8).  A support <= 8 window covers each residue
// pair at most once, so for every sample each lane has at most one tap, and
// no two lanes of a wavefront ever touch the same cell.  A lane keeps the
// running sum of its current cell in registers and adds it to the LDS
// subgrid only when its cell changes; consecutive integrations of one
// baseline move slowly across the grid, so most taps never leave the
// register file.  The add is a no-return LDS atomic (ds_add_f32/f64): no two
// lanes share a cell, but the read-modify-write then happens inside the LDS
// unit and the wavefront does not wait for it.  FP32_KAHAN keeps FP64 cells
// (one conversion and one add per flush); the tile is written to HBM once at
// the end.
//
// The W wavefronts of a workgroup split the tile's records: wavefront w
// grids the w-th contiguous share of the bin's records into every pol, with
// all of the tile's accumulators in its own registers, and stages, reads and
// runs its share on its own, so no barrier stands between sample batches
// (the staging arrays are per wavefront; the LDS executes one wavefront's
// instructions in order, and a wavefront-scope fence keeps the compiler from
// moving its reads above its stores).  Cells that two shares both reach are
// added by both wavefronts: the flush is an LDS atomic add.  The subgrid is
// one per workgroup, so W raises the wavefronts per CU at the LDS cost of the
// staging arrays only.  (Splitting the pols instead, every wavefront running
// every record for its share of the accumulators, repeated the per-sample
// product and staging and held the wavefronts together at two barriers per
// batch: 15.0 ms against 10.9 ms for 4 record shares at 627M samples into
// 256 planes of 2457^2.)
//
// Per-sample work is done once per sample and axis by the lanes that stage
// it, in per-residue layout: for residue r the tap weight of the lanes owning
// u = r (mod 8), with the sign bit set when that residue's cell differs from
// the previous sample's (the weights are non-negative), and the cell; the
// same for v.  In the broadcast loop a lane then reads its two records at a
// fixed offset, ORs the sign bits to decide a flush, and multiplies the
// magnitudes (abs is an operand modifier).  Samples are read from LDS in
// double-buffered groups, so a group's reads do not queue behind the
// previous group's adds; the next batch's entries are loaded from HBM while
// the current batch is gridded.  TU (8, 16 or 32 rows) trades LDS per
// workgroup, hence occupancy, against the share of samples binned to two
// tiles.
constexpr int kBuckets = 40;       // bin-size classes of the largest-first order

// size class of a bin of n records: 0 when empty, else floor(log2 n) + 1
__device__ inline int bin_bucket(unsigned long long n) {
    if (n == 0) return 0;
    const int b = 64 - __clzll((long long)n);
    return b < kBuckets ? b : kBuckets - 1;
}

constexpr int kChunk = 16;        // tiles per work ticket of small bins
constexpr int kEmptyChunk = 8;    // empty tiles zeroed per ticket of the empty-tile queue
constexpr int kChunkBelow = 32;   // records below which a bin is small

// LDS subgrid layout: row-major tile planes of 32 cells per row, with the
// columns of row u XOR-swizzled: cell (u, v) at u 32 + (v ^ X (u mod 8)),
// X = 2 for 8-byte and 4 for 4-byte cells.  The lanes of a wavefront own the
// cells of one 8 x 8 window (u = a, v = b mod 8); in this layout the lanes
// that the LDS serves together fall on distinct banks for every window
// position, and so do consecutive cells of one row as the write-out reads
// them, without padding (which would cost occupancy).
template <int CELL_BYTES>
__device__ inline int subgrid_index(int u, int v) {
    constexpr int X = CELL_BYTES == 8 ? 2 : 4;
    return u * kTile + (v ^ (X * (u & 7)));
}

// cell of residue r for a window starting at w: the first u >= w with u = r (mod 8)
__device__ inline int residue_cell(int w, int r) { return w + ((r - w) & 7); }

// Samples per staging batch of one wavefront: one lane per (sample, axis)
// stages all 8 taps.  Measured at 627M samples into 256 planes of 2457^2
// (4 wavefronts, 16-row tiles): batches of 8 samples 11.0 ms (4 workgroups
// per CU), 16 samples 10.4 ms and 32 samples 10.0 ms (3 workgroups per CU by
// LDS, but half or a quarter of the per-batch staging work per sample).
constexpr int kSplitBatch = 32;
constexpr int kSplitGroup = 4;   // samples per double-buffered read group

template <int MODE, int NP, typename ACC, bool KAHAN, int TU, int W>
__global__ void __launch_bounds__(64 * W)
    __attribute__((amdgpu_waves_per_eu((W >= 4 && sizeof(ACC) == 4) ? 3 : 2)))
std_grid_tiled_kernel(StdArgs a, BinGeom g, const unsigned long long *__restrict__ offsets,
                      const unsigned int *__restrict__ order,
                      const unsigned int *__restrict__ hist, int *queue, int n_items,
                      unsigned long long cap, const char *__restrict__ entries,
                      const double *__restrict__ cgk, int cgk_len, void *grid_out,
                      int accumulate) {
    constexpr int NC = (MODE == NGC_GRID_IMAGE) ? 2 : 1;
    constexpr int NPC = NP * NC;
    using CELL = typename std::conditional<KAHAN, double, ACC>::type;
    constexpr int PLANE = TU * kTile;
    constexpr int SB = kSplitBatch;
    constexpr int KG = kSplitGroup;
    constexpr int NT = 64 * W;
    using E = Entry<NP, NC, ACC>;
    constexpr int EB = EntrySize<NP, NC, ACC>::bytes;
    struct WD {
        ACC v[NPC];
    };
    __shared__ CELL s_acc[NPC * PLANE];
    constexpr int AXROW = 9;
    __shared__ ACC s_ax[W][2][SB][AXROW];
    // per staged sample its window start, u in the low and v in the high 16 bits
    __shared__ int s_c0[W][SB];
    __shared__ WD s_wd[W][SB];
    __shared__ ACC s_k[513];
    __shared__ int s_item;
    __shared__ int s_nz, s_e;
    __shared__ int64_t s_tq[kChunk][3];
    // the 64 KB subgrid of 4 complex pols in 32-row FP64 cells (the default
    // shape of 4-pol images in the single-precision forms) leaves just room
    // for two workgroups per CU
    static_assert(!(NPC * PLANE * sizeof(CELL) == 65536 && W == 4 && sizeof(ACC) == 4) ||
                      sizeof(s_acc) + sizeof(s_ax) + sizeof(s_c0) + sizeof(s_wd) + sizeof(s_k) +
                              sizeof(s_tq) + 3 * sizeof(int) <= 81920,
                  "two workgroups of a 64 KB subgrid fit the 160 KB of a CU");

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    {
        constexpr int KL = (512 + NT - 1) / NT;
        ACC kv[KL];
#pragma unroll
        for (int k = 0; k < KL; k++) {
            const int i = tid + k * NT;
            kv[k] = i < cgk_len && i < 512 ? (ACC)cgk[i] : (ACC)0;
        }
#pragma unroll
        for (int k = 0; k < KL; k++)
            if (tid + k * NT < 512) s_k[tid + k * NT] = kv[k];
        if (tid == 0) s_k[512] = (ACC)0;
    }
#pragma unroll 4
    for (int i = tid; i < NPC * PLANE; i += NT) s_acc[i] = CELL{};
    if (tid == 0) {
        int nb = 0;
        for (int b = bin_bucket(kChunkBelow); b < kBuckets; b++) nb += (int)hist[b];
        s_item = nb;
        s_nz = n_items - (int)hist[0];
    }
    __syncthreads();
    const int n_single = s_item;
    const int n_nz = s_nz;
    using OUT = typename std::conditional<NC == 2,
        typename std::conditional<std::is_same<ACC, float>::value, float2, double2>::type,
        ACC>::type;
    constexpr int F = NP * TU * kTile;
    bool zero_more = !accumulate;
    auto zero_chunk = [&]() -> bool {
        __syncthreads();
        if (tid == 0) s_e = atomicAdd(queue + 2, 1);
        __syncthreads();
        const int64_t first = (int64_t)n_nz + (int64_t)s_e * kEmptyChunk;
        if (first >= n_items) return false;
        const int64_t last = min(first + kEmptyChunk, (int64_t)n_items);
        unsigned eb[kEmptyChunk];
#pragma unroll
        for (int c = 0; c < kEmptyChunk; c++) eb[c] = order[min(first + c, last - 1)];
        const int64_t tiles = g.ntu * g.ntv;
#pragma unroll
        for (int c = 0; c < kEmptyChunk; c++) {
            if (first + c >= last) break;
            const int64_t ach = eb[c] / tiles;
            const int64_t trem = eb[c] - ach * tiles;
            const int64_t u0 = (trem / g.ntv) * TU, v0 = (trem % g.ntv) * kTile;
#pragma unroll 4
            for (int k = tid; k < F; k += NT) {
                const int64_t gu = u0 + (k / kTile) % TU, gv = v0 + k % kTile;
                const int p = k / (kTile * TU);
                if (gu < a.n_u && gv < a.n_v)
                    reinterpret_cast<OUT *>(grid_out)[((ach * a.n_imag_pol + p) * a.n_u + gu) *
                                                          a.n_v + gv] = OUT{};
            }
        }
        return true;
    };
    // no compiler motion of LDS accesses across this point; in hardware one
    // wavefront's LDS instructions complete in order
    auto wave_sync = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    int it_next = 0, it_end = 0, it_first = 0;
    for (;;) {
    __syncthreads();
    if (it_next >= it_end) {
        if (tid == 0) s_item = atomicAdd(queue, 1);
        __syncthreads();
        const int t = s_item;
        if (t < n_single) {
            it_next = t;
            it_end = t + 1;
        } else {
            const int64_t first = (int64_t)n_single + (int64_t)(t - n_single) * kChunk;
            it_next = first < n_nz ? (int)first : n_nz;
            it_end = it_next + kChunk;
        }
        if (it_end > n_nz) it_end = n_nz;
        if (it_next >= n_nz) break;
        it_first = it_next;
        if (tid < it_end - it_next) {
            const int64_t b = order[it_next + tid];
            s_tq[tid][0] = b;
            s_tq[tid][1] = (int64_t)min(offsets[b], cap);
            s_tq[tid][2] = (int64_t)min(offsets[b + 1], cap);
        }
        __syncthreads();
    }
    const int ci = it_next - it_first;
    it_next++;
    const int64_t bin = s_tq[ci][0];
    const int64_t tiles = g.ntu * g.ntv;
    const int64_t a_chan = bin / tiles;
    const int64_t trem = bin - a_chan * tiles;
    const int u_lo = (int)((trem / g.ntv) * TU);
    const int v_lo = (int)((trem % g.ntv) * kTile);
    // this wavefront's share of the bin's records
    const int64_t b_start = s_tq[ci][1], b_n = s_tq[ci][2] - b_start;
    const int64_t start = b_start + b_n * wave / W;
    const int64_t end = b_start + b_n * (wave + 1) / W;

    const int sc = (int)(a.support / 2);
    const int S = (int)a.support;
    const int os = (int)a.os;
    const int ra = lane & 7, rb = lane >> 3;

    int cur_u = -64, cur_v = -64;
    ACC acc[NPC];
#pragma unroll
    for (int q = 0; q < NPC; q++) acc[q] = (ACC)0;

    auto flush = [&]() {
        if ((unsigned)cur_u < (unsigned)TU && (unsigned)cur_v < (unsigned)kTile) {
            const int o = subgrid_index<(int)sizeof(CELL)>(cur_u, cur_v);
#pragma unroll
            for (int q = 0; q < NPC; q++)
                __hip_atomic_fetch_add(&s_acc[q * PLANE + o], (CELL)acc[q], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    };

    // staging: lane l stages taps h RPL .. h RPL + RPL - 1 of axis
    // (l / SB) % 2 of sample l % SB of the batch, h = l / (2 SB)
    constexpr int H = 64 / (2 * SB);
    constexpr int RPL = 8 / H;
    static_assert(H * RPL == 8, "the taps of an axis split evenly over its lanes");
    const int js = lane % SB, axis = (lane / SB) & 1, h = lane / (2 * SB);
    const bool wd_lane = axis == 0 && h == 0;
    const int shift = axis ? 16 : 0;
    const int lo = (axis ? v_lo : u_lo) + sc;
    const int it0 = h * RPL;
    // table offset of each of the lane's taps before the sub-cell shift; taps
    // past the support get one that clamps to the zero entry s_k[512]
    int toff[RPL];
#pragma unroll
    for (int i = 0; i < RPL; i++) toff[i] = it0 + i < S ? os * (it0 + i - sc) : (1 << 20);
    int2 hdr = make_int2(0, 0);
    unsigned int puv = 0;
    WD wdl;
    auto load_batch = [&](int64_t b) {
        const int64_t e = b + js;
        if (e < end) {
            const char *p = entries + e * EB;
            hdr = *reinterpret_cast<const int2 *>(p);
            puv = e > start ? *reinterpret_cast<const unsigned int *>(p - EB) : 0u;
            if (wd_lane) {
                const E *ep = reinterpret_cast<const E *>(p);
#pragma unroll
                for (int q = 0; q < NPC; q++) wdl.v[q] = ep->wd[q / NC][q % NC];
            }
        }
    };
    load_batch(start);
    for (int64_t b0 = start; b0 < end; b0 += SB) {
        const int64_t e = b0 + js;
        const bool valid = e < end;
        unsigned chg;
        {
            int c0 = -1000, dlt = 0, d = 0;
            if (valid) {
                c0 = (int)(((unsigned)hdr.x >> shift) & 0xffffu) - lo;
                d = (int)(short)(((unsigned)hdr.y >> shift) & 0xffffu);
                dlt = e > start ? c0 - ((int)((puv >> shift) & 0xffffu) - lo) : 1000;
            }
            ACC tk[RPL];
#pragma unroll
            for (int i = 0; i < RPL; i++) {
                const int x = toff[i] + d;
                tk[i] = s_k[min(x < 0 ? -x : x, 512)];
            }
#pragma unroll
            for (int i = 0; i < RPL; i++) {
                const int it = it0 + i;
                const ACC t = tk[i];
                const bool moved = (unsigned)(dlt + it) >= 8u;
                s_ax[wave][axis][js][(c0 + it) & 7] = moved ? -t - (ACC)0 : t;
            }
            if (h == 0) reinterpret_cast<short *>(&s_c0[wave][js])[axis] = (short)c0;
            if (wd_lane) {
                if (end - b0 < SB) {   // the share's last batch: records past its end add nothing
#pragma unroll
                    for (int q = 0; q < NPC; q++) wdl.v[q] = valid ? wdl.v[q] : (ACC)0;
                }
                s_wd[wave][js] = wdl;
            }
            // wave-uniform: samples of the batch whose window moved on either
            // axis, from the first-tap lanes (h = 0: lanes 0 .. 2 SB - 1)
            const uint64_t mv = __ballot(dlt != 0) & (~0ull >> (64 - 2 * SB));
            chg = (unsigned)((mv | (mv >> SB)) & ((1ull << SB) - 1));
        }
        const int64_t nleft = end - b0;
        const int nh = nleft >= SB ? SB : (int)nleft;
        wave_sync();

        const int ng = (nh + KG - 1) / KG;
        ACC au0[KG], av0[KG], au1[KG], av1[KG];
        WD wd0[KG], wd1[KG];
        int c00, c01;
        auto load_group = [&](int gi, ACC (&au)[KG], ACC (&av)[KG], WD (&wd)[KG], int &c0l) {
#pragma unroll
            for (int jj = 0; jj < KG; jj++) {
                au[jj] = s_ax[wave][0][gi * KG + jj][ra];
                av[jj] = s_ax[wave][1][gi * KG + jj][rb];
                wd[jj] = s_wd[wave][gi * KG + jj];
            }
            c0l = s_c0[wave][gi * KG + (lane & (KG - 1))];
        };
        auto run_group = [&](int gi, const ACC (&au)[KG], const ACC (&av)[KG],
                             const WD (&wd)[KG], const int &c0l) {
#pragma unroll
            for (int jj = 0; jj < KG; jj++) {
                if ((chg >> (gi * KG + jj)) & 1u) {
                    const int w2 = __builtin_amdgcn_readlane(c0l, jj);
                    const int wu = (int)(short)(w2 & 0xffff), wv = w2 >> 16;
                    if (signbit(au[jj]) || signbit(av[jj])) {
                        flush();
                        cur_u = residue_cell(wu, ra);
                        cur_v = residue_cell(wv, rb);
#pragma unroll
                        for (int q = 0; q < NPC; q++) acc[q] = (ACC)0;
                    }
                }
                const ACC wt = fabs(au[jj]) * fabs(av[jj]);
#pragma unroll
                for (int q = 0; q < NPC; q++) acc[q] += wt * wd[jj].v[q];
            }
        };
        constexpr int kLastGroup = SB / KG - 1;
        load_group(0, au0, av0, wd0, c00);
        if (b0 + SB < end) load_batch(b0 + SB);
        for (int gi = 0; gi < ng; gi += 2) {
            load_group(min(gi + 1, kLastGroup), au1, av1, wd1, c01);
            run_group(gi, au0, av0, wd0, c00);
            if (gi + 1 >= ng) break;
            load_group(min(gi + 2, kLastGroup), au0, av0, wd0, c00);
            run_group(gi + 1, au1, av1, wd1, c01);
        }
        wave_sync();
    }
    flush();
    __syncthreads();

#pragma unroll 4
    for (int k = tid; k < F; k += NT) {
        const int lv = k % kTile;
        const int lu = (k / kTile) % TU;
        const int p = k / (kTile * TU);
        const int o = subgrid_index<(int)sizeof(CELL)>(lu, lv);
        OUT val;
        if constexpr (NC == 2) {
            CELL *re = &s_acc[(2 * p) * PLANE + o];
            CELL *im = &s_acc[(2 * p + 1) * PLANE + o];
            val.x = (ACC)*re;
            val.y = (ACC)*im;
            *re = CELL{};
            *im = CELL{};
        } else {
            CELL *c = &s_acc[p * PLANE + o];
            val = (ACC)*c;
            *c = CELL{};
        }
        const int64_t gu = u_lo + lu, gv = v_lo + lv;
        if (gu >= a.n_u || gv >= a.n_v) continue;
        OUT *gp = reinterpret_cast<OUT *>(grid_out) +
                  ((a_chan * a.n_imag_pol + p) * a.n_u + gu) * a.n_v + gv;
        if (accumulate) {
            OUT g0 = *gp;
            if constexpr (NC == 2) {
                g0.x += val.x;
                g0.y += val.y;
            } else {
                g0 += val;
            }
            *gp = g0;
        } else {
            *gp = val;
        }
    }
    if (zero_more) zero_more = zero_chunk();
    }  // work queue
    while (zero_more) zero_more = zero_chunk();
}

// ---------------------------------------------------------------------------
// graph-safe helpers: fills, a three-pass exclusive scan, and the
// largest-first bin order (bins grouped by floor(log2(count)), descending)
// ---------------------------------------------------------------------------
constexpr int kScanChunk = 1024;  // elements per scan block (256 threads x 4)

template <typename T> __global__ void fill_kernel(T *p, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

template <typename T> __global__ void copy_kernel(T *dst, const T *src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__device__ inline unsigned long long block_exclusive_scan_256(unsigned long long x,
                                                              unsigned long long *total) {
    __shared__ unsigned long long s_w[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned long long incl = x;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) s_w[w] = incl;
    __syncthreads();
    unsigned long long before = 0, all = 0;
    for (int i = 0; i < 4; i++) {
        if (i < w) before += s_w[i];
        all += s_w[i];
    }
    __syncthreads();
    *total = all;
    return before + incl - x;
}

__global__ void __launch_bounds__(256)
scan_reduce_kernel(const unsigned long long *in, int64_t n, unsigned long long *block_sums) {
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + threadIdx.x * 4;
    unsigned long long x = 0;
    for (int k = 0; k < 4; k++)
        if (base + k < n) x += in[base + k];
    unsigned long long total;
    block_exclusive_scan_256(x, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the block sums, in chunks of 1024
__global__ void __launch_bounds__(256)
scan_blocks_kernel(unsigned long long *block_sums, int64_t nb) {
    unsigned long long carry = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += kScanChunk) {
        const int64_t base = c0 + threadIdx.x * 4;
        unsigned long long v[4], x = 0;
        for (int k = 0; k < 4; k++) {
            v[k] = base + k < nb ? block_sums[base + k] : 0ull;
            x += v[k];
        }
        unsigned long long total;
        unsigned long long e = block_exclusive_scan_256(x, &total) + carry;
        for (int k = 0; k < 4; k++) {
            if (base + k < nb) block_sums[base + k] = e;
            e += v[k];
        }
        carry += total;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256)
scan_apply_kernel(const unsigned long long *in, int64_t n, const unsigned long long *block_off,
                  unsigned long long *out) {
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + threadIdx.x * 4;
    unsigned long long v[4], x = 0;
    for (int k = 0; k < 4; k++) {
        v[k] = base + k < n ? in[base + k] : 0ull;
        x += v[k];
    }
    unsigned long long total;
    unsigned long long e = block_exclusive_scan_256(x, &total) + block_off[blockIdx.x];
    for (int k = 0; k < 4; k++) {
        if (base + k < n) out[base + k] = e;
        e += v[k];
    }
}

// hist[b] = number of bins in bucket b (hist zeroed beforehand)
__global__ void __launch_bounds__(256)
order_hist_kernel(const unsigned long long *offsets, int64_t n_bins, unsigned int *hist) {
    __shared__ unsigned int s_h[kBuckets];
    for (int i = threadIdx.x; i < kBuckets; i += blockDim.x) s_h[i] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_bins;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&s_h[bin_bucket(offsets[i + 1] - offsets[i])], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kBuckets; i += blockDim.x)
        if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// cursor[b] = first position of bucket b, buckets in decreasing order
__global__ void order_prefix_kernel(const unsigned int *hist, unsigned int *cursor) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned int pos = 0;
        for (int b = kBuckets - 1; b >= 0; b--) {
            cursor[b] = pos;
            pos += hist[b];
        }
    }
}

__global__ void __launch_bounds__(256)
order_scatter_kernel(const unsigned long long *offsets, int64_t n_bins, unsigned int *cursor,
                     unsigned int *order) {
    __shared__ unsigned int s_h[kBuckets];
    __shared__ unsigned int s_base[kBuckets];
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < n_bins;
         i0 += (int64_t)gridDim.x * blockDim.x) {
        for (int i = threadIdx.x; i < kBuckets; i += blockDim.x) s_h[i] = 0;
        __syncthreads();
        const int64_t i = i0 + threadIdx.x;
        int b = 0;
        unsigned int rank = 0;
        if (i < n_bins) {
            b = bin_bucket(offsets[i + 1] - offsets[i]);
            rank = atomicAdd(&s_h[b], 1u);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < kBuckets; k += blockDim.x)
            s_base[k] = s_h[k] ? atomicAdd(&cursor[k], s_h[k]) : 0u;
        __syncthreads();
        if (i < n_bins) order[s_base[b] + rank] = (unsigned int)i;
        __syncthreads();
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

// the support-1 density grid
static int density_runs(hipStream_t s, const StdArgs &a, const double *cgk, int64_t cgk_len,
                        double *g, double *sum_weight) {
    const int64_t n_thr = a.n_baseline * a.n_chan * ceil_div(a.n_time, (int64_t)64) * 64;
    std_density_runs_kernel<<<launch_blocks(n_thr), kBlock, 0, s>>>(a, cgk, cgk_len, g,
                                                                     sum_weight);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
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

// ---------------------------------------------------------------------------
// tiled pipeline: count -> scan -> fill -> order -> persistent grid.  Every
// step is a kernel launch (no memset, no host synchronisation between the
// passes of one phase), so the same enqueue functions serve the eager call
// and a captured hipGraph.
// ---------------------------------------------------------------------------
struct TiledCfg {
    int tu = 16;        // tile rows (8, 16 or 32); tiles are 32 columns wide
    int ch_lanes = 2;   // channels per binning wavefront (1, 2, 4 or 8)
    int waves = 0;      // wavefronts per tile, splitting its records (0: default)
};

struct TiledGeom {
    TiledCfg cfg;
    BinGeom g;
    int64_t n_bins, n_waves, bin_blocks, scan_nb;
};

struct TiledBufs {
    unsigned long long *counts = nullptr, *offsets = nullptr, *block_sums = nullptr;
    char *entries = nullptr;
    unsigned long long cap = 0;           // entries capacity (records)
    unsigned int *order = nullptr, *hist = nullptr;  // hist: [2][kBuckets] (counts, cursors)
    int *flags = nullptr;                 // [0] work queue, [1] entries overflow, [2] empty-tile queue
    const double *ntab = nullptr, *cgk = nullptr;
    int64_t cgk_len = 0;
    int half = 0;
};

// Tile rows of a call: the caller's choice, else 32 rows where the call has
// at least 65,536 (image channel, tile) bins of 32 x 32 cells, else 16.  With
// fewer bins the two workgroups per CU that a 32-row 4-pol tile allows have
// too few tiles to balance (C1's 614^2 grid, 16 channels, 6,400 bins: image
// call 0.89 ms with 16 rows, 1.02 ms with 32); with 380K bins (C2, C3) the
// taller tiles bin fewer footprints twice and win.
static int tile_rows_for(const ngc_context *ctx, const StdArgs &a) {
    if (ctx->tile_rows) return ctx->tile_rows;
    const int64_t bins32 = a.n_imag_chan * ceil_div(a.n_u, (int64_t)32) * ceil_div(a.n_v, (int64_t)32);
    return bins32 >= 65536 ? 32 : 16;
}

static TiledGeom tiled_geom(const StdArgs &a, const TiledCfg &cfg) {
    TiledGeom t;
    t.cfg = cfg;
    t.g.tu = cfg.tu;
    t.g.tu_log2 = 0;
    while ((1 << t.g.tu_log2) < cfg.tu) t.g.tu_log2++;
    t.g.ntu = ceil_div(a.n_u, cfg.tu);
    t.g.ntv = ceil_div(a.n_v, kTile);
    int ch_log2 = 0;
    while ((1 << (ch_log2 + 1)) <= std::min<int64_t>(a.n_chan, cfg.ch_lanes)) ch_log2++;
    t.g.ch_log2 = ch_log2;
    t.g.n_cb = ceil_div(a.n_chan, int64_t(1) << ch_log2);
    t.g.n_tb = ceil_div(a.n_time, int64_t(64) >> ch_log2);
    t.n_bins = a.n_imag_chan * t.g.ntu * t.g.ntv;
    t.n_waves = a.n_baseline * t.g.n_tb * t.g.n_cb;
    t.g.div_cb = FastDiv::of((uint32_t)t.g.n_cb);
    t.g.div_tb = FastDiv::of((uint32_t)t.g.n_tb);
    t.bin_blocks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(t.n_waves, kBlock / 64),
                                                           1 << 20));
    t.scan_nb = ceil_div(t.n_bins + 1, kScanChunk);
    return t;
}

static int tiled_supported(const StdArgs &a, bool pol_identity, int64_t cgk_len) {
    NGC_CHECK(a.support <= kMaxTiledSupport && a.support >= 2, NGC_ERR_UNSUPPORTED,
              "tiled gridder supports 2 <= support <= 8");
    NGC_CHECK(pol_identity, NGC_ERR_UNSUPPORTED, "tiled gridder needs pol_map = arange(n_pol)");
    NGC_CHECK(a.n_pol <= 4, NGC_ERR_UNSUPPORTED, "tiled gridder supports up to 4 pols");
    NGC_CHECK(a.n_u < 65536 && a.n_v < 65536, NGC_ERR_UNSUPPORTED, "grid axis >= 65536");
    NGC_CHECK(a.os < 30000, NGC_ERR_UNSUPPORTED, "oversampling too large");
    NGC_CHECK(cgk_len <= 512, NGC_ERR_UNSUPPORTED, "cgk_1d longer than 512");
    return NGC_OK;
}

// FP64 tap-sum table for sum_weight; into the context workspace, or into
// `owned` (a plan's buffer list) when given
static int tiled_tables(ngc_context *ctx, hipStream_t s, const StdArgs &a, const double *d_cgk,
                        const double *cgk, int64_t cgk_len, std::vector<void *> *owned,
                        TiledBufs &bf) {
    bf.half = (int)(a.os / 2 + 1);
    std::vector<double> ntab;
    norm_table(cgk, cgk_len, a.support, a.os, bf.half, ntab);
    void *d_ntab;
    if (owned) {
        NGC_HIP(hipMalloc(&d_ntab, ntab.size() * sizeof(double)));
        owned->push_back(d_ntab);
        NGC_HIP(hipMemcpyAsync(d_ntab, ntab.data(), ntab.size() * sizeof(double),
                               hipMemcpyHostToDevice, s));
        NGC_HIP(hipStreamSynchronize(s));
    } else {
        NGC_TRY(upload(ctx, s, "norm_tab", ntab.data(), ntab.size() * sizeof(double), &d_ntab));
    }
    bf.ntab = (const double *)d_ntab;
    bf.cgk = d_cgk;
    bf.cgk_len = cgk_len;
    return NGC_OK;
}

static int tiled_workspace(ngc_context *ctx, const TiledGeom &tg, TiledBufs &bf) {
    NGC_CHECK(tg.n_waves < (int64_t(1) << 32), NGC_ERR_UNSUPPORTED,
              "too many (baseline, time block, channel block) slots for one call");
    void *p;
    NGC_TRY(workspace(ctx, "bin_counts", (tg.n_bins + 1) * 8, &p));
    bf.counts = (unsigned long long *)p;
    NGC_TRY(workspace(ctx, "bin_offsets", (tg.n_bins + 1) * 8, &p));
    bf.offsets = (unsigned long long *)p;
    NGC_TRY(workspace(ctx, "scan_blocks", tg.scan_nb * 8, &p));
    bf.block_sums = (unsigned long long *)p;
    NGC_TRY(workspace(ctx, "bin_order", tg.n_bins * 4 + 4, &p));
    bf.order = (unsigned int *)p;
    NGC_TRY(workspace(ctx, "bin_hist", 2 * kBuckets * 4, &p));
    bf.hist = (unsigned int *)p;
    NGC_TRY(workspace(ctx, "tiled_flags", 4 * sizeof(int), &p));
    bf.flags = (int *)p;
    return NGC_OK;
}

static int fill_blocks(int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), 4096));
}

// count pass and scan: offsets[n_bins] becomes the number of records
template <int MODE, int NP, typename ACC>
static int enqueue_count_scan(ngc_context *ctx, hipStream_t s, const StdArgs &a,
                              const TiledGeom &tg, const TiledBufs &bf, double *sum_weight,
                              int accumulate) {
    const int64_t n_sw = a.n_imag_chan * a.n_imag_pol;
    fill_kernel<unsigned long long><<<fill_blocks(tg.n_bins + 1), 256, 0, s>>>(
        bf.counts, tg.n_bins + 1, 0ull);
    if (!accumulate) fill_kernel<double><<<fill_blocks(n_sw), 256, 0, s>>>(sum_weight, n_sw, 0.0);
    NGC_HIP(hipGetLastError());
    phase_begin(ctx, s, "bin_count");
    if (tg.n_waves > 0) {
        auto launch = [&](auto fast_c) {
            std_bin_kernel<false, MODE, NP, ACC, decltype(fast_c)::value>
                <<<(int)tg.bin_blocks, kBlock, 0, s>>>(a, tg.g, bf.counts, nullptr, 0ull, nullptr,
                                                       bf.ntab, bf.half, sum_weight, 0);
        };
        if constexpr (NP == 1 || NP == 2 || NP == 4) {
            if (bin_fast_loads(a, MODE, NP))
                launch(std::true_type{});
            else
                launch(std::false_type{});
        } else {
            launch(std::false_type{});
        }
        NGC_HIP(hipGetLastError());
    }
    phase_end(ctx, s);
    phase_begin(ctx, s, "bin_scan");
    scan_reduce_kernel<<<(int)tg.scan_nb, 256, 0, s>>>(bf.counts, tg.n_bins + 1, bf.block_sums);
    scan_blocks_kernel<<<1, 256, 0, s>>>(bf.block_sums, tg.scan_nb);
    scan_apply_kernel<<<(int)tg.scan_nb, 256, 0, s>>>(bf.counts, tg.n_bins + 1, bf.block_sums,
                                                      bf.offsets);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}

// fill pass, bin order and the persistent gridding kernel
template <int MODE, int NP, typename ACC, bool KAHAN, int TU, int W>
static int enqueue_fill_grid(ngc_context *ctx, hipStream_t s, const StdArgs &a,
                             const TiledGeom &tg, const TiledBufs &bf, void *grid,
                             double *sum_weight, int accumulate) {
    const int64_t n_sw = a.n_imag_chan * a.n_imag_pol;
    const int sw_lds = n_sw <= 4096 ? 1 : 0;
    const size_t sw_shmem = sw_lds ? n_sw * sizeof(double) : 0;
    copy_kernel<unsigned long long><<<fill_blocks(tg.n_bins + 1), 256, 0, s>>>(
        bf.counts, bf.offsets, tg.n_bins + 1);
    fill_kernel<int><<<1, 64, 0, s>>>(bf.flags, 4, 0);
    fill_kernel<unsigned int><<<1, 64, 0, s>>>(bf.hist, 2 * kBuckets, 0u);
    NGC_HIP(hipGetLastError());
    phase_begin(ctx, s, "bin_fill");
    if (tg.n_waves > 0) {
        auto launch = [&](auto fast_c) {
            std_bin_kernel<true, MODE, NP, ACC, decltype(fast_c)::value>
                <<<(int)tg.bin_blocks, kBlock, sw_shmem, s>>>(
                    a, tg.g, bf.counts, bf.entries, bf.cap, bf.flags + 1, bf.ntab, bf.half,
                    sum_weight, sw_lds);
        };
        if constexpr (NP == 1 || NP == 2 || NP == 4) {
            if (bin_fast_loads(a, MODE, NP))
                launch(std::true_type{});
            else
                launch(std::false_type{});
        } else {
            launch(std::false_type{});
        }
        NGC_HIP(hipGetLastError());
    }
    phase_end(ctx, s);
    phase_begin(ctx, s, "bin_order");
    order_hist_kernel<<<fill_blocks(tg.n_bins), 256, 0, s>>>(bf.offsets, tg.n_bins, bf.hist);
    order_prefix_kernel<<<1, 64, 0, s>>>(bf.hist, bf.hist + kBuckets);
    order_scatter_kernel<<<fill_blocks(tg.n_bins), 256, 0, s>>>(bf.offsets, tg.n_bins,
                                                                bf.hist + kBuckets, bf.order);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    phase_begin(ctx, s, "std_grid_tiled");
    auto kern = std_grid_tiled_kernel<MODE, NP, ACC, KAHAN, TU, W>;
    int per_cu = 1;
    NGC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * W, 0));
    const int64_t wgs = std::max<int64_t>(1, std::min<int64_t>(tg.n_bins,
                                                               (int64_t)per_cu * ctx->n_cu));
    kern<<<(unsigned)wgs, 64 * W, 0, s>>>(a, tg.g, bf.offsets, bf.order, bf.hist, bf.flags,
                                          (int)tg.n_bins,
                                      bf.cap, bf.entries, bf.cgk, (int)bf.cgk_len, grid,
                                      accumulate);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}

// Wavefronts per tile (record shares): `req` when it is 1, 2 or 4, else the
// measured best: 4 for the single-precision forms; 2 for FP64 tiled, whose
// accumulators and staged values take twice the registers.
static int tile_waves(bool single, int req) {
    if (req == 1 || req == 2 || req == 4) return req;
    return single ? 4 : 2;
}

// Calls f(mode_c, np_c, acc_c, kahan_c, tu_c, w_c) with the compile-time form of
// (mode, n_pol, precision, tile rows, wavefronts per tile).
template <typename F>
static int tiled_dispatch(int mode, int64_t n_pol, int precision, int tu, int waves, F &&f) {
    struct F32 { using type = float; };
    struct F64 { using type = double; };
    auto by_w = [&](auto m, auto np, auto acc, auto k, auto t) -> int {
        const int w = tile_waves(std::is_same<typename decltype(acc)::type, float>::value, waves);
        if (w == 4) return f(m, np, acc, k, t, std::integral_constant<int, 4>{});
        if (w == 2) return f(m, np, acc, k, t, std::integral_constant<int, 2>{});
        return f(m, np, acc, k, t, std::integral_constant<int, 1>{});
    };
    auto by_tu = [&](auto m, auto np, auto acc, auto k) -> int {
        if (tu == 32) return by_w(m, np, acc, k, std::integral_constant<int, 32>{});
        if (tu == 8) return by_w(m, np, acc, k, std::integral_constant<int, 8>{});
        return by_w(m, np, acc, k, std::integral_constant<int, 16>{});
    };
    auto by_prec = [&](auto m, auto np) -> int {
        if (precision == NGC_PREC_FP32_KAHAN) return by_tu(m, np, F32{}, std::true_type{});
        if (precision == NGC_PREC_FP32) return by_tu(m, np, F32{}, std::false_type{});
        return by_tu(m, np, F64{}, std::false_type{});
    };
    auto by_pol = [&](auto m) -> int {
        switch (n_pol) {
        case 1: return by_prec(m, std::integral_constant<int, 1>{});
        case 2: return by_prec(m, std::integral_constant<int, 2>{});
        case 3: return by_prec(m, std::integral_constant<int, 3>{});
        default: return by_prec(m, std::integral_constant<int, 4>{});
        }
    };
    if (mode == NGC_GRID_IMAGE) return by_pol(std::integral_constant<int, NGC_GRID_IMAGE>{});
    return by_pol(std::integral_constant<int, NGC_GRID_PSF>{});
}

template <int MODE, int NP, typename ACC> constexpr int entry_bytes() {
    return EntrySize<NP, (MODE == NGC_GRID_IMAGE) ? 2 : 1, ACC>::bytes;
}

// Eager run: count+scan, read the record count, size the entries, then
// fill+order+grid.  With `entries_total` the record count is returned.
static int tiled_run_eager(ngc_context *ctx, hipStream_t s, int mode, int precision,
                           const StdArgs &a, const TiledGeom &tg, TiledBufs &bf, void *grid,
                           double *sum_weight, int accumulate,
                           unsigned long long *entries_total) {
    return tiled_dispatch(mode, a.n_pol, precision, tg.cfg.tu, tg.cfg.waves,
                          [&](auto m, auto np, auto acc, auto k, auto tu, auto w) -> int {
        constexpr int MODE = decltype(m)::value;
        constexpr int NP = decltype(np)::value;
        using ACC = typename decltype(acc)::type;
        constexpr bool KAHAN = decltype(k)::value;
        constexpr int TU = decltype(tu)::value;
        constexpr int W = decltype(w)::value;
        NGC_TRY((enqueue_count_scan<MODE, NP, ACC>(ctx, s, a, tg, bf, sum_weight, accumulate)));
        unsigned long long total = 0;
        NGC_HIP(hipMemcpyAsync(&total, bf.offsets + tg.n_bins, sizeof(total),
                               hipMemcpyDeviceToHost, s));
        NGC_HIP(hipStreamSynchronize(s));
        if (entries_total) *entries_total = total;
        void *p;
        NGC_TRY(workspace(ctx, "bin_entries", (size_t)(total + 1) * entry_bytes<MODE, NP, ACC>(),
                          &p));
        bf.entries = (char *)p;
        bf.cap = total + 1;
        return enqueue_fill_grid<MODE, NP, ACC, KAHAN, TU, W>(ctx, s, a, tg, bf, grid, sum_weight,
                                                           accumulate);
    });
}

static int std_validate(int mode, int precision, const StdArgs &a, const double *freq,
                        const int32_t *chan_map, const int32_t *pol_map, const double *cgk,
                        int64_t cgk_len, const void *grid, const double *sum_weight,
                        bool *pol_identity_out) {
    NGC_CHECK(mode == NGC_GRID_IMAGE || mode == NGC_GRID_PSF || mode == NGC_GRID_DENSITY,
              NGC_ERR_ARG, "mode must be IMAGE, PSF or DENSITY");
    NGC_CHECK(precision >= NGC_PREC_FP64 && precision <= NGC_PREC_FP32, NGC_ERR_ARG,
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
    *pol_identity_out = pol_identity;
    return NGC_OK;
}

static int standard_grid_impl(ngc_context *ctx, hipStream_t s, int mode, int precision,
                              StdArgs a, const double *freq, const int32_t *chan_map,
                              const int32_t *pol_map, const double *cgk, int64_t cgk_len,
                              int accumulate, void *grid, double *sum_weight,
                              double delta_l, double delta_m) {
    bool pol_identity = false;
    NGC_TRY(std_validate(mode, precision, a, freq, chan_map, pol_map, cgk, cgk_len, grid,
                         sum_weight, &pol_identity));

    std::vector<double> scale;
    uv_scale_host(freq, a.n_chan, delta_l, delta_m, a.n_u, a.n_v, scale);
    void *d_scale, *d_cmap, *d_pmap, *d_cgk;
    NGC_TRY(upload(ctx, s, "uv_scale", scale.data(), scale.size() * sizeof(double), &d_scale));
    NGC_TRY(upload(ctx, s, "chan_map", padded_map(chan_map, a.n_chan).data(),
                   a.n_chan * sizeof(int32_t) + 4, &d_cmap));
    NGC_TRY(upload(ctx, s, "pol_map", pol_map, a.n_pol * sizeof(int32_t), &d_pmap));
    NGC_TRY(upload(ctx, s, "cgk", cgk, cgk_len * sizeof(double), &d_cgk));
    a.uv_scale = (const double *)d_scale;
    a.chan_map = (const int *)d_cmap;
    a.pol_map = (const int *)d_pmap;

    const bool single = (precision == NGC_PREC_FP32_KAHAN || precision == NGC_PREC_FP32);
    const bool compensated = (precision == NGC_PREC_FP32_KAHAN);
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
        else if (a.support == 1)
            NGC_TRY(density_runs(s, a, (const double *)d_cgk, cgk_len, g, sum_weight));
        else
            std_grid_plain_kernel<NGC_GRID_DENSITY><<<blocks, kBlock, shmem, s>>>(
                a, (const double *)d_cgk, cgk_len, g, sum_weight, sw_lds);
        NGC_HIP(hipGetLastError());
        phase_end(ctx, s);
        return NGC_OK;
    }

    // ---------------- tiled path ----------------
    NGC_TRY(tiled_supported(a, pol_identity, cgk_len));
    TiledCfg cfg;
    cfg.tu = tile_rows_for(ctx, a);
    cfg.ch_lanes = ctx->bin_chan_lanes;
    cfg.waves = ctx->tile_waves;
    const TiledGeom tg = tiled_geom(a, cfg);
    TiledBufs bf{};
    NGC_TRY(tiled_tables(ctx, s, a, (const double *)d_cgk, cgk, cgk_len, nullptr, bf));
    NGC_TRY(tiled_workspace(ctx, tg, bf));
    return tiled_run_eager(ctx, s, mode, precision, a, tg, bf, grid, sum_weight, accumulate,
                           nullptr);
}

// ---------------------------------------------------------------------------
// Allocation-free standard gridding (ngc_standard_grid_noalloc).  Every device
// buffer a call needs beyond its inputs and outputs -- the uploaded parameter
// tables, the FP64 tap-sum table, the bin counts, offsets, order and queues,
// and the bin records -- is carved out of `scratch`, one file-scope device
// array of NGC_NOALLOC_SCRATCH_BYTES that each GPU holds from the library's
// load until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: its device address on each GPU comes from
// hipGetSymbolAddress, once per device at context creation
// (noalloc_scratch_setup, with that device current), and only that address
// is ever passed to kernels or copies.  Calls on one GPU share the array and
// find the previous call's contents in it; everything that must start at zero
// is cleared by fill kernels in the call itself.  Calls that overlap on one
// GPU (different streams or contexts without ordering between them) race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[NGC_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxScratchDevices] = {};

// Bump allocation of 256-byte aligned regions of one call's scratch.
struct ScratchCarve {
    char *base;
    size_t cap, used;
    void *take(size_t bytes) {
        const size_t off = (used + 255) & ~size_t(255);
        if (bytes > cap || off > cap - bytes) return nullptr;
        used = off + bytes;
        return base + off;
    }
    size_t left() const {
        const size_t off = (used + 255) & ~size_t(255);
        return off < cap ? cap - off : 0;
    }
};
}  // namespace

int noalloc_scratch_setup(int device) {
    NGC_CHECK(device >= 0 && device < kMaxScratchDevices, NGC_ERR_UNSUPPORTED,
              "no-alloc scratch: device index above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_dev[device]) return NGC_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return NGC_ERR_HIP;
    }
    g_scratch_dev[device] = p;
    return NGC_OK;
}

static void *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_dev[device];
}

#define NGC_TAKE(ptr, carve, bytes, what)                                                   \
    do {                                                                                     \
        (ptr) = (carve).take(bytes);                                                         \
        if (!(ptr)) {                                                                        \
            set_error(std::string("ngc_standard_grid_noalloc: ") + (what) + " needs " +      \
                      std::to_string((size_t)(bytes)) + " bytes beyond the " +              \
                      std::to_string((carve).used) + " already taken of the " +             \
                      std::to_string((size_t)NGC_NOALLOC_SCRATCH_BYTES) +                   \
                      "-byte scratch; split the call or use ngc_standard_grid");             \
            return NGC_ERR_ARG;                                                              \
        }                                                                                    \
    } while (0)

static int standard_grid_noalloc_impl(ngc_context *ctx, hipStream_t s, int mode, int precision,
                                      StdArgs a, const double *freq, const int32_t *chan_map,
                                      const int32_t *pol_map, const double *cgk,
                                      int64_t cgk_len, int accumulate, void *grid,
                                      double *sum_weight, double delta_l, double delta_m) {
    bool pol_identity = false;
    NGC_TRY(std_validate(mode, precision, a, freq, chan_map, pol_map, cgk, cgk_len, grid,
                         sum_weight, &pol_identity));
    char *base = (char *)noalloc_scratch(ctx->device);
    NGC_CHECK(base, NGC_ERR_HIP, "no-alloc scratch is not set up on this device");
    ScratchCarve cv{base, (size_t)NGC_NOALLOC_SCRATCH_BYTES, 0};

    // parameter tables (the caller's host arrays are read for their exact sizes)
    std::vector<double> scale;
    uv_scale_host(freq, a.n_chan, delta_l, delta_m, a.n_u, a.n_v, scale);
    const std::vector<int32_t> cmap = padded_map(chan_map, a.n_chan);
    void *d_scale, *d_cmap, *d_pmap, *d_cgk;
    NGC_TAKE(d_scale, cv, scale.size() * sizeof(double), "the uv scale table");
    NGC_TAKE(d_cmap, cv, cmap.size() * sizeof(int32_t), "chan_map");
    NGC_TAKE(d_pmap, cv, (size_t)a.n_pol * sizeof(int32_t), "pol_map");
    NGC_TAKE(d_cgk, cv, (size_t)cgk_len * sizeof(double), "cgk_1d");
    const bool tiled = precision != NGC_PREC_FP64 && mode != NGC_GRID_DENSITY;
    const bool single = (precision == NGC_PREC_FP32_KAHAN || precision == NGC_PREC_FP32);
    NGC_CHECK(tiled || !single, NGC_ERR_UNSUPPORTED,
              "density gridding is FP64 only (precision NGC_PREC_FP64 or FP64_TILED)");
    TiledGeom tg{};
    TiledBufs bf{};
    std::vector<double> ntab;
    void *d_ntab = nullptr;
    if (tiled) {
        NGC_TRY(tiled_supported(a, pol_identity, cgk_len));
        TiledCfg cfg;
        cfg.tu = tile_rows_for(ctx, a);
        cfg.ch_lanes = ctx->bin_chan_lanes;
        cfg.waves = ctx->tile_waves;
        tg = tiled_geom(a, cfg);
        NGC_CHECK(tg.n_waves < (int64_t(1) << 32), NGC_ERR_UNSUPPORTED,
                  "too many (baseline, time block, channel block) slots for one call");
        bf.half = (int)(a.os / 2 + 1);
        norm_table(cgk, cgk_len, a.support, a.os, bf.half, ntab);
        NGC_TAKE(d_ntab, cv, ntab.size() * sizeof(double), "the tap-sum table");
        void *p;
        NGC_TAKE(p, cv, (size_t)(tg.n_bins + 1) * 8, "the bin counts");
        bf.counts = (unsigned long long *)p;
        NGC_TAKE(p, cv, (size_t)(tg.n_bins + 1) * 8, "the bin offsets");
        bf.offsets = (unsigned long long *)p;
        NGC_TAKE(p, cv, (size_t)tg.scan_nb * 8, "the scan partials");
        bf.block_sums = (unsigned long long *)p;
        NGC_TAKE(p, cv, (size_t)tg.n_bins * 4 + 4, "the bin order");
        bf.order = (unsigned int *)p;
        NGC_TAKE(p, cv, 2 * kBuckets * 4, "the bin-size histogram");
        bf.hist = (unsigned int *)p;
        NGC_TAKE(p, cv, 4 * sizeof(int), "the work-queue counters");
        bf.flags = (int *)p;
        bf.ntab = (const double *)d_ntab;
        bf.cgk = (const double *)d_cgk;
        bf.cgk_len = cgk_len;
    }
    // every size is known to fit before anything is enqueued (the bin records
    // are checked after the counting pass, which writes only scratch)
    NGC_HIP(hipMemcpyAsync(d_scale, scale.data(), scale.size() * sizeof(double),
                           hipMemcpyHostToDevice, s));
    NGC_HIP(hipMemcpyAsync(d_cmap, cmap.data(), cmap.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, s));
    NGC_HIP(hipMemcpyAsync(d_pmap, pol_map, (size_t)a.n_pol * sizeof(int32_t),
                           hipMemcpyHostToDevice, s));
    NGC_HIP(hipMemcpyAsync(d_cgk, cgk, (size_t)cgk_len * sizeof(double), hipMemcpyHostToDevice,
                           s));
    if (tiled)
        NGC_HIP(hipMemcpyAsync(d_ntab, ntab.data(), ntab.size() * sizeof(double),
                               hipMemcpyHostToDevice, s));
    a.uv_scale = (const double *)d_scale;
    a.chan_map = (const int *)d_cmap;
    a.pol_map = (const int *)d_pmap;
    const int64_t n_sw = a.n_imag_chan * a.n_imag_pol;
    profile_new_call(ctx, s);

    if (!tiled) {
        const int64_t nc = (mode == NGC_GRID_IMAGE) ? 2 : 1;
        const int64_t n_rc = a.n_time * a.n_baseline * a.n_chan;
        if (!accumulate) {
            const int64_t n_g = a.n_imag_chan * a.n_imag_pol * a.n_u * a.n_v * nc;
            fill_kernel<double><<<fill_blocks(n_g), 256, 0, s>>>((double *)grid, n_g, 0.0);
            fill_kernel<double><<<fill_blocks(n_sw), 256, 0, s>>>(sum_weight, n_sw, 0.0);
            NGC_HIP(hipGetLastError());
        }
        if (n_rc == 0) return NGC_OK;
        const int sw_lds = n_sw <= 4096 ? 1 : 0;
        const size_t shmem = sw_lds ? n_sw * sizeof(double) : 0;
        const int blocks = launch_blocks(n_rc);
        double *g = (double *)grid;
        if (mode == NGC_GRID_IMAGE)
            std_grid_plain_kernel<NGC_GRID_IMAGE><<<blocks, kBlock, shmem, s>>>(
                a, (const double *)d_cgk, cgk_len, g, sum_weight, sw_lds);
        else if (mode == NGC_GRID_PSF)
            std_grid_plain_kernel<NGC_GRID_PSF><<<blocks, kBlock, shmem, s>>>(
                a, (const double *)d_cgk, cgk_len, g, sum_weight, sw_lds);
        else if (a.support == 1)
            NGC_TRY(density_runs(s, a, (const double *)d_cgk, cgk_len, g, sum_weight));
        else
            std_grid_plain_kernel<NGC_GRID_DENSITY><<<blocks, kBlock, shmem, s>>>(
                a, (const double *)d_cgk, cgk_len, g, sum_weight, sw_lds);
        NGC_HIP(hipGetLastError());
        return NGC_OK;
    }

    return tiled_dispatch(mode, a.n_pol, precision, tg.cfg.tu, tg.cfg.waves,
                          [&](auto m, auto np, auto acc, auto k, auto tu, auto w) -> int {
        constexpr int MODE = decltype(m)::value;
        constexpr int NP = decltype(np)::value;
        using ACC = typename decltype(acc)::type;
        constexpr bool KAHAN = decltype(k)::value;
        constexpr int TU = decltype(tu)::value;
        constexpr int W = decltype(w)::value;
        // counting pass with accumulate = 1: it leaves sum_weight untouched, so
        // a call refused below has not modified the outputs
        NGC_TRY((enqueue_count_scan<MODE, NP, ACC>(ctx, s, a, tg, bf, sum_weight, 1)));
        unsigned long long total = 0;
        NGC_HIP(hipMemcpyAsync(&total, bf.offsets + tg.n_bins, sizeof(total),
                               hipMemcpyDeviceToHost, s));
        NGC_HIP(hipStreamSynchronize(s));
        constexpr size_t EB = (size_t)entry_bytes<MODE, NP, ACC>();
        void *p;
        NGC_TAKE(p, cv, (size_t)(total + 1) * EB, "the bin records");
        bf.entries = (char *)p;
        bf.cap = total + 1;
        if (!accumulate) {
            fill_kernel<double><<<fill_blocks(n_sw), 256, 0, s>>>(sum_weight, n_sw, 0.0);
            NGC_HIP(hipGetLastError());
        }
        return enqueue_fill_grid<MODE, NP, ACC, KAHAN, TU, W>(ctx, s, a, tg, bf, grid,
                                                           sum_weight, accumulate);
    });
}
#undef NGC_TAKE
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

extern "C" int ngc_standard_grid_noalloc(ngc_context *ctx, void *stream, int mode,
                                         int precision, const double *uvw, const void *vis,
                                         int vis_dtype, const void *weight, int weight_dtype,
                                         const uint8_t *flag, const double *freq_chan,
                                         const int32_t *chan_map, const int32_t *pol_map,
                                         const double *cgk_1d, int64_t cgk_len, int64_t n_time,
                                         int64_t n_baseline, int64_t n_chan, int64_t n_pol,
                                         int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
                                         int64_t n_v, double delta_l, double delta_m,
                                         int64_t support, int64_t oversampling, int accumulate,
                                         void *grid, double *sum_weight) {
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
    return ngc::standard_grid_noalloc_impl(ctx, s, mode, precision, a, freq_chan, chan_map,
                                           pol_map, cgk_1d, cgk_len, accumulate, grid,
                                           sum_weight, delta_l, delta_m);
}

// ---------------------------------------------------------------------------
// gridding plans: the whole tiled pipeline of one ngc_standard_grid call,
// autotuned once and captured in a hipGraph
// ---------------------------------------------------------------------------
namespace ngc {

void plan_free(ngc_grid_plan *p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    for (void *q : p->owned) (void)hipFree(q);
    delete p;
}

int plan_alloc(ngc_grid_plan *p, size_t bytes, void **out) {
    void *q = nullptr;
    NGC_HIP(hipMalloc(&q, bytes ? bytes : 256));
    p->owned.push_back(q);
    *out = q;
    return NGC_OK;
}

int plan_upload(ngc_grid_plan *p, hipStream_t s, const void *host, size_t bytes,
                void **out) {
    NGC_TRY(plan_alloc(p, bytes, out));
    NGC_HIP(hipMemcpyAsync(*out, host, bytes, hipMemcpyHostToDevice, s));
    return NGC_OK;
}

static std::string shape_key(int mode, int precision, const StdArgs &a) {
    char buf[256];
    snprintf(buf, sizeof(buf), "%d/%d/%lld/%lld/%lld/%lld/%lld/%lld/%lld/%lld/%lld", mode,
             precision, (long long)a.n_time, (long long)a.n_baseline, (long long)a.n_chan,
             (long long)a.n_pol, (long long)a.n_imag_chan, (long long)a.n_u, (long long)a.n_v,
             (long long)a.support, (long long)a.os);
    return buf;
}

static int plan_create(ngc_context *ctx, hipStream_t s, int mode, int precision, StdArgs a,
                       const double *freq, const int32_t *chan_map, const int32_t *pol_map,
                       const double *cgk, int64_t cgk_len, double delta_l, double delta_m,
                       int accumulate, void *grid, double *sum_weight, int autotune,
                       ngc_grid_plan *p) {
    bool pol_identity = false;
    NGC_TRY(std_validate(mode, precision, a, freq, chan_map, pol_map, cgk, cgk_len, grid,
                         sum_weight, &pol_identity));
    NGC_CHECK(mode == NGC_GRID_IMAGE || mode == NGC_GRID_PSF, NGC_ERR_UNSUPPORTED,
              "gridding plans cover image and PSF modes");
    NGC_CHECK(precision != NGC_PREC_FP64, NGC_ERR_UNSUPPORTED,
              "gridding plans run the tiled forms (FP64_TILED, FP32_KAHAN, FP32)");
    NGC_TRY(tiled_supported(a, pol_identity, cgk_len));

    // plan-owned tables
    std::vector<double> scale;
    uv_scale_host(freq, a.n_chan, delta_l, delta_m, a.n_u, a.n_v, scale);
    void *d_scale, *d_cmap, *d_pmap, *d_cgk;
    NGC_TRY(plan_upload(p, s, scale.data(), scale.size() * sizeof(double), &d_scale));
    NGC_TRY(plan_upload(p, s, padded_map(chan_map, a.n_chan).data(), a.n_chan * sizeof(int32_t) + 4,
                        &d_cmap));
    NGC_TRY(plan_upload(p, s, pol_map, a.n_pol * sizeof(int32_t), &d_pmap));
    NGC_TRY(plan_upload(p, s, cgk, cgk_len * sizeof(double), &d_cgk));
    a.uv_scale = (const double *)d_scale;
    a.chan_map = (const int *)d_cmap;
    a.pol_map = (const int *)d_pmap;
    TiledBufs tables{};
    NGC_TRY(tiled_tables(ctx, s, a, (const double *)d_cgk, cgk, cgk_len, &p->owned, tables));

    const bool saved_profile = ctx->profile;
    ctx->profile = false;
    struct Restore {
        ngc_context *c;
        bool v;
        ~Restore() { c->profile = v; }
    } restore{ctx, saved_profile};

    // autotuner: time the whole eager pipeline for every candidate shape on
    // these inputs (one warm-up and one timed run each) and keep the fastest
    const std::string key = shape_key(mode, precision, a);
    TiledCfg best;
    best.tu = tile_rows_for(ctx, a);
    best.ch_lanes = ctx->bin_chan_lanes;
    best.waves = ctx->tile_waves;
    std::string tune_log;
    auto it = ctx->tuned.find(key);
    if (autotune || it == ctx->tuned.end()) {
        if (autotune) {
            hipEvent_t e0, e1;
            NGC_HIP(hipEventCreate(&e0));
            NGC_HIP(hipEventCreate(&e1));
            float best_ms = 1e30f;
            for (int tu : {8, 16, 32})
                for (int cl : {1, 2, 4, 8}) {
                    TiledCfg cfg;
                    cfg.tu = tu;
                    cfg.ch_lanes = cl;
                    cfg.waves = ctx->tile_waves;
                    const TiledGeom tg = tiled_geom(a, cfg);
                    TiledBufs bf = tables;
                    NGC_TRY(tiled_workspace(ctx, tg, bf));
                    float ms = 0.f;
                    for (int rep = 0; rep < 2; rep++) {
                        NGC_HIP(hipEventRecord(e0, s));
                        NGC_TRY(tiled_run_eager(ctx, s, mode, precision, a, tg, bf, grid,
                                                sum_weight, 0, nullptr));
                        NGC_HIP(hipEventRecord(e1, s));
                        NGC_HIP(hipEventSynchronize(e1));
                        NGC_HIP(hipEventElapsedTime(&ms, e0, e1));
                    }
                    char b[64];
                    snprintf(b, sizeof(b), "%s%dx32/%d:%.3f", tune_log.empty() ? "" : " ", tu,
                             cl, ms);
                    tune_log += b;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best = cfg;
                    }
                }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            ctx->tuned[key] = {best.tu, best.ch_lanes};
        }
    } else {
        best.tu = it->second.first;
        best.ch_lanes = it->second.second;
    }

    // plan-owned pipeline buffers sized by one eager count pass
    const TiledGeom tg = tiled_geom(a, best);
    TiledBufs bf = tables;
    void *q;
    NGC_TRY(plan_alloc(p, (tg.n_bins + 1) * 8, &q));
    bf.counts = (unsigned long long *)q;
    NGC_TRY(plan_alloc(p, (tg.n_bins + 1) * 8, &q));
    bf.offsets = (unsigned long long *)q;
    NGC_TRY(plan_alloc(p, tg.scan_nb * 8, &q));
    bf.block_sums = (unsigned long long *)q;
    NGC_TRY(plan_alloc(p, tg.n_bins * 4 + 4, &q));
    bf.order = (unsigned int *)q;
    NGC_TRY(plan_alloc(p, 2 * kBuckets * 4, &q));
    bf.hist = (unsigned int *)q;
    NGC_TRY(plan_alloc(p, 4 * sizeof(int), &q));
    bf.flags = (int *)q;
    p->flags = bf.flags;
    int rc = tiled_dispatch(mode, a.n_pol, precision, best.tu, best.waves,
                            [&](auto m, auto np, auto acc, auto k, auto tu, auto w) -> int {
        constexpr int MODE = decltype(m)::value;
        constexpr int NP = decltype(np)::value;
        using ACC = typename decltype(acc)::type;
        constexpr bool KAHAN = decltype(k)::value;
        constexpr int TU = decltype(tu)::value;
        constexpr int W = decltype(w)::value;
        NGC_TRY((enqueue_count_scan<MODE, NP, ACC>(ctx, s, a, tg, bf, sum_weight, accumulate)));
        unsigned long long total = 0;
        NGC_HIP(hipMemcpyAsync(&total, bf.offsets + tg.n_bins, sizeof(total),
                               hipMemcpyDeviceToHost, s));
        NGC_HIP(hipStreamSynchronize(s));
        // room for the record count to grow by 1/16 (inputs rewritten in place)
        bf.cap = total + total / 16 + 1024;
        void *e;
        NGC_TRY(plan_alloc(p, (size_t)bf.cap * entry_bytes<MODE, NP, ACC>(), &e));
        bf.entries = (char *)e;
        NGC_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        NGC_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        int r = enqueue_count_scan<MODE, NP, ACC>(ctx, p->cap_stream, a, tg, bf, sum_weight,
                                                  accumulate);
        if (r == NGC_OK)
            r = enqueue_fill_grid<MODE, NP, ACC, KAHAN, TU, W>(ctx, p->cap_stream, a, tg, bf, grid,
                                                            sum_weight, accumulate);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(p->cap_stream, &graph);
        p->graph = graph;
        if (r != NGC_OK) return r;
        NGC_CHECK(ec == hipSuccess && graph, NGC_ERR_HIP, "hipStreamEndCapture failed");
        NGC_HIP(hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0));
        char b[256];
        snprintf(b, sizeof(b),
                 "tile %dx32, %d wavefronts per tile, %d channel lanes, %lld bins, %llu records "
                 "(cap %llu)",
                 best.tu, W, best.ch_lanes, (long long)tg.n_bins, total, bf.cap);
        p->info = b;
        return NGC_OK;
    });
    if (!tune_log.empty()) p->info += "; tuning ms " + tune_log;
    return rc;
}

}  // namespace ngc

extern "C" {

int ngc_grid_plan_create(ngc_context *ctx, void *stream, int mode, int precision,
                         const double *uvw, const void *vis, int vis_dtype, const void *weight,
                         int weight_dtype, const uint8_t *flag, const double *freq_chan,
                         const int32_t *chan_map, const int32_t *pol_map, const double *cgk_1d,
                         int64_t cgk_len, int64_t n_time, int64_t n_baseline, int64_t n_chan,
                         int64_t n_pol, int64_t n_imag_chan, int64_t n_imag_pol, int64_t n_u,
                         int64_t n_v, double delta_l, double delta_m, int64_t support,
                         int64_t oversampling, int accumulate, void *grid, double *sum_weight,
                         int autotune, ngc_grid_plan **plan) {
    NGC_CHECK(ctx && plan, NGC_ERR_ARG, "ctx or plan is NULL");
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
    ngc_grid_plan *p = new ngc_grid_plan();
    p->ctx = ctx;
    const int rc = ngc::plan_create(ctx, s, mode, precision, a, freq_chan, chan_map, pol_map,
                                    cgk_1d, cgk_len, delta_l, delta_m, accumulate, grid,
                                    sum_weight, autotune, p);
    if (rc != NGC_OK) {
        (void)hipStreamSynchronize(s);
        ngc::plan_free(p);
        return rc;
    }
    *plan = p;
    return NGC_OK;
}

int ngc_grid_plan_execute(ngc_grid_plan *plan, void *stream) {
    NGC_CHECK(plan && plan->exec, NGC_ERR_ARG, "plan is NULL");
    NGC_HIP(hipSetDevice(plan->ctx->device));
    NGC_HIP(hipGraphLaunch(plan->exec, (hipStream_t)stream));
    return NGC_OK;
}

int ngc_grid_plan_status(ngc_grid_plan *plan, void *stream) {
    NGC_CHECK(plan, NGC_ERR_ARG, "plan is NULL");
    NGC_HIP(hipSetDevice(plan->ctx->device));
    int flags[4] = {0, 0, 0, 0};
    NGC_HIP(hipMemcpyAsync(flags, plan->flags, sizeof(flags), hipMemcpyDeviceToHost,
                           (hipStream_t)stream));
    NGC_HIP(hipStreamSynchronize((hipStream_t)stream));
    NGC_CHECK(flags[1] == 0, NGC_ERR_ARG,
              "the inputs now bin more records than the plan holds; make a new plan");
    NGC_CHECK(flags[3] == 0, NGC_ERR_FIELD, "a row's field id is not in field_id");
    return NGC_OK;
}

int ngc_grid_plan_describe(ngc_grid_plan *plan, char *buf, int64_t len) {
    NGC_CHECK(plan && buf && len > 0, NGC_ERR_ARG, "bad arguments");
    const size_t k = std::min<size_t>(plan->info.size(), (size_t)len - 1);
    memcpy(buf, plan->info.data(), k);
    buf[k] = 0;
    return NGC_OK;
}

int ngc_grid_plan_destroy(ngc_grid_plan *plan) {
    if (!plan) return NGC_OK;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipDeviceSynchronize();
    ngc::plan_free(plan);
    return NGC_OK;
}

}  // extern "C"
