// Tiled aperture (mosaic A-term) gridder for gfx950: the NGC_PREC_FP64_TILED,
// NGC_PREC_FP32_KAHAN and NGC_PREC_FP32 forms behind ngc_aperture_grid().
//
// Replaces _aperture_grid_jit (ngcasa/imaging/_imaging_utils/_aperture_grid.py:377-513)
// and _aperture_weight_grid_jit (:181-291).
//
// Image and PSF modes:
//  1. ap_norm_kernel: the complex tap sum norm = sum CK*PG of every
//     (convolution function, field, sub-cell offset pair), in the reference's
//     tap order and FP64, so every sample's sum_weight term (w Re(norm^2) or
//     w Re(norm), :508-511) is the plain path's to the last bit.
//  2. ap_taps_kernel: the kernel re-laid per sub-cell offset pair as an R x R
//     block (R = max support), so the taps of one sample are contiguous.
//     ap_pg_kernel: the phase gradient split into its u and v factors
//     (PG[x][y] = PG[x][c]*PG[c][y] for the reference's exp(i(x px + y py)));
//     a table that does not factor is refused.
//  3. ap_bin_kernel (count, scan, fill): every gridded (row, channel) goes to
//     the bin of the (support class, image channel, T x T tile) holding its
//     centre cell, in time order per baseline; the fill pass adds the FP64
//     sum_weight term.  A support class gathers the convolution functions
//     whose largest support needs the same number of lane wavefronts.
//  4. ap_tiled_kernel, one launch per support class: one workgroup of R*R
//     lanes per bin, R the class's largest support.  Lane (a, b) owns every
//     cell of the tile's halo region (T + R - 1 cells square) with u = a,
//     v = b (mod R); a footprint of support <= R meets each residue pair at
//     most once, so a lane has at most one tap per sample and no two lanes
//     share a cell (J. W. Romein, ICS 2012).  The lane keeps its cell's
//     running sum in registers and flushes it into the LDS region (FP64
//     cells with FP32_KAHAN) when its cell changes.  The region is written,
//     in the grid's precision, to a scratch slot of the bin.
//  5. ap_merge_kernel: only non-empty bins have scratch slots; every cell of
//     a tile next to a non-empty bin adds, in a fixed order and FP64, the
//     regions that cover it (at most four per support class; the grid is
//     zeroed first unless accumulating), so a chunk of a large cube touches
//     only its part of the grid.
//
// Weight mode: every sample lands on the same centre taps of its (image
// channel, pol, convolution function, field) key, so the weight grid is
// sum_key W_key * CK*PG[centre taps] with W_key the key's sum of imaging
// weights: ap_wsum_kernel reduces W per key in FP64 (and adds w Re(norm) to
// sum_weight), ap_wstamp_kernel stamps the keys.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <functional>
#include <type_traits>

#include "ngc_internal.hpp"

namespace ngc {

constexpr int kApBatch = 32;   // samples staged per batch
constexpr int kApGroup = 8;    // samples whose tap loads are issued together
constexpr int kApBinBlock = 256;
constexpr int kApMaxCls = 16;  // support classes (lane wavefronts 1..16)
// Dword 3 of a raw buffer resource on gfx9 parts (32-bit data format, no
// index stride or swizzle): range checking then compares the byte offset with
// num_records, and a load past it returns zero.
constexpr int kBufferWord3 = 0x00020000;

struct ApT {
    const double *uvw;
    const void *vis;
    const void *weight;
    const uint8_t *flag;
    const int *field;      // [T, B]
    const int *field_id;   // [n_field]
    const int *field_map;  // [field_map_len]: index in field_id of each id, or -1
    int64_t field_map_len; // 0: search field_id instead
    const double *uv_scale;  // [2][n_chan]
    const int *chan_map, *cfb, *cfc;
    const int *support;    // [n_cf][2]
    int64_t n_field, n_cfc, n_cfp, cfp0, n_cf;
    int64_t n_time, n_baseline, n_chan, n_pol, n_imag_chan, n_imag_pol, n_u, n_v;
    int64_t os_u, os_v, k_u, k_v;
    int R, T, W, hl;       // R: max(SUPPORT), the tap blocks' edge
    int Rl;                // the tiled kernel's lanes per axis (its class's support)
    int64_t ntu, ntv;
    int64_t per_cls;       // bins per support class: n_imag_chan ntu ntv
    int64_t slot_elems;    // elements of a scratch slot: 2 n_pol W_max^2
    const int *cf_cls;     // [n_cf] support class of each convolution function
    int ch_log2;
    int64_t n_tb, n_cb;
    FastDiv div_cb, div_tb;  // by n_cb and n_tb (n_baseline n_tb n_cb < 2^32)
    int half_u, half_v;    // sub-cell offset index range [-half, half]
    int vis_dtype, weight_dtype;
    int *error;
};

__device__ inline double apt_weight(const ApT &a, int64_t i) {
    double w = a.weight_dtype == NGC_F32 ? (double)((const float *)a.weight)[i]
                                         : ((const double *)a.weight)[i];
    if (a.flag && a.flag[i]) w = __longlong_as_double(0x7ff8000000000000LL);
    return w;
}

__device__ inline double2 apt_vis(const ApT &a, int64_t i) {
    if (a.vis_dtype == NGC_C64) {
        const float2 v = ((const float2 *)a.vis)[i];
        return make_double2(v.x, v.y);
    }
    return ((const double2 *)a.vis)[i];
}

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
__global__ void ap_norm_kernel(ApT a, const double *ck, const double2 *pg, double2 *ntab) {
#pragma clang fp contract(off)
    const int nwu = 2 * a.half_u + 1, nwv = 2 * a.half_v + 1;
    const int64_t n = a.n_cf * a.n_field * nwu * nwv;
    const int64_t kplane = a.k_u * a.k_v;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ov = i % nwv - a.half_v;
        const int64_t ou = (i / nwv) % nwu - a.half_u;
        const int64_t f = (i / ((int64_t)nwu * nwv)) % a.n_field;
        const int64_t cf = i / ((int64_t)nwu * nwv * a.n_field);
        const int64_t su = a.support[cf * 2], sv = a.support[cf * 2 + 1];
        const int64_t scu = su / 2, scv = sv / 2;
        const int64_t u_off = ou + a.k_u / 2, v_off = ov + a.k_v / 2;
        double nr = 0.0, ni = 0.0;
        for (int64_t iv = -scv; iv < sv - scv; iv++) {
            const int64_t cf_v = a.os_v * iv + v_off;
            for (int64_t iu = -scu; iu < su - scu; iu++) {
                const int64_t cf_u = a.os_u * iu + u_off;
                if (cf_u < 0 || cf_u >= a.k_u || cf_v < 0 || cf_v >= a.k_v) continue;
                const int64_t k = cf_u * a.k_v + cf_v;
                const double kv = ck[cf * kplane + k];
                const double2 g = pg[f * kplane + k];
                nr = nr + kv * g.x;
                ni = ni + kv * g.y;
            }
        }
        ntab[i] = make_double2(nr, ni);
    }
}

// ckr[cf][ou][ov][ru][rv] = CK[cf][os_u (ru - su/2) + ou + k_u/2][...], zero
// outside the support
template <typename ACC>
__global__ void ap_taps_kernel(ApT a, const double *ck, ACC *ckr) {
    const int nwu = 2 * a.half_u + 1, nwv = 2 * a.half_v + 1;
    const int R = a.R;
    const int64_t n = a.n_cf * nwu * nwv * R * R;
    const int64_t kplane = a.k_u * a.k_v;
    // element n is zero: the end of the table, where the tiled kernel's taps
    // outside a sample's support point (its loads past the end read zero)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (i == n) {
            ckr[i] = (ACC)0;
            continue;
        }
        const int rv = (int)(i % R), ru = (int)((i / R) % R);
        const int64_t o = i / ((int64_t)R * R);
        const int64_t ov = o % nwv - a.half_v;
        const int64_t ou = (o / nwv) % nwu - a.half_u;
        const int64_t cf = o / ((int64_t)nwu * nwv);
        const int64_t su = a.support[cf * 2], sv = a.support[cf * 2 + 1];
        double v = 0.0;
        if (ru < su && rv < sv) {
            const int64_t cf_u = a.os_u * (ru - su / 2) + ou + a.k_u / 2;
            const int64_t cf_v = a.os_v * (rv - sv / 2) + ov + a.k_v / 2;
            if (cf_u >= 0 && cf_u < a.k_u && cf_v >= 0 && cf_v < a.k_v)
                v = ck[cf * kplane + cf_u * a.k_v + cf_v];
        }
        ckr[i] = (ACC)v;
    }
}

template <typename ACC2>
__global__ void ap_pg_kernel(ApT a, const double2 *pg, ACC2 *pgu, ACC2 *pgv, int *bad) {
#pragma clang fp contract(off)
    const int64_t cu = a.k_u / 2, cv = a.k_v / 2;
    const int64_t kplane = a.k_u * a.k_v;
    const int64_t n = a.n_field * kplane;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = i / kplane, r = i - f * kplane;
        const int64_t x = r / a.k_v, y = r - x * a.k_v;
        const double2 gx = pg[f * kplane + x * a.k_v + cv];
        const double2 gy = pg[f * kplane + cu * a.k_v + y];
        const double2 g = pg[i];
        const double pr = gx.x * gy.x - gx.y * gy.y, pi = gx.x * gy.y + gx.y * gy.x;
        if (fabs(pr - g.x) > 1e-9 || fabs(pi - g.y) > 1e-9) atomicOr(bad, 1);
        if (y == cv) {
            ACC2 o;
            o.x = gx.x;
            o.y = gx.y;
            pgu[f * a.k_u + x] = o;
        }
        if (x == cu) {
            ACC2 o;
            o.x = gy.x;
            o.y = gy.y;
            pgv[f * a.k_v + y] = o;
        }
    }
}

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
template <int NP, int NW, typename ACC> struct ApEntry {
    int uv;    // uc | vc << 16
    int off;   // (ou & 0xffff) | ov << 16
    int meta;  // cf | field index << 16
    int pad;
    ACC wd[NP][NW];
};
template <int NP, int NW, typename ACC> struct ApEntrySize {
    static constexpr int bytes = ((int)sizeof(ApEntry<NP, NW, ACC>) + 15) / 16 * 16;
};

// Support classes of the tiled kernel: the convolution functions whose
// supports need the same number of lane wavefronts (ceil(R^2 / 64)) share a
// class, gridded by R x R lanes with R the class's largest support.
struct ApCls {
    int n;
    int R[kApMaxCls], W[kApMaxCls], hl[kApMaxCls];
};

struct ApSample {
    int64_t t, b, ch, row, rc;
};

__device__ inline bool ap_sample_at(const ApT &a, int64_t wave_id, int lane, ApSample &sm) {
    const int chl = 1 << a.ch_log2;
    const int tl = 64 >> a.ch_log2;
    // wave_id is the same in every lane: the divisions are scalar work
    const uint32_t wid = (uint32_t)wave_id;
    const uint32_t r = a.div_cb.div(wid);
    const uint32_t cb = wid - r * a.div_cb.d;
    const uint32_t b = a.div_tb.div(r);
    const uint32_t tb = r - b * a.div_tb.d;
    sm.t = (int64_t)tb * tl + (lane >> a.ch_log2);
    sm.ch = cb * chl + (lane & (chl - 1));
    sm.b = b;
    if (b >= a.n_baseline || sm.t >= a.n_time || sm.ch >= a.n_chan) return false;
    sm.row = sm.t * a.n_baseline + b;
    sm.rc = sm.row * a.n_chan + sm.ch;
    return true;
}

// field index, centre cell and offsets of a sample; false when not gridded
__device__ inline bool ap_place(const ApT &a, const ApSample &sm, int &fidx, Pix &px, Pix &pxv) {
    const int fval = a.field[sm.row];
    if (!(fval > -1)) return false;
    fidx = -1;
    if (a.field_map_len > 0) {
        if (fval < a.field_map_len) fidx = a.field_map[fval];
    } else {
        for (int64_t f = 0; f < a.n_field; f++)
            if (a.field_id[f] == fval) {
                fidx = (int)f;
                break;
            }
    }
    if (fidx < 0) {
        atomicOr(a.error, 1);
        return false;
    }
    const double *uvw = a.uvw + sm.row * 3;
    if (!pixel_position(uvw[0], uvw[1], a.uv_scale[sm.ch], a.uv_scale[a.n_chan + sm.ch], a.n_u,
                        a.n_v, a.os_u, px))
        return false;
    pixel_position(uvw[0], uvw[1], a.uv_scale[sm.ch], a.uv_scale[a.n_chan + sm.ch], a.n_u, a.n_v,
                   a.os_v, pxv);
    const int64_t ms = a.R;
    if (!((px.uc + ms < a.n_u) && (px.vc + ms < a.n_v) && (px.uc - ms >= 0) &&
          (px.vc - ms >= 0)))
        return false;
    if (a.flag) {
        bool any = false;
        for (int p = 0; p < a.n_pol; p++) any |= (a.flag[sm.rc * a.n_pol + p] == 0);
        if (!any) return false;
    }
    return true;
}

template <bool FILL, int MODE, int NP, typename ACC>
__global__ void __launch_bounds__(kApBinBlock)
ap_bin_kernel(ApT a, unsigned long long *counts, char *entries, const double2 *ntab,
              double *sum_weight, int sw_copies, const int *halt) {
#pragma clang fp contract(off)
    constexpr int NW = (MODE == NGC_GRID_IMAGE) ? 2 : 1;
    using E = ApEntry<NP, NW, ACC>;
    constexpr int EB = ApEntrySize<NP, NW, ACC>::bytes;
    extern __shared__ double s_sw[];
    // a plan whose buffers the inputs outgrew: nothing is written
    if (FILL && halt && *halt) return;
    const int64_t n_sw = a.n_imag_chan * a.n_imag_pol;
    // sum_weight terms: sw_copies > 0 copies in LDS (a lane adds to copy
    // lane mod sw_copies, so fewer lanes of a wavefront meet on one address),
    // summed into sum_weight at the end; 0: atomics on sum_weight itself
    if (FILL && sw_copies) {
        for (int64_t i = threadIdx.x; i < n_sw * sw_copies; i += blockDim.x) s_sw[i] = 0.0;
        __syncthreads();
    }
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t n_waves = a.n_baseline * a.n_tb * a.n_cb;
    const int waves_per_block = blockDim.x >> 6;
    const int nwu = 2 * a.half_u + 1, nwv = 2 * a.half_v + 1;
    for (int64_t w0 = (int64_t)blockIdx.x * waves_per_block; w0 < n_waves;
         w0 += (int64_t)gridDim.x * waves_per_block) {
        const int64_t wave_id = w0 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        ApSample sm{};
        Pix px{}, pxv{};
        int fidx = 0;
        const bool ok = (wave_id < n_waves) && ap_sample_at(a, wave_id, lane, sm) &&
                        ap_place(a, sm, fidx, px, pxv);
        int64_t key = -1;
        E rec;
        if (ok) {
            const int64_t a_chan = a.chan_map[sm.ch];
            const int64_t cf = ((int64_t)a.cfb[sm.b] * a.n_cfc + a.cfc[sm.ch]) * a.n_cfp + a.cfp0;
            key = a.cf_cls[cf] * a.per_cls + (a_chan * a.ntu + px.uc / a.T) * a.ntv + pxv.vc / a.T;
            if (FILL) {
                // |offset| <= 1/2 cell, so |ou| <= os_u/2 < half_u; the clamp
                // only bounds the table indices
                const int ou = (int)min(max(px.du, (int64_t)-a.half_u), (int64_t)a.half_u);
                const int ov = (int)min(max(pxv.dv, (int64_t)-a.half_v), (int64_t)a.half_v);
                rec.uv = (int)((px.uc & 0xffff) | (pxv.vc << 16));
                rec.off = (ou & 0xffff) | (ov << 16);
                rec.meta = (int)(cf | ((int64_t)fidx << 16));
                rec.pad = 0;
                const double2 nm =
                    ntab[((cf * a.n_field + fidx) * nwu + ou + a.half_u) * nwv + ov + a.half_v];
                const int64_t base = sm.rc * a.n_pol;
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const double w = apt_weight(a, base + p);
                    double wr = w, wi = 0.0;
                    if (MODE == NGC_GRID_IMAGE) {
                        const double2 v = apt_vis(a, base + p);
                        wr = v.x * w;
                        wi = v.y * w;
                    }
                    const bool use = !(isnan(wr) || isnan(wi)) && !(wr == 0.0 && wi == 0.0);
                    rec.wd[p][0] = use ? (ACC)wr : (ACC)0;
                    if (NW == 2) rec.wd[p][NW - 1] = use ? (ACC)wi : (ACC)0;
                    if (use) {
                        const double term = MODE == NGC_GRID_IMAGE
                                                ? w * (nm.x * nm.x - nm.y * nm.y)
                                                : w * nm.x;
                        const int64_t swi = a_chan * a.n_imag_pol + p;
                        if (sw_copies)
                            atomicAdd(&s_sw[swi * sw_copies + (lane & (sw_copies - 1))], term);
                        else
                            atomicAdd(&sum_weight[swi], term);
                    }
                }
            }
        }
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
                    *reinterpret_cast<E *>(entries + pos * EB) = rec;
                }
            }
            if (key == lkey) key = -1;
            pending &= ~same;
        }
    }
    if (FILL && sw_copies) {
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n_sw; i += blockDim.x) {
            double t = 0.0;
            for (int c = 0; c < sw_copies; c++) t += s_sw[i * sw_copies + c];
            if (t != 0.0) atomicAdd(&sum_weight[i], t);
        }
    }
}

// ---------------------------------------------------------------------------
// tiled gridding kernel
// ---------------------------------------------------------------------------
// Byte offsets of the tiled kernel's dynamic LDS block (shared by the
// kernel and the host's size computation).
// Row strides (entries) of the residue-major staging tables.  A wavefront's
// lanes span ~64/R u-residues and all R v-residues; the padding puts the rows
// they read together on different LDS banks: u rows (4-byte entries, read 16
// bytes at a time) start 36 words apart, v rows of 4-byte entries (read 8
// bytes at a time) 33 words apart, v rows of complex entries (read 16 bytes
// at a time) 68 words apart, q rows (NP complex values per entry) 33 entries
// apart.
constexpr int kApSU = kApBatch + 4;
constexpr int kApSV = kApBatch + 1;
constexpr int kApSP = kApBatch + 2;
struct ApLds {
    int fo, pos, base, ui, vi, uc, vc, q, pv, wd, bytes;
};
__host__ __device__ inline int ap_align16(int x) { return (x + 15) & ~15; }
__host__ __device__ inline ApLds ap_lds_layout(int R, int W, int NP, int NW, int acc_bytes,
                                               bool kahan) {
    ApLds L;
    const int nacc = 2 * NP;
    L.fo = ap_align16((kahan ? 2 : 1) * nacc * W * (W + 1) * acc_bytes);
    L.pos = L.fo + kApBatch * 16;
    L.base = L.pos + kApBatch * 16;
    L.ui = ap_align16(L.base + kApBatch * 4);
    L.vi = ap_align16(L.ui + kApSU * R * 4);
    L.uc = ap_align16(L.vi + kApSV * R * 4);
    L.vc = ap_align16(L.uc + kApSU * R * 4);
    L.q = ap_align16(L.vc + kApSV * R * 4);
    L.pv = ap_align16(L.q + kApSV * R * NP * 2 * acc_bytes);
    L.wd = ap_align16(L.pv + kApSP * R * 2 * acc_bytes);
    L.bytes = ap_align16(L.wd + kApBatch * NP * NW * acc_bytes);
    return L;
}

template <typename ACC> struct Cplx {
    ACC x, y;
};

template <int MODE, int NP, typename ACC, bool KAHAN>
__global__ void __launch_bounds__(1024)
ap_tiled_kernel(ApT a, const unsigned long long *__restrict__ offsets,
                const unsigned int *__restrict__ order, const char *__restrict__ entries,
                const ACC *__restrict__ ckr, const Cplx<ACC> *__restrict__ pgu,
                const Cplx<ACC> *__restrict__ pgv,
                ACC *__restrict__ scratch,
                const int *__restrict__ slot, const int *__restrict__ halt) {
    constexpr int NW = (MODE == NGC_GRID_IMAGE) ? 2 : 1;
    constexpr int NACC = NP * 2;  // the grid is complex in every mode
    using E = ApEntry<NP, NW, ACC>;
    constexpr int EB = ApEntrySize<NP, NW, ACC>::bytes;
    using SC = ACC;  // scratch regions in the grid's precision
    const int R = a.Rl, W = a.W;  // lanes per axis and region edge of this class
    const int RT = a.R;            // edge of the tap blocks (max support)
    const int PLANE = W * (W + 1);

    // One dynamic LDS block, carved by integer byte offsets (no pointer round
    // trip through integers, so every access stays an LDS access):
    //   region  [NACC][W][W+1] (+ the same again for the compensation words)
    //   s_fo    [kApBatch] int4   staging: field, offsets, supports
    //   s_pos   [kApBatch] int4   staging: start cells and their residues
    //   s_base  [kApBatch] uint   first tap index of the sample's R x R block
    //   s_ui    [R][kApBatch] uint u-residue a: byte offset of the sample's tap
    //                         block plus ru*R taps; the table's end Z if no tap
    //   s_vi    [R][kApBatch] uint v-residue b: rv taps in bytes; Z if no tap
    //                         (ui + vi is the tap's byte offset; >= Z reads zero)
    //   s_uc    [R][kApBatch] int cell_u*(W+1), or -2^30 if no tap
    //   s_vc    [R][kApBatch] int cell_v, or -2^30 if no tap (uc + vc < 0: no tap)
    //   s_q     [R][kApBatch][NP] u phase factor x weighted visibility per pol
    //   s_pv    [R][kApBatch] v phase factor
    //   s_wd    [kApBatch][NP*NW]
    // The per-residue tables are residue-major and always hold a whole batch
    // (samples past the end of the bin as "no tap"), so that a lane's reads of
    // consecutive samples are one base address plus immediate offsets.
    extern __shared__ __align__(16) char smem_raw[];
    const ApLds L = ap_lds_layout(R, W, NP, NW, (int)sizeof(ACC), KAHAN);
    // FP64 cells for the compensated form: the same 8 bytes as a (sum,
    // compensation) pair of floats, at least its accuracy, and a flush is one
    // conversion and one LDS add instead of a TwoSum
    using CELL = typename std::conditional<KAHAN, double, ACC>::type;
    CELL *s_acc = reinterpret_cast<CELL *>(smem_raw);
    int4 *s_fo = reinterpret_cast<int4 *>(smem_raw + L.fo);
    int4 *s_pos = reinterpret_cast<int4 *>(smem_raw + L.pos);
    unsigned int *s_base = reinterpret_cast<unsigned int *>(smem_raw + L.base);
    unsigned *s_ui = reinterpret_cast<unsigned *>(smem_raw + L.ui);
    unsigned *s_vi = reinterpret_cast<unsigned *>(smem_raw + L.vi);
    int *s_uc = reinterpret_cast<int *>(smem_raw + L.uc);
    int *s_vc = reinterpret_cast<int *>(smem_raw + L.vc);
    // byte size of the tap table up to its end element (the host keeps the
    // table below 2^31 bytes, so ui + vi < 2^32)
    const unsigned zb = (unsigned)(a.n_cf * (2 * a.half_u + 1) * (2 * a.half_v + 1) * RT * RT) *
                        (unsigned)sizeof(ACC);
    // the re-laid taps as a raw buffer of zb bytes: loads at offsets >= zb
    // return zero (kBufferWord3: 32-bit data format, no stride or swizzle)
    const __amdgpu_buffer_rsrc_t ck_rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)ckr, 0, (int)zb, kBufferWord3);
    Cplx<ACC> *s_q = reinterpret_cast<Cplx<ACC> *>(smem_raw + L.q);
    Cplx<ACC> *s_pv = reinterpret_cast<Cplx<ACC> *>(smem_raw + L.pv);
    ACC *s_wd = reinterpret_cast<ACC *>(smem_raw + L.wd);

    __shared__ unsigned s_mv;   // batch samples whose window or support moved

    const int tid = threadIdx.x;
    const int64_t bin = order[blockIdx.x];
    const int64_t start = (int64_t)offsets[bin], end = (int64_t)offsets[bin + 1];
    if (end <= start || (halt && *halt)) return;
    const int64_t tiles = a.ntu * a.ntv;
    const int64_t trem = bin % tiles;
    const int org_u = (int)((trem / a.ntv) * a.T) - a.hl;
    const int org_v = (int)((trem % a.ntv) * a.T) - a.hl;

    for (int i = tid; i < NACC * PLANE; i += blockDim.x) {
        s_acc[i] = CELL{};
    }

    const bool active = tid < R * R;
    const int ra = active ? tid / R : 0, rb = active ? tid % R : 0;
    const int nwv = 2 * a.half_v + 1;
    // this thread's staging items (sample j, residue r < R) of each axis are
    // tid, tid + blockDim, ...: the first divided out once, the rest stepped
    // by blockDim = st_dj R + st_dr
    const int st_j0 = tid / R, st_r0 = tid - st_j0 * R;
    const int st_dj = (int)blockDim.x / R, st_dr = (int)blockDim.x - st_dj * R;
    int cur = -1;  // current cell: u*(W+1) + v in the region, -1 = none
    ACC acc[NACC];
#pragma unroll
    for (int q = 0; q < NACC; q++) acc[q] = (ACC)0;

    auto flush = [&]() {
        if (cur >= 0) {
#pragma unroll
            for (int q = 0; q < NACC; q++) {
                // no two lanes share a cell: the atomic only spares the
                // wavefront the read round trip (the add happens in the LDS)
                __hip_atomic_fetch_add(&s_acc[q * PLANE + cur], (CELL)acc[q], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    };

    // A batch's records are read into registers (threads tid < kApBatch)
    // while the batch before it builds its tables, so that the record load
    // and the tables' phase-gradient loads are in flight together.
    E en{};
    auto fetch = [&](int64_t b0) {
        if (tid < kApBatch && b0 + tid < end)
            en = *reinterpret_cast<const E *>(entries + (b0 + tid) * EB);
    };
    auto stage = [&](int64_t b0) {
        if (tid < kApBatch && b0 + tid < end) {
            const int uc = en.uv & 0xffff, vc = (en.uv >> 16) & 0xffff;
            const int ou = (int)(short)(en.off & 0xffff), ov = (int)(short)(en.off >> 16);
            const int cf = en.meta & 0xffff, fidx = (en.meta >> 16) & 0xffff;
            const int su = a.support[cf * 2], sv = a.support[cf * 2 + 1];
            const int us = uc - su / 2 - org_u, vs = vc - sv / 2 - org_v;
            // (fo.x, fo.y: the sample's phase-gradient factor index at tap
            // offset 0 of its support, so a residue's factor is one 32-bit add)
            s_fo[tid] = make_int4(fidx * (int)a.k_u + ou + (int)(a.k_u / 2) - (int)a.os_u * (su / 2),
                                  fidx * (int)a.k_v + ov + (int)(a.k_v / 2) - (int)a.os_v * (sv / 2),
                                  0, su | (sv << 16));
            s_pos[tid] = make_int4(us, us % R, vs, vs % R);
            s_base[tid] = (unsigned)(((cf * (2 * a.half_u + 1) + ou + a.half_u) * nwv + ov +
                                      a.half_v) * RT * RT);
#pragma unroll
            for (int q = 0; q < NP; q++)
#pragma unroll
                for (int c = 0; c < NW; c++) s_wd[tid * NP * NW + q * NW + c] = en.wd[q][c];
        }
    };
    fetch(start);
    stage(start);

    for (int64_t b0 = start; b0 < end; b0 += kApBatch) {
        const int nh = (int)min((int64_t)kApBatch, end - b0);
        // this batch's staging slots written; the previous batch's tables
        // consumed (and the zeroing done)
        __syncthreads();
        fetch(b0 + kApBatch);
        // samples whose window start or support differs from the previous
        // sample's (and the first of the batch): only there can a lane's cell
        // or tap validity change, so only there do the lanes test their cell
        if (tid < 64) {
            bool m = false;
            if (tid < nh) {
                const int4 p = s_pos[tid], f = s_fo[tid];
                if (tid == 0) {
                    m = true;
                } else {
                    const int4 pp = s_pos[tid - 1], fp = s_fo[tid - 1];
                    m = p.x != pp.x || p.z != pp.z || f.w != fp.w;
                }
            }
            const uint64_t bm = __ballot(m);
            if (tid == 0) s_mv = (unsigned)bm;
        }
        // per-sample, per-residue tables of both axes in one pass (the two
        // phase-gradient factor loads of an item in flight together); samples
        // from nh on are "no tap"
        for (int j = st_j0, r = st_r0; j < kApBatch;
             j += st_dj + (r + st_dr >= R), r += st_dr - (r + st_dr >= R ? R : 0)) {
            const int eu = r * kApSU + j, ev = r * kApSV + j, ep = r * kApSP + j;
            if (j >= nh) {
                s_ui[eu] = zb;
                s_uc[eu] = -(1 << 30);
                s_vi[ev] = zb;
                s_vc[ev] = -(1 << 30);
                s_pv[ep] = Cplx<ACC>{(ACC)0, (ACC)0};
#pragma unroll
                for (int p = 0; p < NP; p++) s_q[ev * NP + p] = Cplx<ACC>{(ACC)0, (ACC)0};
                continue;
            }
            const int4 fo = s_fo[j];
            const int4 ps = s_pos[j];
            const int su = fo.w & 0xffff, sv = fo.w >> 16;
            int ru = r - ps.y;
            ru += ru < 0 ? R : 0;
            int rv = r - ps.w;
            rv += rv < 0 ? R : 0;
            const bool valid_u = ru < su, valid_v = rv < sv;
            Cplx<ACC> pu{(ACC)0, (ACC)0}, pv{(ACC)0, (ACC)0};
            if (valid_u) pu = pgu[fo.x + (int)a.os_u * ru];
            if (valid_v) pv = pgv[fo.y + (int)a.os_v * rv];
            s_ui[eu] = valid_u ? (s_base[j] + (unsigned)(ru * RT)) * (unsigned)sizeof(ACC) : zb;
            s_uc[eu] = valid_u ? (ps.x + ru) * (W + 1) : -(1 << 30);
            s_vi[ev] = valid_v ? (unsigned)rv * (unsigned)sizeof(ACC) : zb;
            s_vc[ev] = valid_v ? ps.z + rv : -(1 << 30);
            s_pv[ep] = pv;
#pragma unroll
            for (int p = 0; p < NP; p++) {
                Cplx<ACC> qv;
                if (NW == 2) {
                    const ACC wr = s_wd[j * NP * NW + p * NW];
                    const ACC wi = s_wd[j * NP * NW + p * NW + 1];
                    qv.x = pu.x * wr - pu.y * wi;
                    qv.y = pu.x * wi + pu.y * wr;
                } else {
                    const ACC w = s_wd[j * NP * NW + p];
                    qv.x = pu.x * w;
                    qv.y = pu.y * w;
                }
                s_q[ev * NP + p] = qv;
            }
        }
        // tables visible; the staging slots consumed, so the next batch's
        // records go in now
        __syncthreads();
        stage(b0 + kApBatch);
        const unsigned moved = __builtin_amdgcn_readfirstlane(s_mv);
        if (active) {
            // Groups of G samples.  A group's tap-table loads (global memory,
            // one per lane and sample) are issued together one group ahead of
            // its arithmetic, branch-free: a tap outside the sample's support
            // (or past the end of the bin) reads zero.  The next
            // group's loads are unconditional (clamped to the last group of
            // the batch) so that the wait for this group's taps counts only
            // the loads issued after them.  A lane's table reads are its
            // residue row plus the sample index: one address per group.
            constexpr int G = kApGroup;
            const int ng = (nh + G - 1) / G;
            const unsigned *ui_r = s_ui + ra * kApSU;
            const unsigned *vi_r = s_vi + rb * kApSV;
            const Cplx<ACC> *pv_r = s_pv + rb * kApSP;
            const Cplx<ACC> *q_r = s_q + ra * kApSV * NP;
            ACC tA[G], tB[G];
            auto load = [&](int j0, ACC (&t)[G]) {
                j0 = min(j0, kApBatch - G);
                unsigned ui[G], vi[G];
#pragma unroll
                for (int g = 0; g < G; g++) {
                    ui[g] = ui_r[j0 + g];
                    vi[g] = vi_r[j0 + g];
                }
#pragma unroll
                for (int g = 0; g < G; g++) {
                    // The tap's byte offset from the table's base: a buffer
                    // load, so that a tap outside the support (offset >= zb:
                    // past the resource's end) reads zero without a clamp.
                    const unsigned off = ui[g] + vi[g];
                    if constexpr (sizeof(ACC) == 4) {
                        t[g] = __builtin_bit_cast(
                            ACC, __builtin_amdgcn_raw_buffer_load_b32(ck_rsrc, off, 0, 0));
                    } else {
                        t[g] = __builtin_bit_cast(
                            ACC, __builtin_amdgcn_raw_buffer_load_b64(ck_rsrc, off, 0, 0));
                    }
                }
            };
            auto run = [&](int j0, const ACC (&t)[G]) {
                Cplx<ACC> pv[G], qv[G][NP];
#pragma unroll
                for (int g = 0; g < G; g++) {
                    pv[g] = pv_r[j0 + g];
#pragma unroll
                    for (int p = 0; p < NP; p++) qv[g][p] = q_r[(j0 + g) * NP + p];
                }
#pragma unroll
                for (int g = 0; g < G; g++) {
                    const int j = j0 + g;
                    if (j >= nh) break;
                    if ((moved >> j) & 1u) {
                        // the lane's cell for this sample (< 0: no tap)
                        const int cell = s_uc[ra * kApSU + j] + s_vc[rb * kApSV + j];
                        if (cell >= 0 && cell != cur) {
                            flush();
                            cur = cell;
#pragma unroll
                            for (int q = 0; q < NACC; q++) acc[q] = (ACC)0;
                        }
                    }
                    // (a tap outside the support is zero: it adds exact zeros
                    // to the current cell)
                    const ACC k = t[g];
                    const ACC tr = k * pv[g].x, ti = k * pv[g].y;
#pragma unroll
                    for (int p = 0; p < NP; p++) {
                        acc[2 * p] = fma(tr, qv[g][p].x, acc[2 * p]);
                        acc[2 * p] = fma(-ti, qv[g][p].y, acc[2 * p]);
                        acc[2 * p + 1] = fma(tr, qv[g][p].y, acc[2 * p + 1]);
                        acc[2 * p + 1] = fma(ti, qv[g][p].x, acc[2 * p + 1]);
                    }
                }
            };
            load(0, tA);
            for (int gi = 0; gi < ng; gi += 2) {
                load((gi + 1) * G, tB);
                run(gi * G, tA);
                if (gi + 1 >= ng) break;
                load((gi + 2) * G, tA);
                run((gi + 1) * G, tB);
            }
        }
    }
    flush();
    __syncthreads();
    // the region goes to this bin's scratch slot: [NACC][W][W]
    SC *out = scratch + (int64_t)slot[bin] * a.slot_elems;
    for (int i = tid; i < NACC * W * W; i += blockDim.x) {
        const int q = i / (W * W), r = i - q * W * W;
        const int lu = r / W, lv = r - lu * W;
        const int o = q * PLANE + lu * (W + 1) + lv;
        out[i] = (SC)s_acc[o];
    }
}

// Tiles whose 3 x 3 neighbourhood holds a non-empty bin (of any support
// class) get the regions that cover their cells added (class by class; in
// each, home tile first, then the neighbours in fixed order; summed in FP64);
// every other cell keeps its value (zero unless accumulating).
// nonempty[bin]: 1 when the bin holds records (bins of every class);
// slot[bin]: its scratch slot; near[tile]: 1 when the tile is merged (tiles
// of one class's index range); tiles[]: the listed tiles, *n_tiles of them.
__global__ void ap_flags_kernel(ApT a, ApCls k, const unsigned long long *offsets, int64_t n_bins,
                                int *nonempty, int *near) {
    const int64_t n = n_bins > a.per_cls ? n_bins : a.per_cls;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b <= n;
         b += (int64_t)gridDim.x * blockDim.x) {
        // the scans' sentinels are zero
        if (b <= n_bins) nonempty[b] = b < n_bins && offsets[b + 1] > offsets[b] ? 1 : 0;
        if (b > a.per_cls) continue;
        if (b == a.per_cls) {
            near[b] = 0;
            continue;
        }
        const int64_t tiles = a.ntu * a.ntv;
        const int64_t ch = b / tiles, r = b - ch * tiles;
        const int64_t tu = r / a.ntv, tv = r - tu * a.ntv;
        int any = 0;
        for (int c = 0; c < k.n; c++)
            for (int du = -1; du <= 1; du++)
                for (int dv = -1; dv <= 1; dv++) {
                    const int64_t u = tu + du, v = tv + dv;
                    if (u < 0 || u >= a.ntu || v < 0 || v >= a.ntv) continue;
                    const int64_t nb = c * a.per_cls + (ch * a.ntu + u) * a.ntv + v;
                    any |= offsets[nb + 1] > offsets[nb];
                }
        near[b] = any;
    }
}

__global__ void ap_list_kernel(int64_t n_tiles, const int *near, const int *near_pos, int *tiles) {
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < n_tiles;
         b += (int64_t)gridDim.x * blockDim.x)
        if (near[b]) tiles[near_pos[b]] = (int)b;
}

// One workgroup per listed tile (grid-stride): the slots of the tile's 3 x 3
// neighbourhood in every support class are looked up once, then each thread
// sums its cells.
template <typename SC, typename OUT>
__global__ void __launch_bounds__(256)
ap_merge_kernel(ApT a, ApCls k, const int *__restrict__ nonempty, const int *__restrict__ slot,
                const int *__restrict__ tiles, const int *__restrict__ n_tiles_p,
                const SC *__restrict__ scratch, OUT *grid, const int *__restrict__ halt) {
    if (halt && *halt) return;
    __shared__ int s_slot[kApMaxCls * 9];  // [class][du + 1][dv + 1]: slot, or -1
    const int T = a.T;
    const int n_tiles = *n_tiles_p;
    const int64_t tiles_per_chan = a.ntu * a.ntv;
    const int cells = (int)a.n_imag_pol * T * T;
    for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const int64_t home = tiles[t];
        const int64_t ch = home / tiles_per_chan, hr = home - ch * tiles_per_chan;
        const int htu = (int)(hr / a.ntv), htv = (int)(hr - (int64_t)htu * a.ntv);
        __syncthreads();  // the previous tile's slots consumed
        for (int i = threadIdx.x; i < k.n * 9; i += blockDim.x) {
            const int cl = i / 9, d = i - cl * 9;
            const int tu = htu + d / 3 - 1, tv = htv + d % 3 - 1;
            int sl = -1;
            if (tu >= 0 && tu < a.ntu && tv >= 0 && tv < a.ntv) {
                const int64_t bin = cl * a.per_cls + (ch * a.ntu + tu) * a.ntv + tv;
                if (nonempty[bin]) sl = slot[bin];
            }
            s_slot[i] = sl;
        }
        __syncthreads();
        for (int c = threadIdx.x; c < cells; c += blockDim.x) {
            const int pol = c / (T * T), r = c - pol * T * T;
            const int lu0 = r / T, lv0 = r - lu0 * T;
            const int64_t u = (int64_t)htu * T + lu0, v = (int64_t)htv * T + lv0;
            if (u >= a.n_u || v >= a.n_v) continue;
            double re = 0.0, im = 0.0;
            // class by class; home tile first, then the neighbours
            for (int cl = 0; cl < k.n; cl++) {
                const int W = k.W[cl], hl = k.hl[cl];
                const int64_t plane = (int64_t)W * W;
                for (int di = 0; di < 3; di++) {
                    const int du = di == 0 ? 0 : (di == 1 ? -1 : 1);
                    const int lu = lu0 + hl - du * T;  // u - ((htu + du) T - hl)
                    if (lu < 0 || lu >= W) continue;
                    for (int dj = 0; dj < 3; dj++) {
                        const int dv = dj == 0 ? 0 : (dj == 1 ? -1 : 1);
                        const int lv = lv0 + hl - dv * T;
                        if (lv < 0 || lv >= W) continue;
                        const int sl = s_slot[cl * 9 + (du + 1) * 3 + dv + 1];
                        if (sl < 0) continue;
                        const SC *sp = scratch + (int64_t)sl * a.slot_elems + lu * W + lv;
                        re += (double)sp[(2 * pol) * plane];
                        im += (double)sp[(2 * pol + 1) * plane];
                    }
                }
            }
            const int64_t gi = ((ch * a.n_imag_pol + pol) * a.n_u + u) * a.n_v + v;
            grid[2 * gi] = (OUT)((double)grid[2 * gi] + re);
            grid[2 * gi + 1] = (OUT)((double)grid[2 * gi + 1] + im);
        }
    }
}

// ---------------------------------------------------------------------------
// weight mode
// ---------------------------------------------------------------------------
// W[key] with key = ((a_chan*n_imag_pol + p)*n_cf + cf)*n_field + f
__global__ void __launch_bounds__(256)
ap_wsum_kernel(ApT a, const double2 *ntab, double *wkey, double *sum_weight, int keys_in_lds,
               int64_t n_keys) {
#pragma clang fp contract(off)
    extern __shared__ double s_w[];
    if (keys_in_lds) {
        for (int64_t i = threadIdx.x; i < n_keys; i += blockDim.x) s_w[i] = 0.0;
        __syncthreads();
    }
    const int nwu = 2 * a.half_u + 1, nwv = 2 * a.half_v + 1;
    const int64_t n_rc = a.n_time * a.n_baseline * a.n_chan;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_rc;
         i += (int64_t)gridDim.x * blockDim.x) {
        ApSample sm;
        sm.ch = i % a.n_chan;
        sm.row = i / a.n_chan;
        sm.b = sm.row % a.n_baseline;
        sm.t = sm.row / a.n_baseline;
        sm.rc = i;
        int fidx = 0;
        Pix px{}, pxv{};
        if (!ap_place(a, sm, fidx, px, pxv)) continue;
        const int64_t cf = ((int64_t)a.cfb[sm.b] * a.n_cfc + a.cfc[sm.ch]) * a.n_cfp + a.cfp0;
        const int64_t a_chan = a.chan_map[sm.ch];
        // centre taps: the zero offset row of the norm table
        const double2 nm = ntab[((cf * a.n_field + fidx) * nwu + a.half_u) * nwv + a.half_v];
        for (int p = 0; p < a.n_pol; p++) {
            const double w = apt_weight(a, i * a.n_pol + p);
            if (isnan(w) || w == 0.0) continue;
            const int64_t key = ((a_chan * a.n_imag_pol + p) * a.n_cf + cf) * a.n_field + fidx;
            if (keys_in_lds)
                atomicAdd(&s_w[key], w);
            else
                atomicAdd(&wkey[key], w);
            atomicAdd(&sum_weight[a_chan * a.n_imag_pol + p], w * nm.x);
        }
    }
    if (keys_in_lds) {
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n_keys; i += blockDim.x)
            if (s_w[i] != 0.0) atomicAdd(&wkey[i], s_w[i]);
    }
}

// one thread per (a_chan, pol, centre tap): sum_key W * CK*PG at the tap
template <typename OUT>
__global__ void ap_wstamp_kernel(ApT a, const double *ck, const double2 *pg, const double *wkey,
                                 OUT *grid, int accumulate) {
#pragma clang fp contract(off)
    const int R = a.R;
    const int64_t n = a.n_imag_chan * a.n_imag_pol * R * R;
    const int64_t kplane = a.k_u * a.k_v;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int rv = (int)(i % R), ru = (int)((i / R) % R);
        const int64_t cp = i / ((int64_t)R * R);  // a_chan * n_imag_pol + pol
        const int64_t iu = ru - R / 2, iv = rv - R / 2;
        double re = 0.0, im = 0.0;
        bool touched = false;
        for (int64_t cf = 0; cf < a.n_cf; cf++) {
            const int64_t su = a.support[cf * 2], sv = a.support[cf * 2 + 1];
            if (iu < -(su / 2) || iu >= su - su / 2 || iv < -(sv / 2) || iv >= sv - sv / 2)
                continue;
            const int64_t k = (a.os_u * iu + a.k_u / 2) * a.k_v + a.os_v * iv + a.k_v / 2;
            for (int64_t f = 0; f < a.n_field; f++) {
                const double wk = wkey[(cp * a.n_cf + cf) * a.n_field + f];
                if (wk == 0.0) continue;
                const double kv = ck[cf * kplane + k];
                const double2 g = pg[f * kplane + k];
                re += (kv * g.x) * wk;
                im += (kv * g.y) * wk;
                touched = true;
            }
        }
        if (!touched) continue;
        const int64_t cell = (cp * a.n_u + a.n_u / 2 + iu) * a.n_v + a.n_v / 2 + iv;
        if (accumulate) {
            grid[2 * cell] = (OUT)((double)grid[2 * cell] + re);
            grid[2 * cell + 1] = (OUT)((double)grid[2 * cell + 1] + im);
        } else {
            grid[2 * cell] = (OUT)re;
            grid[2 * cell + 1] = (OUT)im;
        }
    }
}

__global__ void ap_zero_kernel(double *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

template <typename T> __global__ void ap_zero_grid_kernel(T *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (T)0;
}

// Largest-first tile order without a sort (no library call, so plans can
// capture it): tiles go into size classes of 1/8 octave, classes are laid out
// in decreasing order, and each tile takes the next place of its class.
constexpr int kApOrderClasses = 512;

// 0 for an empty tile, else 1 + 8 floor(log2 n) + the three bits below n's
// leading one (non-decreasing in n; a class spans at most 1/8 of its sizes)
__device__ inline int ap_size_class(unsigned long long n) {
    if (n == 0) return 0;
    const int l = 63 - __clzll((long long)n);
    const unsigned long long m = l >= 3 ? (n >> (l - 3)) & 7ull : (n << (3 - l)) & 7ull;
    const int c = 1 + 8 * l + (int)m;
    return c < kApOrderClasses ? c : kApOrderClasses - 1;
}

// hist[k][c] += bins of support class k (blockIdx.y) in size class c (hist
// zeroed beforehand)
__global__ void __launch_bounds__(256)
ap_order_hist_kernel(const unsigned long long *offsets, int64_t per_cls, unsigned int *hist) {
    __shared__ unsigned int s_n[kApOrderClasses];
    const int64_t b0 = (int64_t)blockIdx.y * per_cls;
    for (int k = threadIdx.x; k < kApOrderClasses; k += blockDim.x) s_n[k] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < per_cls;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&s_n[ap_size_class(offsets[b0 + i + 1] - offsets[b0 + i])], 1u);
    __syncthreads();
    for (int k = threadIdx.x; k < kApOrderClasses; k += blockDim.x)
        if (s_n[k]) atomicAdd(&hist[blockIdx.y * kApOrderClasses + k], s_n[k]);
}

// cursor[k][c] = first place of size class c in support class k's segment
// (k per_cls onwards), the largest size class first
__global__ void ap_order_start_kernel(const unsigned int *hist, unsigned int *cursor, int n_cls,
                                      int64_t per_cls) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int k = 0; k < n_cls; k++) {
            unsigned int pos = (unsigned int)(k * per_cls);
            for (int c = kApOrderClasses - 1; c >= 0; c--) {
                cursor[k * kApOrderClasses + c] = pos;
                pos += hist[k * kApOrderClasses + c];
            }
        }
    }
}

__global__ void __launch_bounds__(256)
ap_order_place_kernel(const unsigned long long *offsets, int64_t per_cls, unsigned int *cursor,
                      unsigned int *order) {
    __shared__ unsigned int s_n[kApOrderClasses];
    __shared__ unsigned int s_base[kApOrderClasses];
    const int64_t b0 = (int64_t)blockIdx.y * per_cls;
    unsigned int *cur = cursor + blockIdx.y * kApOrderClasses;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < per_cls;
         i0 += (int64_t)gridDim.x * blockDim.x) {
        for (int k = threadIdx.x; k < kApOrderClasses; k += blockDim.x) s_n[k] = 0;
        __syncthreads();
        const int64_t i = i0 + threadIdx.x;
        int c = 0;
        unsigned int rank = 0;
        if (i < per_cls) {
            c = ap_size_class(offsets[b0 + i + 1] - offsets[b0 + i]);
            rank = atomicAdd(&s_n[c], 1u);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < kApOrderClasses; k += blockDim.x)
            if (s_n[k]) s_base[k] = atomicAdd(&cur[k], s_n[k]);
        __syncthreads();
        if (i < per_cls) order[s_base[c] + rank] = (unsigned int)(b0 + i);
        __syncthreads();
    }
}

template <typename T> __global__ void ap_copy_kernel(T *dst, const T *src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <typename T> __global__ void ap_fill_kernel(T *p, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// A plan's buffers against the current inputs: *halt = 1 when they bin more
// records, or fill more tiles, than the plan holds (the fill, tiled and merge
// kernels then write nothing).
__global__ void ap_cap_kernel(const unsigned long long *offsets, const int *slot, int64_t n_bins,
                              unsigned long long ent_cap, int64_t slot_cap, int *halt) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *halt = (offsets[n_bins] > ent_cap || (int64_t)slot[n_bins] > slot_cap) ? 1 : 0;
}

static int grid_blocks(int64_t n, int bs = 256) {
    int64_t b = ceil_div(n, bs);
    if (b > (1 << 20)) b = 1 << 20;
    if (b < 1) b = 1;
    return (int)b;
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
// Sizes of one call.
struct ApGeom {
    bool single = false, kahan = false;
    int acc_bytes = 8, nwu = 0, nwv = 0;
    int64_t n_sw = 0, n_rc = 0, cells = 0, n_ntab = 0;
    // image and PSF modes
    int T = 0;
    size_t lds = 0;                  // the largest class's workgroup LDS
    size_t lds_cls[kApMaxCls] = {};
    ApCls cls{};
    std::vector<int> cf_cls;         // support class of each convolution function
    int64_t n_bins = 0, n_ckr = 0, n_pg = 0, n_waves = 0, bin_blocks = 1;
};

// Buffers of the tiled pipeline: the context's workspace for one call, owned
// allocations for a plan.
struct ApBufs {
    double2 *ntab = nullptr;
    void *ckr = nullptr, *pgu = nullptr, *pgv = nullptr;
    unsigned long long *counts = nullptr, *off = nullptr;
    void *scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    int *ne = nullptr, *near = nullptr, *slot = nullptr, *npos = nullptr, *tiles = nullptr;
    int *cf_cls = nullptr;
    void *itmp = nullptr;
    size_t itmp_bytes = 0;
    unsigned int *ord_hist = nullptr, *ord_cursor = nullptr, *order = nullptr;
    char *ent = nullptr;
    unsigned long long ent_cap = 0;  // records the entry buffer holds
    void *scr = nullptr;
    int64_t slot_cap = 0;            // scratch slots
    int *halt = nullptr;             // a plan's capacity flag (nullptr for a call)
};

using ApAlloc = std::function<int(const char *, size_t, void **)>;

// Validates the call, fills the kernel arguments and the sizes; the tile edge
// is T_req (16, 24 or 32) or, for 0, the one whose workgroup LDS lets the most
// workgroups share a CU (the kernel is latency-bound), the smaller on a tie
// (more tiles to balance over the CUs; measured faster than the larger one's
// fewer halo cells on the 25-pointing mosaic).
static int ap_setup(const ApTiledCall &c, int T_req, ApT &a, ApGeom &g) {
    NGC_CHECK(c.pol_identity, NGC_ERR_UNSUPPORTED,
              "tiled aperture gridder needs pol_map = arange(n_pol)");
    NGC_CHECK(c.cfp_uniform, NGC_ERR_UNSUPPORTED,
              "tiled aperture gridder needs one convolution-function pol for all pols");
    NGC_CHECK(c.n_pol >= 1 && c.n_pol <= 4 && c.n_pol != 3, NGC_ERR_UNSUPPORTED,
              "tiled aperture gridder supports 1, 2 or 4 pols");
    NGC_CHECK(c.max_support >= 1 && c.max_support <= 32, NGC_ERR_UNSUPPORTED,
              "tiled aperture gridder supports max(SUPPORT) <= 32");
    NGC_CHECK(c.n_u < 32768 && c.n_v < 32768, NGC_ERR_UNSUPPORTED, "grid axis >= 32768");
    NGC_CHECK(c.n_cf < 65536 && c.n_field < 65536, NGC_ERR_UNSUPPORTED,
              "too many convolution functions or fields");
    NGC_CHECK(c.os_u < 16384 && c.os_v < 16384, NGC_ERR_UNSUPPORTED, "oversampling too large");
    NGC_CHECK(T_req == 0 || T_req == 16 || T_req == 24 || T_req == 32, NGC_ERR_ARG,
              "tile edge must be 16, 24 or 32");

    a.uvw = c.uvw;
    a.vis = c.vis;
    a.weight = c.weight;
    a.flag = c.flag;
    a.field = c.field;
    a.field_id = c.d_field_id;
    a.field_map = c.d_field_map;
    a.field_map_len = c.field_map_len;
    a.uv_scale = c.d_uv_scale;
    a.chan_map = c.d_chan_map;
    a.cfb = c.d_cfb;
    a.cfc = c.d_cfc;
    a.support = c.d_support;
    a.n_field = c.n_field;
    a.n_cfc = c.n_cfc;
    a.n_cfp = c.n_cfp;
    a.cfp0 = c.cfp0;
    a.n_cf = c.n_cf;
    a.n_time = c.n_time;
    a.n_baseline = c.n_baseline;
    a.n_chan = c.n_chan;
    a.n_pol = c.n_pol;
    a.n_imag_chan = c.n_imag_chan;
    a.n_imag_pol = c.n_imag_pol;
    a.n_u = c.n_u;
    a.n_v = c.n_v;
    a.os_u = c.os_u;
    a.os_v = c.os_v;
    a.k_u = c.k_u;
    a.k_v = c.k_v;
    a.R = (int)c.max_support;
    a.hl = a.R / 2;
    a.half_u = (int)(c.os_u / 2 + 1);
    a.half_v = (int)(c.os_v / 2 + 1);
    a.vis_dtype = c.vis_dtype;
    a.weight_dtype = c.weight_dtype;
    int ch_log2 = 0;
    while ((1 << (ch_log2 + 1)) <= std::min<int64_t>(c.n_chan, 2)) ch_log2++;
    a.ch_log2 = ch_log2;
    a.n_cb = ceil_div(c.n_chan, int64_t(1) << ch_log2);
    a.n_tb = ceil_div(c.n_time, int64_t(64) >> ch_log2);
    NGC_CHECK(c.n_baseline * a.n_tb * a.n_cb < (int64_t(1) << 32), NGC_ERR_UNSUPPORTED,
              "too many (baseline, time block, channel block) slots for one call");
    a.div_cb = FastDiv::of((uint32_t)a.n_cb);
    a.div_tb = FastDiv::of((uint32_t)a.n_tb);

    g.single = c.precision != NGC_PREC_FP64_TILED;
    g.kahan = c.precision == NGC_PREC_FP32_KAHAN;
    g.acc_bytes = g.single ? 4 : 8;
    g.n_sw = c.n_imag_chan * c.n_imag_pol;
    g.n_rc = c.n_time * c.n_baseline * c.n_chan;
    g.cells = g.n_sw * c.n_u * c.n_v;
    g.nwu = 2 * a.half_u + 1;
    g.nwv = 2 * a.half_v + 1;
    g.n_ntab = c.n_cf * c.n_field * g.nwu * g.nwv;
    if (c.mode == NGC_GRID_WEIGHT) return NGC_OK;

    // support classes: convolution functions whose largest support needs
    // the same number of lane wavefronts share one, gridded with R the
    // class's largest support
    {
        int cls_of_waves[kApMaxCls + 1];
        int r_of_waves[kApMaxCls + 1] = {};
        for (int64_t f = 0; f < c.n_cf; f++) {
            const int r = std::max(c.support_host[2 * f], c.support_host[2 * f + 1]);
            const int wv = (r * r + 63) / 64;
            r_of_waves[wv] = std::max(r_of_waves[wv], r);
        }
        g.cls.n = 0;
        for (int wv = 1; wv <= kApMaxCls; wv++) {
            cls_of_waves[wv] = -1;
            if (r_of_waves[wv] > 0) {
                cls_of_waves[wv] = g.cls.n;
                g.cls.R[g.cls.n++] = r_of_waves[wv];
            }
        }
        g.cf_cls.resize((size_t)c.n_cf);
        for (int64_t f = 0; f < c.n_cf; f++) {
            const int r = std::max(c.support_host[2 * f], c.support_host[2 * f + 1]);
            g.cf_cls[(size_t)f] = cls_of_waves[(r * r + 63) / 64];
        }
    }
    const int nw = c.mode == NGC_GRID_IMAGE ? 2 : 1;
    auto lds_for = [&](int T) {
        return (size_t)ap_lds_layout(a.R, T + a.R - 1, (int)c.n_pol, nw, g.acc_bytes, g.kahan)
            .bytes;
    };
    int T = T_req;
    if (T == 0) {
        size_t best_per_cu = 0;
        for (int cand : {16, 24, 32}) {
            const size_t per_cu = (160 * 1024) / lds_for(cand);
            if (per_cu > best_per_cu) {
                best_per_cu = per_cu;
                T = cand;
            }
        }
    }
    g.T = T;
    g.lds = lds_for(T);
    NGC_CHECK(g.lds <= 160 * 1024, NGC_ERR_UNSUPPORTED, "aperture tile does not fit in LDS");
    a.T = T;
    a.W = T + a.R - 1;
    a.Rl = a.R;
    a.ntu = ceil_div(c.n_u, T);
    a.ntv = ceil_div(c.n_v, T);
    a.per_cls = c.n_imag_chan * a.ntu * a.ntv;
    a.slot_elems = 2 * c.n_pol * (int64_t)a.W * a.W;
    for (int k = 0; k < g.cls.n; k++) {
        g.cls.W[k] = T + g.cls.R[k] - 1;
        g.cls.hl[k] = g.cls.R[k] / 2;
        g.lds_cls[k] = (size_t)ap_lds_layout(g.cls.R[k], g.cls.W[k], (int)c.n_pol, nw,
                                             g.acc_bytes, g.kahan).bytes;
    }
    g.n_bins = g.cls.n * a.per_cls;
    NGC_CHECK(g.n_bins < (int64_t(1) << 31), NGC_ERR_UNSUPPORTED, "too many tiles");
    g.n_ckr = c.n_cf * g.nwu * g.nwv * a.R * a.R;
    // the tiled kernel addresses the re-laid taps with 32-bit byte offsets
    NGC_CHECK((g.n_ckr + 1) * g.acc_bytes < (int64_t(1) << 31), NGC_ERR_UNSUPPORTED,
              "re-laid convolution kernels exceed 2 GiB");
    // the tiled kernel indexes the factor tables with 32-bit integers
    NGC_CHECK(c.n_field * c.k_u < (int64_t(1) << 31) && c.n_field * c.k_v < (int64_t(1) << 31),
              NGC_ERR_UNSUPPORTED, "phase-gradient factor tables exceed 2^31 entries");
    g.n_pg = c.n_field * c.k_u * c.k_v;
    g.n_waves = c.n_baseline * a.n_tb * a.n_cb;
    g.bin_blocks = std::max<int64_t>(
        1, std::min<int64_t>(ceil_div(g.n_waves, kApBinBlock / 64), 1 << 20));
    return NGC_OK;
}

// Every buffer of the image/PSF pipeline except the entries and the scratch
// slots (sized from the inputs' counts).
static int ap_alloc(const ApTiledCall &c, const ApGeom &g, const ApAlloc &alloc, ApBufs &b) {
    void *q;
    const int64_t nb1 = g.n_bins + 1;
    NGC_TRY(alloc("ap_ntab", g.n_ntab * sizeof(double2), &q));
    b.ntab = (double2 *)q;
    NGC_TRY(alloc("ap_ckr", (g.n_ckr + 1) * g.acc_bytes, &b.ckr));
    NGC_TRY(alloc("ap_pgu", c.n_field * c.k_u * 2 * g.acc_bytes, &b.pgu));
    NGC_TRY(alloc("ap_pgv", c.n_field * c.k_v * 2 * g.acc_bytes, &b.pgv));
    NGC_TRY(alloc("ap_bin_counts", nb1 * sizeof(unsigned long long), &q));
    b.counts = (unsigned long long *)q;
    NGC_TRY(alloc("ap_bin_offsets", nb1 * sizeof(unsigned long long), &q));
    b.off = (unsigned long long *)q;
    NGC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b.scan_tmp_bytes, b.counts, b.off, nb1));
    NGC_TRY(alloc("ap_scan_tmp", b.scan_tmp_bytes, &b.scan_tmp));
    int **ints[] = {&b.ne, &b.near, &b.slot, &b.npos, &b.tiles};
    const char *names[] = {"ap_nonempty", "ap_near", "ap_slot", "ap_near_pos", "ap_tiles"};
    for (int i = 0; i < 5; i++) {
        NGC_TRY(alloc(names[i], nb1 * sizeof(int), &q));
        *ints[i] = (int *)q;
    }
    NGC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b.itmp_bytes, (const int *)b.ne, b.slot,
                                             nb1));
    NGC_TRY(alloc("ap_int_scan_tmp", b.itmp_bytes, &b.itmp));
    NGC_TRY(alloc("ap_order_hist", 2 * g.cls.n * kApOrderClasses * sizeof(unsigned int), &q));
    b.ord_hist = (unsigned int *)q;
    b.ord_cursor = b.ord_hist + g.cls.n * kApOrderClasses;
    NGC_TRY(alloc("ap_cf_cls", c.n_cf * sizeof(int), &q));
    b.cf_cls = (int *)q;
    NGC_TRY(alloc("ap_order", g.n_bins * sizeof(unsigned int) + 4, &q));
    b.order = (unsigned int *)q;
    return NGC_OK;
}

// The re-laid taps and the separable phase-gradient factors; *bad is set when
// the phase gradient does not factor.
static int ap_enqueue_tables(hipStream_t s, const ApTiledCall &c, const ApT &a, const ApGeom &g,
                             const ApBufs &b, int *bad) {
    if (g.single) {
        ap_taps_kernel<float><<<grid_blocks(g.n_ckr + 1), 256, 0, s>>>(a, c.conv_kernel,
                                                                   (float *)b.ckr);
        ap_pg_kernel<float2><<<grid_blocks(g.n_pg), 256, 0, s>>>(
            a, (const double2 *)c.phase_gradient, (float2 *)b.pgu, (float2 *)b.pgv, bad);
    } else {
        ap_taps_kernel<double><<<grid_blocks(g.n_ckr + 1), 256, 0, s>>>(a, c.conv_kernel,
                                                                    (double *)b.ckr);
        ap_pg_kernel<double2><<<grid_blocks(g.n_pg), 256, 0, s>>>(
            a, (const double2 *)c.phase_gradient, (double2 *)b.pgu, (double2 *)b.pgv, bad);
    }
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

// Count pass, bin offsets, the non-empty tiles' scratch slots and the list of
// tiles the merge covers.  No host synchronisation (captured by plans).
template <int MODE, int NP, typename ACC>
static int ap_enqueue_count(ngc_context *ctx, hipStream_t s, const ApTiledCall &c, const ApT &a,
                            const ApGeom &g, const ApBufs &b) {
    const int64_t nb1 = g.n_bins + 1;
    ap_fill_kernel<unsigned long long><<<grid_blocks(nb1), 256, 0, s>>>(b.counts, nb1, 0ull);
    NGC_HIP(hipGetLastError());
    phase_begin(ctx, s, "ap_bin_count");
    if (g.n_waves > 0) {
        ap_bin_kernel<false, MODE, NP, ACC><<<(int)g.bin_blocks, kApBinBlock, 0, s>>>(
            a, b.counts, nullptr, b.ntab, c.sum_weight, 0, nullptr);
        NGC_HIP(hipGetLastError());
    }
    phase_end(ctx, s);
    phase_begin(ctx, s, "ap_bin_scan");
    size_t tb = b.scan_tmp_bytes;
    NGC_HIP(hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, tb, b.counts, b.off, nb1, s));
    phase_end(ctx, s);
    ap_flags_kernel<<<grid_blocks(nb1), 256, 0, s>>>(a, g.cls, b.off, g.n_bins, b.ne, b.near);
    NGC_HIP(hipGetLastError());
    size_t it = b.itmp_bytes;
    NGC_HIP(hipcub::DeviceScan::ExclusiveSum(b.itmp, it, (const int *)b.ne, b.slot, nb1, s));
    it = b.itmp_bytes;
    NGC_HIP(hipcub::DeviceScan::ExclusiveSum(b.itmp, it, (const int *)b.near, b.npos,
                                             a.per_cls + 1, s));
    ap_list_kernel<<<grid_blocks(a.per_cls), 256, 0, s>>>(a.per_cls, (const int *)b.near,
                                                          (const int *)b.npos, b.tiles);
    NGC_HIP(hipGetLastError());
    return NGC_OK;
}

// Fill pass, largest-first tile order (the longest-processing-time-first
// order: workgroups are dispatched roughly in block order, so the heaviest
// tiles start first and the tail is short), tiled gridding and merge.  No
// host synchronisation; set_attr = false while a plan is being captured (the
// LDS attribute is set before the capture).
template <int MODE, int NP, typename ACC, bool KAHAN>
static int ap_enqueue_fill(ngc_context *ctx, hipStream_t s, const ApTiledCall &c, const ApT &a,
                           const ApGeom &g, const ApBufs &b, bool set_attr) {
    using SC = ACC;  // scratch regions in the grid's precision
    const int64_t nb1 = g.n_bins + 1;
    const int sw_copies = g.n_sw * 8 <= 4096 ? 8 : (g.n_sw <= 4096 ? 1 : 0);
    const size_t sw_shmem = g.n_sw * sw_copies * sizeof(double);
    ap_copy_kernel<unsigned long long><<<grid_blocks(nb1), 256, 0, s>>>(b.counts, b.off, nb1);
    NGC_HIP(hipGetLastError());
    phase_begin(ctx, s, "ap_bin_fill");
    if (g.n_waves > 0) {
        ap_bin_kernel<true, MODE, NP, ACC><<<(int)g.bin_blocks, kApBinBlock, sw_shmem, s>>>(
            a, b.counts, b.ent, b.ntab, c.sum_weight, sw_copies, b.halt);
        NGC_HIP(hipGetLastError());
    }
    phase_end(ctx, s);
    phase_begin(ctx, s, "ap_bin_order");
    ap_fill_kernel<unsigned int><<<grid_blocks(g.cls.n * kApOrderClasses), 256, 0, s>>>(
        b.ord_hist, g.cls.n * kApOrderClasses, 0u);
    const dim3 og((unsigned)std::min<int64_t>(grid_blocks(a.per_cls), 4096), (unsigned)g.cls.n);
    ap_order_hist_kernel<<<og, 256, 0, s>>>(b.off, a.per_cls, b.ord_hist);
    ap_order_start_kernel<<<1, 64, 0, s>>>(b.ord_hist, b.ord_cursor, g.cls.n, a.per_cls);
    ap_order_place_kernel<<<og, 256, 0, s>>>(b.off, a.per_cls, b.ord_cursor, b.order);
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    phase_begin(ctx, s, "ap_grid_tiled");
    auto kern = ap_tiled_kernel<MODE, NP, ACC, KAHAN>;
    if (set_attr)
        NGC_HIP(hipFuncSetAttribute((const void *)kern,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)g.lds));
    // one launch per support class, each with its own lanes and region
    for (int k = 0; k < g.cls.n; k++) {
        ApT ak = a;
        ak.Rl = g.cls.R[k];
        ak.W = g.cls.W[k];
        ak.hl = g.cls.hl[k];
        const int threads = (ak.Rl * ak.Rl + 63) / 64 * 64;
        kern<<<(unsigned)a.per_cls, threads, g.lds_cls[k], s>>>(
            ak, b.off, b.order + k * a.per_cls, (const char *)b.ent, (const ACC *)b.ckr,
            (const Cplx<ACC> *)b.pgu, (const Cplx<ACC> *)b.pgv, (SC *)b.scr, (const int *)b.slot,
            b.halt);
        NGC_HIP(hipGetLastError());
    }
    phase_end(ctx, s);
    phase_begin(ctx, s, "ap_merge");
    const int merge_blocks = (int)std::min<int64_t>(a.per_cls, 8192);
    const int *ntp = (const int *)b.npos + a.per_cls;  // number of listed tiles
    if (g.single) {
        if (!c.accumulate)
            ap_zero_grid_kernel<float><<<grid_blocks(2 * g.cells), 256, 0, s>>>(
                (float *)c.grid, 2 * g.cells);
        ap_merge_kernel<SC, float><<<merge_blocks, 256, 0, s>>>(
            a, g.cls, (const int *)b.ne, (const int *)b.slot, (const int *)b.tiles, ntp,
            (const SC *)b.scr, (float *)c.grid, b.halt);
    } else {
        if (!c.accumulate)
            ap_zero_grid_kernel<double><<<grid_blocks(2 * g.cells), 256, 0, s>>>(
                (double *)c.grid, 2 * g.cells);
        ap_merge_kernel<SC, double><<<merge_blocks, 256, 0, s>>>(
            a, g.cls, (const int *)b.ne, (const int *)b.slot, (const int *)b.tiles, ntp,
            (const SC *)b.scr, (double *)c.grid, b.halt);
    }
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}

template <int NP, int NW, typename ACC> constexpr int ap_entry_bytes() {
    return ApEntrySize<NP, NW, ACC>::bytes;
}

// f(mode, pols, accumulator type, compensated) with compile-time arguments
template <typename F> static int ap_dispatch(const ApTiledCall &c, F &&f) {
    using I0 = std::integral_constant<int, NGC_GRID_IMAGE>;
    using I1 = std::integral_constant<int, NGC_GRID_PSF>;
    struct F32 {
        using type = float;
    };
    struct F64 {
        using type = double;
    };
    using KT = std::true_type;
    using KF = std::false_type;
    auto by_prec = [&](auto mode_c, auto np_c) -> int {
        if (c.precision == NGC_PREC_FP32_KAHAN) return f(mode_c, np_c, F32{}, KT{});
        if (c.precision == NGC_PREC_FP32) return f(mode_c, np_c, F32{}, KF{});
        return f(mode_c, np_c, F64{}, KF{});
    };
    auto by_pol = [&](auto mode_c) -> int {
        switch (c.n_pol) {
        case 1: return by_prec(mode_c, std::integral_constant<int, 1>{});
        case 2: return by_prec(mode_c, std::integral_constant<int, 2>{});
        default: return by_prec(mode_c, std::integral_constant<int, 4>{});
        }
    };
    return c.mode == NGC_GRID_IMAGE ? by_pol(I0{}) : by_pol(I1{});
}

static int ap_weight_run(ngc_context *ctx, hipStream_t s, const ApTiledCall &c, const ApT &a,
                         const ApGeom &g, const double2 *ntab) {
    const int64_t n_keys = g.n_sw * c.n_cf * c.n_field;
    void *d_w;
    NGC_TRY(workspace(ctx, "ap_wkey", n_keys * sizeof(double), &d_w));
    ap_zero_kernel<<<grid_blocks(n_keys), 256, 0, s>>>((double *)d_w, n_keys);
    NGC_HIP(hipGetLastError());
    const int in_lds = n_keys <= 4096 ? 1 : 0;
    phase_begin(ctx, s, "ap_weight_sum");
    if (g.n_rc > 0) {
        ap_wsum_kernel<<<grid_blocks(g.n_rc, 256) > 4096 ? 4096 : grid_blocks(g.n_rc, 256), 256,
                         in_lds ? n_keys * sizeof(double) : 0, s>>>(a, ntab, (double *)d_w,
                                                                   c.sum_weight, in_lds, n_keys);
        NGC_HIP(hipGetLastError());
    }
    phase_end(ctx, s);
    phase_begin(ctx, s, "ap_weight_stamp");
    if (g.single) {
        if (!c.accumulate) {
            ap_zero_grid_kernel<float><<<grid_blocks(2 * g.cells), 256, 0, s>>>(
                (float *)c.grid, 2 * g.cells);
            NGC_HIP(hipGetLastError());
        }
        ap_wstamp_kernel<float><<<grid_blocks(g.n_sw * a.R * a.R), 256, 0, s>>>(
            a, c.conv_kernel, (const double2 *)c.phase_gradient, (const double *)d_w,
            (float *)c.grid, c.accumulate);
    } else {
        if (!c.accumulate) {
            ap_zero_grid_kernel<double><<<grid_blocks(2 * g.cells), 256, 0, s>>>(
                (double *)c.grid, 2 * g.cells);
            NGC_HIP(hipGetLastError());
        }
        ap_wstamp_kernel<double><<<grid_blocks(g.n_sw * a.R * a.R), 256, 0, s>>>(
            a, c.conv_kernel, (const double2 *)c.phase_gradient, (const double *)d_w,
            (double *)c.grid, c.accumulate);
    }
    NGC_HIP(hipGetLastError());
    phase_end(ctx, s);
    return NGC_OK;
}

// One eager call with tile edge T_req (0: by occupancy).
static int ap_tiled_run_t(ngc_context *ctx, hipStream_t s, const ApTiledCall &c, int T_req) {
    ApT a{};
    ApGeom g{};
    NGC_TRY(ap_setup(c, T_req, a, g));
    a.error = ctx->d_error;
    const ApAlloc ws = [ctx](const char *name, size_t bytes, void **q) {
        return workspace(ctx, name, bytes, q);
    };
    if (!c.accumulate) {
        ap_zero_kernel<<<grid_blocks(g.n_sw), 256, 0, s>>>(c.sum_weight, g.n_sw);
        NGC_HIP(hipGetLastError());
    }
    ApBufs b{};
    if (c.mode == NGC_GRID_WEIGHT) {
        void *q;
        NGC_TRY(ws("ap_ntab", g.n_ntab * sizeof(double2), &q));
        b.ntab = (double2 *)q;
    } else {
        NGC_TRY(ap_alloc(c, g, ws, b));
        NGC_HIP(hipMemcpyAsync(b.cf_cls, g.cf_cls.data(), c.n_cf * sizeof(int),
                               hipMemcpyHostToDevice, s));
        a.cf_cls = b.cf_cls;
    }
    // FP64 tap-sum table (sum_weight) for every mode
    phase_begin(ctx, s, "ap_tables");
    ap_norm_kernel<<<grid_blocks(g.n_ntab), 256, 0, s>>>(a, c.conv_kernel,
                                                         (const double2 *)c.phase_gradient,
                                                         b.ntab);
    NGC_HIP(hipGetLastError());
    if (c.mode == NGC_GRID_WEIGHT) {
        phase_end(ctx, s);
        return ap_weight_run(ctx, s, c, a, g, b.ntab);
    }
    NGC_HIP(hipMemsetAsync(ctx->d_error + 1, 0, sizeof(int), s));
    NGC_TRY(ap_enqueue_tables(s, c, a, g, b, ctx->d_error + 1));
    phase_end(ctx, s);
    int bad = 0;
    NGC_HIP(hipMemcpyAsync(&bad, ctx->d_error + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    NGC_HIP(hipStreamSynchronize(s));
    NGC_CHECK(bad == 0, NGC_ERR_UNSUPPORTED,
              "tiled aperture gridder needs a phase gradient of the form exp(i(x px + y py))");

    return ap_dispatch(c, [&](auto mode_c, auto np_c, auto acc_c, auto kahan_c) -> int {
        constexpr int MODE = decltype(mode_c)::value;
        constexpr int NP = decltype(np_c)::value;
        using ACC = typename decltype(acc_c)::type;
        constexpr bool KAHAN = decltype(kahan_c)::value;
        constexpr int NW = (MODE == NGC_GRID_IMAGE) ? 2 : 1;
        using SC = ACC;  // scratch regions in the grid's precision
        NGC_TRY((ap_enqueue_count<MODE, NP, ACC>(ctx, s, c, a, g, b)));
        unsigned long long total = 0;
        int n_nonempty = 0;
        NGC_HIP(hipMemcpyAsync(&total, b.off + g.n_bins, sizeof(total), hipMemcpyDeviceToHost,
                               s));
        NGC_HIP(hipMemcpyAsync(&n_nonempty, b.slot + g.n_bins, sizeof(int),
                               hipMemcpyDeviceToHost, s));
        NGC_HIP(hipStreamSynchronize(s));
        void *q;
        NGC_TRY(ws("ap_bin_entries", (size_t)(total + 1) * ap_entry_bytes<NP, NW, ACC>(), &q));
        b.ent = (char *)q;
        b.ent_cap = total + 1;
        NGC_TRY(ws("ap_scratch", (size_t)(n_nonempty + 1) * a.slot_elems * sizeof(SC), &b.scr));
        b.slot_cap = n_nonempty + 1;
        return ap_enqueue_fill<MODE, NP, ACC, KAHAN>(ctx, s, c, a, g, b, true);
    });
}

int ap_tiled_run(ngc_context *ctx, hipStream_t s, const ApTiledCall &c) {
    return ap_tiled_run_t(ctx, s, c, 0);
}

static std::string ap_shape_key(const ApTiledCall &c) {
    char buf[320];
    snprintf(buf, sizeof(buf), "ap/%d/%d/%lld/%lld/%lld/%lld/%lld/%lld/%lld/%lld/%lld/%lld/%lld",
             c.mode, c.precision, (long long)c.n_time, (long long)c.n_baseline,
             (long long)c.n_chan, (long long)c.n_pol, (long long)c.n_imag_chan,
             (long long)c.n_u, (long long)c.n_v, (long long)c.max_support, (long long)c.os_u,
             (long long)c.os_v, (long long)c.n_cf);
    return buf;
}

int ap_plan_create(ngc_context *ctx, hipStream_t s, const ApTiledCall &c, int autotune,
                   ngc_grid_plan *p) {
    NGC_CHECK(c.mode == NGC_GRID_IMAGE || c.mode == NGC_GRID_PSF, NGC_ERR_UNSUPPORTED,
              "gridding plans cover image and PSF modes");
    NGC_CHECK(c.precision != NGC_PREC_FP64, NGC_ERR_UNSUPPORTED,
              "gridding plans run the tiled forms (FP64_TILED, FP32_KAHAN, FP32)");
    {
        ApT a0{};
        ApGeom g0{};
        NGC_TRY(ap_setup(c, 0, a0, g0));
    }
    const bool saved_profile = ctx->profile;
    ctx->profile = false;
    struct Restore {
        ngc_context *c;
        bool v;
        ~Restore() { c->profile = v; }
    } restore{ctx, saved_profile};

    // tile edge: timed on these inputs (the whole eager pipeline, one warm-up
    // and one timed run per candidate) or taken from the context's cache
    const std::string key = ap_shape_key(c);
    int T = 0;
    std::string tune_log;
    auto it = ctx->tuned.find(key);
    if (autotune) {
        hipEvent_t e0, e1;
        NGC_HIP(hipEventCreate(&e0));
        NGC_HIP(hipEventCreate(&e1));
        float best_ms = 1e30f;
        for (int cand : {32, 24, 16}) {
            ApT at{};
            ApGeom gt{};
            if (ap_setup(c, cand, at, gt) != NGC_OK) continue;  // e.g. the tile exceeds LDS
            float ms = 0.f;
            for (int rep = 0; rep < 2; rep++) {
                NGC_HIP(hipEventRecord(e0, s));
                const int rc = ap_tiled_run_t(ctx, s, c, cand);
                if (rc != NGC_OK) {
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    return rc;
                }
                NGC_HIP(hipEventRecord(e1, s));
                NGC_HIP(hipEventSynchronize(e1));
                NGC_HIP(hipEventElapsedTime(&ms, e0, e1));
            }
            char bb[48];
            snprintf(bb, sizeof(bb), "%s%d:%.3f", tune_log.empty() ? "" : " ", cand, ms);
            tune_log += bb;
            if (ms < best_ms) {
                best_ms = ms;
                T = cand;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ctx->tuned[key] = {T, 0};
    } else if (it != ctx->tuned.end()) {
        T = it->second.first;
    }

    ApT a{};
    ApGeom g{};
    NGC_TRY(ap_setup(c, T, a, g));
    void *q;
    // flags: [1] capacity exceeded, [3] a row's field id not in field_id,
    // [4] the phase gradient does not factor
    NGC_TRY(plan_alloc(p, 8 * sizeof(int), &q));
    p->flags = (int *)q;
    ap_fill_kernel<int><<<1, 64, 0, s>>>(p->flags, 8, 0);
    NGC_HIP(hipGetLastError());
    a.error = p->flags + 3;
    const ApAlloc own = [p](const char *, size_t bytes, void **out) {
        return plan_alloc(p, bytes, out);
    };
    ApBufs b{};
    NGC_TRY(ap_alloc(c, g, own, b));
    NGC_HIP(hipMemcpyAsync(b.cf_cls, g.cf_cls.data(), c.n_cf * sizeof(int),
                           hipMemcpyHostToDevice, s));
    a.cf_cls = b.cf_cls;
    // the kernels' tables are built once, here: the plan reads conv_kernel and
    // phase_gradient at creation only
    ap_norm_kernel<<<grid_blocks(g.n_ntab), 256, 0, s>>>(a, c.conv_kernel,
                                                         (const double2 *)c.phase_gradient,
                                                         b.ntab);
    NGC_HIP(hipGetLastError());
    NGC_TRY(ap_enqueue_tables(s, c, a, g, b, p->flags + 4));
    int bad = 0;
    NGC_HIP(hipMemcpyAsync(&bad, p->flags + 4, sizeof(int), hipMemcpyDeviceToHost, s));
    NGC_HIP(hipStreamSynchronize(s));
    NGC_CHECK(bad == 0, NGC_ERR_UNSUPPORTED,
              "tiled aperture gridder needs a phase gradient of the form exp(i(x px + y py))");

    return ap_dispatch(c, [&](auto mode_c, auto np_c, auto acc_c, auto kahan_c) -> int {
        constexpr int MODE = decltype(mode_c)::value;
        constexpr int NP = decltype(np_c)::value;
        using ACC = typename decltype(acc_c)::type;
        constexpr bool KAHAN = decltype(kahan_c)::value;
        constexpr int NW = (MODE == NGC_GRID_IMAGE) ? 2 : 1;
        using SC = ACC;  // scratch regions in the grid's precision
        // buffers sized by one eager count pass, with room for the records and
        // the non-empty tiles to grow by 1/16 (inputs rewritten in place)
        NGC_TRY((ap_enqueue_count<MODE, NP, ACC>(ctx, s, c, a, g, b)));
        unsigned long long total = 0;
        int n_nonempty = 0;
        NGC_HIP(hipMemcpyAsync(&total, b.off + g.n_bins, sizeof(total), hipMemcpyDeviceToHost,
                               s));
        NGC_HIP(hipMemcpyAsync(&n_nonempty, b.slot + g.n_bins, sizeof(int),
                               hipMemcpyDeviceToHost, s));
        NGC_HIP(hipStreamSynchronize(s));
        b.ent_cap = total + total / 16 + 1024;
        void *e;
        NGC_TRY(plan_alloc(p, (size_t)b.ent_cap * ap_entry_bytes<NP, NW, ACC>(), &e));
        b.ent = (char *)e;
        b.slot_cap = std::min<int64_t>(g.n_bins, n_nonempty + n_nonempty / 16 + 64);
        NGC_TRY(plan_alloc(p, (size_t)b.slot_cap * a.slot_elems * sizeof(SC), &b.scr));
        b.halt = p->flags + 1;
        auto kern = ap_tiled_kernel<MODE, NP, ACC, KAHAN>;
        NGC_HIP(hipFuncSetAttribute((const void *)kern,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)g.lds));
        NGC_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
        NGC_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
        hipStream_t cs = p->cap_stream;
        if (!c.accumulate)
            ap_fill_kernel<double><<<grid_blocks(g.n_sw), 256, 0, cs>>>(c.sum_weight, g.n_sw,
                                                                        0.0);
        ap_fill_kernel<int><<<1, 64, 0, cs>>>(p->flags + 3, 1, 0);
        int r = ap_enqueue_count<MODE, NP, ACC>(ctx, cs, c, a, g, b);
        if (r == NGC_OK) {
            ap_cap_kernel<<<1, 64, 0, cs>>>(b.off, b.slot, g.n_bins, b.ent_cap, b.slot_cap,
                                            b.halt);
            r = ap_enqueue_fill<MODE, NP, ACC, KAHAN>(ctx, cs, c, a, g, b, false);
        }
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(cs, &graph);
        p->graph = graph;
        if (r != NGC_OK) return r;
        NGC_CHECK(ec == hipSuccess && graph, NGC_ERR_HIP, "hipStreamEndCapture failed");
        NGC_HIP(hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0));
        char info[320];
        snprintf(info, sizeof(info),
                 "aperture: tile %dx%d, %lld bins, %llu records (cap %llu), "
                 "%d non-empty tiles (cap %lld); lanes per support class:",
                 g.T, g.T, (long long)g.n_bins, total, b.ent_cap, n_nonempty,
                 (long long)b.slot_cap);
        p->info = info;
        for (int k = 0; k < g.cls.n; k++)
            p->info += " " + std::to_string(g.cls.R[k]) + "x" + std::to_string(g.cls.R[k]);
        if (!tune_log.empty()) p->info += "; tuning ms " + tune_log;
        return NGC_OK;
    });
}

}  // namespace ngc
