"""ctypes binding of libngcasa_hip.so, the gfx950 gridding library.

Every entry point of ``include/ngcasa_gridding.h`` is wrapped here.  Large
arrays travel as torch CUDA (HIP) tensors; their device pointers and the
current torch stream are handed to the library.  There is no fallback: if the
library is missing or the device is unavailable, the call raises.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libngcasa_hip.so")

# constants mirrored from include/ngcasa_gridding.h
NGC_OK = 0
ERRORS = {-1: "NGC_ERR_ARG", -2: "NGC_ERR_HIP", -3: "NGC_ERR_UNSUPPORTED",
          -4: "NGC_ERR_FIELD", -5: "NGC_ERR_FFT"}
GRID_IMAGE, GRID_PSF, GRID_DENSITY, GRID_WEIGHT = 0, 1, 2, 3
PREC_FP64, PREC_FP64_TILED, PREC_FP32_KAHAN, PREC_FP32 = 0, 1, 2, 3
PRECISIONS = {"fp64": PREC_FP64, "fp64_tiled": PREC_FP64_TILED, "fp32_kahan": PREC_FP32_KAHAN,
              "fp32": PREC_FP32}
C64, C128, F32, F64, I32, I64, BOOL = 0, 1, 2, 3, 4, 5, 6
WEIGHT_UNIFORM, WEIGHT_BRIGGS = 1, 2

EXPORTS = ("ngc_context_create", "ngc_context_destroy", "ngc_last_error", "ngc_version",
           "ngc_standard_grid", "ngc_briggs_factors", "ngc_imaging_weight_degrid",
           "ngc_aperture_grid", "ngc_grid_to_image", "ngc_normalize_mosaic",
           "ngc_profile_enable", "ngc_profile_read", "ngc_set_tile_rows",
           "ngc_gcf_patterns", "ngc_gcf_resize", "ngc_phase_gradient",
           "ngc_grid_plan_create", "ngc_aperture_grid_plan_create", "ngc_grid_plan_execute",
           "ngc_grid_plan_status",
           "ngc_grid_plan_describe", "ngc_grid_plan_destroy", "ngc_bf16_gemm",
           "ngc_tanh_backward", "ngc_col_sum", "ngc_mse_grad", "ngc_adam_step",
           "ngc_direction_rotate", "ngc_make_pb", "ngc_apply_flags", "ngc_set_tile_waves",
           "ngc_model_to_grid", "ngc_standard_degrid", "ngc_aperture_degrid", "ngc_mlp_sample_rows",
           "ngc_adam_schedule", "ngc_adam_step_sched", "ngc_set_bin_lanes",
           "ngc_standard_grid_noalloc", "ngc_set_fft_mode", "ngc_set_degrid_mode")


class NativeError(RuntimeError):
    """A libngcasa_hip call returned an error status."""

    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with {ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None
_lib_lock = threading.Lock()

_vp, _i64, _i32, _d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double

_PROTOS = {
    "ngc_context_create": (_i32, [_i32, ctypes.POINTER(_vp)]),
    "ngc_context_destroy": (_i32, [_vp]),
    "ngc_last_error": (ctypes.c_char_p, []),
    "ngc_version": (ctypes.c_char_p, []),
    "ngc_standard_grid": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _vp,
                                 _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64,
                                 _i64, _i64, _d, _d, _i64, _i64, _i32, _vp, _vp]),
    "ngc_standard_grid_noalloc": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _i32, _vp, _i32, _vp,
                                         _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64,
                                         _i64, _i64, _i64, _i64, _d, _d, _i64, _i64, _i32, _vp,
                                         _vp]),
    "ngc_briggs_factors": (_i32, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i32, _d, _vp]),
    "ngc_imaging_weight_degrid": (_i32, [_vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp,
                                         _vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _i64,
                                         _d, _d, _vp, _i32]),
    "ngc_aperture_grid": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _vp,
                                 _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64,
                                 _i64, _i64, _i64, _vp, _vp, _i64, _i64, _i64, _i64, _i64,
                                 _i64, _i64, _i64, _d, _d, _i64, _i64, _i32, _vp, _vp]),
    "ngc_aperture_grid_plan_create": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _i32, _vp, _i32,
                                             _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                             _vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp, _i64,
                                             _i64, _i64, _i64, _i64, _i64, _i64, _i64, _d, _d,
                                             _i64, _i64, _i32, _vp, _vp, _i32,
                                             ctypes.POINTER(_vp)]),
    "ngc_grid_to_image": (_i32, [_vp, _vp, _vp, _i32, _i64, _i64, _i64, _i64, _i64, _vp, _vp,
                                 _vp, _vp]),
    "ngc_profile_enable": (_i32, [_vp, _i32]),
    "ngc_set_tile_rows": (_i32, [_vp, _i32]),
    "ngc_set_tile_waves": (_i32, [_vp, _i32]),
    "ngc_set_bin_lanes": (_i32, [_vp, _i32]),
    "ngc_set_fft_mode": (_i32, [_vp, _i32]),
    "ngc_set_degrid_mode": (_i32, [_vp, _i32]),
    "ngc_gcf_patterns": (_i32, [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _i64, _i32, _i32, _d,
                                _d, _i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "ngc_gcf_resize": (_i32, [_vp, _vp, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _i64,
                              _vp]),
    "ngc_phase_gradient": (_i32, [_vp, _vp, _vp, _i64, _i64, _i64, _vp]),
    "ngc_apply_flags": (_i32, [_vp, _vp, _vp, _i32, _vp, _i64, _vp]),
    "ngc_model_to_grid": (_i32, [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp, _vp,
                                 _i32]),
    "ngc_standard_degrid": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64,
                                   _i64, _i64, _i64, _i64, _i64, _d, _d, _i64, _i64, _vp, _i32,
                                   _vp, _i32, _i32]),
    "ngc_aperture_degrid": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp,
                                    _vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp, _i64, _i64,
                                    _i64, _i64, _i64, _i64, _i64, _i64, _d, _d, _i64, _i64,
                                    _vp, _i32, _vp, _i32, _i32]),
    "ngc_make_pb": (_i32, [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _i32, _i32, _i64, _i64, _i64,
                           _i64, _d, _d, _i64, _vp]),
    "ngc_profile_read": (_i32, [_vp, ctypes.c_char_p, _i64, _vp, _i64]),
    "ngc_normalize_mosaic": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64,
                                    _i64, _d, _i32]),
    "ngc_grid_plan_create": (_i32, [_vp, _vp, _i32, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _vp,
                                    _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64,
                                    _i64, _i64, _d, _d, _i64, _i64, _i32, _vp, _vp, _i32,
                                    ctypes.POINTER(_vp)]),
    "ngc_grid_plan_execute": (_i32, [_vp, _vp]),
    "ngc_grid_plan_status": (_i32, [_vp, _vp]),
    "ngc_grid_plan_describe": (_i32, [_vp, ctypes.c_char_p, _i64]),
    "ngc_grid_plan_destroy": (_i32, [_vp]),
    "ngc_bf16_gemm": (_i32, [_vp, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp,
                             _i32, _i32, _vp, _i64, _vp, _vp]),
    "ngc_tanh_backward": (_i32, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "ngc_col_sum": (_i32, [_vp, _vp, _i64, _i64, _vp, _vp]),
    "ngc_mse_grad": (_i32, [_vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "ngc_adam_step": (_i32, [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _i64]),
    "ngc_mlp_sample_rows": (_i32, [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, ctypes.c_uint64,
                                   _vp, _vp, _vp]),
    "ngc_adam_schedule": (_i32, [_vp, _vp, _vp, _vp, _d, _d, _i64, _i64, _d, _d]),
    "ngc_adam_step_sched": (_i32, [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _d, _d, _d]),
    "ngc_direction_rotate": (_i32, [_vp, _vp, _vp, _vp, _i32, _vp, _vp, _i64, _vp, _vp, _vp,
                                    _i64, _i64, _i64, _i64, _i32, _i32, _vp, _vp]),
}


def load_library(path=LIB_PATH):
    """Load libngcasa_hip.so and declare its prototypes.  Raises if missing."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f"libngcasa_hip.so not found at {path}; build it with "
                "`python -m cngi_prototype_amd.build` (there is no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def _check(fn, rc):
    if rc != NGC_OK:
        msg = load_library().ngc_last_error().decode(errors="replace")
        raise NativeError(fn, rc, msg)


class Context:
    """Owns one ngc_context (device workspace, rocFFT plans) on one device."""

    def __init__(self, device=0):
        lib = load_library()
        h = _vp()
        _check("ngc_context_create", lib.ngc_context_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = int(device)

    def __del__(self):
        try:
            if self.handle:
                load_library().ngc_context_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


_contexts = {}


def context(device=None):
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible; the MI355X gridding path needs a GPU")
    if device is None:
        device = torch.cuda.current_device()
    if device not in _contexts:
        _contexts[device] = Context(device)
    return _contexts[device]


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------
def _stream():
    import torch
    return _vp(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return None if t is None else _vp(t.data_ptr())


def _host(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(_vp)


def _dt_vis(t):
    import torch
    if t.dtype == torch.complex64:
        return C64
    if t.dtype == torch.complex128:
        return C128
    raise TypeError(f"visibilities must be complex64 or complex128, got {t.dtype}")


def _dt_w(t):
    import torch
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.float64:
        return F64
    raise TypeError(f"weights must be float32 or float64, got {t.dtype}")


def _dt_grid(t):
    import torch
    return {torch.complex64: C64, torch.complex128: C128, torch.float32: F32,
            torch.float64: F64}[t.dtype]


def _dev(t, name):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a device (HIP) tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


def _uvw_rows(uvw, rows=None):
    """(T, B) of a float64 [T, B, 3] uvw tensor; the kernels read uvw[row*3]
    and uvw[row*3 + 1] directly, so any other dtype or shape is refused."""
    import torch
    if uvw.dtype != torch.float64:
        raise TypeError(f"uvw must be float64, got {uvw.dtype}")
    if uvw.dim() != 3 or uvw.shape[2] != 3:
        raise ValueError(f"uvw must be [n_time, n_baseline, 3], got {tuple(uvw.shape)}")
    T, B = int(uvw.shape[0]), int(uvw.shape[1])
    if rows is not None and (T, B) != tuple(int(x) for x in rows):
        raise ValueError(f"uvw is [{T}, {B}, 3] but the data have {tuple(rows)} rows")
    return T, B


def _field_rows(field, T, B):
    """FIELD_ID must be int32 [T, B], one id per row, as the kernels index it."""
    import torch
    if field.dtype != torch.int32:
        raise TypeError("field must be int32")
    if tuple(field.shape) != (T, B):
        raise ValueError(f"field must be [n_time, n_baseline] = [{T}, {B}], "
                         f"got {tuple(field.shape)}")


# ---------------------------------------------------------------------------
# entry points
# ---------------------------------------------------------------------------
def standard_grid(uvw, weight, freq_chan, chan_map, pol_map, cgk_1d, n_imag_chan, n_uv,
                  delta_lm, support, oversampling, mode, precision, vis=None, flag=None,
                  grid=None, sum_weight=None, accumulate=False, n_imag_pol=None,
                  noalloc=False):
    """ngc_standard_grid on device tensors.  Returns (grid, sum_weight).  With
    noalloc the call goes to ngc_standard_grid_noalloc, which allocates no
    device memory (pass grid and sum_weight to keep the whole call
    allocation-free)."""
    import torch
    ctx = context()
    _dev(uvw, "uvw")
    _dev(weight, "weight")
    n_time, n_baseline, n_chan, n_pol = weight.shape
    _uvw_rows(uvw, (n_time, n_baseline))
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    n_imag_pol = n_pol if n_imag_pol is None else int(n_imag_pol)
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    single = prec in (PREC_FP32_KAHAN, PREC_FP32)
    if mode == GRID_IMAGE:
        gdt = torch.complex64 if single else torch.complex128
        _dev(vis, "vis")
        if tuple(vis.shape) != tuple(weight.shape):
            raise ValueError("vis shape does not match weight")
    else:
        gdt = torch.float32 if single else torch.float64
    shape = (int(n_imag_chan), n_imag_pol, n_u, n_v)
    if grid is None:
        grid = torch.empty(shape, dtype=gdt, device=weight.device)
    elif tuple(grid.shape) != shape or grid.dtype != gdt:
        raise ValueError(f"grid must be {shape} {gdt}")
    if sum_weight is None:
        sum_weight = torch.empty(shape[:2], dtype=torch.float64, device=weight.device)
    if flag is not None:
        _dev(flag, "flag")
        if flag.dtype != torch.uint8 or tuple(flag.shape) != tuple(weight.shape):
            raise ValueError("flag must be uint8 shaped like weight")
    fr, frp = _host(freq_chan, np.float64)
    cm, cmp_ = _host(chan_map, np.int32)
    pm, pmp = _host(pol_map, np.int32)
    cg, cgp = _host(cgk_1d, np.float64)
    lib = load_library()
    name = "ngc_standard_grid_noalloc" if noalloc else "ngc_standard_grid"
    rc = getattr(lib, name)(
        ctx.handle, _stream(), int(mode), prec, _ptr(uvw), _ptr(vis),
        _dt_vis(vis) if vis is not None else C64, _ptr(weight), _dt_w(weight), _ptr(flag),
        frp, cmp_, pmp, cgp, len(cg), n_time, n_baseline, n_chan, n_pol, int(n_imag_chan),
        n_imag_pol, n_u, n_v, float(delta_lm[0]), float(delta_lm[1]), int(support),
        int(oversampling), int(bool(accumulate)), _ptr(grid), _ptr(sum_weight))
    _check(name, rc)
    return grid, sum_weight


class _Plan:
    """A captured gridding pipeline (ngc_grid_plan_*): execute() replays it
    with one graph launch on the current stream."""

    handle = None

    def execute(self):
        _check("ngc_grid_plan_execute",
               load_library().ngc_grid_plan_execute(self.handle, _stream()))
        return self.grid, self.sum_weight

    def status(self):
        rc = load_library().ngc_grid_plan_status(self.handle, _stream())
        if rc == -4:
            raise IndexError("a row's FIELD_ID is not in the gcf's field_id list")
        _check("ngc_grid_plan_status", rc)

    def describe(self):
        buf = ctypes.create_string_buffer(4096)
        _check("ngc_grid_plan_describe",
               load_library().ngc_grid_plan_describe(self.handle, buf, 4096))
        return buf.value.decode()

    def close(self):
        if self.handle is not None:
            load_library().ngc_grid_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GridPlan(_Plan):
    """A captured, autotuned ngc_standard_grid call (ngc_grid_plan_*): the
    tensors given at creation are the plan's inputs and outputs; execute()
    replays the whole tiled pipeline with one graph launch."""

    def __init__(self, uvw, weight, freq_chan, chan_map, pol_map, cgk_1d, n_imag_chan, n_uv,
                 delta_lm, support, oversampling, mode, precision, vis=None, flag=None,
                 grid=None, sum_weight=None, autotune=True):
        import torch
        self._ctx = context()
        self.handle = None
        n_time, n_baseline, n_chan, n_pol = weight.shape
        _uvw_rows(_dev(uvw, "uvw"), (n_time, n_baseline))
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        single = prec in (PREC_FP32_KAHAN, PREC_FP32)
        n_u, n_v = int(n_uv[0]), int(n_uv[1])
        if mode == GRID_IMAGE:
            gdt = torch.complex64 if single else torch.complex128
        else:
            gdt = torch.float32 if single else torch.float64
        shape = (int(n_imag_chan), n_pol, n_u, n_v)
        self.grid = torch.empty(shape, dtype=gdt, device=weight.device) if grid is None else grid
        self.sum_weight = (torch.empty(shape[:2], dtype=torch.float64, device=weight.device)
                           if sum_weight is None else sum_weight)
        # the plan holds device pointers: keep the tensors alive with it
        self._keep = (uvw, weight, vis, flag)
        fr, frp = _host(freq_chan, np.float64)
        cm, cmp_ = _host(chan_map, np.int32)
        pm, pmp = _host(pol_map, np.int32)
        cg, cgp = _host(cgk_1d, np.float64)
        h = _vp()
        rc = load_library().ngc_grid_plan_create(
            self._ctx.handle, _stream(), int(mode), prec, _ptr(uvw), _ptr(vis),
            _dt_vis(vis) if vis is not None else C64, _ptr(weight), _dt_w(weight), _ptr(flag),
            frp, cmp_, pmp, cgp, len(cg), n_time, n_baseline, n_chan, n_pol, int(n_imag_chan),
            n_pol, n_u, n_v, float(delta_lm[0]), float(delta_lm[1]), int(support),
            int(oversampling), 0, _ptr(self.grid), _ptr(self.sum_weight), int(bool(autotune)),
            ctypes.byref(h))
        _check("ngc_grid_plan_create", rc)
        self.handle = h


def briggs_factors(density_grid, sum_weight, weighting, robust=0.5):
    """ngc_briggs_factors.  Returns host float64 [2, n_imag_chan, n_imag_pol]."""
    ctx = context()
    _dev(density_grid, "density_grid")
    _dev(sum_weight, "sum_weight")
    nc, npol, nu, nv = density_grid.shape
    out = np.zeros((2, nc, npol), dtype=np.float64)
    w = {"uniform": WEIGHT_UNIFORM, "briggs": WEIGHT_BRIGGS}[weighting]
    rc = load_library().ngc_briggs_factors(ctx.handle, _stream(), _ptr(density_grid),
                                           _ptr(sum_weight), nc, npol, nu, nv, w,
                                           float(robust), out.ctypes.data_as(_vp))
    _check("ngc_briggs_factors", rc)
    return out


def imaging_weight_degrid(uvw, natural_weight, density_grid, factors, freq_chan, chan_map,
                          pol_map, delta_lm, flag=None, out=None):
    """ngc_imaging_weight_degrid.  Returns the imaging-weight tensor."""
    import torch
    ctx = context()
    _dev(uvw, "uvw")
    _dev(natural_weight, "natural_weight")
    _dev(density_grid, "density_grid")
    n_time, n_baseline, n_chan, n_pol = natural_weight.shape
    _uvw_rows(uvw, (n_time, n_baseline))
    nc, npol, nu, nv = density_grid.shape
    if out is None:
        out = torch.empty_like(natural_weight)
    fac, facp = _host(factors, np.float64)
    fr, frp = _host(freq_chan, np.float64)
    cm, cmp_ = _host(chan_map, np.int32)
    pm, pmp = _host(pol_map, np.int32)
    rc = load_library().ngc_imaging_weight_degrid(
        ctx.handle, _stream(), _ptr(uvw), _ptr(natural_weight), _dt_w(natural_weight),
        _ptr(flag), frp, cmp_, pmp, _ptr(density_grid), facp, n_time, n_baseline, n_chan,
        n_pol, nc, npol, nu, nv, float(delta_lm[0]), float(delta_lm[1]), _ptr(out),
        _dt_w(out))
    _check("ngc_imaging_weight_degrid", rc)
    return out


def _aperture_args(uvw, weight, field, field_id, freq_chan, chan_map, pol_map,
                   cf_baseline_map, cf_chan_map, cf_pol_map, conv_kernel, support,
                   phase_gradient, n_imag_chan, n_uv, delta_lm, oversampling, mode, precision,
                   vis, flag, grid, sum_weight, accumulate):
    """Checked ctypes arguments of ngc_aperture_grid after (ctx, stream), the
    output tensors and the host arrays the pointers refer to."""
    import torch
    for t, n in ((uvw, "uvw"), (weight, "weight"), (field, "field"),
                 (conv_kernel, "conv_kernel"), (phase_gradient, "phase_gradient")):
        _dev(t, n)
    n_time, n_baseline, n_chan, n_pol = weight.shape
    _uvw_rows(uvw, (n_time, n_baseline))
    _field_rows(field, n_time, n_baseline)
    if conv_kernel.dtype != torch.float64 or phase_gradient.dtype != torch.complex128:
        raise TypeError("conv_kernel must be float64 and phase_gradient complex128")
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    n_cfb, n_cfc, n_cfp, k_u, k_v = conv_kernel.shape
    if tuple(phase_gradient.shape[1:]) != (k_u, k_v):
        raise ValueError("phase_gradient must be [n_field, K_u, K_v]")
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    shape = (int(n_imag_chan), n_pol, n_u, n_v)
    gdt = torch.complex64 if prec in (PREC_FP32_KAHAN, PREC_FP32) else torch.complex128
    if grid is None:
        grid = torch.empty(shape, dtype=gdt, device=weight.device)
    if sum_weight is None:
        sum_weight = torch.empty(shape[:2], dtype=torch.float64, device=weight.device)
    fid, fidp = _host(field_id, np.int32)
    fr, frp = _host(freq_chan, np.float64)
    cm, cmp_ = _host(chan_map, np.int32)
    pm, pmp = _host(pol_map, np.int32)
    b, bp = _host(cf_baseline_map, np.int32)
    c, cp = _host(cf_chan_map, np.int32)
    p, pp = _host(cf_pol_map, np.int32)
    s, sp = _host(support, np.int32)
    args = (int(mode), prec, _ptr(uvw), _ptr(vis), _dt_vis(vis) if vis is not None else C64,
            _ptr(weight), _dt_w(weight), _ptr(flag), _ptr(field), fidp, len(fid), frp, cmp_, pmp,
            bp, cp, pp, _ptr(conv_kernel), n_cfb, n_cfc, n_cfp, k_u, k_v, sp,
            _ptr(phase_gradient), n_time, n_baseline, n_chan, n_pol, int(n_imag_chan), n_pol, n_u,
            n_v, float(delta_lm[0]), float(delta_lm[1]), int(oversampling[0]),
            int(oversampling[1]), int(bool(accumulate)), _ptr(grid), _ptr(sum_weight))
    return args, grid, sum_weight, (fid, fr, cm, pm, b, c, p, s)


def aperture_grid(uvw, weight, field, field_id, freq_chan, chan_map, pol_map,
                  cf_baseline_map, cf_chan_map, cf_pol_map, conv_kernel, support,
                  phase_gradient, n_imag_chan, n_uv, delta_lm, oversampling, mode,
                  precision="fp64", vis=None, flag=None, grid=None, sum_weight=None,
                  accumulate=False):
    """ngc_aperture_grid on device tensors.  Returns (complex grid, sum_weight)."""
    ctx = context()
    args, grid, sum_weight, _host_arrays = _aperture_args(
        uvw, weight, field, field_id, freq_chan, chan_map, pol_map, cf_baseline_map,
        cf_chan_map, cf_pol_map, conv_kernel, support, phase_gradient, n_imag_chan, n_uv,
        delta_lm, oversampling, mode, precision, vis, flag, grid, sum_weight, accumulate)
    rc = load_library().ngc_aperture_grid(ctx.handle, _stream(), *args)
    if rc == -4:
        raise IndexError("a row's FIELD_ID is not in the gcf's field_id list")
    _check("ngc_aperture_grid", rc)
    return grid, sum_weight


class ApertureGridPlan(_Plan):
    """A captured ngc_aperture_grid call (ngc_aperture_grid_plan_create): every
    pointing, channel and pol of a mosaic gridded by one graph launch.  The
    tensors given at creation are the plan's inputs and outputs (uvw, vis,
    weight, flag and field may be rewritten in place between replays; the
    convolution kernels and phase gradients are read at creation)."""

    def __init__(self, uvw, weight, field, field_id, freq_chan, chan_map, pol_map,
                 cf_baseline_map, cf_chan_map, cf_pol_map, conv_kernel, support,
                 phase_gradient, n_imag_chan, n_uv, delta_lm, oversampling, mode,
                 precision="fp32_kahan", vis=None, flag=None, grid=None, sum_weight=None,
                 accumulate=False, autotune=True):
        self._ctx = context()
        self.handle = None
        args, self.grid, self.sum_weight, _host_arrays = _aperture_args(
            uvw, weight, field, field_id, freq_chan, chan_map, pol_map, cf_baseline_map,
            cf_chan_map, cf_pol_map, conv_kernel, support, phase_gradient, n_imag_chan, n_uv,
            delta_lm, oversampling, mode, precision, vis, flag, grid, sum_weight, accumulate)
        # the plan holds device pointers: keep the tensors alive with it
        self._keep = (uvw, weight, field, vis, flag)
        h = _vp()
        rc = load_library().ngc_aperture_grid_plan_create(
            self._ctx.handle, _stream(), *args, int(bool(autotune)), ctypes.byref(h))
        if rc == -4:
            raise IndexError("a row's FIELD_ID is not in the gcf's field_id list")
        _check("ngc_aperture_grid_plan_create", rc)
        self.handle = h


def grid_to_image(grid, image_size, sum_weight=None, corr=None):
    """ngc_grid_to_image.  grid [..., n_u, n_v] (leading axes are planes).
    corr = (corr_l, corr_m) host arrays or None.  Returns float64
    [planes..., n_l, n_m] device tensor."""
    import torch
    ctx = context()
    _dev(grid, "grid")
    lead = tuple(grid.shape[:-2])
    n_u, n_v = grid.shape[-2:]
    n_planes = int(np.prod(lead)) if lead else 1
    n_l, n_m = int(image_size[0]), int(image_size[1])
    image = torch.empty(lead + (n_l, n_m), dtype=torch.float64, device=grid.device)
    if sum_weight is not None:
        _dev(sum_weight, "sum_weight")
        if sum_weight.numel() != n_planes or sum_weight.dtype != torch.float64:
            raise ValueError("sum_weight must be float64 with one entry per plane")
    cl = cm = None
    clp = cmp_ = None
    if corr is not None:
        cl, clp = _host(corr[0], np.float64)
        cm, cmp_ = _host(corr[1], np.float64)
    rc = load_library().ngc_grid_to_image(ctx.handle, _stream(), _ptr(grid), _dt_grid(grid),
                                          n_planes, n_u, n_v, n_l, n_m, _ptr(sum_weight), clp,
                                          cmp_, _ptr(image))
    _check("ngc_grid_to_image", rc)
    return image


def normalize_mosaic(image, sum_weight, ps_corr, norm_pb, pb, oversampling, pb_limit,
                     single_precision):
    """ngc_normalize_mosaic, in place on a float64 [planes..., n_l, n_m] tensor."""
    ctx = context()
    _dev(image, "image")
    n_l, n_m = image.shape[-2:]
    n_planes = image.numel() // (n_l * n_m)
    rc = load_library().ngc_normalize_mosaic(
        ctx.handle, _stream(), _ptr(image), _ptr(sum_weight), _ptr(ps_corr), _ptr(norm_pb),
        _ptr(pb), n_planes, n_l, n_m, int(oversampling[0]), int(oversampling[1]),
        float(pb_limit), int(bool(single_precision)))
    _check("ngc_normalize_mosaic", rc)
    return image


def profile_enable(enable=True):
    """Record per-phase HIP-event timings of subsequent calls."""
    ctx = context()
    _check("ngc_profile_enable", load_library().ngc_profile_enable(ctx.handle, int(bool(enable))))


def profile_read():
    """Per-phase durations (ms) of the last profiled call: list of (name, ms)."""
    ctx = context()
    names = ctypes.create_string_buffer(4096)
    ms = np.zeros(64, dtype=np.float64)
    n = load_library().ngc_profile_read(ctx.handle, names, 4096, ms.ctypes.data_as(_vp), 64)
    if n < 0:
        _check("ngc_profile_read", n)
    parts = names.value.decode().split(";") if n > 0 else []
    return list(zip(parts, ms[:n].tolist()))


def set_fft_mode(mode):
    """Image transform of single-precision grids on the current device: 1 (the
    default) the LDS row FFTs where the lengths allow, 0 always rocFFT."""
    ctx = context()
    _check("ngc_set_fft_mode", load_library().ngc_set_fft_mode(ctx.handle, int(mode)))


def set_degrid_mode(mode):
    """Standard degridding on the current device: 1 (the default) LDS tiles
    where the call allows them, 0 always the per-sample gather."""
    ctx = context()
    _check("ngc_set_degrid_mode", load_library().ngc_set_degrid_mode(ctx.handle, int(mode)))


def set_tile_rows(rows):
    """Tile rows (8, 16 or 32; 0, the default: 32 on calls with at least 65,536
    channel-tile bins of 32 x 32 cells, else 16) of the uv-tiled standard
    gridder on the current device."""
    ctx = context()
    _check("ngc_set_tile_rows", load_library().ngc_set_tile_rows(ctx.handle, int(rows)))


def set_tile_waves(waves):
    """Wavefronts per tile of the uv-tiled standard gridder, each gridding a
    share of the tile's records (1, 2 or 4; 0: default, 4 for the
    single-precision forms and 2 for FP64 tiled) on the current device."""
    ctx = context()
    _check("ngc_set_tile_waves", load_library().ngc_set_tile_waves(ctx.handle, int(waves)))


def set_bin_lanes(lanes):
    """Channels per binning wavefront of the uv-tiled standard gridder (1, 2, 4
    or 8) on the current device."""
    ctx = context()
    _check("ngc_set_bin_lanes", load_library().ngc_set_bin_lanes(ctx.handle, int(lanes)))


def gcf_patterns(pb_freq, dish, blockage, pairs, casa, ipower, cell_lm, n_uv):
    """ngc_gcf_patterns.  Returns (device float64 [n_pairs, n_freq, n_u, n_v],
    abs_max, abs_min, centre_row, centre_col) with the host arrays shaped
    [n_pairs, n_freq(, n)]."""
    import torch
    ctx = context()
    fr, frp = _host(pb_freq, np.float64)
    di, dip = _host(dish, np.float64)
    bl, blp = _host(blockage, np.float64)
    pr, prp = _host(pairs, np.int32)
    n_pairs, n_freq = len(pr), len(fr)
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    out = torch.empty((n_pairs, n_freq, n_u, n_v), dtype=torch.float64, device="cuda")
    mx = np.zeros(n_pairs * n_freq)
    mn = np.zeros(n_pairs * n_freq)
    row = np.zeros((n_pairs * n_freq, n_u))
    col = np.zeros((n_pairs * n_freq, n_v))
    rc = load_library().ngc_gcf_patterns(
        ctx.handle, _stream(), frp, n_freq, dip, blp, len(di), prp, n_pairs, int(bool(casa)),
        int(ipower), float(cell_lm[0]), float(cell_lm[1]), n_u, n_v, _ptr(out),
        mx.ctypes.data_as(_vp), mn.ctypes.data_as(_vp), row.ctypes.data_as(_vp),
        col.ctypes.data_as(_vp))
    _check("ngc_gcf_patterns", rc)
    sh = (n_pairs, n_freq)
    return out, mx.reshape(sh), mn.reshape(sh), row.reshape(sh + (n_u,)), col.reshape(sh + (n_v,))


def gcf_resize(full, support, oversampling, resize_conv_size):
    """ngc_gcf_resize on [..., n_u, n_v] planes; support [..., 2]."""
    import torch
    ctx = context()
    _dev(full, "full")
    lead = tuple(full.shape[:-2])
    n_items = int(np.prod(lead)) if lead else 1
    k_u, k_v = int(resize_conv_size[0]), int(resize_conv_size[1])
    sp, spp = _host(np.asarray(support).reshape(n_items, 2), np.int32)
    out = torch.empty(lead + (k_u, k_v), dtype=torch.float64, device=full.device)
    rc = load_library().ngc_gcf_resize(ctx.handle, _stream(), _ptr(full), n_items,
                                       full.shape[-2], full.shape[-1], spp,
                                       int(oversampling[0]), int(oversampling[1]), k_u, k_v,
                                       _ptr(out))
    _check("ngc_gcf_resize", rc)
    return out


def phase_gradient(pix, resize_conv_size):
    """ngc_phase_gradient: device complex128 [n_field, K_u, K_v]."""
    import torch
    ctx = context()
    px, pxp = _host(np.asarray(pix).reshape(-1, 2), np.float64)
    k_u, k_v = int(resize_conv_size[0]), int(resize_conv_size[1])
    out = torch.empty((len(px), k_u, k_v), dtype=torch.complex128, device="cuda")
    rc = load_library().ngc_phase_gradient(ctx.handle, _stream(), pxp, len(px), k_u, k_v,
                                           _ptr(out))
    _check("ngc_phase_gradient", rc)
    return out


def apply_flags(data, flag, out=None):
    """ngc_apply_flags: data where flag == 0, else the dtype's NaN value
    (NaN, integer minimum, True).  Device tensors of the same shape; returns
    out (a new tensor unless given; out may be data)."""
    import torch
    _dev(data, "data")
    _dev(flag, "flag")
    codes = {torch.complex64: C64, torch.complex128: C128, torch.float32: F32,
             torch.float64: F64, torch.int32: I32, torch.int64: I64, torch.bool: BOOL}
    if data.dtype not in codes:
        raise TypeError(f"apply_flags does not support {data.dtype}")
    if tuple(flag.shape) != tuple(data.shape):
        raise ValueError("flag and data shapes differ")
    if flag.dtype not in (torch.bool, torch.uint8):
        raise TypeError("flag must be bool or uint8")
    if not (data.is_contiguous() and flag.is_contiguous()):
        raise ValueError("data and flag must be contiguous")
    out = torch.empty_like(data) if out is None else out
    if out.dtype != data.dtype or tuple(out.shape) != tuple(data.shape) or \
            not out.is_contiguous():
        raise ValueError("out must be a contiguous tensor like data")
    rc = load_library().ngc_apply_flags(context().handle, _stream(), _ptr(data),
                                        codes[data.dtype], _ptr(flag), data.numel(), _ptr(out))
    _check("ngc_apply_flags", rc)
    return out


def model_to_grid(image, n_uv, corr_l=None, corr_m=None, single=True, out=None):
    """ngc_model_to_grid: image device float64 [n_chan, n_pol, n_l, n_m] (Jy/pixel)
    -> grid [n_chan, n_pol, n_u, n_v] complex64 (single) or complex128."""
    import torch
    _dev(image, "image")
    if image.dtype != torch.float64 or image.dim() != 4 or not image.is_contiguous():
        raise TypeError("image must be a contiguous 4-D float64 tensor")
    nc, npol, n_l, n_m = image.shape
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    dt = torch.complex64 if single else torch.complex128
    if out is None:
        out = torch.empty((nc, npol, n_u, n_v), dtype=dt, device=image.device)
    if out.dtype != dt or tuple(out.shape) != (nc, npol, n_u, n_v) or not out.is_contiguous():
        raise ValueError("out must be a contiguous grid of the requested precision")
    cl = cm = None
    if corr_l is not None:
        cl, clp = _host(np.asarray(corr_l).ravel(), np.float64)
        cm, cmp_ = _host(np.asarray(corr_m).ravel(), np.float64)
        if len(cl) != n_l or len(cm) != n_m:
            raise ValueError("correction vectors must match the image size")
    rc = load_library().ngc_model_to_grid(
        context().handle, _stream(), _ptr(image), nc * npol, n_l, n_m, n_u, n_v,
        clp if cl is not None else None, cmp_ if cm is not None else None, _ptr(out),
        C64 if single else C128)
    _check("ngc_model_to_grid", rc)
    return out


def standard_degrid(uvw, grid, freq_chan, chan_map, pol_map, cgk_1d, delta_lm, support,
                    oversampling, n_pol=None, normalize=True, single=True, out=None):
    """ngc_standard_degrid: predicted visibilities [T, B, C, P] (complex64 when
    single, else complex128) from grid [n_imag_chan, n_imag_pol, n_u, n_v]."""
    import torch
    _dev(uvw, "uvw")
    _dev(grid, "grid")
    if grid.dtype not in (torch.complex64, torch.complex128) or not grid.is_contiguous():
        raise TypeError("grid must be a contiguous complex64/complex128 tensor")
    T, B = _uvw_rows(uvw)
    fr, frp = _host(np.asarray(freq_chan).ravel(), np.float64)
    cm, cmp_ = _host(np.asarray(chan_map).ravel(), np.int32)
    pm, pmp = _host(np.asarray(pol_map).ravel(), np.int32)
    cg, cgp = _host(np.asarray(cgk_1d).ravel(), np.float64)
    C = len(fr)
    P = len(pm) if n_pol is None else int(n_pol)
    nic, nip, n_u, n_v = grid.shape
    dt = torch.complex64 if single else torch.complex128
    if out is None:
        out = torch.empty((T, B, C, P), dtype=dt, device=uvw.device)
    if out.dtype != dt or tuple(out.shape) != (T, B, C, P) or not out.is_contiguous():
        raise ValueError("out must be a contiguous [T, B, C, P] tensor of the requested dtype")
    rc = load_library().ngc_standard_degrid(
        context().handle, _stream(), _ptr(uvw), frp, cmp_, pmp, cgp, len(cg), T, B, C, P, nic,
        nip, n_u, n_v, float(delta_lm[0]), float(delta_lm[1]), int(support), int(oversampling),
        _ptr(grid), C64 if grid.dtype == torch.complex64 else C128, _ptr(out),
        C64 if single else C128, int(bool(normalize)))
    _check("ngc_standard_degrid", rc)
    return out


def aperture_degrid(uvw, field, field_id, grid, freq_chan, chan_map, pol_map, cf_baseline_map,
                    cf_chan_map, cf_pol_map, conv_kernel, support, phase_gradient, delta_lm,
                    oversampling, n_pol=None, normalize=True, single=True, out=None):
    """ngc_aperture_degrid: visibilities [T, B, C, P] (complex64 when single,
    else complex128) predicted from grid [n_imag_chan, n_imag_pol, n_u, n_v]
    with the A-term kernels (conv_kernel float64 [n_cfb, n_cfc, n_cfp, K_u,
    K_v], phase_gradient complex128 [n_field, K_u, K_v], device tensors)."""
    import torch
    for t, n in ((uvw, "uvw"), (field, "field"), (grid, "grid"), (conv_kernel, "conv_kernel"),
                 (phase_gradient, "phase_gradient")):
        _dev(t, n)
    T, B = _uvw_rows(uvw)
    _field_rows(field, T, B)
    if conv_kernel.dtype != torch.float64 or phase_gradient.dtype != torch.complex128:
        raise TypeError("conv_kernel must be float64 and phase_gradient complex128")
    if grid.dtype not in (torch.complex64, torch.complex128) or not grid.is_contiguous():
        raise TypeError("grid must be a contiguous complex64/complex128 tensor")
    fid, fidp = _host(np.asarray(field_id).ravel(), np.int32)
    fr, frp = _host(np.asarray(freq_chan).ravel(), np.float64)
    cm, cmp_ = _host(np.asarray(chan_map).ravel(), np.int32)
    pm, pmp = _host(np.asarray(pol_map).ravel(), np.int32)
    cb, cbp = _host(np.asarray(cf_baseline_map).ravel(), np.int32)
    cc, ccp = _host(np.asarray(cf_chan_map).ravel(), np.int32)
    cpm, cpp = _host(np.asarray(cf_pol_map).ravel(), np.int32)
    sp_, spp = _host(np.asarray(support).ravel(), np.int32)
    n_cfb, n_cfc, n_cfp, k_u, k_v = conv_kernel.shape
    if tuple(phase_gradient.shape[1:]) != (k_u, k_v):
        raise ValueError("phase_gradient must be [n_field, K_u, K_v]")
    if sp_.size != n_cfb * n_cfc * n_cfp * 2:
        raise ValueError("support must be [n_cfb, n_cfc, n_cfp, 2]")
    C = len(fr)
    P = len(pm) if n_pol is None else int(n_pol)
    nic, nip, n_u, n_v = grid.shape
    dt = torch.complex64 if single else torch.complex128
    if out is None:
        out = torch.empty((T, B, C, P), dtype=dt, device=uvw.device)
    if out.dtype != dt or tuple(out.shape) != (T, B, C, P) or not out.is_contiguous():
        raise ValueError("out must be a contiguous [T, B, C, P] tensor of the requested dtype")
    rc = load_library().ngc_aperture_degrid(
        context().handle, _stream(), _ptr(uvw), _ptr(field), fidp, len(fid), frp, cmp_, pmp,
        cbp, ccp, cpp, _ptr(conv_kernel), n_cfb, n_cfc, n_cfp, k_u, k_v, spp,
        _ptr(phase_gradient), T, B, C, P, nic, nip, n_u, n_v, float(delta_lm[0]),
        float(delta_lm[1]), int(oversampling[0]), int(oversampling[1]), _ptr(grid),
        C64 if grid.dtype == torch.complex64 else C128, _ptr(out), C64 if single else C128,
        int(bool(normalize)))
    if rc == -4:
        raise IndexError("a row's FIELD_ID is not in the gcf's field_id list")
    _check("ngc_aperture_degrid", rc)
    return out


def make_pb(freq_chan, dish, blockage, casa, ipower, image_size, image_center, cell, n_pol):
    """ngc_make_pb: device float64 [n_l, n_m, n_chan, n_pol, n_dish] Airy
    primary beams (cell in radians)."""
    import torch
    ctx = context()
    fr, frp = _host(np.asarray(freq_chan).ravel(), np.float64)
    di, dip = _host(np.asarray(dish).ravel(), np.float64)
    bl, blp = _host(np.asarray(blockage).ravel(), np.float64)
    if len(di) != len(bl):
        raise ValueError("dish and blockage lists differ in length")
    n_l, n_m = int(image_size[0]), int(image_size[1])
    out = torch.empty((n_l, n_m, len(fr), int(n_pol), len(di)), dtype=torch.float64,
                      device="cuda")
    rc = load_library().ngc_make_pb(ctx.handle, _stream(), frp, len(fr), dip, blp, len(di),
                                    int(bool(casa)), int(ipower), n_l, n_m,
                                    int(image_center[0]), int(image_center[1]),
                                    float(cell[0]), float(cell[1]), int(n_pol), _ptr(out))
    _check("ngc_make_pb", rc)
    return out


# ---------------------------------------------------------------------------
# dense-layer primitives of the kernel surrogate (float32 device tensors)
# ---------------------------------------------------------------------------
def bf16_gemm(A, B, bias=None, act=0, out=None, pre=None, accumulate=False, trans_a=False,
              trans_b=False, aux=None):
    """out = act(op(A) @ op(B) + bias) on the matrix cores (bf16 operands, FP32
    accumulation).  op(X) = X.T when trans_x; act 0 none, 1 tanh, 2 times
    1 - aux^2 (aux M x N).  Returns out (M x N float32)."""
    import torch
    for t, n in ((A, "A"), (B, "B")):
        _dev(t, n)
        if t.dtype != torch.float32 or t.dim() != 2:
            raise TypeError(f"{n} must be a 2-D float32 tensor")
    M, K = (A.shape[1], A.shape[0]) if trans_a else (A.shape[0], A.shape[1])
    Kb, N = (B.shape[1], B.shape[0]) if trans_b else (B.shape[0], B.shape[1])
    if K != Kb:
        raise ValueError("inner dimensions differ")
    sam, sak = (A.stride(1), A.stride(0)) if trans_a else (A.stride(0), A.stride(1))
    sbk, sbn = (B.stride(1), B.stride(0)) if trans_b else (B.stride(0), B.stride(1))
    if out is None:
        out = torch.empty((M, N), dtype=torch.float32, device=A.device)
    if not out.is_contiguous() or tuple(out.shape) != (M, N):
        raise ValueError("out must be a contiguous M x N tensor")
    if pre is not None and (not pre.is_contiguous() or tuple(pre.shape) != (M, N)):
        raise ValueError("pre must be a contiguous M x N tensor")
    if (act == 2) != (aux is not None):
        raise ValueError("aux is given exactly with act=2")
    if aux is not None and (not aux.is_contiguous() or tuple(aux.shape) != (M, N)):
        raise ValueError("aux must be a contiguous M x N tensor")
    rc = load_library().ngc_bf16_gemm(context().handle, _stream(), M, N, K, _ptr(A), sam, sak,
                                      _ptr(B), sbk, sbn, _ptr(bias), int(act),
                                      int(bool(accumulate)), _ptr(out), N, _ptr(pre),
                                      _ptr(aux))
    _check("ngc_bf16_gemm", rc)
    return out


def tanh_backward(dA, A, out=None):
    import torch
    out = torch.empty_like(dA) if out is None else out
    _check("ngc_tanh_backward", load_library().ngc_tanh_backward(
        context().handle, _stream(), dA.numel(), _ptr(dA), _ptr(A), _ptr(out)))
    return out


def col_sum(X, out=None):
    import torch
    out = torch.empty(X.shape[1], dtype=torch.float32, device=X.device) if out is None else out
    _check("ngc_col_sum", load_library().ngc_col_sum(context().handle, _stream(), X.shape[0],
                                                     X.shape[1], _ptr(X), _ptr(out)))
    return out


def mse_grad(y, t, g=None, loss=None):
    """Returns (gradient, device scalar loss)."""
    import torch
    g = torch.empty_like(y) if g is None else g
    loss = torch.empty(1, dtype=torch.float64, device=y.device) if loss is None else loss
    _check("ngc_mse_grad", load_library().ngc_mse_grad(context().handle, _stream(), y.numel(),
                                                       _ptr(y), _ptr(t), _ptr(g), _ptr(loss)))
    return g, loss


def adam_step(w, g, m, v, lr, step, beta1=0.9, beta2=0.999, eps=1e-8):
    _check("ngc_adam_step", load_library().ngc_adam_step(
        context().handle, _stream(), w.numel(), _ptr(w), _ptr(g), _ptr(m), _ptr(v), float(lr),
        float(beta1), float(beta2), float(eps), int(step)))


def mlp_sample_rows(x, t, batch, seed, counter, xb, tb):
    """ngc_mlp_sample_rows: the minibatch of the step in `counter` (device int64)."""
    _check("ngc_mlp_sample_rows", load_library().ngc_mlp_sample_rows(
        context().handle, _stream(), _ptr(x), _ptr(t), x.shape[0], x.shape[1], t.shape[1],
        int(batch), int(seed) & ((1 << 64) - 1), _ptr(counter), _ptr(xb), _ptr(tb)))


def adam_schedule(counter, sched, lr, lr_min, steps, step0, beta1=0.9, beta2=0.999):
    """ngc_adam_schedule: rate and bias corrections of the step in `counter`,
    written to sched (device float [3]); advances the counter."""
    _check("ngc_adam_schedule", load_library().ngc_adam_schedule(
        context().handle, _stream(), _ptr(counter), _ptr(sched), float(lr), float(lr_min),
        int(steps), int(step0), float(beta1), float(beta2)))


def adam_step_sched(w, g, m, v, sched, beta1=0.9, beta2=0.999, eps=1e-8):
    """ngc_adam_step_sched: Adam with the rate and corrections read from sched."""
    _check("ngc_adam_step_sched", load_library().ngc_adam_step_sched(
        context().handle, _stream(), w.numel(), _ptr(w), _ptr(g), _ptr(m), _ptr(v),
        _ptr(sched), float(beta1), float(beta2), float(eps)))


def direction_rotate(uvw, vis, field, field_ids, uvw_rotmat, phase_rotation, freq_chan,
                     common_tangent=True, single_precision=True):
    """ngc_direction_rotate on device tensors.  Returns (uvw_rot float64,
    vis_rot complex64 if single_precision else complex128)."""
    import torch
    ctx = context()
    for t, n in ((uvw, "uvw"), (vis, "vis"), (field, "field")):
        _dev(t, n)
    T, B, C, P = vis.shape
    _uvw_rows(uvw, (T, B))
    _field_rows(field, T, B)
    uvw_out = torch.empty_like(uvw)
    vis_out = torch.empty(vis.shape, dtype=torch.complex64 if single_precision else
                          torch.complex128, device=vis.device)
    ids, idp = _host(field_ids, np.int32)
    m, mp = _host(np.asarray(uvw_rotmat).reshape(-1, 9), np.float64)
    ph, php = _host(np.asarray(phase_rotation).reshape(-1, 3), np.float64)
    fr, frp = _host(freq_chan, np.float64)
    rc = load_library().ngc_direction_rotate(
        ctx.handle, _stream(), _ptr(uvw), _ptr(vis), _dt_vis(vis), _ptr(field), idp, len(ids),
        mp, php, frp, T, B, C, P, int(bool(common_tangent)), int(bool(single_precision)),
        _ptr(uvw_out), _ptr(vis_out))
    if rc == -4:
        raise AssertionError("direction_rotate only supports xds where field_id remains constant "
                             "over baseline (and is listed in FIELD)")
    _check("ngc_direction_rotate", rc)
    return uvw_out, vis_out
