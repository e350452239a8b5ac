"""ORACLE — test infrastructure only.

CPU restatement of ngCASA's imaging hot path, used as the checker for the HIP
library.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product
(``cngi_prototype_amd``) never does.

Two halves:

* ``liboracle.so`` (``grid_oracle.c``): the four numba loops restated in plain
  C, FP64, reference loop order — wrapped here by :func:`standard_grid`,
  :func:`imaging_weight_degrid`, :func:`aperture_grid`,
  :func:`aperture_weight_grid` and :func:`standard_grid_chunked`.
* numpy restatements of the host-side arithmetic: the prolate-spheroidal
  function and kernel tables, Briggs factors, the A-term GCF (Airy patterns,
  FFT, support cut, resize/normalise, cf maps), the phase gradient, and the
  post-grid FFT/crop/normalisation.

Each function cites the reference file:line it follows.  How each one is
pinned (golden vectors, closed forms, invariants, brute force) is listed in
``tests/test_oracle_pinning.py``; functions with no independent pin say so in
their docstring ("parity unpinned").
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

C_LIGHT = 299792458.0
ARCSEC_TO_RAD = np.pi / (3600 * 180)


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def use_native_build(directory=None):
    """Compile the restatement with -O3 -march=native for this host (into
    `directory`, a temporary one by default) and route every call of this
    module through that build.  Used for the timed CPU baseline only: the
    results are those of the default build (no fast-math, no contraction).
    Returns the library path."""
    global _LIB
    import subprocess
    import tempfile
    d = directory or tempfile.mkdtemp(prefix="oracle_native_")
    os.makedirs(d, exist_ok=True)
    subprocess.run(["make", "-s", "-C", _HERE, "native", f"NATIVE_DIR={d}"], check=True)
    path = os.path.join(d, "liboracle_native.so")
    _LIB = ctypes.CDLL(path)
    return path


def cpu_model():
    """The host CPU's model name and the cores this process may use."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return name, avail


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _c128_view(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


# --------------------------------------------------------------------------
# Gridding loops (C)
# --------------------------------------------------------------------------
def standard_grid(vis, uvw, weight, freq_chan, cgk_1d, n_uv, delta_lm, support,
                  oversampling, do_psf, do_imaging_weight=False, chan_mode="cube",
                  complex_grid=None):
    """_standard_grid_jit plus its wrappers' set-up.

    Follows _standard_grid.py:123-177 (image), :180-236 (PSF) and :243-371
    (the loop).  Returns (grid [n_imag_chan, n_pol, n_u, n_v], sum_weight
    [n_imag_chan, n_pol]) in the kernel's own orientation.
    """
    weight = _f64(weight)
    uvw = _f64(uvw)
    n_time, n_baseline, n_chan, n_pol = weight.shape
    if chan_mode == "cube":
        n_imag_chan = n_chan
        chan_map = np.arange(n_chan, dtype=np.int64)
    else:
        n_imag_chan = 1
        chan_map = np.zeros(n_chan, dtype=np.int64)
    pol_map = np.arange(n_pol, dtype=np.int64)
    if complex_grid is None:
        complex_grid = not do_psf
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    shape = (n_imag_chan, n_pol, n_u, n_v)
    grid = np.zeros(shape, dtype=np.complex128 if complex_grid else np.float64)
    sum_weight = np.zeros((n_imag_chan, n_pol), dtype=np.float64)
    visc = None if do_psf else _c128_view(vis)
    L = _lib()
    L.oracle_standard_grid.restype = None
    L.oracle_standard_grid(
        _p(grid), _p(sum_weight), ctypes.c_int(int(do_psf)),
        ctypes.c_int(int(do_imaging_weight)), ctypes.c_int(int(complex_grid)),
        _p(visc), _p(uvw), _p(_f64(freq_chan)), _p(chan_map), _p(pol_map),
        _p(weight), _p(_f64(cgk_1d)), ctypes.c_int64(n_time),
        ctypes.c_int64(n_baseline), ctypes.c_int64(n_chan), ctypes.c_int64(n_pol),
        ctypes.c_int64(n_imag_chan), ctypes.c_int64(n_pol), ctypes.c_int64(n_u),
        ctypes.c_int64(n_v), ctypes.c_double(float(delta_lm[0])),
        ctypes.c_double(float(delta_lm[1])), ctypes.c_int64(int(support)),
        ctypes.c_int64(int(oversampling)))
    return grid, sum_weight


def standard_grid_chunked(vis, uvw, weight, freq_chan, cgk_1d, n_uv, delta_lm,
                          support, oversampling, do_psf, time_chunk, n_threads=0,
                          chan_mode="cube"):
    """The reference's graph layer on the CPU: one task per time chunk with a
    private grid, then a pairwise tree-sum (_standard_grid.py:23-120).
    Returns (grid, sum_weight, threads_used).  This is the timed CPU baseline."""
    weight = _f64(weight)
    uvw = _f64(uvw)
    n_time, n_baseline, n_chan, n_pol = weight.shape
    if chan_mode == "cube":
        n_imag_chan = n_chan
        chan_map = np.arange(n_chan, dtype=np.int64)
    else:
        n_imag_chan = 1
        chan_map = np.zeros(n_chan, dtype=np.int64)
    pol_map = np.arange(n_pol, dtype=np.int64)
    complex_grid = not do_psf
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    grid = np.zeros((n_imag_chan, n_pol, n_u, n_v),
                    dtype=np.complex128 if complex_grid else np.float64)
    sum_weight = np.zeros((n_imag_chan, n_pol), dtype=np.float64)
    visc = None if do_psf else _c128_view(vis)
    L = _lib()
    L.oracle_standard_grid_chunked.restype = ctypes.c_int
    used = L.oracle_standard_grid_chunked(
        _p(grid), _p(sum_weight), ctypes.c_int(int(do_psf)),
        ctypes.c_int(int(complex_grid)), _p(visc), _p(uvw), _p(_f64(freq_chan)),
        _p(chan_map), _p(pol_map), _p(weight), _p(_f64(cgk_1d)),
        ctypes.c_int64(n_time), ctypes.c_int64(n_baseline), ctypes.c_int64(n_chan),
        ctypes.c_int64(n_pol), ctypes.c_int64(n_imag_chan), ctypes.c_int64(n_pol),
        ctypes.c_int64(n_u), ctypes.c_int64(n_v),
        ctypes.c_double(float(delta_lm[0])), ctypes.c_double(float(delta_lm[1])),
        ctypes.c_int64(int(support)), ctypes.c_int64(int(oversampling)),
        ctypes.c_int64(int(time_chunk)), ctypes.c_int(int(n_threads)))
    return grid, sum_weight, used


def imaging_weight_degrid(density_grid_img, uvw, natural_weight, briggs_factors,
                          freq_chan, n_uv, delta_lm, chan_mode="cube"):
    """_standard_imaging_weight_degrid_jit with its wrapper
    (_standard_grid.py:443-518).  ``density_grid_img`` is in image orientation
    [n_u, n_v, n_imag_chan, n_pol]; ``briggs_factors`` is [2, n_imag_chan,
    n_pol]."""
    natural_weight = _f64(natural_weight)
    n_time, n_baseline, n_chan, n_pol = natural_weight.shape
    if chan_mode == "cube":
        chan_map = np.arange(n_chan, dtype=np.int64)
    else:
        chan_map = np.zeros(n_chan, dtype=np.int64)
    pol_map = np.arange(n_pol, dtype=np.int64)
    g = _f64(density_grid_img)
    n_imag_chan = g.shape[2]
    out = np.zeros(natural_weight.shape, dtype=np.float64)
    L = _lib()
    L.oracle_imaging_weight_degrid.restype = None
    L.oracle_imaging_weight_degrid(
        _p(out), _p(g), _p(_f64(briggs_factors)), _p(_f64(uvw)),
        _p(_f64(freq_chan)), _p(chan_map), _p(pol_map), _p(natural_weight),
        ctypes.c_int64(n_time), ctypes.c_int64(n_baseline), ctypes.c_int64(n_chan),
        ctypes.c_int64(n_pol), ctypes.c_int64(n_imag_chan), ctypes.c_int64(n_pol),
        ctypes.c_int64(int(n_uv[0])), ctypes.c_int64(int(n_uv[1])),
        ctypes.c_double(float(delta_lm[0])), ctypes.c_double(float(delta_lm[1])))
    return out


def _aperture_common(uvw, weight, field, field_id, cf_baseline_map, cf_chan_map,
                     cf_pol_map, kernel, support, phase_gradient, chan_mode):
    weight = _f64(weight)
    n_time, n_baseline, n_chan, n_pol = weight.shape
    if chan_mode == "cube":
        n_imag_chan = n_chan
        chan_map = np.arange(n_chan, dtype=np.int64)
    else:
        n_imag_chan = 1
        chan_map = np.zeros(n_chan, dtype=np.int64)
    pol_map = np.arange(n_pol, dtype=np.int64)
    kernel = _f64(kernel)
    if kernel.ndim != 5:
        raise ValueError("kernel must be [n_cfb, n_cfc, n_cfp, K_u, K_v]")
    pg = np.ascontiguousarray(phase_gradient, dtype=np.complex128)
    return dict(weight=weight, dims=(n_time, n_baseline, n_chan, n_pol),
                n_imag_chan=n_imag_chan, chan_map=chan_map, pol_map=pol_map,
                kernel=kernel, pg=pg, support=_i64(support), field=_i64(field),
                field_id=_i64(field_id), cfb=_i64(cf_baseline_map),
                cfc=_i64(cf_chan_map), cfp=_i64(cf_pol_map), uvw=_f64(uvw))


def aperture_grid(vis, uvw, weight, field, field_id, cf_baseline_map, cf_chan_map,
                  cf_pol_map, conv_kernel, weight_support, phase_gradient,
                  freq_chan, n_uv, delta_lm, oversampling, do_psf,
                  chan_mode="cube"):
    """_aperture_grid_jit with its wrappers (_aperture_grid.py:294-513).
    Returns (complex grid [n_imag_chan, n_pol, n_u, n_v], sum_weight)."""
    c = _aperture_common(uvw, weight, field, field_id, cf_baseline_map,
                         cf_chan_map, cf_pol_map, conv_kernel, weight_support,
                         phase_gradient, chan_mode)
    n_time, n_baseline, n_chan, n_pol = c["dims"]
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    grid = np.zeros((c["n_imag_chan"], n_pol, n_u, n_v), dtype=np.complex128)
    sw = np.zeros((c["n_imag_chan"], n_pol), dtype=np.float64)
    k = c["kernel"]
    visc = None if do_psf else _c128_view(vis)
    L = _lib()
    L.oracle_aperture_grid.restype = ctypes.c_int
    rc = L.oracle_aperture_grid(
        _p(grid), _p(sw), ctypes.c_int(int(do_psf)), _p(visc), _p(c["uvw"]),
        _p(_f64(freq_chan)), _p(c["chan_map"]), _p(c["pol_map"]), _p(c["cfb"]),
        _p(c["cfc"]), _p(c["cfp"]), _p(c["weight"]), _p(k),
        ctypes.c_int64(k.shape[0]), ctypes.c_int64(k.shape[1]),
        ctypes.c_int64(k.shape[2]), ctypes.c_int64(k.shape[3]),
        ctypes.c_int64(k.shape[4]), ctypes.c_int64(n_time),
        ctypes.c_int64(n_baseline), ctypes.c_int64(n_chan), ctypes.c_int64(n_pol),
        ctypes.c_int64(c["n_imag_chan"]), ctypes.c_int64(n_pol),
        ctypes.c_int64(n_u), ctypes.c_int64(n_v),
        ctypes.c_double(float(delta_lm[0])), ctypes.c_double(float(delta_lm[1])),
        _p(c["support"]), ctypes.c_int64(int(oversampling[0])),
        ctypes.c_int64(int(oversampling[1])), _p(c["field"]), _p(c["field_id"]),
        ctypes.c_int64(len(c["field_id"])), _p(c["pg"]))
    if rc != 0:
        raise IndexError("field id not found in the gcf field_id list")
    return grid, sw


def aperture_degrid(grid, uvw, field, field_id, cf_baseline_map, cf_chan_map, cf_pol_map,
                    conv_kernel, weight_support, phase_gradient, freq_chan, n_uv, delta_lm,
                    oversampling, chan_mode="cube", normalize=True, n_pol=None):
    """Visibility prediction with the A-term kernels: the adjoint of
    aperture_grid's loop with unit weights (same rows, fields, cells, offsets
    and CK*PG taps), V = sum conj(CK*PG) G[cell], divided by the tap sum
    sum CK (without the phase gradient) when normalize.  grid [n_imag_chan, n_pol, n_u, n_v] complex; returns vis
    [T, B, C, P] complex128 (0 for rows and samples that are not gridded)."""
    grid = np.ascontiguousarray(grid, dtype=np.complex128)
    uvw = _f64(uvw)
    freq_chan = _f64(freq_chan)
    n_time, n_baseline, n_chan = uvw.shape[0], uvw.shape[1], len(freq_chan)
    n_pol = grid.shape[1] if n_pol is None else int(n_pol)
    c = _aperture_common(uvw, np.zeros((n_time, n_baseline, n_chan, n_pol)), field, field_id,
                         cf_baseline_map, cf_chan_map, cf_pol_map, conv_kernel,
                         weight_support, phase_gradient, chan_mode)
    k = c["kernel"]
    vis = np.zeros((n_time, n_baseline, n_chan, n_pol), dtype=np.complex128)
    L = _lib()
    L.oracle_aperture_degrid.restype = ctypes.c_int
    rc = L.oracle_aperture_degrid(
        _p(vis.view(np.float64)), _p(grid.view(np.float64)), _p(c["uvw"]), _p(freq_chan),
        _p(c["chan_map"]), _p(c["pol_map"]), _p(c["cfb"]), _p(c["cfc"]), _p(c["cfp"]), _p(k),
        ctypes.c_int64(k.shape[0]), ctypes.c_int64(k.shape[1]), ctypes.c_int64(k.shape[2]),
        ctypes.c_int64(k.shape[3]), ctypes.c_int64(k.shape[4]), ctypes.c_int64(n_time),
        ctypes.c_int64(n_baseline), ctypes.c_int64(n_chan), ctypes.c_int64(n_pol),
        ctypes.c_int64(grid.shape[1]), ctypes.c_int64(int(n_uv[0])), ctypes.c_int64(int(n_uv[1])),
        ctypes.c_double(float(delta_lm[0])), ctypes.c_double(float(delta_lm[1])),
        _p(c["support"]), ctypes.c_int64(int(oversampling[0])),
        ctypes.c_int64(int(oversampling[1])), _p(c["field"]), _p(c["field_id"]),
        ctypes.c_int64(len(c["field_id"])), _p(c["pg"]), ctypes.c_int(int(bool(normalize))))
    if rc != 0:
        raise IndexError("field id not found in the gcf field_id list")
    return vis


def aperture_weight_grid(uvw, weight, field, field_id, cf_baseline_map,
                         cf_chan_map, cf_pol_map, weight_conv_kernel,
                         weight_support, phase_gradient, freq_chan, n_uv,
                         delta_lm, oversampling, chan_mode="cube"):
    """_aperture_weight_grid_jit with its wrapper (_aperture_grid.py:146-291)."""
    c = _aperture_common(uvw, weight, field, field_id, cf_baseline_map,
                         cf_chan_map, cf_pol_map, weight_conv_kernel,
                         weight_support, phase_gradient, chan_mode)
    n_time, n_baseline, n_chan, n_pol = c["dims"]
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    grid = np.zeros((c["n_imag_chan"], n_pol, n_u, n_v), dtype=np.complex128)
    sw = np.zeros((c["n_imag_chan"], n_pol), dtype=np.float64)
    k = c["kernel"]
    L = _lib()
    L.oracle_aperture_weight_grid.restype = ctypes.c_int
    rc = L.oracle_aperture_weight_grid(
        _p(grid), _p(sw), _p(c["uvw"]), _p(_f64(freq_chan)), _p(c["chan_map"]),
        _p(c["pol_map"]), _p(c["cfb"]), _p(c["cfc"]), _p(c["cfp"]),
        _p(c["weight"]), _p(k), ctypes.c_int64(k.shape[0]),
        ctypes.c_int64(k.shape[1]), ctypes.c_int64(k.shape[2]),
        ctypes.c_int64(k.shape[3]), ctypes.c_int64(k.shape[4]),
        ctypes.c_int64(n_time), ctypes.c_int64(n_baseline), ctypes.c_int64(n_chan),
        ctypes.c_int64(n_pol), ctypes.c_int64(c["n_imag_chan"]),
        ctypes.c_int64(n_pol), ctypes.c_int64(n_u), ctypes.c_int64(n_v),
        ctypes.c_double(float(delta_lm[0])), ctypes.c_double(float(delta_lm[1])),
        _p(c["support"]), ctypes.c_int64(int(oversampling[0])),
        ctypes.c_int64(int(oversampling[1])), _p(c["field"]), _p(c["field_id"]),
        ctypes.c_int64(len(c["field_id"])), _p(c["pg"]))
    if rc != 0:
        raise IndexError("field id not found in the gcf field_id list")
    return grid, sw


# --------------------------------------------------------------------------
# Prolate spheroidal kernel (_gridding_convolutional_kernels.py)
# --------------------------------------------------------------------------
_PS_P = np.array([[8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1],
                  [4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2]])
_PS_Q = np.array([[1.0000000e0, 8.212018e-1, 2.078043e-1],
                  [1.0000000e0, 9.599102e-1, 2.918724e-1]])


def prolate_spheroidal_function(u):
    """Schwab's rational approximation, M=6, alpha=1
    (_gridding_convolutional_kernels.py:101-148).  Returns (grdsf,
    (1-u^2)*grdsf)."""
    u = np.abs(np.asarray(u, dtype=np.float64))
    uend = np.zeros(u.shape, dtype=np.float64)
    part = np.zeros(u.shape, dtype=np.int64)
    part[(u >= 0.0) & (u < 0.75)] = 0
    part[(u >= 0.75) & (u <= 1.0)] = 1
    uend[(u >= 0.0) & (u < 0.75)] = 0.75
    uend[(u >= 0.75) & (u <= 1.0)] = 1.0
    delusq = u ** 2 - uend ** 2
    top = _PS_P[part, 0].copy()
    for k in range(1, _PS_P.shape[1]):
        top += _PS_P[part, k] * np.power(delusq, k)
    bot = _PS_Q[part, 0].copy()
    for k in range(1, _PS_Q.shape[1]):
        bot += _PS_Q[part, k] * np.power(delusq, k)
    grdsf = np.zeros(u.shape, dtype=np.float64)
    ok = bot > 0.0
    grdsf[ok] = top[ok] / bot[ok]
    grdsf[np.abs(u > 1.0)] = 0.0
    return grdsf, (1 - u ** 2) * grdsf


def ps_kernel_1d(oversampling, support):
    """_create_prolate_spheroidal_kernel_1D (:151-157)."""
    sc = support // 2
    u = np.arange(oversampling * sc) / (sc * oversampling)
    k = np.zeros(oversampling * (sc + 1))
    _, k[0:oversampling * sc] = prolate_spheroidal_function(u)
    return k


def _coordinates(n):
    return (np.arange(n) - n // 2) / n


def ps_correction_image(n_uv):
    """Correction image of _create_prolate_spheroidal_kernel (:87-93)."""
    cu = prolate_spheroidal_function(np.abs(2.0 * _coordinates(int(n_uv[0]))))[0]
    cv = prolate_spheroidal_function(np.abs(2.0 * _coordinates(int(n_uv[1]))))[0]
    return np.outer(cu, cv)


def remove_padding(image, image_size):
    """_remove_padding (_remove_padding.py:20-30), leading two axes."""
    padded = np.array(image.shape[0:2])
    image_size = np.asarray(image_size)
    start = padded // 2 - image_size // 2
    end = start + image_size
    return image[start[0]:end[0], start[1]:end[1]]


def grid_to_image(grid_kernel_orient, image_size):
    """FFT + crop + scale as make_psf.py:117-121 / make_image.py:116-120.
    Input [n_chan, n_pol, n_u, n_v]; output [N_l, N_m, n_chan, n_pol] real,
    not yet divided by sum_weight or the correction image."""
    g = np.moveaxis(grid_kernel_orient, [0, 1], [-2, -1])
    img = np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(g, axes=(0, 1)), axes=(0, 1)),
                          axes=(0, 1))
    npad = g.shape[0:2]
    return remove_padding(img, image_size).real * (npad[0] * npad[1])


def model_to_grid(model, n_uv, corr=True):
    """Model image -> uv grid for prediction, the inverse placement of
    grid_to_image: pad the image [N_l, N_m, n_chan, n_pol] (Jy/pixel),
    divided by the cropped correction image (when corr; the A-term path
    takes the model as it is), into the centre of the n_uv grid, then
    fftshift(fft2(ifftshift(.))) (unnormalised forward transform, so a pixel
    of flux S gives |G| = S).  Returns [n_chan, n_pol, n_u, n_v] complex in
    the gridder's orientation.  No reference counterpart:
    predict_modelvis_image.py:20 is a stub."""
    model = np.asarray(model, dtype=np.float64)
    n_u, n_v = int(n_uv[0]), int(n_uv[1])
    size = np.array(model.shape[0:2])
    pad = np.zeros((n_u, n_v) + model.shape[2:], dtype=np.float64)
    start = np.array([n_u, n_v]) // 2 - size // 2
    if corr:
        model = model / remove_padding(ps_correction_image(n_uv), size)[:, :, None, None]
    pad[start[0]:start[0] + size[0], start[1]:start[1] + size[1]] = model
    g = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(pad, axes=(0, 1)), axes=(0, 1)),
                        axes=(0, 1))
    return np.ascontiguousarray(np.moveaxis(g, [-2, -1], [0, 1]))


def standard_degrid(grid, uvw, freq_chan, cgk_1d, n_uv, delta_lm, support, oversampling,
                    chan_mode="cube", normalize=True, n_pol=None):
    """Visibility prediction by degridding: the adjoint of standard_grid's
    loop with unit weights (same cells, offsets, bounds and taps), divided by
    the tap sum when normalize.  grid [n_imag_chan, n_pol, n_u, n_v] complex;
    returns vis [T, B, C, P] complex128 (0 for samples off the grid)."""
    grid = np.ascontiguousarray(grid, dtype=np.complex128)
    uvw = _f64(uvw)
    freq_chan = _f64(freq_chan)
    n_time, n_baseline = uvw.shape[0], uvw.shape[1]
    n_chan = len(freq_chan)
    n_imag_chan, n_imag_pol = grid.shape[0], grid.shape[1]
    n_pol = n_imag_pol if n_pol is None else int(n_pol)
    chan_map = (np.arange(n_chan, dtype=np.int64) if chan_mode == "cube"
                else np.zeros(n_chan, dtype=np.int64))
    pol_map = np.arange(n_pol, dtype=np.int64)
    vis = np.zeros((n_time, n_baseline, n_chan, n_pol), dtype=np.complex128)
    L = _lib()
    L.oracle_standard_degrid.restype = None
    L.oracle_standard_degrid(
        _p(vis.view(np.float64)), _p(grid.view(np.float64)), _p(uvw), _p(freq_chan),
        _p(chan_map), _p(pol_map), _p(_f64(cgk_1d)), ctypes.c_int64(n_time),
        ctypes.c_int64(n_baseline), ctypes.c_int64(n_chan), ctypes.c_int64(n_pol),
        ctypes.c_int64(n_imag_chan), ctypes.c_int64(n_imag_pol), ctypes.c_int64(int(n_uv[0])),
        ctypes.c_int64(int(n_uv[1])), ctypes.c_double(float(delta_lm[0])),
        ctypes.c_double(float(delta_lm[1])), ctypes.c_int64(int(support)),
        ctypes.c_int64(int(oversampling)), ctypes.c_int(int(bool(normalize))))
    return vis


def correct_standard_image(uncorrected, sum_weight, n_uv, image_size):
    """correct_image of make_psf.py:124-131 / make_image.py:123-130."""
    corr = remove_padding(ps_correction_image(n_uv), image_size)
    sw = np.array(sum_weight, dtype=np.float64, copy=True)
    sw[sw == 0] = 1
    return (uncorrected / sw[None, None, :, :]) / corr[:, :, None, None]


# --------------------------------------------------------------------------
# Briggs weighting (make_imaging_weight.py:144-255)
# --------------------------------------------------------------------------
def briggs_factors(density_grid_img, sum_weight, weighting, robust=0.5):
    """calculate_briggs_parms (make_imaging_weight.py:198-213).  Input grid in
    image orientation [n_u, n_v, n_chan, n_pol].  Returns [2, n_chan, n_pol]."""
    if weighting == "briggs":
        f = np.ones((2,) + sum_weight.shape)
        sq = np.sum(density_grid_img ** 2, axis=(0, 1))
        f[0] = np.square(5.0 * 10.0 ** (-robust)) / (sq / sum_weight)
    elif weighting == "uniform":
        f = np.zeros((2,) + sum_weight.shape)
        f[0] = 1.0
    else:
        raise ValueError("weighting " + str(weighting))
    return f


# --------------------------------------------------------------------------
# A-term GCF (make_gridding_convolution_function.py, _make_pb_symmetric.py)
# --------------------------------------------------------------------------
def airy_rorder(freq_chan, n_pol, dish, blockage, ipower, cell, image_size,
                image_center, casa=True):
    """_casa_airy_disk_rorder (_make_pb_symmetric.py:187-235) and
    _airy_disk_rorder (:135-183).  Returns [n_dish, n_freq, n_pol, Nx, Ny]."""
    from scipy.special import jn
    freq_chan = np.asarray(freq_chan, dtype=np.float64)
    k = (2 * np.pi * freq_chan) / C_LIGHT
    x = np.arange(-image_center[0], image_size[0] - image_center[0]) * cell[0]
    y = np.arange(-image_center[1], image_size[1] - image_center[1]) * cell[1]
    out = np.zeros((len(dish), len(freq_chan), 1, image_size[0], image_size[1]))
    for i, (d, bl) in enumerate(zip(dish, blockage)):
        aperture = d / 2
        xg, yg = np.meshgrid(x, y, indexing="ij")
        r = np.moveaxis(np.sqrt(xg ** 2 + yg ** 2)[:, :, None] * k * aperture, 2, 0)
        r[:, image_center[0], image_center[1]] = 1.0
        if bl == 0.0:
            out[i, :, 0] = (2.0 * jn(1, r) / r) ** ipower
        elif casa:
            ar = (d / bl) ** 2
            lr = d / bl
            out[i, :, 0] = ((ar * 2.0 * jn(1, r) / r - 2.0 * jn(1, r * lr) / (r * lr))
                            / (ar - 1.0)) ** ipower
        else:
            e = bl / d
            out[i, :, 0] = ((2.0 * jn(1, r) / r - 2.0 * e * jn(1, r * e) / r)
                            / (1.0 - e ** 2)) ** ipower
    out[:, :, 0, image_center[0], image_center[1]] = 1.0
    return np.tile(out, (1, 1, n_pol, 1, 1))


def pb_image(freq_chan, n_pol, dish, blockage, ipower, cell, image_size, image_center,
             casa=True):
    """_casa_airy_disk / _airy_disk (_make_pb_symmetric.py:24-130), the
    pattern function of make_pb (make_pb.py:86-104): the same values as the
    _rorder forms above, laid out [Nx, Ny, n_freq, n_pol, n_dish]."""
    return np.transpose(airy_rorder(freq_chan, n_pol, dish, blockage, ipower, cell, image_size,
                                    image_center, casa), (3, 4, 1, 2, 0))


def cf_baseline_map(unique_ant_indx, baseline_ant, n_unique_ant):
    """create_cf_baseline_map (make_gridding_convolution_function.py:512-528)."""
    n_pairs = int((n_unique_ant ** 2 + n_unique_ant) / 2)
    pairs = np.zeros((n_pairs, 2), dtype=int)
    k = 0
    for i in range(n_unique_ant):
        for j in range(i, n_unique_ant):
            pairs[k] = [i, j]
            k += 1
    unique_ant_indx = np.asarray(unique_ant_indx)
    baseline_ant = np.asarray(baseline_ant)
    m = np.zeros((baseline_ant.shape[0],), dtype=int)
    bu = np.stack([unique_ant_indx[baseline_ant[:, 0]], unique_ant_indx[baseline_ant[:, 1]]], 1)
    for k, ij in enumerate(pairs):
        m[(bu[:, 0] == ij[0]) & (bu[:, 1] == ij[1])] = k
    return m, pairs


def cf_chan_map(freq_chan, chan_tolerance_factor):
    """create_cf_chan_map (make_gridding_convolution_function.py:536-564)."""
    freq_chan = np.asarray(freq_chan, dtype=np.float64)
    n_chan = len(freq_chan)
    tol = np.max(freq_chan) * chan_tolerance_factor
    n_pb_chan = int(np.floor((np.max(freq_chan) - np.min(freq_chan)) / tol) + 0.5)
    if n_pb_chan == 0:
        n_pb_chan = 1
    if n_pb_chan >= n_chan:
        return np.arange(n_chan), freq_chan
    d = (np.max(freq_chan) - np.min(freq_chan)) / n_pb_chan
    pb_freq = np.arange(n_pb_chan) * d + np.min(freq_chan) + d / 2
    m = np.zeros((n_chan,), dtype=int)
    for i in range(n_chan):
        m[i] = (np.abs(pb_freq - freq_chan[i])).argmin()
    return m, pb_freq


def calc_conv_size(sub_a_term, imsize, support_cut_level, oversampling, max_support):
    """calc_conv_size (make_gridding_convolution_function.py:414-457)."""
    a = np.abs(sub_a_term)
    mx = a[np.unravel_index(np.argmax(a), a.shape)]
    cut = support_cut_level * mx
    assert np.min(a) < cut
    ix, iy = imsize[0] // 2, imsize[1] // 2
    while sub_a_term[ix, iy] > cut:
        ix += 1
        assert ix < imsize[0]
    sx = (int(0.5 + (ix - imsize[0] // 2) / oversampling[0]) + 1) * 2 + 1
    ix, iy = imsize[0] // 2, imsize[1] // 2
    while sub_a_term[ix, iy] > cut:
        iy += 1
        assert iy < imsize[1]
    sy = (int(0.5 + (iy - imsize[1] // 2) / oversampling[1]) + 1) * 2 + 1
    assert sx < max_support[0] and sy < max_support[1]
    s = max(sx, sy)
    return [s, s]


def make_gcf(freq_chan, n_pol, unique_ant_indx, baseline_ant, dish, blockage,
             cell_size_rad, image_size_padded, oversampling=(10, 10),
             max_support=(15, 15), support_cut_level=0.025,
             chan_tolerance_factor=0.005, casa=True):
    """The a_term branch of make_gridding_convolution_function
    (make_gridding_convolution_function.py:161-280, 361-412).  ``cell_size_rad``
    is already converted (cell[0] negated).  Returns dict with CONV_KERNEL,
    WEIGHT_CONV_KERNEL, SUPPORT, CF_BASELINE_MAP, CF_CHAN_MAP, CF_POL_MAP."""
    os_ = np.asarray(oversampling, dtype=int)
    max_support = np.asarray(max_support, dtype=int)
    npad = np.asarray(image_size_padded, dtype=int)
    cfb, pairs = cf_baseline_map(unique_ant_indx, baseline_ant, len(dish))
    cfc, pb_freq = cf_chan_map(freq_chan, chan_tolerance_factor)
    cfp = np.zeros(n_pol, dtype=int)
    resize = (max_support + 1) * os_
    pb_cell = np.asarray(cell_size_rad) * os_
    centre = npad // 2
    pats = []
    for ipower in (1, 2):
        p = airy_rorder(pb_freq, 1, dish, blockage, ipower, pb_cell, npad, centre, casa)
        bp = np.zeros((len(pairs), len(pb_freq), 1, npad[0], npad[1]))
        for k, (i, j) in enumerate(pairs):
            for f in range(len(pb_freq)):
                bp[k, f, 0] = p[i, f, 0] * p[j, f, 0]
        pats.append(np.real(np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(bp, axes=(3, 4)),
                                                        axes=(3, 4)), axes=(3, 4))))
    ck, wck = pats
    shape = ck.shape[0:3]
    support = np.zeros(shape + (2,), dtype=int)
    start = npad // 2 - resize // 2
    end = start + resize
    rck = np.zeros(shape + tuple(resize))
    rwck = np.zeros(shape + tuple(resize))
    import itertools
    for idx in itertools.product(*[range(s) for s in shape]):
        support[idx] = calc_conv_size(wck[idx], npad, support_cut_level, os_, max_support)
        emb = (support[idx] + 1) * os_
        es = resize // 2 - emb // 2
        ee = es + emb
        rck[idx] = ck[idx[0], idx[1], idx[2], start[0]:end[0], start[1]:end[1]]
        nf = np.real(np.sum(rck[idx[0], idx[1], idx[2], es[0]:ee[0], es[1]:ee[1]]) / (os_[0] * os_[1]))
        rck[idx] = rck[idx] / nf
        rwck[idx] = wck[idx[0], idx[1], idx[2], start[0]:end[0], start[1]:end[1]]
        nf = np.real(np.sum(rwck[idx[0], idx[1], idx[2], es[0]:ee[0], es[1]:ee[1]]) / (os_[0] * os_[1]))
        rwck[idx] = rwck[idx] / nf
    return dict(CONV_KERNEL=rck, WEIGHT_CONV_KERNEL=rwck, SUPPORT=support,
                CF_BASELINE_MAP=cfb, CF_CHAN_MAP=cfc, CF_POL_MAP=cfp,
                pb_freq=pb_freq, pairs=pairs)


def sin_world_to_pixel_offset(ra, dec, ra0, dec0, cell_rad):
    """Offset (pixels) of a direction from the reference pixel under
    RA---SIN/DEC--SIN with LONPOLE=180 (Calabretta & Greisen 2002, zenithal
    projection).  Replaces astropy's all_world2pix minus crpix, as used at
    make_gridding_convolution_function.py:339-347.  Angles in radians."""
    ra = np.asarray(ra, dtype=np.float64)
    dec = np.asarray(dec, dtype=np.float64)
    dra = ra - ra0
    l = np.cos(dec) * np.sin(dra)
    m = np.sin(dec) * np.cos(dec0) - np.cos(dec) * np.sin(dec0) * np.cos(dra)
    rad_to_deg = 180 / np.pi
    return (l * rad_to_deg) / (cell_rad[0] * rad_to_deg), (m * rad_to_deg) / (cell_rad[1] * rad_to_deg)


def phase_gradient(field_phase_dir, phase_center, cell_size_rad, image_size_padded,
                   oversampling, resize_conv_size):
    """make_phase_gradient (make_gridding_convolution_function.py:331-359).
    Returns [n_field, K_u, K_v] complex128."""
    npad = np.asarray(image_size_padded)
    os_ = np.asarray(oversampling)
    fpd = np.asarray(field_phase_dir, dtype=np.float64)
    dx, dy = sin_world_to_pixel_offset(fpd[:, 0], fpd[:, 1], phase_center[0],
                                       phase_center[1], cell_size_rad)
    pix_dist = np.stack([dx, dy], 1)
    pix = -pix_dist * 2 * np.pi / (npad * os_)
    K = np.asarray(resize_conv_size)
    c = K // 2
    x = np.arange(-c[0], K[0] - c[0])
    y = np.arange(-c[1], K[1] - c[1])
    xg, yg = np.meshgrid(x, y, indexing="ij")
    return np.moveaxis(np.exp(1j * (xg[:, :, None] * pix[:, 0] + yg[:, :, None] * pix[:, 1])), 2, 0)


def normalize_mosaic(image, sum_weight, ps_corr_image, pb, weight_pb, oversampling,
                     norm_type="flat_sky", pb_limit=0.2, single_precision=True):
    """_normalize, 'forward' direction (_normalize.py:17-89).  image, pb,
    weight_pb in [l, m, chan, pol]."""
    sw = np.array(sum_weight, dtype=np.float64, copy=True)
    sw[sw == 0] = 1
    if norm_type == "flat_noise":
        nimg = ps_corr_image[:, :, None, None] * pb
    elif norm_type == "flat_sky":
        nimg = ps_corr_image[:, :, None, None] * weight_pb
    else:
        nimg = ps_corr_image[:, :, None, None]
    n = np.array(image.shape[0:2])
    c = n // 2
    sx = np.sinc(np.arange(-c[0], n[0] - c[0]) / (n[0] * oversampling[0]))
    sy = np.sinc(np.arange(-c[1], n[1] - c[1]) / (n[1] * oversampling[1]))
    osc = np.dot(sx[:, None], sy[None, :])
    out = (image / sw[None, None]) / (osc[:, :, None, None] * nimg)
    if pb_limit > 0:
        out[pb < pb_limit] = 0.0
    if single_precision:
        out = out.astype(np.float32).astype(np.float64)
    return out
