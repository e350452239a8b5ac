"""Elliptical Gaussian fit of the PSF main lobe (cngi/image/fit_gaussian.py:25-231).

Fits the centre ``npix_window`` pixels of every (time, chan, pol) plane: the
window is resampled onto a ``sampling`` grid with a bivariate spline
(scipy's RectBivariateSpline, the routine behind the reference's
interpn(method='splinef2d')), samples below ``cutoff`` are masked, and
(width_x, width_y, rotation) are found by minimising the squared residual
(scipy.optimize.minimize, BFGS, start [2.5, 2.5, 0]).  The result is
[major FWHM, minor FWHM (arcsec), position angle (deg)] per plane, stored as
``beam_set_name`` with dims (time, chan, pol, elps_index).  Host-side: a 9x9
window per plane.
"""
import numpy as np


def _gaussian(params, x, y):
    """The elliptical Gaussian at the sample offsets x, y."""
    width_x, width_y, rotation = params
    rot = np.deg2rad(90 - rotation)
    c, s = np.cos(rot), np.sin(rot)
    xp = x * c - y * s
    yp = x * s + y * c
    return np.exp(-((xp / width_x) ** 2 + (yp / width_y) ** 2) / 2.0)


def _chi2(params, pm, xm, ym):
    """Squared residual over the unmasked samples pm at offsets xm, ym."""
    return np.sum((_gaussian(params, xm, ym) - pm) ** 2)


def fit_planes(img, npix_window, sampling, cutoff, delta):
    """img [l, m, time, chan, pol] -> [time, chan, pol, 3].

    The sample offsets are built once, and each plane's unmasked samples once
    (the objective then evaluates the Gaussian only there: the same values
    summed in the same order as evaluating everywhere and masking).  Planes
    that are bit-identical (e.g. the pols of a PSF with equal weights) share
    one fit."""
    from scipy import optimize
    from scipy.interpolate import RectBivariateSpline
    npix_window = np.asarray(npix_window)
    sampling = np.asarray(sampling)
    out = np.zeros(img.shape[2:5] + (3,))
    size = np.array(img.shape[0:2])
    c = size // 2
    s0 = c - npix_window // 2
    e0 = c + npix_window // 2 + 1
    win = img[s0[0]:e0[0], s0[1]:e0[1]]
    d0 = np.arange(0, npix_window[0]) * np.abs(delta[0])
    d1 = np.arange(0, npix_window[1]) * np.abs(delta[1])
    i0 = np.linspace(0, npix_window[0] - 1, sampling[0]) * np.abs(delta[0])
    i1 = np.linspace(0, npix_window[1] - 1, sampling[1]) * np.abs(delta[1])
    xp, yp = np.meshgrid(i0, i1, indexing="ij")
    xp, yp = np.ravel(xp), np.ravel(yp)
    half = sampling // 2
    gx, gy = np.meshgrid(np.arange(half[0] * 2 + 1) - half[0],
                         np.arange(half[1] * 2 + 1) - half[1], indexing="ij")
    gx, gy = np.ravel(gx), np.ravel(gy)
    done = {}
    for t in range(win.shape[2]):
        for ch in range(win.shape[3]):
            for p in range(win.shape[4]):
                plane = np.ascontiguousarray(win[:, :, t, ch, p], dtype=np.float64)
                key = plane.tobytes()
                if key in done:
                    out[t, ch, p] = done[key]
                    continue
                if not np.all(np.isfinite(plane)):
                    out[t, ch, p] = np.nan
                    done[key] = out[t, ch, p].copy()
                    continue
                spl = RectBivariateSpline(d0, d1, plane)
                vals = spl.ev(xp, yp)
                interp = np.reshape(vals, [sampling[1], sampling[0]]).T
                interp[interp < cutoff] = np.nan
                pr = np.ravel(interp)
                mask = ~np.isnan(pr)
                res = optimize.minimize(_chi2, [2.5, 2.5, 0.0],
                                        args=(pr[mask], gx[mask], gy[mask]))
                rx = res.x
                phi = rx[2] - 90.0
                if phi < -90.0:
                    phi += 180.0
                out[t, ch, p, 0] = np.max(np.abs(rx[0:2])) * np.abs(delta[0]) * 2.355 / (
                    sampling[0] / npix_window[0])
                out[t, ch, p, 1] = np.min(np.abs(rx[0:2])) * np.abs(delta[1]) * 2.355 / (
                    sampling[1] / npix_window[1])
                out[t, ch, p, 2] = -phi
                done[key] = out[t, ch, p].copy()
    return out


def fit_gaussian(xds, dv='PSF', beam_set_name='RESTORE_PARMS', npix_window=[9, 9],
                 sampling=[9, 9], cutoff=0.35):
    from ..xds import DataArray
    from .._utils._io import copy_structure
    _xds = copy_structure(xds)
    l = np.asarray(_xds.coords['l'].values)
    m = np.asarray(_xds.coords['m'].values)
    delta = np.array([l[1] - l[0], m[1] - m[0]]) * 3600 * 180 / np.pi
    img = np.asarray(_xds[dv].values)
    parms = fit_planes(img, npix_window, sampling, cutoff, delta)
    _xds[beam_set_name] = DataArray(parms, dims=['time', 'chan', 'pol', 'elps_index'])
    return _xds
