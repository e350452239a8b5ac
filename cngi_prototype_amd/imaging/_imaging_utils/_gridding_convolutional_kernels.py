"""Prolate-spheroidal gridding kernel tables and grid-correction factors.

Schwab's rational approximation of the spheroidal function (M = 6, alpha = 1)
with the coefficient table and the two-branch evaluation of
ngcasa/imaging/_imaging_utils/_gridding_convolutional_kernels.py:101-148, so
that the oversampled kernel (:151-157) and the correction image (:87-93) are
the same FP64 numbers the reference produces.  These tables are set-up data
for the GPU gridder (a few hundred values per call).
"""
import numpy as np

# Schwab, "Optimal gridding of visibility data in radio interferometry"
# (Indirect Imaging, 1984): numerator and denominator coefficients for the
# intervals 0 <= nu < 0.75 and 0.75 <= nu <= 1.
_P = np.array([[8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1],
               [4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2]])
_Q = np.array([[1.0000000e0, 8.212018e-1, 2.078043e-1],
               [1.0000000e0, 9.599102e-1, 2.918724e-1]])


def prolate_spheroidal(nu):
    """Return (grdsf(nu), (1 - nu^2)*grdsf(nu)) for |nu| <= 1, zero beyond."""
    nu = np.abs(np.asarray(nu, dtype=np.float64))
    branch = (nu >= 0.75).astype(np.int64)
    nu_end = np.where(branch == 1, 1.0, 0.75)
    inside = nu <= 1.0
    branch = np.where(inside, branch, 0)
    nu_end = np.where(inside, nu_end, 0.0)
    d = nu ** 2 - nu_end ** 2
    num = _P[branch, 0].copy()
    for k in range(1, _P.shape[1]):
        num += _P[branch, k] * np.power(d, k)
    den = _Q[branch, 0].copy()
    for k in range(1, _Q.shape[1]):
        den += _Q[branch, k] * np.power(d, k)
    grdsf = np.zeros(nu.shape, dtype=np.float64)
    ok = den > 0.0
    grdsf[ok] = num[ok] / den[ok]
    grdsf[~inside] = 0.0
    return grdsf, (1 - nu ** 2) * grdsf


def prolate_spheroidal_kernel_1d(oversampling, support):
    """Oversampled half kernel, length oversampling*(support//2 + 1), sampled at
    nu = k/(support//2*oversampling) with a zero tail of one oversampling step."""
    half = support // 2
    nu = np.arange(oversampling * half) / (half * oversampling)
    kernel = np.zeros(oversampling * (half + 1))
    kernel[:oversampling * half] = prolate_spheroidal(nu)[1]
    return kernel


def correction_factors(n_uv):
    """Separable grid-correction factors grdsf(|2*(k - n//2)/n|) per padded
    axis; their outer product is the reference's correcting image."""
    out = []
    for n in (int(n_uv[0]), int(n_uv[1])):
        coords = (np.arange(n) - n // 2) / n
        out.append(prolate_spheroidal(np.abs(2.0 * coords))[0])
    return out


def cropped_correction_factors(n_uv, image_size):
    """correction_factors cropped to the image (_remove_padding.py:20-30)."""
    cu, cv = correction_factors(n_uv)
    su = int(n_uv[0]) // 2 - int(image_size[0]) // 2
    sv = int(n_uv[1]) // 2 - int(image_size[1]) // 2
    return cu[su:su + int(image_size[0])], cv[sv:sv + int(image_size[1])]
