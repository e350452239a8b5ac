"""The LDS transforms of single-precision grids (lds_fft.hip) against the
oracle's numpy transforms and against rocFFT on the same inputs: every radix
(2, 3, 4, 5, 7, 8, 11, 13), odd and non-square planes (ifftshift != fftshift).
Inverse (grid_to_image): complex64 and float32 grids, sum_weight and the
correction factors; forward (model_to_grid): padded odd and even models,
correction division, point sources."""
import numpy as np
import pytest

import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu


def _dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t if dtype is None else t.to(dtype)


@pytest.mark.parametrize("n_u,n_v,N_l,N_m", [
    (120, 120, 100, 100),      # 8 3 5
    (2457, 2457, 2048, 2048),  # 3^3 7 13: the C2 plane
    (1001, 77, 800, 60),       # 7 11 13 and 7 11, odd, not square
    (130, 96, 101, 81),        # 2 5 13 and 8 4 3
    (64, 1024, 64, 1000),      # 8 8 and 8 8 8 2
])
@pytest.mark.parametrize("real", [False, True])
def test_lds_transform_matches_oracle_and_rocfft(gpu, n_u, n_v, N_l, N_m, real):
    import torch
    import cngi_prototype_amd._native as nat
    rng = np.random.default_rng(n_u + n_v + real)
    n_planes = 2 if n_u * n_v > 1e6 else 3
    if real:
        g = rng.standard_normal((n_planes, 1, n_u, n_v))
    else:
        g = rng.standard_normal((n_planes, 1, n_u, n_v)) + 1j * rng.standard_normal(
            (n_planes, 1, n_u, n_v))
    g = g.astype(np.float32 if real else np.complex64)
    sw = rng.uniform(0.5, 2.0, size=(n_planes, 1))
    sw[0, 0] = 0.0   # a zero sum_weight divides by 1
    corr = (rng.uniform(0.5, 1.5, N_l), rng.uniform(0.5, 1.5, N_m))
    ref = oracle.grid_to_image(g.astype(np.complex128), [N_l, N_m])   # [N_l, N_m, planes, 1]
    swr = np.where(sw == 0.0, 1.0, sw)
    ref = ref / swr[None, None] / (corr[0][:, None] * corr[1][None, :])[:, :, None, None]
    ref = np.moveaxis(ref, [0, 1], [-2, -1])
    gd = _dev(g)
    out = {}
    for mode in (1, 0):
        nat.set_fft_mode(mode)
        out[mode] = nat.grid_to_image(gd, [N_l, N_m], sum_weight=_dev(sw), corr=corr).cpu().numpy()
    nat.set_fft_mode(1)
    # FP32 transforms: the image bound of the single-precision forms
    assert rel_err(out[1], ref) < 1e-5
    assert rel_err(out[0], ref) < 1e-5
    assert rel_err(out[1], out[0]) < 1e-5


def test_lds_transform_of_a_point_source(gpu):
    """A single grid cell at the uv centre is a flat image of value
    n_u n_v g / sw; one at (du, dv) from the centre a plane wave
    cos(2 pi (du l / n_u + dv m / n_v)) after the crop's shifts."""
    import torch
    import cngi_prototype_amd._native as nat
    n_u, n_v, N = 2457, 2457, 2048
    g = np.zeros((1, 1, n_u, n_v), dtype=np.complex64)
    g[0, 0, n_u // 2, n_v // 2] = 2.0
    img = nat.grid_to_image(_dev(g), [N, N]).cpu().numpy()[0, 0]
    assert np.allclose(img, 2.0, rtol=2e-6, atol=0)
    g[:] = 0
    du, dv = 5, -3
    g[0, 0, n_u // 2 + du, n_v // 2 + dv] = 1.0
    img = nat.grid_to_image(_dev(g), [N, N]).cpu().numpy()[0, 0]
    l = np.arange(N) - N // 2
    want = np.cos(2 * np.pi * (du * l[:, None] / n_u + dv * l[None, :] / n_v))
    assert np.abs(img - want).max() < 2e-5


@pytest.mark.parametrize("n_u,n_v,N_l,N_m", [
    (120, 120, 100, 100),      # 8 3 5, even model
    (63, 65, 50, 41),          # 7 9 and 5 13, odd grid and model
    (1001, 77, 800, 60),       # 7 11 13 and 7 11
    (64, 1024, 63, 1000),      # 8 8 and 8 8 8 2
    (2457, 2457, 2048, 2048),  # the C2 plane
])
def test_lds_model_to_grid_matches_oracle_and_rocfft(gpu, n_u, n_v, N_l, N_m):
    """The forward placement (pad, ifftshift, fft2, fftshift) of the LDS path
    against the oracle's numpy FFT and against rocFFT, with the correction
    division."""
    import cngi_prototype_amd._native as nat
    from cngi_prototype_amd.imaging._imaging_utils._gridding_convolutional_kernels import \
        cropped_correction_factors
    rng = np.random.default_rng(n_u * 7 + n_v)
    n_planes = 2 if n_u * n_v > 1e6 else 3
    model = rng.normal(size=(N_l, N_m, n_planes, 1))
    cl, cm = cropped_correction_factors([n_u, n_v], [N_l, N_m])
    ref = oracle.model_to_grid(model, [n_u, n_v])   # divides by the correction image
    md = _dev(np.moveaxis(model, [0, 1], [-2, -1]))
    out = {}
    for mode in (1, 0):
        nat.set_fft_mode(mode)
        out[mode] = nat.model_to_grid(md, [n_u, n_v], cl, cm, single=True).cpu().numpy()
    nat.set_fft_mode(1)
    assert out[1].shape == ref.shape
    assert rel_err(out[1], ref) < 2e-6
    assert rel_err(out[0], ref) < 2e-6
    assert rel_err(out[1], out[0]) < 2e-6


def test_lds_model_to_grid_of_a_point_source(gpu):
    """A unit pixel at the image centre is a flat grid of ones; one at
    (dl, dm) from the centre the plane wave exp(-2 pi i (du dl / n_u +
    dv dm / n_v)) in (du, dv) from the grid centre."""
    import cngi_prototype_amd._native as nat
    n_u, n_v, N = 2457, 2457, 2048
    m = np.zeros((1, 1, N, N))
    m[0, 0, N // 2, N // 2] = 1.0
    g = nat.model_to_grid(_dev(m), [n_u, n_v], single=True).cpu().numpy()[0, 0]
    assert np.allclose(g, 1.0, rtol=0, atol=1e-6)
    m[:] = 0
    dl, dm = 7, -4
    m[0, 0, N // 2 + dl, N // 2 + dm] = 1.0
    g = nat.model_to_grid(_dev(m), [n_u, n_v], single=True).cpu().numpy()[0, 0]
    du = np.arange(n_u) - n_u // 2
    dv = np.arange(n_v) - n_v // 2
    want = np.exp(-2j * np.pi * (du[:, None] * dl / n_u + dv[None, :] * dm / n_v))
    assert np.abs(g - want).max() < 2e-5


@pytest.mark.parametrize("n_u,n_v,N_l,N_m", [
    (2197, 64, 2000, 60),      # 13^3: no column split N1 <= N2 <= 128, transposes
    (8192, 40, 8000, 33),      # 64 x 128 split, 8 8 8 8 2 rows of 40 = 8 5
    (48, 3969, 40, 3500),      # rows of 63^2 = 3^4 7^2 (16 loads per thread)
    (36, 8192, 30, 8000),      # rows of 2^13 (32 loads per thread)
])
def test_lds_transforms_at_long_columns(gpu, n_u, n_v, N_l, N_m):
    """Column lengths at the ends of the four-step split and past it (the
    transpose path), and rows long enough for 16 and 32 loads per thread,
    both directions, against the oracle."""
    import cngi_prototype_amd._native as nat
    from cngi_prototype_amd.imaging._imaging_utils._gridding_convolutional_kernels import \
        cropped_correction_factors
    rng = np.random.default_rng(n_u + 3 * n_v)
    g = (rng.standard_normal((2, 1, n_u, n_v)) +
         1j * rng.standard_normal((2, 1, n_u, n_v))).astype(np.complex64)
    ref = np.moveaxis(oracle.grid_to_image(g.astype(np.complex128), [N_l, N_m]), [0, 1], [-2, -1])
    got = nat.grid_to_image(_dev(g), [N_l, N_m]).cpu().numpy()
    assert rel_err(got, ref) < 1e-5
    model = rng.normal(size=(N_l, N_m, 2, 1))
    cl, cm = cropped_correction_factors([n_u, n_v], [N_l, N_m])
    ref = oracle.model_to_grid(model, [n_u, n_v])
    got = nat.model_to_grid(_dev(np.moveaxis(model, [0, 1], [-2, -1])), [n_u, n_v], cl, cm,
                            single=True).cpu().numpy()
    assert rel_err(got, ref) < 2e-6
