"""Visibility prediction on the GPU (ngc_model_to_grid, ngc_standard_degrid,
ngc_aperture_degrid, predict_modelvis_image) against the oracle restatement, which the CPU tests
pin by adjointness with the gridder, closed forms and point-source
visibilities (the reference itself has no implementation)."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from helpers import small_case

pytestmark = pytest.mark.gpu

ARCSEC = np.pi / (180 * 3600)


def _t(a, dtype=None):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(a)).cuda()
    return t.to(dtype) if dtype is not None else t


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("chan_mode", ["cube", "continuum"])
@pytest.mark.parametrize("normalize", [True, False])
def test_degrid_matches_oracle(gpu, single, chan_mode, normalize):
    import torch
    import cngi_prototype_amd._native as nat
    c = small_case(n_time=20, n_chan=3, n_pol=2)
    npad = c["npad"]
    rng = np.random.default_rng(4)
    n_ic = 3 if chan_mode == "cube" else 1
    G = rng.normal(size=(n_ic, 2, npad, npad)) + 1j * rng.normal(size=(n_ic, 2, npad, npad))
    if single:
        G = G.astype(np.complex64).astype(np.complex128)
    cgk = oracle.ps_kernel_1d(100, 7)
    ref = oracle.standard_degrid(G, c["uvw"], c["freq"], cgk, [npad, npad], c["delta"], 7, 100,
                                 chan_mode=chan_mode, normalize=normalize, n_pol=2)
    chan_map = np.arange(3) if chan_mode == "cube" else np.zeros(3, dtype=int)
    got = nat.standard_degrid(_t(c["uvw"]), _t(G, torch.complex64 if single else torch.complex128),
                              c["freq"], chan_map, np.arange(2), cgk, c["delta"], 7, 100,
                              normalize=normalize, single=single)
    got = got.cpu().numpy()
    assert np.count_nonzero(ref) > 0
    # zero exactly where the oracle places nothing
    assert np.array_equal(got == 0, ref == 0)
    assert rel_err(got, ref) < (1e-6 if single else 1e-12)


@pytest.mark.parametrize("n_pol,support,chan_mode", [
    (1, 7, "cube"), (2, 8, "continuum"), (4, 7, "cube"), (4, 5, "continuum"), (2, 3, "cube")])
@pytest.mark.parametrize("vis_single", [True, False])
def test_tiled_degrid_equals_the_gather(gpu, n_pol, support, chan_mode, vis_single):
    """The LDS-tile degridder against the per-sample gather on the same
    complex64 grid: the same bits for every sample, over a partial last time
    block, samples off a grid too small for the coverage (zeros), NaN uvw
    rows, odd grid lengths and both channel modes; and against the oracle."""
    import torch
    import cngi_prototype_amd._native as nat
    c = small_case(n_time=130, n_chan=3, n_pol=n_pol, image_size=96)
    n_u, n_v = 77, 91            # smaller than the coverage: edge tiles, off-grid samples
    uvw = c["uvw"].copy()
    uvw[5, :7] = np.nan
    rng = np.random.default_rng(60 + 7 * n_pol + support)
    n_ic = 3 if chan_mode == "cube" else 1
    G = (rng.normal(size=(n_ic, n_pol, n_u, n_v)) +
         1j * rng.normal(size=(n_ic, n_pol, n_u, n_v))).astype(np.complex64)
    chan_map = np.arange(3) if chan_mode == "cube" else np.zeros(3, dtype=int)
    cgk = oracle.ps_kernel_1d(100, support)
    gd = _t(G)
    out = {}
    for mode in (1, 0):
        nat.set_degrid_mode(mode)
        out[mode] = nat.standard_degrid(_t(uvw), gd, c["freq"], chan_map, np.arange(n_pol), cgk,
                                        c["delta"], support, 100,
                                        single=vis_single).cpu().numpy()
    nat.set_degrid_mode(1)
    assert np.array_equal(out[1], out[0])
    ref = oracle.standard_degrid(G.astype(np.complex128), uvw, c["freq"], cgk, [n_u, n_v],
                                 c["delta"], support, 100, chan_mode=chan_mode, n_pol=n_pol)
    assert np.count_nonzero(ref == 0) > 0 and np.count_nonzero(ref) > 0
    assert np.array_equal(out[1] == 0, ref == 0)
    assert rel_err(out[1], ref) < 1e-6


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("n_pol,support", [(1, 7), (4, 7), (4, 9), (1, 13), (2, 16)])
def test_degrid_shapes_match_oracle(gpu, single, n_pol, support):
    """130 integrations (two full 64-lane time blocks and a partial third),
    1, 2 or 4 pols (the unrolled pol paths) and supports 7 to 16 (the
    support-8 and support-16 kernel forms) against the oracle.  Supports above
    9 use oversampling 50, which keeps the kernel table within the 512 entries
    the degridder holds (os * (S//2 + 1))."""
    import torch
    import cngi_prototype_amd._native as nat
    c = small_case(n_time=130, n_chan=2, n_pol=n_pol, image_size=96)
    npad = c["npad"]
    rng = np.random.default_rng(40 + support)
    G = rng.normal(size=(2, n_pol, npad, npad)) + 1j * rng.normal(size=(2, n_pol, npad, npad))
    if single:
        G = G.astype(np.complex64).astype(np.complex128)
    os_ = 100 if support <= 9 else 50
    cgk = oracle.ps_kernel_1d(os_, support)
    ref = oracle.standard_degrid(G, c["uvw"], c["freq"], cgk, [npad, npad], c["delta"], support,
                                 os_, n_pol=n_pol)
    got = nat.standard_degrid(_t(c["uvw"]), _t(G, torch.complex64 if single else torch.complex128),
                              c["freq"], np.arange(2), np.arange(n_pol), cgk, c["delta"], support,
                              os_, single=single).cpu().numpy()
    assert np.count_nonzero(ref) > 0
    assert np.array_equal(got == 0, ref == 0)
    assert rel_err(got, ref) < (1e-6 if single else 1e-12)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("n_uv", [[61, 60], [60, 61], [64, 64]])
def test_model_to_grid_matches_oracle(gpu, single, n_uv):
    """Odd and even grid lengths on both axes (fftshift and the padding
    offsets differ between them)."""
    import cngi_prototype_amd._native as nat
    from cngi_prototype_amd.imaging._imaging_utils._gridding_convolutional_kernels import \
        cropped_correction_factors
    rng = np.random.default_rng(8)
    N = 50
    model = rng.normal(size=(N, N, 2, 2))
    ref = oracle.model_to_grid(model, n_uv)
    cl, cm = cropped_correction_factors(n_uv, [N, N])
    got = nat.model_to_grid(_t(np.moveaxis(model, [0, 1], [-2, -1])), n_uv, cl, cm, single=single)
    assert rel_err(got.cpu().numpy(), ref) < (2e-6 if single else 1e-12)


def test_predict_modelvis_image_point_sources(gpu):
    """A cube model of two point sources (Jy/pixel) predicts their
    visibilities: exactly for the source at the phase centre, within the
    spheroidal kernel's aliasing error for the other; incremental adds; the
    vis dataset is not modified."""
    import torch
    from cngi_prototype_amd.imaging import predict_modelvis_image
    from cngi_prototype_amd.xds import Dataset
    rng = np.random.default_rng(11)
    n, cell = 64, 20 * ARCSEC
    freq = np.array([1.4e9, 1.45e9])
    lam = 299792458.0 / freq
    T, B = 30, 40
    uvw = np.zeros((T, B, 3))
    uvw[..., :2] = rng.uniform(-0.45 / cell, 0.45 / cell, size=(T, B, 2)) * lam[0] * 0.95
    vis_ds = Dataset(data_vars={"UVW": (["time", "baseline", "uvw_index"], _t(uvw))},
                     coords={"chan": (["chan"], freq), "pol": (["pol"], np.array([9, 12]))})
    model = np.zeros((n, n, 1, 2, 2))
    model[n // 2, n // 2, 0, :, :] = 2.0            # 2 Jy at the phase centre
    model[n // 2 - 5, n // 2 - 9, 0, :, 1] = 0.5    # 0.5 Jy in the second pol only
    img = Dataset(data_vars={"MODEL": (["l", "m", "time", "chan", "pol"], model)})
    gp = {"image_size": [n, n], "cell_size": [20.0, 20.0], "fft_padding": 1.2,
          "chan_mode": "cube"}
    out = predict_modelvis_image(img, vis_ds, gp, {})
    v = out["MODEL"].data.cpu().numpy()
    assert v.shape == (T, B, 2, 2) and "MODEL" not in vis_ds.data_vars
    u = uvw[:, :, 0, None] / lam
    w = uvw[:, :, 1, None] / lam
    # pixel (n//2 - 5, n//2 - 9) is (l, m) = (5, -9) cells: cell_l is negative
    exp2 = 2.0 + 0.5 * np.exp(2j * np.pi * (u * 5 * cell + w * (-9) * cell))
    assert np.abs(v[..., 0] - 2.0).max() < 2e-6
    assert np.abs(v[..., 1] - exp2).max() < 1e-2
    out2 = predict_modelvis_image(img, out, gp, {"incremental": True})
    assert np.allclose(out2["MODEL"].data.cpu().numpy(), 2 * v, rtol=1e-12, atol=1e-12)


def _ap_case(seed, pg_form="random", cf_per_pol=False):
    """Two convolution functions per pol function (supports 5 and 7, or 5/7
    and 3/7 with a function per pol), three fields, random real CK, a row on
    field -1.  pg_form "random": unit phase gradients of random phase (no
    u/v factorisation, the general kernel); "separable": the gcf builder's
    exp(i((x - K_u/2) px + (y - K_v/2) py)) per field (the separable kernel)."""
    rng = np.random.default_rng(seed)
    c = small_case(n_time=16, n_chan=3, n_pol=2)
    T, B = c["uvw"].shape[:2]
    os_, K = [4, 5], [40, 48]
    n_cfp = 2 if cf_per_pol else 1
    ck = rng.uniform(0.1, 1.0, size=(2, 1, n_cfp, K[0], K[1]))
    sup = np.array([[5, 5], [7, 7], [3, 7], [7, 5]])[:2 * n_cfp].reshape(2, n_cfp, 2)
    sup = np.ascontiguousarray(np.swapaxes(sup.reshape(n_cfp, 2, 2), 0, 1)).reshape(2, 1, n_cfp, 2)
    if pg_form == "random":
        pg = np.exp(1j * rng.uniform(-np.pi, np.pi, size=(3, K[0], K[1])))
    else:
        x = np.arange(K[0])[:, None] - K[0] // 2
        y = np.arange(K[1])[None, :] - K[1] // 2
        slopes = rng.uniform(-0.3, 0.3, size=(3, 2))
        pg = np.exp(1j * (x[None] * slopes[:, 0, None, None] + y[None] * slopes[:, 1, None, None]))
    field = rng.integers(0, 3, size=(T, B)) + 10
    field[2, 3] = -1
    maps = (rng.integers(0, 2, size=B), np.zeros(3, int), np.arange(2) if cf_per_pol else np.zeros(2, int))
    return c, ck, sup, pg, field, [10, 11, 12], maps, os_


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("form", ["random", "separable", "separable_cf_per_pol"])
def test_aperture_degrid_matches_oracle(gpu, single, normalize, form):
    """Both kernels against the oracle: the general one (a phase gradient
    that does not factor) and the separable one, with pols sharing one
    convolution function or each with its own; the profile phases show
    which kernel ran."""
    import torch
    import cngi_prototype_amd._native as nat
    c, ck, sup, pg, field, fid, maps, os_ = _ap_case(6, pg_form=form.split("_")[0],
                                                     cf_per_pol=form.endswith("per_pol"))
    npad = c["npad"]
    rng = np.random.default_rng(9)
    G = rng.normal(size=(3, 2, npad, npad)) + 1j * rng.normal(size=(3, 2, npad, npad))
    if single:
        G = G.astype(np.complex64).astype(np.complex128)
    ref = oracle.aperture_degrid(G, c["uvw"], field, fid, *maps, ck, sup, pg, c["freq"],
                                 [npad, npad], c["delta"], os_, normalize=normalize)
    nat.profile_enable(True)
    try:
        got = nat.aperture_degrid(_t(c["uvw"]), _t(field, torch.int32), fid,
                                  _t(G, torch.complex64 if single else torch.complex128),
                                  c["freq"], np.arange(3), np.arange(2), *maps, _t(ck), sup,
                                  _t(pg), c["delta"], os_, normalize=normalize, single=single)
        phases = [name for name, _ in nat.profile_read()]
    finally:
        nat.profile_enable(False)
    assert ("aperture_degrid_tables" in phases) == (form != "random"), phases
    got = got.cpu().numpy()
    assert np.count_nonzero(ref) > 0
    assert np.array_equal(got == 0, ref == 0)
    assert rel_err(got, ref) < (1e-6 if single else 1e-12)


@pytest.mark.parametrize("n_pol", [1, 4])
def test_aperture_degrid_pols_and_time_blocks(gpu, n_pol):
    """The separable A-term kernel with 1 pol and with 4 pols sharing one
    convolution function (the NP = 4 pass), over 130 integrations (a partial
    last time block), against the oracle."""
    import torch
    import cngi_prototype_amd._native as nat
    rng = np.random.default_rng(50 + n_pol)
    c = small_case(n_time=130, n_chan=2, n_pol=n_pol, image_size=96)
    T, B = c["uvw"].shape[:2]
    os_, K = [4, 5], [40, 48]
    ck = rng.uniform(0.1, 1.0, size=(2, 1, 1, K[0], K[1]))
    sup = np.array([[5, 5], [7, 7]]).reshape(2, 1, 1, 2)
    x = np.arange(K[0])[:, None] - K[0] // 2
    y = np.arange(K[1])[None, :] - K[1] // 2
    slopes = rng.uniform(-0.3, 0.3, size=(3, 2))
    pg = np.exp(1j * (x[None] * slopes[:, 0, None, None] + y[None] * slopes[:, 1, None, None]))
    field = rng.integers(0, 3, size=(T, B)) + 10
    field[5, 2] = -1
    maps = (rng.integers(0, 2, size=B), np.zeros(2, int), np.zeros(n_pol, int))
    npad = c["npad"]
    G = rng.normal(size=(2, n_pol, npad, npad)) + 1j * rng.normal(size=(2, n_pol, npad, npad))
    ref = oracle.aperture_degrid(G, c["uvw"], field, [10, 11, 12], *maps, ck, sup, pg, c["freq"],
                                 [npad, npad], c["delta"], os_)
    nat.profile_enable(True)
    try:
        got = nat.aperture_degrid(_t(c["uvw"]), _t(field, torch.int32), [10, 11, 12], _t(G),
                                  c["freq"], np.arange(2), np.arange(n_pol), *maps, _t(ck), sup,
                                  _t(pg), c["delta"], os_, single=False)
        phases = [name for name, _ in nat.profile_read()]
    finally:
        nat.profile_enable(False)
    assert "aperture_degrid_tables" in phases, phases
    got = got.cpu().numpy()
    assert np.count_nonzero(ref) > 0
    assert np.array_equal(got == 0, ref == 0)
    assert rel_err(got, ref) < 1e-12


def test_aperture_degrid_unknown_field_raises(gpu):
    import torch
    import cngi_prototype_amd._native as nat
    c, ck, sup, pg, field, fid, maps, os_ = _ap_case(6)
    npad = c["npad"]
    field = field.copy()
    field[1, 1] = 99
    G = np.ones((3, 2, npad, npad), dtype=np.complex128)
    with pytest.raises(IndexError):
        nat.aperture_degrid(_t(c["uvw"]), _t(field, torch.int32), fid, _t(G), c["freq"],
                            np.arange(3), np.arange(2), *maps, _t(ck), sup, _t(pg), c["delta"],
                            os_, single=False)


@pytest.mark.parametrize("pg_form", ["random", "separable"])
def test_predict_modelvis_image_with_gcf(gpu, pg_form):
    """predict_modelvis_image with a gcf dataset: the model, transformed
    without the spheroidal correction, is degridded with the A-term kernels,
    fields and maps of the gcf, as the oracle degrids the same grid (both
    A-term kernels: a phase gradient that does not factor, and the gcf
    builder's form)."""
    import torch
    import cngi_prototype_amd._native as nat
    from cngi_prototype_amd.imaging import predict_modelvis_image
    from cngi_prototype_amd.xds import Dataset
    c, ck, sup, pg, field, fid, maps, os_ = _ap_case(12, pg_form=pg_form)
    n, npad = c["image_size"], c["npad"]
    rng = np.random.default_rng(13)
    T, B = c["uvw"].shape[:2]
    vis_ds = Dataset(data_vars={"UVW": (["time", "baseline", "uvw_index"], _t(c["uvw"])),
                                "FIELD_ID": (["time", "baseline"], field)},
                     coords={"chan": (["chan"], c["freq"]), "pol": (["pol"], np.array([9, 12]))})
    gcf = Dataset(data_vars={"CONV_KERNEL": (["cf_b", "cf_c", "cf_p", "k_u", "k_v"], ck),
                             "SUPPORT": (["cf_b", "cf_c", "cf_p", "d"], sup),
                             "PHASE_GRADIENT": (["field_id", "k_u", "k_v"], pg),
                             "CF_BASELINE_MAP": (["baseline"], maps[0]),
                             "CF_CHAN_MAP": (["chan"], maps[1]),
                             "CF_POL_MAP": (["pol"], maps[2])},
                  coords={"field_id": (["field_id"], np.array(fid))},
                  attrs={"oversampling": np.array(os_)})
    model = np.zeros((n, n, 1, 3, 2))
    for _ in range(6):
        model[rng.integers(n // 4, 3 * n // 4), rng.integers(n // 4, 3 * n // 4), 0, :, :] += \
            rng.uniform(0.1, 2.0)
    img = Dataset(data_vars={"MODEL": (["l", "m", "time", "chan", "pol"], model)})
    cell = c["cell_arcsec"]
    gp = {"image_size": [n, n], "cell_size": [cell, cell], "fft_padding": 1.2,
          "chan_mode": "cube"}
    out = predict_modelvis_image(img, vis_ds, gp, {}, gcf_dataset=gcf)
    v = out["MODEL"].data.cpu().numpy()
    m = _t(np.ascontiguousarray(np.moveaxis(model[:, :, 0], [0, 1], [-2, -1])))
    G = nat.model_to_grid(m, [npad, npad], single=False).cpu().numpy()
    ref = oracle.aperture_degrid(G, c["uvw"], field, fid, *maps, ck, sup, pg, c["freq"],
                                 [npad, npad], c["delta"], os_)
    assert np.count_nonzero(ref) > 0
    assert rel_err(v, ref) < 1e-12


def test_builder_phase_gradient_takes_the_tap_parallel_kernel(gpu):
    """The phase gradient that ngc_phase_gradient builds (the gcf's
    exp(i(x px + y py))) factors into its u and v parts to rounding, so A-term
    prediction with a built gcf runs the tap-parallel kernel, and that kernel
    matches the oracle on it."""
    import torch
    import cngi_prototype_amd._native as nat
    c, ck, sup, _, field, fid, maps, os_ = _ap_case(21)
    K = ck.shape[-2:]
    rng = np.random.default_rng(22)
    pix = rng.uniform(-400.0, 400.0, size=(3, 2))
    pg_d = nat.phase_gradient(pix, K)
    pg = pg_d.cpu().numpy()
    cu, cv = K[0] // 2, K[1] // 2
    fac = pg[:, :, cv][:, :, None] * pg[:, cu, :][:, None, :]
    assert np.abs(fac - pg).max() < 1e-12
    npad = c["npad"]
    G = rng.normal(size=(3, 2, npad, npad)) + 1j * rng.normal(size=(3, 2, npad, npad))
    nat.profile_enable(True)
    try:
        got = nat.aperture_degrid(_t(c["uvw"]), _t(field, torch.int32), fid, _t(G), c["freq"],
                                  np.arange(3), np.arange(2), *maps, _t(ck), sup, pg_d,
                                  c["delta"], os_, single=False)
        phases = [name for name, _ in nat.profile_read()]
    finally:
        nat.profile_enable(False)
    assert "aperture_degrid_tables" in phases, phases
    ref = oracle.aperture_degrid(G, c["uvw"], field, fid, *maps, ck, sup, pg, c["freq"],
                                 [npad, npad], c["delta"], os_)
    assert np.count_nonzero(ref) > 0
    # the factors round the phase as x px and y py apart, not x px + y py:
    # about 1e-16 |phase| per tap, amplified where the steep ramps of these
    # offsets (up to 400 pixels) make the tap sums cancel
    assert rel_err(got.cpu().numpy(), ref) < 1e-11
