"""The drop-in API end to end (make_imaging_weight, make_psf, make_image,
make_grid) against the oracle chain on the same host inputs."""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from cngi_prototype_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def _mxds(n_pol=2, seed=1, natural=False):
    cfg = syn.scaled(syn.CONFIGS["C1"], n_time=24, n_chan=3, n_pol=n_pol, n_src=4,
                     flag_frac=0.1, seed=seed)
    mxds = syn.make_mxds(cfg)
    if natural:
        # the imaging functions read the data group's imaging_weight, as the
        # reference's do (make_psf.py:104)
        from cngi_prototype_amd.imaging import make_imaging_weight
        mxds = make_imaging_weight(mxds, {"weighting": "natural"}, {}, {"xds": "xds0"})
    return cfg, mxds


def _grid_parms(cfg, N=96, chan_mode="cube", precision=None):
    gp = {"image_size": [N, N], "cell_size": [cfg.cell_arcsec * 4, cfg.cell_arcsec * 4],
          "fft_padding": 1.2, "chan_mode": chan_mode, "phase_center": syn.phase_center(cfg)}
    if precision:
        gp["precision"] = precision
    return gp


def _oracle_inputs(mxds, weight_var):
    xds = mxds.attrs["xds0"]
    return (np.asarray(xds.DATA.values), np.asarray(xds.UVW.values),
            np.asarray(xds[weight_var].values), np.asarray(xds.coords["chan"].values))


@pytest.mark.parametrize("weighting,n_pol", [("natural", 2), ("briggs", 2), ("uniform", 2),
                                             ("briggs", 4), ("uniform", 4)])
def test_make_imaging_weight(gpu, weighting, n_pol):
    """n_pol = 4 goes through the full_pol opt-in (the C2 workload's Briggs
    chain); without it the reference's assert refuses more than 2 pols."""
    from cngi_prototype_amd.imaging import make_imaging_weight
    cfg, mxds = _mxds(n_pol=n_pol)
    gp = _grid_parms(cfg)
    iparms = {"weighting": weighting, "robust": 0.5}
    if n_pol > 2:
        with pytest.raises(AssertionError, match="Full polarization"):
            make_imaging_weight(mxds, dict(iparms), dict(gp), {"xds": "xds0"})
        iparms["full_pol"] = True
    out = make_imaging_weight(mxds, iparms, dict(gp), {"xds": "xds0"})
    xds = out.attrs["xds0"]
    dg = xds.attrs["data_groups"][0]["0"]
    if weighting == "natural":
        assert dg["imaging_weight"] == "DATA_WEIGHT"
        return
    iw = np.asarray(xds[dg["imaging_weight"]].values)
    _, uvw, w, fr = _oracle_inputs(mxds, "DATA_WEIGHT")
    cell = np.array([-gp["cell_size"][0], gp["cell_size"][1]]) * syn.ARCSEC
    N = gp["image_size"][0]
    g, sw = oracle.standard_grid(None, uvw, w, fr, np.ones(1), [N, N], cell, 1, 0, do_psf=True,
                                 do_imaging_weight=True)
    gi = np.moveaxis(g, [0, 1], [-2, -1])
    f = oracle.briggs_factors(gi, sw, weighting, 0.5)
    ref = oracle.imaging_weight_degrid(gi, uvw, w, f, fr, [N, N], cell)
    both = np.isnan(iw) & np.isnan(ref)
    assert np.array_equal(np.isnan(iw), np.isnan(ref))
    assert rel_err(np.where(both, 0, iw), np.where(both, 0, ref)) < 1e-12
    # the input dataset is not modified
    assert "imaging_weight" not in mxds.attrs["xds0"].attrs["data_groups"][0]["0"]


@pytest.mark.parametrize("chan_mode", ["cube", "continuum"])
@pytest.mark.parametrize("precision,tol", [(None, 1e-10), ("fp32_kahan", 1e-5)])
def test_make_psf_and_make_image(gpu, chan_mode, precision, tol):
    from cngi_prototype_amd.imaging import make_image, make_psf
    from cngi_prototype_amd.xds import Dataset
    cfg, mxds = _mxds(natural=True)
    gp = _grid_parms(cfg, chan_mode=chan_mode, precision=precision)
    img = make_psf(mxds, Dataset(), gp, {"xds": "xds0"}, {})
    img = make_image(mxds, img, gp, {"xds": "xds0"}, {})
    vis, uvw, w, fr = _oracle_inputs(mxds, "DATA_WEIGHT")
    N = gp["image_size"][0]
    npad = int(1.2 * N)
    cell = np.array([-gp["cell_size"][0], gp["cell_size"][1]]) * syn.ARCSEC
    cgk = oracle.ps_kernel_1d(100, 7)
    if precision == "fp32_kahan":
        vis = vis.astype(np.complex64).astype(np.complex128)
        w = w.astype(np.float32).astype(np.float64)
    for psf, var, swvar in ((True, "PSF", "PSF_SUM_WEIGHT"), (False, "IMAGE", "SUM_WEIGHT")):
        g, sw = oracle.standard_grid(None if psf else vis, uvw, w, fr, cgk, [npad, npad], cell,
                                     7, 100, do_psf=psf, chan_mode=chan_mode)
        ref = oracle.correct_standard_image(oracle.grid_to_image(g, [N, N]), sw, [npad, npad],
                                            [N, N])
        got = np.asarray(img[var].values)
        assert got.shape == (N, N, 1, g.shape[0], 2)
        assert rel_err(got[:, :, 0], ref) < tol, var
        assert rel_err(np.asarray(img[swvar].values)[0], sw) < 1e-12, swvar
    g0 = oracle.prolate_spheroidal_function(np.array([0.0]))[0][0]
    psf = np.asarray(img["PSF"].values)
    assert np.allclose(psf[N // 2, N // 2], 1 / g0 ** 2, rtol=1e-5)
    assert "PSF_FIT" in img
    assert np.asarray(img["PSF_FIT"].values).shape[-1] == 3
    # sky coordinates of the image
    assert np.asarray(img.coords["right_ascension"].values).shape == (N, N)


def test_make_grid(gpu):
    from cngi_prototype_amd.imaging import make_grid
    from cngi_prototype_amd.xds import Dataset
    cfg, mxds = _mxds(natural=True)
    gp = _grid_parms(cfg)
    out = make_grid(mxds, Dataset(), gp, {"xds": "xds0"}, {})
    vis, uvw, w, fr = _oracle_inputs(mxds, "DATA_WEIGHT")
    npad = int(1.2 * gp["image_size"][0])
    cell = np.array([-gp["cell_size"][0], gp["cell_size"][1]]) * syn.ARCSEC
    g, sw = oracle.standard_grid(vis, uvw, w, fr, oracle.ps_kernel_1d(100, 7), [npad, npad],
                                 cell, 7, 100, do_psf=False)
    got = np.asarray(out["GRID"].values)[:, :, 0]
    assert rel_err(got, np.moveaxis(g, [0, 1], [-2, -1])) < 1e-12
    assert rel_err(np.asarray(out["SUM_WEIGHT"].values)[0], sw) < 1e-12


def test_make_image_from_a_vis_zarr_store(gpu, tmp_path):
    """The imaging chain on a multi-dataset read back from a vis.zarr store
    (write_vis -> read_vis, zstd chunks over time and channel) gives the
    image of the in-memory multi-dataset (to the FP64 atomics' summation
    order)."""
    from cngi_prototype_amd.dio import read_vis, write_vis
    from cngi_prototype_amd.imaging import make_image, make_imaging_weight, make_psf
    from cngi_prototype_amd.xds import Dataset
    cfg, mxds = _mxds()
    store = str(tmp_path / "c1slice.vis.zarr")
    write_vis(mxds, store, chunks_on_disk={"time": 5, "chan": 2})
    back = read_vis(store)
    gp = _grid_parms(cfg)
    imgs = []
    for m in (mxds, back):
        m = make_imaging_weight(m, {"weighting": "briggs", "robust": 0.5}, dict(gp),
                                {"xds": "xds0"})
        img = make_psf(m, Dataset(), dict(gp), {"xds": "xds0"}, {})
        imgs.append(make_image(m, img, dict(gp), {"xds": "xds0"}, {}))
    for var in ("IMAGE", "PSF", "SUM_WEIGHT"):
        assert rel_err(np.asarray(imgs[1][var].values), np.asarray(imgs[0][var].values)) < 1e-12, var
