"""Host-side logic of the drop-in API (CPU only): parameter normalisation,
data-group resolution, the natural-weighting path, the PSF Gaussian fit and
the synthetic workloads' invariants."""
import copy

import numpy as np
import pytest

from cngi_prototype_amd import synthetic as syn
from cngi_prototype_amd.imaging._imaging_utils._check_imaging_parms import (
    _check_gcf_parms, _check_grid_parms, _check_imaging_weights_parms, _check_norm_parms)
from cngi_prototype_amd._utils._check_parms import _check_sel_parms

ARCSEC = np.pi / (180 * 3600)


def test_grid_parms_conversion_and_defaults():
    """_check_imaging_parms.py:23-41: arcsec -> rad, cell[0] negated,
    image_size_padded = int(fft_padding * image_size), defaults 1.2 / cube
    / centre N//2."""
    gp = {'image_size': [200, 101], 'cell_size': [0.5, 0.25]}
    assert _check_grid_parms(gp)
    assert np.allclose(gp['cell_size'], [-0.5 * ARCSEC, 0.25 * ARCSEC], rtol=1e-15)
    assert gp['fft_padding'] == 1.2 and gp['chan_mode'] == 'cube'
    assert list(gp['image_size_padded']) == [240, 121]
    assert list(gp['image_center']) == [100, 50]
    gp = {'image_size': [2048, 2048], 'cell_size': [0.1, 0.1], 'fft_padding': 1.2}
    assert _check_grid_parms(gp)
    assert list(gp['image_size_padded']) == [2457, 2457]   # truncation, not rounding


@pytest.mark.parametrize("bad", [
    {'image_size': [200], 'cell_size': [0.5, 0.5]},
    {'image_size': [200, 200], 'cell_size': [0.5, 0.5], 'chan_mode': 'mfs'},
    {'image_size': [200, 200], 'cell_size': [0.5, 0.5], 'fft_padding': 0.5},
    {'image_size': [200.0, 200], 'cell_size': [0.5, 0.5]},
    {'image_size': [200, 200], 'cell_size': [0.5, 0.5], 'precision': 'fp16'},
])
def test_grid_parms_rejects(bad):
    assert not _check_grid_parms(copy.deepcopy(bad))


def test_weight_norm_gcf_parms_defaults():
    iw = {'weighting': 'briggs'}
    assert _check_imaging_weights_parms(iw) and iw['robust'] == 0.5
    assert not _check_imaging_weights_parms({'weighting': 'briggs', 'robust': 2.5})
    assert not _check_imaging_weights_parms({'weighting': 'superuniform'})
    nm = {}
    assert _check_norm_parms(nm)
    assert nm == {'norm_type': 'flat_sky', 'single_precision': True, 'pb_limit': 0.2}
    gcf = {'list_dish_diameters': [7.0, 12.0], 'list_blockage_diameters': [0.75, 0.75],
           'unique_ant_indx': [0, 1], 'freq_chan': [1e9], 'pol': [9, 12]}
    assert _check_gcf_parms(gcf)
    assert gcf['function'] == 'casa_airy'
    assert list(gcf['oversampling']) == [10, 10] and list(gcf['max_support']) == [15, 15]
    assert gcf['support_cut_level'] == 0.025 and gcf['chan_tolerance_factor'] == 0.005
    bad = dict(gcf, list_blockage_diameters=[0.75])
    assert not _check_gcf_parms(bad)


def _xds():
    cfg = syn.scaled(syn.CONFIGS["C1"], n_time=3, n_chan=2, n_ant=4, n_src=1)
    return syn.make_mxds(cfg).attrs['xds0']


def test_sel_parms_default_groups():
    """No ids given: data_group_in is the first group; outputs are appended
    to it when append_to_in_id, else go to a new id max+1."""
    xds = _xds()
    sp = {}
    _check_sel_parms(xds, sp, new_or_modified_data_variables={'imaging_weight': 'IMAGING_WEIGHT'},
                     append_to_in_id=True)
    assert sp['data_group_in']['data'] == 'DATA'
    assert sp['data_group_out']['id'] == '0'
    assert sp['data_group_out']['imaging_weight'] == 'IMAGING_WEIGHT'
    sp = {}
    _check_sel_parms(xds, sp, new_or_modified_data_variables={'image': 'IMAGE'})
    assert sp['data_group_out']['id'] == '1'
    sp = {'data_group_in_id': 3}
    with pytest.raises(AssertionError):
        _check_sel_parms(xds, sp)


def test_make_imaging_weight_natural_reuses_weight():
    """make_imaging_weight.py:95-104: natural weighting only points the
    output group's imaging_weight at the input weight variable; 4 pols are
    refused unless full_pol is requested."""
    from cngi_prototype_amd.imaging import make_imaging_weight
    cfg = syn.scaled(syn.CONFIGS["C1"], n_time=3, n_chan=2, n_ant=4, n_src=1)
    mxds = syn.make_mxds(cfg)
    out = make_imaging_weight(mxds, {'weighting': 'natural'}, {}, {'xds': 'xds0'})
    dg = out.attrs['xds0'].attrs['data_groups'][0]['0']
    assert dg['imaging_weight'] == 'DATA_WEIGHT'
    assert 'imaging_weight' not in mxds.attrs['xds0'].attrs['data_groups'][0]['0']
    cfg4 = syn.scaled(cfg, n_pol=4)
    with pytest.raises(AssertionError):
        make_imaging_weight(syn.make_mxds(cfg4), {'weighting': 'natural'}, {}, {'xds': 'xds0'})


def test_fit_gaussian_recovers_beam():
    """An elliptical Gaussian of known FWHM and position angle sampled on
    the image grid is recovered to a few per cent (the reference's 9x9
    window, spline resampling and BFGS fit)."""
    from cngi_prototype_amd.image import fit_gaussian, make_empty_sky_image
    from cngi_prototype_amd.xds import DataArray, Dataset
    n = 64
    cell = 0.5 * ARCSEC
    xds = make_empty_sky_image(Dataset(), [0.0, 0.3], [n, n], [-cell, cell], [1e9], [1e6], [9],
                               [0.0])
    major, minor, pa = 3.0, 2.0, 30.0            # arcsec, arcsec, deg
    i = (np.arange(n) - n // 2)[:, None] * 0.5
    j = (np.arange(n) - n // 2)[None, :] * 0.5
    th = np.deg2rad(pa)
    a = i * np.cos(th) + j * np.sin(th)
    b = -i * np.sin(th) + j * np.cos(th)
    s = 2 * np.sqrt(2 * np.log(2))
    g = np.exp(-0.5 * ((a / (major / s)) ** 2 + (b / (minor / s)) ** 2))
    xds['PSF'] = DataArray(g[:, :, None, None, None], dims=['l', 'm', 'time', 'chan', 'pol'])
    out = fit_gaussian(xds, dv='PSF', beam_set_name='PSF_FIT')
    fit = np.asarray(out['PSF_FIT'].values)[0, 0, 0]
    assert abs(fit[0] - major) / major < 0.05
    assert abs(fit[1] - minor) / minor < 0.05


def test_mosaic_synthetic_layout():
    """Type-0 antennas first (type(ant1) <= type(ant2) on every baseline, so
    cf_baseline_map never meets a reversed pair), fields visited in equal
    blocks, 25 hex pointings centred on the phase centre."""
    cfg = syn.CONFIGS["C4"]
    types, dish = syn.dish_types(cfg)
    a1, a2 = syn.baseline_pairs(cfg.n_ant)
    assert np.all(types[a1] <= types[a2])
    assert list(dish) == [7.0, 12.0] and (types == 0).sum() == 12
    sched = syn.field_schedule(cfg)
    counts = np.bincount(sched)
    assert len(counts) == 25 and counts.sum() == cfg.n_time
    assert counts.max() - counts.min() <= 1
    f = syn.mosaic_fields(cfg)
    pc = syn.phase_center(cfg)
    assert f.shape == (25, 2)
    assert np.allclose(f.mean(axis=0), pc, atol=1e-6)
    assert cfg.n_baseline == 1326 and cfg.n_vis == 1440 * 1326 * 32 * 2


def test_uvw_and_field_shapes_are_checked():
    """The kernels index uvw as float64 [T, B, 3] and FIELD_ID as int32 [T, B]
    per row; the bindings refuse anything else before any device call."""
    import torch
    from cngi_prototype_amd import _native as nat
    assert nat._uvw_rows(torch.zeros((4, 5, 3), dtype=torch.float64)) == (4, 5)
    with pytest.raises(TypeError):
        nat._uvw_rows(torch.zeros((4, 5, 3), dtype=torch.float32))
    with pytest.raises(ValueError):
        nat._uvw_rows(torch.zeros((4, 5, 2), dtype=torch.float64))
    with pytest.raises(ValueError):
        nat._uvw_rows(torch.zeros((20, 3), dtype=torch.float64))
    with pytest.raises(ValueError):
        nat._uvw_rows(torch.zeros((4, 5, 3), dtype=torch.float64), (4, 6))
    nat._field_rows(torch.zeros((4, 5), dtype=torch.int32), 4, 5)
    with pytest.raises(ValueError):
        nat._field_rows(torch.zeros((4,), dtype=torch.int32), 4, 5)
    with pytest.raises(TypeError):
        nat._field_rows(torch.zeros((4, 5), dtype=torch.int64), 4, 5)
