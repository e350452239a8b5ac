"""make_gridding_convolution_function on the MI355X path
(ngcasa/imaging/make_gridding_convolution_function.py:36-319)."""
import copy

import numpy as np

from .. import _native
from .._utils._check_parms import _check_sel_parms
from .._utils._io import copy_structure
from ..xds import DataArray, Dataset
from ._imaging_utils._check_imaging_parms import _check_gcf_parms, _check_grid_parms
from ._imaging_utils._gcf_maps import cf_baseline_map, cf_chan_map, field_pixel_offsets, \
    gcf_kernels


def make_gridding_convolution_function(mxds, gcf_parms, grid_parms, sel_parms):
    """A-term gridding convolution functions for heterogeneous arrays (Airy
    or CASA-Airy primary beams).  Same parameters, defaults and output
    variables as the reference (CONV_KERNEL, WEIGHT_CONV_KERNEL, SUPPORT,
    PS_CORR_IMAGE, PHASE_GRADIENT, CF_BASELINE_MAP, CF_CHAN_MAP, CF_POL_MAP;
    attrs cell_uv, oversampling).  The patterns, their FFTs, the support cut's
    reductions, the crop/normalisation and the phase gradients are computed
    on the GPU; the kernel arrays of the returned dataset stay on the device."""
    print('######################### Start make_gridding_convolution_function #########################')
    _mxds = copy_structure(mxds)
    _gcf_parms = copy.deepcopy(gcf_parms)
    _grid_parms = copy.deepcopy(grid_parms)
    _sel_parms = copy.deepcopy(sel_parms)
    assert ('xds' in _sel_parms), "######### ERROR: xds must be specified in sel_parms"
    _vis_dataset = _mxds.attrs[sel_parms['xds']]
    _check_sel_parms(_vis_dataset, _sel_parms)

    _gcf_parms['basline_ant'] = np.array([np.asarray(_vis_dataset.ANTENNA1.values),
                                          np.asarray(_vis_dataset.ANTENNA2.values)]).T
    _gcf_parms['freq_chan'] = np.asarray(_vis_dataset.coords['chan'].values)
    _gcf_parms['pol'] = np.asarray(_vis_dataset.coords['pol'].values)
    field = mxds.attrs['FIELD']
    field_phase_dir = np.asarray(field['PHASE_DIR'].values)[:, 0, :]
    field_id = np.asarray(field.coords['field_id'].values)

    assert (_check_gcf_parms(_gcf_parms)), "######### ERROR: gcf_parms checking failed"
    assert (_check_grid_parms(_grid_parms)), "######### ERROR: grid_parms checking failed"
    if _gcf_parms['function'] not in ('airy', 'casa_airy'):
        raise AssertionError("######### ERROR: Only airy and casa_airy function has been implemented")
    casa = _gcf_parms['function'] == 'casa_airy'

    os_ = _gcf_parms['oversampling']
    resize = (_gcf_parms['max_support'] + 1) * os_
    npad = _grid_parms['image_size_padded']
    n_unique = len(_gcf_parms['list_dish_diameters'])
    cfb, pairs = cf_baseline_map(_gcf_parms['unique_ant_indx'], _gcf_parms['basline_ant'], n_unique)
    cfc, pb_freq = cf_chan_map(_gcf_parms['freq_chan'], _gcf_parms['chan_tolerance_factor'])
    cfp = np.zeros(len(_gcf_parms['pol']), dtype=int)
    pb_cell = _grid_parms['cell_size'] * os_

    weight_conv_kernel, conv_kernel, supports = gcf_kernels(
        pb_freq, _gcf_parms['list_dish_diameters'], _gcf_parms['list_blockage_diameters'], pairs,
        casa, pb_cell, npad, os_, _gcf_parms['max_support'], _gcf_parms['support_cut_level'])
    surrogate = _gcf_parms.get('surrogate')
    if surrogate is not None:
        # per-channel convolution kernels from the learned surrogate (the
        # exact kernel of the channel's bin plus the predicted change to the
        # channel's own frequency), each checked against the exact kernel when
        # surrogate_check is set; the weight kernels stay the bin's exact ones
        from ._imaging_utils._kernel_surrogate import per_channel_kernels
        conv_kernel, weight_conv_kernel, report = per_channel_kernels(
            surrogate, _gcf_parms['freq_chan'], cfc, pb_freq, _gcf_parms['list_dish_diameters'],
            _gcf_parms['list_blockage_diameters'], pairs, casa, pb_cell, npad, os_,
            _gcf_parms['max_support'], supports, conv_kernel, weight_conv_kernel,
            check=_gcf_parms.get('surrogate_check', True))
        supports = supports[:, cfc]
        cfc = np.arange(len(_gcf_parms['freq_chan']))

    pix_dist = field_pixel_offsets(field_phase_dir, _gcf_parms['phase_center'],
                                   _grid_parms['cell_size'])
    pix = -pix_dist * 2 * np.pi / (npad * os_)
    phase_gradient = _native.phase_gradient(pix, resize)

    dims5 = ('conv_baseline', 'conv_chan', 'conv_pol', 'u', 'v')
    gcf = Dataset(
        data_vars={
            'CONV_KERNEL': (dims5, conv_kernel),
            'WEIGHT_CONV_KERNEL': (dims5, weight_conv_kernel),
            'SUPPORT': (('conv_baseline', 'conv_chan', 'conv_pol', 'xy'), supports),
            'PS_CORR_IMAGE': (('l', 'm'), np.ones(tuple(_grid_parms['image_size']))),
            'PHASE_GRADIENT': (('field_id', 'u', 'v'), phase_gradient),
            'CF_BASELINE_MAP': (('baseline',), cfb),
            'CF_CHAN_MAP': (('chan',), cfc),
            'CF_POL_MAP': (('pol',), cfp),
        },
        coords={'u': (('u',), np.arange(resize[0])), 'v': (('v',), np.arange(resize[1])),
                'xy': (('xy',), np.arange(2)), 'field_id': (('field_id',), field_id),
                'l': (('l',), np.arange(_grid_parms['image_size'][0])),
                'm': (('m',), np.arange(_grid_parms['image_size'][1]))})
    gcf.attrs['cell_uv'] = 1 / (npad * _grid_parms['cell_size'] * os_)
    gcf.attrs['oversampling'] = os_
    if surrogate is not None:
        gcf.attrs['surrogate_report'] = report
    print('#########################  Created graph for make_gridding_convolution_function #########################')
    return gcf
