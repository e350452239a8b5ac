"""make_imaging_weight on the MI355X gridding path
(ngcasa/imaging/make_imaging_weight.py:20-255)."""
import copy

import numpy as np

from .. import _native
from .._utils._check_parms import _check_sel_parms
from .._utils._io import copy_structure
from ..xds import DataArray
from ._imaging_utils._check_imaging_parms import _check_grid_parms, _check_imaging_weights_parms
from ._imaging_utils._standard_grid import degrid_imaging_weight, grid_standard


def make_imaging_weight(vis_mxds, imaging_weights_parms, grid_parms, sel_parms):
    """Imaging weights (time, baseline, chan, pol) for natural, uniform or
    Briggs weighting.  As in the reference, more than two polarisations are
    refused ("Full polarization is not supported") unless
    imaging_weights_parms['full_pol'] is True, in which case the reference's
    density rule is applied unchanged: the density grid of every pol is built
    from (w_0 + w_1)/2 and each pol's natural weight is divided by
    f0*g + f1.  'briggs_abs' fails as in the reference, whose branch names an
    undefined variable."""
    print('######################### Start make_imaging_weights #########################')
    _mxds = copy_structure(vis_mxds)
    _imaging_weights_parms = copy.deepcopy(imaging_weights_parms)
    _grid_parms = copy.deepcopy(grid_parms)
    _sel_parms = copy.deepcopy(sel_parms)

    assert ('xds' in _sel_parms), "######### ERROR: xds must be specified in sel_parms"
    _vis_xds = _mxds.attrs[sel_parms['xds']]
    if not _imaging_weights_parms.get('full_pol', False):
        assert _vis_xds.dims['pol'] <= 2, "Full polarization is not supported."
    _check_sel_parms(_vis_xds, _sel_parms,
                     new_or_modified_data_variables={'imaging_weight': 'IMAGING_WEIGHT'},
                     append_to_in_id=True)
    assert (_check_imaging_weights_parms(_imaging_weights_parms)), \
        "######### ERROR: imaging_weights_parms checking failed"
    if _imaging_weights_parms['weighting'] != 'natural':
        assert (_check_grid_parms(_grid_parms)), "######### ERROR: grid_parms checking failed"
    else:
        _sel_parms['data_group_out']['imaging_weight'] = _sel_parms['data_group_in']['weight']
        _vis_xds.attrs['data_groups'][0] = {**_vis_xds.attrs['data_groups'][0],
                                            **{_sel_parms['data_group_out']['id']:
                                               _sel_parms['data_group_out']}}
        print("Since weighting is natural input weight will be reused as imaging weight.")
        print('######################### Created graph for make_imaging_weight #########################')
        return _mxds

    _vis_xds[_sel_parms['data_group_out']['imaging_weight']] = \
        _vis_xds[_sel_parms['data_group_in']['weight']]
    _sel_parms['data_group_in']['imaging_weight'] = _sel_parms['data_group_out']['imaging_weight']
    calc_briggs_weights(_vis_xds, _imaging_weights_parms, _grid_parms, _sel_parms)
    _vis_xds.attrs['data_groups'][0] = {**_vis_xds.attrs['data_groups'][0],
                                        **{_sel_parms['data_group_out']['id']:
                                           _sel_parms['data_group_out']}}
    print('######################### Created graph for make_imaging_weight #########################')
    return _mxds


def calc_briggs_weights(vis_xds, imaging_weights_parms, grid_parms, sel_parms):
    """Density grid (nearest cell + conjugate, no padding), Briggs factors and
    degridded imaging weights, all on the GPU (make_imaging_weight.py:144-255).
    The imaging weights stay resident on the device in the returned dataset."""
    if imaging_weights_parms['weighting'] == 'briggs_abs':
        raise NameError("name 'briggs_factor' is not defined (briggs_abs is not functional in "
                        "the reference, make_imaging_weight.py:204-208)")
    grid_parms['image_size_padded'] = grid_parms['image_size']
    grid_parms['oversampling'] = 0
    grid_parms['support'] = 1
    grid_parms['do_psf'] = True
    grid_parms['complex_grid'] = False
    grid_parms['do_imaging_weight'] = True
    cgk_1d = np.ones((1))
    prec = grid_parms.get('precision', 'fp64')
    density_prec = 'fp64' if prec == 'fp64' else 'fp64_tiled'
    density, sum_weight = grid_standard(vis_xds, cgk_1d, grid_parms, sel_parms, 'density',
                                        precision=density_prec)
    factors = _native.briggs_factors(density, sum_weight, imaging_weights_parms['weighting'],
                                     imaging_weights_parms.get('robust', 0.5))
    imaging_weight = degrid_imaging_weight(vis_xds, density, factors, grid_parms, sel_parms)
    del density
    dims = vis_xds[sel_parms['data_group_in']['data']].dims
    vis_xds[sel_parms['data_group_in']['imaging_weight']] = DataArray(imaging_weight, dims=dims)
    return factors
