"""make_grid on the MI355X path (ngcasa/imaging/make_grid.py:27-143): the
un-FFT'd standard grid and its sum of weights."""
import copy

import numpy as np

from .._utils._check_parms import _check_sel_parms
from .._utils._io import copy_structure
from ..device import to_host
from ..xds import DataArray
from ._common import sum_weight_dataarray, update_data_groups
from ._imaging_utils._check_imaging_parms import _check_grid_parms
from ._imaging_utils._gridding_convolutional_kernels import prolate_spheroidal_kernel_1d
from ._imaging_utils._standard_grid import grid_standard


def make_grid(vis_mxds, img_xds, grid_parms, vis_sel_parms, img_sel_parms):
    """Complex grid of the visibilities (u, v, time, chan, pol) with the
    prolate-spheroidal kernel (support 7, oversampling 100), and SUM_WEIGHT."""
    print('######################### Start make_grid #########################')
    _mxds = copy_structure(vis_mxds)
    _img_xds = copy_structure(img_xds)
    _vis_sel_parms = copy.deepcopy(vis_sel_parms)
    _img_sel_parms = copy.deepcopy(img_sel_parms)
    _grid_parms = copy.deepcopy(grid_parms)
    assert (_check_grid_parms(_grid_parms)), "######### ERROR: grid_parms checking failed"
    assert ('xds' in _vis_sel_parms), "######### ERROR: xds must be specified in sel_parms"
    _vis_xds = _mxds.attrs[_vis_sel_parms['xds']]
    _check_sel_parms(_vis_xds, _vis_sel_parms)
    _check_sel_parms(_img_xds, _img_sel_parms,
                     new_or_modified_data_variables={'sum_weight': 'SUM_WEIGHT', 'grid': 'GRID'},
                     append_to_in_id=True)
    _grid_parms['oversampling'] = 100
    _grid_parms['support'] = 7
    cgk_1d = prolate_spheroidal_kernel_1d(_grid_parms['oversampling'], _grid_parms['support'])
    grid, sum_weight = grid_standard(_vis_xds, cgk_1d, _grid_parms, _vis_sel_parms, 'image')
    g = np.moveaxis(to_host(grid), [0, 1], [-2, -1])[:, :, None, :, :]
    out = _img_sel_parms['data_group_out']
    _img_xds[out['sum_weight']] = sum_weight_dataarray(sum_weight)
    _img_xds[out['grid']] = DataArray(g, dims=['u', 'v', 'time', 'chan', 'pol'])
    update_data_groups(_img_xds, _img_sel_parms)
    print('######################### Created graph for make_grid #########################')
    return _img_xds
