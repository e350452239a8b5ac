"""make_mosaic_pb on the MI355X path (ngcasa/imaging/make_mosaic_pb.py:30-162)."""
import copy

import numpy as np

from .._utils._check_parms import _check_sel_parms
from .._utils._io import copy_structure
from ..device import to_host
from ..xds import DataArray
from ._common import new_image, plane_image, sum_weight_dataarray, update_data_groups
from ._imaging_utils._aperture_grid import grid_aperture
from ._imaging_utils._check_imaging_parms import _check_grid_parms


def make_mosaic_pb(mxds, gcf_dataset, img_dataset, vis_sel_parms, img_sel_parms, grid_parms):
    """Mosaic primary beam (PB) and weight image (WEIGHT_PB) from the weight
    convolution kernels gridded at the grid centre for every visibility."""
    print('######################### Start make_mosaic_pb #########################')
    _mxds = copy_structure(mxds)
    _img_dataset = copy_structure(img_dataset)
    _vis_sel_parms = copy.deepcopy(vis_sel_parms)
    _img_sel_parms = copy.deepcopy(img_sel_parms)
    _grid_parms = copy.deepcopy(grid_parms)
    assert ('xds' in _vis_sel_parms), "######### ERROR: xds must be specified in sel_parms"
    _vis_dataset = _mxds.attrs[_vis_sel_parms['xds']]
    assert (_check_grid_parms(_grid_parms)), "######### ERROR: grid_parms checking failed"
    _check_sel_parms(_vis_dataset, _vis_sel_parms)
    _check_sel_parms(_img_dataset, _img_sel_parms,
                     new_or_modified_data_variables={'pb': 'PB', 'weight_pb': 'WEIGHT_PB',
                                                     'weight_pb_sum_weight': 'WEIGHT_PB_SUM_WEIGHT'},
                     append_to_in_id=True)
    _grid_parms['grid_weights'] = True
    _grid_parms['do_psf'] = False
    _grid_parms['oversampling'] = np.array(gcf_dataset.attrs['oversampling'])
    grid, sum_weight = grid_aperture(_vis_dataset, gcf_dataset, _grid_parms, _vis_sel_parms,
                                     'weight')
    weight = plane_image(grid, _grid_parms, sum_weight=sum_weight)
    del grid
    w = np.moveaxis(to_host(weight), [0, 1], [-2, -1])[:, :, None, :, :]
    pb = np.sqrt(np.abs(w))
    _img_dataset = new_image(_img_dataset, _vis_dataset, _grid_parms)
    out = _img_sel_parms['data_group_out']
    dims = ['l', 'm', 'time', 'chan', 'pol']
    _img_dataset[out['pb']] = DataArray(pb, dims=dims)
    _img_dataset[out['weight_pb']] = DataArray(w, dims=dims)
    _img_dataset[out['weight_pb_sum_weight']] = sum_weight_dataarray(sum_weight)
    update_data_groups(_img_dataset, _img_sel_parms)
    print('#########################  Created graph for make_mosaic_pb #########################')
    return _img_dataset
