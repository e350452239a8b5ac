"""make_image_with_gcf on the MI355X path (ngcasa/imaging/make_image_with_gcf.py:20-191)."""
import copy

import numpy as np

from .._utils._check_parms import _check_sel_parms
from .._utils._io import copy_structure
from ._common import image_dataarray, new_image, plane_image, sum_weight_dataarray, \
    update_data_groups
from ._imaging_utils._aperture_grid import grid_aperture
from ._imaging_utils._check_imaging_parms import _check_grid_parms, _check_norm_parms
from ._imaging_utils._normalize import normalize


def make_image_with_gcf(mxds, gcf_dataset, img_dataset, grid_parms, norm_parms, vis_sel_parms,
                        img_sel_parms):
    """A-projected (mosaic) dirty image: aperture gridding with the gcf
    dataset's kernels and phase gradients, FFT, then PB normalisation
    (flat_sky, flat_noise or none)."""
    print('######################### Start make_image_with_gcf #########################')
    _mxds = copy_structure(mxds)
    _img_dataset = copy_structure(img_dataset)
    _vis_sel_parms = copy.deepcopy(vis_sel_parms)
    _img_sel_parms = copy.deepcopy(img_sel_parms)
    _grid_parms = copy.deepcopy(grid_parms)
    _norm_parms = copy.deepcopy(norm_parms)
    assert ('xds' in _vis_sel_parms), "######### ERROR: xds must be specified in sel_parms"
    _vis_dataset = _mxds.attrs[_vis_sel_parms['xds']]
    assert (_check_grid_parms(_grid_parms)), "######### ERROR: grid_parms checking failed"
    assert (_check_norm_parms(_norm_parms)), "######### ERROR: norm_parms checking failed"
    _check_sel_parms(_vis_dataset, _vis_sel_parms)
    _check_sel_parms(_img_dataset, _img_sel_parms,
                     new_or_modified_data_variables={'sum_weight': 'SUM_WEIGHT', 'image': 'IMAGE'},
                     required_data_variables={'pb': 'PB', 'weight_pb': 'WEIGHT_PB'},
                     append_to_in_id=False)
    _grid_parms['grid_weights'] = False
    _grid_parms['do_psf'] = False
    _grid_parms['oversampling'] = np.array(gcf_dataset.attrs['oversampling'])
    grid, sum_weight = grid_aperture(_vis_dataset, gcf_dataset, _grid_parms, _vis_sel_parms,
                                     'image')
    image = plane_image(grid, _grid_parms)
    del grid
    print('grid sizes', _grid_parms['image_size_padded'][0], _grid_parms['image_size_padded'][1])
    normalize(image, sum_weight, img_dataset, gcf_dataset, 'forward', _norm_parms, _img_sel_parms)
    _img_dataset = new_image(_img_dataset, _vis_dataset, _grid_parms)
    out = _img_sel_parms['data_group_out']
    _img_dataset[out['sum_weight']] = sum_weight_dataarray(sum_weight)
    _img_dataset[out['image']] = image_dataarray(image)
    update_data_groups(_img_dataset, _img_sel_parms)
    print('#########################  Created graph for make_mosaic_with_gcf #########################')
    return _img_dataset
