"""make_psf_with_gcf on the MI355X path (ngcasa/imaging/make_psf_with_gcf.py:20-181)."""
import copy

import numpy as np

from .._utils._check_parms import _check_sel_parms
from .._utils._io import copy_structure
from ..image import fit_gaussian
from ._common import image_dataarray, new_image, plane_image, sum_weight_dataarray, \
    update_data_groups
from ._imaging_utils._aperture_grid import grid_aperture
from ._imaging_utils._check_imaging_parms import _check_grid_parms, _check_norm_parms
from ._imaging_utils._normalize import normalize


def make_psf_with_gcf(mxds, gcf_dataset, img_dataset, grid_parms, norm_parms, vis_sel_parms,
                      img_sel_parms):
    """A-projected PSF: aperture gridding of the imaging weights, FFT, PB
    normalisation, then division by the centre pixel of every plane
    (make_psf_with_gcf.py:140); adds PSF, PSF_SUM_WEIGHT and PSF_FIT."""
    print('######################### Start make_psf_with_gcf #########################')
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
                     new_or_modified_data_variables={'sum_weight': 'PSF_SUM_WEIGHT', 'psf': 'PSF',
                                                     'psf_fit': 'PSF_FIT'},
                     required_data_variables={'pb': 'PB', 'weight_pb': 'WEIGHT_PB'},
                     append_to_in_id=False)
    _grid_parms['grid_weights'] = False
    _grid_parms['do_psf'] = True
    _grid_parms['oversampling'] = np.array(gcf_dataset.attrs['oversampling'])
    grid, sum_weight = grid_aperture(_vis_dataset, gcf_dataset, _grid_parms, _vis_sel_parms,
                                     'psf')
    image = plane_image(grid, _grid_parms)
    del grid
    normalize(image, sum_weight, img_dataset, gcf_dataset, 'forward', _norm_parms, _img_sel_parms)
    c = _grid_parms['image_center']
    image /= image[:, :, int(c[0]):int(c[0]) + 1, int(c[1]):int(c[1]) + 1].clone()
    _img_dataset = new_image(_img_dataset, _vis_dataset, _grid_parms)
    out = _img_sel_parms['data_group_out']
    _img_dataset[out['sum_weight']] = sum_weight_dataarray(sum_weight)
    _img_dataset[out['psf']] = image_dataarray(image)
    update_data_groups(_img_dataset, _img_sel_parms)
    _img_dataset = fit_gaussian(_img_dataset, dv=out['psf'], beam_set_name=out['psf_fit'])
    print('#########################  Created graph for make_psf_with_gcf #########################')
    return _img_dataset
