"""make_psf on the MI355X gridding path (ngcasa/imaging/make_psf.py:27-163)."""
import copy

from .._utils._check_parms import _check_sel_parms
from .._utils._io import copy_structure
from ..image import fit_gaussian
from ._common import image_dataarray, new_image, plane_image, sum_weight_dataarray, \
    update_data_groups
from ._imaging_utils._check_imaging_parms import _check_grid_parms
from ._imaging_utils._gridding_convolutional_kernels import cropped_correction_factors, \
    prolate_spheroidal_kernel_1d
from ._imaging_utils._standard_grid import grid_standard


def make_psf(vis_mxds, img_xds, grid_parms, vis_sel_parms, img_sel_parms):
    """Point spread function from uvw and imaging weights, prolate-spheroidal
    gridding (support 7, oversampling 100).  Same parameters, defaults and
    outputs (PSF, PSF_SUM_WEIGHT, PSF_FIT) as the reference; the gridding,
    FFT and correction run on the GPU.  grid_parms may carry the optional key
    'precision' ('fp64' default, 'fp64_tiled', 'fp32_kahan')."""
    print('######################### Start make_psf #########################')
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
                     new_or_modified_data_variables={'sum_weight': 'PSF_SUM_WEIGHT', 'psf': 'PSF',
                                                     'psf_fit': 'PSF_FIT'},
                     append_to_in_id=True)

    _grid_parms['oversampling'] = 100
    _grid_parms['support'] = 7
    cgk_1d = prolate_spheroidal_kernel_1d(_grid_parms['oversampling'], _grid_parms['support'])
    corr = cropped_correction_factors(_grid_parms['image_size_padded'], _grid_parms['image_size'])

    grid, sum_weight = grid_standard(_vis_xds, cgk_1d, _grid_parms, _vis_sel_parms, 'psf')
    image = plane_image(grid, _grid_parms, sum_weight=sum_weight, corr=corr)
    del grid

    _img_xds = new_image(_img_xds, _vis_xds, _grid_parms)
    out = _img_sel_parms['data_group_out']
    _img_xds[out['sum_weight']] = sum_weight_dataarray(sum_weight)
    _img_xds[out['psf']] = image_dataarray(image)
    update_data_groups(_img_xds, _img_sel_parms)
    _img_xds = fit_gaussian(_img_xds, dv=out['psf'], beam_set_name=out['psf_fit'])
    print('######################### Created graph for make_psf #########################')
    return _img_xds
