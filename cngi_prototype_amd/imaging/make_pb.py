"""make_pb on the MI355X path (ngcasa/imaging/make_pb.py:28-123).

The rotationally symmetric Airy primary beam of every dish type, per image
pixel and channel (_casa_airy_disk / _airy_disk,
_imaging_utils/_make_pb_symmetric.py:24-130), evaluated by one GPU kernel
(ngc_make_pb) instead of a dask map_blocks over channel chunks.
"""
import copy

import numpy as np

from .. import _native
from .._utils._check_parms import _check_sel_parms
from .._utils._io import copy_structure
from ..device import to_host
from ..xds import DataArray
from ._common import update_data_groups
from ._imaging_utils._check_imaging_parms import _check_grid_parms, _check_pb_parms


def make_pb(img_xds, pb_parms, grid_parms, sel_parms):
    """Primary beam image PB with dims (l, m, time, chan, pol, dish_type).

    pb_parms: list_dish_diameters, list_blockage_diameters, function
    ('casa_airy' (default) or 'airy').  grid_parms: image_size, cell_size
    (arcsec), image_center (default image_size // 2), fft_padding, chan_mode.
    sel_parms: data_group_in_id / data_group_out_id, pb (default 'PB')."""
    print('######################### Start make_pb #########################')
    _img_xds = copy_structure(img_xds)
    _grid_parms = copy.deepcopy(grid_parms)
    _pb_parms = copy.deepcopy(pb_parms)
    _sel_parms = copy.deepcopy(sel_parms)
    _check_sel_parms(_img_xds, _sel_parms, new_or_modified_data_variables={'pb': 'PB'},
                     append_to_in_id=True)
    assert (_check_pb_parms(_img_xds, _pb_parms)), \
        "######### ERROR: user_imaging_weights_parms checking failed"
    assert (_check_grid_parms(_grid_parms)), "######### ERROR: grid_parms checking failed"
    if _pb_parms['function'] == 'airy':
        casa = False
    elif _pb_parms['function'] == 'casa_airy':
        casa = True
    else:
        # the reference prints this and then fails on its unset pattern function
        print('Only the airy function has been implemented')
        raise UnboundLocalError("no primary beam function for " + repr(_pb_parms['function']))
    _pb_parms['ipower'] = 2
    freq = np.asarray(_img_xds.coords['chan'].values, dtype=np.float64)
    pol = np.asarray(_img_xds.pol.values)
    pb = _native.make_pb(freq, _pb_parms['list_dish_diameters'],
                         _pb_parms['list_blockage_diameters'], casa, _pb_parms['ipower'],
                         _grid_parms['image_size'], _grid_parms['image_center'],
                         _grid_parms['cell_size'], len(pol))
    pb = to_host(pb)[:, :, None]
    out = _sel_parms['data_group_out']
    _img_xds[out['pb']] = DataArray(pb, dims=['l', 'm', 'time', 'chan', 'pol', 'dish_type'])
    _img_xds = _img_xds.assign_coords(
        {'dish_type': np.arange(len(_pb_parms['list_dish_diameters']))})
    update_data_groups(_img_xds, _sel_parms)
    print('######################### Created graph for make_pb #########################')
    return _img_xds
