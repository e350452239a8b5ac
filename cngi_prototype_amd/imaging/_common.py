"""Steps shared by the imaging entry points: image coordinates and the
post-grid FFT/crop/correction on the device."""
import numpy as np

from .. import _native
from ..device import to_host
from ..image import make_empty_sky_image
from ..xds import DataArray


def image_coords(vis_xds, grid_parms):
    """Frequency/width/pol/time coordinates of the output image
    (make_psf.py:134-151)."""
    if grid_parms['chan_mode'] == 'continuum':
        freq_coords = [np.mean(vis_xds.coords['chan'].values)]
        chan_width = np.array([np.mean(np.asarray(vis_xds['chan_width'].values))])
    else:
        freq_coords = vis_xds.coords['chan'].values
        chan_width = np.asarray(vis_xds['chan_width'].values)
    pol_coords = vis_xds.coords['pol'].values if 'pol' in vis_xds.coords else np.arange(
        vis_xds.dims['pol'])
    time_coords = [np.mean(vis_xds.coords['time'].values)] if 'time' in vis_xds.coords else [0.0]
    return freq_coords, chan_width, pol_coords, time_coords


def new_image(img_xds, vis_xds, grid_parms):
    f, w, p, t = image_coords(vis_xds, grid_parms)
    return make_empty_sky_image(img_xds, grid_parms['phase_center'], grid_parms['image_size'],
                                grid_parms['cell_size'], f, w, p, t)


def plane_image(grid, grid_parms, sum_weight=None, corr=None):
    """Device grid [C, P, n_u, n_v] -> device image [C, P, N_l, N_m]."""
    return _native.grid_to_image(grid, grid_parms['image_size'], sum_weight=sum_weight, corr=corr)


def image_dataarray(img_dev):
    """Device [C, P, l, m] -> host DataArray (l, m, time, chan, pol)."""
    a = to_host(img_dev)
    a = np.moveaxis(a, [0, 1], [-2, -1])[:, :, None, :, :]
    return DataArray(a, dims=['l', 'm', 'time', 'chan', 'pol'])


def sum_weight_dataarray(sw_dev):
    return DataArray(to_host(sw_dev)[None, :, :], dims=['time', 'chan', 'pol'])


def update_data_groups(xds, sel_parms):
    xds.attrs['data_groups'][0] = {**xds.attrs['data_groups'][0],
                                   **{sel_parms['data_group_out']['id']: sel_parms['data_group_out']}}
