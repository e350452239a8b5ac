"""Graph layer of the aperture (A-term) gridder, MI355X form.

Replaces _graph_aperture_grid and the numpy wrappers of
ngcasa/imaging/_imaging_utils/_aperture_grid.py:25-371: the whole selection
is gridded by one ngc_aperture_grid call on the GPU (the reference's
per-chunk private grids and tree-sum only change FP64 summation order).
"""
import numpy as np

from ... import _native
from ...device import to_device


def grid_aperture(vis_xds, gcf_xds, grid_parms, sel_parms, mode):
    """mode: 'image', 'psf' or 'weight'.  Returns device (grid complex
    [n_imag_chan, n_pol, n_u, n_v], sum_weight [n_imag_chan, n_pol])."""
    import torch
    dg = sel_parms['data_group_in']
    weight = to_device(vis_xds[dg['imaging_weight']].data)
    uvw = to_device(vis_xds[dg['uvw']].data)
    field = to_device(vis_xds['FIELD_ID'].data, dtype=torch.int32)
    freq = np.asarray(vis_xds.coords['chan'].values, dtype=np.float64)
    n_chan, n_pol = weight.shape[2], weight.shape[3]
    if grid_parms['chan_mode'] == 'cube':
        chan_map, n_imag_chan = np.arange(n_chan), n_chan
    else:
        chan_map, n_imag_chan = np.zeros(n_chan, dtype=int), 1
    kernel_name = 'WEIGHT_CONV_KERNEL' if mode == 'weight' else 'CONV_KERNEL'
    kernel = to_device(gcf_xds[kernel_name].data, dtype=torch.float64)
    pg = to_device(gcf_xds['PHASE_GRADIENT'].data, dtype=torch.complex128)
    native_mode = {'image': _native.GRID_IMAGE, 'psf': _native.GRID_PSF,
                   'weight': _native.GRID_WEIGHT}[mode]
    vis = to_device(vis_xds[dg['data']].data) if mode == 'image' else None
    return _native.aperture_grid(
        uvw, weight, field, np.asarray(gcf_xds.coords['field_id'].values), freq, chan_map,
        np.arange(n_pol), np.asarray(gcf_xds['CF_BASELINE_MAP'].values),
        np.asarray(gcf_xds['CF_CHAN_MAP'].values), np.asarray(gcf_xds['CF_POL_MAP'].values),
        kernel, np.asarray(gcf_xds['SUPPORT'].values), pg, n_imag_chan,
        grid_parms['image_size_padded'], grid_parms['cell_size'], grid_parms['oversampling'],
        native_mode, grid_parms.get('precision', 'fp64'), vis=vis)
