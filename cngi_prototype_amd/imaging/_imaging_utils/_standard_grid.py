"""Graph layer of the standard gridder, MI355X form.

Replaces _graph_standard_grid / _graph_standard_degrid and the numpy wrappers
of ngcasa/imaging/_imaging_utils/_standard_grid.py:23-177, 381-464.  The
reference splits the data into dask chunks, grids each chunk into a private
grid with numba and tree-sums the grids; here the whole selection goes to the
GPU in one call (``ngc_standard_grid``), which bins the samples into uv tiles
and writes each grid cell once.  Chunking therefore only changes the FP64
summation order, never the result beyond rounding.

Grids stay on the device in the kernel's orientation
[n_imag_chan, n_pol, n_u, n_v]; the image-orientation view (u, v, chan, pol)
that the reference builds with moveaxis (:100-104) is produced only where a
caller asks for the grid itself (make_grid).
"""
import numpy as np

from ... import _native
from ...device import to_device


def _chan_map(n_chan, chan_mode):
    if chan_mode == 'cube':
        return np.arange(n_chan, dtype=np.int32), n_chan
    return np.zeros(n_chan, dtype=np.int32), 1


def grid_standard(vis_xds, cgk_1d, grid_parms, sel_parms, mode, precision=None):
    """Grid a visibility dataset.  mode: 'image', 'psf' or 'density'.

    Returns (grid, sum_weight) as device tensors, grid [n_imag_chan, n_pol,
    n_u, n_v] and sum_weight [n_imag_chan, n_pol] float64.
    """
    dg = sel_parms['data_group_in']
    weight = to_device(vis_xds[dg['imaging_weight']].data)
    uvw = to_device(vis_xds[dg['uvw']].data)
    freq = np.asarray(vis_xds.coords['chan'].values, dtype=np.float64)
    n_chan = weight.shape[2]
    n_pol = weight.shape[3]
    chan_map, n_imag_chan = _chan_map(n_chan, grid_parms['chan_mode'])
    pol_map = np.arange(n_pol, dtype=np.int32)
    if precision is None:
        precision = grid_parms.get('precision', 'fp64')
    vis = None
    if mode == 'image':
        native_mode = _native.GRID_IMAGE
        vis = to_device(vis_xds[dg['data']].data)
    elif mode == 'psf':
        native_mode = _native.GRID_PSF
    elif mode == 'density':
        native_mode = _native.GRID_DENSITY
        if precision == 'fp32_kahan':
            precision = 'fp64_tiled'
    else:
        raise ValueError(mode)
    return _native.standard_grid(
        uvw, weight, freq, chan_map, pol_map, cgk_1d, n_imag_chan,
        grid_parms['image_size_padded'], grid_parms['cell_size'], grid_parms['support'],
        grid_parms['oversampling'], native_mode, precision, vis=vis)


def degrid_imaging_weight(vis_xds, density_grid, factors, grid_parms, sel_parms):
    """_graph_standard_degrid + _standard_imaging_weight_degrid_jit
    (_standard_grid.py:381-518) on the GPU.  Returns a device tensor shaped like
    the natural weights."""
    dg = sel_parms['data_group_in']
    weight = to_device(vis_xds[dg['imaging_weight']].data)
    uvw = to_device(vis_xds[dg['uvw']].data)
    freq = np.asarray(vis_xds.coords['chan'].values, dtype=np.float64)
    n_chan, n_pol = weight.shape[2], weight.shape[3]
    chan_map, _ = _chan_map(n_chan, grid_parms['chan_mode'])
    pol_map = np.arange(n_pol, dtype=np.int32)
    return _native.imaging_weight_degrid(uvw, weight, density_grid, factors, freq, chan_map,
                                         pol_map, grid_parms['cell_size'])
