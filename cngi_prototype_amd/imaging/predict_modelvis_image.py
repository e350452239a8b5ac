"""predict_modelvis_image on the MI355X path.

The reference's function (ngcasa/imaging/predict_modelvis_image.py:20-44) is
a stub: its docstring describes the operation (model visibilities from a
model image in Jy/pixel, saved as 'MODEL', optionally added to an existing
model) and its comments name the steps (gridding correction, Fourier
transform, degrid).  Here those steps run on the GPU with the standard
gridder's own conventions, so that predicting from a model and imaging the
prediction are consistent:

- the model image is divided by the prolate-spheroidal correction image,
  padded to image_size_padded and transformed (ngc_model_to_grid), the
  inverse placement of make_image's FFT and crop;
- every visibility is the tap-weighted mean of the grid over the footprint
  the gridder would give it (ngc_standard_degrid, the adjoint of the
  gridder with unit weights, normalised by the tap sum); samples off the grid
  are 0.

Parameters as the reference names them: img_dataset holds the model image
(`storage_parms['image_name']`, default 'MODEL', dims (l, m, time, chan,
pol)); vis_dataset holds UVW and the chan coordinate; grid_parms as for
make_image (image_size, cell_size, fft_padding, chan_mode, optional
precision); storage_parms: arr_name (default 'MODEL'), incremental
(default False: overwrite; True: add to the existing variable).

With `gcf_dataset` (the output of make_gridding_convolution_function) the
prediction uses the mosaic A-term kernels instead: the model is transformed
without the prolate-spheroidal correction and every visibility is
sum conj(CK PG) G / sum CK over the cells, offsets and taps the aperture
gridder would give it (ngc_aperture_degrid, the adjoint of the aperture
gridder with unit weights, normalised so that the beam is 1 at each field's
pointing centre: a source there predicts its flux, a source elsewhere its
flux times the beam); vis_dataset then needs FIELD_ID.
"""
import copy

import numpy as np

from .. import _native
from .._utils._io import copy_structure
from ..device import to_device
from ..xds import DataArray
from ._imaging_utils._check_imaging_parms import _check_grid_parms
from ._imaging_utils._gridding_convolutional_kernels import cropped_correction_factors, \
    prolate_spheroidal_kernel_1d


def predict_modelvis_image(img_dataset, vis_dataset, grid_parms, storage_parms, gcf_dataset=None):
    """Model visibilities [time, baseline, chan, pol] from a model image cube
    (or a one-channel continuum model with chan_mode 'continuum'); with
    gcf_dataset, through the A-term kernels."""
    import torch
    print('######################### Start predict_modelvis_image #########################')
    _grid_parms = copy.deepcopy(grid_parms)
    _storage_parms = copy.deepcopy(storage_parms) if storage_parms else {}
    assert (_check_grid_parms(_grid_parms)), "######### ERROR: grid_parms checking failed"
    image_name = _storage_parms.get('image_name', 'MODEL')
    arr_name = _storage_parms.get('arr_name', 'MODEL')
    incremental = bool(_storage_parms.get('incremental', False))
    assert image_name in img_dataset, \
        "######### ERROR: the image dataset has no model image " + str(image_name)

    model = img_dataset[image_name]
    m = model.data
    m = m[:, :, 0] if isinstance(m, torch.Tensor) else np.asarray(m)[:, :, 0]
    # (l, m, chan, pol) -> (chan, pol, l, m), float64 on the device
    m = to_device(m, torch.float64).permute(2, 3, 0, 1).contiguous()
    n_mchan, n_mpol = int(m.shape[0]), int(m.shape[1])
    assert tuple(m.shape[2:]) == tuple(int(x) for x in _grid_parms['image_size']), \
        "######### ERROR: the model image size differs from grid_parms['image_size']"

    freq = np.asarray(vis_dataset.coords['chan'].values, dtype=np.float64)
    uvw = to_device(vis_dataset['UVW'].data, torch.float64)
    n_chan = len(freq)
    if _grid_parms['chan_mode'] == 'cube':
        assert n_mchan == n_chan, \
            "######### ERROR: a cube model needs one plane per visibility channel"
        chan_map = np.arange(n_chan)
    else:
        assert n_mchan == 1, "######### ERROR: a continuum model has one channel"
        chan_map = np.zeros(n_chan, dtype=int)

    single = _grid_parms.get('precision', 'fp64') in ('fp32_kahan', 'fp32')
    if gcf_dataset is None:
        oversampling, support = 100, 7
        cgk_1d = prolate_spheroidal_kernel_1d(oversampling, support)
        corr_l, corr_m = cropped_correction_factors(_grid_parms['image_size_padded'],
                                                    _grid_parms['image_size'])
        grid = _native.model_to_grid(m, _grid_parms['image_size_padded'], corr_l, corr_m,
                                     single=single)
        vis = _native.standard_degrid(uvw, grid, freq, chan_map, np.arange(n_mpol), cgk_1d,
                                      _grid_parms['cell_size'], support, oversampling,
                                      single=single)
    else:
        assert 'FIELD_ID' in vis_dataset, \
            "######### ERROR: A-term prediction needs FIELD_ID in the visibility dataset"
        grid = _native.model_to_grid(m, _grid_parms['image_size_padded'], single=single)
        vis = _native.aperture_degrid(
            uvw, to_device(vis_dataset['FIELD_ID'].data, torch.int32),
            np.asarray(gcf_dataset.coords['field_id'].values), grid, freq, chan_map,
            np.arange(n_mpol), np.asarray(gcf_dataset['CF_BASELINE_MAP'].values),
            np.asarray(gcf_dataset['CF_CHAN_MAP'].values),
            np.asarray(gcf_dataset['CF_POL_MAP'].values),
            to_device(gcf_dataset['CONV_KERNEL'].data, torch.float64),
            np.asarray(gcf_dataset['SUPPORT'].values),
            to_device(gcf_dataset['PHASE_GRADIENT'].data, torch.complex128),
            _grid_parms['cell_size'], np.asarray(gcf_dataset.attrs['oversampling']),
            single=single)
    del grid
    out = copy_structure(vis_dataset)
    if incremental and arr_name in out.data_vars:
        prev = to_device(out[arr_name].data)
        vis = vis + prev.to(vis.dtype)
    out[arr_name] = DataArray(vis, dims=['time', 'baseline', 'chan', 'pol'])
    print('######################### Created graph for predict_modelvis_image '
          '#########################')
    return out
