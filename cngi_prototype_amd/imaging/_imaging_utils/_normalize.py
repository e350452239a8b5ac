"""Mosaic normalisation, 'forward' direction of _normalize
(ngcasa/imaging/_imaging_utils/_normalize.py:17-89), on the GPU."""
import numpy as np

from ... import _native
from ...device import to_device


def _plane_major(a):
    """[l, m, chan, pol] host/device array -> contiguous device [chan, pol, l, m]."""
    import torch
    t = to_device(a, dtype=torch.float64)
    return t.permute(2, 3, 0, 1).contiguous()


def normalize(image_dev, sum_weight_dev, img_dataset, gcf_dataset, direction, norm_parms,
              sel_parms):
    """image_dev: device float64 [chan, pol, l, m] (modified in place)."""
    if direction != 'forward':
        print('reverse operation not yet implemented not yet implemented')
        return None
    norm_type = norm_parms['norm_type']
    dg = sel_parms['data_group_in']
    pb = _plane_major(np.asarray(img_dataset[dg['pb']].values)[:, :, 0, :, :])
    ps = to_device(np.ascontiguousarray(np.asarray(gcf_dataset['PS_CORR_IMAGE'].values),
                                        dtype=np.float64))
    if norm_type == 'flat_noise':
        norm_pb = pb
    elif norm_type == 'flat_sky':
        norm_pb = _plane_major(np.asarray(img_dataset[dg['weight_pb']].values)[:, :, 0, :, :])
    else:
        print('in normalize none ')
        norm_pb = None
    return _native.normalize_mosaic(image_dev, sum_weight_dev, ps, norm_pb, pb,
                                    np.asarray(gcf_dataset.attrs['oversampling']),
                                    norm_parms['pb_limit'] if norm_parms['pb_limit'] > 0 else 0.0,
                                    norm_parms['single_precision'])
