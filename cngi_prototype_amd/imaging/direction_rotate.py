"""direction_rotate on the MI355X path (ngcasa/imaging/direction_rotate.py:28-248).

Rotates uvw and phase-rotates the visibilities of every field to a new phase
centre (joint mosaics use common-tangent reprojection).  The 3 x 3 rotation
matrices and phase-rotation vectors are computed on the host in closed form
(the reference builds them with scipy's Euler rotations, :156-194); the
per-sample rotation and phasor run on the GPU in FP64
(ngc_direction_rotate).  The new variables (UVW_ROT, DATA_ROT by default)
stay in device memory.
"""
import copy

import numpy as np

from .. import _native
from .._utils._check_parms import _check_parms, _check_sel_parms
from .._utils._io import copy_structure
from ..device import to_device
from ..xds import DataArray


def _rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def directional_cosine(phase_center):
    """Unit vector of (ra, dec) in the equatorial frame (direction_rotate.py:199-210)."""
    ra, dec = phase_center
    return np.array([np.cos(ra) * np.cos(dec), np.sin(ra) * np.cos(dec), np.sin(dec)])


def rotation_matrices(field_phase_dirs, new_phase_center, common_tangent_reprojection=True):
    """uvw rotation matrices [n, 3, 3] and phase-rotation vectors [n, 3]
    (calc_rotation_mats, :148-196).  The frame of the new phase centre is
    Rx(pi/2 - dec) Rz(pi/2 - ra), that of a field Rz(ra - pi/2) Rx(dec - pi/2)
    (intrinsic X-Z and Z-X rotations); the uvw matrix is the transpose of
    their product, with its w row's u and v terms dropped under common-tangent
    reprojection (CASA FTMachine::girarUVW)."""
    ra_i, dec_i = float(new_phase_center[0]), float(new_phase_center[1])
    rot_new = _rx(np.pi / 2 - dec_i) @ _rz(np.pi / 2 - ra_i)
    cos_new = directional_cosine((ra_i, dec_i))
    fpd = np.asarray(field_phase_dirs, dtype=np.float64).reshape(-1, 2)
    mats = np.zeros((len(fpd), 3, 3))
    vecs = np.zeros((len(fpd), 3))
    for i, (ra, dec) in enumerate(fpd):
        rot_f = _rz(ra - np.pi / 2) @ _rx(dec - np.pi / 2)
        mats[i] = (rot_new @ rot_f).T
        if common_tangent_reprojection:
            mats[i, 2, 0:2] = 0.0
        vecs[i] = rot_new @ (cos_new - directional_cosine((ra, dec)))
    return mats, vecs


def _check_rotation_parms(rotation_parms):
    """_check_imaging_parms.py:96-106."""
    import numbers
    ok = True
    if not _check_parms(rotation_parms, 'new_phase_center', [list],
                        list_acceptable_data_types=[numbers.Number], list_len=2):
        ok = False
    if not _check_parms(rotation_parms, 'common_tangent_reprojection', [bool], default=True):
        ok = False
    if not _check_parms(rotation_parms, 'single_precision', [bool], default=True):
        ok = False
    return ok


def direction_rotate(mxds, rotation_parms, sel_parms):
    """Same parameters, defaults and outputs as the reference."""
    import torch
    print('######################### Start direction_rotate #########################')
    _mxds = copy_structure(mxds)
    _sel_parms = copy.deepcopy(sel_parms)
    _rotation_parms = copy.deepcopy(rotation_parms)
    assert (_check_rotation_parms(_rotation_parms)), \
        "######### ERROR: rotation_parms checking failed"
    assert ('xds' in _sel_parms), "######### ERROR: xds must be specified in sel_parms"
    _vis_dataset = _mxds.attrs[sel_parms['xds']]
    _check_sel_parms(_vis_dataset, _sel_parms,
                     new_or_modified_data_variables={'uvw': 'UVW_ROT', 'data': 'DATA_ROT'})
    dg_in, dg_out = _sel_parms['data_group_in'], _sel_parms['data_group_out']
    field = to_device(_vis_dataset['FIELD_ID'].data, dtype=torch.int32)
    valid = field[field > -1]
    field_ids = torch.unique(valid).cpu().numpy() if valid.numel() else np.array([0])
    fds = _mxds.attrs['FIELD']
    fid_coord = np.asarray(fds.coords['field_id'].values)
    phase_dir = np.asarray(fds['PHASE_DIR'].values)[:, 0, :]
    sel = [int(np.nonzero(fid_coord == f)[0][0]) for f in field_ids]
    mats, vecs = rotation_matrices(phase_dir[sel], _rotation_parms['new_phase_center'],
                                   _rotation_parms['common_tangent_reprojection'])
    uvw = to_device(_vis_dataset[dg_in['uvw']].data)
    vis = to_device(_vis_dataset[dg_in['data']].data)
    uvw_rot, vis_rot = _native.direction_rotate(
        uvw, vis, field, field_ids, mats, vecs, np.asarray(_vis_dataset.coords['chan'].values),
        _rotation_parms['common_tangent_reprojection'], _rotation_parms['single_precision'])
    _vis_dataset[dg_out['uvw']] = DataArray(uvw_rot, dims=_vis_dataset[dg_in['uvw']].dims)
    _vis_dataset[dg_out['data']] = DataArray(vis_rot, dims=_vis_dataset[dg_in['data']].dims)
    _vis_dataset.attrs['data_groups'][0] = {**_vis_dataset.attrs['data_groups'][0],
                                            **{dg_out['id']: dg_out}}
    print('######################### Created graph for direction_rotate #########################')
    return _mxds
