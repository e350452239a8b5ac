"""ORACLE — test infrastructure only (see oracle/__init__.py).

numpy restatement of direction_rotate (ngcasa/imaging/direction_rotate.py):
the rotation matrices of calc_rotation_mats (:148-196, scipy Rotation as in
the reference), the directional cosine (:199-210), apply_rotation_matrix
(:213-234) and apply_phasor (:238-268), in float64.
"""
import numpy as np

C_LIGHT = 299792458.0


def directional_cosine(phase_center):
    ra, dec = phase_center
    return np.array([np.cos(ra) * np.cos(dec), np.sin(ra) * np.cos(dec), np.sin(dec)])


def rotation_mats(field_ids, phase_dirs, new_phase_center, common_tangent_reprojection=True):
    """uvw_rotmat [n, 3, 3] and phase_rotation [n, 3] for the fields listed
    (phase_dirs [n, 2] radians), as calc_rotation_mats (:148-196)."""
    from scipy.spatial.transform import Rotation as R
    ra_i, dec_i = new_phase_center
    rot_new = R.from_euler('XZ', [[np.pi / 2 - dec_i, -ra_i + np.pi / 2]]).as_matrix()[0]
    cos_new = directional_cosine(np.array([ra_i, dec_i]))
    n = len(field_ids)
    uvw_rotmat = np.zeros((n, 3, 3))
    phase_rotation = np.zeros((n, 3))
    for i in range(n):
        fpc = np.asarray(phase_dirs[i], dtype=np.float64)
        rot_f = R.from_euler('ZX', [[-np.pi / 2 + fpc[0], fpc[1] - np.pi / 2]]).as_matrix()[0]
        uvw_rotmat[i] = np.matmul(rot_new, rot_f).T
        if common_tangent_reprojection:
            uvw_rotmat[i, 2, 0:2] = 0.0
        phase_rotation[i] = np.matmul(rot_new, cos_new - directional_cosine(fpc))
    return uvw_rotmat, phase_rotation


def direction_rotate(vis, uvw, field, freq_chan, field_ids, uvw_rotmat, phase_rotation,
                     common_tangent_reprojection=True, single_precision=True):
    """apply_rotation_matrix + apply_phasor.  field [T, B] int; every time's
    valid field ids (> -1) must be one value."""
    T, B = uvw.shape[:2]
    uvw_rot = np.zeros(uvw.shape)
    pdir = np.zeros((T, B))
    end = 2 if common_tangent_reprojection else 3
    for t in range(T):
        ft = field[t]
        u = np.unique(ft[ft > -1])
        assert len(u) == 1, "field_id must be constant over baseline"
        k = int(np.nonzero(np.asarray(field_ids) == u[0])[0][0])
        uvw_rot[t] = uvw[t] @ uvw_rotmat[k]
        pdir[t] = uvw_rot[t, :, 0:end] @ phase_rotation[k, 0:end]
    ph = np.exp(2.0j * np.pi * pdir[:, :, None] * np.asarray(freq_chan)[None, None, :] / C_LIGHT)
    out = np.asarray(vis) * ph[:, :, :, None]
    if single_precision:
        out = out.astype(np.complex64).astype(np.complex128)
    return uvw_rot, out
