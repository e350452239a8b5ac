"""direction_rotate on the GPU against the oracle (oracle/rotate.py)."""
import numpy as np
import pytest

from conftest import rel_err
from oracle import rotate

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ct", [True, False])
@pytest.mark.parametrize("single", [True, False])
def test_direction_rotate_matches_oracle(gpu, ct, single):
    from helpers import mosaic_case
    from cngi_prototype_amd.imaging import direction_rotate
    c = mosaic_case()
    mxds = c["mxds"]
    new = [float(c["phase_center"][0]) + 2e-5, float(c["phase_center"][1]) - 1e-5]
    out = direction_rotate(mxds, {"new_phase_center": new, "common_tangent_reprojection": ct,
                                  "single_precision": single}, {"xds": "xds0"})
    xds = out.attrs["xds0"]
    dg = xds.attrs["data_groups"][0]
    gid = max(dg, key=int)
    assert dg[gid]["uvw"] == "UVW_ROT" and dg[gid]["data"] == "DATA_ROT"
    x0 = mxds.attrs["xds0"]
    field = np.asarray(x0.FIELD_ID.values)
    fids = np.unique(field[field > -1])
    pdir = np.asarray(mxds.attrs["FIELD"]["PHASE_DIR"].values)[fids, 0, :]
    M, P = rotate.rotation_mats(fids, pdir, new, ct)
    uvw_ref, vis_ref = rotate.direction_rotate(np.asarray(x0.DATA.values),
                                               np.asarray(x0.UVW.values), field,
                                               x0.coords["chan"].values, fids, M, P, ct, single)
    uvw_g = np.asarray(xds["UVW_ROT"].values)
    vis_g = np.asarray(xds["DATA_ROT"].values).astype(np.complex128)
    assert rel_err(uvw_g, uvw_ref) < 1e-14
    fin = np.isfinite(vis_ref)
    assert np.array_equal(fin, np.isfinite(vis_g))
    assert rel_err(vis_g[fin], vis_ref[fin]) < (1e-7 if single else 1e-9)
    # the input dataset keeps its groups
    assert "DATA_ROT" not in mxds.attrs["xds0"]


def test_direction_rotate_refuses_mixed_fields(gpu):
    from helpers import mosaic_case
    from cngi_prototype_amd.imaging import direction_rotate
    from cngi_prototype_amd.xds import DataArray
    c = mosaic_case()
    x0 = c["mxds"].attrs["xds0"]
    f = np.array(x0.FIELD_ID.values, copy=True)
    f[0, 3] = (f[0, 0] + 1) % 4
    x0["FIELD_ID"] = DataArray(f, dims=["time", "baseline"])
    with pytest.raises(AssertionError):
        direction_rotate(c["mxds"], {"new_phase_center": list(map(float, c["phase_center"]))},
                         {"xds": "xds0"})
