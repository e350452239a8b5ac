"""apply_flags on the GPU (ngc_apply_flags) against the reference's
semantics restated in numpy: where(flag == 0) then astype(original dtype)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reference_flagging(data, flag):
    """xarray's dv.where(fv == 0).astype(dv.dtype), in numpy."""
    with np.errstate(invalid="ignore"):
        return np.where(flag == 0, data, np.nan).astype(data.dtype)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128, np.float32, np.float64,
                                   np.int32, np.int64, np.bool_])
def test_apply_flags_kernel_matches_where_astype(gpu, dtype):
    import torch
    import cngi_prototype_amd._native as nat
    rng = np.random.default_rng(5)
    shape = (37, 11, 5, 2)
    if np.issubdtype(dtype, np.complexfloating):
        data = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(dtype)
    elif dtype == np.bool_:
        data = rng.uniform(size=shape) < 0.5
    else:
        data = (rng.normal(size=shape) * 1000).astype(dtype)
    flag = rng.uniform(size=shape) < 0.3
    got = nat.apply_flags(torch.as_tensor(data).cuda(), torch.as_tensor(flag).cuda())
    got = got.cpu().numpy()
    ref = _reference_flagging(data, flag)
    if np.issubdtype(dtype, np.inexact):
        assert np.array_equal(np.isnan(got), np.isnan(ref))
        assert np.array_equal(got[~flag], ref[~flag])
    else:
        assert np.array_equal(got, ref)
    # in place
    d = torch.as_tensor(data).cuda()
    nat.apply_flags(d, torch.as_tensor(flag.astype(np.uint8)).cuda(), out=d)
    assert np.array_equal(d.cpu().numpy(), got, equal_nan=np.issubdtype(dtype, np.inexact))


def test_apply_flags_entry_point(gpu):
    """Only variables with FLAG's dims are flagged; the flags, UVW and the
    input dataset are untouched; device variables stay on the device."""
    import torch
    from cngi_prototype_amd.vis import apply_flags
    from cngi_prototype_amd.xds import Dataset
    rng = np.random.default_rng(2)
    T, B, C, P = 6, 10, 4, 2
    vis = (rng.normal(size=(T, B, C, P)) + 1j * rng.normal(size=(T, B, C, P))).astype(np.complex64)
    w = rng.uniform(0.5, 2, size=(T, B, C, P)).astype(np.float32)
    flag = rng.uniform(size=(T, B, C, P)) < 0.2
    uvw = rng.normal(size=(T, B, 3))
    dims4 = ["time", "baseline", "chan", "pol"]
    xds = Dataset(data_vars={"DATA": (dims4, torch.as_tensor(vis).cuda()),
                             "DATA_WEIGHT": (dims4, w), "FLAG": (dims4, flag),
                             "UVW": (["time", "baseline", "uvw_index"], uvw)})
    mxds = Dataset(attrs={"xds0": xds})
    out = apply_flags(mxds, "xds0", flags="FLAG")
    fx = out.attrs["xds0"]
    d = fx["DATA"].data
    assert isinstance(d, torch.Tensor) and d.is_cuda
    d = d.cpu().numpy()
    assert np.all(np.isnan(d[flag])) and np.array_equal(d[~flag], vis[~flag])
    wt = fx["DATA_WEIGHT"].data
    assert isinstance(wt, np.ndarray) and wt.dtype == np.float32
    assert np.all(np.isnan(wt[flag])) and np.array_equal(wt[~flag], w[~flag])
    assert np.array_equal(fx["FLAG"].data, flag)
    assert np.array_equal(fx["UVW"].data, uvw)
    # the input is not modified
    assert not np.isnan(mxds.attrs["xds0"]["DATA_WEIGHT"].data).any()
