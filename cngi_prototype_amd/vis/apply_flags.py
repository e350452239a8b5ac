"""apply_flags on the MI355X path (cngi/vis/apply_flags.py:21-55).

Every data variable whose dims equal a flag variable's dims is replaced by a
copy that holds the dtype's NaN value wherever the flag is set (xarray's
``where(flag == 0)`` followed by ``astype`` of the original dtype): NaN for
floating and complex variables, the integer minimum for integers, True for
bool.  Device tensors go through one HIP pass (ngc_apply_flags) each; host
numpy variables are copied to the device for it and the result comes back to
the host, so the output keeps each variable's placement.

The gridders here also take the flag mask directly (the ``flag`` argument of
the native calls), which avoids writing NaN-poisoned copies at all.
"""
import numpy as np

from .. import _native
from .._utils._io import copy_structure
from ..device import to_device, to_host
from ..xds import DataArray


def _flagged(data, flag):
    import torch
    d = to_device(data)
    f = to_device(flag)
    if f.dtype not in (torch.bool, torch.uint8):
        f = f != 0        # the reference keeps samples where flag == 0 (NaN flags: flagged)
    out = _native.apply_flags(d, f)
    if not isinstance(data, torch.Tensor):
        return to_host(out)
    return out if data.is_cuda else out.cpu()


def apply_flags(mxds, vis, flags='FLAG'):
    """Apply flag variables to the other data variables of mxds.attrs[vis];
    returns a new mxds (the input is not modified)."""
    xds = mxds.attrs[vis]
    flags = np.atleast_1d(flags)
    flagged_xds = xds.copy(deep=False)
    for fv in flags:
        for dv in list(xds.data_vars):
            if dv == fv:
                continue   # the flags themselves stay
            if flagged_xds[dv].dims == flagged_xds[fv].dims:
                src = flagged_xds[dv]
                flagged_xds[dv] = DataArray(_flagged(src.data, flagged_xds[fv].data), src.dims,
                                            src.coords, src.attrs, src.name)
    out = copy_structure(mxds)
    out.attrs[vis] = flagged_xds
    return out
