"""Minimal stand-in for the xarray objects the ngCASA imaging API exchanges.

The reference passes ``xarray.Dataset`` objects: a multi-dataset *mxds* whose
``attrs`` hold per-DDI visibility datasets (``xds0`` ...) plus subtables, and
image datasets with dims ``(l, m, time, chan, pol)``
(cngi/_utils/_io.py:37-64, cngi/image/make_empty_sky_image.py:18-82).
xarray is not required here: this module gives the small subset of the
Dataset/DataArray interface that the imaging functions use (item and
attribute access, ``coords``, ``attrs``, ``dims``, ``copy``,
``assign_coords``, ``.values``/``.data``), so the same calls work on these
objects and on real xarray objects alike.

A DataArray's ``data`` may be a numpy array or a torch tensor already resident
on the GPU; the imaging functions accept either.
"""
import copy as _copy

import numpy as np


def _to_numpy(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:
        pass
    return np.asarray(x)


class DataArray:
    def __init__(self, data, dims=None, coords=None, attrs=None, name=None):
        if isinstance(data, DataArray):
            dims = data.dims if dims is None else dims
            data = data.data
        if not _is_tensor(data):
            data = np.asarray(data)
        if dims is None:
            dims = tuple(f"dim_{i}" for i in range(len(data.shape)))
        if isinstance(dims, str):
            dims = (dims,)
        dims = tuple(dims)
        if len(dims) != len(data.shape):
            raise ValueError(f"dims {dims} do not match shape {tuple(data.shape)}")
        self.data = data
        self.dims = dims
        self.attrs = dict(attrs or {})
        self.name = name
        self.coords = dict(coords or {})

    @property
    def values(self):
        return _to_numpy(self.data)

    @property
    def shape(self):
        return tuple(self.data.shape)

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def sizes(self):
        return dict(zip(self.dims, self.shape))

    def __len__(self):
        return self.shape[0]

    def __array__(self, dtype=None, copy=None):
        v = self.values
        return v.astype(dtype) if dtype is not None else v

    def __getitem__(self, idx):
        return DataArray(self.values[idx])

    def mean(self):
        return DataArray(np.asarray(np.mean(self.values)), dims=())

    def copy(self, deep=True):
        d = self.data
        if deep:
            d = d.clone() if _is_tensor(d) else np.array(d, copy=True)
        return DataArray(d, self.dims, self.coords, _copy.deepcopy(self.attrs), self.name)

    def __repr__(self):
        return f"<DataArray {self.name or ''} dims={self.dims} shape={self.shape} dtype={self.dtype}>"


def _is_tensor(x):
    try:
        import torch
        return isinstance(x, torch.Tensor)
    except ImportError:
        return False


def _as_dataarray(v):
    if isinstance(v, DataArray):
        return v
    if isinstance(v, tuple) and len(v) == 2 and isinstance(v[0], (tuple, list, str)):
        return DataArray(v[1], v[0])
    if hasattr(v, "dims") and hasattr(v, "data"):
        return DataArray(v.data, v.dims)
    return DataArray(v)


class Dataset:
    def __init__(self, data_vars=None, coords=None, attrs=None):
        object.__setattr__(self, "data_vars", {})
        object.__setattr__(self, "coords", {})
        object.__setattr__(self, "attrs", dict(attrs or {}))
        for k, v in (coords or {}).items():
            self.coords[k] = _as_dataarray(v) if not np.isscalar(v) else DataArray(np.asarray(v), ())
        for k, v in (data_vars or {}).items():
            self.data_vars[k] = _as_dataarray(v)

    # -- item / attribute access ------------------------------------------
    def __getitem__(self, key):
        if key in self.data_vars:
            return self.data_vars[key]
        if key in self.coords:
            return self.coords[key]
        raise KeyError(key)

    def __setitem__(self, key, value):
        self.data_vars[key] = _as_dataarray(value)

    def __contains__(self, key):
        return key in self.data_vars or key in self.coords

    def __getattr__(self, key):
        if key.startswith("__"):
            raise AttributeError(key)
        dv = object.__getattribute__(self, "data_vars")
        co = object.__getattribute__(self, "coords")
        at = object.__getattribute__(self, "attrs")
        if key in dv:
            return dv[key]
        if key in co:
            return co[key]
        if key in at:
            return at[key]
        raise AttributeError(key)

    def __setattr__(self, key, value):
        if key in ("data_vars", "coords", "attrs"):
            object.__setattr__(self, key, value)
        else:
            raise AttributeError("use item assignment to add variables")

    # -- xarray-like helpers -----------------------------------------------
    @property
    def dims(self):
        d = {}
        for v in list(self.data_vars.values()) + list(self.coords.values()):
            for name, n in zip(v.dims, v.shape):
                d.setdefault(name, n)
        return d

    @property
    def sizes(self):
        return self.dims

    def assign_coords(self, coords):
        out = self.copy(deep=False)
        for k, v in coords.items():
            if isinstance(v, tuple) and len(v) == 2 and isinstance(v[0], (str, tuple, list)):
                out.coords[k] = DataArray(_to_numpy(v[1]) if not _is_tensor(v[1]) else v[1], v[0])
            else:
                arr = np.asarray(_to_numpy(v))
                out.coords[k] = DataArray(arr, (k,) if arr.ndim == 1 else tuple(
                    f"{k}_{i}" for i in range(arr.ndim)))
        return out

    def copy(self, deep=True):
        out = Dataset()
        for k, v in self.data_vars.items():
            out.data_vars[k] = v.copy(deep)
        for k, v in self.coords.items():
            out.coords[k] = v.copy(deep)
        attrs = {}
        for k, v in self.attrs.items():
            if isinstance(v, Dataset):
                attrs[k] = v.copy(deep)
            else:
                attrs[k] = _copy.deepcopy(v) if deep else v
        object.__setattr__(out, "attrs", attrs)
        return out

    def __repr__(self):
        lines = ["<Dataset>", "  dims: " + str(self.dims)]
        for k, v in self.coords.items():
            lines.append(f"  coord {k}: {v.dims} {v.shape}")
        for k, v in self.data_vars.items():
            lines.append(f"  var   {k}: {v.dims} {v.shape} {v.dtype}")
        for k in self.attrs:
            lines.append(f"  attr  {k}")
        return "\n".join(lines)
