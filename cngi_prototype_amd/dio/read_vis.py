"""read_vis: a vis.zarr directory into a multi-dataset (cngi/dio/read_vis.py:21-203).

Partitions are the store's sub-directories (``xds0``, ``xdsa0``, ...); the
partition ``global`` stands for every subtable under ``global/`` (``ANTENNA``,
``FIELD``, ...), which are named without the prefix.  Each partition is a
zarr v2 group read by :mod:`._zarr` into a Dataset whose arrays are numpy
arrays, loaded whole: there is no dask here, so ``chunks`` and
``overwrite_encoded_chunks`` only keep the reference's signature.  S3 stores
are not read (no S3 client in this package).
"""
import os

import numpy as np

from ..xds import Dataset
from .._utils._io import vis_xds_packager
from . import _zarr


def open_group(path, consolidated=True):
    """One zarr group as a Dataset: variables named like a dimension, or named
    by a 'coordinates' attribute (a variable's or the group's), are
    coordinates."""
    variables, gattrs = _zarr.read_group(path, consolidated)
    gattrs = dict(gattrs)
    coord_names = set(str(gattrs.pop("coordinates", "")).split())
    for dims, _, attrs in variables.values():
        coord_names.update(str(attrs.pop("coordinates", "")).split())
    all_dims = {d for dims, _, _ in variables.values() for d in dims}
    data_vars, coords = {}, {}
    for name, (dims, arr, attrs) in variables.items():
        (coords if (name in all_dims or name in coord_names) else data_vars)[name] = (dims, arr)
    ds = Dataset(data_vars=data_vars, coords=coords, attrs=gattrs)
    for name, (dims, arr, attrs) in variables.items():
        ds[name].attrs.update(attrs)
    return ds


def read_vis(infile, partition=None, chunks=None, consolidated=True,
             overwrite_encoded_chunks=True, **kwargs):
    """Read a vis.zarr store into a multi-dataset (attrs: partition name ->
    Dataset, with the subtables' id and name coordinates)."""
    print('overwrite_encoded_chunks', overwrite_encoded_chunks)
    if str(infile).lower().startswith("s3"):
        raise NotImplementedError("S3 stores are not supported (no S3 client here)")
    infile = os.path.expanduser(infile)
    if partition is None:
        partition = os.listdir(infile)
    partition = [str(p) for p in np.atleast_1d(partition)]
    if "global" in partition and os.path.isdir(os.path.join(infile, "global")):
        partition += sorted("global/" + t for t in os.listdir(os.path.join(infile, "global")))
    xds_list = []
    for part in partition:
        if part == "global":
            continue
        path = os.path.join(infile, part)
        try:
            if os.path.isdir(path):
                xds_list.append((part.replace("global/", ""), open_group(path, consolidated)))
        except Exception:
            print('Can not open ', part)
    return vis_xds_packager(xds_list)
