"""write_vis: a multi-dataset to a vis.zarr directory (cngi/dio/write_vis.py:26-91).

Partitions whose name contains ``xds`` go to ``outfile/<name>``, the others
(subtables) to ``outfile/global/<name>``; the store is replaced.  Each is a
zarr v2 group as xarray writes it (``_ARRAY_DIMENSIONS``, non-dimension
coordinates named in the group's ``coordinates`` attribute, datetime64
values as CF integer nanoseconds since 1970-01-01), with the time the write took under
``<graph_name>_time`` in the group attributes and, with ``consolidated``, the
consolidated ``.zmetadata``.  The default compressor is zstd level 2 as a
plain zstd frame per chunk (numcodecs ``{"id": "zstd", "level": 2}``); the
reference's default wraps the same zstd level 2 in a Blosc frame.  A
numcodecs-style configuration dict (or a codec object with ``get_config``)
selects another codec among those ``_zarr`` encodes.
"""
import os
import shutil
import time

import numpy as np

from . import _zarr


def _config(compressor):
    if compressor is None:
        return {"id": "zstd", "level": 2}
    if hasattr(compressor, "get_config"):
        return compressor.get_config()
    return dict(compressor)


def dataset_variables(xds):
    """{name: (dims, numpy array, attrs)} of a Dataset's coordinates and data
    variables, and the names of its non-dimension coordinates."""
    out = {}
    for name, da in list(xds.coords.items()) + list(xds.data_vars.items()):
        out[name] = (tuple(da.dims), np.asarray(da.values), dict(getattr(da, "attrs", {})))
    dims = set(xds.dims)
    extra = [n for n in xds.coords if n not in dims]
    return out, extra


def write_group(path, xds, compressor=None, chunks=None, consolidated=True, extra_attrs=None):
    variables, extra = dataset_variables(xds)
    attrs = {k: v for k, v in xds.attrs.items() if not hasattr(v, "data_vars")}
    if extra:
        attrs["coordinates"] = " ".join(extra)
    attrs.update(extra_attrs or {})
    _zarr.write_group(path, variables, attrs, _config(compressor), chunks, consolidated)


def write_vis(mxds, outfile, chunks_on_disk=None, partition=None, consolidated=True,
              compressor=None, graph_name='write_zarr'):
    if partition is None:
        partition = list(mxds.attrs.keys())
    partition = [str(p) for p in np.atleast_1d(partition)]
    if os.path.isdir(outfile):
        shutil.rmtree(outfile)
    os.makedirs(outfile)
    for name in partition:
        xds = mxds.attrs[name]
        path = os.path.join(outfile, name) if "xds" in name else \
            os.path.join(outfile, "global", name)
        start = time.time()
        write_group(path, xds, compressor, chunks_on_disk, False)
        took = time.time() - start
        print('Time to store and execute graph for ', name, graph_name, took)
        # the timing goes into the group attributes before consolidation
        _zarr.update_group_attrs(path, {graph_name + '_time': took})
        if consolidated:
            _zarr.consolidate(path)
