"""describe_vis: one summary row per visibility partition of a vis.zarr
store (cngi/dio/describe_vis.py:21-57)."""
import os

import numpy as np

from .read_vis import open_group


def describe_vis(infile):
    """pandas DataFrame indexed by partition: spw_id, pol_id, times,
    baselines, chans, pols and size_MB (all arrays, rounded up to MiB)."""
    import pandas as pd
    infile = os.path.expanduser(infile)
    rows = []
    parts = sorted(d for d in os.listdir(infile) if os.path.isdir(os.path.join(infile, d)))
    for part in parts:
        if part.startswith('global'):
            continue
        xds = open_group(os.path.join(infile, part))
        nbytes = sum(np.asarray(v.values).nbytes
                     for v in list(xds.data_vars.values()) + list(xds.coords.values()))
        dims = xds.dims
        rows.append({'xds': part, 'spw_id': xds['spw_id'].values[0],
                     'pol_id': xds['pol_id'].values[0], 'times': dims['time'],
                     'baselines': dims['baseline'], 'chans': dims['chan'], 'pols': dims['pol'],
                     'size_MB': int(np.ceil(nbytes / 1024 ** 2))})
    return pd.DataFrame(rows).set_index('xds').sort_index()
