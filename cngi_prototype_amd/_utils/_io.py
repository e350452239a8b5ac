"""Dataset copying without copying array data.

The reference starts every entry point with ``xds.copy(deep=True)`` on
dask-backed datasets, which copies structure and metadata but not the lazily
held visibility arrays.  ``copy_structure`` does the same for eagerly held
arrays (numpy or device tensors): new dataset objects, deep-copied attributes
(data_groups and subtables), shared array data.
"""
import copy


def copy_structure(ds):
    out = ds.copy(deep=False)
    attrs = {}
    for k, v in ds.attrs.items():
        if hasattr(v, 'data_vars') and hasattr(v, 'attrs'):
            attrs[k] = copy_structure(v)
        else:
            attrs[k] = copy.deepcopy(v)
    out.attrs = attrs
    return out
