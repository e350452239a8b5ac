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


# (multi-dataset coordinate, subtable, its column): the id coordinates the
# reference's mxds carries for the subtables present, and for some of them the
# name coordinate over those ids (cngi/_utils/_io.py:37-64)
_MXDS_IDS = [("antenna_ids", "ANTENNA", "antenna_id"), ("field_ids", "FIELD", "field_id"),
             ("feed_ids", "FEED", "feed_id"), ("observation_ids", "OBSERVATION", "observation_id"),
             ("polarization_ids", "POLARIZATION", "d0"), ("source_ids", "SOURCE", "source_id"),
             ("spw_ids", "SPECTRAL_WINDOW", "spectral_window_id"), ("state_ids", "STATE", "state_id")]
_MXDS_NAMES = {"ANTENNA": ("antennas", "NAME"), "FIELD": ("fields", "NAME"),
               "OBSERVATION": ("observations", "PROJECT"), "SOURCE": ("sources", "NAME")}


def vis_xds_packager(xds_list):
    """The multi-dataset of (name, dataset) pairs: the datasets as attributes,
    plus id (and name) coordinates for the subtables among them."""
    from ..xds import Dataset
    mxds = Dataset(attrs=dict(xds_list))
    coords = {}
    for coord, table, column in _MXDS_IDS:
        if table not in mxds.attrs:
            continue
        sub = mxds.attrs[table]
        coords[coord] = ([coord], sub[column].values)
        if table in _MXDS_NAMES:
            name_coord, name_col = _MXDS_NAMES[table]
            coords[name_coord] = ([coord], sub[name_col].values)
    return mxds.assign_coords(coords) if coords else mxds
