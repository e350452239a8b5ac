"""Visibility stores on disk: the vis.zarr directory layout that the
reference's convert_ms and write_vis produce, read into and written from the
package's Dataset objects (cngi/dio/read_vis.py, write_vis.py,
describe_vis.py)."""
from .read_vis import read_vis
from .write_vis import write_vis
from .describe_vis import describe_vis

__all__ = ["read_vis", "write_vis", "describe_vis"]
