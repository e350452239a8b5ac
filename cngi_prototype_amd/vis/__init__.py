"""Visibility-dataset operations on the MI355X path (cngi.vis)."""
from .apply_flags import apply_flags
