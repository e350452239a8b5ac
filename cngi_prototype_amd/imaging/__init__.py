"""ngCASA imaging entry points on the MI355X gridding path.

Same names, signatures, defaults and outputs as ngcasa.imaging for the
gridding hot path; the gridding, weight-density, degridding, FFT and
correction steps run on the GPU through libngcasa_hip.
"""
from .make_psf import make_psf
from .make_image import make_image
from .make_imaging_weight import make_imaging_weight
