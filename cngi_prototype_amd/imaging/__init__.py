"""ngCASA imaging entry points on the MI355X gridding path.

Same names, signatures, defaults and outputs as ngcasa.imaging for the
gridding hot path; the gridding, weight-density, degridding, GCF generation,
FFT and correction steps run on the GPU through libngcasa_hip.
"""
from .make_psf import make_psf
from .make_image import make_image
from .make_imaging_weight import make_imaging_weight
from .make_grid import make_grid
from .make_gridding_convolution_function import make_gridding_convolution_function
from .make_mosaic_pb import make_mosaic_pb
from .make_image_with_gcf import make_image_with_gcf
from .make_psf_with_gcf import make_psf_with_gcf
from .make_pb import make_pb
from .predict_modelvis_image import predict_modelvis_image
from .direction_rotate import direction_rotate
