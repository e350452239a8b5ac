from .make_empty_sky_image import make_empty_sky_image
from .fit_gaussian import fit_gaussian
