"""MI355X-native implementation of ngCASA's imaging hot path.

The drop-in entry points live in :mod:`cngi_prototype_amd.imaging`
(make_psf, make_image, make_imaging_weight, the mosaic/A-term functions);
they run on the gfx950 library ``libngcasa_hip.so`` through
:mod:`cngi_prototype_amd._native`.
"""
__version__ = "0.1.0"
