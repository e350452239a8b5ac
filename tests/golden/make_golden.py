"""Writes the golden fixtures of this folder.

sin_projection_kat.json: the casacore toPixelWCS known-answer case quoted in
ngcasa/imaging/make_gridding_convolution_function.py:569-581 (a comment block
of that file; the numbers are transcribed here as data).  It fixes the
RA---SIN/DEC--SIN world-to-pixel map used for the mosaic phase gradients
(crpix, cdelt in degrees, LONPOLE 180, pixel indices 1-based as in WCS).
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    kat = {
        "source": "ngcasa/imaging/make_gridding_convolution_function.py:569-581",
        "crpix": [120.0, 120.0],
        "cdelt_deg": [-5e-05, 5e-05],
        "crval_deg": [180.46846189999996568, -18.863873247222226581],
        "world_deg": [-179.5337374791666889, -18.863873258333338612],
        "pixel": [161.6249842951184803, 119.99951947142589859],
    }
    with open(os.path.join(HERE, "sin_projection_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)


if __name__ == "__main__":
    main()
