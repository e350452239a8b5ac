"""Image coordinates for a new sky image (cngi/image/make_empty_sky_image.py:18-82).

Dims (l, m, time, chan, pol); l/m coordinates are (index - N//2)*cell, and the
right_ascension/declination coordinates come from the SIN (orthographic)
projection about the phase centre, evaluated in closed form (the reference
uses astropy's WCS with crpix = N//2 and origin 1, which is the same map).
"""
import numpy as np


def sin_pixel_to_world(l, m, ra0, dec0):
    """Inverse SIN projection: direction cosines (l, m) about (ra0, dec0) to
    (ra, dec) in radians (Calabretta & Greisen 2002, zenithal, LONPOLE = 180)."""
    l = np.asarray(l, dtype=np.float64)
    m = np.asarray(m, dtype=np.float64)
    n = np.sqrt(1.0 - l * l - m * m)
    dec = np.arcsin(m * np.cos(dec0) + n * np.sin(dec0))
    ra = ra0 + np.arctan2(l, n * np.cos(dec0) - m * np.sin(dec0))
    return ra, dec


def make_empty_sky_image(xds, phase_center, image_size, cell_size, chan_coords, chan_width,
                         pol_coords, time_coords, direction_reference='FK5', projection='SIN',
                         spectral_reference='lsrk', velocity_type='radio', unit='Jy/beam'):
    if projection != 'SIN':
        raise NotImplementedError("only the SIN projection is provided")
    image_size = np.asarray(image_size)
    cell_size = np.asarray(cell_size, dtype=np.float64)
    image_center = image_size // 2
    l = np.arange(-image_center[0], image_size[0] - image_center[0]) * cell_size[0]
    m = np.arange(-image_center[1], image_size[1] - image_center[1]) * cell_size[1]
    L, M = np.meshgrid(l, m, indexing='ij')
    ra, dec = sin_pixel_to_world(L, M, float(phase_center[0]), float(phase_center[1]))
    coords = {'time': np.asarray(time_coords), 'chan': np.asarray(chan_coords),
              'pol': np.asarray(pol_coords), 'chan_width': ('chan', np.asarray(chan_width)),
              'l': l, 'm': m, 'right_ascension': (('l', 'm'), ra),
              'declination': (('l', 'm'), dec)}
    xds.attrs['axis_units'] = ['rad', 'rad', 'time', 'Hz', 'pol']
    xds = xds.assign_coords(coords)
    xds.attrs['direction_reference'] = direction_reference
    xds.attrs['spectral_reference'] = spectral_reference
    xds.attrs['velocity_type'] = velocity_type
    xds.attrs['unit'] = unit
    return xds
