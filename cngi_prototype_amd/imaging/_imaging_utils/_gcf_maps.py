"""Host-side integer logic and geometry of the A-term GCF.

* cf_baseline_map: dish-type pair of every baseline
  (make_gridding_convolution_function.py:512-528).  As in the reference, only
  baselines whose (type(ant1), type(ant2)) equals an ordered pair (i <= j) are
  matched; a reversed pair keeps the default 0.
* cf_chan_map: frequencies at which convolution functions are computed and
  the nearest one for every channel (:536-564).
* conv_support: the support cut of calc_conv_size (:414-457), walking the
  weight kernel from the centre along +u and +v until it drops to
  support_cut_level of its peak.
* field_pixel_offsets: pixel offset of each field centre from the image phase
  centre under the SIN projection (the reference's astropy all_world2pix
  with crpix = N//2, origin 1), in closed form (Calabretta & Greisen 2002).
"""
import numpy as np


def cf_baseline_map(unique_ant_indx, baseline_ant, n_unique_ant):
    n_pairs = (n_unique_ant * n_unique_ant + n_unique_ant) // 2
    pairs = np.array([(i, j) for i in range(n_unique_ant) for j in range(i, n_unique_ant)],
                     dtype=int).reshape(n_pairs, 2)
    types = np.asarray(unique_ant_indx)[np.asarray(baseline_ant)]  # [n_baseline, 2]
    cmap = np.zeros(len(types), dtype=int)
    for k, (i, j) in enumerate(pairs):
        cmap[(types[:, 0] == i) & (types[:, 1] == j)] = k
    return cmap, pairs


def cf_chan_map(freq_chan, chan_tolerance_factor):
    freq_chan = np.asarray(freq_chan, dtype=np.float64)
    fmax, fmin = np.max(freq_chan), np.min(freq_chan)
    n_pb = int(np.floor((fmax - fmin) / (fmax * chan_tolerance_factor)) + 0.5)
    n_pb = max(n_pb, 1)
    if n_pb >= len(freq_chan):
        return np.arange(len(freq_chan)), freq_chan
    width = (fmax - fmin) / n_pb
    pb_freq = np.arange(n_pb) * width + fmin + width / 2
    cmap = np.array([int(np.argmin(np.abs(pb_freq - f))) for f in freq_chan], dtype=int)
    return cmap, pb_freq


def conv_support(centre_row, centre_col, abs_max, abs_min, imsize, support_cut_level,
                 oversampling, max_support):
    """Support [s, s] of one weight-kernel plane from its centre row/column."""
    cut = support_cut_level * abs_max
    assert abs_min < cut, "######### ERROR: support_cut_level too small or imsize too small."
    supports = []
    for line, n, os_, ms in ((centre_row, imsize[0], oversampling[0], max_support[0]),
                             (centre_col, imsize[1], oversampling[1], max_support[1])):
        i = n // 2
        while line[i] > cut:
            i += 1
            assert i < n, "######### ERROR: support_cut_level too small or imsize too small."
        s = (int(0.5 + (i - n // 2) / os_) + 1) * 2 + 1
        assert s < ms, ("######### ERROR: support_cut_level too small or imsize too small."
                        + str(s) + ",*," + str(ms))
        supports.append(s)
    s = max(supports)
    return [s, s]


def field_pixel_offsets(field_dir, phase_center, cell_size):
    """(dx, dy) pixel offsets of field directions [n_field, 2] (radians)."""
    fd = np.asarray(field_dir, dtype=np.float64).reshape(-1, 2)
    ra0, dec0 = float(phase_center[0]), float(phase_center[1])
    dra = fd[:, 0] - ra0
    l = np.cos(fd[:, 1]) * np.sin(dra)
    m = np.sin(fd[:, 1]) * np.cos(dec0) - np.cos(fd[:, 1]) * np.sin(dec0) * np.cos(dra)
    deg = 180 / np.pi
    return np.stack([(l * deg) / (cell_size[0] * deg), (m * deg) / (cell_size[1] * deg)], 1)


def gcf_kernels(pb_freq, dish, blockage, pairs, casa, pb_cell, npad, oversampling, max_support,
                support_cut_level, supports=None):
    """The a_term branch's kernels on the GPU (make_gridding_convolution_function.py:
    181-280): weight kernels (power patterns) first, whose support cut decides
    the supports unless given, then the convolution kernels (voltage-pattern
    products), each cropped and normalised.  Returns device float64
    (WEIGHT_CONV_KERNEL, CONV_KERNEL) [n_pairs, n_freq, 1, K, K] and SUPPORT
    [n_pairs, n_freq, 1, 2]."""
    from ... import _native
    os_ = np.asarray(oversampling)
    resize = (np.asarray(max_support) + 1) * os_
    kernels = []
    for ipower in (2, 1):
        full, amax, amin, row, col = _native.gcf_patterns(pb_freq, dish, blockage, pairs, casa,
                                                          ipower, pb_cell, npad)
        if ipower == 2 and supports is None:
            supports = np.zeros((len(pairs), len(pb_freq), 1, 2), dtype=int)
            for k in range(len(pairs)):
                for f in range(len(pb_freq)):
                    supports[k, f, 0] = conv_support(row[k, f], col[k, f], amax[k, f],
                                                     amin[k, f], npad, support_cut_level, os_,
                                                     max_support)
        resized = _native.gcf_resize(full, supports[:, :, 0], os_, resize)
        del full
        kernels.append(resized.unsqueeze(2))
    return kernels[0], kernels[1], supports
