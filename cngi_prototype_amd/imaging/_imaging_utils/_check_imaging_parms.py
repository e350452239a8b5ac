"""Imaging parameter normalisation (ngcasa/imaging/_imaging_utils/_check_imaging_parms.py).

Same keys, defaults, ranges and unit conversions as the reference:
``_check_grid_parms`` (:23-41) converts cell sizes from arcsec to radians,
negates cell[0] and sets ``image_size_padded = int(fft_padding*image_size)``.
One optional key is added, ``grid_parms['precision']`` (absent by default),
choosing the GPU form: 'fp64' (plain FP64 drop-in, the default),
'fp64_tiled', 'fp32_kahan' or 'fp32'.
"""
import numbers

import numpy as np

from ..._utils._check_parms import INT_TYPES, _check_parms

ARCSEC_TO_RAD = np.pi / (3600 * 180)
PRECISIONS = ['fp64', 'fp64_tiled', 'fp32_kahan', 'fp32']


def _check_grid_parms(grid_parms):
    """_check_imaging_parms.py:23-41."""
    parms_passed = True
    if not (_check_parms(grid_parms, 'image_size', [list], list_acceptable_data_types=[INT_TYPES],
                         list_len=2)):
        parms_passed = False
    if not (_check_parms(grid_parms, 'image_center', [list],
                         list_acceptable_data_types=[INT_TYPES], list_len=2,
                         default=np.array(grid_parms.get('image_size', [0, 0])) // 2)):
        parms_passed = False
    if not (_check_parms(grid_parms, 'cell_size', [list],
                         list_acceptable_data_types=[numbers.Number], list_len=2)):
        parms_passed = False
    if not (_check_parms(grid_parms, 'fft_padding', [numbers.Number], default=1.2,
                         acceptable_range=[1, 10])):
        parms_passed = False
    if not (_check_parms(grid_parms, 'chan_mode', [str], acceptable_data=['cube', 'continuum'],
                         default='cube')):
        parms_passed = False
    if 'precision' in grid_parms:
        if not (_check_parms(grid_parms, 'precision', [str], acceptable_data=PRECISIONS)):
            parms_passed = False
    if parms_passed:
        grid_parms['image_size'] = np.array(grid_parms['image_size']).astype(int)
        grid_parms['image_size_padded'] = (grid_parms['fft_padding'] *
                                           grid_parms['image_size']).astype(int)
        grid_parms['image_center'] = np.array(grid_parms['image_center'])
        grid_parms['cell_size'] = ARCSEC_TO_RAD * np.array(grid_parms['cell_size'])
        grid_parms['cell_size'][0] = -grid_parms['cell_size'][0]
    return parms_passed


def _check_gcf_parms(gcf_parms):
    """_check_imaging_parms.py:45-84."""
    parms_passed = True
    if not (_check_parms(gcf_parms, 'function', [str], acceptable_data=['casa_airy', 'airy'],
                         default='casa_airy')):
        parms_passed = False
    for key in ('freq_chan', 'list_dish_diameters', 'list_blockage_diameters', 'unique_ant_indx',
                'pol'):
        if not (_check_parms(gcf_parms, key, [list, np.array],
                             list_acceptable_data_types=[numbers.Number], list_len=-1)):
            parms_passed = False
    if not (_check_parms(gcf_parms, 'chan_tolerance_factor', [numbers.Number], default=0.005)):
        parms_passed = False
    if not (_check_parms(gcf_parms, 'oversampling', [list, np.array],
                         list_acceptable_data_types=[INT_TYPES], list_len=2, default=[10, 10])):
        parms_passed = False
    if not (_check_parms(gcf_parms, 'max_support', [list, np.array],
                         list_acceptable_data_types=[INT_TYPES], list_len=2, default=[15, 15])):
        parms_passed = False
    if not (_check_parms(gcf_parms, 'support_cut_level', [numbers.Number], default=2.5 * 10 ** -2)):
        parms_passed = False
    if not (_check_parms(gcf_parms, 'a_chan_num_chunk', [INT_TYPES], default=3)):
        parms_passed = False
    if gcf_parms.get('function') in ('airy', 'casa_airy'):
        if len(gcf_parms.get('list_dish_diameters', [])) != len(
                gcf_parms.get('list_blockage_diameters', [])):
            print('######### ERROR:Parameter ',
                  'list_dish_diameters and list_blockage_diameters must be the same length.')
            parms_passed = False
    if parms_passed:
        gcf_parms['oversampling'] = np.array(gcf_parms['oversampling']).astype(int)
        gcf_parms['max_support'] = np.array(gcf_parms['max_support']).astype(int)
        for key in ('freq_chan', 'list_dish_diameters', 'list_blockage_diameters',
                    'unique_ant_indx', 'pol'):
            gcf_parms[key] = np.array(gcf_parms[key])
    return parms_passed


def _check_pb_parms(img_dataset, pb_parms):
    """_check_imaging_parms.py:148-164."""
    parms_passed = True
    if not (_check_parms(pb_parms, 'function', [str], default='casa_airy')):
        parms_passed = False
    if not (_check_parms(pb_parms, 'list_dish_diameters', [list],
                         list_acceptable_data_types=[numbers.Number], list_len=-1)):
        parms_passed = False
    if not (_check_parms(pb_parms, 'list_blockage_diameters', [list],
                         list_acceptable_data_types=[numbers.Number], list_len=-1)):
        parms_passed = False
    if len(pb_parms['list_dish_diameters']) != len(pb_parms['list_blockage_diameters']):
        print('######### ERROR:Parameter ', 'list_dish_diameters and list_blockage_diameters '
              'must be the same length.')
        parms_passed = False
    return parms_passed


def _check_mosaic_pb_parms(pb_mosaic_parms):
    parms_passed = True
    if not (_check_parms(pb_mosaic_parms, 'pb_name', [str], default='PB')):
        parms_passed = False
    if not (_check_parms(pb_mosaic_parms, 'weight_name', [str], default='WEIGHT_PB')):
        parms_passed = False
    return parms_passed


def _check_norm_parms(norm_parms):
    """_check_imaging_parms.py:110-120."""
    parms_passed = True
    if not (_check_parms(norm_parms, 'norm_type', [str], default='flat_sky',
                         acceptable_data=['flat_noise', 'flat_sky', 'none'])):
        parms_passed = False
    if not (_check_parms(norm_parms, 'single_precision', [bool], default=True)):
        parms_passed = False
    if not (_check_parms(norm_parms, 'pb_limit', [numbers.Number], default=0.2)):
        parms_passed = False
    return parms_passed


def _check_imaging_weights_parms(imaging_weights_parms):
    """_check_imaging_parms.py:123-146."""
    parms_passed = True
    if not (_check_parms(imaging_weights_parms, 'weighting', [str],
                         acceptable_data=['natural', 'uniform', 'briggs', 'briggs_abs'],
                         default='natural')):
        parms_passed = False
    if imaging_weights_parms.get('weighting') == 'briggs_abs':
        if not (_check_parms(imaging_weights_parms, 'briggs_abs_noise', [numbers.Number],
                             default=1.0)):
            parms_passed = False
    if imaging_weights_parms.get('weighting') in ('briggs', 'briggs_abs'):
        if not (_check_parms(imaging_weights_parms, 'robust', [numbers.Number], default=0.5,
                             acceptable_range=[-2, 2])):
            parms_passed = False
    return parms_passed
