"""Parameter and data-group checking used by every imaging entry point.

Restates cngi/_utils/_check_parms.py:18-260: defaults are filled in place and
printed ("Setting default ..."), invalid parameters print an error and make
the check return False, and ``sel_parms`` data-group ids are resolved to
variable names.  The reference's ``np.int`` (an alias of Python ``int`` that
numpy 2 removed) is written here as ``INT_TYPES``.
"""
import copy

import numpy as np

INT_TYPES = (int, np.integer)


def _is_seq_type(acceptable_data_types):
    return (list in acceptable_data_types) or (np.array in acceptable_data_types) or (
        np.ndarray in acceptable_data_types)


def _isinstance(x, types):
    for t in types:
        if t is np.array:
            t = np.ndarray
        if isinstance(t, tuple):
            if isinstance(x, t):
                return True
        elif isinstance(x, t):
            return True
    return False


def _check_parms(parm_dict, string_key, acceptable_data_types, acceptable_data=None,
                 acceptable_range=None, list_acceptable_data_types=None, list_len=None,
                 default=None):
    """cngi/_utils/_check_parms.py:18-110."""
    if string_key in parm_dict:
        if _is_seq_type(acceptable_data_types):
            if (len(parm_dict[string_key]) != list_len) and (list_len != -1):
                print('######### ERROR:Parameter ', string_key, 'must be a list of ',
                      list_acceptable_data_types, ' and length', list_len, '. Wrong length.')
                return False
            list_len = len(parm_dict[string_key])
            for i in range(list_len):
                if not _isinstance(parm_dict[string_key][i], list_acceptable_data_types):
                    print('######### ERROR:Parameter ', string_key, 'must be a list of ',
                          list_acceptable_data_types, ' and length', list_len, '. Wrong type.')
                    return False
                if acceptable_data is not None:
                    if not (parm_dict[string_key][i] in acceptable_data):
                        print('######### ERROR: Invalid', string_key, '. Can only be one of ',
                              acceptable_data, '.')
                        return False
                if acceptable_range is not None:
                    if (parm_dict[string_key][i] < acceptable_range[0]) or (
                            parm_dict[string_key][i] > acceptable_range[1]):
                        print('######### ERROR: Invalid', string_key,
                              '. Must be within the range ', acceptable_range, '.')
                        return False
        elif dict in acceptable_data_types:
            parms_passed = True
            if default is None:
                print('######### ERROR:Dictionary parameters must have a default. Please report bug.')
                return False
            for default_element in default:
                if default_element in parm_dict[string_key]:
                    if not (_check_parms(parm_dict[string_key], default_element,
                                         [type(default[default_element])],
                                         default=default[default_element])):
                        parms_passed = False
                else:
                    parm_dict[string_key][default_element] = default[default_element]
                    print('Setting default', string_key, '[\'', default_element, '\']', ' to ',
                          default[default_element])
            return parms_passed
        else:
            if not _isinstance(parm_dict[string_key], acceptable_data_types):
                print('######### ERROR:Parameter ', string_key, 'must be of type ',
                      acceptable_data_types)
                return False
            if acceptable_data is not None:
                if not (parm_dict[string_key] in acceptable_data):
                    print('######### ERROR: Invalid', string_key, '. Can only be one of ',
                          acceptable_data, '.')
                    return False
            if acceptable_range is not None:
                if (parm_dict[string_key] < acceptable_range[0]) or (
                        parm_dict[string_key] > acceptable_range[1]):
                    print('######### ERROR: Invalid', string_key, '. Must be within the range ',
                          acceptable_range, '.')
                    return False
    else:
        if default is not None:
            parm_dict[string_key] = default
            print('Setting default', string_key, ' to ', parm_dict[string_key])
        else:
            print('######### ERROR:Parameter ', string_key, 'must be specified')
            return False
    return True


def _check_dataset(vis_dataset, data_variable_name):
    try:
        vis_dataset[data_variable_name]
    except Exception:
        print('######### ERROR Data array ', data_variable_name, 'can not be found in dataset.')
        return False
    return True


def _check_existence_sel_parms(dataset, sel_parms):
    """cngi/_utils/_check_parms.py:251-260."""
    parms_passed = True
    for sel in sel_parms:
        if isinstance(sel_parms[sel], dict):
            if sel != 'properties':
                _check_existence_sel_parms(dataset, sel_parms[sel])
        else:
            if (sel != 'id') and (sel != 'properties'):
                if not (_check_dataset(dataset, sel_parms[sel])):
                    parms_passed = False
    return parms_passed


def _check_sub_sel_parms(sel_parms, select_defaults):
    parms_passed = True
    for sel in select_defaults:
        if not (_check_parms(sel_parms, sel, [type(select_defaults[sel])],
                             default=select_defaults[sel])):
            parms_passed = False
    return parms_passed


def _check_sel_parms(xds, sel_parms, new_or_modified_data_variables={}, required_data_variables={},
                     drop_description_out_keys=[], append_to_in_id=False):
    """Resolve data groups (cngi/_utils/_check_parms.py:122-223)."""
    if not ('data_groups' in xds.attrs):
        if 'data_group_out_id' in sel_parms:
            xds.attrs['data_groups'] = [{str(sel_parms['data_group_out_id']): {
                'id': str(sel_parms['data_group_out_id'])}}]
        else:
            xds.attrs['data_groups'] = [{'0': {'id': '0'}}]
        append_to_in_id = True

    data_group_ids = np.array(list(xds.attrs['data_groups'][0])).astype(int)

    if 'data_group_in_id' in sel_parms:
        sel_parms['data_group_in'] = {'id': str(sel_parms['data_group_in_id'])}
    if 'data_group_out_id' in sel_parms:
        sel_parms['data_group_out'] = {'id': str(sel_parms['data_group_out_id'])}

    for mdv in new_or_modified_data_variables:
        if mdv in sel_parms:
            if 'data_group_out' in sel_parms:
                sel_parms['data_group_out'] = {**sel_parms['data_group_out'], **{mdv: sel_parms[mdv]}}
            else:
                sel_parms['data_group_out'] = {mdv: sel_parms[mdv]}

    new_data_group_id = None
    if 'data_group_in' in sel_parms:
        if 'id' in sel_parms['data_group_in']:
            assert (int(sel_parms['data_group_in']['id']) in data_group_ids), \
                "######### ERROR: data_group_in id does not exist in " + str(sel_parms.get('xds'))
            sel_parms['data_group_in'] = copy.deepcopy(
                xds.attrs['data_groups'][0][str(sel_parms['data_group_in']['id'])])
            print("Setting data_group_in  to ", sel_parms['data_group_in'])
            data_group_in_default = sel_parms['data_group_in']
            if append_to_in_id:
                new_data_group_id = sel_parms['data_group_in']['id']
            else:
                new_data_group_id = np.max(data_group_ids) + 1
        else:
            data_group_in_default = dict(list(xds.attrs['data_groups'][0].values())[0])
            data_group_in_default['id'] = -1  # custom
    else:
        data_group_in_default = list(xds.attrs['data_groups'][0].values())[0]
        if append_to_in_id:
            new_data_group_id = data_group_in_default['id']
        else:
            new_data_group_id = np.max(data_group_ids) + 1

    if 'data_group_out_id' in sel_parms:
        new_data_group_id = sel_parms['data_group_out_id']

    if 'data_group_out' in sel_parms:
        temp_sel_parms = copy.deepcopy(sel_parms)
        for sel in temp_sel_parms['data_group_out']:
            if not (sel in new_or_modified_data_variables) and (sel != 'id'):
                sel_parms['data_group_out'].pop(sel, None)

    data_group_in_default = {**required_data_variables, **data_group_in_default}
    data_group_out_defaults = {**{'id': str(new_data_group_id)}, **new_or_modified_data_variables}
    data_group_out_defaults = {**data_group_in_default, **data_group_out_defaults}
    sel_defaults = {'data_group_in': data_group_in_default,
                    'data_group_out': data_group_out_defaults}
    assert (_check_sub_sel_parms(sel_parms, sel_defaults)), \
        "######### ERROR: sel_parms checking failed"
    sel_check = {'data_group_in': sel_parms['data_group_in']}
    assert (_check_existence_sel_parms(xds, sel_check)), "######### ERROR: sel_parms checking failed"

    for mdv in new_or_modified_data_variables:
        for d_id in xds.attrs['data_groups'][0]:
            if mdv in xds.attrs['data_groups'][0][d_id]:
                assert (sel_parms['data_group_out']['id'] == d_id) or (
                    xds.attrs['data_groups'][0][d_id][mdv] != sel_parms['data_group_out'][mdv]), \
                    "Data variables, that are modified by the function, can not be replaced if " \
                    "they are referenced in another data_group"
    return True
