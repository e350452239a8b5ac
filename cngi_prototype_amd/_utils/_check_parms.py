"""Checks of the parameter dictionaries the imaging functions take.

`_check_parms` validates one entry of a parameter dictionary and writes its
default in when it is missing; `_check_sel_parms` turns the data groups that
`sel_parms` names into the variable names a function reads
(`data_group_in`) and writes (`data_group_out`).

Callers of the drop-in API see the behaviour of cngi/_utils/_check_parms.py
(:18-260): the values written back into their dictionaries, the return
values and assertion texts, and the "Setting default ..." lines.  That
behaviour, including two quirks (a failing entry of a dict parameter is
reported without failing the check, and an input group naming a variable the
dataset lacks is reported, not refused), is tabulated case by case in
tests/test_check_parms.py; this module is written against that table.
Integer parameters accept Python and numpy integers (`INT_TYPES`; the
reference's `np.int` is Python's `int`, which numpy 2 no longer aliases).
"""
import copy

import numpy as np

INT_TYPES = (int, np.integer)

_ERROR = '######### ERROR:'


def _is_one_of(value, types):
    """isinstance against a list of types, in which `np.array` stands for
    `np.ndarray` and an entry may itself be a tuple of types."""
    return isinstance(value, tuple(np.ndarray if t is np.array else t for t in types))


def _refuse(*words):
    print(_ERROR, *words)
    return False


def _admissible(name, value, acceptable_data, acceptable_range):
    if acceptable_data is not None and value not in acceptable_data:
        return _refuse('invalid', name, '=', value, '; allowed values:', acceptable_data)
    if acceptable_range is not None:
        low, high = acceptable_range
        if value < low or value > high:
            return _refuse('invalid', name, '=', value, '; allowed range:', acceptable_range)
    return True


def _check_parms(parm_dict, string_key, acceptable_data_types, acceptable_data=None,
                 acceptable_range=None, list_acceptable_data_types=None, list_len=None,
                 default=None):
    """Check parm_dict[string_key]; returns True when it is acceptable.

    - Missing: `default` is written in (and printed), else the check fails.
    - `list` or `np.array` among the types: the value must have `list_len`
      elements (any number for -1), each of one of
      `list_acceptable_data_types`, in `acceptable_data` and inside
      `acceptable_range` when those are given.
    - `dict` among the types: `default` must be a dict; its keys missing
      from the value are written in one by one (and printed), the present
      ones are checked against the type of their default.  A present entry
      that fails is printed but does not fail the check.
    - Otherwise the value must be of one of the types, in `acceptable_data`
      and inside `acceptable_range` when those are given.
    """
    if string_key not in parm_dict:
        if default is None:
            return _refuse('parameter', string_key, 'must be specified')
        parm_dict[string_key] = default
        print('Setting default', string_key, ' to ', default)
        return True

    value = parm_dict[string_key]
    kinds = list(acceptable_data_types)
    if list in kinds or np.array in kinds:
        if list_len != -1 and len(value) != list_len:
            return _refuse('parameter', string_key, 'must hold', list_len, 'elements of',
                           list_acceptable_data_types)
        for item in value:
            if not _is_one_of(item, list_acceptable_data_types):
                return _refuse('parameter', string_key, 'must hold elements of',
                               list_acceptable_data_types)
            if not _admissible(string_key, item, acceptable_data, acceptable_range):
                return False
        return True

    if dict in kinds:
        if default is None:
            return _refuse('dict parameter', string_key, 'has no default dict')
        for entry, entry_default in default.items():
            if entry in value:
                _check_parms(value, entry, [type(entry_default)], default=entry_default)
            else:
                value[entry] = entry_default
                print('Setting default', string_key, "['", entry, "']", ' to ', entry_default)
        return True

    if not _is_one_of(value, kinds):
        return _refuse('parameter', string_key, 'must be of type', kinds)
    return _admissible(string_key, value, acceptable_data, acceptable_range)


def _check_dataset(vis_dataset, data_variable_name):
    """True when the dataset has the variable; otherwise prints and returns False."""
    try:
        vis_dataset[data_variable_name]
    except Exception:
        print('######### ERROR Data array ', data_variable_name, 'can not be found in dataset.')
        return False
    return True


def _report_missing_variables(xds, group):
    """Print every variable a data group names (its 'id' and 'properties'
    entries aside, nested groups included) that the dataset lacks."""
    for key, name in group.items():
        if key in ('id', 'properties'):
            continue
        if isinstance(name, dict):
            _report_missing_variables(xds, name)
        else:
            _check_dataset(xds, name)


def _input_group(groups, ids, sel_parms, append_to_in_id):
    """(input group defaults, id the outputs go to) of a resolved sel_parms."""
    given = sel_parms.get('data_group_in')
    if given is None:
        group = next(iter(groups.values()))
    elif 'id' in given:
        # (the reference builds this message from sel_parms['xds'] and raises
        # KeyError instead when an image dataset's sel_parms has no 'xds')
        assert int(given['id']) in ids, \
            "######### ERROR: data_group_in id does not exist in " + str(
                sel_parms.get('xds', 'the dataset'))
        group = copy.deepcopy(groups[str(given['id'])])
        sel_parms['data_group_in'] = group
        print("Setting data_group_in  to ", group)
    else:
        raise TypeError("data_group_in must name a stored data group by its 'id'")
    out_id = group['id'] if append_to_in_id else max(ids) + 1
    return group, out_id


def _check_sel_parms(xds, sel_parms, new_or_modified_data_variables={},
                     required_data_variables={}, drop_description_out_keys=[],
                     append_to_in_id=False):
    """Resolve sel_parms['data_group_in'] and ['data_group_out'] in place.

    - A dataset without data groups gets one, `{'id': id}` with id the given
      data_group_out_id or '0', and the outputs then join it.
    - data_group_in: the stored group data_group_in_id names (a copy), or the
      first group; `required_data_variables` fill what it lacks.
    - data_group_out: the input group's variables plus
      `new_or_modified_data_variables` (a sel_parms key of the same name
      renames that output), under data_group_out_id, the input group's id
      (append_to_in_id) or a new id max(ids) + 1.  Entries of a given
      data_group_out that the function does not write are dropped.
    - Refused: an unknown input id; an output variable name that another
      data group already uses for the same variable.
    """
    outputs = new_or_modified_data_variables
    if 'data_groups' not in xds.attrs:
        first = str(sel_parms['data_group_out_id']) if 'data_group_out_id' in sel_parms else '0'
        xds.attrs['data_groups'] = [{first: {'id': first}}]
        append_to_in_id = True
    groups = xds.attrs['data_groups'][0]
    ids = [int(k) for k in groups]

    if 'data_group_in_id' in sel_parms:
        sel_parms['data_group_in'] = {'id': str(sel_parms['data_group_in_id'])}
    if 'data_group_out_id' in sel_parms:
        sel_parms['data_group_out'] = {'id': str(sel_parms['data_group_out_id'])}
    renamed = {name: sel_parms[name] for name in outputs if name in sel_parms}
    if renamed:
        sel_parms['data_group_out'] = {**sel_parms.get('data_group_out', {}), **renamed}

    in_group, out_id = _input_group(groups, ids, sel_parms, append_to_in_id)
    if 'data_group_out_id' in sel_parms:
        out_id = sel_parms['data_group_out_id']
    given_out = sel_parms.get('data_group_out')
    if given_out is not None:
        for key in [k for k in given_out if k != 'id' and k not in outputs]:
            del given_out[key]

    in_defaults = {**required_data_variables, **in_group}
    out_defaults = {**in_defaults, 'id': str(out_id), **outputs}
    passed = [_check_parms(sel_parms, key, [dict], default=d)
              for key, d in (('data_group_in', in_defaults), ('data_group_out', out_defaults))]
    assert all(passed), "######### ERROR: sel_parms checking failed"
    _report_missing_variables(xds, sel_parms['data_group_in'])

    out = sel_parms['data_group_out']
    for name in outputs:
        for gid, group in groups.items():
            if name in group:
                assert out['id'] == gid or group[name] != out[name], \
                    "Data variables, that are modified by the function, can not be replaced " \
                    "if they are referenced in another data_group"
    return True
