"""Behaviour table of the parameter checks every imaging entry point runs
(the observable contract of cngi/_utils/_check_parms.py:18-260: return
values, values written back into the dictionaries, the "Setting default"
messages, assertion texts).  CPU only."""
import numpy as np
import pytest

from cngi_prototype_amd._utils._check_parms import (INT_TYPES, _check_dataset, _check_parms,
                                                    _check_sel_parms)
from cngi_prototype_amd.xds import Dataset


# ---------------------------------------------------------------------------
# _check_parms: one parameter of a dictionary
# ---------------------------------------------------------------------------
SCALAR_CASES = [
    # (given dict, key, kwargs, returns, dict afterwards)
    ({}, "pb_limit", dict(acceptable_data_types=[float], default=0.2), True, {"pb_limit": 0.2}),
    ({}, "norm_type", dict(acceptable_data_types=[str]), False, {}),
    ({"n": 3}, "n", dict(acceptable_data_types=[INT_TYPES]), True, {"n": 3}),
    ({"n": np.int64(3)}, "n", dict(acceptable_data_types=[INT_TYPES]), True, {"n": 3}),
    ({"n": 3.0}, "n", dict(acceptable_data_types=[INT_TYPES]), False, {"n": 3.0}),
    ({"m": "cube"}, "m", dict(acceptable_data_types=[str], acceptable_data=["cube", "continuum"]),
     True, {"m": "cube"}),
    ({"m": "mfs"}, "m", dict(acceptable_data_types=[str], acceptable_data=["cube", "continuum"]),
     False, {"m": "mfs"}),
    ({"r": 2.5}, "r", dict(acceptable_data_types=[float], acceptable_range=[-2, 2]), False,
     {"r": 2.5}),
    ({"r": -2}, "r", dict(acceptable_data_types=[int], acceptable_range=[-2, 2]), True,
     {"r": -2}),
    # a given value wins over the default, and is still checked
    ({"s": True}, "s", dict(acceptable_data_types=[bool], default=False), True, {"s": True}),
    ({"s": 1.0}, "s", dict(acceptable_data_types=[bool], default=False), False, {"s": 1.0}),
]


@pytest.mark.parametrize("given,key,kw,ret,after", SCALAR_CASES)
def test_check_parms_scalars(given, key, kw, ret, after):
    d = dict(given)
    types = kw.pop("acceptable_data_types")
    assert _check_parms(d, key, types, **kw) is ret
    assert d == after


LIST_CASES = [
    ({"size": [10, 20]}, dict(list_acceptable_data_types=[INT_TYPES], list_len=2), True),
    ({"size": [10]}, dict(list_acceptable_data_types=[INT_TYPES], list_len=2), False),
    ({"size": [10, 2.0]}, dict(list_acceptable_data_types=[INT_TYPES], list_len=2), False),
    ({"size": [1, 2, 3, 4, 5]}, dict(list_acceptable_data_types=[INT_TYPES], list_len=-1), True),
    ({"size": np.array([4, 5])}, dict(list_acceptable_data_types=[np.integer], list_len=2),
     True),
    ({"size": [1, 9]}, dict(list_acceptable_data_types=[int], list_len=2,
                            acceptable_range=[0, 8]), False),
    ({"size": [1, 8]}, dict(list_acceptable_data_types=[int], list_len=2,
                            acceptable_range=[0, 8]), True),
    ({"size": ["a", "c"]}, dict(list_acceptable_data_types=[str], list_len=2,
                                acceptable_data=["a", "b"]), False),
]


@pytest.mark.parametrize("given,kw,ret", LIST_CASES)
def test_check_parms_lists(given, kw, ret):
    d = dict(given)
    assert _check_parms(d, "size", [list, np.array], **kw) is ret
    assert d.keys() == given.keys()


def test_check_parms_list_default_is_filled():
    d = {}
    assert _check_parms(d, "oversampling", [list], list_acceptable_data_types=[int], list_len=2,
                        default=[10, 10])
    assert d == {"oversampling": [10, 10]}


def test_check_parms_dict_fills_missing_keys_and_prints(capsys):
    """A dict parameter needs a default dict; missing keys are filled one by
    one, present keys are checked against their default's type.  A nested
    failure is printed but, as in the reference, does not make the check
    return False."""
    d = {"group": {"data": "DATA"}}
    assert _check_parms(d, "group", [dict], default={"id": "0", "data": "X", "uvw": "UVW"})
    assert d["group"] == {"data": "DATA", "id": "0", "uvw": "UVW"}
    out = capsys.readouterr().out
    assert "Setting default" in out and "uvw" in out
    d = {"group": {"data": 5}}
    assert _check_parms(d, "group", [dict], default={"data": "DATA"}) is True
    assert "ERROR" in capsys.readouterr().out
    assert _check_parms({"group": {}}, "group", [dict]) is False
    d = {}
    assert _check_parms(d, "group", [dict], default={"id": "1"})
    assert d == {"group": {"id": "1"}}


def test_check_dataset():
    ds = Dataset(data_vars={"DATA": (["t"], np.zeros(2))})
    assert _check_dataset(ds, "DATA") is True
    assert _check_dataset(ds, "MODEL") is False


# ---------------------------------------------------------------------------
# _check_sel_parms: data-group resolution
# ---------------------------------------------------------------------------
VIS_GROUP = {"id": "0", "data": "DATA", "uvw": "UVW", "weight": "DATA_WEIGHT"}


def _vis(groups=None, extra=()):
    names = ["DATA", "UVW", "DATA_WEIGHT"] + list(extra)
    ds = Dataset(data_vars={n: (["t"], np.zeros(2)) for n in names})
    ds.attrs["data_groups"] = [dict(groups) if groups is not None else {"0": dict(VIS_GROUP)}]
    return ds


def test_first_group_in_new_group_out():
    """No ids: data_group_in is the first group; without append_to_in_id the
    output group is a new id, max(ids) + 1, holding the input group's
    variables (the function adds none)."""
    sp = {"xds": "xds0"}
    assert _check_sel_parms(_vis(), sp) is True
    assert sp["data_group_in"] == VIS_GROUP
    assert sp["data_group_out"] == {**VIS_GROUP, "id": "1"}
    assert list(sp["data_group_out"])[0] == "id"


def test_append_to_the_input_group():
    """append_to_in_id: the output group keeps the input group's id and adds
    the modified variables under their default names."""
    sp = {"xds": "xds0"}
    _check_sel_parms(_vis(), sp, new_or_modified_data_variables={"imaging_weight":
                                                                 "IMAGING_WEIGHT"},
                     append_to_in_id=True)
    assert sp["data_group_out"] == {**VIS_GROUP, "imaging_weight": "IMAGING_WEIGHT"}


def test_ids_and_renamed_outputs(capsys):
    """data_group_in_id selects a stored group (a copy, printed); a sel_parms
    key naming a modified variable renames that output; the output group's
    given keys come first, then the defaults it lacks, in default order."""
    groups = {"0": dict(VIS_GROUP), "2": {"id": "2", "data": "CORRECTED", "uvw": "UVW",
                                          "weight": "DATA_WEIGHT"}}
    xds = _vis(groups, extra=["CORRECTED"])
    sp = {"xds": "xds0", "data_group_in_id": 2, "imaging_weight": "MY_IW"}
    _check_sel_parms(xds, sp, new_or_modified_data_variables={"imaging_weight":
                                                              "IMAGING_WEIGHT"})
    assert sp["data_group_in"] == groups["2"]
    assert sp["data_group_in"] is not xds.attrs["data_groups"][0]["2"]
    assert sp["data_group_out"] == {"imaging_weight": "MY_IW", "id": "3", "data": "CORRECTED",
                                    "uvw": "UVW", "weight": "DATA_WEIGHT"}
    assert list(sp["data_group_out"]) == ["imaging_weight", "id", "data", "uvw", "weight"]
    assert "Setting data_group_in" in capsys.readouterr().out
    # append_to_in_id with an explicit input id keeps that id
    sp = {"xds": "xds0", "data_group_in_id": 2}
    _check_sel_parms(xds, sp, new_or_modified_data_variables={"image": "IMAGE"},
                     append_to_in_id=True)
    assert sp["data_group_out"]["id"] == "2"


def test_output_id_and_unrelated_keys():
    """data_group_out_id fixes the output id; keys of a given data_group_out
    that the function does not produce are dropped (the id stays)."""
    sp = {"xds": "xds0", "data_group_out_id": 7,
          "data_group_out": {"id": "5", "image": "IMG", "junk": "X"}}
    _check_sel_parms(_vis(), sp, new_or_modified_data_variables={"image": "IMAGE"})
    assert sp["data_group_out"]["id"] == "7"
    assert "junk" not in sp["data_group_out"]
    assert sp["data_group_out"]["image"] == "IMAGE"
    sp = {"xds": "xds0", "data_group_out": {"id": "5", "image": "IMG", "junk": "X"}}
    _check_sel_parms(_vis(), sp, new_or_modified_data_variables={"image": "IMAGE"})
    assert sp["data_group_out"] == {"id": "5", "image": "IMG", **{
        k: v for k, v in VIS_GROUP.items() if k != "id"}}


def test_dataset_without_groups_gets_one():
    """An image dataset with no data_groups gets [{'0': {'id': '0'}}] (or the
    data_group_out_id given), and the outputs are appended to that group
    whatever append_to_in_id says."""
    img = Dataset()
    sp = {}
    _check_sel_parms(img, sp, new_or_modified_data_variables={"image": "IMAGE",
                                                              "sum_weight": "SUM_WEIGHT"})
    assert img.attrs["data_groups"] == [{"0": {"id": "0"}}]
    assert sp["data_group_in"] == {"id": "0"}
    assert sp["data_group_out"] == {"id": "0", "image": "IMAGE", "sum_weight": "SUM_WEIGHT"}
    img = Dataset()
    sp = {"data_group_out_id": 4}
    _check_sel_parms(img, sp, new_or_modified_data_variables={"psf": "PSF"})
    assert img.attrs["data_groups"] == [{"4": {"id": "4"}}]
    assert sp["data_group_out"] == {"id": "4", "psf": "PSF"}


def test_required_variables_join_the_input_defaults(capsys):
    """required_data_variables enter the input group's defaults.  Variables
    the input group names but the dataset lacks are reported, not refused
    (the reference's existence check discards its result for the nested
    group)."""
    sp = {"xds": "xds0"}
    _check_sel_parms(_vis(extra=["FLAG"]), sp, required_data_variables={"flag": "FLAG"})
    assert sp["data_group_in"]["flag"] == "FLAG"
    assert sp["data_group_out"]["flag"] == "FLAG"
    capsys.readouterr()
    assert _check_sel_parms(_vis(), {"xds": "xds0"}, required_data_variables={"flag": "FLAG"})
    out = capsys.readouterr().out
    assert "FLAG" in out and "can not be found" in out
    bad = _vis({"0": {"id": "0", "data": "NOPE"}})
    assert _check_sel_parms(bad, {"xds": "xds0"})
    assert "NOPE" in capsys.readouterr().out


def test_refusals():
    """An unknown input id, and a modified variable name that another group
    already uses, are refused."""
    with pytest.raises(AssertionError, match="data_group_in id does not exist in xds0"):
        _check_sel_parms(_vis(), {"xds": "xds0", "data_group_in_id": 3})
    # an image dataset's sel_parms has no 'xds': still an AssertionError
    with pytest.raises(AssertionError, match="data_group_in id does not exist"):
        _check_sel_parms(_vis(), {"data_group_in_id": 3})
    groups = {"0": {**VIS_GROUP, "imaging_weight": "IMAGING_WEIGHT"},
              "1": dict(VIS_GROUP, id="1")}
    with pytest.raises(AssertionError, match="referenced in another data_group"):
        _check_sel_parms(_vis(groups, extra=["IMAGING_WEIGHT"]),
                         {"xds": "xds0", "data_group_in_id": 1},
                         new_or_modified_data_variables={"imaging_weight": "IMAGING_WEIGHT"})
    # the same name in the group being written is fine
    sp = {"xds": "xds0", "data_group_in_id": 0}
    _check_sel_parms(_vis(groups, extra=["IMAGING_WEIGHT"]), sp,
                     new_or_modified_data_variables={"imaging_weight": "IMAGING_WEIGHT"},
                     append_to_in_id=True)
    assert sp["data_group_out"]["id"] == "0"


def test_input_group_without_id_is_refused():
    """A data_group_in dictionary without an id is not usable (the
    reference fails on it with a TypeError)."""
    with pytest.raises(TypeError):
        _check_sel_parms(_vis(), {"xds": "xds0", "data_group_in": {"data": "DATA"}})
