"""The drop-in surface: every public name that the reference's solver, CLI and tests import from the
sdpvibe modules exists in sdpvibe_amd under the same module name, with the reference's parameter
names in the reference's order (tests/golden/dropin_api.json, written from sdpvibe/*.py). Names this
build does not provide are listed there with the reason. CPU only: nothing is launched."""
import importlib
import inspect
import json
import os

import pytest

SPEC = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "dropin_api.json")))
CASES = [(mod, name, params) for mod, names in SPEC["modules"].items() for name, params in names.items()]


@pytest.mark.parametrize("mod,name,params", CASES, ids=[f"{m}.{n}" for m, n, _ in CASES])
def test_name_and_parameters(mod, name, params):
    m = importlib.import_module(f"sdpvibe_amd.{mod}")
    assert hasattr(m, name), f"sdpvibe_amd.{mod} lacks {name}"
    if params is None:
        return
    obj = getattr(m, name)
    sig = inspect.signature(obj.__init__ if inspect.isclass(obj) else obj)
    got = [p for p in sig.parameters if p != "self"]
    assert got[:len(params)] == params, (got, params)


def test_not_provided_are_documented():
    for key in SPEC["not_provided"]:
        mod, name = key.split(".")
        m = importlib.import_module(f"sdpvibe_amd.{mod}")
        assert not hasattr(m, name), key


def test_package_lists_its_modules():
    import sdpvibe_amd
    for mod in SPEC["modules"]:
        assert mod in sdpvibe_amd.__all__, mod
