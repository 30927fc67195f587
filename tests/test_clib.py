"""The C-ABI library loads on any host and exports every symbol include/*.h declares."""
import ctypes
import glob
import os
import re

from carcassonne_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names.update(re.findall(r"^\s*(?:int|const char\*)\s+(cc_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_all_declared_symbols():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 10
    for name in sorted(names):
        assert hasattr(lib, name), name


def test_binding_table_matches_header():
    assert set(_lib.SIGNATURES) | {"cc_last_error"} == declared_symbols()


def test_version_and_error_string():
    lib = _lib.load()
    assert lib.cc_version() >= 1
    assert isinstance(lib.cc_last_error(), bytes)


def test_status_codes_map_to_reference_exceptions():
    import pytest
    from carcassonne_amd.errors import RelaxFailed
    with pytest.raises(ValueError):
        _lib.check(_lib.CC_ERR_ARG, "x")
    with pytest.raises(RelaxFailed):
        _lib.check(_lib.CC_ERR_NOT_CONVERGED, "x")
    with pytest.raises(NotImplementedError):
        _lib.check(_lib.CC_ERR_UNSUPPORTED, "x")
