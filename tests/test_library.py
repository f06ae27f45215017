"""The C-ABI library builds, loads and exports every entry point that
include/mddem_hip.h declares. No compute calls: this runs without a GPU."""
import os
import re

import pytest

from mddem_amd import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(mdd_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_systems():
    names = declared_functions()
    for must in ("mdd_create", "mdd_bin_neighbor_list", "mdd_lj_force",
                 "mdd_hertz_mindlin_contact_force", "mdd_nh_pre_initial", "mdd_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = hip.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert missing == []


def test_bindings_cover_every_declared_symbol():
    assert sorted(hip.SIGNATURES) == declared_functions()


def test_version():
    assert hip.lib().mdd_version() >= 1


def test_create_fails_cleanly_without_a_gfx950(monkeypatch):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(hip.MddError) as e:
        hip.System(hip.MODEL_LJ)
    assert e.value.code == -6
