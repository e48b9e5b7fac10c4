"""The C-ABI library loads (without a device) and exports every symbol the header declares;
host-only entry points agree with the oracle."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cpm_mi355x.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cpm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from corepotts import load_library
    lib = load_library()
    names = declared_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    assert lib.cpm_abi_version() == 1


def test_ensemble_seed_matches_oracle():
    # sciml/interface.jl:1009-1026 restated by the oracle; the product computes it on the host
    from corepotts import ensemble_seed
    from oracle import bridge
    lib = bridge.load()
    for master in (0, 1234, 0xDEADBEEFCAFEF00D):
        for traj in (1, 2, 77, 4096):
            for rerun in (1, 3):
                assert ensemble_seed(master, traj, rerun) == lib.orc_ensemble_seed(master, traj, rerun)


def test_ensemble_seed_domain_checks():
    import pytest
    from corepotts import ensemble_seed
    with pytest.raises(ValueError):
        ensemble_seed(1, 0)
    with pytest.raises(ValueError):
        ensemble_seed(1, 1, 257)


def test_create_reports_errors_without_crashing():
    # invalid descriptor: rejected on the host before any device work
    from corepotts.native import CpmModelDesc, load_library
    lib = load_library()
    desc = CpmModelDesc()
    desc.abi_version = 99
    handle = ctypes.c_void_p()
    rc = lib.cpm_create(ctypes.byref(desc), ctypes.byref(handle))
    assert rc != 0
    assert not handle.value
    assert lib.cpm_last_error(None)
