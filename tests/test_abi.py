"""The C ABI: the built library loads (no GPU needed) and exports every
function include/*.h declares; the error path works without a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")

DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9]*\s*\**\s*(hfx_[a-z_0-9]+)\s*\(", re.M)


def declared():
    names = {}
    for f in sorted(os.listdir(INCLUDE)):
        if f.endswith(".h"):
            with open(os.path.join(INCLUDE, f)) as fh:
                text = fh.read()
            for m in DECL.finditer(text):
                names[m.group(1)] = f
    return names


def test_headers_declare_the_boundary():
    names = declared()
    # the reference-facing core (ERI layout, Fock build), the SCF driver and the server
    for must in ("hfx_basis_create", "hfx_eri_unique", "hfx_fock_build", "hfx_scf_run", "hfx_server_submit"):
        assert must in names
    assert len(names) >= 30


@pytest.mark.parametrize("name", sorted(declared()))
def test_library_exports(name):
    from hartree_fock_cpp_amd import _native as N
    assert hasattr(N.lib, name), f"{name} declared in include/{declared()[name]} but not exported"


def test_no_device_is_an_error_not_a_crash():
    import torch
    from hartree_fock_cpp_amd import _native as N
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    assert N.lib.hfx_device_count() == 0
    h = C.c_void_p()
    from hartree_fock_cpp_amd.serve import ServerOpts
    rc = N.lib.hfx_server_create(C.byref(ServerOpts(2, 0, 1)), C.byref(h))
    assert rc != 0 and N.lib.hfx_last_error()
