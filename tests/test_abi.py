"""The C-ABI library builds for gfx950, loads, and exports exactly the entry
points include/pmnb.h declares (no GPU needed: nothing is computed)."""
import ctypes
import os
import re

from protomoladdon_amd import _build, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "pmnb.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pmnb_[a-z_]+)\s*\(", src)))


def test_library_is_built_and_loads():
    assert os.path.exists(_build.LIB_PATH), "run __graft_entry__.build()"
    lib = _lib.load()
    assert lib.pmnb_version() == 100


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_build.LIB_PATH)
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.EXPORTED) == names


def test_library_targets_gfx950():
    data = open(_build.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_null_arguments_are_rejected_without_a_device():
    lib = _lib.load()
    assert lib.pmnb_create(None, None) == _lib.PMNB_ERR_ARG
    assert lib.pmnb_last_error(None) == b"null context"
    lib.pmnb_destroy(None)


def test_protomol_adapter_compiles_against_the_reference_headers():
    """The SystemForce adapter builds with ProtoMol's flags (skipped when the
    reference tree is not present)."""
    import shutil
    import subprocess
    import pytest
    src = "/root/reference/protomol/src"
    if not os.path.isdir(src) or not shutil.which("g++"):
        pytest.skip("ProtoMol sources not present")
    code = ('#include "protomoladdon_amd/adapter/PmnbCutoffSystemForce.h"\n'
            "template class ProtoMol::PmnbCutoffSystemForce<"
            "ProtoMol::C2SwitchingFunction>;\n"
            "template class ProtoMol::PmnbCutoffSystemForce<"
            "ProtoMol::C1SwitchingFunction>;\n"
            "template class ProtoMol::PmnbCutoffSystemForce<"
            "ProtoMol::C2SwitchingFunction, ProtoMol::PmnbEwaldRealForce>;\n"
            "template class ProtoMol::PmnbCutoffSystemForce<"
            "ProtoMol::C1SwitchingFunction, ProtoMol::PmnbEwaldRealForce>;\n")
    r = subprocess.run(["g++", "-std=c++0x", "-fpermissive", "-w",
                        "-fsyntax-only", "-I" + src,
                        "-I" + os.path.join(ROOT, "include"), "-I" + ROOT,
                        "-x", "c++", "-"], input=code, text=True,
                       capture_output=True)
    assert r.returncode == 0, r.stderr[-2000:]
