"""The C-ABI library loads without a GPU and exports every declared symbol."""
import re
from pathlib import Path

from cie_amd import _lib

INCLUDE = Path(__file__).resolve().parent.parent / "include"


def declared_functions():
    names = set()
    for header in INCLUDE.glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names.update(re.findall(r"\b(cie_(?:hip|sys)_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.library()
    declared = declared_functions()
    assert len(declared) > 40
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_binding_covers_declarations():
    assert declared_functions() == set(_lib.exported_names())


def test_status_strings_and_no_device_path():
    lib = _lib.library()
    assert lib.cie_hip_status_string(_lib.ERROR_NOT_SPD)
    count = _lib.ctypes.c_int(-1)
    assert lib.cie_hip_device_count(_lib.byref(count)) == 0
    assert count.value >= 0


def test_binding_constants_match_the_header():
    """The constants the Python mirror restates equal the header's macros."""
    header = (Path(__file__).resolve().parent.parent / "include" / "cie_hip.h").read_text()
    macros = dict(re.findall(r"#define\s+(CIE_HIP_\w+)\s+([-\w.]+)", header))
    assert float(macros["CIE_HIP_DEFAULT_REFINEMENT_FACTOR"]) == _lib.DEFAULT_REFINEMENT_FACTOR
    assert int(macros["CIE_HIP_MIXED_PLAIN_FP32"]) == _lib.MIXED_PLAIN_FP32
