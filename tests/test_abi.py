"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/pe.h declares, and the ctypes mirror matches the header."""
import ctypes as C
import os
import re

from pricing_amd import abi, engine

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pe.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pe_[a-z0-9_]+)\s*\(", text)))


def test_header_declarations_are_bound():
    names = declared_functions()
    assert "pe_plan_compile" in names and "pe_plan_evaluate" in names
    assert sorted(abi.PE_SYMBOLS) == names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(engine.library_path())
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.pe_abi_version() == abi.PE_ABI_VERSION


def test_struct_sizes_match_c_layout():
    # pe_result: 18 estimates (4 doubles + u64) + 3x18 doubles + 2 doubles + 3 u64 + 2 u32 + double + u64
    assert C.sizeof(abi.pe_estimate) == 40
    assert C.sizeof(abi.pe_result) == 18 * 40 + 3 * 18 * 8 + 2 * 8 + 3 * 8 + 2 * 4 + 8 + 8
