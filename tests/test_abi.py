"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/pe.h declares, and the ctypes mirror matches the header: the
size and every field offset of every struct, as a C compiler lays them out
from pe.h."""
import ctypes as C
import os
import re
import subprocess
import tempfile

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


def _mirrors():
    return {name: cls for name, cls in vars(abi).items()
            if name.startswith("pe_") and isinstance(cls, type) and issubclass(cls, C.Structure)}


def test_every_header_struct_has_a_ctypes_mirror():
    text = open(HEADER).read()
    declared = set(re.findall(r"typedef struct (pe_[a-z0-9_]+) \{", text))
    assert declared and declared == set(_mirrors()), declared ^ set(_mirrors())


def test_struct_layout_matches_the_c_compiler():
    """Compiles a probe against pe.h with the host C compiler and compares
    sizeof / offsetof of every field with the ctypes mirror."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "pe.h"', "int main(void) {"]
    expected = []
    for name, cls in sorted(_mirrors().items()):
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        expected.append(f"{name} {C.sizeof(cls)}")
        for field in cls._fields_:
            f = field[0]
            lines.append(f'printf("{name}.{f} %zu\\n", offsetof({name}, {f}));')
            expected.append(f"{name}.{f} {getattr(cls, f).offset}")
    lines += ["return 0;", "}"]
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "layout.c")
        exe = os.path.join(tmp, "layout")
        with open(src, "w") as fh:
            fh.write("\n".join(lines))
        subprocess.check_call(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), src, "-o", exe])
        got = subprocess.check_output([exe], text=True).split("\n")[:-1]
    assert got == expected, [(g, e) for g, e in zip(got, expected) if g != e]
