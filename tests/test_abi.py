"""The C ABI: the library builds for gfx950, loads, and exports every entry
point include/boitata_gpu.h declares, with the struct layouts the ctypes
mirror assumes. No GPU call."""
import ctypes as C
import re
from pathlib import Path

from boitata_amd import _abi, _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "boitata_gpu.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(bk_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _lib.load()
    names = declared()
    assert "bk_krige" in names and "bk_loo_loss_grad" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.exported_symbols())


def test_version_needs_no_device():
    assert b"gfx950" in _lib.load().bk_version()


def test_struct_layouts():
    # Offsets fixed by the header (all fields naturally aligned).
    assert C.sizeof(_abi.Structure) == 48
    assert C.sizeof(_abi.Variogram) == 8 + 8 + 48 * _abi.MAX_STRUCTURES + 40
    assert _abi.Search.radius.offset == 16 and _abi.Search.azimuth.offset == 32
    assert C.sizeof(_abi.Search) == 16 + 8 + 8 + 40 + 7 * 8 + 16
    assert _abi.Search.calibration_target.offset == 136
    assert C.sizeof(_abi.Block) == 24 + 16 + 24


def test_struct_layouts_match_the_compiler(tmp_path):
    # Compile the header with the host compiler and compare sizeof/offsetof.
    import subprocess

    src = tmp_path / "layout.c"
    src.write_text(f'#include "{HEADER}"\n#include <stdio.h>\n#include <stddef.h>\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(bk_structure),"
                   " sizeof(bk_variogram), sizeof(bk_search), sizeof(bk_block), offsetof(bk_search, max_per_hole),"
                   " offsetof(bk_variogram, azimuth)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(_abi.Structure), C.sizeof(_abi.Variogram), C.sizeof(_abi.Search), C.sizeof(_abi.Block),
            _abi.Search.max_per_hole.offset, _abi.Variogram.azimuth.offset]
    assert got == want


def test_library_is_built_for_gfx950():
    data = _lib.LIB_PATH.read_bytes()
    assert b"gfx950" in data
