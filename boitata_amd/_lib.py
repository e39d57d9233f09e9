"""Loads libboitata_gpu.so (built in-tree by boitata_amd.build).

There is no CPU path: if the library or a HIP device is missing, the
engine raises instead of computing anything.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from . import _abi

LIB_PATH = Path(__file__).resolve().parent / "libboitata_gpu.so"
_lib = None


class EngineUnavailable(RuntimeError):
    """The HIP library is not built, or cannot be loaded."""


def load() -> C.CDLL:
    """The HIP library with every signature declared; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise EngineUnavailable(f"{LIB_PATH} is not built; run `python -m boitata_amd.build`")
    try:
        lib = C.CDLL(str(LIB_PATH))
    except OSError as e:  # pragma: no cover - depends on the machine
        raise EngineUnavailable(f"cannot load {LIB_PATH}: {e}") from e
    for name, res, args in _abi.SIGNATURES + _abi.EXTRA_SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return [n for n, _, _ in _abi.SIGNATURES + _abi.EXTRA_SIGNATURES]
