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


def _share_hip_runtime():
    """One HIP runtime per process: PyTorch-ROCm ships its own libamdhip64
    (same SONAME, requested by an unversioned name), so if this library bound
    /opt/rocm's first, importing torch afterwards would bring up a second HIP
    and HSA runtime that sees no GPU. Loading torch's copy globally first makes
    both bind to it; without torch, /opt/rocm's is used."""
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    hip = Path(spec.origin).parent / "lib" / "libamdhip64.so"
    if hip.exists():
        try:
            C.CDLL(str(hip), mode=C.RTLD_GLOBAL)
        except OSError:
            pass


def load() -> C.CDLL:
    """The HIP library with every signature declared; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise EngineUnavailable(f"{LIB_PATH} is not built; run `python -m boitata_amd.build`")
    _share_hip_runtime()
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
