"""Errors of the drop-in estimators, as boitata.errors (python/boitata/errors.py,
crates/py/src/lib.rs `invalid()`, crates/py/src/io.rs:16-21): parameters out
of range raise InvalidInput at construction, fit or predict; per-target
failures are NaN, not exceptions; a file that cannot be opened raises
FileError."""


class BoitataError(Exception):
    """Base class of every error raised by the estimators."""


class InvalidInput(BoitataError, ValueError):
    """A parameter or an input array is out of range."""


class FileError(BoitataError, OSError):
    """A file could not be read or written."""
