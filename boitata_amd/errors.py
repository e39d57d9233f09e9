"""Errors of the drop-in estimators, as boitata.errors (crates/py/src/lib.rs
`invalid()`): parameters out of range raise InvalidInput at construction, fit
or predict; per-target failures are NaN, not exceptions."""


class InvalidInput(ValueError):
    """A parameter or an input array is out of range."""
