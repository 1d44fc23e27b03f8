"""Extension loading (epcq/utils.py:2-15 of the reference)."""


def _get_c_extension():
    """Returns the `epcq.csrc` op module.

    Unlike the reference, which warns and returns None, a missing or broken
    HIP library is an error here: there is no slow fallback path."""
    try:
        import epcq.csrc as _C
    except Exception as e:  # pragma: no cover - exercised only without the library
        raise RuntimeError("HIP extension epcq.csrc could not be loaded: " + str(e)) from e
    if not hasattr(_C, "scatter"):
        raise RuntimeError("HIP extension epcq.csrc is incomplete (no `scatter` op)")
    return _C
