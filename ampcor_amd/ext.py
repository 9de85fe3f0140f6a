"""Discovery of the compiled correlator.

Mirrors pkg/ampcor/ext/__init__.py:19-51 of the reference: the accelerated
module is ``None`` when it cannot be imported.  Unlike the reference there is
no silent CPU fallback behind it: :func:`require` raises, so a GPU path can
never run on anything but the HIP library.
"""

try:
    from . import ampcor_hip as libampcor_hip
except ImportError as error:  # not built
    libampcor_hip = None
    import_error = error
else:
    import_error = None


def require():
    """Return the ``ampcor_hip`` module or raise if the library is not built."""
    if libampcor_hip is None:
        raise ImportError(
            "the ampcor_hip extension is not built (run __graft_entry__.build()): "
            f"{import_error}")
    return libampcor_hip


def device_count():
    """Number of HIP devices the library sees (0 when not built)."""
    if libampcor_hip is None:
        return 0
    return libampcor_hip.device_count()
