"""Exceptions raised with the same class and messages as the reference.

Restates halotools/custom_exceptions.py:8-13 (HalotoolsError).
"""

__all__ = ("HalotoolsError",)


class HalotoolsError(Exception):
    """Base class of Halotools-specific exceptions (halotools/custom_exceptions.py:8-13)."""

    def __init__(self, message):
        super(HalotoolsError, self).__init__(message)
