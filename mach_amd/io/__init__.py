"""Readers of acquisition files that exist locally (the reference's ``mach.io``; nothing here downloads)."""

from . import must, uff

__all__ = ["must", "uff"]
