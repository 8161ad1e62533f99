"""Measurement operators and gate constants."""
from . import gates
from .ment import Ment, Measurement

__all__ = ["Ment", "Measurement", "gates"]
