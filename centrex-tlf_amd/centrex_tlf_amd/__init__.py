"""MI355X-native Lindblad / optical-Bloch solver for TlF molecular-beam simulations."""

from . import lindblad

__all__ = ["lindblad"]
__version__ = "0.1.0"
