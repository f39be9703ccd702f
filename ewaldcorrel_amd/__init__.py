"""MI355X-native SPME reciprocal-space engine (drop-in for the libsfac.so C ABI
and the ewald.py entry points of frobnitzem/EwaldCorrel)."""
from .ewald import (Sfac, SfacBatch, SfacServer, LofS, SofL, ewald_f,  # noqa: F401
                    ewald_kernel, describe_sfac, last_error, SFAC_FP64, SFAC_FP32, SFAC_FIXED)

__all__ = ["Sfac", "SfacBatch", "SfacServer", "LofS", "SofL", "ewald_f", "ewald_kernel",
           "describe_sfac", "last_error", "SFAC_FP64", "SFAC_FP32", "SFAC_FIXED"]
