"""MI355X-native batched shading executor for Open Shading Language groups.

Host side of the C ABI declared in include/oslmi.h; compute runs in the HIP
library liboslmi.so (gfx950). PyTorch is used only for device buffers and
streams.
"""
from ._capi import OSLMIError, lib, check  # noqa: F401

__all__ = ["OSLMIError", "lib", "check"]
