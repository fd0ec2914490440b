"""cftamd - the Python side of the MI355X engine for cft-fp256's hot path.

libcft.so (cft-fp256_amd/lib) exports the reference library's C ABI for the
elementwise binary128/binary256 arithmetic and the sequencer executor; this
package binds it with ctypes, as the reference's own binding does.
"""

from . import _lib
from ._lib import (CftError, FP32, FP64, FP128, FP256, RNE, RTZ, RDN, RUP, RMM,
                   check, lib, library_path)
from .core import (Buffer, Device, DeviceArray, Program, apply_knobs, check_image, last_plan,
                   run_multi, tune_batch, tune_multi)

__all__ = ["CftError", "Device", "Buffer", "DeviceArray", "Program", "check_image", "run_multi", "last_plan",
           "tune_multi", "tune_batch", "apply_knobs",
           "FP32", "FP64", "FP128", "FP256", "RNE", "RTZ", "RDN", "RUP", "RMM", "check", "lib",
           "library_path", "_lib"]
