"""MI355X-native engine for the DNSPIR server response path.

`dnspir_amd.pir` mirrors the `dnspir-pir` crate API over libdnspir.so
(hand-written HIP kernels for gfx950 behind the C ABI in include/dnspir.h).
"""
from ._lib import LANES, N, DnspirError, lib  # noqa: F401
from . import pir  # noqa: F401
