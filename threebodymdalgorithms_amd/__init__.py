"""MI355X-native three-body (Axilrod-Teller) force engine for the NATA / AUTA /
P3BCA triplet hot path.

The compute path is libtb3.so (hand-written gfx950 HIP kernels behind the C ABI
in include/tb3.h); this package is a thin host binding over it.
"""
from ._abi import (ALGORITHMS, AUTA, FP32_MIXED, FP64, FP64_EXACT, NATA, P3BCA, PRECISIONS, TB3Error, CLI_PATH,
                   LIB_PATH)
from .engine import Engine, device_count, read_csv, write_csv

__all__ = ["Engine", "TB3Error", "device_count", "read_csv", "write_csv", "NATA", "P3BCA", "AUTA", "FP64_EXACT",
           "FP64", "FP32_MIXED", "ALGORITHMS", "PRECISIONS", "LIB_PATH", "CLI_PATH"]
