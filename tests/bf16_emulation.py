"""numpy emulation of the BF16x3 operand split used by the packed GEMM:
x = hi + lo with hi = bf16(x), lo = bf16(x - hi), both rounded to nearest even
from FP32."""
import numpy as np


def bf16_rne(x):
    """FP32 -> BF16 round-to-nearest-even, returned as float64 values."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16x3_split(x):
    hi = bf16_rne(np.asarray(x, dtype=np.float64).astype(np.float32))
    lo = bf16_rne((np.asarray(x, dtype=np.float64) - hi).astype(np.float32))
    return hi, lo
