"""frank_amd: MI355X-native BLR-LU / BLR-QR hot path of FRANK (gfx950 HIP kernels behind a C ABI)."""
from ._lib import FrankError, lib  # noqa: F401  (loads libfrank_amd.so; raises if it is missing)
from .blr import (BlrMatrix, Context, compress, default_context, gemm, getrf_nopiv_batched,  # noqa: F401
                  lr_add_batched, trsm)
