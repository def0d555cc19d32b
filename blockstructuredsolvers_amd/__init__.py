"""blockstructuredsolvers_amd -- MI355X-native recursive Schur-complement
block-tridiagonal Cholesky solver (the batched GPU path of BlockStructuredSolvers).

The compute path is libbss.so (hand-written HIP kernels for gfx950, C ABI in
include/bss.h); this package is the thin host mirror of the reference API.
"""
from .solver import (BlockTriDiagData_batch, BlockTriDiagData_batched, BlockTriDiagData_group,
                     BlockTriDiagData_seq, build_smoother, factorize, factorize_,
                     factorize_solve, initialize_batch, initialize_batched,
                     initialize_batched_group, initialize_seq, reference_schedule, solve,
                     solve_)

__all__ = ["BlockTriDiagData_batched", "BlockTriDiagData_group", "BlockTriDiagData_seq",
           "BlockTriDiagData_batch", "initialize_batch",
           "initialize_batched", "initialize_batched_group", "initialize_seq", "factorize",
           "solve", "factorize_solve", "factorize_", "solve_", "reference_schedule", "build_smoother"]
