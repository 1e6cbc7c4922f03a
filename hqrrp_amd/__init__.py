"""MI355X-native HQRRP (Householder QR with randomized column pivoting).

Python face of libhqrrp_amd.so.  It mirrors the reference's two entry points
(``dgeqp4`` and ``NoFLA_HQRRP_WY_blk_var4``, host arrays) and exposes the
device-pointer interface on torch tensors.  All numerical work runs in the
library's HIP kernels; nothing here computes a factorization.
"""
from .api import (BatchPlan, Handle, PackedBatch, dgeqp4, fixed_pivot_qr,
                  geqp3_device, geqrf_device, hqrrp, hqrrp_batched, hqrrp_device, hqrrp_ut,
                  hqrrp_ut_device,
                  omega_carry, orgqr_device, ormqr_device, qr_fixed, qr_fixed_adjoint,
                  reference_omega, set_omega_carry, to_colmajor, from_colmajor)
from ._lib import HQRRPError, Options, lib

__all__ = ["BatchPlan", "Handle", "PackedBatch", "hqrrp_batched", "dgeqp4", "fixed_pivot_qr", "geqp3_device", "geqrf_device", "hqrrp",
           "hqrrp_device", "hqrrp_ut", "hqrrp_ut_device", "orgqr_device", "ormqr_device", "qr_fixed", "qr_fixed_adjoint",
           "reference_omega", "omega_carry", "set_omega_carry", "to_colmajor", "from_colmajor", "HQRRPError",
           "Options", "lib"]
