"""qed_fermion_amd: MI355X-native HMC fermion hot path of qed_fermion.

HIP/CDNA4 kernels (qed_fermion_amd/csrc, built into libqf_hip.so) behind a
C ABI (include/qed_fermion_amd.h), driven from PyTorch-ROCm through ctypes.
"""
from ._lib import QfError  # noqa: F401

__all__ = ["QfError"]
