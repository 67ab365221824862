"""MI355X UW shallow-cumulus column physics (compute_uwshcu) behind a C ABI.

Modules:
  generator -- seeded synthetic physics chunks (numpy)
  _capi     -- ctypes binding of include/uwshcu.h
  device    -- compute_uwshcu_inv on device-resident torch tensors
  build     -- in-tree hipcc build of libuwshcu_hip.so
"""
__all__ = ["generator", "build"]
