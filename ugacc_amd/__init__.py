"""MI355X (gfx950) closed-shell CCSD / CCSD(T) behind ugacc's plugin interface.

- ``ugacc_amd.ugx``     ctypes binding of the C ABI (include/ugx.h, libugx.so)
- ``ugacc_amd.plugin``  the host-side mirror of the reference plugin entry:
  ``read_options`` / ``ugacc`` and the ``CCWfn`` object (ugacc.cc:54-117,
  ccwfn.h)
- ``ugacc_amd.build``   in-tree hipcc build of libugx.so for gfx950
"""
