"""ghive_amd: MI355X-native (gfx950) implementation of the operators GHive offloads from Hive's
ql operator tree -- filter with stream compaction, hash-join build and probe, hash group-by and
key sort -- behind a C ABI (include/ghive/ghive.h, libghive_mi355x.so).

`ghive_amd.abi` holds the ctypes structures and loads nothing. `ghive_amd.ops` (operator
classes; needs torch and the built HIP library) is imported explicitly by its users.
"""
from . import abi  # noqa: F401
