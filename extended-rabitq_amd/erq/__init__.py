"""erq -- MI355X-native Extended-RaBitQ IVF search (host side).

The search itself runs in lib/liberq.so (HIP, gfx950) behind the C-ABI declared in
include/erq.h; this package holds the Python view of that ABI, the HA index format
(inc/index/IVF.hpp:219-293 of the reference) and the synthetic workload generators.
"""
from .index import IVFIndex, build_index, random_rotation  # noqa: F401
