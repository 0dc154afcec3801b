"""flowunsteady_amd -- MI355X-native rVPM particle-interaction step.

The compute path is libvpmx.so (hand-written HIP for gfx950 behind the C ABI
in include/vpmx.h).  `vpm` mirrors the FLOWVPM particle-field interface that
FLOWUnsteady drives; `workloads` generates the seeded synthetic particle sets.
"""
from . import vpm, workloads  # noqa: F401

__version__ = "0.1.0"
