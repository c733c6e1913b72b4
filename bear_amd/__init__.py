"""MI355X-native forward-model and likelihood engine for BeAR.

The compute path is the HIP library bear_amd/libbear_hip.so (C ABI in
include/bear_hip.h); bear_amd.api is its Python binding and bear_amd.synth
builds seeded synthetic workloads in the reference's file formats.
"""

__all__ = ["api", "synth"]
