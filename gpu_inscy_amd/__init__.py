"""MI355X-native INSCY subspace clustering (HIP engine behind a C ABI).

`gpu_inscy_amd.inscy` mirrors the reference's python/inscy.py API.
"""

__version__ = "0.1.0"
