"""MI355X-native multi-way INT32 equi-join engine: a drop-in for the reference's
Contest::execute (src/execute.cpp:598) with HIP kernels for gfx950.

Import as importlib.import_module("hash-join-optimizer_amd") with the repository root on sys.path.
"""
from . import abi  # noqa: F401
from .abi import EngineMissing  # noqa: F401
from .plan import (CacheError, Codec, Column, ColumnarTable, DataType, JoinNode,  # noqa: F401
                   Plan, PlanNode, ScanNode, Table, read_cache, sorted_rows, write_cache)
from .engine import (Batch, Context, DeviceResult, DeviceTable, Directory,  # noqa: F401
                     EngineError, Server)

from .build import hash_algo as _hash_algo

# the reference's CMake HASH_ALGO value this library is (CMakeLists.txt:173-187): "hip", the GPU
# tables (bucketized open addressing for left-deep chains whose build keys are unique, the
# unchained directory for every other join); the built library reports the same through
# hj_hash_algo()
HASH_ALGO = _hash_algo("hip")
