"""rknn_amd -- Python mirror of the gfx950 RkNN engine's C ABI.

See include/rknn.h for the semantics.  The compute path is lib/librknn.so
(hand-written HIP kernels); this package only marshals numpy buffers.
"""
from ._lib import LIB_PATH, load  # noqa: F401
from .engine import (  # noqa: F401
    Batch,
    Context,
    Dataset,
    RkError,
    Server,
    default_seed,
    device_count,
    generate,
    generate_queries,
    load_text,
    write_text,
)
