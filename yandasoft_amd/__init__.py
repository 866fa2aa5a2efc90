"""yandasoft_amd -- MI355X-native convolutional gridding for yandasoft.

The product is ``libygrid.so`` (HIP kernels for gfx950 behind the C ABI in
``include/ygrid.h``).  This package is its Python binding (ctypes), used by the tests and
the benchmark; it mirrors the IVisGridder surface (initialiseGrid / grid / finaliseGrid /
finaliseWeights / initialiseDegrid / degrid).  Importing it never compiles anything; if the
library has not been built, loading fails loudly.
"""
from .abi import (  # noqa: F401
    YGRID_MEM_DEVICE,
    YGRID_MEM_HOST,
    Batch,
    Chunk,
    Context,
    Gridder,
    ImageDesc,
    LearnedCF,
    YgridError,
    imaging_chunk,
    lib,
    library_path,
    uvw_rotate,
)
