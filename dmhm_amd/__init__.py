"""MI355X-native H-matrix arithmetic (HMat2d / HMat3d hot path of dmhm).

The compute lives in libdmhm_amd.so (hand-written gfx950 HIP kernels behind
the C-ABI in include/dmhm_amd.h); this package is the Python host mirror of
the reference interface plus the synthetic workload generators.
"""
from ._lib import (DMHM_F64, DMHM_F32, DmhmError, NonConformalError, NotConvergedError,  # noqa: F401
                   init, load, set_tile_engine)
from .hmat import (HMat2d, HMat3d, HBatch, EstimateTwoNorm, RoundedUpdate, Compress, RankModel,  # noqa: F401
                   last_product_stats,
                   build_natural_to_hierarchical_map, build_natural_to_hierarchical_map_device,
                   geometry_block_counts, geometry_product_terms,
                   geometry_leaves, check_direct_invert_program)
from . import workloads  # noqa: F401
