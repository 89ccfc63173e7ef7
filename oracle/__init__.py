"""TEST INFRASTRUCTURE — the BM3D oracle package (see oracle/bm3d_oracle.cpp)."""
from .oracle import *  # noqa: F401,F403
from .oracle import build, lib, make_params, transform, transform_matrix, block_distance, filter_group  # noqa: F401
from .oracle import num_blocks, bm3d, match, match_blocks, aggregate, agg_z, vaggregate, OracleParams  # noqa: F401
