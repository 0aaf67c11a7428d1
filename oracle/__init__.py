"""The oracle: a CPU restatement of model3d's MarchingCubesSearch path.

TEST INFRASTRUCTURE ONLY.  tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py use it as the checker; the product package
(model3d_amd) never imports it.  See mc_oracle.c for what pins each part.
"""
from .binding import OracleBackend, oracle_bounds, oracle_case_table, oracle_contains, oracle_search  # noqa: F401
