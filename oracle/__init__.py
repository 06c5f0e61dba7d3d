"""Oracle: CPU restatement of the reference's hot path (liboracle.so, built from the C++
sources in this directory by the Makefile).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
use it as the checker / CPU baseline; the product (hashdag_amd) never imports it.
"""
from .binding import Oracle, OracleColors, OracleDAG, build, orc  # noqa: F401
