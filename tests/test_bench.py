"""bench.py's roofline reads the committed rocprofv3 traffic of the kernel it times: the name it
looks up must be the one the profile holds, or `roofline.traffic` would come out null."""
import os
import sys

from spheral_amd import abi

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_bench_finds_the_committed_traffic_of_its_kernel():
    for prec in (abi.SPX_PREC_FP64, abi.SPX_PREC_MIXED):
        t = bench.measured_traffic(bench.sph_kernel(prec))
        assert t is not None and t > 1e8, prec
