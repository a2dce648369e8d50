"""The compiler's resource report of the built library (kept by
physbam_amd/Makefile beside the objects): the hot kernels of the FP32 3D
step must not spill to scratch memory.  A spill in the fused G2P->P2G kernel
once slowed the whole step down 2.2x without changing a result, so it is
checked here rather than found in a benchmark.  Skipped when the library was
built elsewhere (no report)."""
import os
import re

import pytest

REPORT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "physbam_amd", "_build",
                      "mpmx_step.resources.txt")
# kernel (mangled name prefix) -> largest scratch bytes per lane allowed
HOT = {
    "_ZN4mpmx12g2p2g_kernelIfLi3ELb0E": 0,     # fused G2P -> P2G, the step's main kernel
    "_ZN4mpmx12g2p2g_kernelIfLi3ELb1E": 0,     # the same with wall friction
    "_ZN4mpmx10p2g_kernelIfLi3EE": 8,         # unfused P2G (first step of a run)
    "_ZN4mpmx10g2p_kernelIfLi3ELb0E": 0,
    "_ZN4mpmx10g2p_kernelIfLi3ELb1E": 0,
    "_ZN4mpmx11grid_kernelIfLi3ELb0E": 0,
    "_ZN4mpmx11grid_kernelIfLi3ELb1E": 0,
    "_ZN4mpmx16p2g_extra_kernelIfLi3EE": 0,   # unfused mover / slow-list P2G
    "_ZN4mpmx17fused_side_kernelIfLi3ELb0E": 0,  # fused step's movers and slow list
    "_ZN4mpmx16g2p_mover_kernelIfLi3ELb0EE": 0,
    "_ZN4mpmx28reflect_mass_momentum_kernelIfLi3E": 0,
    "_ZN4mpmx23reflect_velocity_kernelIfLi3E": 0,
}


def report():
    if not os.path.isfile(REPORT):
        pytest.skip("no compiler resource report (library not built here)")
    res, cur = {}, None
    with open(REPORT) as f:
        for line in f:
            m = re.search(r"Function Name: (\S+)", line)
            if m:
                cur = m.group(1)
                res[cur] = {}
                continue
            m = re.search(r"remark:\s+([A-Za-z ]+?)(?: \[[^\]]*\])?: (\d+)", line)
            if m and cur:
                res[cur][m.group(1).strip()] = int(m.group(2))
    return res


@pytest.mark.parametrize("prefix", sorted(HOT))
def test_hot_kernel_does_not_spill(prefix):
    res = report()
    names = [k for k in res if k.startswith(prefix)]
    assert names, f"{prefix} not in the report"
    for k in names:
        scratch = res[k].get("ScratchSize", 0)
        assert scratch <= HOT[prefix], (k, res[k])
