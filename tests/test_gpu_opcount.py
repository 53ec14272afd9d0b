"""The op-counting run of the FP32 screening node (gra_hip_screen_opcount): its
tallies cover every node of the counted points, and one accepted node's count
equals an independent count of the same node from the gfx950 assembly of the
production kernel, taken by hand along an accepted node's path (offshellFF = 2,
series SolvePz step): 49 FMA, 18 multiplies, 13 adds or subtractions and 9
transcendental instructions, 138 FLOP with FMA = 2 (each phi row adds its four
direction projections, hoisted out of the row's nodes). Contraction changes
instruction counts (a mul and an add become one FMA) but not FLOP, so the
source-level tally and the assembly must agree in FLOP."""
import ctypes as C
import sys

import numpy as np
import pytest

from graniitti_amd import abi
from graniitti_amd.process import MContinuum
from conftest import small_tables

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_opcount_of_the_fp32_screening_node():
    cfg = small_tables(abi.cfg_default(2), 1000)
    say("case: op count of the FP32 screening node, preset 2, one VEGAS pass of 2^17")
    with MContinuum(cfg) as P:
        grid = abi.GraGrid()
        P._chk(P.lib.gra_hip_grid_init(C.byref(grid), cfg.dim, 128))
        st = P.vegas_iterate(grid, 1 << 17, 5, 0)
        oc = (C.c_double * 12)()
        P._chk(P.lib.gra_hip_screen_opcount(P.h, 4096, oc))
        # a process without the FP32 Regge node is refused
        bad = small_tables(abi.cfg_default(2), 1000)
        bad.amp_precision = abi.GRA_AMP_FP64
        with MContinuum(bad) as Q:
            Q.vegas_iterate(grid, 1 << 12, 5, 0)
            assert Q.lib.gra_hip_screen_opcount(Q.h, 16, oc) == abi.GRA_E_STATE
        P._chk(P.lib.gra_hip_screen_opcount(P.h, 4096, oc))
    o = np.array(oc[:])
    nodes = (cfg.n_loop_phi + 1) * (cfg.n_loop_kt + 1)
    flop = o[3] + o[4] + 2 * o[5] + o[6]
    say(f"points {o[0]:.0f} nodes {o[1]:.0f} accepted {o[2]:.0f}: add {o[3]:.0f} mul {o[4]:.0f} "
        f"fma {o[5]:.0f} trans {o[6]:.0f} cmp {o[7]:.0f}; {flop / o[1]:.2f} FLOP per node; one "
        f"accepted node: add {o[8]:.0f} mul {o[9]:.0f} fma {o[10]:.0f} trans {o[11]:.0f}")
    assert st.n_valid > 4096 and o[0] == 4096 and o[1] == o[0] * nodes
    assert 0.5 * o[1] < o[2] <= o[1]
    assert o[8] + o[9] + 2 * o[10] + o[11] == 138
    assert o[11] == 9
    # every node runs at least the SolvePz discriminant test (4 fma, 1 sub, 1 add,
    # 2 mul, 1 compare), an accepted one at most 138 FLOP; each row adds its
    # projections (4 mul, 4 fma) and two FP64 adds, the point epilogue 15
    rows = cfg.n_loop_phi + 1
    assert o[1] * 10 <= flop <= o[1] * 138 + o[0] * (14 * rows + 15)
    assert o[7] >= o[1]
