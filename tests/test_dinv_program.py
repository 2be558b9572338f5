"""The lowering of HMat2d/HMat3d::DirectInvert (src/HMat2d/Invert-incl.hpp:
17-125) into merged launches, on the host (no GPU): every task of every step
stays inside the arena and the workspace, and the program holds exactly the
recursion's products. For a NODE with nc children the reference's
Gauss-Jordan makes nc(nc-1) row-scaling products, nc(nc-1)nc/2 eliminations
and nc(nc-1)nc/2 back substitutions, nc(nc-1)(nc+1) in all, grouped into
nc + 2(nc-1) independent steps; every diagonal dense leaf is inverted once."""
import pytest

import dmhm_amd


@pytest.mark.parametrize("dim,sizes,L,rank,strong", [(2, [16, 16], 3, 16, True), (2, [16, 16], 3, 12, False),
                                                     (2, [20, 12], 3, 8, True), (3, [8, 8, 8], 3, 8, True),
                                                     (3, [6, 5, 4], 2, 4, True), (2, [64, 64], 4, 8, True),
                                                     (2, [9, 9], 1, 4, True)])
def test_program_counts_and_bounds(dim, sizes, L, rank, strong):
    st = dmhm_amd.check_direct_invert_program(dim, sizes, L, rank, strong)
    nc = 1 << dim
    # Diagonal NODE blocks: one per box on levels 0 .. L-2; diagonal leaves:
    # one per box on level L-1 (every box of a level has a diagonal block).
    boxes = [1]
    for _ in range(1, L):
        boxes.append(boxes[-1] * nc)
    nodes = sum(boxes[:-1])
    assert st["products"] == nodes * nc * (nc - 1) * (nc + 1)
    assert st["plans"] == nodes * (nc + 2 * (nc - 1))
    assert st["leaf_inversions"] == boxes[-1]


def test_program_rejects_too_wide_sketches():
    with pytest.raises(dmhm_amd.DmhmError):
        dmhm_amd.check_direct_invert_program(2, [16, 16], 3, 62, True, oversample=4)
