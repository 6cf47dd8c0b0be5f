"""mose_get_state_noalloc, the variant of mose_get_state that stages the conversion through a
fixed device array instead of a buffer it allocates: bit-identical to mose_get_state on two calls
in a row (the second on a smaller block), from two contexts on one GPU, and a block beyond the
array refused with MOSE_E_SIZE."""
import numpy as np
import pytest

from mose_amd import abi, cases
from mose_amd.lib import GpuSolver, MoseError

pytestmark = pytest.mark.gpu

MAX_CELLS = (1 << 22) // 5      # padded cells the 32 MiB staging array holds


def _padded_cells(dims):
    return int(np.prod([d + 4 for d in dims]))


def _same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_noalloc_equals_get_state_twice_second_smaller():
    case = cases.multiblock_chain(nblocks=4, seed=23, lo=12, hi=40)
    sizes = [_padded_cells(b.dims) for b in case.blocks]
    big, small = int(np.argmax(sizes)) + 1, int(np.argmin(sizes)) + 1
    assert sizes[small - 1] < sizes[big - 1] <= MAX_CELLS
    print("noalloc: blocks", big, small, "padded cells", sizes[big - 1], sizes[small - 1], flush=True)
    g = GpuSolver(case)
    g.advance(3)
    first = g.get_state_noalloc(big)          # first call on this device in a fresh process
    second = g.get_state_noalloc(small)       # smaller: the array keeps the first call's tail
    assert _same_bits(first, g.get_state(big))
    assert _same_bits(second, g.get_state(small))
    assert np.all(np.isfinite(first)) and np.all(first[2:-2, 2:-2, 2:-2, 0] > 0)   # [k][j][i][v]
    g.close()


def test_noalloc_two_contexts_one_gpu():
    ga = GpuSolver(cases.vortex(n=24, riemann="HLLC"))
    gb = GpuSolver(cases.vortex(n=16, riemann="AUSM+"))
    ga.advance(2)
    gb.advance(2)
    for g in (ga, gb, ga):
        assert _same_bits(g.get_state_noalloc(1), g.get_state(1))
    ga.close()
    gb.close()


def test_noalloc_refuses_block_beyond_array():
    dims = (96, 96, 92)
    assert _padded_cells(dims) > MAX_CELLS
    g = GpuSolver(cases.vortex(dims=dims, riemann="HLLC"))
    with pytest.raises(MoseError) as ei:
        g.get_state_noalloc(1)
    assert ei.value.code == abi.E_SIZE
    g.close()
