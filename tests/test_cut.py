"""Uniform block cut positions (fvm_cut_uniform, the reference's
ClassGeomBlock%cut_uniform, src/base/defs_block_geom.f90:757-769): host
logic, no GPU.

Pinned by what the rule fixes: the pieces tile the range with shared
interface planes, their cell counts differ by at most one and never
decrease (each cut takes the floor of the remaining cells over the remaining
pieces), which determines the cut uniquely."""
import itertools

import pytest

import fvmgpu


@pytest.mark.parametrize("n,ncut", list(itertools.product([2, 3, 7, 10, 11, 64, 101, 1000],
                                                           [1, 2, 3, 4, 7, 16])))
def test_cut_positions(n, ncut):
    if ncut > n - 1:
        with pytest.raises(fvmgpu.FvmError):
            fvmgpu.cut_uniform(n, ncut)
        return
    r = fvmgpu.cut_uniform(n, ncut)
    assert len(r) == ncut and r[0][0] == 0 and r[-1][1] == n - 1
    for (a0, a1), (b0, b1) in zip(r, r[1:]):
        assert a1 == b0                      # shared interface plane
    cells = [b - a for a, b in r]
    assert sum(cells) == n - 1 and min(cells) >= 1
    assert max(cells) - min(cells) <= 1
    assert cells == sorted(cells)            # smaller pieces first


def test_cut_examples():
    # 10 cells in 3 pieces: floor(10/3) = 3, floor(7/2) = 3, then 4
    assert fvmgpu.cut_uniform(11, 3) == [(0, 3), (3, 6), (6, 10)]
    assert fvmgpu.cut_uniform(5, 4) == [(0, 1), (1, 2), (2, 3), (3, 4)]
    with pytest.raises(fvmgpu.FvmError):
        fvmgpu.cut_uniform(5, 0)
