"""CPU restatement of the tile order of the sparse-level Laplace M2L kernel
(k_lap_m2l_packed in wbfmm_amd/csrc/m2l_mfma.hip).

A block holds 32 targets.  For each of the 216 offsets k of the 6x6x6
neighbourhood of the parents (216 slots, of which the 27 near ones are never
set) it lists the targets whose source box is occupied (lists[k][0..ncols[k]),
in target order) and runs them in 16-column tiles; `anym` has bit k set when
any target uses offset k.  The loop starts from next_tile(-1, 0) and, while
the current tile runs, looks up the next one with next_tile(k, t0).

These tests run that control flow on the host with every array access
bounds-checked, over random and extreme occupancy patterns, and check that
every (offset, target) pair with an occupied source is visited exactly once,
in increasing offset order, and that no lookup reads outside ncols/lists.
A lookup that started from ncols[-1] (reading before the array) fails here.
"""

import numpy as np
import pytest

NOFF = 216
PACK_T = 32


class Checked:
    """array whose every read is range-checked (no negative indices)"""

    def __init__(self, a):
        self.a = a

    def __getitem__(self, i):
        if not 0 <= i < len(self.a):
            raise IndexError(f"read at {i} outside [0, {len(self.a)})")
        return self.a[i]


def build_lists(vm):
    """vm[j][k]: target j has an occupied source at offset k"""
    ncols = np.zeros(NOFF, dtype=np.int64)
    lists = np.full((NOFF, PACK_T), -1, dtype=np.int64)
    for k in range(NOFF):
        n = 0
        for j in range(PACK_T):
            if vm[j, k]:
                lists[k, n] = j
                n += 1
        ncols[k] = n
    anym = vm.any(axis=0)
    return Checked(ncols), [Checked(r) for r in lists], Checked(anym)


def next_tile(k, t0, ncols, anym):
    # the kernel's next_tile: the next 16 columns of this offset, else the
    # first tile of the next used offset (216 when there is none)
    if 0 <= k < NOFF and t0 + 16 < ncols[k]:
        return k, t0 + 16
    k += 1
    while k < NOFF and not anym[k]:
        k += 1
    return k, 0


def run_tiles(vm):
    ncols, lists, anym = build_lists(vm)
    visited = []
    k, t0 = next_tile(-1, 0, ncols, anym)
    while k < NOFF:
        for col in range(16):
            sl = t0 + col
            if sl < ncols[k]:
                j = lists[k][sl]
                assert 0 <= j < PACK_T
                visited.append((k, j))
        k, t0 = next_tile(k, t0, ncols, anym)
    return visited


def expected(vm):
    return [(k, j) for k in range(NOFF) for j in range(PACK_T) if vm[j, k]]


def near_mask():
    # offsets of the 27 near neighbours of a child (never listed); the kernel
    # skips them for every class, the pattern below uses class 0
    m = np.zeros(NOFF, dtype=bool)
    for k in range(NOFF):
        dx, dy, dz = k % 6 - 2, (k // 6) % 6 - 2, k // 36 - 2
        if -1 <= dx <= 1 and -1 <= dy <= 1 and -1 <= dz <= 1:
            m[k] = True
    return m


@pytest.mark.parametrize("seed", range(12))
@pytest.mark.parametrize("density", [0.0, 0.02, 0.38, 0.6, 1.0])
def test_every_pair_once_in_offset_order(seed, density):
    rng = np.random.default_rng(1000 * seed + int(100 * density))
    vm = rng.uniform(size=(PACK_T, NOFF)) < density
    vm[:, near_mask()] = False
    got = run_tiles(vm)
    assert got == expected(vm)


@pytest.mark.parametrize("k_only", [0, 1, 100, 214, 215])
@pytest.mark.parametrize("ntgt", [1, 15, 16, 17, 31, 32])
def test_single_offset_and_tile_boundaries(k_only, ntgt):
    # one used offset at either end of the range, with 1..32 columns:
    # one tile up to 16 columns, two above
    vm = np.zeros((PACK_T, NOFF), dtype=bool)
    vm[:ntgt, k_only] = True
    assert run_tiles(vm) == expected(vm)


def test_no_used_offset_runs_no_tile():
    vm = np.zeros((PACK_T, NOFF), dtype=bool)
    assert run_tiles(vm) == []


def test_lookup_from_before_the_array_is_caught():
    # the first lookup of an earlier version tested ncols[k] with k = -1;
    # the checked arrays turn that read into an error
    vm = np.zeros((PACK_T, NOFF), dtype=bool)
    vm[3, 5] = True
    ncols, _, anym = build_lists(vm)

    def next_tile_unguarded(k, t0):
        if t0 + 16 < ncols[k]:
            return k, t0 + 16
        return next_tile(k, t0, ncols, anym)

    with pytest.raises(IndexError):
        next_tile_unguarded(-1, 0)
