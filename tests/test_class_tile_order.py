"""CPU restatement of the tile order of the class-packed sparse-level
Laplace M2L kernel (k_lap_m2l_packed_cls in wbfmm_amd/csrc/m2l_lap_impl.h).

Offsets with the same dz and the same dx^2 + dy^2 share the polar angle and
the distance, hence all three stage tables of the rotate / coaxial / rotate
back translation; only the azimuthal phases differ, and those are applied
per column.  For each such class (in the order of the key (dz + 3) 32 +
dx^2 + dy^2, members in cube-index order) a block of 32 targets lists its
(target, member) pairs with an occupied source in target-major order, and
deals them round-robin over T = max(ceil(n / 16), most pairs of one
target) tiles: position p goes to tile p mod T, column p div T.  A target's
pairs are consecutive positions, at most T of them, so a tile holds each
target at most once, and the lanes accumulating a target find its column
in tile t from (r, q, cnt) = (s mod T, s div T, count), s its first
position.

These tests run that construction on the host over random and extreme
occupancy patterns and check: every (offset, target) pair with an occupied
source is visited exactly once; a tile has at most 16 columns and distinct
targets; the column lookup finds exactly the target's column; the sizes fit
the kernel's encodings (pair = target | member << 5 in one byte, T <= 16,
at most 64 classes of at most 8 members).
"""

import numpy as np
import pytest

NOFF = 216
PACK_T = 32


def offset_classes(px, py, pz):
    """the kernel's class table of one parity: [(key, [k, ...]), ...]"""
    cls = {}
    for k in range(NOFF):
        dx, dy, dz = k % 6 - 2 - px, (k // 6) % 6 - 2 - py, k // 36 - 2 - pz
        if max(abs(dx), abs(dy), abs(dz)) <= 1:
            continue
        cls.setdefault((dz + 3) * 32 + dx * dx + dy * dy, []).append(k)
    return [(key, cls[key]) for key in sorted(cls)]


def build(vm, classes):
    """per class: n, T, pairs (target-major), first position and count per
    target"""
    out = []
    for key, mem in classes:
        cnt = np.array([sum(int(vm[j, k]) for k in mem) for j in range(PACK_T)])
        s = np.concatenate([[0], np.cumsum(cnt)[:-1]])
        n = int(cnt.sum())
        T = max(-(-n // 16), int(cnt.max())) if n else 0
        pairs = [(j, i) for j in range(PACK_T) for i, k in enumerate(mem) if vm[j, k]]
        out.append(dict(mem=mem, cnt=cnt, s=s, n=n, T=T, pairs=pairs))
    return out


def record(c, j):
    """the kernel's 16-bit record of target j in class c: r | q << 4 |
    cnt << 8, and 0 for a target without pairs (whose s may reach n = 16 T,
    where q = 16 would spill into the count bits)"""
    T, s, cnt = c["T"], int(c["s"][j]), int(c["cnt"][j])
    if not cnt:
        return 0
    r, q = s % T, s // T
    assert r < 16 and q < 16 and cnt < 16
    return r | (q << 4) | (cnt << 8)


def column_of(c, j, tau):
    """the kernel's lookup: column of target j in tile tau, or -1"""
    T = c["T"]
    z = record(c, j)
    r, q, cnt = z & 15, (z >> 4) & 15, z >> 8
    d = tau - r
    if d < 0:
        d += T
    if d >= cnt:
        return -1
    return q + (1 if r + d >= T else 0)


def run(vm, px, py, pz):
    classes = offset_classes(px, py, pz)
    assert len(classes) <= 64 and max(len(m) for _, m in classes) <= 8
    seen = np.zeros_like(vm, dtype=np.int64)
    tiles = 0
    for c in build(vm, classes):
        if c["n"] == 0:
            continue
        assert c["n"] <= 256 and c["T"] <= 16
        for tau in range(c["T"]):
            cols = [c["pairs"][p] for p in range(tau, c["n"], c["T"])]
            assert 1 <= len(cols) <= 16
            targets = [j for j, _ in cols]
            assert len(set(targets)) == len(targets)
            for col, (j, i) in enumerate(cols):
                assert j | (i << 5) < 256
                seen[j, c["mem"][i]] += 1
            for j in range(PACK_T):
                want = targets.index(j) if j in targets else -1
                assert column_of(c, j, tau) == want, (j, tau)
            tiles += 1
    assert np.array_equal(seen, vm.astype(np.int64))
    return tiles


def near(k, px, py, pz):
    dx, dy, dz = k % 6 - 2 - px, (k // 6) % 6 - 2 - py, k // 36 - 2 - pz
    return max(abs(dx), abs(dy), abs(dz)) <= 1


@pytest.mark.parametrize("parity", range(8))
@pytest.mark.parametrize("density", [0.0, 0.05, 0.38, 0.9, 1.0])
def test_every_pair_once_distinct_targets_per_tile(parity, density):
    px, py, pz = parity & 1, (parity >> 1) & 1, (parity >> 2) & 1
    rng = np.random.default_rng(100 * parity + int(100 * density))
    for _ in range(4):
        vm = rng.uniform(size=(PACK_T, NOFF)) < density
        for k in range(NOFF):
            if near(k, px, py, pz):
                vm[:, k] = False
        run(vm, px, py, pz)


def test_class_table_shares_angle_and_distance():
    # members of a class have the same polar angle and distance: the same
    # dz and dx^2 + dy^2; every far offset is in exactly one class
    for parity in range(8):
        px, py, pz = parity & 1, (parity >> 1) & 1, (parity >> 2) & 1
        classes = offset_classes(px, py, pz)
        ks = sorted(k for _, m in classes for k in m)
        assert ks == [k for k in range(NOFF) if not near(k, px, py, pz)]
        for key, mem in classes:
            d = [(k % 6 - 2 - px, (k // 6) % 6 - 2 - py, k // 36 - 2 - pz) for k in mem]
            assert len({(dz, dx * dx + dy * dy) for dx, dy, dz in d}) == 1
            assert mem == sorted(mem)


def test_class_records_at_the_limits():
    # every target holding every offset (n = 256 in a class of 8, T = 16);
    # and the first 16 targets full, the rest empty: in a class of 8,
    # n = 128 = 16 T, so the empty targets start at s = 16 T (q = 16)
    for parity in range(8):
        px, py, pz = parity & 1, (parity >> 1) & 1, (parity >> 2) & 1
        vm = np.zeros((PACK_T, NOFF), dtype=bool)
        for k in range(NOFF):
            vm[:, k] = not near(k, px, py, pz)
        run(vm, px, py, pz)
        vm[16:, :] = False
        run(vm, px, py, pz)


def test_single_target_with_every_offset():
    # one target holding all 189 offsets: T is set by its count per class,
    # so a class of 8 members takes 8 tiles of one column each
    for parity in range(8):
        px, py, pz = parity & 1, (parity >> 1) & 1, (parity >> 2) & 1
        vm = np.zeros((PACK_T, NOFF), dtype=bool)
        for k in range(NOFF):
            vm[5, k] = not near(k, px, py, pz)
        assert run(vm, px, py, pz) == 189
