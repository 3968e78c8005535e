"""The link-skipping rule of the candidate labeling (k_cand_local / k_cand_merge): a voxel
links a foreground forward neighbour only when no neighbour already known to be in its set
is adjacent to it (the two offsets differ by at most 1 on every axis), visiting (0, 0, +1)
first; in the merge the candidates inside the tile count as known.

This is a CPU model of that rule over the same tile-local and merge link sets (tiles
clipped to the volume, wrapped linear offsets for links that leave the tile), checked
against the oracle's labels on random volumes with ragged tiles, wrap and small widths.
A model that drops a needed link (for example one that counts a neighbour as known before
its set is joined) fails it.  The tile-local kernel's gathering of the 13 forward
neighbours' mask bits from five rows is checked against the definition (a neighbour inside
the clipped tile that is a candidate)."""
import numpy as np
import pytest

import oracle


def fwd(k):  # forward offset k as (dx, dy, dz), the order of ccl.hip's fwd_offset
    if k < 9:
        return (k % 3 - 1, k // 3 - 1, 1)
    if k < 12:
        return (k - 10, 1, 0)
    return (1, 0, 0)


OFF = [fwd(k) for k in range(13)]
ADJ = [sum(1 << p for p in range(13)
           if p != q and all(abs(OFF[p][i] - OFF[q][i]) <= 1 for i in range(3)))
       for q in range(13)]
ORDER = [4, 10, 0, 1, 2, 3, 5, 6, 7, 8, 9, 11, 12]


def label_with_rule(m, tile, skip=True, wrong=False):
    D, H, W = m.shape
    TX, TY, TZ = tile
    N = m.size
    flat = m.ravel().astype(bool)
    par = np.arange(N)

    def find(x):
        while par[x] != x:
            par[x] = par[par[x]]
            x = par[x]
        return x

    def unite(a, b):
        a, b = find(a), find(b)
        if a != b:
            par[max(a, b)] = min(a, b)

    links = 0
    for z0 in range(0, D, TZ):
        for y0 in range(0, H, TY):
            for x0 in range(0, W, TX):
                ex, ey, ez = min(TX, W - x0), min(TY, H - y0), min(TZ, D - z0)
                for lz in range(ez):
                    for ly in range(ey):
                        for lx in range(ex):
                            x, y, z = x0 + lx, y0 + ly, z0 + lz
                            v = (z * H + y) * W + x
                            if v == 0 or not flat[v]:
                                continue
                            fin, fout, at = 0, 0, {}
                            for q, (dx, dy, dz) in enumerate(OFF):
                                nx, ny, nz = lx + dx, ly + dy, lz + dz
                                if 0 <= nx < ex and 0 <= ny < ey and nz < ez:
                                    u = ((z + dz) * H + y + dy) * W + x + dx
                                    if flat[u]:
                                        fin |= 1 << q
                                        at[q] = u
                                    continue
                                u = v + dz * W * H + dy * W + dx  # wrapped linear offset
                                if u < N and flat[u]:
                                    fout |= 1 << q
                                    at[q] = u
                            known = 0  # tile-local pass
                            for q in ORDER:
                                if (fin >> q) & 1:
                                    if not (skip and ADJ[q] & known):
                                        unite(v, at[q])
                                        links += 1
                                    known |= 1 << q
                            # merge pass: candidates inside the tile are in v's set already
                            # (the planted mistake counts every neighbour as known up front)
                            known = (fin | fout) if wrong else fin
                            for q in ORDER:
                                if (fout >> q) & 1:
                                    if not (skip and ADJ[q] & known):
                                        unite(v, at[q])
                                        links += 1
                                    known |= 1 << q
    lab = np.array([find(i) if i and flat[i] else 0 for i in range(N)], np.uint32)
    return lab.reshape(m.shape), links


def test_adjacency_table():
    # (0, 0, +1) touches every other forward neighbour; (+1, 0, 0) and (-1, +1, +1) do not
    assert ADJ[4] == 0x1FFF & ~(1 << 4)
    assert not (ADJ[12] >> 6) & 1 and OFF[6] == (-1, 1, 1) and OFF[12] == (1, 0, 0)
    assert sorted(ORDER) == list(range(13))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rule_matches_oracle_labels(seed):
    rng = np.random.default_rng(seed)
    kept = full = 0
    for _ in range(12):
        D, H, W = (int(v) for v in rng.integers(2, 9, 3))
        m = (rng.random((D, H, W)) < rng.uniform(0.15, 0.7)).astype(np.uint8)
        tile = tuple(int(v) for v in rng.integers(1, 5, 3))
        got, n1 = label_with_rule(m, tile)
        np.testing.assert_array_equal(got, oracle.ccl3d(m))
        _, n0 = label_with_rule(m, tile, skip=False)
        kept, full = kept + n1, full + n0
    assert kept < 0.6 * full  # the rule removes most links on these densities


def test_a_wrong_rule_is_caught():
    # counting an outside neighbour as known before it is linked splits regions
    rng = np.random.default_rng(7)
    bad = 0
    for _ in range(20):
        D, H, W = (int(v) for v in rng.integers(2, 8, 3))
        m = (rng.random((D, H, W)) < 0.45).astype(np.uint8)
        tile = tuple(int(v) for v in rng.integers(1, 4, 3))
        got, _ = label_with_rule(m, tile, wrong=True)
        bad += not np.array_equal(got, oracle.ccl3d(m))
    assert bad > 0


def five_row_mask(bits, vx, vy, vz, ey, ez, RTY=8):
    """k_cand_local's forward-neighbour bits from five mask rows: x-1, x, x+1 of rows
    (y-1, z+1), (y, z+1), (y+1, z+1), (y+1, z) in three bits each, then x+1 of (y, z)."""
    M = (1 << 64) - 1

    def tri(row, ok):
        w = bits[row if ok else 0]
        t = ((w >> (vx - 1)) & 7) if vx else ((w << 1) & M) & 7
        return t if ok else 0

    yl, yr, zr = vy > 0, vy + 1 < ey, vz + 1 < ez
    r0 = vy + RTY * vz
    return (tri(r0 - 1 + RTY, yl and zr) | tri(r0 + RTY, zr) << 3 |
            tri(r0 + 1 + RTY, yr and zr) << 6 | tri(r0 + 1, yr) << 9 |
            ((bits[r0] >> vx >> 1) & 1) << 12)


@pytest.mark.parametrize("seed", [0, 1])
def test_five_row_neighbour_mask_equals_the_definition(seed):
    # bit q set exactly when forward neighbour q (OFF[q]) is a candidate inside the tile
    # clipped to (ex, ey, ez); candidates exist only inside it, as the records guarantee
    rng = np.random.default_rng(seed)
    RTX, RTY, RTZ = 64, 8, 16
    for _ in range(40):
        ex, ey, ez = (int(rng.integers(1, n + 1)) for n in (RTX, RTY, RTZ))
        cand = np.zeros((RTZ, RTY, RTX), bool)
        cand[:ez, :ey, :ex] = rng.random((ez, ey, ex)) < 0.4
        bits = [int(sum(1 << x for x in range(RTX) if cand[z, y, x]))
                for z in range(RTZ) for y in range(RTY)]
        for _ in range(200):
            vx, vy, vz = int(rng.integers(0, ex)), int(rng.integers(0, ey)), int(rng.integers(0, ez))
            want = 0
            for q, (dx, dy, dz) in enumerate(OFF):
                x, y, z = vx + dx, vy + dy, vz + dz
                if 0 <= x < ex and 0 <= y < ey and 0 <= z < ez and cand[z, y, x]:
                    want |= 1 << q
            assert five_row_mask(bits, vx, vy, vz, ey, ez) == want, (vx, vy, vz, ex, ey, ez)
