"""The line-ownership rule of the packed matrix-core binary tiles, on the CPU.

R's packed 'dist' order stores column j (the pairs (i, j), i > j) as one
run.  A tile of QB rows i0 .. i0 + QB - 1 writes, of column j, the rows

    [j + 1, i0 + QB + e_j)        if j + 1 >= i0 (the run starts in the tile)
    [i0 + e_j, i0 + QB + e_j)     otherwise
    nothing                       if j + 1 >= i0 + QB

clipped to i < n, where e_j = (-(F + base_j + i0)) mod 32 is the distance
from row i0 to the next 128-byte line start (F: the output's address in
floats).  Two properties must hold for every n, tile side and address:
every packed element is written by exactly one tile, and a 32-float line
is shared by two tiles only if it holds the end of one column's run.
"""
import numpy as np
import pytest


def packed_base(n, j):
    return j * n - j * (j + 1) // 2 - j - 1


def owners(n, QB, F):
    T = (n + QB - 1) // QB
    total = n * (n - 1) // 2
    owner = np.full(total, -1, np.int64)
    count = np.zeros(total, np.int64)
    for I in range(T):
        i0 = I * QB
        for j in range(min(i0 + QB, n)):   # columns of tiles J <= I
            if j + 1 >= i0 + QB:
                continue
            e = (-(F + packed_base(n, j) + i0)) & 31
            lo = j + 1 if j + 1 >= i0 else i0 + e
            hi = min(n, i0 + QB + e)
            assert hi <= i0 + QB + 32           # within the 32 extra rows
            if hi > lo:
                idx = packed_base(n, j) + np.arange(lo, hi)
                count[idx] += 1
                owner[idx] = I
    return owner, count


@pytest.mark.parametrize("n,QB,F", [(2, 128, 0), (33, 128, 5), (300, 128, 0), (300, 128, 17),
                                    (513, 128, 12), (257, 256, 29), (700, 256, 31), (1025, 256, 3)])
def test_every_element_once_and_lines_whole(n, QB, F):
    owner, count = owners(n, QB, F)
    assert count.min() == 1 and count.max() == 1
    lines = (F + np.arange(owner.size)) // 32
    ends = {(F + packed_base(n, j) + n - 1) // 32 for j in range(n - 1)}
    order = np.argsort(lines, kind="stable")
    ls, os_ = lines[order], owner[order]
    starts = np.flatnonzero(np.r_[True, ls[1:] != ls[:-1]])
    for a, b in zip(starts, np.r_[starts[1:], ls.size]):
        if np.unique(os_[a:b]).size > 1:
            assert ls[a] in ends
