"""Oracle pins for the pencil exchange copies memcpy_v1_h1 / memcpy_v1_h2
(src/transpose_cuda.cu:97-197).  The reference holds no vectors for them;
they are pinned by what the layouts fix, with numpy slicing independent of
the oracle's loops:
  * pack: the block for rank p is the column slice [start_p, start_p + n_p)
    of every row of every batch, in (h, i) order;
  * unpack: a buffer built by concatenating, per rank, the row slices of a
    known array reproduces that array;
  * a P-rank pencil transpose (pack on every rank, exchange blocks, unpack)
    leaves rank p with the global columns of its range, ragged splits and
    empty ranks included."""
import numpy as np
import pytest

from conftest import rng


def _split(n, P, empty=None):
    sizes = [n // P + (1 if q < n % P else 0) for q in range(P)]
    if empty is not None and P > 1:  # move one rank's share to its neighbour
        sizes[(empty + 1) % P] += sizes[empty]
        sizes[empty] = 0
    starts = list(np.cumsum([0] + sizes[:-1]))
    return sizes, starts


@pytest.mark.parametrize("howmany,n0,N1,t,P", [(1, 4, 10, 3, 3), (3, 5, 7, 2, 4), (2, 3, 9, 1, 2)])
def test_pack_blocks_are_column_slices(orc, howmany, n0, N1, t, P):
    a = rng(61).standard_normal((howmany, n0, N1, t))
    sizes, starts = _split(N1, P)
    out = orc.pencil_pack(P, howmany, n0, t, sizes, a, n0 * N1 * t, N1, starts)
    off = 0
    for p in range(P):
        blk = a[:, :, starts[p]:starts[p] + sizes[p], :].ravel()
        assert np.array_equal(out[off:off + blk.size], blk)
        off += blk.size
    assert off == out.size == a.size


@pytest.mark.parametrize("howmany,N0,n1,t,P", [(1, 10, 4, 3, 3), (3, 7, 5, 2, 4), (2, 9, 3, 1, 2)])
def test_unpack_inverts_row_concatenation(orc, howmany, N0, n1, t, P):
    want = rng(62).standard_normal((howmany, N0, n1, t))
    sizes, starts = _split(N0, P)
    recv = np.concatenate([want[:, starts[p]:starts[p] + sizes[p]].ravel() for p in range(P)])
    got = orc.pencil_unpack(P, howmany, starts, sizes, np.zeros(want.size), N0 * n1 * t, n1, t,
                            recv)
    assert np.array_equal(got, want.ravel())


def simulate_pencil_transpose(pack, unpack, G, P, empty=None):
    """P virtual ranks own row ranges of G[h][N0][N1][t]; returns what each
    rank holds after pack -> all-to-all -> unpack."""
    howmany, N0, N1, t = G.shape
    rows, rstart = _split(N0, P, empty)
    cols, cstart = _split(N1, P, empty)
    sends = []
    for r in range(P):
        loc = np.ascontiguousarray(G[:, rstart[r]:rstart[r] + rows[r]])
        sends.append(pack(P, howmany, rows[r], t, cols, loc, rows[r] * N1 * t, N1, cstart))
    out = []
    for p in range(P):
        parts = []
        for r in range(P):
            off = t * howmany * rows[r] * sum(cols[:p])
            parts.append(sends[r][off:off + t * howmany * rows[r] * cols[p]])
        recv = np.concatenate(parts)
        res = unpack(P, howmany, rstart, rows, np.zeros(howmany * N0 * cols[p] * t), N0 * cols[p] * t,
                     cols[p], t, recv)
        out.append((res, cstart[p], cols[p]))
    return out


@pytest.mark.parametrize("shape,P,empty", [((2, 12, 10, 3), 3, None), ((1, 9, 14, 2), 4, None),
                                           ((3, 8, 8, 1), 4, 1), ((2, 5, 6, 4), 1, None)])
def test_pencil_transpose_gives_column_ranges(orc, shape, P, empty):
    G = rng(63).standard_normal(shape)
    for res, cs, cn in simulate_pencil_transpose(orc.pencil_pack, orc.pencil_unpack, G, P, empty):
        assert np.array_equal(res, G[:, :, cs:cs + cn, :].ravel())
