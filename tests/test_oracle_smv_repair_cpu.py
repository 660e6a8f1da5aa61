"""The oracle's smv zero-row repair (createMeasurements_smv.cu:263-383) against the
unrepaired matrix of the same draws: no empty row is left, every column keeps p
distinct rows, exactly one entry moves into each formerly empty row, each moved
entry keeps its value and leaves a row that still holds an entry, the donors are
the earliest repeated occurrences in entry order, and the empty rows receive
them in ascending order.  Matrices with no empty row are left unchanged."""
import numpy as np
import pytest

SEED = 20260101


def _columns(rows, vals, n, p):
    R = rows.reshape(n, p)
    V = vals.reshape(n, p)
    return [dict(zip(R[j].tolist(), V[j].tolist())) for j in range(n)]


@pytest.mark.parametrize("m,n,p,ens,inst", [(1000, 1024, 2, 2, 3), (60, 64, 1, 2, 1),
                                            (1024, 1024, 3, 1, 5), (900, 1024, 2, 3, 2),
                                            (512, 600, 1, 2, 7)])
def test_repair(oracle, m, n, p, ens, inst):
    raw_r, raw_v, e0 = oracle.smv_matrix(m, n, p, ens, SEED, inst, repair=False)
    rep_r, rep_v, e1 = oracle.smv_matrix(m, n, p, ens, SEED, inst, repair=True)
    assert e0 == e1
    raw_cnt = np.bincount(raw_r, minlength=m)
    empty = np.flatnonzero(raw_cnt == 0)
    assert empty.size == e0 > 0  # the cases are chosen to have empty rows
    # no empty row afterwards; columns: p distinct rows in ascending order
    assert np.all(np.bincount(rep_r, minlength=m) > 0)
    assert np.all(np.diff(rep_r.reshape(n, p), axis=1) > 0)
    # what moved: per column, (row, value) pairs removed and added
    raw_c, rep_c = _columns(raw_r, raw_v, n, p), _columns(rep_r, rep_v, n, p)
    moves = []  # (column, old row, new row)
    for j in range(n):
        gone = sorted(set(raw_c[j]) - set(rep_c[j]))
        came = sorted(set(rep_c[j]) - set(raw_c[j]))
        assert len(gone) == len(came)
        for r in set(raw_c[j]) & set(rep_c[j]):
            assert raw_c[j][r] == rep_c[j][r]
        for g, c in zip(gone, came):
            moves.append((j, g, c))
    assert len(moves) == empty.size
    assert sorted(c for _, _, c in moves) == empty.tolist()  # one entry per empty row
    # each donor row still holds an entry and held at least two before
    rep_cnt = np.bincount(rep_r, minlength=m)
    for j, g, c in moves:
        assert raw_cnt[g] >= 2 and rep_cnt[g] >= 1
        assert raw_c[j][g] == rep_c[j][c]  # the value moves with the entry
    # donors: the earliest repeated occurrences in entry order, matched in order to
    # the empty rows in ascending order
    seen = np.zeros(m, bool)
    repeated = []
    for e, r in enumerate(raw_r.tolist()):
        if seen[r]:
            repeated.append((e // p, r))
        else:
            seen[r] = True
    donors = repeated[:empty.size]
    got = sorted((j, g) for j, g, _ in moves)
    assert got == sorted(donors)
    by_donor = {(j, g): c for j, g, c in moves}
    assert [by_donor[d] for d in donors] == empty.tolist()


def test_no_empty_rows_no_change(oracle):
    raw_r, raw_v, e0 = oracle.smv_matrix(256, 2048, 7, 2, SEED, 1, repair=False)
    rep_r, rep_v, _ = oracle.smv_matrix(256, 2048, 7, 2, SEED, 1, repair=True)
    assert e0 == 0
    np.testing.assert_array_equal(raw_r, rep_r)
    np.testing.assert_array_equal(raw_v, rep_v)
