"""CPU checks of the work-list schedule: every column of every stamp's wrapped grid and every
output row is covered by exactly one item, tiles respect the LDS limit, and the largest FFT
sizes come first."""
import numpy as np
import pytest

from galsim_amd import _lib
from galsim_amd.batch import STAGE_LDS, BatchPlan, lds_line, work_lists


def _recs(spec):
    recs = np.zeros(len(spec), dtype=_lib.STAMP_DTYPE)
    for k, (N, ny) in enumerate(spec):
        recs[k]['N'] = N
        recs[k]['ny'] = ny
    groups = []
    for N in np.unique(recs['N']):
        idx = np.nonzero(recs['N'] == N)[0]
        groups.append((int(N), int(idx[0]), len(idx), int(recs['ny'][idx].max())))
    return recs, groups


@pytest.mark.parametrize("cols,rows,chunks", [
    ({}, {}, None), ({128: 3, 768: 1}, {128: 5, 384: 7}, None),
    ({}, {}, {128: 1, 192: 1, 384: 1, 768: 1}),
    ({128: 3}, {}, {128: 4, 192: 1000, 768: 7})])
def test_work_lists_cover_every_column_and_row_once(cols, rows, chunks):
    recs, groups = _recs([(128, 48), (128, 128), (192, 77), (384, 200), (768, 9)])
    recs['Nk'] = 2 * recs['N']
    fold, rw, lds_f, lds_r = work_lists(recs, groups, cols, rows, 8, chunks=chunks)
    for s in range(len(recs)):
        N, ny = int(recs['N'][s]), int(recs['ny'][s])
        f = fold[fold[:, 0] == s]
        seen = np.zeros(N // 2, int)            # kx = N/2 travels with column 0
        staged = (2 * int(f[0, 3]) * N + N + 3 * N + 2) * 8 <= STAGE_LDS
        k = (chunks or {}).get(N) if staged else 1
        for _, c0, cnt, per in f:
            assert cnt >= 1 and per >= 1
            if k is not None:
                assert cnt <= k * per                     # at most k chunks per item
            seen[c0:c0 + cnt] += 1
        assert np.all(seen == 1)
        # items split the columns evenly: every item but the last spans a whole number of chunks
        assert all(cnt % per == 0 for _, _, cnt, per in f[:-1])
        r = rw[rw[:, 0] == s]
        seen = np.zeros(ny, int)
        for _, r0, cnt, per in r:
            assert cnt >= 1 and per >= 1
            seen[r0:r0 + cnt] += 1
        assert np.all(seen == 1)
        r = r[np.argsort(r[:, 1])]
        assert all(cnt % per == 0 for _, _, cnt, per in r[:-1])
    Ns = recs['N'][fold[:, 0]]
    assert np.all(np.diff(Ns) <= 0)                       # largest FFT sizes first
    # the row items run in the reverse order of the fold's: the stamps folded last come first
    assert np.array_equal(np.unique(rw[:, 0], return_index=True)[0], np.unique(fold[:, 0]))
    first_row = {}
    for k, s in enumerate(rw[:, 0]):
        first_row.setdefault(int(s), k)
    last_fold = {}
    for k, s in enumerate(fold[:, 0]):
        last_fold[int(s)] = k
    order_rows = sorted(first_row, key=first_row.get)
    assert order_rows == sorted(last_fold, key=last_fold.get, reverse=True)
    assert np.all(np.diff(recs['N'][rw[:, 0]]) >= 0)
    need = []
    for (_, _, _, per), N in zip(fold, recs['N'][fold[:, 0]]):
        # transform lines are padded by one slot per 32 elements in LDS
        staged = (2 * per * lds_line(N) + N + 3 * N + 2) * 8
        need.append(staged if staged <= STAGE_LDS else 2 * per * lds_line(N) * 8)
    assert lds_f == max(need)
    assert lds_r == max((2 * per * lds_line(int(recs['N'][s]) // 2) + int(recs['N'][s])) * 8
                        for s, _, _, per in rw)


def test_work_lists_reject_tiles_beyond_lds():
    recs, groups = _recs([(1024, 10)])
    with pytest.raises(_lib.GalSimAMDError):
        work_lists(recs, groups, {1024: 8}, {}, 16)
    assert BatchPlan.default_cols(128) == 8 and BatchPlan.default_cols(2048) == 1
    assert BatchPlan.default_rows(128) == 16 and BatchPlan.default_rows(4096) == 1


def test_float_reciprocal_index_division_is_exact():
    """The kernels' index arithmetic divides by FFT sizes and tile widths with
    int((t + 0.5) * (1.f / d)) in FP32; it equals t // d over the ranges they use."""
    t = np.arange(0, 20000, dtype=np.int64)
    for d in list(range(1, 1100)) + [1536, 2048, 3072, 4096, 6144, 8192]:
        inv = np.float32(1.0) / np.float32(d)
        q = ((t.astype(np.float32) + np.float32(0.5)) * inv).astype(np.int64)
        assert np.array_equal(q, t // d), d


def test_program_shape_detection():
    """A batch whose stamps all lower to one specialised shape gets its id; a mixed batch, or a
    shape without a specialised kernel, gets 0 (the interpreter)."""
    from galsim_amd.batch import _program_shape
    L = _lib
    ops = np.zeros(8, dtype=L.OP_DTYPE)
    for k, (c, a) in enumerate([(1, 0), (1, 1), (3, 0), (1, 2), (2, 0), (1, 0), (1, 1), (2, 0)]):
        ops[k]['code'], ops[k]['arg'] = c, a
    leaves = np.zeros(7, dtype=L.LEAF_DTYPE)
    for k, t in enumerate([L.LEAF_EXPONENTIAL, L.LEAF_SERSIC, L.LEAF_KOLMOGOROV, L.LEAF_BOX,
                           L.LEAF_EXPONENTIAL, L.LEAF_MOFFAT, L.LEAF_BOX]):
        leaves[k]['type'] = t
    recs = np.zeros(2, dtype=L.STAMP_DTYPE)
    recs[0]['op_begin'], recs[0]['op_count'], recs[0]['leaf_begin'] = 0, 5, 0
    recs[1]['op_begin'], recs[1]['op_count'], recs[1]['leaf_begin'] = 5, 3, 4
    assert _program_shape(recs[:1], ops, leaves) == 1
    assert _program_shape(recs[1:], ops, leaves) == 2
    assert _program_shape(recs, ops, leaves) == 0
    leaves[5]['type'] = L.LEAF_GAUSSIAN
    assert _program_shape(recs[1:], ops, leaves) == 0
