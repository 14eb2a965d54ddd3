"""The search drivers replayed from journals, on the CPU (no device): the EM
trajectory on hand-built timings, and the crash/hang recovery rules of the
journal (GatlasAppUtil.cpp:202-228): a key whose last state is
BUILD_IN_PROGRESS, BUILD_OK or RUN_IN_PROGRESS is bad and skipped, a trial
time of 0 (failed check) excludes the key, a truncated last record is
ignored, and a candidate with no record is attempted (and, with no device,
recorded as a failed build/launch instead of ending the search)."""
import pytest

from gatlas_amd import _capi
from gatlas_amd.interface import BUFFER, MATMUL, MATVEC, NATIVE, SAXPY, Journal, Kernel

M = N = K = 64


def _cands():
    k = Kernel(MATMUL, _capi.GAT_F32, 4, BUFFER)
    return k, k.search(M, N, K, vl=4)


def test_candidate_set_hand_derived():
    # wg in 10..9 fails 64 % (wg*4) (largest group is min(10, sqrt(1024))),
    # so only the 8x8 group remains, with bh in {4, 8} (64 % (8*bh) == 0)
    # and all 12 extra parameters: 24 candidates in (bh, xp) order.
    k, c = _cands()
    assert len(c) == 24
    assert [(p[7], p[9], p[11]) for p in c[:3]] == [(8, 4, 0), (8, 4, 1), (8, 4, 2)]
    assert {p[9] for p in c} == {4, 8}
    assert all(p[:7] == [1, 0, 64, 64, 64, 0, 0] and p[8] == 8 and p[10] == 4 for p in c)


def _write(path, k, rows):
    j = Journal(path)
    for p, vals in rows:
        for v in vals:
            assert j.takeMemo(k, p, v)


def _time(p):
    bh, xp = p[9], p[11]
    if bh == 4:
        return 500 - (100 if xp == 7 else 0)
    return 300 - (50 if xp == 3 else 0)


def test_em_converges_on_journal(tmp_path):
    # step 0 (xp 0): bh 8 (300) beats bh 4 (500); step 1 (bh 8): xp 3 (250);
    # step 0 again (xp 3): (8, 8, 3) stays best -> converged.  The global
    # optimum over this set is the same point.
    k, c = _cands()
    path = str(tmp_path / "j.txt")
    _write(path, k, [(p, [-1, -2, -3, -4, _time(p)]) for p in c])
    best = k.em_search(path, M, N, K, vl=4)
    assert best is not None and (best[7], best[9], best[11]) == (8, 8, 3)


@pytest.mark.parametrize("last_state", [-1, -2, -3])
def test_crashed_or_hung_key_is_skipped(tmp_path, last_state):
    # (8,8,3) crashed during build (-1), during argument setup (-2) or hung
    # while running (-3): it is never retried and the search continues.  The
    # other bh-8 keys tie at 300; ties go to the last index (xp 11).
    k, c = _cands()
    path = str(tmp_path / "j.txt")
    rows = []
    for p in c:
        if p[9] == 8 and p[11] == 3:
            rows.append((p, [-1, -2, -3, -4, 250, -1, -2, -3][: 5 + {-1: 1, -2: 2, -3: 3}[last_state]]))
        else:
            rows.append((p, [-1, -2, -3, -4, _time(p)]))
    _write(path, k, rows)
    j = Journal(path)
    j.loadMemo()
    bad = [p for p in c if p[9] == 8 and p[11] == 3][0]
    assert j.memoRunState(k, bad) == last_state
    best = k.em_search(path, M, N, K, vl=4)
    assert (best[7], best[9], best[11]) == (8, 8, 11)


def test_failed_check_time_zero_is_excluded(tmp_path):
    k, c = _cands()
    path = str(tmp_path / "j.txt")
    _write(path, k, [(p, [-1, -2, -3, -4, 0 if (p[9] == 8 and p[11] == 3) else _time(p)]) for p in c])
    best = k.em_search(path, M, N, K, vl=4)
    assert (best[9], best[11]) == (8, 11)


def test_truncated_last_record_is_ignored(tmp_path):
    k, c = _cands()
    path = str(tmp_path / "j.txt")
    _write(path, k, [(p, [-1, -2, -3, -4, _time(p)]) for p in c])
    with open(path, "a") as f:
        f.write("matmulbufferfloat4_1_0_64_64_64_0_0_8_8_8_4_3_")  # killed mid-write, no value
    j = Journal(path)
    assert j.loadMemo()
    assert j.memoGood() == len(c)
    best = k.em_search(path, M, N, K, vl=4)
    assert (best[9], best[11]) == (8, 3)


def test_missing_candidate_without_device_is_recorded_bad(tmp_path):
    # One candidate has no record: the search attempts it, the attempt fails
    # (no device here), the journal records its BUILD states and the search
    # still finishes on the journaled candidates.
    k, c = _cands()
    path = str(tmp_path / "j.txt")
    missing = [p for p in c if p[9] == 4 and p[11] == 0][0]
    _write(path, k, [(p, [-1, -2, -3, -4, _time(p)]) for p in c if p != missing])
    best = k.em_search(path, M, N, K, vl=4)
    assert (best[9], best[11]) == (8, 3)
    j = Journal(path)
    j.loadMemo()
    assert j.memoRunState(k, missing) in (-1, -2)


def test_replay_is_deterministic_and_journal_unchanged(tmp_path):
    k, c = _cands()
    path = str(tmp_path / "j.txt")
    _write(path, k, [(p, [-1, -2, -3, -4, _time(p)]) for p in c])
    before = open(path).read()
    b1 = k.em_search(path, M, N, K, vl=4)
    b2 = k.em_search(path, M, N, K, vl=4)
    assert b1 == b2
    assert open(path).read() == before


def test_matvec_and_saxpy_em_replay(tmp_path):
    # matvecbuffer: candidates for M = N = 1024, VL 4; fastest = (gs 64, bh 16, xp 1)
    kv = Kernel(MATVEC, _capi.GAT_F32, 4, BUFFER)
    cv = kv.search(1024, 1024, vl=4, max_block_height=64, max_group_size=256)
    assert cv
    target = [p for p in cv if p[5] == 64 and p[6] == 16 and p[8] == 1]
    assert target
    path = str(tmp_path / "mv.txt")
    _write(path, kv, [(p, [-1, -2, -3, -4, 100 if p == target[0] else 200 + p[6]]) for p in cv])
    best = kv.em_search(path, 1024, 1024, vl=4, max_block_height=64, max_group_size=256)
    assert best == target[0]
    # saxpynative: 1-D, n = 2^20 floats as M = 2^18 vectors of 4
    ks = Kernel(SAXPY, _capi.GAT_F32, 4, NATIVE)
    cs = ks.search(1 << 18, 4)
    assert len(cs) == 5 * 4 * 2
    want = [p for p in cs if p[3] == 256 and p[5] == 4 and p[8] == 1][0]
    path = str(tmp_path / "sx.txt")
    # step 0 (bh 4, xp 0 fixed) must pick the 256 group; step 1 then finds want
    _write(path, ks, [(p, [-1, -2, -3, -4, 50 if p == want else 90 + p[5] + (0 if p[3] == 256 else 20)])
                      for p in cs])
    best = ks.em_search(path, 1 << 18, 4, vl=4)
    assert best == want


def test_library_search_keeps_stdout_clean(tmp_path, capfd):
    # the command-line tools print every trial; the library entry point
    # prints nothing, so a caller's own output (one JSON line) stays intact
    k, c = _cands()
    path = str(tmp_path / "j.txt")
    _write(path, k, [(p, [-1, -2, -3, -4, _time(p)]) for p in c])
    capfd.readouterr()
    best = k.em_search(path, M, N, K, vl=4)
    out = capfd.readouterr().out
    assert best is not None
    assert out == ""
