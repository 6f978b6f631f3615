"""Host side of the autotuner: candidate choices stay within the kernels'
limits, the tuning table takes and clears entries through the C ABI, and
the cache keeps entries per device model."""
import json

from qimpy_amd import _lib, autotune


def test_dot_candidates_within_limits():
    for batch, M, N, G in ((1, 16, 16, 300), (16, 64, 32, 3000), (7, 144, 72, 9000),
                           (2, 288, 144, 20000), (1, 1, 1, 1)):
        cands = autotune.dot_candidates(batch, M, N, G)
        assert cands and len(set(cands)) == len(cands)
        M16, N16 = -(-M // 16), -(-N // 16)
        for kern, bm, bn, ch in cands:
            assert kern in (0, 1) and ch >= 1
            assert ch <= max(1, -(-G // 256))
            if kern == 0:
                assert 1 <= bm <= M16 and 1 <= bn <= N16
                assert bm + bn <= 14 and bm * bn <= 48


def test_table_set_clear_and_buckets():
    lib = _lib.lib()
    autotune.clear()
    assert int(lib.qmx_tune_count()) == 0
    autotune._set(autotune.TUNE_DOT, (0, 16, 64, 32, 3444), (0, 4, 2, 8))
    autotune._set(autotune.TUNE_ROTATE, (0, 0, 0, 72, 144, 3444), (2,))
    assert int(lib.qmx_tune_count()) == 2
    autotune.clear()
    assert int(lib.qmx_tune_count()) == 0
    b = [int(lib.qmx_tune_bucket(g)) for g in (1, 64, 65, 1000, 1024, 1025)]
    assert b[0] == 64 and b[1] == 64 and b == sorted(b) and b[4] == 1024 and b[5] > 1024
    assert [int(lib.qmx_tune_batch_bucket(x)) for x in (1, 3, 16, 17)] == [1, 4, 16, 32]


def test_cache_roundtrip_and_device(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "device_name", lambda: "test_device")
    entries = [{"op": 1, "key": [0, 16, 64, 32, 3444], "choice": [0, 4, 2, 8]},
               {"op": 2, "key": [0, 0, 0, 72, 144, 3444], "choice": None}]
    p = tmp_path / "c.json"
    autotune.save(p, entries)
    assert json.loads(p.read_text())["device"] == "test_device"
    autotune.clear()
    assert autotune.load(p) == 1  # entries without a choice keep the default
    monkeypatch.setattr(autotune, "device_name", lambda: "another_device")
    autotune.clear()
    assert autotune.load(p) == 0
    assert autotune.load(tmp_path / "missing.json") == 0
    autotune.clear()
