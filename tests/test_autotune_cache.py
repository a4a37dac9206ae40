"""Autotuner cache (host side): keys, atomic store, lookup."""

import json
import threading


def test_cache_store_and_lookup(tmp_path):
    from DoseCUDA import autotune as A
    path = str(tmp_path / "sub" / "tune.json")
    sig = {"arch": "gfx950", "beams": 3, "bricks": 33, "spots": 27, "spots_per_layer": 13, "cull": True,
           "wet": "reach"}
    assert A.lookup(sig, path) is None
    t = {"adj_cs": 32, "adj_wg_target": 65536, "adj_min_bricks": 32, "adj_max_slabs": 0, "fwd_blocks": 2048,
         "ray_blocks": 0}
    A.store(sig, t, {"adj_cs": {"32": 1.0, "64": 1.2}}, path)
    assert A.lookup(sig, path) == t
    other = dict(sig, beams=4)
    assert A.lookup(other, path) is None
    data = json.load(open(path))
    assert len(data) == 1 and list(data.values())[0]["signature"] == sig


def test_cache_concurrent_writers(tmp_path):
    from DoseCUDA import autotune as A
    path = str(tmp_path / "tune.json")
    base = {"arch": "gfx950", "beams": 1, "bricks": 1, "spots": 1, "spots_per_layer": 1, "cull": True,
            "wet": "full"}

    def writer(i):
        A.store(dict(base, beams=i), {"adj_cs": i}, {}, path)

    th = [threading.Thread(target=writer, args=(i,)) for i in range(16)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    data = json.load(open(path))
    assert len(data) == 16
    for i in range(16):
        assert A.lookup(dict(base, beams=i), path) == {"adj_cs": i}


def test_bucket_is_monotone():
    from DoseCUDA.autotune import _bucket
    vals = [_bucket(x) for x in (1, 2, 3, 10, 100, 1e4, 1e6, 1e8)]
    assert vals == sorted(vals)
    assert _bucket(1000) == _bucket(1001)
