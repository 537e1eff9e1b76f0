"""The batched engine's autotuner: class grid and JSON cache handling (CPU only; the tuner's
measurements run on the GPU in bench.py)."""
import json

from empanada_amd import native, tune


def test_classes_cover_the_batched_workload():
    from empanada_amd import gen
    seen = set()
    for i in range(4096, 4096 + 64):
        p = gen.batched_set_params(i)
        k = tune.class_of(p["L"], p["n"])
        assert k is not None
        seen.add(k)
    assert len(seen) >= 6
    assert tune.class_of(256, 10) is None and tune.class_of(100, (1 << 23) + 1) is None


def test_cache_is_reused_only_when_it_fits(tmp_path):
    path = str(tmp_path / "tuning.json")
    c = {"format": tune.FORMAT, "library": native.version(), "grid": {"read_len_max": list(tune.READ_LEN_MAX), "n_max": list(tune.N_MAX)},
         "bucket_cap": 1024, "classes": [dict(read_len_max=100, n_max=32768, fine_load=512, coarse_load=16384)]}
    with open(path, "w") as f:
        json.dump(c, f)
    assert tune.load(path)["bucket_cap"] == 1024
    got, cached = tune.load_or_tune(path, lambda: (_ for _ in ()).throw(AssertionError("measured")))
    assert cached and got["classes"] == c["classes"]
    c["library"] = "another build"
    with open(path, "w") as f:
        json.dump(c, f)
    assert tune.load(path) is None
    with open(path, "w") as f:
        f.write("{not json")
    assert tune.load(path) is None
    assert tune.load(str(tmp_path / "missing.json")) is None
