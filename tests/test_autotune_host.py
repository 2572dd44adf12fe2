"""The autotuner's host logic on the CPU, with stand-in engines whose speed depends on the tile
setting: it picks the fastest candidate per bucket, reuses the choice for the same bucket, keeps
separate buckets apart and persists its table."""
import json

import pytest

from micrographia_amd import autotune


class FakeEngine:
    speed = {}  # (variant, block) -> seconds per step
    made = 0

    def __init__(self):
        FakeEngine.made += 1
        self.setting = None

    def tune(self, variant, block):
        self.setting = (variant, block)

    def step(self, k):
        import time

        time.sleep(FakeEngine.speed[self.setting] * k)

    def close(self):
        pass


@pytest.fixture
def fake():
    FakeEngine.made = 0
    FakeEngine.speed = {(1, 4096): 4e-4, (1, 2048): 4e-4, (1, 1024): 2e-4, (1, 512): 1e-4, (1, 256): 3e-4, (0, 256): 5e-4}
    return FakeEngine


def test_engine_tuner_picks_the_fastest_and_caches_per_bucket(fake, tmp_path):
    path = str(tmp_path / "t.json")
    t = autotune.EngineTuner(path=path, steps=5, warmup=1)
    assert t.choose(3, 5000, 1, fake) == (1, 512)
    made = fake.made
    assert t.choose(3, 6000, 1, fake) == (1, 512)   # same bucket: floor(log2 n) = 12
    assert fake.made == made
    fake.speed[(0, 256)] = 1e-5
    assert t.choose(3, 9000, 1, fake) == (0, 256)   # bucket 13 is timed afresh
    assert t.choose(2, 5000, 1, fake) == (0, 256)
    saved = json.load(open(path))["engine"]
    assert saved["3d/12/one"] == [1, 512] and saved["3d/13/one"] == [0, 256]
    again = autotune.EngineTuner(path=path)
    assert again.choose(3, 5000, 1, fake) == (1, 512) and fake.made == made + 2 * len(autotune.ENGINE_CANDIDATES)


def test_size_classes():
    assert [autotune.size_class(n) for n in (1, 2, 3, 255, 256, 65535, 65536)] == [0, 1, 1, 7, 8, 15, 16]
    assert autotune.BatchTuner.bucket(3, 300) == "3d/8"


def test_batch_tuner_installs_its_table():
    calls = []

    class E:
        def set_tuning(self, *a):
            calls.append(a)

    t = autotune.BatchTuner()
    t.table = {"3d/12": [0, 1.25], "2d/9": [1, 1.5, 1024]}
    t.install(E())
    assert sorted(calls) == [(2, 9, 1, 1.5, 1024), (3, 12, 0, 1.25, 256)]
