"""The schedule cache of msosm_amd.tune (host logic only; the measurement itself runs on the GPU)."""
import json

from msosm_amd import tune


class FakePlan:
    def __init__(self):
        self.M, self.D, self.B, self.in_format, self.out_flags = 1 << 15, 512, 32, 1, 0
        self.applied, self.tuned = None, 0

    def autotune(self, input_dev, pushes):
        self.tuned += 1
        self.applied = (256, 4)
        return 256, 4, 5.4

    def set_schedule(self, w, o):
        self.applied = (w, o)


def test_measures_once_then_reuses(tmp_path):
    cache = str(tmp_path / "sched.json")
    p = FakePlan()
    a = tune.tuned_schedule(p, 0, "MI355X", nbins=1024, cache=cache)
    assert not a["cached"] and p.tuned == 1 and p.applied == (256, 4)
    q = FakePlan()
    b = tune.tuned_schedule(q, 0, "MI355X", nbins=1024, cache=cache)
    assert b["cached"] and q.tuned == 0 and q.applied == (256, 4)
    with open(cache) as fh:
        table = json.load(fh)
    assert list(table) == [tune.plan_key("MI355X", p, 1024)]
    # another shape or device is measured separately
    r = FakePlan()
    r.D = 64
    assert not tune.tuned_schedule(r, 0, "MI355X", nbins=1024, cache=cache)["cached"]
    assert not tune.tuned_schedule(FakePlan(), 0, "other", nbins=1024, cache=cache)["cached"]


def test_unreadable_cache_is_ignored(tmp_path):
    cache = tmp_path / "sched.json"
    cache.write_text("{not json")
    p = FakePlan()
    assert not tune.tuned_schedule(p, 0, "MI355X", cache=str(cache))["cached"]
