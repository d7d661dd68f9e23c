"""CPU tests of the autotuner's search and cache (no GPU: a fake timer)."""
import json

from dwba_amd import autotune as AT


def test_picks_fastest_and_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("DWBA_AUTOTUNE_CACHE", str(tmp_path / "c.json"))
    cost = {(0, 8): 3.0, (16, 8): 1.5, (32, 4): 0.9, (64, 4): 2.0}
    cands = [{"seg": s, "lb": b} for (s, b) in cost]
    calls = []

    def run(c):
        calls.append((c["seg"], c["lb"]))

    def timer(fn, steps):
        fn()
        return cost[calls[-1]]

    t = AT.Autotuner()
    ch = t.tune("k|dev|n=1", cands, run, timer=timer, steps=3, warmup=1)
    assert ch.params == {"seg": 32, "lb": 4} and ch.ms == 0.9 and not ch.cached
    assert len(ch.timings) == 4
    data = json.load(open(tmp_path / "c.json"))
    assert data["k|dev|n=1"]["params"] == {"seg": 32, "lb": 4}
    n = len(calls)
    ch2 = AT.Autotuner().tune("k|dev|n=1", cands, run, timer=timer)
    assert ch2.cached and ch2.params == ch.params and len(calls) == n


def test_disabled_never_hits_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("DWBA_AUTOTUNE_CACHE", str(tmp_path / "c.json"))
    t = AT.Autotuner()
    t.store("a", AT.Choice(params={"seg": 1}, ms=1.0))
    assert t.lookup("a").params == {"seg": 1}
    monkeypatch.setenv("DWBA_AUTOTUNE", "off")
    assert AT.Autotuner().lookup("a") is None


def test_signature_buckets():
    assert AT.bucket(9500) == 8192 and AT.bucket(12000) == 16384 and AT.bucket(0) == 0
    k1 = AT.signature("semisep_f64", "MI355X", jobs=AT.bucket(25600), n=AT.bucket(9500), L=30)
    k2 = AT.signature("semisep_f64", "MI355X", L=30, n=AT.bucket(9000), jobs=AT.bucket(26000))
    assert k1 == k2


def test_unwritable_cache_is_harmless(tmp_path, monkeypatch):
    monkeypatch.setenv("DWBA_AUTOTUNE_CACHE", str(tmp_path / "nodir" / "x" / "c.json"))
    (tmp_path / "nodir").write_text("a file, not a directory")
    t = AT.Autotuner()
    ch = t.tune("k", [{"seg": 1}], lambda c: None, timer=lambda fn, s: 1.0)
    assert ch.params == {"seg": 1}


def test_autotune_off_keeps_defaults(monkeypatch):
    """DWBA_AUTOTUNE=off: lookups miss (and ScanBatch.tune() returns None
    before timing anything, keeping the library defaults)."""
    from dwba_amd import autotune as AT
    monkeypatch.setenv("DWBA_AUTOTUNE", "off")
    assert not AT.enabled()
    t = AT.Autotuner(persist=False)
    t.store("k", AT.Choice(params={"seg": 1}, ms=1.0, timings=[]))
    assert t.lookup("k") is None


def test_autotune_margin_prefers_earlier_candidate(tmp_path):
    """With a margin, the first candidate within it of the fastest wins."""
    from dwba_amd import autotune as AT
    t = AT.Autotuner(path=str(tmp_path / "t.json"))
    times = {0: 10.0, 1: 9.9, 2: 9.0}
    timer = lambda fn, steps: (fn(), times[cur[0]])[1]
    cur = [0]
    run = lambda c: cur.__setitem__(0, c["i"])
    c = t.tune("k1", [{"i": 0}, {"i": 1}], run, timer=timer, warmup=0, margin=0.02)
    assert c.params == {"i": 0}            # 10.0 is within 2 % of 9.9
    c = t.tune("k2", [{"i": 0}, {"i": 2}], run, timer=timer, warmup=0, margin=0.02)
    assert c.params == {"i": 2}            # 9.0 beats 10.0 by more than 2 %
    c = t.tune("k3", [{"i": 0}, {"i": 1}], run, timer=timer, warmup=0)
    assert c.params == {"i": 1}            # no margin: the fastest
