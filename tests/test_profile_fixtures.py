"""The HBM traffic that bench.py reports per launch (the roofline's `traffic`)
comes from small JSON fixtures under profiles/: rocprofv3 FETCH_SIZE and
WRITE_SIZE medians (KiB per dispatch) with the copy-kernel calibration
factors of the same runs.  Each fixture names the committed summary it was
read from; that summary must hold the same counter values and factors, so a
refreshed profile cannot leave bench.py reporting stale traffic, and every
fixture bench.py asks for must exist."""
import glob
import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURES = sorted(glob.glob(os.path.join(ROOT, "profiles", "*traffic*.json")))


def _load(path):
    with open(path) as f:
        return json.load(f)


def test_bench_fixtures_exist():
    with open(os.path.join(ROOT, "bench.py")) as f:
        names = re.findall(r'measured_traffic\("([^"]+)"\)', f.read())
    assert names, "bench.py reads no traffic fixture"
    for name in names:
        assert os.path.exists(os.path.join(ROOT, "profiles", name)), name


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_fixture_matches_its_summary(path):
    t = _load(path)
    src = os.path.join(ROOT, t["source"])
    assert os.path.exists(src), t["source"]
    with open(src) as f:
        text = f.read()
    for key in ("fetch_kib", "write_kib"):
        assert t[key] > 0
        assert f"{t[key]:,.1f}" in text, (key, t[key], t["source"])
    assert f"x {t['fetch_factor']:.1f}" in text
    assert f"x {t['write_factor']:.1f}" in text


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_bench_scales_the_counters(path):
    import bench
    t = _load(path)
    got = bench.measured_traffic(os.path.basename(path))
    # KiB x calibration factor, in bytes
    want = int((t["fetch_kib"] * t["fetch_factor"] + t["write_kib"] * t["write_factor"]) * 1024)
    assert got == want
    assert got > t["write_kib"] * 1024  # reads and writes both counted
