"""bench_deflate (the drop-in parallel_deflate.h's CPU throughput table): runs
on the CPU through tests/cpp/facade_check, prints one row per level 0, 1, 3,
and level 0 stores (output a little over the input)."""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_deflate_table():
    exe = os.path.join(REPO, "tests", "cpp", "facade_check")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/facade_check was not built (run __graft_entry__.build())")
    out = subprocess.run([exe, "bench-deflate", str(1 << 20), "2"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = {}
    for line in out.stdout.splitlines():
        m = re.match(r"\s+(\d)\s+([\d.]+)\s+([\d.]+)\s+([\d.]+)\s+(\d+) us", line)
        if m:
            rows[int(m.group(1))] = (float(m.group(2)), float(m.group(3)), float(m.group(4)))
    assert sorted(rows) == [0, 1, 3], out.stdout
    assert 1.0 <= rows[0][2] < 1.01              # stored blocks
    assert rows[3][2] < 1.0 and all(r[0] > 0 and r[1] > 0 for r in rows.values())
