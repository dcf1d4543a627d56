"""bench.py contract checks that need no GPU."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_bench_help_lists_the_contract_flags():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--help"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--end-to-end"):
        assert flag in r.stdout


def test_profile_figures_parse_and_are_complete():
    """Every per-launch figure line in profiles/*.md parses, each (kernel,
    workload, precision) carries all four figures bench.py reads, and the
    default workload's dominant kernel has them."""
    import glob
    import re
    sys.path.insert(0, ROOT)
    import bench
    names = ("traffic_bytes_per_launch", "valu_insts_per_launch",
             "mfma_busy_cycles_per_launch", "profile_avg_launch_ms")
    pat = re.compile(r"^(\w+)\[(\w+),(\w+),(\w+)\] = ([0-9.]+)\s*$")
    seen = {}
    for path in glob.glob(os.path.join(ROOT, "profiles", "*.md")):
        for line in open(path):
            if "_per_launch[" in line or "profile_avg_launch_ms[" in line:
                m = pat.match(line)
                assert m, (path, line)
                name, key = m.group(1), m.group(2, 3, 4)
                assert name in names, (path, line)
                assert key + (name,) not in seen, ("duplicate", path, line)
                seen[key + (name,)] = float(m.group(5))
    keys = {k[:3] for k in seen}
    assert keys
    for key in keys:
        for name in names:
            assert key + (name,) in seen, (key, name)
            assert bench.profiled(name, *key) == seen[key + (name,)]
        assert seen[key + ("profile_avg_launch_ms",)] > 0
        assert seen[key + ("valu_insts_per_launch",)] > 0
    assert ("search_kernel", "cfg2_1m_d200", "bf16") in keys
