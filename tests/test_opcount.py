"""The op-counting build of the oracle (oracle/opcount.cpp) runs and its
WENO5/ADER3 counts agree with the flop figures the benchmark reports
(bench.FLOPS_PER_CELL_SWEEP / FLOPS_PER_CELL_STEP, counted at 256x256x128;
a smaller grid differs only by its share of the n + 1 faces per line)."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "opcount"], check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "opcount")] + [str(a) for a in args], check=True,
                         capture_output=True, text=True).stdout
    return out


def test_opcount_matches_bench_figures():
    sys.path.insert(0, ROOT)
    import bench
    out = _run([64, 64, 32, 5, 3])
    rows = {m.group(1): [float(v) for v in m.group(2).split()]
            for m in re.finditer(r"^\s+(x|y|z|mean sweep)\s+([-0-9. ]+)$", out, re.M)}
    for d, name in enumerate("xyz"):
        add, mul, div, sq, pw, other, flops = rows[name]
        assert abs(add + mul + div + sq + pw - flops) < 0.2
        assert other == 0.0
        # one sqrt and (with ngll = 3) five pow per cell and face
        assert 1.0 <= sq <= 1.04 and 5.0 <= pw <= 5.1
        assert abs(flops - bench.FLOPS_PER_CELL_SWEEP[d]) <= 0.01 * bench.FLOPS_PER_CELL_SWEEP[d]
    step = float(re.search(r"step .*: ([0-9.]+) flops per cell", out).group(1))
    assert abs(step - bench.FLOPS_PER_CELL_STEP) <= 0.01 * bench.FLOPS_PER_CELL_STEP
