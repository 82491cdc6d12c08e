"""CPU tests of bench.py's contract and of the byte model behind its roofline
figures (no GPU: the module is imported, not run, except for the --gpus check,
which exits before anything touches a device).

  * the metric is the headline metric named for the benchmark (the one line of
    JSON must carry it verbatim);
  * `--gpus N` with N > 1 exits with status 2 and says the solver runs on one GPU;
  * the per-kernel plane counts add up to the per-site figures the documentation
    derives independently: 182 / 150 / 150 reals with U in nine planes
    (correlated position-space noise / momentum-space noise / white position-space
    noise) and 164 / 132 / 132 with U as its rows 0 and 1 (six planes).
"""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, REPO)
    import bench as b
    return b


def test_metric_is_the_benchmark_metric(bench):
    with open(os.path.join(REPO, "BASELINE.json")) as f:
        assert bench.METRIC == json.load(f)["metric"]


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("kind,u,reals", [("correlated", 9, 182), ("kspace", 9, 150), ("white", 9, 150),
                                          ("correlated", 6, 164), ("kspace", 6, 132), ("white", 6, 132)])
def test_reals_per_site(bench, kind, u, reals):
    assert bench.reals_of(kind, u) == reals


def test_plane_counts_per_kernel(bench):
    # momentum-space noise, U in six planes: column A writes A (4) and the noise's row
    # spectra (8), column B reads 8 and writes 4, U xi~ U^dagger reads xi~ (8) and U and
    # writes V (8), the update reads A, B (4 + 4) and U and writes U
    t = bench.planes_of("kspace", 6)
    assert t == {"row_noise_fwd": 0, "col_kmult": (12, 12), "row_uxiu_fwd": 6 + 16, "row_update": 8 + 12}
    # position-space correlated noise: the noise pass writes 8 planes and column A
    # also reads them
    t = bench.planes_of("correlated", 6)
    assert t["row_noise_fwd"] == 8 and t["col_kmult"] == (20, 12)
