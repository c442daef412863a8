"""bench.py's argument handling accepts the exact command line the benchmark
is run with (`bench.py --gpus 1 --steps 20 --warmup 5`) and plans a run of
whole exchange rounds from it; anything it cannot run exits with status 2
and a message.  CPU only: nothing here touches a GPU."""
import os
import subprocess
import sys

import pytest

from conftest import REPO

sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_driver_argv_plans_rounds():
    args = bench.parse_args(["--gpus", "1", "--steps", "20", "--warmup", "5"])
    pl = bench.plan(args)
    assert pl["rounds"] == 20 and pl["warm_rounds"] == 5
    assert pl["timed_mc_steps_per_replica"] == 20 * args.steps_per_launch
    assert pl["record_steps"] == args.cpu_steps  # the CPU leg's length does not follow --warmup


def test_defaults_plan_a_short_run():
    pl = bench.plan(bench.parse_args([]))
    assert pl["rounds"] >= 1 and pl["warm_rounds"] >= 1
    assert pl["timed_mc_steps_per_replica"] <= 100000


def test_zero_warmup_still_warms_one_round():
    pl = bench.plan(bench.parse_args(["--warmup", "0", "--steps", "3"]))
    assert pl["warm_rounds"] == 1 and pl["rounds"] == 3


@pytest.mark.parametrize("argv", [["--gpus", "2"], ["--gpus", "8"], ["--steps", "0"], ["--steps-per-launch", "0"]])
def test_unrunnable_arguments_raise(argv):
    with pytest.raises(SystemExit):
        bench.plan(bench.parse_args(argv))


def test_multi_gpu_request_exits_with_message():
    p = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2", "--steps", "20"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
