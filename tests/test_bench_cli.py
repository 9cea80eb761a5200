"""CPU checks of bench.py's command-line contract (no GPU is touched):
--gpus/--steps/--warmup are accepted, and any GPU count but one is refused
with a message before anything is initialised."""
import os
import subprocess
import sys

BENCH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py")


def run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True, timeout=120)


def test_help_lists_the_contract_flags():
    r = run("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup"):
        assert flag in r.stdout


def test_more_than_one_gpu_is_refused_with_a_message():
    for n in ("2", "8"):
        r = run("--gpus", n, "--steps", "1", "--warmup", "0")
        assert r.returncode == 2
        assert "one GPU" in r.stderr
        assert r.stdout == ""  # no JSON line for a run that did not happen
