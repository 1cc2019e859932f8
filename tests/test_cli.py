"""The `cusmc` command line (main.cu:10-110) without a GPU: help, argument
errors (each exits 0 after its message, as main.cu:27 does) and a model
that fails to load (exit 2).  The runs themselves are in test_gpu_cli.py."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "cusmc_amd" / "lib" / "cusmc"

USAGE = [
    "Use -m or --model, followed by a path, for inputting a path the the model XML file.",
    "Use -s or --seed, followed by a number, to initialize curand with a constant seed. (0 = random seed)",
    "Use -q or --query, followed by : number of simulation, time bound, a variable threshold, a variable name "
    "(id), Max=1 or Min=0, Estimate=1 or Probability=0.",
]


def _cli(*args, env=None):
    if not CLI.exists():
        pytest.fail(f"{CLI} not built")
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=60, env=env)


def test_help_prints_usage_and_exits_zero():
    for flag in ("-h", "--help"):
        r = _cli(flag)
        assert r.returncode == 0
        assert r.stdout.splitlines() == USAGE


@pytest.mark.parametrize("args,message", [
    (("-m",), "Error: -i option requires a value"),
    (("--model",), "Error: -i option requires a value"),
    (("-s",), "Error: -s option requires a value"),
    (("-s", "abc"), "Error: invalid argument: stoi"),
    (("-s", "99999999999"), "Error: number too big: stoi"),
    (("-q", "10", "5", "1", "0", "1", "1"), "Error: -r option requires a value"),
    (("-q", "10", "x", "1", "0", "1", "1", "P.l"), "Error: invalid argument: stoi"),
    (("--bogus",), "Error: unknown option: --bogus"),
    (("-m", "a.xml", "-x"), "Error: unknown option: -x"),
])
def test_argument_errors(args, message):
    r = _cli(*args)
    assert r.returncode == 0 and r.stdout == ""
    assert r.stderr.splitlines() == [message]


def test_missing_model_file_fails_before_the_run(tmp_path):
    r = _cli("-m", tmp_path / "none.xml", "-q", "10", "5", "1", "0", "1", "1", " ")
    assert r.returncode == 2
    assert r.stderr.startswith("Error:") and r.stdout == ""


def test_variable_count_line_precedes_device_setup(models_dir):
    """The parsed model's variable count is printed before the device model
    is built (smc.cu:317); without a GPU the run then stops with status 2."""
    import os
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")  # hide any GPU: stop at device setup
    r = _cli("-m", models_dir / "condTest.xml", "-q", "100", "10", "1", "-1", "1", "0", "Template.l3", env=env)
    assert r.stdout.splitlines()[0] == "Set number of variables to: 2"
    assert r.returncode == 2 and r.stderr.startswith("Error:")
