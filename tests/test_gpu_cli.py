"""`cusmc` runs on the GPU: the printed blocks of smc.cu:345-373 (variable
queries) and smc.cu:436-462 with statistics.cuh:165-169 (location queries),
the `Time taken:` lines (smc.cu:252-256) and the results5.csv timing matrix
(simulation/write_csv.cu:15-99), here written to SMC_RESULTS_CSV."""
from __future__ import annotations

import math
import os
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "cusmc_amd" / "lib" / "cusmc"


def _run(args, csv):
    env = dict(os.environ, SMC_RESULTS_CSV=str(csv))
    print(f"[cli] cusmc {' '.join(map(str, args))}", flush=True)
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=120, env=env)


def _check_block(lines, model, sims, seed, bound, comps, logging):
    assert lines[0].startswith("Set number of variables to: ")
    assert lines[1:12] == ["=================", "", "Running SMC with the following settings:",
                           f"- Number of simulations: {sims}", f"- Model: {model}", f"- Random seed: {seed}",
                           f"- Time bound: {bound}", f"- Num components: {comps}", logging, "=================", ""]
    assert re.fullmatch(r"Time taken: [0-9.e+-]+ ms on file " + re.escape(str(model)), lines[12])
    assert re.fullmatch(r"Time taken: [0-9]+ on file " + re.escape(str(model)), lines[13])
    assert re.fullmatch(r"Time taken: [0-9]+ ns \([0-9.e+-]+ ms\)", lines[14])


def test_location_query_block_and_csv(models_dir, tmp_path):
    csv = tmp_path / "results5.csv"
    model = models_dir / "condTest.xml"
    r = _run(["-m", model, "-s", 7, "-q", 20000, 10, 0, -1, 1, 0, "Template.l3"], csv)
    assert r.returncode == 1, r.stderr
    lines = r.stdout.splitlines()
    _check_block(lines, model, 20000, 7, 10, 1, "- Logging type: location query on query: Template.l3")
    assert lines[15] == "Number of simulations: 20000"
    hits = int(re.fullmatch(r"Number of hits: ([0-9]+)", lines[16]).group(1))
    m = re.fullmatch(r"The answer to Template\.l3 is ([0-9.e+-]+)", lines[17])
    assert float(m.group(1)) == pytest.approx(float(np.float32(hits) / np.float32(20000)), rel=1e-5)
    p = 0.75 - 2 * math.log(2) * 2 ** -10.0   # condTest closed form
    assert abs(hits / 20000 - p) < 5 * math.sqrt(p * (1 - p) / 20000)
    assert len(lines) == 18
    rows = [l.split(";") for l in csv.read_text().splitlines()]
    assert rows[0] == ["", "MC1_Sim20000_Time10.00"]
    assert rows[1][0] == "condTest.xml" and float(rows[1][1]) > 0


def test_variable_query_block_and_csv_matrix(models_dir, tmp_path):
    csv = tmp_path / "results5.csv"
    model = models_dir / "RaceTest.xml"
    r = _run(["-m", model, "-q", 20000, 10, -1, 0, 1, 1, " "], csv)
    assert r.returncode == 1, r.stderr
    lines = r.stdout.splitlines()
    _check_block(lines, model, 20000, 1235, 10, 9, "- Logging type: variable query. Finding max of variable with ID 0")
    res = float(re.fullmatch(r"Result: ([0-9.e+-]+)", lines[15]).group(1))
    assert abs(res - (8 * (1 - 2 ** -10) + (4 / 9) * 2 ** -10)) < 0.05
    assert len(lines) == 16
    # a second model and a second configuration extend the matrix
    r2 = _run(["-m", models_dir / "condTest.xml", "-q", 1000, 10, 0, -1, 1, 0, "Template.l3"], csv)
    assert r2.returncode == 1, r2.stderr
    r3 = _run(["-m", model, "-q", 1000, 10, -1, 0, 1, 1, " "], csv)
    assert r3.returncode == 1, r3.stderr
    rows = [l.split(";") for l in csv.read_text().splitlines()]
    assert rows[0] == ["", "MC9_Sim20000_Time10.00", "MC1_Sim1000_Time10.00", "MC9_Sim1000_Time10.00"]
    assert [r[0] for r in rows[1:]] == ["RaceTest.xml", "condTest.xml"]
    assert rows[1][1] and rows[1][2] == "" and rows[1][3]
    assert rows[2][1] == "" and rows[2][2] and len(rows[2]) == 4
