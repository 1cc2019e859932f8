"""The C ABI library: loads, exports every declared symbol, reports errors
through status codes, and the host-side helpers compute what they claim."""
from __future__ import annotations

import ctypes as C
import math
import re
import subprocess
from pathlib import Path

import pytest

import cusmc_amd as cs
from cusmc_amd import _capi as A

from conftest import ROOT, gpu_available


def _declared(header: Path) -> set[str]:
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return set(re.findall(r"\b(smc_[a-z0-9_]+)\s*\(", text))


def test_header_declarations_are_exported():
    decl = _declared(ROOT / "include" / "smc_capi.h")
    assert {"smc_model_create", "smc_run", "smc_parse_uppaal", "smc_last_error"} <= decl
    out = subprocess.run(["nm", "-D", "--defined-only", str(A.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (smc_[a-z0-9_]+)$", out, flags=re.M))
    assert decl <= exported, decl - exported
    assert decl <= set(A.SIGNATURES), decl - set(A.SIGNATURES)


def test_library_loads_and_reports_version():
    assert b"gfx950" in A.lib().smc_version()


def test_ch_epsilon_values():
    assert cs.ch_epsilon(10_000) == pytest.approx(math.sqrt(math.log(40) / 20_000))
    assert cs.ch_epsilon(10_000) == pytest.approx(0.0136, abs=5e-5)
    assert cs.ch_epsilon(1_000_000) == pytest.approx(0.00136, abs=5e-6)
    assert math.isnan(A.lib().smc_ch_epsilon(0, 0.05))


def test_null_arguments_are_rejected():
    L = A.lib()
    assert L.smc_model_create(None, 0, None) == A.SMC_E_INVALID
    assert L.smc_run(None, None, 0, 0, 1, None, None, None) == A.SMC_E_INVALID
    assert L.smc_parse_uppaal(None, None, None) == A.SMC_E_INVALID
    assert b"null" in L.smc_last_error()


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU path")
def test_model_create_without_gpu_fails_cleanly(models_dir):
    net = cs.ParsedNetwork.from_file(models_dir / "condTest.xml", "Template.l3")
    with pytest.raises(cs.SmcError) as e:
        cs.GpuSimulator(net)
    assert e.value.code == A.SMC_E_DEVICE


def test_invalid_network_is_rejected_before_the_device(models_dir):
    net = cs.ParsedNetwork.from_file(models_dir / "condTest.xml", "Template.l3")
    n = net.net
    bad = A.Network.from_buffer_copy(n)
    bad.n_components = 0
    h = C.c_void_p()
    assert A.lib().smc_model_create(C.byref(bad), 0, C.byref(h)) == A.SMC_E_INVALID
    assert b"no components" in A.lib().smc_last_error()
