"""The C ABI driven from compiled C alone (tests/c_harness/dqmc_harness.c): the reference's
initialize_run -> sweep_once! order against the config-A fixture."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_c_harness_config_a():
    exe = os.path.join(ROOT, "tests", "c_harness", "dqmc_harness")
    fixture = os.path.join(ROOT, "tests", "golden", "config_a_raw.bin")
    r = subprocess.run([exe, fixture], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
