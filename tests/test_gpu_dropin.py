"""The GPU drop-in end to end: the reference's own C driver and CLI linked over
libslosh_ftn.so (the Fortran-ABI entry points over the HIP time step) writes
the same env and rex files, byte for byte, as the same driver over the FP32
oracle (oracle/orc_ftn.c + liborc32).  Against the reference FP64 build the
FP32 envelope of the 48 h polar run agrees on wet/dry for >= 99 % of cells
(measured 99.5 %); the two precisions take different branches at wetting
fronts (see test_oracle_golden), so no cell-wise tolerance is asserted."""
import os

import numpy as np
import pytest

import helpers
from helpers import synth
from test_dropin_cli import CASES

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (os.path.exists(helpers.GPU_CLI) and os.path.exists(helpers.ORC_CLI["32"])),
                                 reason="oracle/_ref drop-in binaries not built")]


def _run_pair(b, t, root):
    bsn = synth.write_case(root, b, t)["basin"]
    print("drop-in", bsn, b.imxb, b.jmxb, flush=True)
    env_g, rex_g = helpers.run_cli(helpers.GPU_CLI, root, bsn, "gpu")
    env_o, rex_o = helpers.run_cli(helpers.ORC_CLI["32"], root, bsn, "orc32")
    return bsn, env_g, rex_g, env_o, rex_o


@pytest.mark.parametrize("cid,b,t", CASES, ids=[c[0] for c in CASES])
def test_gpu_cli_matches_oracle32_cli(cid, b, t, tmp_path):
    _, env_g, rex_g, env_o, rex_o = _run_pair(b, t, str(tmp_path))
    assert helpers.same_cli_outputs(env_g, rex_g, env_o, rex_o) == []


def test_gpu_cli_polar_100x90_vs_reference(tmp_path):
    """Polar 100x90, 48 h: identical bytes to the FP32 oracle through the same
    driver, and an envelope that agrees with the reference FP64 CLI."""
    b, t = synth.config1()
    root = str(tmp_path)
    bsn, env_g, rex_g, env_o, rex_o = _run_pair(b, t, root)
    assert helpers.same_cli_outputs(env_g, rex_g, env_o, rex_o) == []
    env_r, _ = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref")
    g, r = helpers.read_env(env_g)[3], helpers.read_env(env_r)[3]
    wet_g, wet_r = g != 999, r != 999
    both = wet_g & wet_r
    d = np.abs(g.astype(int) - r.astype(int))[both]
    print("fp32 vs reference fp64: wet/dry agreement %.4f, within 0.1 ft %.4f, max %.1f ft"
          % ((wet_g == wet_r).mean(), (d <= 1).mean(), d.max() / 10.0))
    assert (wet_g == wet_r).mean() >= 0.99


def _run_gpu64(root, bsn):
    old = os.environ.get("SLG_PRECISION")
    os.environ["SLG_PRECISION"] = "64"
    try:
        return helpers.run_cli(helpers.GPU_CLI, root, bsn, "gpu64")
    finally:
        if old is None:
            del os.environ["SLG_PRECISION"]
        else:
            os.environ["SLG_PRECISION"] = old


@pytest.mark.parametrize("cid,b,t", CASES, ids=[c[0] for c in CASES])
def test_gpu_cli_fp64_matches_reference_cli(cid, b, t, tmp_path):
    """With the FP64 step the drop-in writes the reference CLI's own env and rex
    bytes."""
    root = str(tmp_path)
    bsn = synth.write_case(root, b, t)["basin"]
    env_g, rex_g = _run_gpu64(root, bsn)
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref")
    assert helpers.same_cli_outputs(env_g, rex_g, env_r, rex_r) == []


def test_gpu_cli_fp64_polar_100x90_matches_reference_cli(tmp_path):
    b, t = synth.config1()
    root = str(tmp_path)
    bsn = synth.write_case(root, b, t)["basin"]
    env_g, rex_g = _run_gpu64(root, bsn)
    env_r, rex_r = helpers.run_cli(helpers.REF_CLI, root, bsn, "ref")
    assert helpers.same_cli_outputs(env_g, rex_g, env_r, rex_r) == []
