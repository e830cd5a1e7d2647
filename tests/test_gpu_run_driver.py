"""The `rbmd -j` driver on the GPU against the same driver on the CPU oracle:
the final thermo values within 1e-8, and every column of energy.rbmd and
temperature.rbmd except the wall-clock ones equal to the printed precision
(6 significant digits, so within one unit of the 6th digit)."""
import os

import numpy as np
import pytest

from rbmd_amd.lib import GpuSystem
from tests.test_run_driver import INBUILD, REF, drive, rows

pytestmark = pytest.mark.gpu


def gpu_factory(ctx):
    return lambda system, run: GpuSystem(ctx, system, run)


def close_printed(a, b):
    a, b = float(a), float(b)
    return abs(a - b) <= 1.5e-5 * max(abs(a), abs(b)) + 1e-300


@pytest.mark.parametrize("path", [REF, INBUILD], ids=["water_reference_input", "lj_inbuild"])
def test_gpu_driver_matches_oracle(ctx, tmp_path, path):
    print(f"[case] rbmd -j {os.path.basename(path)}", flush=True)
    steps = 6 if path == REF else None
    _, _, lo, _ = drive(path, tmp_path / "cpu", steps)
    _, _, lg, _ = drive(path, tmp_path / "gpu", steps, sim_factory=gpu_factory(ctx))
    for k in ("temperature", "pressure", "lj", "near", "far", "pe", "ke", "total"):
        assert abs(lg[k] - lo[k]) <= 1e-8 * max(1.0, abs(lo[k])), (k, lg[k], lo[k])
    for name, skip in (("energy.rbmd", (1, 9)), ("temperature.rbmd", ())):
        ho, bo = rows(tmp_path / "cpu" / name)
        hg, bg = rows(tmp_path / "gpu" / name)
        assert ho == hg and len(bo) == len(bg)
        for ro, rg in zip(bo, bg):
            assert len(ro) == len(rg)
            for c, (a, b) in enumerate(zip(ro, rg)):
                if c not in skip:
                    assert close_printed(a, b), (name, ro[0], c, a, b)
