"""Steps without thermo sums.  A call of n steps (rbmd_md_step,
rbmd_md_run_graph) forms the virial and energy sums on its last step only;
the other n - 1 steps compute the same forces and the kinetic energy the
thermostat needs.  The trajectory, the Nose-Hoover chain and the thermo
values read after the call must equal n one-step calls (every step with
the sums) to the bit: FP64 and mixed precision, the list and the fused
cell-stencil real space, eager steps and graph replays."""
import numpy as np
import pytest

from rbmd_amd.lib import GpuSystem
from tests.systems import salt, salt_run, water, water_run

pytestmark = pytest.mark.gpu

KEYS = ("step", "temperature", "pressure", "e_lj", "e_coul", "e_bond", "e_angle")


def same(a, b):
    for u, v in zip(a.state(), b.state()):
        assert np.array_equal(u, v)
    ta, tb = a.thermo(), b.thermo()
    for k in KEYS:
        assert ta[k] == tb[k], k
    assert list(ta["virial"]) == list(tb["virial"])
    assert list(ta["nhc_eta"]) == list(tb["nhc_eta"])


CASES = [
    ("salt", "fp64", "lists"),
    ("salt", "fp64", "cells"),
    ("water", "fp64", "lists"),
    ("water", "mixed", "lists"),
    ("water", "fp64", "cells"),
]


@pytest.mark.parametrize("kind,precision,path", CASES)
def test_call_of_n_steps_equals_n_single_steps(ctx, kind, precision, path):
    if kind == "salt":
        s, run = salt(n=1000, seed=4), lambda: salt_run(precision=precision, pair_path=path)
    else:
        s, run = water(nmol=300, seed=31), lambda: water_run(precision=precision, pair_path=path)
    print(f"[case] thermo-free steps {kind} {precision} {path}", flush=True)
    one, many, graph = (GpuSystem(ctx, s, run()) for _ in range(3))
    for g in (one, many, graph):
        g.md_init()
    for n in (1, 9, 14):
        for _ in range(n):
            one.md_step(1)
        many.md_step(n)
        graph.md_run_graph(n)
        same(one, many)
        same(one, graph)
    for g in (one, many, graph):
        g.close()
