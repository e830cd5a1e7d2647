"""rbmd_md_energy_noalloc: the kmax-grid Ewald energy without device
allocation (the k grid and structure-factor partials in the library's fixed
per-device scratch) against rbmd_md_energy, which allocates them per call.

Same kernels and the same reduction order, so the energies agree bit for bit:
two calls in a row, the second on a smaller system, then the first system
again (the scratch keeps the previous call's contents and must not leak into
the result); a kmax grid beyond the scratch is refused on the host.
"""
import pytest

from rbmd_amd.lib import GpuSystem, RbmdError
from tests.systems import salt, salt_run, water, water_run

pytestmark = pytest.mark.gpu


def _both(g):
    g.md_energy()
    e_alloc = g.thermo()["e_long"]
    g.md_energy_noalloc()
    e_noalloc = g.thermo()["e_long"]
    return e_alloc, e_noalloc


def test_noalloc_energy_matches_allocating_entry_point(ctx):
    big = GpuSystem(ctx, water(), water_run(kmax=[8, 8, 8]))
    small = GpuSystem(ctx, salt(n=512, seed=3), salt_run(kmax=[4, 4, 4]))
    print(f"[case] noalloc energy: water {big.n} then salt {small.n}", flush=True)
    big.md_init()
    small.md_init()
    a1, n1 = _both(big)
    assert a1 != 0.0
    assert n1 == a1
    a2, n2 = _both(small)  # second call: a smaller input
    assert a2 != 0.0
    assert n2 == a2
    big.md_step(3)
    a3, n3 = _both(big)
    assert n3 == a3
    assert a3 != a1


def test_noalloc_refuses_grid_beyond_scratch(ctx):
    g = GpuSystem(ctx, salt(n=512, seed=3), salt_run(kmax=[50, 50, 50]))
    g.md_init()
    with pytest.raises(RbmdError) as e:
        g.md_energy_noalloc()
    assert "scratch" in str(e.value)
    g.md_energy()  # the allocating entry point has no such limit
    assert g.thermo()["e_long"] != 0.0
