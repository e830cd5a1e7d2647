"""ctypes binding of the CPU FP64 oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
as the checker, and by bench.py's cpu_baseline leg.  The product package
(rbmd_amd) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from rbmd_amd.abi import DescHolder, Params, Thermo, make_params

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build(force: bool = False):
    if force or not os.path.exists(_SO):
        subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        vp = C.c_void_p
        L.orc_create.restype = C.c_int
        L.orc_destroy.argtypes = [vp]
        L.orc_rbe_S.restype = C.c_double
        L.orc_rbe_S.argtypes = [C.c_double, C.POINTER(C.c_double)]
        L.orc_rbe_S_direct.restype = C.c_double
        L.orc_rbe_S_direct.argtypes = [C.c_double, C.POINTER(C.c_double), C.c_int32]
        L.orc_self_energy_total.restype = C.c_double
        L.orc_self_energy_total.argtypes = [vp]
        L.orc_rbe_sample.argtypes = [C.c_double, C.POINTER(C.c_double), C.c_int32, C.c_uint64, C.c_uint64,
                                     C.POINTER(C.c_double)]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_md_step.argtypes = [vp, C.c_int64]
        _lib = L
    return _lib


def _p(a, t=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def set_threads(n: int) -> int:
    return lib().orc_set_threads(int(n))


def static_table_gnear(dis):
    """G of the RBL erfc step table at each distance (ExecStaticTable.h:96-135)."""
    d = np.ascontiguousarray(dis, np.float64)
    g = np.zeros_like(d)
    rc = lib().orc_static_table_gnear(C.c_int32(len(d)), _p(d), _p(g))
    assert rc == 0
    return g


def rbe_S(alpha, box):
    b = np.ascontiguousarray(box, np.float64)
    return lib().orc_rbe_S(alpha, _p(b))


def rbe_S_direct(alpha, box, mmax):
    b = np.ascontiguousarray(box, np.float64)
    return lib().orc_rbe_S_direct(alpha, _p(b), int(mmax))


def rbe_sample(alpha, box, p, seed, stream):
    b = np.ascontiguousarray(box, np.float64)
    m = np.zeros(3 * p)
    rc = lib().orc_rbe_sample(alpha, _p(b), int(p), int(seed), int(stream), _p(m))
    assert rc == 0
    return m.reshape(p, 3)


class OracleSystem:
    """One system in the oracle; method names follow the product's C ABI."""

    def __init__(self, system, run):
        self.L = lib()
        self.n = system.n_atoms
        self._desc = DescHolder(system)
        self._params = make_params(run)
        h = C.c_void_p()
        rc = self.L.orc_create(C.byref(self._desc.desc), C.byref(self._params), C.byref(h))
        if rc != 0:
            raise RuntimeError(f"orc_create failed {rc}")
        self.h = h
        x = np.ascontiguousarray(system.x, np.float64).reshape(-1)
        v = np.ascontiguousarray(system.v, np.float64).reshape(-1)
        self.L.orc_set_state(self.h, _p(x), _p(v), None)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_state(self, x=None, v=None):
        xx = None if x is None else np.ascontiguousarray(x, np.float64).reshape(-1)
        vv = None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1)
        self.L.orc_set_state(self.h, _p(xx), _p(vv), None)

    def state(self):
        x = np.zeros(3 * self.n)
        v = np.zeros(3 * self.n)
        f = np.zeros(3 * self.n)
        img = np.zeros(3 * self.n, np.int32)
        self.L.orc_get_state(self.h, _p(x), _p(v), _p(f), _p(img, C.c_int32))
        return x.reshape(-1, 3), v.reshape(-1, 3), f.reshape(-1, 3), img.reshape(-1, 3)

    def geometry(self):
        dims = np.zeros(3, np.int32)
        nc = C.c_int32()
        dx = np.zeros(3)
        self.L.orc_locator_geometry(self.h, _p(dims, C.c_int32), C.byref(nc), _p(dx))
        return dims, nc.value, dx

    def locator(self):
        self.L.orc_locator_build(self.h)
        dims, _, _ = self.geometry()
        nc = int(np.prod(dims))
        cell = np.zeros(self.n, np.int32)
        pid = np.zeros(self.n, np.int32)
        lo = np.zeros(nc, np.int32)
        up = np.zeros(nc, np.int32)
        self.L.orc_locator_export(self.h, _p(cell, C.c_int32), _p(pid, C.c_int32), _p(lo, C.c_int32),
                                  _p(up, C.c_int32))
        return cell, pid, lo, up

    def neighbors(self, cap):
        rc = self.L.orc_neighbors_build(self.h, int(cap))
        if rc != 0:
            raise RuntimeError(f"orc_neighbors_build {rc}")
        cap = self.L.orc_neighbors_capacity(self.h)
        ids = np.zeros(self.n * cap, np.int32)
        off = np.zeros(3 * self.n * cap)
        nn = 2 * self.n if self._params.neighbor_type == 1 else self.n
        num = np.zeros(nn, np.int32)
        self.L.orc_neighbors_export(self.h, int(cap), _p(ids, C.c_int32), _p(off), _p(num, C.c_int32))
        return ids.reshape(self.n, cap), off.reshape(self.n, cap, 3), num

    def bruteforce_counts(self, rc):
        num = np.zeros(self.n, np.int32)
        self.L.orc_bruteforce_counts(self.h, C.c_double(rc), _p(num, C.c_int32))
        return num

    def force_real(self):
        f = np.zeros(3 * self.n)
        vir = np.zeros(6 * self.n)
        elj = np.zeros(self.n)
        ec = np.zeros(self.n)
        rc = self.L.orc_force_real(self.h, _p(f), _p(vir), _p(elj), _p(ec))
        if rc != 0:
            raise RuntimeError(f"orc_force_real {rc}")
        return f.reshape(-1, 3), vir.reshape(-1, 6), elj, ec

    def set_kvectors(self, m):
        m = np.ascontiguousarray(m, np.float64).reshape(-1)
        self.L.orc_rbe_set_kvectors(self.h, len(m) // 3, _p(m))

    def structure_factor(self, p):
        rho = np.zeros(2 * p)
        self.L.orc_rbe_structure_factor(self.h, _p(rho))
        return rho.reshape(p, 2)

    def rbe_force(self):
        f = np.zeros(3 * self.n)
        vir = np.zeros(6 * self.n)
        self.L.orc_rbe_force(self.h, _p(f), _p(vir))
        return f.reshape(-1, 3), vir.reshape(-1, 6)

    def ewald_force(self, kmax):
        km = np.ascontiguousarray(kmax, np.int32)
        f = np.zeros(3 * self.n)
        vir = np.zeros(6 * self.n)
        e = C.c_double()
        self.L.orc_ewald_force(self.h, _p(km, C.c_int32), _p(f), _p(vir), C.byref(e))
        return f.reshape(-1, 3), vir.reshape(-1, 6), e.value

    def self_energy(self):
        return self.L.orc_self_energy_total(self.h)

    def force_bonded(self):
        f = np.zeros(3 * self.n)
        eb = C.c_double()
        ea = C.c_double()
        self.L.orc_force_bonded(self.h, _p(f), C.byref(eb), C.byref(ea))
        return f.reshape(-1, 3), eb.value, ea.value

    def md_init(self):
        return self.L.orc_md_init(self.h)

    def md_step(self, n):
        return self.L.orc_md_step(self.h, int(n))

    def thermo(self):
        t = Thermo()
        self.L.orc_md_thermo(self.h, C.byref(t))
        return t.as_dict()

    def md_energy(self):
        return self.L.orc_md_energy(self.h)
