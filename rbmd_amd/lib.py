"""ctypes binding of the HIP library (rbmd_amd/_build/librbmd_mi355x.so).

This is the only compute path of the package: every force, list and step
call goes through the C ABI of include/rbmd_mi355x.h.  There is no CPU
fallback; if the library is missing, ``load()`` raises.
"""
import ctypes as C
import os
import weakref

import numpy as np

from .abi import DescHolder, Params, SysDesc, Thermo, make_params

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "librbmd_mi355x.so")
_lib = None

# every symbol include/rbmd_mi355x.h declares
EXPORTS = [
    "rbmd_ctx_create", "rbmd_ctx_destroy", "rbmd_ctx_synchronize", "rbmd_last_error", "rbmd_version",
    "rbmd_sys_create", "rbmd_sys_destroy", "rbmd_sys_upload_state", "rbmd_sys_download_state",
    "rbmd_sys_footprint", "rbmd_footprint_estimate", "rbmd_locator_geometry", "rbmd_locator_build",
    "rbmd_locator_export", "rbmd_neighbors_build", "rbmd_neighbors_count", "rbmd_neighbors_capacity", "rbmd_neighbors_export", "rbmd_force_real",
    "rbmd_rbe_set_kvectors", "rbmd_rbe_sample", "rbmd_rbe_sample_host", "rbmd_rbe_S",
    "rbmd_rbe_structure_factor", "rbmd_rbe_force", "rbmd_ewald_force", "rbmd_force_bonded", "rbmd_md_init",
    "rbmd_md_step", "rbmd_md_run_graph", "rbmd_md_enqueue", "rbmd_md_wait", "rbmd_md_query", "rbmd_md_thermo", "rbmd_md_energy",
    "rbmd_md_energy_noalloc", "rbmd_profile_enable",
    "rbmd_profile_reset", "rbmd_profile_get", "rbmd_abi_sizeof", "rbmd_pack_create", "rbmd_pack_destroy",
    "rbmd_pack_step", "rbmd_pack_enqueue", "rbmd_pack_wait", "rbmd_static_table_gnear",
]


class RbmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rbmd error {code}: {msg}")
        self.code = code


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    pd = C.POINTER(C.c_double)
    pi = C.POINTER(C.c_int32)
    L.rbmd_last_error.restype = C.c_char_p
    L.rbmd_version.restype = C.c_char_p
    L.rbmd_ctx_create.argtypes = [C.c_int32, C.POINTER(vp)]
    L.rbmd_ctx_destroy.argtypes = [vp]
    L.rbmd_ctx_synchronize.argtypes = [vp]
    L.rbmd_sys_create.argtypes = [vp, C.POINTER(SysDesc), C.POINTER(Params), C.POINTER(vp)]
    L.rbmd_sys_destroy.argtypes = [vp]
    L.rbmd_sys_upload_state.argtypes = [vp, pd, pd, pi]
    L.rbmd_sys_download_state.argtypes = [vp, pd, pd, pd, pi]
    L.rbmd_sys_footprint.argtypes = [vp, C.POINTER(C.c_int64)]
    L.rbmd_footprint_estimate.argtypes = [C.POINTER(SysDesc), C.POINTER(Params), C.POINTER(C.c_int64)]
    L.rbmd_locator_geometry.argtypes = [vp, pi, pi, pd]
    L.rbmd_locator_build.argtypes = [vp]
    L.rbmd_locator_export.argtypes = [vp, pi, pi, pi, pi]
    L.rbmd_neighbors_build.argtypes = [vp, C.c_int32]
    L.rbmd_neighbors_export.argtypes = [vp, C.c_int32, pi, pd, pi]
    L.rbmd_neighbors_count.argtypes = [vp, C.POINTER(C.c_int64)]
    L.rbmd_neighbors_capacity.argtypes = [vp, C.POINTER(C.c_int32)]
    L.rbmd_force_real.argtypes = [vp, pd, pd, pd, pd]
    L.rbmd_rbe_set_kvectors.argtypes = [vp, C.c_int32, pd]
    L.rbmd_rbe_sample.argtypes = [vp, C.c_uint64, C.c_uint64, pd]
    L.rbmd_rbe_sample_host.argtypes = [C.c_double, pd, C.c_int32, C.c_uint64, C.c_uint64, pd]
    L.rbmd_rbe_S.argtypes = [C.c_double, pd, pd]
    L.rbmd_rbe_structure_factor.argtypes = [vp, pd]
    L.rbmd_rbe_force.argtypes = [vp, pd, pd]
    L.rbmd_ewald_force.argtypes = [vp, pi, pd, pd, pd]
    L.rbmd_force_bonded.argtypes = [vp, pd, pd, pd]
    L.rbmd_md_init.argtypes = [vp]
    L.rbmd_md_step.argtypes = [vp, C.c_int64]
    L.rbmd_md_run_graph.argtypes = [vp, C.c_int64]
    L.rbmd_md_enqueue.argtypes = [vp, C.c_int64]
    L.rbmd_md_wait.argtypes = [vp]
    L.rbmd_md_query.argtypes = [vp, C.POINTER(C.c_int32)]
    L.rbmd_md_thermo.argtypes = [vp, C.POINTER(Thermo)]
    L.rbmd_md_energy.argtypes = [vp]
    L.rbmd_md_energy_noalloc.argtypes = [vp]
    L.rbmd_profile_enable.argtypes = [vp, C.c_int32]
    L.rbmd_profile_reset.argtypes = [vp]
    L.rbmd_profile_get.argtypes = [vp, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_int64), pd]
    L.rbmd_pack_create.argtypes = [vp, C.POINTER(vp), C.c_int32, C.POINTER(vp)]
    L.rbmd_pack_destroy.argtypes = [vp]
    L.rbmd_pack_step.argtypes = [vp, C.c_int64]
    L.rbmd_pack_enqueue.argtypes = [vp, C.c_int64]
    L.rbmd_pack_wait.argtypes = [vp, pi]
    L.rbmd_static_table_gnear.argtypes = [vp, C.c_int32, pd, pd]
    L.rbmd_abi_sizeof.argtypes = [C.c_int32]
    L.rbmd_abi_sizeof.restype = C.c_int32
    for name in EXPORTS:
        f = getattr(L, name)
        if f.restype is None or f.restype is C.c_int:
            if name not in ("rbmd_last_error", "rbmd_version"):
                f.restype = C.c_int32
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise RbmdError(rc, load().rbmd_last_error().decode())
    return rc


def _p(a, t=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(t))


def rbe_S(alpha, box):
    b = np.ascontiguousarray(box, np.float64)
    out = C.c_double()
    check(load().rbmd_rbe_S(alpha, _p(b), C.byref(out)))
    return out.value


def rbe_sample_host(alpha, box, p, seed, stream):
    b = np.ascontiguousarray(box, np.float64)
    m = np.zeros(3 * p)
    check(load().rbmd_rbe_sample_host(alpha, _p(b), int(p), int(seed), int(stream), _p(m)))
    return m.reshape(p, 3)


def static_table_gnear(ctx, dis):
    """the RBL erfc step table looked up on the device (ExecStaticTable.h:96-135)"""
    d = np.ascontiguousarray(dis, np.float64)
    g = np.zeros_like(d)
    check(load().rbmd_static_table_gnear(ctx.h, C.c_int32(len(d)), _p(d), _p(g)))
    return g


class Context:
    """One HIP device + stream (rbmd_ctx)."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = C.c_void_p()
        check(self.L.rbmd_ctx_create(int(device), C.byref(h)))
        self.h = h
        self._systems = weakref.WeakSet()
        self._packs = weakref.WeakSet()

    def close(self):
        if getattr(self, "h", None):
            for p in list(self._packs):
                p.close()
            for s in list(self._systems):
                s.close()
            self.L.rbmd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def synchronize(self):
        check(self.L.rbmd_ctx_synchronize(self.h))

    def profile(self, on=True):
        check(self.L.rbmd_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        check(self.L.rbmd_profile_reset(self.h))

    def profile_records(self):
        out = {}
        i = 0
        while True:
            name = C.c_char_p()
            n = C.c_int64()
            ms = C.c_double()
            if self.L.rbmd_profile_get(self.h, i, C.byref(name), C.byref(n), C.byref(ms)) != 0:
                break
            out[name.value.decode()] = (n.value, ms.value)
            i += 1
        return out


class GpuSystem:
    """One system on the device; method names mirror oracle.OracleSystem."""

    def __init__(self, ctx: Context, system, run):
        self.ctx = ctx
        self.L = ctx.L
        self.n = system.n_atoms
        self._desc = DescHolder(system)
        self._params = make_params(run, system.n_atoms)
        h = C.c_void_p()
        check(self.L.rbmd_sys_create(ctx.h, C.byref(self._desc.desc), C.byref(self._params), C.byref(h)))
        self.h = h
        ctx._systems.add(self)
        self.set_state(system.x, system.v)

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.L.rbmd_sys_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def footprint(self):
        b = C.c_int64()
        check(self.L.rbmd_sys_footprint(self.h, C.byref(b)))
        return b.value

    @staticmethod
    def estimate_footprint(system, run):
        """device bytes a system would hold (rbmd_footprint_estimate; no device needed)"""
        d = DescHolder(system)
        p = make_params(run)
        b = C.c_int64()
        check(load().rbmd_footprint_estimate(C.byref(d.desc), C.byref(p), C.byref(b)))
        return b.value

    def set_state(self, x=None, v=None):
        xx = None if x is None else np.ascontiguousarray(x, np.float64).reshape(-1)
        vv = None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1)
        check(self.L.rbmd_sys_upload_state(self.h, _p(xx), _p(vv), None))

    def state(self):
        x = np.zeros(3 * self.n)
        v = np.zeros(3 * self.n)
        f = np.zeros(3 * self.n)
        img = np.zeros(3 * self.n, np.int32)
        check(self.L.rbmd_sys_download_state(self.h, _p(x), _p(v), _p(f), _p(img, C.c_int32)))
        return x.reshape(-1, 3), v.reshape(-1, 3), f.reshape(-1, 3), img.reshape(-1, 3)

    def geometry(self):
        dims = np.zeros(3, np.int32)
        nc = C.c_int32()
        dx = np.zeros(3)
        check(self.L.rbmd_locator_geometry(self.h, _p(dims, C.c_int32), C.byref(nc), _p(dx)))
        return dims, nc.value, dx

    def locator(self):
        check(self.L.rbmd_locator_build(self.h))
        dims, _, _ = self.geometry()
        nc = int(np.prod(dims))
        cell = np.zeros(self.n, np.int32)
        pid = np.zeros(self.n, np.int32)
        lo = np.zeros(nc, np.int32)
        up = np.zeros(nc, np.int32)
        check(self.L.rbmd_locator_export(self.h, _p(cell, C.c_int32), _p(pid, C.c_int32), _p(lo, C.c_int32),
                                         _p(up, C.c_int32)))
        return cell, pid, lo, up

    def neighbors(self, cap):
        check(self.L.rbmd_neighbors_build(self.h, int(cap)))
        c = C.c_int32()
        check(self.L.rbmd_neighbors_capacity(self.h, C.byref(c)))
        cap = c.value
        ids = np.zeros(self.n * cap, np.int32)
        off = np.zeros(3 * self.n * cap)
        nn = 2 * self.n if self._params.neighbor_type == 1 else self.n
        num = np.zeros(nn, np.int32)
        check(self.L.rbmd_neighbors_export(self.h, int(cap), _p(ids, C.c_int32), _p(off), _p(num, C.c_int32)))
        return ids.reshape(self.n, cap), off.reshape(self.n, cap, 3), num

    def build_neighbors(self, cap=0):
        check(self.L.rbmd_neighbors_build(self.h, int(cap)))

    def neighbor_entries(self):
        t = C.c_int64()
        check(self.L.rbmd_neighbors_count(self.h, C.byref(t)))
        return t.value

    def force_real(self):
        f = np.zeros(3 * self.n)
        vir = np.zeros(6 * self.n)
        elj = np.zeros(self.n)
        ec = np.zeros(self.n)
        check(self.L.rbmd_force_real(self.h, _p(f), _p(vir), _p(elj), _p(ec)))
        return f.reshape(-1, 3), vir.reshape(-1, 6), elj, ec

    def set_kvectors(self, m):
        m = np.ascontiguousarray(m, np.float64).reshape(-1)
        check(self.L.rbmd_rbe_set_kvectors(self.h, len(m) // 3, _p(m)))

    def sample_kvectors(self, seed, stream):
        m = np.zeros(3 * self._params.rbe_p)
        check(self.L.rbmd_rbe_sample(self.h, int(seed), int(stream), _p(m)))
        return m.reshape(-1, 3)

    def structure_factor(self, p):
        rho = np.zeros(2 * p)
        check(self.L.rbmd_rbe_structure_factor(self.h, _p(rho)))
        return rho.reshape(p, 2)

    def rbe_force(self):
        f = np.zeros(3 * self.n)
        vir = np.zeros(6 * self.n)
        check(self.L.rbmd_rbe_force(self.h, _p(f), _p(vir)))
        return f.reshape(-1, 3), vir.reshape(-1, 6)

    def ewald_force(self, kmax):
        km = np.ascontiguousarray(kmax, np.int32)
        f = np.zeros(3 * self.n)
        vir = np.zeros(6 * self.n)
        e = C.c_double()
        check(self.L.rbmd_ewald_force(self.h, _p(km, C.c_int32), _p(f), _p(vir), C.byref(e)))
        return f.reshape(-1, 3), vir.reshape(-1, 6), e.value

    def force_bonded(self):
        f = np.zeros(3 * self.n)
        eb = C.c_double()
        ea = C.c_double()
        check(self.L.rbmd_force_bonded(self.h, _p(f), C.byref(eb), C.byref(ea)))
        return f.reshape(-1, 3), eb.value, ea.value

    def md_init(self):
        return check(self.L.rbmd_md_init(self.h))

    def md_step(self, n):
        return check(self.L.rbmd_md_step(self.h, int(n)))

    def md_run_graph(self, n):
        return check(self.L.rbmd_md_run_graph(self.h, int(n)))

    def md_enqueue(self, n):
        """enqueue n graph-replayed steps and return at once (see md_wait)"""
        return check(self.L.rbmd_md_enqueue(self.h, int(n)))

    def md_wait(self):
        return check(self.L.rbmd_md_wait(self.h))

    def md_idle(self):
        idle = C.c_int32()
        check(self.L.rbmd_md_query(self.h, C.byref(idle)))
        return bool(idle.value)

    def thermo(self):
        t = Thermo()
        check(self.L.rbmd_md_thermo(self.h, C.byref(t)))
        return t.as_dict()

    def md_energy(self):
        return check(self.L.rbmd_md_energy(self.h))

    def md_energy_noalloc(self):
        return check(self.L.rbmd_md_energy_noalloc(self.h))


class Pack:
    """Systems of one shape stepped together, one launch per kernel of the
    step (rbmd_pack_*).  Each system's results are the same as stepping it
    alone; the systems stay usable (thermo, state) through their GpuSystem."""

    def __init__(self, ctx: Context, systems):
        self.ctx = ctx
        self.L = ctx.L
        self.systems = list(systems)
        arr = (C.c_void_p * len(self.systems))(*[s.h.value for s in self.systems])
        h = C.c_void_p()
        check(self.L.rbmd_pack_create(ctx.h, arr, len(self.systems), C.byref(h)))
        self.h = h
        ctx._packs.add(self)

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.L.rbmd_pack_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def step(self, n):
        """n packed steps, launched kernel by kernel; waits and checks every system"""
        check(self.L.rbmd_pack_step(self.h, int(n)))
        return self.wait()

    def enqueue(self, n):
        """enqueue n graph-replayed packed steps and return at once"""
        check(self.L.rbmd_pack_enqueue(self.h, int(n)))

    def wait(self, raise_on_error=True):
        """wait for the pack; returns the status word of every system"""
        st = np.zeros(len(self.systems), np.int32)
        rc = self.L.rbmd_pack_wait(self.h, _p(st, C.c_int32))
        if rc != 0 and raise_on_error:
            check(rc)
        return st
