"""Force-error report of the random-batch Ewald path on the GPU.

Two error sources are separated, each against an FP64 reference computed on
the same configuration:

* method error: the RBE k-space force with P sampled k vectors against the
  converged FP64 Ewald k-space force at the same alpha (rbmd_ewald_force over
  a k grid whose truncated Gaussian weight e^{-k^2/4 alpha} is below `ktol`).
  The real-space erfc part is the same pair sum in both, so the k-space
  difference is the whole method error at that cutoff.  It is reported per
  k set (one MD step's force) and for the average over M independent k sets,
  which converges to the Ewald force like 1/sqrt(M P) because the RBE force is
  unbiased (E[F_RBE] = F_Ewald,k).  The reference's own Ewald path is
  ComputeNewFarElectrostaticsWorklet (RunWorklet.cpp:2492-2540); the RBE
  force is ComputeNewRBEForceWorklet (RunWorklet.cpp:2086-2141).
* precision error: the mixed-precision path against the FP64 path on the same
  lists and the same k sets (real-space, RBE and total force).

Errors are relative RMS: sqrt(mean_i |dF_i|^2) / sqrt(mean_i |F_i|^2), with
the denominator either the total nonbonded force (real-space LJ + erfc pair
force plus the Ewald k-space force, "_total") or the Ewald k-space force
alone ("_kspace").
"""
import math

import numpy as np

from .lib import GpuSystem, rbe_sample_host


def rms(a):
    a = np.asarray(a).reshape(len(a), -1)
    return float(np.sqrt(np.mean(np.sum(a * a, 1))))


def converged_kmax(alpha, box, ktol=1e-12):
    """per-axis |m| such that every omitted k = 2 pi m / L has e^{-k^2/4 alpha} < ktol"""
    kc = math.sqrt(4.0 * alpha * math.log(1.0 / ktol))
    return [int(math.ceil(kc * float(L) / (2.0 * math.pi))) for L in box]


def rbe_method_error(ctx, system, run, n_sets=16, seed=2024, ktol=1e-12, kmax=None, checkpoints=()):
    """RBE k-space force of n_sets independent k sets against converged Ewald.

    checkpoints: set counts M (<= n_sets) at which the error of the average
    over the first M sets is also reported ("averaged_rel_total_by_m"), with
    the least-squares slope of log(error) against log(M) and the M at which
    that line reaches a relative error of 1e-3."""
    g = GpuSystem(ctx, system, run)
    try:
        km = kmax or converged_kmax(run.alpha, system.box, ktol)
        fe, _, _ = g.ewald_force(km)
        g.build_neighbors(0)
        fr, _, _, _ = g.force_real()
        ftot = fr + fe
        acc = np.zeros_like(fe)
        per = []
        by_m = {}
        marks = set(int(m) for m in checkpoints if 0 < int(m) <= n_sets)
        for s in range(n_sets):
            g.set_kvectors(rbe_sample_host(run.alpha, system.box, run.rbe_p, seed, s + 1))
            fk, _ = g.rbe_force()
            per.append(rms(fk - fe))
            acc += fk
            if s + 1 in marks:
                by_m[s + 1] = rms(acc / (s + 1) - fe) / rms(ftot)
        avg = acc / n_sets
        out = {
            "atoms": system.n_atoms, "P": run.rbe_p, "alpha": run.alpha, "kmax": km, "k_sets": n_sets,
            "rms_force_total": rms(ftot), "rms_force_ewald_k": rms(fe),
            "per_step_rel_total": float(np.mean(per)) / rms(ftot),
            "per_step_rel_kspace": float(np.mean(per)) / rms(fe),
            "averaged_rel_total": rms(avg - fe) / rms(ftot),
            "averaged_rel_kspace": rms(avg - fe) / rms(fe),
        }
        if by_m:
            out["averaged_rel_total_by_m"] = {str(m): by_m[m] for m in sorted(by_m)}
        if len(by_m) >= 2:
            lm = np.log(np.array(sorted(by_m), dtype=float))
            le = np.log(np.array([by_m[m] for m in sorted(by_m)]))
            slope, icpt = np.polyfit(lm, le, 1)
            out["log_slope"] = float(slope)
            out["m_for_1e-3"] = float(np.exp((math.log(1e-3) - icpt) / slope)) if slope < 0 else None
        return out
    finally:
        g.close()


def precision_error(ctx, system, run, seed=2024):
    """mixed precision against FP64 on identical lists and one identical k set"""
    g64 = GpuSystem(ctx, system, type(run)(**{**run.__dict__, "precision": "fp64"}))
    g32 = GpuSystem(ctx, system, type(run)(**{**run.__dict__, "precision": "mixed"}))
    try:
        out = {"atoms": system.n_atoms}
        g64.build_neighbors(0)
        g32.build_neighbors(0)
        r64, v64, el64, ec64 = g64.force_real()
        r32, v32, el32, ec32 = g32.force_real()
        out["real_rel_rms"] = rms(r32 - r64) / rms(r64)
        out["real_virial_rel"] = float(np.abs(v32.sum(0) - v64.sum(0)).max() / np.abs(v64.sum(0)).max())
        out["e_lj_rel"] = float(abs(el32.sum() - el64.sum()) / max(abs(el64.sum()), 1e-300))
        out["e_coul_rel"] = float(abs(ec32.sum() - ec64.sum()) / max(abs(ec64.sum()), 1e-300))
        if run.coulomb_type == "RBE":
            m = rbe_sample_host(run.alpha, system.box, run.rbe_p, seed, 1)
            g64.set_kvectors(m)
            g32.set_kvectors(m)
            k64, kv64 = g64.rbe_force()
            k32, kv32 = g32.rbe_force()
            out["rbe_rel_rms"] = rms(k32 - k64) / rms(k64)
            out["total_rel_rms"] = rms((r32 + k32) - (r64 + k64)) / rms(r64 + k64)
            out["kspace_virial_rel"] = float(np.abs(kv32.sum(0) - kv64.sum(0)).max() / np.abs(kv64.sum(0)).max())
        return out
    finally:
        g64.close()
        g32.close()
