/*
 * rbmd_mi355x.h -- C ABI of the MI355X-native RBMD force path.
 *
 * The reference (RBMD 2.2.0) exposes this path as C++ seams, not a C ABI:
 *   ContPointLocator       framework/locator/ContPointLocator.h:40-127
 *   ContForceFunction      framework/forceFunction/ContForceFunction.h:33-83
 *   ContTopology           framework/topology/ContTopology.h:36-87
 *   RunWorklet operators   framework/run/worklet/RunWorklet.h:40-527
 *   ExecutionMD / NVT      framework/run/ExecutionMD.h:36-96, ExecutionNVT.h:33-180
 *   Executioner loop       framework/executioner/Executioner.h:36-71
 * Each entry point below names the seam it replaces.  Everything is plain C:
 * host pointers and sizes, FP64 on the host side, int32 atom ids (the
 * reference builds with VTKm_USE_64BIT_IDS=OFF, README.md:94).
 *
 * Ownership: the library owns every device buffer.  Host arrays are borrowed
 * for the duration of the call only.  A context owns one HIP stream; calls on
 * a context enqueue on that stream and are not thread-safe.  Functions that
 * return data to host memory synchronise the stream.
 *
 * Errors: every function returns an rbmd_status (0 = ok, negative = error);
 * rbmd_last_error() returns a message for the last failure on this thread.
 * The library never exits the process (the reference's console::Error calls
 * exit(0), framework/Logging.h:44-51).
 */
#ifndef RBMD_MI355X_H
#define RBMD_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rbmd_status;

#define RBMD_OK 0
#define RBMD_ERR_ARG (-1)         /* invalid argument or descriptor */
#define RBMD_ERR_OOM (-2)         /* device allocation failed */
#define RBMD_ERR_HIP (-3)         /* HIP runtime error */
#define RBMD_ERR_SHAKE (-4)       /* SHAKE did not converge / singular */
#define RBMD_ERR_NONFINITE (-5)   /* NaN or Inf detected in state */
#define RBMD_ERR_STATE (-6)       /* call out of order (e.g. step before init) */
#define RBMD_ERR_CAPACITY (-7)    /* a bounded buffer would overflow */
#define RBMD_ERR_UNSUPPORTED (-8) /* option not built */
#define RBMD_ERR_NODEVICE (-9)    /* no HIP device visible */

/* Unit systems: ExecutionMD::InitParameters, framework/run/ExecutionMD.cpp:329-360 */
enum { RBMD_UNIT_LJ = 0, RBMD_UNIT_REAL = 1, RBMD_UNIT_METAL = 2 };
/* force_field.type: framework/run/ExecutionNVT.cpp:195-235 */
enum { RBMD_FF_LJ_CUT = 0, RBMD_FF_LJ_CUT_COUL_LONG = 1, RBMD_FF_CVFF = 2 };
/* neighbor.type: framework/hyperparameters/Neighbor.cpp:34-49 */
enum { RBMD_NEIGHBOR_VERLET = 0, RBMD_NEIGHBOR_RBL = 1 };
/* coulomb.type: framework/hyperparameters/Coulomb.cpp:34-41 */
enum { RBMD_COULOMB_NONE = 0, RBMD_COULOMB_RBE = 1, RBMD_COULOMB_EWALD = 2 };
/* extend.fix_shake as read by GetJudge: absent -> "null" (Configuration.h:98-110) */
enum { RBMD_SHAKE_NULL = 0, RBMD_SHAKE_TRUE = 1, RBMD_SHAKE_FALSE = 2 };
/* execution.temp_ctrl_type: framework/run/ExecutionNVT.cpp:90-135.
 * NONE is plain velocity Verlet (the reference's NVE is a stub,
 * framework/run/ExecutionNVE.cpp:36-41). */
enum { RBMD_TEMP_NOSE_HOOVER = 0, RBMD_TEMP_BERENDSEN = 1, RBMD_TEMP_RESCALE = 2, RBMD_TEMP_NONE = 3 };
/* Arithmetic of the pair / structure-factor kernels.  FP64: everything in
 * double.  MIXED: the pair scalar (1/r, LJ, the erfc/exp table) and the
 * sin/cos of every rho(k) and k-space force term in FP32, each FP32 result
 * summed in FP64 (exact per term, no worse than two-sum compensation); pair
 * vectors, phases, positions, velocities, SHAKE and all reductions stay FP64.
 * Built for the VERLET list path (pair_path LISTS, n_types <= 16); other
 * combinations return RBMD_ERR_UNSUPPORTED.  rbmd_ewald_force is FP64 in
 * both. */
enum { RBMD_PREC_FP64 = 0, RBMD_PREC_MIXED = 1 };
/* Real-space erfc kernel: analytic G(r) (ExecForceFunction.h:682-689) or the
 * fixed alpha=0.1 step table of the RBL path (ExecStaticTable.h:96-135). */
enum { RBMD_ERFC_ANALYTIC = 0, RBMD_ERFC_TABLE = 1 };

/* VERLET real-space path.  CELLS: the neighbour search and the pair term in
 * one fused pass over the cell stencil (no stored list; same pair set and
 * pair term, force sums in a different order).  LISTS: build the Verlet lists
 * (rbmd_neighbors_*), then evaluate the pair term over them. */
enum { RBMD_PAIR_LISTS = 0, RBMD_PAIR_CELLS = 1 };

/* Atom/topology description of one periodic system.
 * Mirrors what ModelFileInitCondition puts into Para
 * (framework/InitCondition/ModelFileInitCondition.cpp:406-478, 736-956). */
typedef struct rbmd_sys_desc {
  int32_t n_atoms;
  int32_t n_types;
  double lo[3];                   /* box range min (xlo ylo zlo) */
  double hi[3];                   /* box range max (xhi yhi zhi) */
  const int32_t* type;            /* [n_atoms] 0-based atom type */
  const double* charge;           /* [n_atoms] */
  const double* mass;             /* [n_atoms] */
  const int32_t* mol_id;          /* [n_atoms] 0-based molecule id, or NULL */
  const double* eps;              /* [n_types] LJ epsilon */
  const double* sigma;            /* [n_types] LJ sigma */
  /* special-bond exclusion lists, CSR over atoms (all three NULL if absent):
   * SetSpecialBonds, ModelFileInitCondition.cpp:848-956 */
  const int32_t* special_offsets; /* [n_atoms+1] */
  const int32_t* special_ids;     /* [special_offsets[n_atoms]] */
  const double* special_weights;  /* [special_offsets[n_atoms]] */
  /* harmonic bonds: E = k (r - r0)^2 (MolecularWorklet.h:58-160) */
  int32_t n_bonds;
  const int32_t* bond_atoms;      /* [2*n_bonds] */
  const double* bond_k;           /* [n_bonds] */
  const double* bond_r0;          /* [n_bonds] */
  /* harmonic angles: E = k (theta - theta0)^2, theta0 in degrees
   * (MolecularWorklet.h:219-397); SHAKE triplets are the angle list (H,O,H) */
  int32_t n_angles;
  const int32_t* angle_atoms;     /* [3*n_angles] */
  const double* angle_k;          /* [n_angles] */
  const double* angle_theta0;     /* [n_angles] degrees */
} rbmd_sys_desc;

/* Run parameters: the `rbmd -j` JSON keys of the force path
 * (MDApplication.cpp:104-244, ExecutionNVT::InitParameters :818-839). */
typedef struct rbmd_params {
  int32_t unit;                   /* RBMD_UNIT_* */
  int32_t force_field;            /* RBMD_FF_* */
  int32_t neighbor_type;          /* RBMD_NEIGHBOR_* */
  int32_t coulomb_type;           /* RBMD_COULOMB_* */
  double cut_off;                 /* neighbor.cut_off */
  double r_core;                  /* neighbor.r_core (RBL); = cut_off for VERLETLIST */
  int32_t neighbor_sample_num;    /* neighbor.neighbor_sample_num (RBL) */
  int32_t rbe_p;                  /* coulomb.coulomb_sample_num */
  double alpha;                   /* coulomb.alpha (= beta^2 of erfc(beta r)) */
  int32_t kmax[3];                /* coulomb.kmax */
  int32_t fix_shake;              /* RBMD_SHAKE_* */
  int32_t special_bonds;          /* 1 when extend.special_bonds is present */
  int32_t inbuild;                /* 1 for init_configuration.inbuild */
  int32_t temp_ctrl;              /* RBMD_TEMP_* */
  double temperature[3];          /* [t_start, t_stop, t_damp] */
  double dt;                      /* execution.timestep */
  int64_t num_steps;              /* execution.num_steps (loop runs num_steps+1) */
  int32_t precision;              /* RBMD_PREC_* */
  int32_t erfc_mode;              /* RBMD_ERFC_* */
  uint64_t seed;                  /* k-vector sampler seed */
  int32_t energy_every;           /* k-grid Ewald energy every n steps, 0 = never */
  double rbl_rho;                 /* number density used by RBL capacities; 0 = N/V */
  int32_t pair_path;              /* VERLET real space: RBMD_PAIR_* */
  /* tile choices (0 = built-in default): lanes per atom of the pair kernel
   * (4, 8, 16; default by atom count), wavefronts splitting the k vectors
   * of the k-space force kernel (2, 4, 8; default 4), k vectors per lane of
   * the structure-factor kernel (1, 2, 4; default 1).  rbmd_amd/tune.py
   * measures them per kind, precision and size and rbmd_amd/tiles.json holds
   * the result. */
  int32_t pair_lanes;
  int32_t kforce_lanes;
  int32_t rho_kblock;
  /* SHAKE convergence tolerance on the Lagrange-multiplier change between
   * iterations; 0 = the reference's 1e-5 (MolecularWorklet.h:2507-2508) */
  double shake_tol;
  /* Verlet-list skin for MD steps (VERLET lists path, both boxes edges >= 2 (rc + skin)):
   * 0 = the reference's list rebuilt every step (ExecutionMD.cpp:878-957);
   * > 0 = one list to rc + skin kept over steps and rebuilt on the device when
   * an atom has moved more than skin / 2 since the last build.  The pair
   * kernel applies the reference's test rc^2 - d^2 > 1e-4 to every entry,
   * so the pairs are the same; only the summation order differs. */
  double skin;
} rbmd_params;

/* Thermodynamic scalars after a step; per-atom averages as in Para
 * (ExecutionMD.cpp:742-749, ExecutionNVT.cpp:736-773, 1258-1276). */
typedef struct rbmd_thermo {
  int64_t step;
  double temperature;             /* PARA_TEMPT */
  double pressure;                /* PARA_PRESSURE */
  double ke_sum;                  /* PARA_TEMPT_SUM = sum m v^2 mvv2e */
  double e_lj;                    /* PARA_LJ_ENERGY (sum / N) */
  double e_coul;                  /* PARA_COUL_ENERGY (sum / N) */
  double e_long;                  /* PARA_EWALD_LONG_ENERGY (grid + self, / N); valid when energy_step == step */
  double e_bond;                  /* PARA_BOND_ENERGY (sum / n_bonds) */
  double e_angle;                 /* PARA_ANGLE_ENERGY (sum / n_angles) */
  double virial[6];               /* xx yy zz xy xz yz, as ComputeVirial */
  double nhc_eta[3];
  double nhc_eta_dot[3];
  int64_t energy_step;            /* step at which e_long was last evaluated */
  int32_t shake_max_iter;         /* most SHAKE iterations of any molecule this step */
  int32_t status_flags;           /* bit0 nonfinite, bit1 shake not converged, bit2 capacity */
  int32_t misfit_atoms;           /* atoms listed from their own stencil in the last list build */
  int32_t list_builds;            /* cell lists sorted (neighbour lists built) since creation */
} rbmd_thermo;

typedef struct rbmd_ctx rbmd_ctx;
typedef struct rbmd_sys rbmd_sys;
typedef struct rbmd_pack rbmd_pack;

/* ---- context (replaces the VTK-m device tag switch, Application.cpp:38-130) ---- */
rbmd_status rbmd_ctx_create(int32_t device, rbmd_ctx** out);
rbmd_status rbmd_ctx_destroy(rbmd_ctx* ctx);
rbmd_status rbmd_ctx_synchronize(rbmd_ctx* ctx);
const char* rbmd_last_error(void);
const char* rbmd_version(void);

/* ---- system (replaces Para fields + ContTopology::Set*, ContForceFunction::SetParameters) ---- */
rbmd_status rbmd_sys_create(rbmd_ctx* ctx, const rbmd_sys_desc* desc, const rbmd_params* params, rbmd_sys** out);
rbmd_status rbmd_sys_destroy(rbmd_sys* sys);
/* x, v: AoS [3*n] FP64; image: [3*n] int32 or NULL */
rbmd_status rbmd_sys_upload_state(rbmd_sys* sys, const double* x, const double* v, const int32_t* image);
rbmd_status rbmd_sys_download_state(rbmd_sys* sys, double* x, double* v, double* force, int32_t* image);
/* bytes of device memory this system holds (admission by footprint) */
rbmd_status rbmd_sys_footprint(const rbmd_sys* sys, int64_t* bytes);
/* estimate before creation */
rbmd_status rbmd_footprint_estimate(const rbmd_sys_desc* desc, const rbmd_params* params, int64_t* bytes);

/* ---- locator: ContPointLocator::SetRange/SetCutOff/SetRs/SetPosition->Build
 *      (ContPointLocator.h:44-94, ContPointLocator.cxx:38-92) ---- */
rbmd_status rbmd_locator_geometry(const rbmd_sys* sys, int32_t dims[3], int32_t* num_cycles, double dxdydz[3]);
rbmd_status rbmd_locator_build(rbmd_sys* sys);
/* cell_of_atom [n], point_ids [n], cell_lower/upper [n_cells]; any may be NULL */
rbmd_status rbmd_locator_export(rbmd_sys* sys, int32_t* cell_of_atom, int32_t* point_ids, int32_t* cell_lower,
                                int32_t* cell_upper);

/* ---- neighbour lists: ComputeNeighboursWorklet (RunWorklet.cpp:296-355) and
 *      ComputeRBLNeighboursOnceWorklet (RunWorklet.cpp:906-980).  Export for parity:
 *      ids [n*cap], offsets [3*n*cap], num [n] (VERLET) or [2*n] (RBL: core, shell kept). ---- */
rbmd_status rbmd_neighbors_build(rbmd_sys* sys, int32_t cap_per_atom);
rbmd_status rbmd_neighbors_export(rbmd_sys* sys, int32_t cap_per_atom, int32_t* ids, double* offsets, int32_t* num);

/* capacity (slots per atom) of the lists last built */
rbmd_status rbmd_neighbors_capacity(const rbmd_sys* sys, int32_t* cap);
/* total list entries (VERLET: pairs counted from both sides; RBL: core + kept shell) */
rbmd_status rbmd_neighbors_count(rbmd_sys* sys, int64_t* entries);

/* ---- real-space force: ComputeNearForceVerlet(Weight)VirialWorklet (RunWorklet.cpp:2767-2985),
 *      ComputeNearForceRBLERF(SpecialBonds)Worklet (RunWorklet.cpp:982-1266).
 *      Uses the current locator; outputs per atom (any may be NULL). ---- */
rbmd_status rbmd_force_real(rbmd_sys* sys, double* force, double* virial6, double* e_lj, double* e_coul);

/* ---- RBE k-space: RBEPSAMPLE (RBEPSample.h:40-177), ComputePnumberChargeStructureFactor
 *      (RunWorklet.cpp:2209-2291), ComputeNewRBEForceWorklet (RunWorklet.cpp:2086-2141) ---- */
rbmd_status rbmd_rbe_set_kvectors(rbmd_sys* sys, int32_t p, const double* m); /* m: [3*p] integer-valued */
rbmd_status rbmd_rbe_sample(rbmd_sys* sys, uint64_t seed, uint64_t stream, double* m_out); /* draws p vectors */
rbmd_status rbmd_rbe_sample_host(double alpha, const double box[3], int32_t p, uint64_t seed, uint64_t stream,
                                 double* m_out);
rbmd_status rbmd_rbe_S(double alpha, const double box[3], double* S);
rbmd_status rbmd_rbe_structure_factor(rbmd_sys* sys, double* rho_out); /* [2*p]: Re, Im */
/* force already multiplied by qqr2e (UnitRescale, ExecutionNVT.cpp:718); virial not (ExecutionMD.cpp:1309) */
rbmd_status rbmd_rbe_force(rbmd_sys* sys, double* force, double* virial6);

/* ---- Ewald (ComputeNewFarElectrostaticsWorklet RunWorklet.cpp:2492-2540, ComputeEwaldEnergy
 *      ExecutionMD.cpp:405-443, ComputeSelfEnergy :544-558) over the kmax grid ---- */
rbmd_status rbmd_ewald_force(rbmd_sys* sys, const int32_t kmax[3], double* force, double* virial6,
                             double* energy_total);

/* ---- the RBL path's erfc step table: ExecStaticTable::Extract + TableGnearValue
 *      (ExecStaticTable.h:96-135) over the reference's tabulated G at alpha = 0.1
 *      (ERFTable.h).  Evaluates the device lookup for n distances (host arrays). ---- */
rbmd_status rbmd_static_table_gnear(rbmd_ctx* ctx, int32_t n, const double* dis, double* g_out);

/* ---- bonded terms (MolecularWorklet.h:58-397); for CVFF + RBL + special bonds also the
 *      intramolecular Coulomb correction SpecialCoulGeneral (RunWorklet.cpp:2012-2084) ---- */
rbmd_status rbmd_force_bonded(rbmd_sys* sys, double* force, double* e_bond_sum, double* e_angle_sum);

/* ---- MD loop: ExecutionNVT::Init (ExecutionNVT.cpp:60-74) then Solve per step (:90-135).
 *      rbmd_md_init evaluates the first force; rbmd_md_step enqueues n steps (PreSolve+Solve
 *      each) on the context stream without host round trips; rbmd_md_run_graph runs the same
 *      steps by replaying one captured step graph (captured on first use and again after
 *      anything it captured changes), with results bit-identical to rbmd_md_step. ---- */
rbmd_status rbmd_md_init(rbmd_sys* sys);
rbmd_status rbmd_md_step(rbmd_sys* sys, int64_t n_steps);
rbmd_status rbmd_md_run_graph(rbmd_sys* sys, int64_t n_steps);
/* Enqueue-only variant for many systems in flight (no counterpart in the reference,
 * whose executioner runs one system): n graph-replayed steps are enqueued on the
 * context stream and the call returns at once; rbmd_md_wait synchronises and
 * returns the status of everything enqueued since; rbmd_md_query reports 1 when
 * the context stream is idle, 0 while work is pending. */
rbmd_status rbmd_md_enqueue(rbmd_sys* sys, int64_t n_steps);
rbmd_status rbmd_md_wait(rbmd_sys* sys);
rbmd_status rbmd_md_query(rbmd_sys* sys, int32_t* idle);
rbmd_status rbmd_md_thermo(rbmd_sys* sys, rbmd_thermo* out);
rbmd_status rbmd_md_energy(rbmd_sys* sys); /* evaluate e_long now (ComputeEwaldEnergy) */
/* rbmd_md_energy without device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and result, bit for bit; the k grid and the structure-factor
 * partials go to a fixed per-device scratch of 32 MiB held until the process
 * ends, so a kmax grid that needs more (8 bytes x (3 K + 2 K + 2 K x atom
 * blocks), K = (2 kmax + 1)^3 - 1) returns RBMD_ERR_ARG.  Calls on one device
 * share the scratch: calls that overlap on one device (two contexts, two
 * threads) race. */
rbmd_status rbmd_md_energy_noalloc(rbmd_sys* sys);

/* ---- packs: many systems, one launch per kernel of the step.  No counterpart
 * in the reference (its executioner, Executioner.cpp:62-129, runs one system);
 * each system's step is the ExecutionNVT::Solve sequence of rbmd_md_step
 * (ExecutionNVT.cpp:90-135) with the same bits as stepping it alone.
 * Systems of a pack share the context, have run rbmd_md_init, use the VERLET
 * list path (pair_path LISTS) with at most 16 types, NOSE_HOOVER or NONE
 * thermostats, energy_every 0, sampled (not injected) RBE k vectors or Ewald,
 * and agree on force field, special-bond mode, precision, SHAKE and
 * thermostat; sizes, boxes and step counts may differ.  A pack keeps the
 * systems' device pointers: do not resize, re-initialise or destroy a system
 * while a pack holds it. ---- */
rbmd_status rbmd_pack_create(rbmd_ctx* ctx, rbmd_sys* const* systems, int32_t n, rbmd_pack** out);
rbmd_status rbmd_pack_destroy(rbmd_pack* pack);
rbmd_status rbmd_pack_step(rbmd_pack* pack, int64_t n_steps);    /* n packed steps, launched kernel by kernel */
rbmd_status rbmd_pack_enqueue(rbmd_pack* pack, int64_t n_steps); /* n replays of one captured packed step */
/* wait for the pack's work; status_out[k] (optional) = status word of system k
 * (bit0 non-finite, bit1 SHAKE, bit2 capacity); returns the first failing
 * system's error */
rbmd_status rbmd_pack_wait(rbmd_pack* pack, int32_t* status_out);

/* ---- per-kernel timing with HIP events on the context stream ---- */
rbmd_status rbmd_profile_enable(rbmd_ctx* ctx, int32_t on);
rbmd_status rbmd_profile_reset(rbmd_ctx* ctx);
/* i-th kernel record: name, launches, total milliseconds; RBMD_ERR_ARG past the end */
rbmd_status rbmd_profile_get(rbmd_ctx* ctx, int32_t i, const char** name, int64_t* launches, double* total_ms);

/* sizeof of rbmd_sys_desc (0), rbmd_params (1), rbmd_thermo (2): lets a binding check its struct mirror */
int32_t rbmd_abi_sizeof(int32_t which);

#ifdef __cplusplus
}
#endif
#endif /* RBMD_MI355X_H */
