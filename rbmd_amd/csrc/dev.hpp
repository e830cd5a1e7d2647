// dev.hpp -- shared device-side types and kernel launcher declarations of the
// MI355X force path.  Kernels live in k_*.hip; api.cpp orchestrates them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct rbmd_ctx;  // the C ABI's context (include/rbmd_mi355x.h), defined in api.cpp

namespace rbmd {

constexpr double kPi = 3.14159265358979323846;
constexpr int kBlock = 256;     // threads per workgroup (4 wavefronts of 64)
constexpr int kWave = 64;       // CDNA wavefront width
constexpr int kChunk = 64;      // i-atoms per neighbour-build work item (one wavefront)
constexpr int kMaxSpecial = 8;  // special-list entries held in registers per atom (RBL kernels)
constexpr int kShiftBits = 9;   // 3 bits per axis image shift in packed neighbour entries
constexpr int kMaxLanes = 16;   // most lanes sharing one atom's neighbour list
constexpr int kCoulM = 32;      // Coulomb table intervals over [0, rc]
constexpr int kCoulD = 9;       // polynomial degree per interval
constexpr int kCoulPairs = (kCoulD + 1) / 2;           // coefficient pairs per interval
constexpr int kCoulTab = 2 * kCoulPairs * kCoulM;      // double2 entries (P and Q)
constexpr int kCoulD32 = 5;     // degree of the FP32 table of the mixed-precision path
constexpr int kCoulPairs32 = (kCoulD32 + 1) / 2;
constexpr int kCoulTab32 = 2 * kCoulPairs32 * kCoulM;  // float2 entries
constexpr int kMaxTypes = 16;   // type-pair tables held in LDS up to this many types

// Neighbour-list slot layout: one row of `cap` slots per atom (sorted order),
// so a row's consecutive entries are consecutive words.
// |r|^2 of a pair vector as the reference's host code and the CPU oracle
// round it: x*x + y*y, then + z*z, every operation rounded (no fused
// multiply-add), so the cutoff predicate lists exactly the same pairs on
// every path, also for pairs that sit at the cutoff to the last bit.
__device__ __forceinline__ double dist2(double x, double y, double z) {
#pragma clang fp contract(off)
  return x * x + y * y + z * z;
}

__host__ __device__ __forceinline__ size_t nbr_index(int slot, int i, int cap) { return (size_t)i * cap + slot; }

// Everything a kernel needs to know about one system.  Passed by value.
struct SysView {
  int n;            // atoms
  int ntypes;
  int ncell;        // dims0*dims1*dims2
  int dims[3];
  int nc;           // stencil half width (ceil(rc / rs))
  int cap;          // neighbour-list slots per atom (VERLET: nseg * seg_cap)
  int nseg;         // VERLET: one list segment per (x, y) column of the 3x3x3 stencil (9); RBL: 1
  int seg_cap;      // slots per segment (multiple of the lane count)
  int lanes_lg;     // log2 of the lanes per atom in the pair kernel (2..4)
  int k_lanes_lg;   // log2 of the lanes per atom in the k-space force kernel (1..3)
  int rho_kb;       // k vectors per workgroup in the structure-factor kernel (4, 8, 16)
  int sp_max;       // longest special list
  int n_items;      // capacity of the work-item list
  int weighted;     // special-bond weights in the pair term
  int lj_only;      // LJ/CUT force field
  int rbl;          // random-batch list
  int erfc_table;   // step-table erfc (alpha = 0.1)
  int rs_num, pice_num;
  int nslab;        // real-space partials per atom in fn*, vir_near, e_lj, e_coul (fused path 3, else 1)
  int use_min_image;  // some box edge < 2 rc: the pair vector may need the minimum image
  int mixed;        // RBMD_PREC_MIXED: the pair scalar in FP32 (pair vectors and sums stay FP64)
  double lo[3], hi[3], box[3], dxdydz[3];
  // dxdydz[k] * dims[k]: one image shift along axis k as the cell code forms
  // it (kept beside dims and dxdydz whenever they change)
  double shift_len[3];
  double volume, cut, cut2, rs, rs2, alpha, sqrt_alpha, qqr2e;
  // per atom, original order (SoA)
  double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz;
  double *ox, *oy, *oz, *ovx, *ovy, *ovz;
  double *q, *mass;
  int *type, *img;
  // per type pair tables [ntypes*ntypes]
  double *eps_ij, *sig6_ij;
  // Coulomb radial table: kCoulM intervals of width h over [0, rc]; per
  // interval the degree-kCoulD coefficients (in u = r/h - k - 1/2) of
  // P(r) = erfc(beta r) + 2 beta/sqrt(pi) r e^{-beta^2 r^2} (so G(r)/r = P / r^3)
  // and of Q(r) = erfc(beta r), coefficient-major: ctab[(f * kCoulPairs + d/2) *
  // kCoulM + k] holds coefficients d, d+1 (d even) of function f (0 P, 1 Q) on
  // interval k, so lanes at different distances read different LDS banks
  const double2* ctab;  // [2][kCoulPairs][kCoulM]
  double ctab_inv_h;
  const float2* ctab32; // the same functions at degree kCoulD32 in FP32: [2][kCoulPairs32][kCoulM]
  // special lists (CSR) and molecule groups
  int *sp_off, *sp_ids;
  double *sp_w;
  // cell list
  int *cell_of;     // [n] binned cell (clamped)
  int *cell_count;  // [ncell]
  int *cell_start;  // [ncell+1]
  int *cursor;      // [ncell]
  int *point_ids;   // [n] sorted position -> atom id
  int *rank;        // [n] atom id -> sorted position
  int *items;       // [n_items] packed (cell << 8 | chunk)
  int *n_items_dev; // [1]
  int *misfit;      // [n] sorted position flag
  int *misfit_list; // [n] sorted positions of misfit atoms
  int *misfit_count;// [1]
  // sorted copies for the pair kernels
  double4* spos;    // [n] x,y,z,q in sorted order
  int* stype;       // [n]
  // neighbour lists, column-major [cap][n] of packed (j_sorted << 9 | shift code)
  int* nbr;         // layout nbr_index(slot, i, cap): one row per atom
  int lds_bins;     // binning kernels aggregate counts in LDS (ncell small enough)
  int* nbr_num;     // VERLET [nseg][n] counts per segment; RBL [2n] (core, shell kept)
  // per-atom force components (original order)
  double *fnx, *fny, *fnz;  // real space, sorted order, [nslab][n] partials added in order
  double *fkx, *fky, *fkz;  // k space
  double *fbx, *fby, *fbz;  // bonded
  // per-atom optional outputs (sorted order for real space)
  double* vir_near;  // [6n] or null
  double* e_lj;      // [n] or null
  double* e_coul;    // [n] or null
  double* vir_k;     // [6n] or null
  // status word: bit0 nonfinite, bit1 shake, bit2 capacity
  int* status;
  // Verlet list with a skin (skin_on): the list holds every pair within
  // rc + skin (list_cut2 = (rc + skin)^2; without a skin list_cut2 = cut2) and
  // is rebuilt when an atom has moved more than skin / 2 since the build.
  // Between builds atoms keep their sorted order and spos holds each build
  // position plus the minimum-image displacement since.  The conditional
  // kernels of a step do their work when *rebuild_at equals the device step
  // counter *cur_step.
  double list_cut2;
  int skin_on;
  double half_skin2;        // (skin / 2)^2
  const int64_t* cur_step;  // the NHC state's step counter
  int64_t* rebuild_at;      // step of the last requested rebuild
  double *x0, *y0, *z0;     // positions at the last build (original order)
  int* list_builds;         // cell sorts (list builds) since creation
};

// the list and cells of this step are (re)built: always without a skin
__device__ __forceinline__ bool list_rebuild(const SysView& s) {
  return !s.skin_on || *s.rebuild_at == *s.cur_step;
}

// Thermo scalars reduced from per-workgroup partial rows.  Each producer
// writes rows of its own width holding consecutive slots:
//   pair   S_ELJ .. S_VIR_NEAR+5  (8)     k space  S_VIR_K .. +5   (6)
//   bonded S_VIR_BOND .. S_EANGLE (14)    KE       S_KE            (1)
//   SHAKE  S_VIR_SHAKE .. +5      (6)
enum ScalarSlot {
  S_KE = 0,
  S_ELJ = 1,
  S_ECOUL = 2,
  S_VIR_NEAR = 3,   // 6
  S_VIR_K = 9,      // 6
  S_VIR_BOND = 15,  // 6
  S_VIR_ANGLE = 21, // 6
  S_EBOND = 27,
  S_EANGLE = 28,
  S_VIR_SHAKE = 29, // 6
  S_NSLOT = 35
};
constexpr int kRowPair = 8, kRowK = 6, kRowBond = 14, kRowKe = 1, kRowShake = 6;

// One producer's rows: nrows rows of `width` doubles, row k -> slot slot0 + k;
// `groups` is how many thread groups stride the rows in the reduction.
struct RowSeg {
  const double* rows;
  int nrows, width, slot0, groups;
};
struct RowsDesc {
  RowSeg seg[5];
  int nseg;
};

// ---------------------------------------------------------------- host glue
// the message of rbmd_last_error() for this thread; returns code
int32_t api_fail(int32_t code, const char* fmt, ...);
// a context's device and stream (the context type is the C ABI's opaque rbmd_ctx)
int ctx_device(const ::rbmd_ctx* ctx);
hipStream_t ctx_stream(const ::rbmd_ctx* ctx);

// ---------------------------------------------------------------- no-alloc scratch
// bytes of the fixed per-device scratch (noalloc.hip)
constexpr size_t kScratchBytes = size_t(32) << 20;
// take the scratch's device address on `device` (current); once per device
hipError_t scratch_init(int device);
// the scratch's device address on `device` (null before scratch_init)
void* scratch_address(int device);

// ---------------------------------------------------------------- cells
void launch_bin(const SysView& s, hipStream_t st, const double* x, const double* y, const double* z);
void launch_scan(const SysView& s, hipStream_t st);
void launch_scatter(const SysView& s, hipStream_t st);
void launch_cell_sort(const SysView& s, hipStream_t st);
bool scan_in_scatter(const SysView& s);  // few cells: the scan runs inside the scatter launch
void launch_scan_scatter(const SysView& s, hipStream_t st);
void launch_nbr_build(const SysView& s, hipStream_t st);
// skin lists: flag a rebuild when an atom has moved more than skin / 2 since the
// last build, and bin every atom (the counts feed the scan of a rebuild step)
void launch_skin_check(const SysView& s, hipStream_t st);
// request a rebuild at the device step counter's value + ahead
void launch_skin_request(const SysView& s, int ahead, hipStream_t st);
void launch_fill_int(int* p, int v, int n, hipStream_t st);
void launch_fill_double(double* p, double v, int n, hipStream_t st);

// ---------------------------------------------------------------- pair
// partials: [nblocks][kRowPair] block sums (may be null when thermo == 0)
void launch_pair(const SysView& s, hipStream_t st, int thermo, int peratom, double* partials, int* nblocks);
int pair_blocks(const SysView& s);
// RBL energies over a full rc stencil (no list) and the pair thermo rows
// (energies + the mean-removed RBL virial); e_lj/e_coul per atom when peratom
void launch_rbl_energy(const SysView& s, hipStream_t st, int peratom, double* rows, int* nblocks);
int rbl_energy_blocks(const SysView& s);
// the RBL erfc step table at n distances (device arrays)
void launch_static_table(const double* dis, double* g, int n, hipStream_t st);
// fused cell-stencil real-space force (no stored list); fn* get 3 slab partials
void launch_force_cells(const SysView& s, hipStream_t st, int thermo, int peratom, double* partials, int* nblocks);
int force_cells_blocks(const SysView& s);

// ---------------------------------------------------------------- k space
// One k vector as the force pass reads it (written by the rho reduction):
// k, the prefactor 1/(V k^2) (Ewald: times e^{-k^2/4 alpha}), m = k L / 2 pi
// and rho(k); in mixed precision rr and ri hold pre Re rho and pre Im rho.
// 80 bytes, read with scalar loads.
struct KRec {
  double kx, ky, kz, pre, mx, my, mz, rr, ri, pad;
};

struct KView {
  int P;
  double S_over_P;      // S / P
  const double* kvec;   // [3P] k = 2 pi m / L (physical units)
  double* rho_part;     // [nblk_rho][2P]
  double* rho;          // [2P]
  int nblk_rho;
  int f32;              // phases' sin/cos and the force coefficient in FP32 (sums in FP64)
  KRec* krec;           // [P] records for the force pass (null: the reduction writes none)
  int ewald;            // prefactor with the Ewald Gaussian
  int rho_atoms;        // atoms per structure-factor workgroup (= n / nblk_rho rounded up to it)
};
// atoms per structure-factor workgroup for n atoms and kblocks k-vector
// blocks: 1024, halved (down to 64) while the grid has under 2048 workgroups
int rho_atoms_per_block(int n, int kblocks);
int rho_kblocks(const SysView& s, int P);  // k-vector blocks of the structure-factor grid
void launch_rho_partial(const SysView& s, const KView& k, hipStream_t st);
void launch_rho_reduce(const SysView& s, const KView& k, hipStream_t st);
void launch_kforce(const SysView& s, const KView& k, hipStream_t st, int thermo, int peratom, double* partials,
                   int* nblocks);
int kforce_blocks(const SysView& s);
void launch_kvec_slot(double* dst, const double* ring, int R, int per, const int64_t* step, hipStream_t st);

// ---------------------------------------------------------------- integrate
struct IntView {
  double dt, fmt2v, mvv2e;
  int nmol;                 // SHAKE triplets
  const int* angle_atoms;   // [3*nmol]
  // bonded incidence (CSR over atoms)
  int nb, na;
  const int* bond_atoms;
  const double *bond_k, *bond_r0;
  const int* angle_all;     // [3*na]
  const double *angle_k, *angle_t0;
  const int *inc_b_off, *inc_b;   // atom -> bonds
  const int *inc_a_off, *inc_a;   // atom -> angles
  double shake_tol;               // SHAKE convergence on the multiplier change
  double shake_d12;               // H-H distance of the rigid geometry (1 A bonds at 109.47 deg), formed on the host
  // SpecialCoulGeneral (CVFF + RBL + special bonds): atom -> the atoms of its
  // molecule, itself included, in ascending id (CSR over atoms)
  int spec_coul;
  const int *grp_off, *grp_ids;
};
void launch_first_half(const SysView& s, const IntView& iv, hipStream_t st, const double* scale, int do_bin);
void launch_shake(const SysView& s, const IntView& iv, hipStream_t st, double* partials, int* max_iter);
// launch_first_half + launch_shake in one launch, for systems whose every atom
// is in a SHAKE triplet (bit for bit the two)
void launch_first_half_shake(const SysView& s, const IntView& iv, hipStream_t st, const double* scale, int do_bin,
                             double* partials, int* max_iter);
void launch_rattle(const SysView& s, const IntView& iv, hipStream_t st);
// RATTLE of one step and the first half step + SHAKE of the next, one launch
// (systems whose every atom is constrained)
void launch_rattle_first_half_shake(const SysView& s, const IntView& iv, hipStream_t st, const double* scale,
                                    int do_bin, double* rows, int* max_iter);
void launch_bonded(const SysView& s, const IntView& iv, hipStream_t st, int thermo, double* partials, int* nblocks);
void launch_second_half(const SysView& s, const IntView& iv, hipStream_t st, int kick, int has_k, int has_b,
                        double* partials, int* nblocks);
void launch_scale_v(const SysView& s, hipStream_t st, const double* scale);
void launch_ke_rows(const SysView& s, const IntView& iv, hipStream_t st, double* rows);
void launch_unconstrained_copy(const SysView& s, hipStream_t st, const int* free_ids, int nfree, int pos);

// NHC / thermostat scalar state on the device (one thread)
struct NhcState {
  double eta[3], eta_dot[4], eta_dotdot[3], eta_mass[3];
  double t_target, ke_target;
  double tempT, tempT_sum, pressure;
  double factor;         // velocity scale of the step start (phase 1), or the last phase's
  double virial[6];
  double e_lj, e_coul, e_bond, e_angle;
  int64_t step;
  double factor_end;     // velocity scale of the step end (phases 2, 3, 4)
};
struct NhcParams {
  int tdof, mode;        // mode: 0 NHC, 1 Berendsen, 2 rescale, 3 none
  int inbuild, fix_shake;
  int n, nb, na;
  double boltz, kB, nktv2p, qqr2e;
  double t_start, t_stop, t_freq, t_period, dt, dthalf, dt4, dt8, tdrag;
  double volume_range;
  int64_t num_steps;
  int lj_only, cvff, shake_vir;
  int* shake_iter;       // SHAKE iteration maximum of the step, zeroed at the step start (phases 1, 5)
};
// phase 1: step++ and NHC half step (InitialIntegrate); 2: ComputeTempe + pressure + NHC
// (FinalIntegrate); 3: SetUp; 4: Berendsen/rescale end of step; 5: step++ only;
// 6: phase 2 then phase 1 (between two steps of one call)
// ke_only: reduce the kinetic-energy segment only (steps without thermo sums)
void launch_nhc_rows(NhcState* st_dev, const NhcParams& p, const RowsDesc& rd, int phase, hipStream_t st,
                     int ke_only = 0);

// ---------------------------------------------------------------- packs
// A pack runs the MD steps of many independent systems with one launch per
// kernel of the step: the blocks of every system sit back to back in one
// grid, and each block finds its system and its block index within that
// system's own grid, so every system runs exactly the decomposition (and the
// arithmetic) of its own launches.

// One system's arguments: what its own launches pass.
struct PackSys {
  SysView v;
  IntView iv;
  KView kv;               // kv.kvec: this step's k vectors
  NhcState* nhc;
  NhcParams np;
  RowsDesc rd;            // thermo rows, reduced by the NHC phase 2 launch
  double *rows_pair, *rows_k, *rows_bond, *rows_ke, *rows_shake;
  int* shake_iter;
  const double* kring;    // RBE k-vector ring [2][R][3P] (null: fixed k vectors)
  int ring_R;
};

// One packed launch: map[b] = system of block b, first[k] = first block of
// system k (first[nsys] = total)
struct PackLaunch {
  const PackSys* sys;
  const int* map;
  const int* first;
};

// kernels of the packed step, each with its own block map
enum PackKernel {
  PK_NHC,
  PK_KSLOT,
  PK_FIRST_HALF,
  PK_SCAN,
  PK_SCATTER,
  PK_CELL_SORT,
  PK_ONE,  // one small workgroup per system (counter resets, status gather)
  PK_SHAKE,
  PK_NBR,
  PK_PAIR,
  PK_RHO,
  PK_RHO_RED,
  PK_KFORCE,
  PK_BONDED,
  PK_SECOND,
  PK_ATOMS,  // one thread per atom (velocity scale)
  PK_RATTLE,
  PK_COUNT
};

// blocks of system s in each packed kernel (the grids of its own launches)
int pack_blocks(const SysView& s, const IntView& iv, int P, int kernel);

struct PackSig {  // launch-wide choices, equal for every system of a pack
  int pair_mode, pair_sp, f32, has_k, ewald, has_b, k_lanes_lg, rho_kb;
  int thermo;  // this step computes the virial and energy sums (else forces and KE only)
};

// Packed launches of one kernel (nblocks = total over the pack; lds = the
// largest dynamic LDS any system's own launch uses).  arg: PK_NHC the phase
// (+8: reduce the thermo rows), PK_FIRST_HALF 1 = apply the NHC scale,
// PK_SECOND the kick flag, PK_ONE 0 = zero the SHAKE iteration counters,
// 1 = copy every status word to out[k].
void launch_pack_cells(int kernel, const PackLaunch& L, int nblocks, size_t lds, hipStream_t st);
void launch_pack_pair(const PackLaunch& L, int nblocks, const PackSig& g, hipStream_t st);
void launch_pack_kspace(int kernel, const PackLaunch& L, int nblocks, const PackSig& g, hipStream_t st);
void launch_pack_integrate(int kernel, const PackLaunch& L, int nblocks, const PackSig& g, int arg, size_t lds,
                           int* out, hipStream_t st);

}  // namespace rbmd
