/*
 * oracle.h -- CPU FP64 restatement of the RBMD force path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (rbmd_amd/, include/)
 * links, loads or calls this code; it is used by tests/, by the smoke check
 * and by the cpu_baseline leg of bench.py as the checker and the CPU
 * baseline, never as the thing measured or shipped.
 *
 * Every function names the reference file:line it restates.  What pins each
 * part (tests/test_oracle_pins.py unless noted):
 *   locator            canonical order defined here; neighbour sets below
 *   neighbours         all-image brute-force counts, entry validity, no duplicates
 *   force_real         two-atom LJ and erfc-Coulomb closed forms; finite
 *                      differences of the energy (salt)
 *   RBL lists / force  same pair term; capacities and sampling are the
 *                      reference's formulas restated (parity unpinned beyond
 *                      the restatement: no closed form exists for the shell sample)
 *   rbe_S              theta-function identity vs the direct k sum; 734.5889
 *   rbe_sample         marginal frequencies vs the exact discrete Gaussian
 *   rbe_force / rho    momentum conservation; unbiased vs the Ewald k-space force
 *   ewald              Madelung constant 1.747564594633
 *   bonded             finite differences of the bonded energy; one-bond closed form
 *   md_step            velocity-Verlet time reversibility; SHAKE/RATTLE residuals;
 *                      NHC chain restated (parity unpinned beyond the restatement)
 *   reference inputs   tests/test_inputs.py (the reference's own data and JSON files)
 */
#ifndef RBMD_ORACLE_H
#define RBMD_ORACLE_H

#include <stdint.h>
#include "../include/rbmd_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

int orc_create(const rbmd_sys_desc* desc, const rbmd_params* params, orc_sys** out);
void orc_destroy(orc_sys* s);
int orc_set_state(orc_sys* s, const double* x, const double* v, const int32_t* image);
int orc_get_state(orc_sys* s, double* x, double* v, double* f, int32_t* image);
int orc_set_threads(int n);

/* locator: ContPointLocator (ContPointLocator.h:44-94, .cxx:38-92) */
void orc_locator_geometry(const orc_sys* s, int32_t dims[3], int32_t* nc, double dxdydz[3]);
int orc_locator_build(orc_sys* s);
int orc_locator_export(const orc_sys* s, int32_t* cell_of_atom, int32_t* point_ids, int32_t* lower, int32_t* upper);

/* neighbour lists (RunWorklet.cpp:296-355, 906-980) */
int orc_neighbors_build(orc_sys* s, int32_t cap);
int orc_neighbors_capacity(const orc_sys* s);
int orc_neighbors_export(const orc_sys* s, int32_t cap, int32_t* ids, double* offsets, int32_t* num);
/* brute-force O(N^2) all-image enumeration of pairs with 1e-4 < d^2 < rc^2 - 1e-4; counts per atom */
int orc_bruteforce_counts(const orc_sys* s, double rc, int32_t* num);

/* RBL erfc step table: ExecStaticTable::Extract + TableGnearValue (ExecStaticTable.h:96-135) */
int orc_static_table_gnear(int32_t n, const double* dis, double* g);

/* real-space force over the current neighbour lists (RunWorklet.cpp:2767-2985, 982-1266);
 * RBL also forms the energies over a full rc list (ExecutionMD.cpp:670-750) */
int orc_force_real(orc_sys* s, double* f, double* vir, double* elj, double* ecoul);

/* RBE (RBEPSample.h:40-177; RunWorklet.cpp:2086-2141, 2209-2291) */
int orc_rbe_set_kvectors(orc_sys* s, int32_t p, const double* m);
int orc_rbe_sample(double alpha, const double box[3], int32_t p, uint64_t seed, uint64_t stream, double* m);
double orc_rbe_S(double alpha, const double box[3]);
double orc_rbe_S_direct(double alpha, const double box[3], int32_t mmax);
int orc_rbe_structure_factor(orc_sys* s, double* rho);
int orc_rbe_force(orc_sys* s, double* f, double* vir);

/* Ewald over the kmax grid (RunWorklet.cpp:2492-2540; ExecutionMD.cpp:405-443, 544-558) */
int orc_ewald_force(orc_sys* s, const int32_t kmax[3], double* f, double* vir, double* energy_total);
double orc_self_energy_total(const orc_sys* s);

/* bonded (MolecularWorklet.h:58-397) */
int orc_force_bonded(orc_sys* s, double* f, double* ebond_sum, double* eangle_sum);

/* MD loop (ExecutionNVT.cpp:60-135, 1106-1276) */
int orc_md_init(orc_sys* s);
int orc_md_step(orc_sys* s, int64_t n);
int orc_md_thermo(const orc_sys* s, rbmd_thermo* out);
int orc_md_energy(orc_sys* s);

#ifdef __cplusplus
}
#endif
#endif
