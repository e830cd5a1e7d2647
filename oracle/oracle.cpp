// oracle.cpp -- CPU FP64 restatement of the RBMD force path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
// baseline of bench.py.  The product never links or loads it.
//
// The reference stores Real = vtkm::FloatDefault (FP32 in its default build,
// framework/Types.h:48).  This restatement runs every formula in FP64 with
// the reference's operation order, so that the FP64 HIP path can be checked
// to 1e-10 and the mixed path to 1e-5.  Each function cites the reference
// lines it follows.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPi = 3.14159265358979323846;

struct Units {
  double kB, fmt2v, mvv2e, qqr2e, boltz, nktv2p;
};

// ExecutionMD::InitParameters, ExecutionMD.cpp:329-360
Units units_of(int u) {
  Units r;
  if (u == RBMD_UNIT_REAL) {
    r.kB = 1.9872067 * std::pow(10.0, -3);
    r.fmt2v = 4.186 * std::pow(10.0, -4);
    r.mvv2e = 1.0 / (4.186 * std::pow(10.0, -4));
    r.qqr2e = 332.06371;
    r.boltz = 0.0019872067;
    r.nktv2p = 68568.415;
  } else if (u == RBMD_UNIT_METAL) {
    r.kB = 8.617343e-5;
    r.fmt2v = 1.0 / 1.0364269e-4;
    r.mvv2e = 1.0364269e-4;
    r.qqr2e = 14.399645;
    r.boltz = 8.617343e-5;
    r.nktv2p = 1.6021765e6;
  } else {
    r.kB = r.fmt2v = r.mvv2e = r.qqr2e = r.boltz = r.nktv2p = 1.0;
  }
  return r;
}

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double& at(V3& a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }
inline double at(const V3& a, int d) { return d == 0 ? a.x : (d == 1 ? a.y : a.z); }

// splitmix64 counter stream: the seeded replacement of
// std::mt19937(std::random_device{}()) in RandomValue (framework/math/Utiles.h:32-39).
struct Rng {
  uint64_t s;
  Rng(uint64_t seed, uint64_t stream) : s(seed ^ (stream * 0xD1B54A32D192ED03ull)) { next(); }
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

struct orc_sys {
  rbmd_params p;
  Units u;
  int n = 0, ntypes = 0;
  double lo[3], hi[3], box[3], volume = 0;
  std::vector<int> type, mol;
  std::vector<double> q, m, eps, sig;
  bool has_special = false;
  std::vector<int> sp_off, sp_ids;
  std::vector<double> sp_w;
  int nb = 0, na = 0;
  std::vector<int> bond_atoms, angle_atoms;
  std::vector<double> bond_k, bond_r0, angle_k, angle_t0;
  // molecule groups for SpecialCoulGeneral (ModelFileInitCondition.cpp:698-736)
  std::vector<int> grp_off, grp_ids;

  // state
  std::vector<V3> x, v, f, old_x, old_v;
  std::vector<int> img;  // 3 per atom

  // locator
  int dims[3] = {1, 1, 1}, nc = 1;
  double dxdydz[3] = {1, 1, 1};
  std::vector<int> cell_of, point_ids, lower, upper;

  // neighbour lists
  int cap = 0;
  std::vector<int> nb_ids, nb_num;  // VERLET: num[n]; RBL: num[2n]
  std::vector<V3> nb_off;
  int rs_num = 0, pice_num = 1;

  // k-set
  int P = 0;
  std::vector<V3> kset;

  // per-atom outputs
  std::vector<V3> f_near, f_far, f_bond, f_angle, f_spec;
  std::vector<double> vir_near, vir_far, vir_bond, vir_angle, vir_shake, vir_spec;
  std::vector<double> e_lj, e_coul;

  // thermo
  int64_t step = 0;
  int tdof = 0;
  double tempT = 0, tempT_sum = 0, pressure = 0;
  double e_lj_avg = 0, e_coul_avg = 0, e_long_avg = 0, e_bond_avg = 0, e_angle_avg = 0;
  int64_t energy_step = -1;
  double virial[6] = {0, 0, 0, 0, 0, 0};
  int shake_max_iter = 0, status_flags = 0;

  // NHC (ExecutionNVT.cpp:158-186)
  double t_start, t_stop, t_period, t_target = 0, ke_target = 0, t_freq;
  double eta[3], eta_dot[4], eta_dotdot[3], eta_mass[3];
  int mtchain = 3, nc_tchain = 1, eta_mass_flag = 1;
  double dthalf, dt4, dt8, drag, tdrag_factor, factor_eta = 1;
  bool initialized = false;
};

namespace {

// ExecPointLocator::MinDistanceVec, ExecPointLocator.h:225-260 (single image)
inline V3 min_image(const orc_sys* s, V3 a, V3 b) {
  V3 v = a - b;
  for (int d = 0; d < 3; ++d) {
    double& c = at(v, d);
    if (std::fabs(c) > s->box[d] * 0.5) c -= (c > 0 ? s->box[d] : -s->box[d]);
  }
  return v;
}

// vtkm Vec3f -> Id3 conversion: truncation toward zero
inline int trunc_id(double v) { return (int)v; }

// ContPointLocator::SetRange/SetCutOff/SetRs, ContPointLocator.h:44-72
void geometry(orc_sys* s) {
  double rc = s->p.cut_off;
  double rs = s->p.neighbor_type == RBMD_NEIGHBOR_RBL ? s->p.r_core : s->p.cut_off;
  s->nc = (int)std::ceil(rc / rs);
  for (int d = 0; d < 3; ++d) {
    s->dims[d] = trunc_id((s->hi[d] - s->lo[d]) / (rc / s->nc));
    if (s->dims[d] < 1) s->dims[d] = 1;
    // ExecPointLocator ctor: _dxdydz((max - _min) / _dims)
    s->dxdydz[d] = (s->hi[d] - s->lo[d]) / s->dims[d];
  }
}

inline int ncells(const orc_sys* s) { return s->dims[0] * s->dims[1] * s->dims[2]; }

// BinPointsWorklet + SortByKey + Upper/LowerBounds, ContPointLocator.cxx:38-92.
// Within-cell order: ascending atom id (stable sort of the identity ids).
void build_locator(orc_sys* s) {
  int n = s->n, C = ncells(s);
  s->cell_of.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    int ijk[3];
    V3 c = s->x[i];
    for (int d = 0; d < 3; ++d) {
      int t = trunc_id((at(c, d) - s->lo[d]) / s->dxdydz[d]);
      t = std::max(t, 0);
      t = std::min(t, s->dims[d] - 1);
      ijk[d] = t;
    }
    s->cell_of[i] = ijk[0] + ijk[1] * s->dims[0] + ijk[2] * s->dims[0] * s->dims[1];
  }
  std::vector<int> cnt(C + 1, 0);
  for (int i = 0; i < n; ++i) cnt[s->cell_of[i] + 1]++;
  for (int c = 0; c < C; ++c) cnt[c + 1] += cnt[c];
  s->lower.assign(C, 0);
  s->upper.assign(C, 0);
  for (int c = 0; c < C; ++c) {
    s->lower[c] = cnt[c];
    s->upper[c] = cnt[c + 1];
  }
  s->point_ids.assign(n, 0);
  std::vector<int> cur(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n; ++i) s->point_ids[cur[s->cell_of[i]]++] = i;
}

// ExecPointLocator::InWhichCell (unclamped), ExecPointLocator.h:271-274.  A
// NaN coordinate maps to cell 0 and the index is held within +-2e9 so the
// integer conversion is defined (both only reachable after a failed step).
inline void in_which_cell(const orc_sys* s, V3 p, int ijk[3]) {
  for (int d = 0; d < 3; ++d) {
    double t = (at(p, d) - s->lo[d]) / s->dxdydz[d];
    if (!(t == t)) t = 0.0;
    ijk[d] = trunc_id(std::max(std::min(t, 2.0e9), -2.0e9));
  }
}

// Visit the (2nc+1)^3 stencil of atom i in the reference's order
// (x offset outer, z inner; RunWorklet.cpp:322-331); f(j, offset, r_ij, d2).
template <class F>
void visit_stencil(const orc_sys* s, int i, F&& fn) {
  V3 pi = s->x[i];
  int ci[3];
  in_which_cell(s, pi, ci);
  int nc = s->nc;
  for (int a = -nc; a <= nc; ++a)
    for (int b = -nc; b <= nc; ++b)
      for (int c = -nc; c <= nc; ++c) {
        int nb[3] = {ci[0] + a, ci[1] + b, ci[2] + c};
        int w[3];
        V3 off;
        bool inside = true;
        for (int d = 0; d < 3; ++d) {
          w[d] = (nb[d] + s->dims[d]) % s->dims[d];  // PeriodicIndexOffset
          at(off, d) = s->dxdydz[d] * (double)(w[d] - nb[d]);  // PeriodicCoordOffset
          // an atom more than one box away wraps to no cell (the reference
          // would index out of range); only reachable after a failed step
          inside = inside && w[d] >= 0;
        }
        if (!inside) continue;
        int cell = w[0] + w[1] * s->dims[0] + w[2] * s->dims[0] * s->dims[1];
        for (int pidx = s->lower[cell]; pidx < s->upper[cell]; ++pidx) {
          int j = s->point_ids[pidx];
          V3 pj = s->x[j] - off;
          V3 r = pj - pi;
          double d2 = r.x * r.x + r.y * r.y + r.z * r.z;
          fn(j, off, d2);
        }
      }
}

// ComputeNeighboursWorklet, RunWorklet.cpp:296-355
int build_verlet(orc_sys* s, int cap) {
  int n = s->n;
  s->cap = cap;
  s->nb_ids.assign((size_t)n * cap, -1);
  s->nb_off.assign((size_t)n * cap, V3{0, 0, 0});
  s->nb_num.assign(n, 0);
  double rc2 = s->p.cut_off * s->p.cut_off;
  int overflow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : overflow)
  for (int i = 0; i < n; ++i) {
    int idx = 0;
    visit_stencil(s, i, [&](int j, V3 off, double d2) {
      if (rc2 - d2 > 0.0001 && d2 > 0.0001) {
        if (idx < cap) {
          s->nb_ids[(size_t)i * cap + idx] = j;
          s->nb_off[(size_t)i * cap + idx] = off;
        } else {
          overflow = 1;
        }
        idx++;
      }
    });
    s->nb_num[i] = idx;
  }
  return overflow ? RBMD_ERR_CAPACITY : RBMD_OK;
}

// RBL capacities, ExecutionMD.cpp:582-608
void rbl_capacity(orc_sys* s) {
  double rho = s->p.rbl_rho > 0 ? s->p.rbl_rho : s->n / s->volume;
  double rc = s->p.cut_off, rs = s->p.r_core;
  double pif = (double)(float)kPi;
  double coeff_rcs = 1.0 + (0.05 / rho - 0.05);
  long id_coeff = std::lround(coeff_rcs);
  long rs_num = (long)(id_coeff * rho * std::ceil(4.0 / 3.0 * pif * rs * rs * rs)) + 1;
  long rc_num = (long)(id_coeff * rho * std::ceil(4.0 / 3.0 * pif * rc * rc * rc)) + 1;
  long rcs_num = rc_num - rs_num;
  double random_num = s->p.neighbor_sample_num;
  double random_rate = random_num / (double)rcs_num;
  s->pice_num = (int)std::ceil(1.0 / random_rate);
  // The reference sizes the core slots at rs_num (about the mean core count + 1)
  // and silently overwrites past it (hazard Q7); here the core region is twice
  // that plus 32 slots and an overflow is reported.  pice_num is unchanged.
  s->rs_num = (int)(2 * rs_num + 32);
}

// ComputeRBLNeighboursOnceWorklet, RunWorklet.cpp:906-980.  Slots [0, rs_num)
// hold the core, slots [rs_num, cap) the kept shell atoms.  The reference writes
// past its capacity silently (hazard Q7); here an overflow is reported.
int build_rbl(orc_sys* s, int cap) {
  int n = s->n;
  rbl_capacity(s);
  s->cap = cap;
  s->nb_ids.assign((size_t)n * cap, -1);
  s->nb_off.assign((size_t)n * cap, V3{0, 0, 0});
  s->nb_num.assign((size_t)2 * n, 0);
  double rc2 = s->p.cut_off * s->p.cut_off, rs2 = s->p.r_core * s->p.r_core;
  int rs_num = s->rs_num, pice = s->pice_num;
  int overflow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : overflow)
  for (int i = 0; i < n; ++i) {
    int ishell = 0, irand = 0, irs = 0;
    visit_stencil(s, i, [&](int j, V3 off, double d2) {
      if (rc2 - d2 > 0.0001 && d2 - rs2 > 0.0001) {
        if (ishell % pice == 0) {
          int slot = rs_num + irand;
          if (slot < cap) {
            s->nb_ids[(size_t)i * cap + slot] = j;
            s->nb_off[(size_t)i * cap + slot] = off;
          } else {
            overflow = 1;
          }
          irand++;
        }
        ishell++;
      } else if (rs2 - d2 > 0.0001 && d2 > 0.0001) {
        if (irs < rs_num) {
          s->nb_ids[(size_t)i * cap + irs] = j;
          s->nb_off[(size_t)i * cap + irs] = off;
        } else {
          overflow = 1;
        }
        irs++;
      }
    });
    s->nb_num[2 * i] = irs;
    s->nb_num[2 * i + 1] = irand;
  }
  return overflow ? RBMD_ERR_CAPACITY : RBMD_OK;
}

int build_neighbors(orc_sys* s, int cap) {
  if (s->p.neighbor_type == RBMD_NEIGHBOR_RBL) return build_rbl(s, cap);
  return build_verlet(s, cap);
}

int default_cap(const orc_sys* s) {
  // bounded capacity instead of the reference's N x N scratch (ExecutionMD.cpp:885-898)
  double rho = s->n / s->volume;
  double rc = s->p.cut_off;
  double c = rho * 4.0 / 3.0 * kPi * rc * rc * rc;
  int cap = (int)(2.0 * c + 64);
  if (s->p.neighbor_type == RBMD_NEIGHBOR_RBL) {
    rbl_capacity(const_cast<orc_sys*>(s));
    cap = s->rs_num + (int)(2.0 * c / s->pice_num + 64);
  }
  return std::min(cap, std::max(s->n, 1) * (2 * s->nc + 1) * (2 * s->nc + 1) * (2 * s->nc + 1));
}

// ExecForceFunction::Gnear, ExecForceFunction.h:682-689 (erf from libm)
inline double gnear(double alpha, double dis) {
  double erfcx = std::sqrt(alpha) * dis;
  double expx = -alpha * dis * dis;
  return (1.0 - std::erf(erfcx)) / (dis * dis) + 2 * std::sqrt(alpha) * std::exp(expx) / (std::sqrt(kPi) * dis);
}

// ERF step table at alpha = 0.1: the reference's tabulated G values
// (framework/model/ERFTable.h erf_gnear_leq2_v / erf_gnear_geq2_v, extracted by
// tests/golden/make_erf_table.py), indexed as ExecStaticTable::Extract and
// TableGnearValue do (ExecStaticTable.h:96-135): the float bits of the distance
// masked with 2^26 - 2^15 (3 exponent + 8 mantissa bits) and shifted by 15,
// minus the first index of the table chosen by dis < 2.  The reference reads
// out of bounds for dis below 0.0998 or at 12 and beyond (hazard Q3); here the
// bucket is clamped to the table (the product rejects rc >= 12 in table mode).
#define RBMD_ERF_STORAGE static const
#include "erf_table.inc"
#undef RBMD_ERF_STORAGE
inline double gnear_table(double dis) {
  const float fd = (float)dis;
  uint32_t bits;
  std::memcpy(&bits, &fd, 4);
  const int idx = (int)((bits & ((1u << 26) - (1u << 15))) >> 15);
  if (fd < 2.0f) {
    int k = idx - RBMD_ERF_LEQ2_INDEX0;
    k = k < 0 ? 0 : (k >= RBMD_ERF_LEQ2_N ? RBMD_ERF_LEQ2_N - 1 : k);
    return (double)kErfGnearLeq2[k];
  }
  int k = idx - RBMD_ERF_GEQ2_INDEX0;
  k = k < 0 ? 0 : (k >= RBMD_ERF_GEQ2_N ? RBMD_ERF_GEQ2_N - 1 : k);
  return (double)kErfGnearGeq2[k];
}

// ExecForceFunction::ComputeLJForce, ExecForceFunction.h:412-435
inline V3 lj_force(V3 r, double ei, double ej, double si, double sj, double cut) {
  double d2 = r.x * r.x + r.y * r.y + r.z * r.z;
  if (d2 < cut * cut && d2 > 0.0001) {
    double sij = (si + sj) / 2;
    double s6 = sij * sij * sij * sij * sij * sij;
    double d6 = d2 * d2 * d2;
    double sd6 = s6 / d6;
    return (-24 * std::sqrt(ei * ej) * ((2 * sd6 - 1) * sd6) / d2) * r;
  }
  return {0, 0, 0};
}

// ExecForceFunction::ComputeLJForceRcs, ExecForceFunction.h:437-462
inline V3 lj_force_rcs(V3 r, double ei, double ej, double si, double sj, double rc, double rs) {
  double d2 = r.x * r.x + r.y * r.y + r.z * r.z;
  if (d2 < rc * rc && d2 > rs * rs) {
    double sij = (si + sj) / 2;
    double s6 = sij * sij * sij * sij * sij * sij;
    double d6 = d2 * d2 * d2;
    double sd6 = s6 / d6;
    return (-24 * std::sqrt(ei * ej) * ((2 * sd6 - 1) * sd6) / d2) * r;
  }
  return {0, 0, 0};
}

// ExecForceFunction::ComputePotentialEn_fix, ExecForceFunction.h:818-841
inline double lj_energy(V3 r, double ei, double ej, double si, double sj, double cut) {
  double d2 = r.x * r.x + r.y * r.y + r.z * r.z;
  double e = 0;
  if (d2 < cut * cut && d2 > 0.0001) {
    double sij = (si + sj) / 2;
    double s6 = sij * sij * sij * sij * sij * sij;
    double d6 = d2 * d2 * d2;
    double eij = std::sqrt(ei * ej);
    e = 4 * eij * (s6 / d6 - 1) * (s6 / d6);
  }
  return 0.5 * e;
}

inline void add_vir(double* v, V3 r, V3 fp, double scale) {
  V3 f = scale * fp;
  v[0] += r.x * f.x;
  v[1] += r.y * f.y;
  v[2] += r.z * f.z;
  v[3] += r.x * f.y;
  v[4] += r.x * f.z;
  v[5] += r.y * f.z;
}
inline void set_vir(double* v, V3 r, V3 fp, double scale) {
  for (int k = 0; k < 6; ++k) v[k] = 0;
  add_vir(v, r, fp, scale);
}

double special_weight(const orc_sys* s, int i, int j) {
  double w = 1.0;
  for (int k = s->sp_off[i]; k < s->sp_off[i + 1]; ++k)
    if (s->sp_ids[k] == j) w = s->sp_w[k];  // last match wins
  return w;
}

// Plain real space: ComputeNearForceVerletVirialWorklet (RunWorklet.cpp:2890-2985) or,
// with special bonds, ComputeNearForceVerletWeightVirialWorklet1 (:2767-2888);
// LJ/CUT alone: ComputeLJForceVerletWorklet (:498-575).
void force_verlet(orc_sys* s) {
  int n = s->n;
  double rc = s->p.cut_off, alpha = s->p.alpha, qqr2e = s->u.qqr2e;
  bool weighted = s->has_special && s->p.special_bonds;
  bool lj_only = s->p.force_field == RBMD_FF_LJ_CUT;
  s->f_near.assign(n, V3{0, 0, 0});
  s->vir_near.assign((size_t)6 * n, 0.0);
  s->e_lj.assign(n, 0.0);
  s->e_coul.assign(n, 0.0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    V3 pi = s->x[i];
    int ti = s->type[i];
    double ei = s->eps[ti], si = s->sig[ti], qi = s->q[i];
    V3 fsum{0, 0, 0};
    double vir[6] = {0, 0, 0, 0, 0, 0};
    double pelj = 0, pecoul = 0;
    for (int k = 0; k < s->nb_num[i]; ++k) {
      int j = s->nb_ids[(size_t)i * s->cap + k];
      V3 pj = s->x[j] - s->nb_off[(size_t)i * s->cap + k];
      V3 r = min_image(s, pj, pi);
      int tj = s->type[j];
      double ej = s->eps[tj], sj = s->sig[tj], qj = s->q[j];
      V3 flj = lj_force(r, ei, ej, si, sj, rc);
      double elj = lj_energy(r, ei, ej, si, sj, rc);
      if (lj_only) {
        fsum = fsum + flj;
        add_vir(vir, r, flj, 0.5);
        pelj += elj;
        continue;
      }
      double d2 = r.x * r.x + r.y * r.y + r.z * r.z;
      double dis = std::sqrt(d2);
      bool in = d2 < rc * rc && d2 > 0.0001;
      if (!weighted) {
        V3 fc{0, 0, 0};
        double ec = 0;
        if (in) {
          fc = qqr2e * ((-qi * qj * gnear(alpha, dis)) * r / dis);
          ec = qqr2e * (0.5 * qi * qj * (1.0 - std::erf(std::sqrt(alpha) * dis)) / dis);
        }
        V3 fp = fc + flj;
        fsum = fsum + fp;
        add_vir(vir, r, fp, 0.5);  // ComputePairVirial_fix
        pelj += elj;
        pecoul += ec;
      } else {
        double w = special_weight(s, i, j);
        V3 fcf{0, 0, 0}, fc{0, 0, 0};
        double ec = 0, ecf = 0;
        if (in) {
          double d3 = dis * dis * dis;
          fcf = (-qi * qj) * r / d3;
          fc = (-qi * qj * gnear(alpha, dis)) * r / dis;
          ec = 0.5 * qi * qj * (1.0 - std::erf(std::sqrt(alpha) * dis)) / dis;
          ecf = 0.5 * qi * qj / dis;
        }
        fc = fc - (1 - w) * fcf;
        ec = qqr2e * (ec - (1 - w) * ecf);
        V3 fp = qqr2e * fc + w * flj;
        fsum = fsum + fp;
        add_vir(vir, r, fp, -0.5);  // ComputePairVirial
        pelj += w * elj;
        pecoul += ec;
      }
    }
    s->f_near[i] = fsum;
    for (int k = 0; k < 6; ++k) s->vir_near[(size_t)6 * i + k] = vir[k];
    s->e_lj[i] = pelj;
    s->e_coul[i] = pecoul;
  }
}

// RBL real space: ComputeNearForceRBLERFWorklet (RunWorklet.cpp:982-1093) or
// ComputeNearForceRBLERFSpecialBondsWorklet (:1095-1266), then the mean removal
// of ExecutionMD.cpp:663-667 (SumRBLCorrForce/Virial, RunWorklet.cpp:1911-1947).
// The virial keeps the reference's assignment of the running-sum force (hazard Q5).
void force_rbl(orc_sys* s) {
  int n = s->n;
  double rc = s->p.cut_off, rs = s->p.r_core, qqr2e = s->u.qqr2e, alpha = s->p.alpha;
  bool weighted = s->has_special && s->p.special_bonds;
  bool table = s->p.erfc_mode == RBMD_ERFC_TABLE;
  double pice = (double)s->pice_num;
  std::vector<V3> corr(n);
  std::vector<double> cvir((size_t)6 * n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    V3 pi = s->x[i];
    int ti = s->type[i];
    double ei = s->eps[ti], si = s->sig[ti], qi = s->q[i];
    V3 rs_lj{0, 0, 0}, rcs_lj{0, 0, 0}, rs_el{0, 0, 0}, rcs_el{0, 0, 0};
    double rs_vir[6] = {0, 0, 0, 0, 0, 0}, rcs_vir[6] = {0, 0, 0, 0, 0, 0};
    auto pair = [&](int j, V3 off, int flag) {
      V3 pj = s->x[j] - off;
      V3 r = min_image(s, pj, pi);
      int tj = s->type[j];
      double ej = s->eps[tj], sj = s->sig[tj], qj = s->q[j];
      double dis = std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
      double d2 = r.x * r.x + r.y * r.y + r.z * r.z;
      double g = table ? gnear_table(dis) : gnear(alpha, dis);
      double w = weighted ? special_weight(s, i, j) : 1.0;
      if (flag == 1) {
        if (d2 < rc * rc && d2 > 0.0001) rs_el = rs_el + (-qi * qj * g) * r / dis;  // ComputeNearEnergyForceERF
        rs_lj = rs_lj + w * lj_force(r, ei, ej, si, sj, rs);
        set_vir(rs_vir, r, rs_lj + qqr2e * rs_el, -0.5);
      } else {
        bool inshell = weighted ? (d2 < rc * rc && d2 > rs * rs) : (d2 < rc * rc && d2 > 0.0001);
        if (inshell) rcs_el = rcs_el + pice * ((-qi * qj * g) * r / dis);
        rcs_lj = rcs_lj + w * (pice * lj_force_rcs(r, ei, ej, si, sj, rc, rs));
        set_vir(rcs_vir, r, rcs_lj + qqr2e * rcs_el, -0.5);
      }
    };
    for (int k = 0; k < s->nb_num[2 * i]; ++k)
      pair(s->nb_ids[(size_t)i * s->cap + k], s->nb_off[(size_t)i * s->cap + k], 1);
    for (int k = 0; k < s->nb_num[2 * i + 1]; ++k) {
      size_t slot = (size_t)i * s->cap + s->rs_num + k;
      pair(s->nb_ids[slot], s->nb_off[slot], 2);
    }
    corr[i] = (rs_lj + rcs_lj) + qqr2e * (rs_el + rcs_el);
    for (int k = 0; k < 6; ++k) cvir[(size_t)6 * i + k] = rs_vir[k] + rcs_vir[k];
  }
  V3 mean{0, 0, 0};
  double vmean[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    mean = mean + corr[i];
    for (int k = 0; k < 6; ++k) vmean[k] += cvir[(size_t)6 * i + k];
  }
  mean = mean / (double)n;
  for (int k = 0; k < 6; ++k) vmean[k] /= n;
  s->f_near.assign(n, V3{0, 0, 0});
  s->vir_near.assign((size_t)6 * n, 0.0);
  for (int i = 0; i < n; ++i) {
    s->f_near[i] = corr[i] - mean;
    for (int k = 0; k < 6; ++k) s->vir_near[(size_t)6 * i + k] = cvir[(size_t)6 * i + k] - vmean[k];
  }
}

// RBL energies: ExecutionMD::ComputeLJCoulEnergy (ExecutionMD.cpp:670-750)
// lists every pair with rc^2 - d^2 > 1e-4 and d^2 > 1e-4 over the RBL
// locator's stencil (ComputeNeighbours, RunWorklet.cpp:296-355) and sums the
// analytic LJ + erfc energies of LJCoulEnergyVerletVirialWorklet (:3145-3221)
// or, with special bonds, SpecialLJCoulEnergyVerletWeightVirialWorklet
// (:3047-3143): per atom e_lj = sum w E_LJ / 2, e_coul = qqr2e sum (E_erfc - (1-w) q q / 2r).
void rbl_energy(orc_sys* s) {
  int n = s->n;
  double rc = s->p.cut_off, rc2 = rc * rc, alpha = s->p.alpha, qqr2e = s->u.qqr2e;
  bool weighted = s->has_special && s->p.special_bonds;
  s->e_lj.assign(n, 0.0);
  s->e_coul.assign(n, 0.0);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    V3 pi = s->x[i];
    int ti = s->type[i];
    double ei = s->eps[ti], si = s->sig[ti], qi = s->q[i];
    double pelj = 0, pecoul = 0;
    visit_stencil(s, i, [&](int j, V3 off, double d2) {
      if (!(rc2 - d2 > 0.0001 && d2 > 0.0001)) return;
      V3 r = min_image(s, s->x[j] - off, pi);
      int tj = s->type[j];
      double elj = lj_energy(r, ei, s->eps[tj], si, s->sig[tj], rc);
      double dd2 = r.x * r.x + r.y * r.y + r.z * r.z;
      double dis = std::sqrt(dd2);
      bool in = dd2 < rc2 && dd2 > 0.0001;
      double qj = s->q[j];
      double ec = in ? 0.5 * qi * qj * (1.0 - std::erf(std::sqrt(alpha) * dis)) / dis : 0.0;
      if (weighted) {
        double w = special_weight(s, i, j);
        double ecf = in ? 0.5 * qi * qj / dis : 0.0;
        pelj += w * elj;
        pecoul += qqr2e * (ec - (1 - w) * ecf);
      } else {
        pelj += elj;
        pecoul += qqr2e * ec;
      }
    });
    s->e_lj[i] = pelj;
    s->e_coul[i] = pecoul;
  }
}

// SpecialCoulGeneral for the RBL water path, RunWorklet.cpp:2012-2084
// (332.06371 hard-coded; the virial accumulates into an unset output in the
// reference, hazard Q24 -- here it starts from zero).
void force_special_coul(orc_sys* s) {
  int n = s->n;
  s->f_spec.assign(n, V3{0, 0, 0});
  s->vir_spec.assign((size_t)6 * n, 0.0);
  if (s->grp_off.empty()) return;
  for (int i = 0; i < n; ++i) {
    int b = s->grp_off[i], e = s->grp_off[i + 1];
    if (e - b < 3) continue;
    V3 fsum{0, 0, 0};
    for (int k = b; k < e; ++k) {
      int j = s->grp_ids[k];
      if (j == i) continue;
      V3 r = min_image(s, s->x[i], s->x[j]);
      double dis = std::sqrt(dot(r, r));
      double fcomp = -332.06371 * s->q[i] * s->q[j] / std::pow(dis, 3);
      double w = special_weight(s, i, j);
      V3 cp = ((1.0 - w) * fcomp) * r;
      fsum = fsum + cp;
      add_vir(&s->vir_spec[(size_t)6 * i], r, cp, 0.5);
    }
    s->f_spec[i] = fsum;
  }
}

// Compute_H / Compute_S, RBEPSample.h:40-63 = ExecForceFunction.h:774-797
void compute_H(double alpha, const double box[3], double H[3]) {
  for (int d = 0; d < 3; ++d) {
    double factor = -(alpha * box[d] * box[d]);
    H[d] = 0;
    for (int m = -10; m <= 10; ++m) H[d] += std::exp(m * m * factor);
    H[d] *= std::sqrt(-factor / kPi);
  }
}
double compute_S(double alpha, const double box[3]) {
  double H[3];
  compute_H(alpha, box, H);
  return H[0] * H[1] * H[2] - 1;
}

// RBEPSAMPLE::FetchSample_1D (Box-Muller), RBEPSample.h:117-143
double fetch_1d(Rng& g, double mu, double sigma) {
  double U1;
  do {
    U1 = g.uniform();
  } while (U1 < 1e-6);
  double U2 = g.uniform();
  return sigma * std::sqrt(-2.0 * std::log(U1)) * std::cos(2 * kPi * U2) + mu;
}

double dist_P(double x, int d, double alpha, const double box[3], const double H[3]) {
  double t = 2 * kPi * x / box[d];
  return std::exp(-(t * t) / (4 * alpha)) / H[d];
}

double dist_q(double x, int d, double alpha, const double box[3]) {
  double s = std::sqrt(alpha * box[d] * box[d] / (kPi * kPi));
  if (x == 0) return std::erf((1.0 / 2) / s);
  return (std::erf(((1.0 / 2) + std::fabs(x)) / s) - std::erf((std::fabs(x) - (1.0 / 2)) / s)) / 2;
}

// RBEPSAMPLE::MH_Algorithm, RBEPSample.h:65-87
double mh_step(Rng& g, double m, double sigma, int d, double alpha, const double box[3], const double H[3]) {
  double x = fetch_1d(g, 0.0, sigma);
  double mw = std::round(x);
  double prob = (dist_P(mw, d, alpha, box, H) / dist_P(m, d, alpha, box, H)) *
                (dist_q(m, d, alpha, box) / dist_q(mw, d, alpha, box));
  prob = std::min(prob, 1.0);
  double u = g.uniform();
  if (u <= prob) m = mw;
  return m;
}

// RBEPSAMPLE::Fetch_P_Sample, RBEPSample.h:145-176; sigma_k from PreForce (ExecutionNVT.cpp:336-345)
void sample_kset(double alpha, const double box[3], int P, uint64_t seed, uint64_t stream, double* out) {
  Rng g(seed, stream);
  double H[3];
  compute_H(alpha, box, H);
  double sig[3];
  for (int d = 0; d < 3; ++d) sig[d] = std::sqrt(alpha / 2.0) * box[d] / kPi;
  double X0[3];
  do {
    for (int d = 0; d < 3; ++d) X0[d] = std::round(fetch_1d(g, 0.0, sig[d]));
  } while (std::fabs(X0[0]) < 1e-6 && std::fabs(X0[1]) < 1e-6 && std::fabs(X0[2]) < 1e-6);
  for (int d = 0; d < 3; ++d) out[d] = X0[d];
  for (int i = 1; i < P; i++) {
    double X1[3];
    for (int d = 0; d < 3; ++d) X1[d] = mh_step(g, X0[d], sig[d], d, alpha, box, H);
    for (int d = 0; d < 3; ++d) out[3 * i + d] = X1[d];
    for (int d = 0; d < 3; ++d) X0[d] = X1[d];
    if (std::fabs(X1[0]) < 1e-6 && std::fabs(X1[1]) < 1e-6 && std::fabs(X1[2]) < 1e-6) {
      i = i - 1;
      continue;
    }
  }
}

inline V3 kvec(const orc_sys* s, V3 m) {
  return {2 * kPi * m.x / s->box[0], 2 * kPi * m.y / s->box[1], 2 * kPi * m.z / s->box[2]};
}

// Neumaier-compensated sum
struct KSum {
  double s = 0, c = 0;
  void add(double v) {
    double t = s + v;
    if (std::fabs(s) >= std::fabs(v))
      c += (s - t) + v;
    else
      c += (v - t) + s;
    s = t;
  }
  double get() const { return s + c; }
};

// ComputePnumberChargeStructureFactor + ReduceByKey, RunWorklet.cpp:2209-2291, ExecutionMD.cpp:469-501
void rbe_rho(const orc_sys* s, std::vector<double>& rho) {
  int P = s->P;
  rho.assign((size_t)2 * P, 0.0);
#pragma omp parallel for schedule(static)
  for (int l = 0; l < P; ++l) {
    V3 k = kvec(s, s->kset[l]);
    KSum re, im;
    for (int i = 0; i < s->n; ++i) {
      double ph = dot(k, s->x[i]);
      re.add(s->q[i] * std::cos(ph));
      im.add(s->q[i] * std::sin(ph));
    }
    rho[2 * l] = re.get();
    rho[2 * l + 1] = im.get();
  }
}

// ComputeNewRBEForceWorklet, RunWorklet.cpp:2086-2141, with ComputeRBEForceSum_fix
// (ExecForceFunction.h:748-762) and ComputeKspaceVirial (:397-410); then UnitRescale(qqr2e).
void rbe_force(orc_sys* s, const std::vector<double>& rho) {
  int n = s->n, P = s->P;
  double S = compute_S(s->p.alpha, s->box);
  s->f_far.assign(n, V3{0, 0, 0});
  s->vir_far.assign((size_t)6 * n, 0.0);
  std::vector<V3> kk(P);
  std::vector<double> k2(P);
  for (int l = 0; l < P; ++l) {
    kk[l] = kvec(s, s->kset[l]);
    k2[l] = dot(kk[l], kk[l]);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    V3 r = s->x[i];
    double qi = s->q[i];
    V3 f{0, 0, 0};
    double vir[6] = {0, 0, 0, 0, 0, 0};
    for (int l = 0; l < P; ++l) {
      V3 k = kk[l];
      double ph = dot(k, r);
      double fs = (-4 * kPi * qi / (s->volume * k2[l])) * (std::cos(ph) * rho[2 * l + 1] - std::sin(ph) * rho[2 * l]);
      f = f + fs * k;
      vir[0] += (r.x * k.x + r.x * k.x) * fs;
      vir[1] += (r.y * k.y + r.y * k.y) * fs;
      vir[2] += (r.z * k.z + r.z * k.z) * fs;
      vir[3] += (r.x * k.y + r.y * k.x) * fs;
      vir[4] += (r.x * k.z + r.z * k.x) * fs;
      vir[5] += (r.y * k.z + r.z * k.y) * fs;
    }
    f = f * S / (double)P;
    s->f_far[i] = s->u.qqr2e * f;
    for (int k = 0; k < 6; ++k) s->vir_far[(size_t)6 * i + k] = vir[k] * S / (double)P;
  }
}

// Ewald grid in the reference's k order, ExecPointLocator.h:126-145 / ExecutionMD.cpp:363-403
void ewald_grid(const orc_sys* s, const int kmax[3], std::vector<V3>& ks) {
  ks.clear();
  for (int i = -kmax[0]; i <= kmax[0]; ++i)
    for (int j = -kmax[1]; j <= kmax[1]; ++j)
      for (int k = -kmax[2]; k <= kmax[2]; ++k)
        if (i != 0 || j != 0 || k != 0) ks.push_back(kvec(s, V3{(double)i, (double)j, (double)k}));
}

void structure_factor(const orc_sys* s, const std::vector<V3>& ks, std::vector<double>& rho) {
  size_t K = ks.size();
  rho.assign(2 * K, 0.0);
#pragma omp parallel for schedule(dynamic, 4)
  for (long l = 0; l < (long)K; ++l) {
    KSum re, im;
    for (int i = 0; i < s->n; ++i) {
      double ph = dot(ks[l], s->x[i]);
      re.add(s->q[i] * std::cos(ph));
      im.add(s->q[i] * std::sin(ph));
    }
    rho[2 * l] = re.get();
    rho[2 * l + 1] = im.get();
  }
}

// sum over the grid of exp(-k^2/4a) |rho|^2 / k^2 (ComputeEwaldEnergy, ExecutionMD.cpp:405-443)
double ewald_energy_sum(const orc_sys* s, const std::vector<V3>& ks, const std::vector<double>& rho) {
  KSum e;
  for (size_t l = 0; l < ks.size(); ++l) {
    double k2 = dot(ks[l], ks[l]);
    e.add(std::exp(-k2 / (4 * s->p.alpha)) * (rho[2 * l] * rho[2 * l] + rho[2 * l + 1] * rho[2 * l + 1]) / k2);
  }
  return e.get();
}

// ComputeSelfEnergy total (before /N and qqr2e), ExecutionMD.cpp:544-558
double self_energy_raw(const orc_sys* s) {
  KSum q2;
  for (int i = 0; i < s->n; ++i) q2.add(s->q[i] * s->q[i]);
  return -std::sqrt(s->p.alpha / kPi) * q2.get();
}

// e_long = qqr2e * (2 pi / V) * grid / N + qqr2e * self / N (ExecutionMD.cpp:508-519)
double long_energy_avg(const orc_sys* s) {
  std::vector<V3> ks;
  ewald_grid(s, s->p.kmax, ks);
  std::vector<double> rho;
  structure_factor(s, ks, rho);
  double e = ewald_energy_sum(s, ks, rho);
  double avg = s->u.qqr2e * e * (2 * kPi / s->volume) / s->n;
  return avg + s->u.qqr2e * self_energy_raw(s) / s->n;
}

// ComputeNewFarElectrostaticsWorklet over the kmax grid, RunWorklet.cpp:2492-2540
void ewald_force(orc_sys* s, const int kmax[3], double* energy_sum) {
  std::vector<V3> ks;
  ewald_grid(s, kmax, ks);
  std::vector<double> rho;
  structure_factor(s, ks, rho);
  if (energy_sum) *energy_sum = ewald_energy_sum(s, ks, rho);
  int n = s->n;
  size_t K = ks.size();
  s->f_far.assign(n, V3{0, 0, 0});
  s->vir_far.assign((size_t)6 * n, 0.0);
  double alpha = s->p.alpha;
  std::vector<double> pref(K);
  for (size_t l = 0; l < K; ++l) {
    double k2 = dot(ks[l], ks[l]);
    pref[l] = std::exp(-k2 / (4 * alpha)) / (s->volume * k2);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    V3 r = s->x[i];
    double qi = s->q[i];
    V3 f{0, 0, 0};
    double vir[6] = {0, 0, 0, 0, 0, 0};
    for (size_t l = 0; l < K; ++l) {
      V3 k = ks[l];
      double ph = dot(k, r);
      double fs = -4 * kPi * qi * pref[l] * (std::cos(ph) * rho[2 * l + 1] - std::sin(ph) * rho[2 * l]);
      f = f + fs * k;
      vir[0] += (r.x * k.x + r.x * k.x) * fs;
      vir[1] += (r.y * k.y + r.y * k.y) * fs;
      vir[2] += (r.z * k.z + r.z * k.z) * fs;
      vir[3] += (r.x * k.y + r.y * k.x) * fs;
      vir[4] += (r.x * k.z + r.z * k.x) * fs;
      vir[5] += (r.y * k.z + r.z * k.y) * fs;
    }
    s->f_far[i] = s->u.qqr2e * f;
    for (int k = 0; k < 6; ++k) s->vir_far[(size_t)6 * i + k] = vir[k];
  }
}

// ComputeBondHarmonicWorklet + ReduceForceWorklet, MolecularWorklet.h:58-217
double bond_forces(orc_sys* s) {
  int n = s->n;
  s->f_bond.assign(n, V3{0, 0, 0});
  s->vir_bond.assign((size_t)6 * std::max(s->nb, 1), 0.0);
  KSum e;
  for (int b = 0; b < s->nb; ++b) {
    int i = s->bond_atoms[2 * b], j = s->bond_atoms[2 * b + 1];
    V3 r = min_image(s, s->x[i], s->x[j]);
    double d = std::sqrt(dot(r, r));
    double dr = d - s->bond_r0[b];
    double rk = s->bond_k[b] * dr;
    double fb = d > 0.01 ? -2.0 * rk / d : 0.0;
    V3 fi = r * fb;
    e.add(rk * dr);
    set_vir(&s->vir_bond[(size_t)6 * b], r, fi, 1.0);
    s->f_bond[i] = s->f_bond[i] + fi;
    s->f_bond[j] = s->f_bond[j] - fi;
  }
  return e.get();
}

// ComputeAngleHarmonicWorklet, MolecularWorklet.h:219-397 (SMALL = 0.001, :36)
double angle_forces(orc_sys* s) {
  int n = s->n;
  s->f_angle.assign(n, V3{0, 0, 0});
  s->vir_angle.assign((size_t)6 * std::max(s->na, 1), 0.0);
  KSum e;
  for (int a = 0; a < s->na; ++a) {
    int i = s->angle_atoms[3 * a], j = s->angle_atoms[3 * a + 1], k = s->angle_atoms[3 * a + 2];
    V3 rij = min_image(s, s->x[i], s->x[j]);
    V3 rkj = min_image(s, s->x[k], s->x[j]);
    double dij2 = dot(rij, rij), dij = std::sqrt(dij2);
    double dkj2 = dot(rkj, rkj), dkj = std::sqrt(dkj2);
    double c = dot(rij, rkj);
    c /= dij * dkj;
    if (c > 1.0) c = 1.0;
    if (c < -1.0) c = -1.0;
    double sn = std::sqrt(1.0 - c * c);
    if (sn < 0.001) sn = 0.001;
    sn = 1.0 / sn;
    double dth = std::acos(c) - (s->angle_t0[a] * kPi) / 180;
    double tk = s->angle_k[a] * dth;
    e.add(tk * dth);
    double aa = -2.0 * tk * sn;
    double a11 = aa * c / dij2, a12 = -aa / (dij * dkj), a22 = aa * c / dkj2;
    V3 fi = a11 * rij + a12 * rkj;
    V3 fk = a22 * rkj + a12 * rij;
    double* vv = &s->vir_angle[(size_t)6 * a];
    vv[0] = rij.x * fi.x + rkj.x * fk.x;
    vv[1] = rij.y * fi.y + rkj.y * fk.y;
    vv[2] = rij.z * fi.z + rkj.z * fk.z;
    vv[3] = rij.x * fi.y + rkj.x * fk.y;
    vv[4] = rij.x * fi.z + rkj.x * fk.z;
    vv[5] = rij.y * fi.z + rkj.y * fk.z;
    s->f_angle[i] = s->f_angle[i] + fi;
    s->f_angle[j] = s->f_angle[j] - (fi + fk);
    s->f_angle[k] = s->f_angle[k] + fk;
  }
  return e.get();
}

// ---------------------------------------------------------------- MD loop

// ExecutionNVT::ComputeTempe, ExecutionNVT.cpp:736-773
void compute_tempe(orc_sys* s) {
  KSum ke;
  for (int i = 0; i < s->n; ++i) ke.add(s->m[i] * dot(s->v[i], s->v[i]) * s->u.mvv2e);
  s->tempT_sum = ke.get();
  double n = s->n, kB = s->u.kB;
  if (s->p.inbuild)
    s->tempT = 0.5 * s->tempT_sum / (3 * n / 2.0);
  else if (s->p.fix_shake == RBMD_SHAKE_NULL)
    s->tempT = 0.5 * s->tempT_sum / ((3 * n - 3) * kB / 2.0);
  else if (s->p.fix_shake == RBMD_SHAKE_TRUE)
    s->tempT = 0.5 * s->tempT_sum / ((3 * n - n - 3) * kB / 2.0);
  else
    s->tempT = 0.5 * s->tempT_sum / ((3 * n) * kB / 2.0);
}

// ExecutionNVT::ComputeTempTarget, ExecutionNVT.cpp:1124-1147
void compute_temp_target(orc_sys* s) {
  if (s->t_stop == s->t_start) {
    s->t_target = s->t_start;
  } else {
    double delta = (double)s->step;
    if (delta != 0.0) delta = delta / (double)s->p.num_steps;
    s->t_target = s->t_start + delta * (s->t_stop - s->t_start);
  }
  s->ke_target = s->tdof * s->u.boltz * s->t_target;
}

// ExecutionNVT::NHCTempIntegrate, ExecutionNVT.cpp:1172-1256
void nhc_integrate(orc_sys* s) {
  double T = s->tempT;
  double boltz = s->u.boltz;
  double kecur = s->tdof * boltz * T;
  if (s->eta_mass_flag) {
    s->eta_mass[0] = s->tdof * boltz * s->t_target / (s->t_freq * s->t_freq);
    for (int c = 1; c < s->mtchain; ++c) s->eta_mass[c] = boltz * s->t_target / (s->t_freq * s->t_freq);
  }
  s->eta_dotdot[0] = s->eta_mass[0] > 0.0 ? (kecur - s->ke_target) / s->eta_mass[0] : 0.0;
  double ncfac = 1.0 / s->nc_tchain, expfac;
  for (int it = 0; it < s->nc_tchain; ++it) {
    for (int c = s->mtchain - 1; c > 0; --c) {
      expfac = std::exp(-ncfac * s->dt8 * s->eta_dot[c + 1]);
      s->eta_dot[c] *= expfac;
      s->eta_dot[c] += s->eta_dotdot[c] * ncfac * s->dt4;
      s->eta_dot[c] *= s->tdrag_factor;
      s->eta_dot[c] *= expfac;
    }
    expfac = std::exp(-ncfac * s->dt8 * s->eta_dot[1]);
    s->eta_dot[0] *= expfac;
    s->eta_dot[0] += s->eta_dotdot[0] * ncfac * s->dt4;
    s->eta_dot[0] *= s->tdrag_factor;
    s->eta_dot[0] *= expfac;
    s->factor_eta = std::exp(-ncfac * s->dthalf * s->eta_dot[0]);
    for (int i = 0; i < s->n; ++i) s->v[i] = s->factor_eta * s->v[i];
    s->tempT = T * s->factor_eta * s->factor_eta;
    kecur = s->tdof * boltz * s->tempT;
    s->eta_dotdot[0] = s->eta_mass[0] > 0.0 ? (kecur - s->ke_target) / s->eta_mass[0] : 0.0;
    for (int c = 0; c < s->mtchain; ++c) s->eta[c] += ncfac * s->dthalf * s->eta_dot[c];
    s->eta_dot[0] *= expfac;
    s->eta_dot[0] += s->eta_dotdot[0] * ncfac * s->dt4;
    s->eta_dot[0] *= expfac;
    for (int c = 1; c < s->mtchain; ++c) {
      expfac = std::exp(-ncfac * s->dt8 * s->eta_dot[c + 1]);
      s->eta_dot[c] *= expfac;
      s->eta_dotdot[c] =
          (s->eta_mass[c - 1] * s->eta_dot[c - 1] * s->eta_dot[c - 1] - boltz * s->t_target) / s->eta_mass[c];
      s->eta_dot[c] += s->eta_dotdot[c] * ncfac * s->dt4;
      s->eta_dot[c] *= expfac;
    }
  }
}

// ExecutionNVT::UpdateVelocity, ExecutionNVT.cpp:237-248 + RunWorklet.cpp:2357-2375
void update_velocity(orc_sys* s) {
  s->old_v = s->v;
  double h = 0.5 * s->p.dt * s->u.fmt2v;
  for (int i = 0; i < s->n; ++i) s->v[i] = s->v[i] + (0.5 * s->f[i] / s->m[i] * s->p.dt * s->u.fmt2v);
  (void)h;
}

// ExecutionNVT::UpdatePosition + ApplyPbcFlag, ExecutionNVT.cpp:250-268, RunWorklet.cpp:2602-2636
void update_position(orc_sys* s) {
  s->old_x = s->x;
  for (int i = 0; i < s->n; ++i) {
    s->x[i] = s->x[i] + s->v[i] * s->p.dt;
    for (int d = 0; d < 3; ++d) {
      double& c = at(s->x[i], d);
      if (c < s->lo[d]) {
        c += s->box[d];
        s->img[3 * i + d] -= 1;
      } else if (c >= s->hi[d]) {
        c -= s->box[d];
        s->img[3 * i + d] += 1;
      }
    }
  }
  build_locator(s);
}

// NewConstraintAWaterBondAngleWorklet (SHAKE), MolecularWorklet.h:2367-2645, on
// (old_x, old_v, F); results copied into x, v (ExecutionNVT.cpp:843-871).
void constraint_a(orc_sys* s) {
  double dt = s->p.dt, fmt2v = s->u.fmt2v;
  s->vir_shake.assign((size_t)6 * std::max(s->na, 1), 0.0);
  int maxit = 0, bad = 0;
  double range[3] = {s->hi[0] - s->lo[0], s->hi[1] - s->lo[1], s->hi[2] - s->lo[2]};
  for (int a = 0; a < s->na; ++a) {
    int i0 = s->angle_atoms[3 * a], i1 = s->angle_atoms[3 * a + 1], i2 = s->angle_atoms[3 * a + 2];
    V3 sp[3];
    int ids[3] = {i0, i1, i2};
    for (int t = 0; t < 3; ++t) {
      int id = ids[t];
      sp[t] = s->old_x[id] + s->old_v[id] * dt + 0.5 * dt * dt * s->f[id] / s->m[id] * fmt2v;
    }
    double bond1 = 1.0, bond2 = 1.0;
    double bond12 = std::sqrt(bond1 * bond1 + bond2 * bond2 - 2.0 * bond1 * bond2 * std::cos((109.4700 / 180.0) * kPi));
    V3 r01 = min_image(s, s->old_x[i1], s->old_x[i0]);
    V3 r12 = min_image(s, s->old_x[i2], s->old_x[i1]);
    V3 r20 = min_image(s, s->old_x[i0], s->old_x[i2]);
    V3 s10 = min_image(s, sp[0], sp[1]);
    V3 s21 = min_image(s, sp[1], sp[2]);
    V3 s02 = min_image(s, sp[2], sp[0]);
    double r01sq = dot(r01, r01), r02sq = dot(r20, r20), r12sq = dot(r12, r12);
    double s01sq = dot(s10, s10), s02sq = dot(s02, s02), s12sq = dot(s21, s21);
    double im0 = 1.0 / s->m[i0], im1 = 1.0 / s->m[i1], im2 = 1.0 / s->m[i2];
    double a11 = 2.0 * (im0 + im1) * dot(s10, r01);
    double a12 = -2.0 * im1 * dot(s10, r12);
    double a13 = -2.0 * im0 * dot(s10, r20);
    double a21 = -2.0 * im1 * dot(s21, r01);
    double a22 = 2.0 * (im1 + im2) * dot(s21, r12);
    double a23 = -2.0 * im2 * dot(s21, r20);
    double a31 = -2.0 * im0 * dot(s02, r01);
    double a32 = -2.0 * im2 * dot(s02, r12);
    double a33 = 2.0 * (im0 + im2) * dot(s02, r20);
    double det = a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a11 * a23 * a32 - a12 * a21 * a33 - a13 * a22 * a31;
    if (std::fabs(det) < 0.0001) bad = 1;
    double di = 1.0 / det;
    double b11 = di * (a22 * a33 - a23 * a32), b12 = -di * (a12 * a33 - a13 * a32), b13 = di * (a12 * a23 - a13 * a22);
    double b21 = -di * (a21 * a33 - a23 * a31), b22 = di * (a11 * a33 - a13 * a31), b23 = -di * (a11 * a23 - a13 * a21);
    double b31 = di * (a21 * a32 - a22 * a31), b32 = -di * (a11 * a32 - a12 * a31), b33 = di * (a11 * a22 - a12 * a21);
    double r0120 = dot(r01, r20), r0112 = dot(r01, r12), r2012 = dot(r20, r12);
    double q1_0101 = (im0 + im1) * (im0 + im1) * r01sq, q1_1212 = im1 * im1 * r12sq, q1_2020 = im0 * im0 * r02sq;
    double q1_0120 = -2.0 * (im0 + im1) * im0 * r0120, q1_0112 = -2.0 * (im0 + im1) * im1 * r0112;
    double q1_2012 = 2.0 * im0 * im1 * r2012;
    double q2_0101 = im1 * im1 * r01sq, q2_1212 = (im1 + im2) * (im1 + im2) * r12sq, q2_2020 = im2 * im2 * r02sq;
    double q2_0120 = 2.0 * im1 * im2 * r0120, q2_0112 = -2.0 * (im1 + im2) * im1 * r0112;
    double q2_2012 = -2.0 * (im1 + im2) * im2 * r2012;
    double q3_0101 = im0 * im0 * r01sq, q3_1212 = im2 * im2 * r12sq, q3_2020 = (im0 + im2) * (im0 + im2) * r02sq;
    double q3_0120 = -2.0 * (im0 + im2) * im0 * r0120, q3_0112 = 2.0 * im0 * im2 * r0112;
    double q3_2012 = -2.0 * (im0 + im2) * im2 * r2012;
    double tol = s->p.shake_tol > 0 ? s->p.shake_tol : 0.00001;  // MolecularWorklet.h:2507-2508
    int max_iter = 5000;
    double l01 = 0, l20 = 0, l12 = 0;
    int niter = 0, done = 0, ovf = 0;
    while (!done && niter < max_iter) {
      double quad1 = q1_0101 * l01 * l01 + q1_2020 * l20 * l20 + q1_1212 * l12 * l12 + q1_0120 * l01 * l20 +
                     q1_0112 * l01 * l12 + q1_2012 * l20 * l12;
      double quad2 = q2_0101 * l01 * l01 + q2_2020 * l20 * l20 + q2_1212 * l12 * l12 + q2_0120 * l01 * l20 +
                     q2_0112 * l01 * l12 + q2_2012 * l20 * l12;
      double quad3 = q3_0101 * l01 * l01 + q3_2020 * l20 * l20 + q3_1212 * l12 * l12 + q3_0120 * l01 * l20 +
                     q3_0112 * l01 * l12 + q3_2012 * l20 * l12;
      double b1 = bond1 * bond1 - s01sq - quad1;
      double b2 = bond2 * bond2 - s12sq - quad2;
      double b3 = bond12 * bond12 - s02sq - quad3;
      double n01 = b11 * b1 + b12 * b2 + b13 * b3;
      double n12 = b21 * b1 + b22 * b2 + b23 * b3;
      double n20 = b31 * b1 + b32 * b2 + b33 * b3;
      done = 1;
      if (std::fabs(n01 - l01) > tol) done = 0;
      if (std::fabs(n20 - l20) > tol) done = 0;
      if (std::fabs(n12 - l12) > tol) done = 0;
      l01 = n01;
      l20 = n20;
      l12 = n12;
      if (std::isnan(l01) || std::isnan(l20) || std::isnan(l12) || std::fabs(l01) > 1e20 || std::fabs(l20) > 1e20 ||
          std::fabs(l12) > 1e20) {
        done = 1;
        ovf = 1;
      }
      niter++;
    }
    if (!done || ovf) bad = 1;
    maxit = std::max(maxit, niter);
    V3 pc0 = l01 * r01 * im0 - l20 * r20 * im0;
    V3 pc1 = l12 * r12 * im1 - l01 * r01 * im1;
    V3 pc2 = l20 * r20 * im2 - l12 * r12 * im2;
    V3 pcs[3] = {pc0, pc1, pc2};
    for (int t = 0; t < 3; ++t) {
      int id = ids[t];
      V3 sv = s->old_v[id] + 0.5 * dt * s->f[id] / s->m[id] * fmt2v;
      s->old_v[id] = sv + pcs[t] / dt;
      s->old_x[id] = sp[t] + pcs[t];
    }
    for (int t = 0; t < 3; ++t) {
      int id = ids[t];
      for (int d = 0; d < 3; ++d) {
        double& c = at(s->old_x[id], d);
        if (c < s->lo[d]) {
          c += range[d];
          s->img[3 * id + d] -= 1;
        }
        if (c > s->hi[d]) {
          c -= range[d];
          s->img[3 * id + d] += 1;
        }
      }
    }
    double dtfsq = 0.5 * dt * dt * fmt2v;
    l01 /= dtfsq;
    l20 /= dtfsq;
    l12 /= dtfsq;
    double* vv = &s->vir_shake[(size_t)6 * a];
    vv[0] = l01 * r01.x * r01.x + l20 * r20.x * r20.x + l12 * r12.x * r12.x;
    vv[1] = l01 * r01.y * r01.y + l20 * r20.y * r20.y + l12 * r12.y * r12.y;
    vv[2] = l01 * r01.z * r01.z + l20 * r20.z * r20.z + l12 * r12.z * r12.z;
    vv[3] = l01 * r01.x * r01.y + l20 * r20.x * r20.y + l12 * r12.x * r12.y;
    vv[4] = l01 * r01.x * r01.z + l20 * r20.x * r20.z + l12 * r12.x * r12.z;
    vv[5] = l01 * r01.y * r01.z + l20 * r20.y * r20.z + l12 * r12.y * r12.z;
  }
  s->x = s->old_x;
  s->v = s->old_v;
  s->shake_max_iter = maxit;
  if (bad) s->status_flags |= 2;
}

// NewConstraintBWaterBondAngleWorklet (RATTLE), MolecularWorklet.h:2647-2832,
// on (x, old_v, F); result copied into v (ExecutionNVT.cpp:873-897).
void constraint_b(orc_sys* s) {
  double dt = s->p.dt, fmt2v = s->u.fmt2v;
  for (int a = 0; a < s->na; ++a) {
    int i0 = s->angle_atoms[3 * a], i1 = s->angle_atoms[3 * a + 1], i2 = s->angle_atoms[3 * a + 2];
    int ids[3] = {i0, i1, i2};
    V3 sv[3];
    for (int t = 0; t < 3; ++t) sv[t] = s->old_v[ids[t]] + 0.5 * dt * s->f[ids[t]] / s->m[ids[t]] * fmt2v;
    V3 r01 = min_image(s, s->x[i1], s->x[i0]);
    V3 r12 = min_image(s, s->x[i2], s->x[i1]);
    V3 r20 = min_image(s, s->x[i0], s->x[i2]);
    V3 sv10 = sv[0] - sv[1], sv21 = sv[1] - sv[2], sv02 = sv[2] - sv[0];
    double im0 = 1.0 / s->m[i0], im1 = 1.0 / s->m[i1], im2 = 1.0 / s->m[i2];
    double A[3][3];
    A[0][0] = (im1 + im0) * dot(r01, r01);
    A[0][1] = -im1 * dot(r01, r12);
    A[0][2] = (-im0) * dot(r01, r20);
    A[1][0] = A[0][1];
    A[1][1] = (im1 + im2) * dot(r12, r12);
    A[1][2] = -(im2)*dot(r20, r12);
    A[2][0] = A[0][2];
    A[2][1] = A[1][2];
    A[2][2] = (im0 + im2) * dot(r20, r20);
    double c[3] = {-dot(sv10, r01), -dot(sv21, r12), -dot(sv02, r20)};
    double det = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] -
                 A[0][0] * A[1][2] * A[2][1] - A[0][1] * A[1][0] * A[2][2] - A[0][2] * A[1][1] * A[2][0];
    if (std::fabs(det) < 0.0001) s->status_flags |= 2;
    double di = 1.0 / det;
    double B[3][3];
    B[0][0] = di * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
    B[0][1] = -di * (A[0][1] * A[2][2] - A[0][2] * A[2][1]);
    B[0][2] = di * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
    B[1][0] = -di * (A[1][0] * A[2][2] - A[1][2] * A[2][0]);
    B[1][1] = di * (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
    B[1][2] = -di * (A[0][0] * A[1][2] - A[0][2] * A[1][0]);
    B[2][0] = di * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    B[2][1] = -di * (A[0][0] * A[2][1] - A[0][1] * A[2][0]);
    B[2][2] = di * (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
    double l[3];
    for (int r = 0; r < 3; ++r) {
      l[r] = 0;
      for (int k = 0; k < 3; ++k) l[r] += B[r][k] * c[k];
    }
    V3 vc0 = l[0] * r01 * im0 - l[2] * r20 * im0;
    V3 vc1 = l[1] * r12 * im1 - l[0] * r01 * im1;
    V3 vc2 = l[2] * r20 * im2 - l[1] * r12 * im2;
    s->old_v[i0] = sv[0] + vc0;
    s->old_v[i1] = sv[1] + vc1;
    s->old_v[i2] = sv[2] + vc2;
  }
  s->v = s->old_v;
}

// ExecutionNVT::ComputeAllForce, ExecutionNVT.cpp:195-235
void compute_force(orc_sys* s) {
  int n = s->n;
  s->f.assign(n, V3{0, 0, 0});
  int ff = s->p.force_field;
  // near force (ComputeVerletlistNearForce / ComputeRBLNearForce)
  if (s->cap == 0) s->cap = default_cap(s);
  if (build_neighbors(s, s->cap) != RBMD_OK) s->status_flags |= 4;
  if (s->p.neighbor_type == RBMD_NEIGHBOR_RBL) {
    force_rbl(s);
    rbl_energy(s);  // ComputeRBLNearForce ends with ComputeLJCoulEnergy (ExecutionMD.cpp:670)
  } else {
    force_verlet(s);
  }
  double elj = 0, ecoul = 0;
  {
    KSum a, b;
    for (int i = 0; i < n; ++i) {
      a.add(s->e_lj[i]);
      b.add(s->e_coul[i]);
    }
    elj = a.get() / n;
    ecoul = b.get() / n;
  }
  s->e_lj_avg = elj;
  s->e_coul_avg = ecoul;
  if (ff == RBMD_FF_LJ_CUT) {
    s->f = s->f_near;
    s->f_far.assign(n, V3{0, 0, 0});
    s->vir_far.assign((size_t)6 * n, 0.0);
    return;
  }
  // far force (EleNewForce, ExecutionNVT.cpp:704-720)
  if (s->p.coulomb_type == RBMD_COULOMB_RBE) {
    std::vector<double> rho;
    rbe_rho(s, rho);
    rbe_force(s, rho);
    if (s->p.energy_every > 0 && s->step % s->p.energy_every == 0) {
      s->e_long_avg = long_energy_avg(s);
      s->energy_step = s->step;
    }
  } else if (s->p.coulomb_type == RBMD_COULOMB_EWALD) {
    double esum = 0;
    ewald_force(s, s->p.kmax, &esum);
    s->e_long_avg = s->u.qqr2e * esum * (2 * kPi / s->volume) / n + s->u.qqr2e * self_energy_raw(s) / n;
    s->energy_step = s->step;
  } else {
    s->f_far.assign(n, V3{0, 0, 0});
    s->vir_far.assign((size_t)6 * n, 0.0);
  }
  for (int i = 0; i < n; ++i) s->f[i] = s->f_far[i] + s->f_near[i];
  if (ff == RBMD_FF_CVFF) {
    if (s->p.neighbor_type == RBMD_NEIGHBOR_RBL && s->p.special_bonds) {
      force_special_coul(s);
      for (int i = 0; i < n; ++i) s->f[i] = s->f[i] + s->f_spec[i];
    }
    double eb = bond_forces(s);
    for (int i = 0; i < n; ++i) s->f[i] = s->f[i] + s->f_bond[i];
    double ea = angle_forces(s);
    for (int i = 0; i < n; ++i) s->f[i] = s->f[i] + s->f_angle[i];
    s->e_bond_avg = s->nb > 0 ? eb / s->nb : 0;
    s->e_angle_avg = s->na > 0 ? ea / s->na : 0;
  }
}

// ExecutionMD::ComputeVirial, ExecutionMD.cpp:1263-1366
void compute_virial(orc_sys* s) {
  double w[6] = {0, 0, 0, 0, 0, 0};
  auto addsum = [&](const std::vector<double>& a, size_t cnt, double scale) {
    for (int k = 0; k < 6; ++k) {
      KSum t;
      for (size_t i = 0; i < cnt; ++i) t.add(a[6 * i + k]);
      w[k] += t.get() * scale;
    }
  };
  int ff = s->p.force_field;
  addsum(s->vir_near, s->n, 1.0);
  if (ff != RBMD_FF_LJ_CUT) addsum(s->vir_far, s->n, s->u.qqr2e);
  if (ff == RBMD_FF_CVFF) {
    if (s->p.neighbor_type == RBMD_NEIGHBOR_RBL && !s->vir_spec.empty()) addsum(s->vir_spec, s->n, 1.0);
    if (s->nb > 0) addsum(s->vir_bond, s->nb, 1.0);
    if (s->na > 0) addsum(s->vir_angle, s->na, 1.0);
    if (s->p.fix_shake == RBMD_SHAKE_TRUE && !s->vir_shake.empty() && s->na > 0) addsum(s->vir_shake, s->na, 1.0);
  }
  for (int k = 0; k < 6; ++k) s->virial[k] = w[k];
}

// ExecutionNVT::Compute_Pressure_Scalar, ExecutionNVT.cpp:1258-1276
void compute_pressure(orc_sys* s) {
  double vol = (s->hi[0] - s->lo[0]) * (s->hi[1] - s->lo[1]) * (s->hi[2] - s->lo[2]);
  compute_virial(s);
  s->pressure = (s->tdof * s->u.boltz * s->tempT + s->virial[0] + s->virial[1] + s->virial[2]) / 3.0 * (1.0 / vol) *
                s->u.nktv2p;
}

void draw_kset(orc_sys* s, uint64_t stream) {
  if (s->p.coulomb_type != RBMD_COULOMB_RBE) return;
  std::vector<double> m((size_t)3 * s->p.rbe_p);
  sample_kset(s->p.alpha, s->box, s->p.rbe_p, s->p.seed, stream, m.data());
  s->P = s->p.rbe_p;
  s->kset.resize(s->P);
  for (int l = 0; l < s->P; ++l) s->kset[l] = V3{m[3 * l], m[3 * l + 1], m[3 * l + 2]};
}

}  // namespace

// ------------------------------------------------------------------ C ABI

extern "C" {

int orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
  return n;
#else
  (void)n;
  return 1;
#endif
}

int orc_static_table_gnear(int32_t n, const double* dis, double* g) {
  if (n < 0 || (n > 0 && (!dis || !g))) return RBMD_ERR_ARG;
  for (int i = 0; i < n; ++i) g[i] = gnear_table(dis[i]);
  return RBMD_OK;
}

int orc_create(const rbmd_sys_desc* d, const rbmd_params* p, orc_sys** out) {
  if (!d || !p || !out || d->n_atoms <= 0 || d->n_types <= 0) return RBMD_ERR_ARG;
  // LJ/CUT over random-batch lists is the reference's separate LJForceRBL path
  // (ExecutionMD::ComputeRBLLJForce, ExecutionMD.cpp:752-830; ComputeLJForceRBLWorklet,
  // RunWorklet.cpp:1817-1910), not built here
  if (p->neighbor_type == RBMD_NEIGHBOR_RBL && p->force_field == RBMD_FF_LJ_CUT) return RBMD_ERR_UNSUPPORTED;
  orc_sys* s = new orc_sys();
  s->p = *p;
  s->u = units_of(p->unit);
  s->n = d->n_atoms;
  s->ntypes = d->n_types;
  for (int k = 0; k < 3; ++k) {
    s->lo[k] = d->lo[k];
    s->hi[k] = d->hi[k];
    s->box[k] = d->hi[k] - d->lo[k];
  }
  s->volume = s->box[0] * s->box[1] * s->box[2];
  int n = s->n;
  s->type.assign(d->type, d->type + n);
  s->q.assign(d->charge, d->charge + n);
  s->m.assign(d->mass, d->mass + n);
  if (d->mol_id)
    s->mol.assign(d->mol_id, d->mol_id + n);
  else {
    s->mol.resize(n);
    for (int i = 0; i < n; ++i) s->mol[i] = i;
  }
  s->eps.assign(d->eps, d->eps + s->ntypes);
  s->sig.assign(d->sigma, d->sigma + s->ntypes);
  if (d->special_offsets && d->special_ids && d->special_weights) {
    s->has_special = true;
    s->sp_off.assign(d->special_offsets, d->special_offsets + n + 1);
    int ns = s->sp_off[n];
    s->sp_ids.assign(d->special_ids, d->special_ids + ns);
    s->sp_w.assign(d->special_weights, d->special_weights + ns);
  }
  s->nb = d->n_bonds;
  if (s->nb > 0) {
    s->bond_atoms.assign(d->bond_atoms, d->bond_atoms + 2 * s->nb);
    s->bond_k.assign(d->bond_k, d->bond_k + s->nb);
    s->bond_r0.assign(d->bond_r0, d->bond_r0 + s->nb);
  }
  s->na = d->n_angles;
  if (s->na > 0) {
    s->angle_atoms.assign(d->angle_atoms, d->angle_atoms + 3 * s->na);
    s->angle_k.assign(d->angle_k, d->angle_k + s->na);
    s->angle_t0.assign(d->angle_theta0, d->angle_theta0 + s->na);
  }
  // molecule groups (SetMolecularGroup, ModelFileInitCondition.cpp:698-736)
  if (d->mol_id) {
    int nm = 0;
    for (int i = 0; i < n; ++i) nm = std::max(nm, s->mol[i] + 1);
    std::vector<std::vector<int>> members(nm);
    for (int i = 0; i < n; ++i) members[s->mol[i]].push_back(i);
    s->grp_off.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) s->grp_off[i + 1] = s->grp_off[i] + (int)members[s->mol[i]].size();
    s->grp_ids.resize(s->grp_off[n]);
    for (int i = 0; i < n; ++i) std::copy(members[s->mol[i]].begin(), members[s->mol[i]].end(), s->grp_ids.begin() + s->grp_off[i]);
  }
  s->x.assign(n, V3{0, 0, 0});
  s->v.assign(n, V3{0, 0, 0});
  s->f.assign(n, V3{0, 0, 0});
  s->img.assign((size_t)3 * n, 0);
  geometry(s);
  *out = s;
  return RBMD_OK;
}

void orc_destroy(orc_sys* s) { delete s; }

int orc_set_state(orc_sys* s, const double* x, const double* v, const int32_t* image) {
  for (int i = 0; i < s->n; ++i) {
    if (x) s->x[i] = V3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    if (v) s->v[i] = V3{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
  }
  if (image) std::copy(image, image + 3 * s->n, s->img.begin());
  return RBMD_OK;
}

int orc_get_state(orc_sys* s, double* x, double* v, double* f, int32_t* image) {
  for (int i = 0; i < s->n; ++i)
    for (int d = 0; d < 3; ++d) {
      if (x) x[3 * i + d] = at(s->x[i], d);
      if (v) v[3 * i + d] = at(s->v[i], d);
      if (f) f[3 * i + d] = at(s->f[i], d);
    }
  if (image) std::copy(s->img.begin(), s->img.end(), image);
  return RBMD_OK;
}

void orc_locator_geometry(const orc_sys* s, int32_t dims[3], int32_t* nc, double dxdydz[3]) {
  for (int d = 0; d < 3; ++d) {
    if (dims) dims[d] = s->dims[d];
    if (dxdydz) dxdydz[d] = s->dxdydz[d];
  }
  if (nc) *nc = s->nc;
}

int orc_locator_build(orc_sys* s) {
  build_locator(s);
  return RBMD_OK;
}

int orc_locator_export(const orc_sys* s, int32_t* cell_of_atom, int32_t* point_ids, int32_t* lower, int32_t* upper) {
  if (s->cell_of.empty()) return RBMD_ERR_STATE;
  if (cell_of_atom) std::copy(s->cell_of.begin(), s->cell_of.end(), cell_of_atom);
  if (point_ids) std::copy(s->point_ids.begin(), s->point_ids.end(), point_ids);
  if (lower) std::copy(s->lower.begin(), s->lower.end(), lower);
  if (upper) std::copy(s->upper.begin(), s->upper.end(), upper);
  return RBMD_OK;
}

int orc_neighbors_build(orc_sys* s, int32_t cap) {
  if (s->cell_of.empty()) build_locator(s);
  if (cap <= 0) cap = default_cap(s);
  return build_neighbors(s, cap);
}

int orc_neighbors_capacity(const orc_sys* s) { return s->cap; }

int orc_neighbors_export(const orc_sys* s, int32_t cap, int32_t* ids, double* offsets, int32_t* num) {
  if (cap != s->cap) return RBMD_ERR_ARG;
  size_t tot = (size_t)s->n * cap;
  if (ids) std::copy(s->nb_ids.begin(), s->nb_ids.end(), ids);
  if (offsets)
    for (size_t k = 0; k < tot; ++k) {
      offsets[3 * k] = s->nb_off[k].x;
      offsets[3 * k + 1] = s->nb_off[k].y;
      offsets[3 * k + 2] = s->nb_off[k].z;
    }
  if (num) std::copy(s->nb_num.begin(), s->nb_num.end(), num);
  return RBMD_OK;
}

int orc_bruteforce_counts(const orc_sys* s, double rc, int32_t* num) {
  // every periodic image within one box length in each direction
  double rc2 = rc * rc;
  for (int i = 0; i < s->n; ++i) {
    int c = 0;
    for (int j = 0; j < s->n; ++j)
      for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
          for (int e = -1; e <= 1; ++e) {
            V3 r = s->x[j] + V3{a * s->box[0], b * s->box[1], e * s->box[2]} - s->x[i];
            double d2 = dot(r, r);
            if (rc2 - d2 > 0.0001 && d2 > 0.0001) c++;
          }
    num[i] = c;
  }
  return RBMD_OK;
}

int orc_force_real(orc_sys* s, double* f, double* vir, double* elj, double* ecoul) {
  if (s->nb_num.empty()) return RBMD_ERR_STATE;
  if (s->p.neighbor_type == RBMD_NEIGHBOR_RBL) {
    force_rbl(s);
    rbl_energy(s);
  } else {
    force_verlet(s);
  }
  for (int i = 0; i < s->n; ++i) {
    if (f) {
      f[3 * i] = s->f_near[i].x;
      f[3 * i + 1] = s->f_near[i].y;
      f[3 * i + 2] = s->f_near[i].z;
    }
    if (vir)
      for (int k = 0; k < 6; ++k) vir[6 * i + k] = s->vir_near[(size_t)6 * i + k];
    if (elj) elj[i] = s->e_lj[i];
    if (ecoul) ecoul[i] = s->e_coul[i];
  }
  return RBMD_OK;
}

int orc_rbe_set_kvectors(orc_sys* s, int32_t p, const double* m) {
  if (p <= 0 || !m) return RBMD_ERR_ARG;
  s->P = p;
  s->kset.resize(p);
  for (int l = 0; l < p; ++l) s->kset[l] = V3{m[3 * l], m[3 * l + 1], m[3 * l + 2]};
  return RBMD_OK;
}

int orc_rbe_sample(double alpha, const double box[3], int32_t p, uint64_t seed, uint64_t stream, double* m) {
  if (p <= 0 || !m || alpha <= 0) return RBMD_ERR_ARG;
  sample_kset(alpha, box, p, seed, stream, m);
  return RBMD_OK;
}

double orc_rbe_S(double alpha, const double box[3]) { return compute_S(alpha, box); }

double orc_rbe_S_direct(double alpha, const double box[3], int32_t mmax) {
  KSum sum;
  for (int i = -mmax; i <= mmax; ++i)
    for (int j = -mmax; j <= mmax; ++j)
      for (int k = -mmax; k <= mmax; ++k) {
        if (!i && !j && !k) continue;
        double kx = 2 * kPi * i / box[0], ky = 2 * kPi * j / box[1], kz = 2 * kPi * k / box[2];
        sum.add(std::exp(-(kx * kx + ky * ky + kz * kz) / (4 * alpha)));
      }
  return sum.get();
}

int orc_rbe_structure_factor(orc_sys* s, double* rho) {
  if (s->P <= 0) return RBMD_ERR_STATE;
  std::vector<double> r;
  rbe_rho(s, r);
  std::copy(r.begin(), r.end(), rho);
  return RBMD_OK;
}

int orc_rbe_force(orc_sys* s, double* f, double* vir) {
  if (s->P <= 0) return RBMD_ERR_STATE;
  std::vector<double> r;
  rbe_rho(s, r);
  rbe_force(s, r);
  for (int i = 0; i < s->n; ++i) {
    if (f) {
      f[3 * i] = s->f_far[i].x;
      f[3 * i + 1] = s->f_far[i].y;
      f[3 * i + 2] = s->f_far[i].z;
    }
    if (vir)
      for (int k = 0; k < 6; ++k) vir[6 * i + k] = s->vir_far[(size_t)6 * i + k];
  }
  return RBMD_OK;
}

int orc_ewald_force(orc_sys* s, const int32_t kmax[3], double* f, double* vir, double* energy_total) {
  double esum = 0;
  ewald_force(s, kmax, &esum);
  if (energy_total) *energy_total = s->u.qqr2e * esum * (2 * kPi / s->volume);
  for (int i = 0; i < s->n; ++i) {
    if (f) {
      f[3 * i] = s->f_far[i].x;
      f[3 * i + 1] = s->f_far[i].y;
      f[3 * i + 2] = s->f_far[i].z;
    }
    if (vir)
      for (int k = 0; k < 6; ++k) vir[6 * i + k] = s->vir_far[(size_t)6 * i + k];
  }
  return RBMD_OK;
}

double orc_self_energy_total(const orc_sys* s) { return s->u.qqr2e * self_energy_raw(s); }

int orc_force_bonded(orc_sys* s, double* f, double* ebond_sum, double* eangle_sum) {
  double eb = bond_forces(s);
  double ea = angle_forces(s);
  // CVFF + RBL + special bonds: the intramolecular Coulomb correction too, as
  // ComputeAllForce adds it beside the bonded terms (ExecutionNVT.cpp:212-214)
  const bool spec = s->p.force_field == RBMD_FF_CVFF && s->p.neighbor_type == RBMD_NEIGHBOR_RBL && s->p.special_bonds;
  if (spec) force_special_coul(s);
  for (int i = 0; i < s->n; ++i) {
    V3 t = s->f_bond[i] + s->f_angle[i];
    if (spec) t = s->f_spec[i] + t;
    if (f) {
      f[3 * i] = t.x;
      f[3 * i + 1] = t.y;
      f[3 * i + 2] = t.z;
    }
  }
  if (ebond_sum) *ebond_sum = eb;
  if (eangle_sum) *eangle_sum = ea;
  return RBMD_OK;
}

// ExecutionNVT::Init + InitialCondition + SetUp, ExecutionNVT.cpp:60-74, 139-187, 1106-1122
int orc_md_init(orc_sys* s) {
  s->step = 0;
  s->status_flags = 0;
  build_locator(s);
  draw_kset(s, 0);
  s->tdof = 3 * s->n - 3;
  if (s->p.fix_shake == RBMD_SHAKE_TRUE) s->tdof -= s->n;
  s->t_start = s->p.temperature[0];
  s->t_stop = s->p.temperature[1];
  s->t_period = s->p.temperature[2];
  s->t_freq = 1 / s->t_period;
  s->mtchain = 3;
  s->nc_tchain = 1;
  s->eta_mass_flag = 1;
  double dt = s->p.dt;
  s->dthalf = 0.5 * dt;
  s->dt4 = 0.25 * dt;
  s->dt8 = 0.125 * dt;
  s->drag = 0.0;
  s->tdrag_factor = 1.0 - (dt * s->t_freq * s->drag / s->nc_tchain);
  for (int c = 0; c < 3; ++c) s->eta[c] = s->eta_dot[c] = s->eta_dotdot[c] = 0.0;
  s->eta_dot[3] = 0;
  compute_tempe(s);
  compute_temp_target(s);
  s->eta_mass[0] = s->tdof * s->u.boltz * s->t_target / (s->t_freq * s->t_freq);
  for (int c = 1; c < s->mtchain; ++c) s->eta_mass[c] = s->u.boltz * s->t_target / (s->t_freq * s->t_freq);
  for (int c = 1; c < s->mtchain; ++c)
    s->eta_dotdot[c] = (s->eta_mass[c - 1] * s->eta_dot[c - 1] * s->eta_dot[c - 1] - s->u.boltz * s->t_target) / s->eta_mass[c];
  s->cap = 0;
  compute_force(s);
  compute_pressure(s);
  s->initialized = true;
  return s->status_flags ? RBMD_ERR_CAPACITY : RBMD_OK;
}

// Executioner::TakeStep -> ExecutionMD::Execute -> ExecutionNVT::PreSolve + Solve,
// Executioner.cpp:112-117, ExecutionNVT.cpp:76-135.  PreSolve's locator rebuild is
// not repeated: nothing reads it before UpdatePosition rebuilds it.
int orc_md_step(orc_sys* s, int64_t nsteps) {
  if (!s->initialized) return RBMD_ERR_STATE;
  bool shake = s->p.fix_shake == RBMD_SHAKE_TRUE && s->na > 0;
  for (int64_t it = 0; it < nsteps; ++it) {
    s->step++;
    draw_kset(s, (uint64_t)s->step);
    int tc = s->p.temp_ctrl;
    if (tc == RBMD_TEMP_NOSE_HOOVER || tc == RBMD_TEMP_NONE) {
      if (tc == RBMD_TEMP_NOSE_HOOVER) {
        compute_temp_target(s);
        nhc_integrate(s);
      }
      update_velocity(s);
      update_position(s);
      if (shake) constraint_a(s);
      compute_force(s);
      update_velocity(s);
      compute_tempe(s);
      compute_pressure(s);
      if (tc == RBMD_TEMP_NOSE_HOOVER) nhc_integrate(s);
      if (shake) constraint_b(s);
    } else {
      update_velocity(s);
      update_position(s);
      if (shake) constraint_a(s);
      compute_force(s);
      update_velocity(s);
      if (shake) constraint_b(s);
      compute_tempe(s);
      // UpdateVelocityByTempConType, ExecutionNVT.cpp:270-291
      double coeff;
      if (tc == RBMD_TEMP_RESCALE)
        coeff = std::sqrt(s->t_start / s->tempT);
      else
        coeff = std::sqrt(1.0 + (s->p.dt / s->t_period) * (s->t_start / s->tempT - 1.0));
      for (int i = 0; i < s->n; ++i) s->v[i] = coeff * s->v[i];
      compute_pressure(s);
    }
    // a failed constraint or a non-finite position ends the run
    for (int i = 0; i < s->n && !(s->status_flags & 1); ++i)
      if (!std::isfinite(s->x[i].x) || !std::isfinite(s->x[i].y) || !std::isfinite(s->x[i].z)) s->status_flags |= 1;
    if (s->status_flags & 3) break;
  }
  if (s->status_flags & 1) return RBMD_ERR_NONFINITE;
  return (s->status_flags & 2) ? RBMD_ERR_SHAKE : ((s->status_flags & 4) ? RBMD_ERR_CAPACITY : RBMD_OK);
}

int orc_md_thermo(const orc_sys* s, rbmd_thermo* o) {
  std::memset(o, 0, sizeof(*o));
  o->step = s->step;
  o->temperature = s->tempT;
  o->pressure = s->pressure;
  o->ke_sum = s->tempT_sum;
  o->e_lj = s->e_lj_avg;
  o->e_coul = s->e_coul_avg;
  o->e_long = s->e_long_avg;
  o->e_bond = s->e_bond_avg;
  o->e_angle = s->e_angle_avg;
  for (int k = 0; k < 6; ++k) o->virial[k] = s->virial[k];
  for (int c = 0; c < 3; ++c) {
    o->nhc_eta[c] = s->eta[c];
    o->nhc_eta_dot[c] = s->eta_dot[c];
  }
  o->energy_step = s->energy_step;
  o->shake_max_iter = s->shake_max_iter;
  o->status_flags = s->status_flags;
  return RBMD_OK;
}

int orc_md_energy(orc_sys* s) {
  if (s->p.coulomb_type == RBMD_COULOMB_NONE) return RBMD_ERR_STATE;
  s->e_long_avg = long_energy_avg(s);
  s->energy_step = s->step;
  return RBMD_OK;
}

}  // extern "C"
