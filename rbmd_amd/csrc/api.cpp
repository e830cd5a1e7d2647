// api.cpp -- the C ABI of include/rbmd_mi355x.h: contexts, systems, device
// memory, the step schedule and per-kernel timing.  Host code only; kernels
// are reached through the launch_* functions of dev.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rbmd_mi355x.h"
#include "dev.hpp"
#include "sampler.hpp"

using namespace rbmd;

namespace {

thread_local std::string g_last_error;

rbmd_status fail(rbmd_status code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) return fail(RBMD_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

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

struct ProfRec {
  std::string name;
  int64_t launches = 0;
  double ms = 0;
};

}  // namespace

struct rbmd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream of a step: the k-space and bonded forces run on it beside
  // the list build and pair forces (both default priority)
  hipStream_t side = nullptr;
  bool prof = false;
  std::vector<ProfRec> recs;
  std::map<std::string, int> index;
  struct Pending {
    int rec;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;

  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
  void flush() {
    if (pending.empty()) return;
    (void)hipStreamSynchronize(stream);
    for (auto& p : pending) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, p.a, p.b);
      recs[p.rec].ms += ms;
      recs[p.rec].launches++;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  template <class F>
  void run(const char* name, F&& f) {
    if (!prof) {
      f();
      return;
    }
    auto it = index.find(name);
    int r;
    if (it == index.end()) {
      r = (int)recs.size();
      index[name] = r;
      recs.push_back(ProfRec{name, 0, 0});
    } else {
      r = it->second;
    }
    hipEvent_t a = ev(), b = ev();
    (void)hipEventRecord(a, stream);
    f();
    (void)hipEventRecord(b, stream);
    pending.push_back({r, a, b});
    if (pending.size() > 4096) flush();
  }
};

// producers of thermo partial rows (row widths kRowPair, kRowK, kRowBond, kRowKe, kRowShake)
enum { SEG_PAIR, SEG_K, SEG_BOND, SEG_KE, SEG_SHAKE, NSEG };
constexpr int kSegWidth[NSEG] = {kRowPair, kRowK, kRowBond, kRowKe, kRowShake};
constexpr int kSegSlot0[NSEG] = {S_ELJ, S_VIR_K, S_VIR_BOND, S_KE, S_VIR_SHAKE};
constexpr int kSegGroups[NSEG] = {64, 16, 16, 64, 16};  // 992 reduction threads in all

struct rbmd_sys {
  rbmd_ctx* ctx = nullptr;
  rbmd_params p;
  Units u;
  int n = 0, ntypes = 0, nb = 0, na = 0, nmol = 0;
  SysView v;
  IntView iv;
  KView kv{};
  std::vector<void*> allocs;
  int64_t bytes = 0;
  // partial-sum rows, one segment per producer; seg_used = rows its last
  // launch wrote (the reduction reads exactly those)
  double* rows = nullptr;
  size_t seg_off[NSEG] = {0, 0, 0, 0, 0};
  int seg_used[NSEG] = {0, 0, 0, 0, 0};
  double* seg_rows(int g) { return rows + seg_off[g]; }
  size_t rows_len = 0;
  bool fused = false;  // VERLET real space through the fused cell-stencil kernel
  NhcState* nhc = nullptr;
  NhcParams np;
  int* shake_iter = nullptr;
  int* free_dev = nullptr;
  int nfree = 0;
  // k space
  KSampler sampler;
  int P = 0;
  double S = 0;
  double* kcur = nullptr;  // [3P] current k vectors
  double* kgraph = nullptr;  // [3P] k vectors of the running step inside a step graph
  // captured MD steps (rbmd_md_run_graph): [0] without, [1] with the thermo sums
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  bool graph_ok = false;           // cleared whenever captured pointers or sizes change
  bool rattle_pending = false;     // within an eager call: the last step's RATTLE runs with the next step's start
  double* kring = nullptr; // [2R][3P]
  double* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;  // side-stream branch of a step
  int ring_R = 256;
  int export_cap = 0;
  int64_t ring_batch = -1;  // batch of steps currently staged in half (batch % 2)
  int64_t ring_have[2] = {-1, -1};
  bool injected = false;
  std::vector<double> k_injected;
  double* rho_part = nullptr;
  KRec* krec = nullptr;  // [P] k records of the force pass
  double* rho = nullptr;
  int rho_part_cap = 0;  // doubles
  // cell geometry of the reference's per-step list (cells of width >= rc) and,
  // for a skin list, of the MD step's list (cells of width >= rc + skin)
  struct Geom {
    int dims[3];
    double dxdydz[3];
    int ncell, n_items, lds_bins;
    double list_cut2;
  };
  Geom g_ref{}, g_skin{};
  bool skin = false;        // MD steps keep a list with a skin (params.skin > 0)
  bool skin_active = false; // v currently holds the skin geometry
  bool skin_ready = false;  // the skin list's buffers and build request are in place for the next step
  // host mirrors
  bool locator_built = false, nbr_built = false, initialized = false;
  double e_long = 0;
  int64_t energy_step = -1;
  int64_t step = 0;
  std::vector<int> host_angles;

  template <class T>
  rbmd_status dalloc(T** ptr, size_t count) {
    size_t b = std::max<size_t>(count, 1) * sizeof(T);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, b);
    if (e != hipSuccess) return fail(RBMD_ERR_OOM, "hipMalloc(%zu bytes): %s", b, hipGetErrorString(e));
    allocs.push_back(q);
    bytes += (int64_t)b;
    *ptr = (T*)q;
    return RBMD_OK;
  }
  template <class T>
  rbmd_status upload(T** ptr, const T* host, size_t count) {
    rbmd_status rc = dalloc(ptr, count);
    if (rc) return rc;
    if (count && host) HIP_TRY(hipMemcpy(*ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
    return RBMD_OK;
  }
};

namespace {

#define RC_TRY(expr)             \
  do {                           \
    rbmd_status _r = (expr);     \
    if (_r != RBMD_OK) return _r; \
  } while (0)

rbmd_status check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(RBMD_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return RBMD_OK;
}

// ContPointLocator::SetRange/SetCutOff/SetRs (ContPointLocator.h:44-72): nc = ceil(rc/rs),
// dims = trunc(L / (rc/nc)), dxdydz = L / dims
void geometry(const rbmd_params& p, const double lo[3], const double hi[3], int dims[3], int* nc, double dx[3]) {
  double rc = p.cut_off;
  double rs = p.neighbor_type == RBMD_NEIGHBOR_RBL ? p.r_core : p.cut_off;
  *nc = (int)std::ceil(rc / rs);
  for (int d = 0; d < 3; ++d) {
    dims[d] = (int)((hi[d] - lo[d]) / (rc / *nc));
    if (dims[d] < 1) dims[d] = 1;
    dx[d] = (hi[d] - lo[d]) / dims[d];
  }
}

// RBL capacities, ExecutionMD.cpp:582-592
void rbl_capacity(const rbmd_params& p, int n, double volume, int* rs_num, int* pice_num) {
  double rho = p.rbl_rho > 0 ? p.rbl_rho : n / volume;
  double rc = p.cut_off, rs = p.r_core;
  double pif = (double)(float)kPi;
  double coeff_rcs = 1.0 + (0.05 / rho - 0.05);
  long id_coeff = std::lround(coeff_rcs);
  long rsn = (long)(id_coeff * rho * std::ceil(4.0 / 3.0 * pif * rs * rs * rs)) + 1;
  long rcn = (long)(id_coeff * rho * std::ceil(4.0 / 3.0 * pif * rc * rc * rc)) + 1;
  double rate = (double)p.neighbor_sample_num / (double)(rcn - rsn);
  *pice_num = (int)std::ceil(1.0 / rate);
  // core slot capacity: twice the reference's rs_num plus 32 (the reference
  // overwrites past rs_num silently, hazard Q7); pice_num keeps its formula
  *rs_num = (int)(2 * rsn + 32);
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Piecewise polynomial table of the Coulomb radial factors over [0, rc]:
// interval k = [k h, (k+1) h], local variable u = r/h - k - 1/2 in [-1/2, 1/2].
// Per interval the degree-`degree` (odd) Chebyshev interpolant at the
// Chebyshev nodes (coefficients by the discrete cosine sums, then converted to
// powers of u), all in long double.
// P(r) = erfc(beta r) + 2 beta/sqrt(pi) r exp(-beta^2 r^2),  Q(r) = erfc(beta r).
// Layout: [f][d / 2][k][d % 2], f = 0 P, 1 Q.  The FP64 path uses degree
// kCoulD; the mixed-precision path rounds a degree-kCoulD32 table to FP32.
std::vector<double> coul_table(double beta, double rc, int degree) {
  const int M = kCoulM, K = degree + 1;
  const long double pi = 3.14159265358979323846264338327950288L;
  const long double h = (long double)rc / M;
  const long double two_over_sqrtpi = 1.1283791670955125738961589031215452L;
  std::vector<double> tab((size_t)2 * K * M);
  // monomial coefficients of T_n(x): tcoef[n][d]
  std::vector<std::vector<long double>> tcoef(K, std::vector<long double>(K, 0.0L));
  tcoef[0][0] = 1;
  if (K > 1) tcoef[1][1] = 1;
  for (int n = 2; n < K; ++n)
    for (int d = 0; d < K; ++d)
      tcoef[n][d] = (d > 0 ? 2 * tcoef[n - 1][d - 1] : 0.0L) - tcoef[n - 2][d];
  for (int k = 0; k < M; ++k) {
    for (int f = 0; f < 2; ++f) {
      std::vector<long double> y(K), x(K);
      for (int j = 0; j < K; ++j) {
        x[j] = cosl(pi * (j + 0.5L) / K);  // node in [-1, 1]; u = x / 2
        const long double r = (k + 0.5L + 0.5L * x[j]) * h;
        const long double br = (long double)beta * r;
        y[j] = erfcl(br);
        if (f == 0) y[j] += two_over_sqrtpi * (long double)beta * r * expl(-br * br);
      }
      std::vector<long double> mono(K, 0.0L);
      for (int nn = 0; nn < K; ++nn) {
        long double a = 0;
        for (int j = 0; j < K; ++j) a += y[j] * cosl(pi * nn * (j + 0.5L) / K);
        a *= 2.0L / K;
        if (nn == 0) a *= 0.5L;
        for (int d = 0; d <= nn; ++d) mono[d] += a * tcoef[nn][d];  // in x = 2u
      }
      long double scale = 1;
      for (int d = 0; d < K; ++d) {  // x^d = 2^d u^d
        tab[(((size_t)f * (K / 2) + d / 2) * M + k) * 2 + (d % 2)] = (double)(mono[d] * scale);
        scale *= 2;
      }
    }
  }
  return tab;
}

// Largest fraction of a cutoff sphere that one stencil column (a prism of
// cross-section wx x wy, infinite along z for this purpose) can hold: the
// overlap of a ball with a symmetric box is largest when the ball sits at the
// box centre.  Midpoint rule over the column cross-section.
double column_fraction(double wx, double wy, double rc) {
  const int m = 128;
  double hx = 0.5 * wx / rc, hy = 0.5 * wy / rc;
  if (hx > 1.0) hx = 1.0;
  if (hy > 1.0) hy = 1.0;
  double acc = 0.0;
  for (int a = 0; a < m; ++a)
    for (int b = 0; b < m; ++b) {
      double x = ((a + 0.5) / m * 2.0 - 1.0) * hx, y = ((b + 0.5) / m * 2.0 - 1.0) * hy;
      double z2 = 1.0 - x * x - y * y;
      if (z2 > 0) acc += 2.0 * std::sqrt(z2);
    }
  return acc * (2.0 * hx / m) * (2.0 * hy / m) / (4.0 / 3.0 * kPi);
}

// Default list capacity.  VERLET: per column segment, 1.15 x the mean count
// in the largest column share of the list's cutoff sphere (rc, or rc + skin)
// + 64; RBL: the core region plus twice the mean shell sample + 64.
int default_cap(const rbmd_sys* s) {
  double rho = s->n / s->v.volume;
  double rc = s->v.rbl ? s->p.cut_off : std::sqrt(s->v.list_cut2);
  double c = rho * 4.0 / 3.0 * kPi * rc * rc * rc;
  int st = (2 * s->v.nc + 1);
  long maxc = (long)std::max(s->n, 1) * st * st * st;
  if (s->v.rbl) return round_up((int)std::min<long>(s->v.rs_num + (int)(2.0 * c / s->v.pice_num + 64), maxc), kMaxLanes);
  double f = column_fraction(s->v.dxdydz[0], s->v.dxdydz[1], rc);
  return (int)std::min<long>((long)(1.15 * f * c + 64), maxc);
}

// cap: the width the caller exports (0: default).  VERLET keeps one segment of
// round_up(cap) slots per stencil column, so no segment can overflow below cap.
rbmd_status ensure_nbr(rbmd_sys* s, int cap) {
  int seg_cap, nseg, export_cap;
  if (s->v.rbl) {
    nseg = 1;
    export_cap = cap > 0 ? cap : default_cap(s);
    seg_cap = round_up(export_cap, kMaxLanes);
    if (seg_cap < s->v.rs_num + 1) return fail(RBMD_ERR_ARG, "RBL capacity %d below the core region %d", cap, s->v.rs_num);
  } else {
    nseg = 9;
    if (cap > 0) {
      seg_cap = round_up(cap, kMaxLanes);
      export_cap = cap;
    } else {
      seg_cap = round_up(default_cap(s), kMaxLanes);
      export_cap = nseg * seg_cap;
    }
  }
  s->export_cap = export_cap;
  if (s->v.nbr && seg_cap == s->v.seg_cap && nseg == s->v.nseg) return RBMD_OK;
  if (s->v.nbr) {
    auto it = std::find(s->allocs.begin(), s->allocs.end(), (void*)s->v.nbr);
    if (it != s->allocs.end()) s->allocs.erase(it);
    s->bytes -= (int64_t)s->v.cap * s->n * 4;
    (void)hipFree(s->v.nbr);
    s->v.nbr = nullptr;
  }
  s->v.nseg = nseg;
  s->v.seg_cap = seg_cap;
  s->v.cap = nseg * seg_cap;
  s->graph_ok = false;
  return s->dalloc(&s->v.nbr, (size_t)s->v.cap * s->n);
}

// Put the reference geometry (per-step list) or the skin geometry into the
// system's view.  The cell and list buffers are shared: switching drops the
// built cells and lists, and a skin list is rebuilt on the next MD step.
rbmd_status use_geom(rbmd_sys* s, bool skin) {
  if (s->skin_active == skin) return RBMD_OK;
  const rbmd_sys::Geom& g = skin ? s->g_skin : s->g_ref;
  for (int d = 0; d < 3; ++d) {
    s->v.dims[d] = g.dims[d];
    s->v.dxdydz[d] = g.dxdydz[d];
    s->v.shift_len[d] = g.dxdydz[d] * (double)g.dims[d];
  }
  s->v.ncell = g.ncell;
  s->v.n_items = g.n_items;
  s->v.lds_bins = g.lds_bins;
  s->v.list_cut2 = g.list_cut2;
  s->v.skin_on = skin ? 1 : 0;
  s->skin_active = skin;
  s->skin_ready = false;
  s->locator_built = s->nbr_built = false;
  s->graph_ok = false;
  // the counters the cell kernels expect at zero (launch_fill: not a memset node)
  launch_fill_int(s->v.cell_count, 0, std::max(s->g_ref.ncell, s->g_skin.ncell), s->ctx->stream);
  launch_fill_int(s->v.cursor, 0, std::max(s->g_ref.ncell, s->g_skin.ncell), s->ctx->stream);
  return ensure_nbr(s, 0);
}

// Before MD steps of a skin-list system: the skin geometry, and a rebuild
// requested for the next step when the list is not current.
rbmd_status prepare_skin(rbmd_sys* s) {
  if (!s->skin || s->skin_ready) return RBMD_OK;
  RC_TRY(use_geom(s, true));
  launch_skin_request(s->v, 1, s->ctx->stream);
  s->skin_ready = true;
  return RBMD_OK;
}

rbmd_status read_status(rbmd_sys* s, int* st) {
  HIP_TRY(hipMemcpyAsync(st, s->v.status, sizeof(int), hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return RBMD_OK;
}

rbmd_status status_to_rc(int st) {
  if (st & 1) return fail(RBMD_ERR_NONFINITE, "non-finite position, velocity or force");
  if (st & 2) return fail(RBMD_ERR_SHAKE, "SHAKE/RATTLE did not converge or hit a singular matrix");
  if (st & 4) return fail(RBMD_ERR_CAPACITY, "a bounded buffer (neighbour list, cell sort, special list) overflowed");
  return RBMD_OK;
}

// --------------------------------------------------------------- schedule pieces

void build_cells(rbmd_sys* s, bool bin_from_x) {
  rbmd_ctx* c = s->ctx;
  if (bin_from_x) c->run("bin", [&] { launch_bin(s->v, c->stream, s->v.x, s->v.y, s->v.z); });
  if (scan_in_scatter(s->v)) {
    c->run("scan_scatter", [&] { launch_scan_scatter(s->v, c->stream); });
  } else {
    c->run("scan", [&] { launch_scan(s->v, c->stream); });
    c->run("scatter", [&] { launch_scatter(s->v, c->stream); });
  }
  c->run("cell_sort", [&] { launch_cell_sort(s->v, c->stream); });
}

// the stored lists (list path) on the sorted copies build_cells wrote
void gather_and_list(rbmd_sys* s) {
  rbmd_ctx* c = s->ctx;
  s->locator_built = true;
  if (s->fused) {
    s->nbr_built = false;
    return;
  }
  c->run("nbr_build", [&] { launch_nbr_build(s->v, c->stream); });
  s->nbr_built = true;
}

// reduction descriptor of the thermo rows (nseg = 0: no reduction)
RowsDesc rows_desc(rbmd_sys* s, bool reduce) {
  RowsDesc d;
  d.nseg = reduce ? NSEG : 0;
  for (int g = 0; g < NSEG; ++g)
    d.seg[g] = RowSeg{s->seg_rows(g), s->seg_used[g], kSegWidth[g], kSegSlot0[g], kSegGroups[g]};
  return d;
}

// real-space force on the current cell list: fused kernel or list pair kernel
void launch_real(rbmd_sys* s, int thermo, int peratom) {
  rbmd_ctx* c = s->ctx;
  int nblk = 0;
  if (s->fused) {
    s->v.nslab = 3;
    c->run("force_cells", [&] { launch_force_cells(s->v, c->stream, thermo, peratom, s->seg_rows(SEG_PAIR), &nblk); });
  } else {
    s->v.nslab = 1;
    c->run("pair", [&] { launch_pair(s->v, c->stream, thermo, peratom, s->seg_rows(SEG_PAIR), &nblk); });
    // RBL: the force kernels write no rows; the energy pass (every step that
    // forms thermo sums, as the reference forms them every step) writes the
    // energies and the mean-removed RBL virial
    if (s->v.rbl && (thermo || peratom))
      c->run("rbl_energy", [&] { launch_rbl_energy(s->v, c->stream, peratom, s->seg_rows(SEG_PAIR), &nblk); });
  }
  if (thermo || peratom) s->seg_used[SEG_PAIR] = nblk;
}

bool has_k(const rbmd_sys* s) {
  return s->p.force_field != RBMD_FF_LJ_CUT && s->p.coulomb_type != RBMD_COULOMB_NONE;
}
bool has_bonded(const rbmd_sys* s) {
  return s->p.force_field == RBMD_FF_CVFF && (s->nb > 0 || s->na > 0 || s->iv.spec_coul);
}
bool shake_on(const rbmd_sys* s) { return s->p.fix_shake == RBMD_SHAKE_TRUE && s->nmol > 0; }

rbmd_status ensure_rho(rbmd_sys* s, int P) {
  const int atoms = rho_atoms_per_block(s->n, rho_kblocks(s->v, P));
  const int nblk = (s->n + atoms - 1) / atoms;
  size_t need = (size_t)nblk * 2 * P;
  if ((size_t)s->rho_part_cap < need) {
    RC_TRY(s->dalloc(&s->rho_part, need));
    RC_TRY(s->dalloc(&s->rho, 2 * (size_t)P));
    RC_TRY(s->dalloc(&s->krec, (size_t)P));
    s->rho_part_cap = (int)need;
  }
  s->kv.nblk_rho = nblk;
  s->kv.rho_atoms = atoms;
  s->kv.rho_part = s->rho_part;
  s->kv.rho = s->rho;
  s->kv.krec = s->krec;
  return RBMD_OK;
}

// Full force evaluation on the current locator (ExecutionNVT::ComputeAllForce,
// ExecutionNVT.cpp:195-235).  kvec: device pointer to this step's k vectors.
// k-space and bonded forces on stream st (they read the unsorted positions
// and charges only, and write their own force arrays and thermo rows)
static void launch_k_bonded(rbmd_sys* s, const double* kvec, int thermo, hipStream_t st) {
  rbmd_ctx* c = s->ctx;
  if (has_k(s)) {
    KView k = s->kv;
    k.P = s->P;
    k.kvec = kvec;
    k.S_over_P = s->p.coulomb_type == RBMD_COULOMB_RBE ? s->S / (double)s->P : 1.0;
    k.f32 = s->v.mixed;
    k.ewald = s->p.coulomb_type == RBMD_COULOMB_EWALD;
    c->run("rho_partial", [&] { launch_rho_partial(s->v, k, st); });
    c->run("rho_reduce", [&] { launch_rho_reduce(s->v, k, st); });
    int nblk = 0;
    c->run("kforce", [&] { launch_kforce(s->v, k, st, thermo, 0, s->seg_rows(SEG_K), &nblk); });
    if (thermo) s->seg_used[SEG_K] = nblk;
  }
  if (has_bonded(s)) {
    int nblk = 0;
    c->run("bonded", [&] { launch_bonded(s->v, s->iv, st, thermo, s->seg_rows(SEG_BOND), &nblk); });
    if (thermo) s->seg_used[SEG_BOND] = nblk;
  }
}

void compute_force(rbmd_sys* s, const double* kvec, int thermo) {
  launch_real(s, thermo, 0);
  launch_k_bonded(s, kvec, thermo, s->ctx->stream);
}

// Refill the k-vector rings of the systems whose ring half for `step` holds
// another batch: the host draws the R steps' k vectors of every such system
// (one thread pool over all (system, step) pairs) into the pinned staging
// half, then each system's ring half is copied on its stream.  The staging
// half is reused only after its previous copy completed (its event).
// steps[k]: the step system k runs next.
rbmd_status ring_fill(rbmd_sys* const* sv, const int64_t* steps, int nsys) {
  std::vector<rbmd_sys*> todo;
  std::vector<int64_t> tstep;
  for (int k = 0; k < nsys; ++k) {
    rbmd_sys* s = sv[k];
    if (!has_k(s) || s->p.coulomb_type != RBMD_COULOMB_RBE || s->injected) continue;
    const int64_t batch = steps[k] / s->ring_R;
    if (s->ring_have[batch % 2] != batch) {
      todo.push_back(s);
      tstep.push_back(steps[k]);
    }
  }
  if (todo.empty()) return RBMD_OK;
  for (size_t k = 0; k < todo.size(); ++k)
    HIP_TRY(hipEventSynchronize(todo[k]->stage_ev[(tstep[k] / todo[k]->ring_R) % 2]));
  std::vector<std::pair<int, int>> work;  // (system, step within the batch)
  for (int k = 0; k < (int)todo.size(); ++k)
    for (int r = 0; r < todo[k]->ring_R; ++r) work.emplace_back(k, r);
  const unsigned nth = std::max(1u, std::min<unsigned>(16u, std::min<unsigned>(std::thread::hardware_concurrency(),
                                                                                  (unsigned)work.size())));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nth; ++t) {
    th.emplace_back([&, t] {
      std::vector<double> m;
      for (size_t w = t; w < work.size(); w += nth) {
        rbmd_sys* s = todo[work[w].first];
        const int r = work[w].second;
        const int64_t batch = tstep[work[w].first] / s->ring_R;
        const size_t per = (size_t)3 * s->P;
        m.resize(per);
        s->sampler.draw(s->p.seed, (uint64_t)(batch * s->ring_R + r), m.data());
        s->sampler.to_k(m.data(), s->stage[batch % 2] + r * per);
      }
    });
  }
  for (auto& t : th) t.join();
  for (size_t k = 0; k < todo.size(); ++k) {
    rbmd_sys* s = todo[k];
    const int64_t batch = tstep[k] / s->ring_R;
    const int half = (int)(batch % 2);
    const size_t per = (size_t)3 * s->P;
    HIP_TRY(hipMemcpyAsync(s->kring + (size_t)half * s->ring_R * per, s->stage[half], per * s->ring_R * sizeof(double),
                           hipMemcpyHostToDevice, s->ctx->stream));
    HIP_TRY(hipEventRecord(s->stage_ev[half], s->ctx->stream));
    s->ring_have[half] = batch;
  }
  return RBMD_OK;
}

// k vectors of one MD step (stream id = step index), staged through a
// double-buffered pinned ring so host sampling overlaps device work
rbmd_status kvec_for_step(rbmd_sys* s, int64_t step, const double** out) {
  if (!has_k(s) || s->p.coulomb_type != RBMD_COULOMB_RBE || s->injected) {
    *out = s->kcur;
    return RBMD_OK;
  }
  RC_TRY(ring_fill(&s, &step, 1));
  const int R = s->ring_R;
  *out = s->kring + ((size_t)((step / R) % 2) * R + (size_t)(step % R)) * ((size_t)3 * s->P);
  return RBMD_OK;
}

void fill_thermo(const rbmd_sys* s, const NhcState& S, int status, int max_iter, rbmd_thermo* o) {
  std::memset(o, 0, sizeof(*o));
  o->step = S.step;
  o->temperature = S.tempT;
  o->pressure = S.pressure;
  o->ke_sum = S.tempT_sum;
  o->e_lj = S.e_lj;
  o->e_coul = S.e_coul;
  o->e_long = s->e_long;
  o->e_bond = S.e_bond;
  o->e_angle = S.e_angle;
  for (int k = 0; k < 6; ++k) o->virial[k] = S.virial[k];
  for (int k = 0; k < 3; ++k) {
    o->nhc_eta[k] = S.eta[k];
    o->nhc_eta_dot[k] = S.eta_dot[k];
  }
  o->energy_step = s->energy_step;
  o->shake_max_iter = max_iter;
  o->status_flags = status;
}

// e_long = qqr2e (2 pi / V) sum_grid e^{-k^2/4a}|rho|^2/k^2 / N + self / N
// (ComputeEwaldEnergy + ComputeSelfEnergy, ExecutionMD.cpp:405-443, 508-558).
// The k grid, the structure-factor partials and the reduced rho live in
// device memory allocated for the call, or, with `scratch` (the no-alloc
// entry point), in that buffer of scratch_bytes, which every launch below
// overwrites in full before reading it (no clearing needed).
rbmd_status long_energy(rbmd_sys* s, const int kmax[3], double* e_avg, double* grid_sum_out,
                        void* scratch = nullptr, size_t scratch_bytes = 0) {
  std::vector<double> kg;
  for (int i = -kmax[0]; i <= kmax[0]; ++i)
    for (int j = -kmax[1]; j <= kmax[1]; ++j)
      for (int k = -kmax[2]; k <= kmax[2]; ++k)
        if (i || j || k) {
          kg.push_back(2 * kPi * i / s->v.box[0]);
          kg.push_back(2 * kPi * j / s->v.box[1]);
          kg.push_back(2 * kPi * k / s->v.box[2]);
        }
  int K = (int)(kg.size() / 3);
  const int atoms = rho_atoms_per_block(s->n, rho_kblocks(s->v, K));
  const int nblk = (s->n + atoms - 1) / atoms;
  // three slices, each starting on a 256-byte boundary
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t b_k = up(kg.size() * sizeof(double)), b_rp = up((size_t)nblk * 2 * K * sizeof(double)),
               b_rr = up((size_t)2 * K * sizeof(double));
  double *kdev = nullptr, *rp = nullptr, *rr = nullptr;
  if (scratch) {
    if (b_k + b_rp + b_rr > scratch_bytes)
      return fail(RBMD_ERR_ARG,
                  "the kmax grid (%d k vectors over %d atom blocks) needs %zu bytes of scratch; the no-alloc "
                  "scratch holds %zu",
                  K, nblk, b_k + b_rp + b_rr, scratch_bytes);
    kdev = static_cast<double*>(scratch);
    rp = reinterpret_cast<double*>(static_cast<char*>(scratch) + b_k);
    rr = reinterpret_cast<double*>(static_cast<char*>(scratch) + b_k + b_rp);
  } else {
    HIP_TRY(hipMalloc(&kdev, b_k + b_rp + b_rr));
    rp = reinterpret_cast<double*>(reinterpret_cast<char*>(kdev) + b_k);
    rr = reinterpret_cast<double*>(reinterpret_cast<char*>(kdev) + b_k + b_rp);
  }
  std::vector<double> rho((size_t)2 * K);
  std::vector<double> q(s->n);
  hipError_t e = hipMemcpyAsync(kdev, kg.data(), kg.size() * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream);
  if (e == hipSuccess) {
    KView k{K, 1.0, kdev, rp, rr, nblk, 0, nullptr, 1, atoms};  // energy only: no force records
    launch_rho_partial(s->v, k, s->ctx->stream);
    launch_rho_reduce(s->v, k, s->ctx->stream);
    e = hipMemcpyAsync(rho.data(), rr, rho.size() * sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(q.data(), s->v.q, s->n * sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
  if (!scratch) (void)hipFree(kdev);
  if (e != hipSuccess) return fail(RBMD_ERR_HIP, "Ewald energy: %s", hipGetErrorString(e));
  double sum = 0, cmp = 0;
  for (int l = 0; l < K; ++l) {
    double kx = kg[3 * l], ky = kg[3 * l + 1], kz = kg[3 * l + 2];
    double k2 = kx * kx + ky * ky + kz * kz;
    double term = std::exp(-k2 / (4 * s->p.alpha)) * (rho[2 * l] * rho[2 * l] + rho[2 * l + 1] * rho[2 * l + 1]) / k2;
    double t = sum + term;
    cmp += std::fabs(sum) >= std::fabs(term) ? (sum - t) + term : (term - t) + sum;
    sum = t;
  }
  sum += cmp;
  double q2 = 0;
  for (int i = 0; i < s->n; ++i) q2 += q[i] * q[i];
  double self = -std::sqrt(s->p.alpha / kPi) * q2;
  if (grid_sum_out) *grid_sum_out = sum;
  *e_avg = s->u.qqr2e * sum * (2 * kPi / s->v.volume) / s->n + s->u.qqr2e * self / s->n;
  return check_launch("long_energy");
}

}  // namespace

namespace rbmd {

int32_t api_fail(int32_t code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int ctx_device(const rbmd_ctx* c) { return c->device; }
hipStream_t ctx_stream(const rbmd_ctx* c) { return c->stream; }

}  // namespace rbmd

// =================================================================== C ABI

extern "C" {

const char* rbmd_last_error(void) { return g_last_error.c_str(); }
const char* rbmd_version(void) { return "rbmd-mi355x 0.1.0 (gfx950)"; }

rbmd_status rbmd_ctx_create(int32_t device, rbmd_ctx** out) {
  if (!out) return fail(RBMD_ERR_ARG, "null out");
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0) return fail(RBMD_ERR_NODEVICE, "no HIP device visible");
  if (device < 0 || device >= cnt) return fail(RBMD_ERR_ARG, "device %d out of range (%d devices)", device, cnt);
  HIP_TRY(hipSetDevice(device));
  // the no-alloc scratch's address on this device (taken once per device)
  const hipError_t se = scratch_init(device);
  if (se != hipSuccess) return fail(RBMD_ERR_HIP, "scratch address on device %d: %s", device, hipGetErrorString(se));
  rbmd_ctx* c = new rbmd_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
  if (e != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(RBMD_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  *out = c;
  return RBMD_OK;
}

rbmd_status rbmd_ctx_destroy(rbmd_ctx* c) {
  if (!c) return RBMD_OK;
  (void)hipSetDevice(c->device);
  c->flush();
  for (auto e : c->pool) (void)hipEventDestroy(e);
  (void)hipStreamSynchronize(c->stream);
  (void)hipStreamSynchronize(c->side);
  (void)hipStreamDestroy(c->stream);
  (void)hipStreamDestroy(c->side);
  delete c;
  return RBMD_OK;
}

rbmd_status rbmd_ctx_synchronize(rbmd_ctx* c) {
  if (!c) return fail(RBMD_ERR_ARG, "null ctx");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RBMD_OK;
}

rbmd_status rbmd_footprint_estimate(const rbmd_sys_desc* d, const rbmd_params* p, int64_t* bytes) {
  if (!d || !p || !bytes || d->n_atoms <= 0) return fail(RBMD_ERR_ARG, "bad arguments");
  int dims[3], nc;
  double dx[3];
  geometry(*p, d->lo, d->hi, dims, &nc, dx);
  double vol = (d->hi[0] - d->lo[0]) * (d->hi[1] - d->lo[1]) * (d->hi[2] - d->lo[2]);
  double c = d->n_atoms / vol * 4.0 / 3.0 * kPi * p->cut_off * p->cut_off * p->cut_off;
  int64_t n = d->n_atoms;
  int64_t cells = (int64_t)dims[0] * dims[1] * dims[2];
  // list slots per atom as ensure_nbr sizes them (VERLET: 9 column segments)
  int64_t slots;
  if (p->neighbor_type == RBMD_NEIGHBOR_RBL) {
    slots = (int64_t)(2.0 * c + 64);
  } else {
    const double f = column_fraction(dx[0], dx[1], p->cut_off);
    slots = 9 * (int64_t)round_up((int)(1.15 * f * c + 64), kMaxLanes);
  }
  int64_t b = n * (8 * 60 + 4 * 24) + cells * 16 + slots * n * 4;
  b += (int64_t)3 * std::max(p->rbe_p, 1) * 8 * 512 + ((n + 1023) / 1024) * 2 * std::max(p->rbe_p, 1) * 8;
  *bytes = b;
  return RBMD_OK;
}

rbmd_status rbmd_sys_create(rbmd_ctx* ctx, const rbmd_sys_desc* d, const rbmd_params* p, rbmd_sys** out) {
  if (!ctx || !d || !p || !out) return fail(RBMD_ERR_ARG, "null argument");
  if (d->n_atoms <= 0 || d->n_types <= 0) return fail(RBMD_ERR_ARG, "n_atoms and n_types must be positive");
  if (!d->type || !d->charge || !d->mass || !d->eps || !d->sigma) return fail(RBMD_ERR_ARG, "missing per-atom arrays");
  if (d->n_atoms >= (1 << (31 - kShiftBits))) return fail(RBMD_ERR_UNSUPPORTED, "n_atoms above %d", 1 << (31 - kShiftBits));
  if (p->cut_off <= 0) return fail(RBMD_ERR_ARG, "cut_off must be positive");
  if (p->precision != RBMD_PREC_FP64 && p->precision != RBMD_PREC_MIXED) return fail(RBMD_ERR_ARG, "bad precision");
  if (p->precision == RBMD_PREC_MIXED) {
    if (p->neighbor_type == RBMD_NEIGHBOR_RBL)
      return fail(RBMD_ERR_UNSUPPORTED, "mixed precision is built for the Verlet-list path only");
    if (p->pair_path == RBMD_PAIR_CELLS)
      return fail(RBMD_ERR_UNSUPPORTED, "mixed precision is built for pair_path lists only");
    if (d->n_types > kMaxTypes) return fail(RBMD_ERR_UNSUPPORTED, "mixed precision needs n_types <= %d", kMaxTypes);
  }
  if (p->neighbor_type == RBMD_NEIGHBOR_RBL && p->force_field == RBMD_FF_LJ_CUT)
    return fail(RBMD_ERR_UNSUPPORTED,
                "LJ/CUT over random-batch lists (the reference's LJForceRBL path, ExecutionMD.cpp:752-830) is not built");
  if (p->neighbor_type == RBMD_NEIGHBOR_RBL && p->erfc_mode == RBMD_ERFC_TABLE && !(p->cut_off < 12.0))
    return fail(RBMD_ERR_UNSUPPORTED, "the erfc step table covers r < 12 (ExecStaticTable.h:96-135); cut_off %g",
                p->cut_off);
  if (p->neighbor_type == RBMD_NEIGHBOR_RBL && !(p->r_core > 0 && p->r_core <= p->cut_off))
    return fail(RBMD_ERR_ARG, "r_core must be in (0, cut_off]");
  if (p->shake_tol < 0) return fail(RBMD_ERR_ARG, "shake_tol must be >= 0");
  for (int k = 0; k < 3; ++k)
    if (!(d->hi[k] > d->lo[k])) return fail(RBMD_ERR_ARG, "empty box along axis %d", k);
  for (int i = 0; i < d->n_atoms; ++i) {
    if (d->type[i] < 0 || d->type[i] >= d->n_types) return fail(RBMD_ERR_ARG, "atom %d has type %d", i, d->type[i]);
    if (!(d->mass[i] > 0)) return fail(RBMD_ERR_ARG, "atom %d has non-positive mass", i);
  }
  if (p->pair_path != RBMD_PAIR_CELLS && p->pair_path != RBMD_PAIR_LISTS) return fail(RBMD_ERR_ARG, "bad pair_path");
  if (p->pair_lanes != 0 && p->pair_lanes != 4 && p->pair_lanes != 8 && p->pair_lanes != 16)
    return fail(RBMD_ERR_ARG, "pair_lanes must be 0, 4, 8 or 16");
  if (p->kforce_lanes != 0 && p->kforce_lanes != 2 && p->kforce_lanes != 4 && p->kforce_lanes != 8)
    return fail(RBMD_ERR_ARG, "kforce_lanes must be 0, 2, 4 or 8");
  if (p->rho_kblock != 0 && p->rho_kblock != 1 && p->rho_kblock != 2 && p->rho_kblock != 4)
    return fail(RBMD_ERR_ARG, "rho_kblock must be 0, 1, 2 or 4");
  if (p->coulomb_type == RBMD_COULOMB_RBE && p->rbe_p <= 0) return fail(RBMD_ERR_ARG, "coulomb_sample_num must be > 0");
  if (p->coulomb_type != RBMD_COULOMB_NONE && !(p->alpha > 0)) return fail(RBMD_ERR_ARG, "alpha must be > 0");
  if (d->special_offsets && d->special_ids && d->special_weights) {
    if (d->special_offsets[0] != 0) return fail(RBMD_ERR_ARG, "special_offsets[0] must be 0");
    for (int i = 0; i < d->n_atoms; ++i)
      if (d->special_offsets[i + 1] < d->special_offsets[i]) return fail(RBMD_ERR_ARG, "special_offsets not monotonic");
    for (int k = 0; k < d->special_offsets[d->n_atoms]; ++k)
      if (d->special_ids[k] < 0 || d->special_ids[k] >= d->n_atoms) return fail(RBMD_ERR_ARG, "special id %d out of range", k);
  }
  HIP_TRY(hipSetDevice(ctx->device));
  rbmd_sys* s = new rbmd_sys();
  s->ctx = ctx;
  s->p = *p;
  if (hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess) {
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    delete s;
    return fail(RBMD_ERR_HIP, "hipEventCreate");
  }
  s->u = units_of(p->unit);
  int n = s->n = d->n_atoms;
  s->ntypes = d->n_types;
  SysView& v = s->v;
  std::memset(&v, 0, sizeof(v));
  v.n = n;
  v.ntypes = d->n_types;
  for (int k = 0; k < 3; ++k) {
    v.lo[k] = d->lo[k];
    v.hi[k] = d->hi[k];
    v.box[k] = d->hi[k] - d->lo[k];
  }
  v.volume = v.box[0] * v.box[1] * v.box[2];
  geometry(*p, d->lo, d->hi, v.dims, &v.nc, v.dxdydz);
  for (int k = 0; k < 3; ++k) v.shift_len[k] = v.dxdydz[k] * (double)v.dims[k];
  v.ncell = v.dims[0] * v.dims[1] * v.dims[2];
  v.cut = p->cut_off;
  v.cut2 = p->cut_off * p->cut_off;
  v.rs = p->neighbor_type == RBMD_NEIGHBOR_RBL ? p->r_core : p->cut_off;
  v.rs2 = v.rs * v.rs;
  v.alpha = p->alpha;
  v.sqrt_alpha = std::sqrt(p->alpha);
  v.qqr2e = s->u.qqr2e;
  v.lj_only = p->force_field == RBMD_FF_LJ_CUT;
  v.rbl = p->neighbor_type == RBMD_NEIGHBOR_RBL;
  v.erfc_table = p->erfc_mode == RBMD_ERFC_TABLE;
  v.mixed = p->precision == RBMD_PREC_MIXED ? 1 : 0;
  // accepted pairs have |r| < rc, so the minimum image is a no-op unless an edge is below 2 rc
  v.use_min_image = (v.box[0] < 2 * v.cut || v.box[1] < 2 * v.cut || v.box[2] < 2 * v.cut) ? 1 : 0;
  // the fused kernel keeps the type-pair tables in LDS; more types take the list path
  s->fused = !v.rbl && p->pair_path == RBMD_PAIR_CELLS && d->n_types <= kMaxTypes;
  bool have_sp = d->special_offsets && d->special_ids && d->special_weights;
  v.weighted = (have_sp && p->special_bonds) ? 1 : 0;
  if (v.rbl) rbl_capacity(*p, n, v.volume, &v.rs_num, &v.pice_num);
  v.n_items = v.ncell + n / kChunk + 1;
  v.lds_bins = v.ncell <= 4096 ? 1 : 0;
  v.list_cut2 = v.cut2;
  s->g_ref = rbmd_sys::Geom{{v.dims[0], v.dims[1], v.dims[2]}, {v.dxdydz[0], v.dxdydz[1], v.dxdydz[2]},
                            v.ncell, v.n_items, v.lds_bins, v.cut2};
  s->g_skin = s->g_ref;
  if (p->skin > 0) {
    // a list to rc + skin over cells at least that wide; every box edge at least
    // 2 (rc + skin), so no pair vector needs the minimum image
    const double rl = p->cut_off + p->skin;
    if (v.rbl || p->pair_path != RBMD_PAIR_LISTS) {
      delete s;
      return fail(RBMD_ERR_UNSUPPORTED, "a list skin is built for the VERLET lists path (pair_path lists)");
    }
    for (int k = 0; k < 3; ++k)
      if (v.box[k] < 2 * rl) {
        delete s;
        return fail(RBMD_ERR_UNSUPPORTED, "a list skin needs box edges >= 2 (cut_off + skin) = %g", 2 * rl);
      }
    rbmd_sys::Geom g{};
    g.ncell = 1;
    for (int k = 0; k < 3; ++k) {
      g.dims[k] = std::max(1, (int)(v.box[k] / rl));
      g.dxdydz[k] = v.box[k] / g.dims[k];
      g.ncell *= g.dims[k];
    }
    g.n_items = g.ncell + n / kChunk + 1;
    g.lds_bins = g.ncell <= 4096 ? 1 : 0;
    g.list_cut2 = rl * rl;
    s->g_skin = g;
    s->skin = true;
    v.half_skin2 = 0.25 * p->skin * p->skin;
  }
  const int ncell_max = std::max(s->g_ref.ncell, s->g_skin.ncell);
  const int items_max = std::max(s->g_ref.n_items, s->g_skin.n_items);
  // lanes per atom in the pair kernel: enough threads to fill 256 CUs for
  // small systems, fewer lanes (longer per-lane loops) for large ones
  v.lanes_lg = n <= 60000 ? 4 : (n <= 250000 ? 3 : 2);
  if (p->pair_lanes) v.lanes_lg = p->pair_lanes == 4 ? 2 : (p->pair_lanes == 8 ? 3 : 4);
  v.k_lanes_lg = p->kforce_lanes ? (p->kforce_lanes == 2 ? 1 : (p->kforce_lanes == 8 ? 3 : 2)) : 2;
  v.rho_kb = p->rho_kblock ? p->rho_kblock : 1;
  v.sp_max = 0;
  if (d->special_offsets && d->special_ids && d->special_weights)
    for (int i = 0; i < n; ++i) v.sp_max = std::max(v.sp_max, d->special_offsets[i + 1] - d->special_offsets[i]);
  if (v.sp_max > 16) {
    delete s;
    return fail(RBMD_ERR_UNSUPPORTED, "special lists longer than 16 entries");
  }

  rbmd_status rc = RBMD_OK;
  auto A = [&](rbmd_status r) {
    if (r != RBMD_OK && rc == RBMD_OK) rc = r;
  };
  // per-atom state
  double** fields[] = {&v.x, &v.y, &v.z, &v.vx, &v.vy, &v.vz, &v.fx, &v.fy, &v.fz, &v.ox, &v.oy, &v.oz,
                       &v.ovx, &v.ovy, &v.ovz, &v.fkx, &v.fky, &v.fkz, &v.fbx, &v.fby, &v.fbz};
  for (double** f : fields) A(s->dalloc(f, n));
  // real-space outputs hold up to 3 slab partials (fused path)
  double** slabbed[] = {&v.fnx, &v.fny, &v.fnz, &v.e_lj, &v.e_coul};
  for (double** f : slabbed) A(s->dalloc(f, 3 * (size_t)n));
  A(s->dalloc(&v.vir_near, 18 * (size_t)n));
  v.nslab = 1;
  A(s->upload(&v.q, d->charge, n));
  A(s->upload(&v.mass, d->mass, n));
  A(s->upload(&v.type, d->type, n));
  A(s->dalloc(&v.img, 3 * (size_t)n));
  // pair tables: eps_ij = sqrt(eps_i eps_j), sigma_ij = (sigma_i + sigma_j)/2 (ExecForceFunction.h:424-431)
  std::vector<double> eij((size_t)v.ntypes * v.ntypes), s6((size_t)v.ntypes * v.ntypes);
  for (int a = 0; a < v.ntypes; ++a)
    for (int b = 0; b < v.ntypes; ++b) {
      eij[a * v.ntypes + b] = std::sqrt(d->eps[a] * d->eps[b]);
      double sij = (d->sigma[a] + d->sigma[b]) / 2;
      s6[a * v.ntypes + b] = sij * sij * sij * sij * sij * sij;
    }
  A(s->upload(&v.eps_ij, eij.data(), eij.size()));
  A(s->upload(&v.sig6_ij, s6.data(), s6.size()));
  {
    const double beta = v.sqrt_alpha > 0 ? v.sqrt_alpha : 0.0;
    std::vector<double> ct = coul_table(beta, p->cut_off, kCoulD);
    double* dct = nullptr;
    A(s->upload(&dct, ct.data(), ct.size()));
    v.ctab = reinterpret_cast<const double2*>(dct);
    v.ctab_inv_h = kCoulM / p->cut_off;
    if (v.mixed) {
      std::vector<double> c32 = coul_table(beta, p->cut_off, kCoulD32);
      std::vector<float> f32(c32.begin(), c32.end());
      float* dft = nullptr;
      A(s->upload(&dft, f32.data(), f32.size()));
      v.ctab32 = reinterpret_cast<const float2*>(dft);
    }
  }
  if (have_sp) {
    int ns = d->special_offsets[n];
    A(s->upload(&v.sp_off, d->special_offsets, (size_t)n + 1));
    A(s->upload(&v.sp_ids, d->special_ids, (size_t)std::max(ns, 1)));
    A(s->upload(&v.sp_w, d->special_weights, (size_t)std::max(ns, 1)));
  }
  // cells and lists
  A(s->dalloc(&v.cell_of, n));
  A(s->dalloc(&v.cell_count, ncell_max));
  A(s->dalloc(&v.cell_start, (size_t)ncell_max + 1));
  A(s->dalloc(&v.cursor, ncell_max));
  A(s->dalloc(&v.point_ids, n));
  A(s->dalloc(&v.rank, n));
  A(s->dalloc(&v.items, items_max));
  A(s->dalloc(&v.n_items_dev, 1));
  if (s->skin) {
    A(s->dalloc(&v.x0, n));
    A(s->dalloc(&v.y0, n));
    A(s->dalloc(&v.z0, n));
    A(s->dalloc(&v.rebuild_at, 1));
  }
  A(s->dalloc(&v.list_builds, 1));
  A(s->dalloc(&v.misfit, n));
  A(s->dalloc(&v.misfit_list, n));
  A(s->dalloc(&v.misfit_count, 1));
  A(s->dalloc(&v.spos, n));
  A(s->dalloc(&v.stype, n));
  A(s->dalloc(&v.nbr_num, 9 * (size_t)n));
  A(s->dalloc(&v.vir_k, 6 * (size_t)n));
  A(s->dalloc(&v.status, 1));
  if (rc == RBMD_OK) A(ensure_nbr(s, 0));
  // bonded terms and SHAKE triplets
  s->nb = d->n_bonds;
  s->na = d->n_angles;
  IntView& iv = s->iv;
  std::memset(&iv, 0, sizeof(iv));
  iv.dt = p->dt;
  iv.fmt2v = s->u.fmt2v;
  iv.mvv2e = s->u.mvv2e;
  iv.nb = s->nb;
  iv.na = s->na;
  iv.shake_tol = p->shake_tol > 0 ? p->shake_tol : 0.00001;
  {  // the rigid geometry's H-H distance, once, as the oracle forms it
    const double b1 = 1.0, b2 = 1.0;
    iv.shake_d12 = std::sqrt(b1 * b1 + b2 * b2 - 2.0 * b1 * b2 * std::cos((109.4700 / 180.0) * kPi));
  }
  // molecule groups of SpecialCoulGeneral: every atom lists the atoms of its
  // molecule in ascending id (SetMolecularGroup, ModelFileInitCondition.cpp:698-736);
  // the correction runs for CVFF over random-batch lists with special bonds
  // (ExecutionNVT.cpp:212-214, SpecialCoulForce :673-702)
  if (v.rbl && p->force_field == RBMD_FF_CVFF && p->special_bonds && d->mol_id) {
    int nm = 0;
    for (int i = 0; i < n; ++i) {
      if (d->mol_id[i] < 0) {
        rc = fail(RBMD_ERR_ARG, "atom %d has a negative molecule id", i);
        break;
      }
      nm = std::max(nm, d->mol_id[i] + 1);
    }
    if (rc == RBMD_OK) {
      std::vector<int> moff(nm + 1, 0);
      for (int i = 0; i < n; ++i) moff[d->mol_id[i] + 1]++;
      for (int m = 0; m < nm; ++m) moff[m + 1] += moff[m];
      std::vector<int> members(n), cur(moff.begin(), moff.end() - 1);
      for (int i = 0; i < n; ++i) members[cur[d->mol_id[i]]++] = i;  // ascending id within a molecule
      std::vector<int> goff(n + 1, 0), gids;
      for (int i = 0; i < n; ++i) {
        const int m = d->mol_id[i];
        gids.insert(gids.end(), members.begin() + moff[m], members.begin() + moff[m + 1]);
        goff[i + 1] = (int)gids.size();
      }
      int* tmp = nullptr;
      A(s->upload(&tmp, goff.data(), goff.size()));
      iv.grp_off = tmp;
      A(s->upload(&tmp, gids.data(), std::max<size_t>(gids.size(), 1)));
      iv.grp_ids = tmp;
      iv.spec_coul = 1;
    }
  }
  {
    std::vector<int> boff(n + 1, 0), aoff(n + 1, 0);
    for (int b = 0; b < s->nb; ++b) {
      int i0 = d->bond_atoms[2 * b], i1 = d->bond_atoms[2 * b + 1];
      if (i0 < 0 || i0 >= n || i1 < 0 || i1 >= n) {
        rc = fail(RBMD_ERR_ARG, "bond %d references atom out of range", b);
        break;
      }
      boff[i0 + 1]++;
      if (i1 != i0) boff[i1 + 1]++;
    }
    for (int a = 0; a < s->na; ++a)
      for (int t = 0; t < 3; ++t) {
        int id = d->angle_atoms[3 * a + t];
        if (id < 0 || id >= n) {
          rc = fail(RBMD_ERR_ARG, "angle %d references atom out of range", a);
          break;
        }
      }
    if (rc == RBMD_OK) {
      std::vector<int> incb, inca;
      for (int i = 0; i < n; ++i) boff[i + 1] += boff[i];
      incb.resize(boff[n]);
      std::vector<int> cur(boff.begin(), boff.end() - 1);
      for (int b = 0; b < s->nb; ++b) {
        int i0 = d->bond_atoms[2 * b], i1 = d->bond_atoms[2 * b + 1];
        incb[cur[i0]++] = b;
        if (i1 != i0) incb[cur[i1]++] = b;
      }
      for (int a = 0; a < s->na; ++a) {
        int ids[3] = {d->angle_atoms[3 * a], d->angle_atoms[3 * a + 1], d->angle_atoms[3 * a + 2]};
        for (int t = 0; t < 3; ++t) {
          bool dup = false;
          for (int u = 0; u < t; ++u) dup |= ids[u] == ids[t];
          if (!dup) aoff[ids[t] + 1]++;
        }
      }
      for (int i = 0; i < n; ++i) aoff[i + 1] += aoff[i];
      inca.resize(aoff[n]);
      std::vector<int> cura(aoff.begin(), aoff.end() - 1);
      for (int a = 0; a < s->na; ++a) {
        int ids[3] = {d->angle_atoms[3 * a], d->angle_atoms[3 * a + 1], d->angle_atoms[3 * a + 2]};
        for (int t = 0; t < 3; ++t) {
          bool dup = false;
          for (int u = 0; u < t; ++u) dup |= ids[u] == ids[t];
          if (!dup) inca[cura[ids[t]]++] = a;
        }
      }
      int* tmp;
      A(s->upload(&tmp, boff.data(), boff.size()));
      iv.inc_b_off = tmp;
      A(s->upload(&tmp, incb.data(), std::max<size_t>(incb.size(), 1)));
      iv.inc_b = tmp;
      A(s->upload(&tmp, aoff.data(), aoff.size()));
      iv.inc_a_off = tmp;
      A(s->upload(&tmp, inca.data(), std::max<size_t>(inca.size(), 1)));
      iv.inc_a = tmp;
      if (s->nb > 0) {
        A(s->upload(&tmp, d->bond_atoms, 2 * (size_t)s->nb));
        iv.bond_atoms = tmp;
        double* tk;
        A(s->upload(&tk, d->bond_k, s->nb));
        iv.bond_k = tk;
        A(s->upload(&tk, d->bond_r0, s->nb));
        iv.bond_r0 = tk;
      }
      if (s->na > 0) {
        A(s->upload(&tmp, d->angle_atoms, 3 * (size_t)s->na));
        iv.angle_all = tmp;
        double* tk;
        A(s->upload(&tk, d->angle_k, s->na));
        iv.angle_k = tk;
        A(s->upload(&tk, d->angle_theta0, s->na));
        iv.angle_t0 = tk;
        s->host_angles.assign(d->angle_atoms, d->angle_atoms + 3 * (size_t)s->na);
      }
    }
  }
  if (p->fix_shake == RBMD_SHAKE_TRUE && s->na > 0) {
    s->nmol = s->na;
    iv.nmol = s->na;
    iv.angle_atoms = iv.angle_all;
    std::vector<char> in(n, 0);
    for (int a = 0; a < s->na; ++a)
      for (int t = 0; t < 3; ++t) in[d->angle_atoms[3 * a + t]] = 1;
    std::vector<int> fr;
    for (int i = 0; i < n; ++i)
      if (!in[i]) fr.push_back(i);
    s->nfree = (int)fr.size();
    if (s->nfree) A(s->upload(&s->free_dev, fr.data(), fr.size()));
  }
  // reduction rows: one segment per producer, sized for its largest launch
  {
    const int nblk = (n + kBlock - 1) / kBlock;
    const int pair_rows = std::max(std::max(pair_blocks(v), force_cells_blocks(v)), v.rbl ? rbl_energy_blocks(v) : 0);
    const int cap[NSEG] = {pair_rows, kforce_blocks(v), nblk, nblk,
                           (std::max(s->nmol, 1) + kBlock - 1) / kBlock};
    size_t off = 0;
    for (int g = 0; g < NSEG; ++g) {
      s->seg_off[g] = off;
      off += (size_t)cap[g] * kSegWidth[g];
    }
    s->rows_len = off;
    A(s->dalloc(&s->rows, off));
  }
  A(s->dalloc(&s->nhc, 1));
  if (s->nhc) v.cur_step = &s->nhc->step;  // a device address: the skin kernels read the step counter there
  A(s->dalloc(&s->shake_iter, 1));
  // k space
  if (p->coulomb_type != RBMD_COULOMB_NONE && !v.lj_only) {
    s->P = p->coulomb_type == RBMD_COULOMB_RBE ? p->rbe_p : 1;
    if (p->coulomb_type == RBMD_COULOMB_RBE) {
      s->sampler = KSampler(p->alpha, v.box, p->rbe_p);
      s->S = s->sampler.S();
      A(s->dalloc(&s->kcur, 3 * (size_t)s->P));
      A(s->dalloc(&s->kring, (size_t)2 * s->ring_R * 3 * s->P));
      for (int h = 0; h < 2; ++h) {
        hipError_t e = hipHostMalloc((void**)&s->stage[h], (size_t)s->ring_R * 3 * s->P * sizeof(double));
        if (e != hipSuccess) A(fail(RBMD_ERR_OOM, "hipHostMalloc staging"));
        e = hipEventCreateWithFlags(&s->stage_ev[h], hipEventDisableTiming);
        if (e != hipSuccess) A(fail(RBMD_ERR_HIP, "hipEventCreate"));
      }
      A(ensure_rho(s, s->P));
    } else {
      // Ewald force over the kmax grid in the reference's k order
      std::vector<double> kg;
      for (int i = -p->kmax[0]; i <= p->kmax[0]; ++i)
        for (int j = -p->kmax[1]; j <= p->kmax[1]; ++j)
          for (int k = -p->kmax[2]; k <= p->kmax[2]; ++k)
            if (i || j || k) {
              kg.push_back(2 * kPi * i / v.box[0]);
              kg.push_back(2 * kPi * j / v.box[1]);
              kg.push_back(2 * kPi * k / v.box[2]);
            }
      s->P = (int)(kg.size() / 3);
      A(s->upload(&s->kcur, kg.data(), kg.size()));
      A(ensure_rho(s, s->P));
    }
  }
  if (rc != RBMD_OK) {
    rbmd_sys_destroy(s);
    return rc;
  }
  hipStream_t st = ctx->stream;
  launch_fill_int(v.cell_count, 0, ncell_max, st);
  launch_fill_int(v.cursor, 0, ncell_max, st);  // the one-launch scan + scatter counts from zero
  launch_fill_int(v.status, 0, 1, st);
  launch_fill_int(v.img, 0, 3 * n, st);
  launch_fill_int(v.list_builds, 0, 1, st);
  if (s->skin) launch_fill_int(reinterpret_cast<int*>(v.rebuild_at), 0, 2, st);  // int64 zero
  launch_fill_double(s->rows, 0.0, (int)s->rows_len, st);
  launch_fill_double(v.fkx, 0.0, n, st);
  launch_fill_double(v.fky, 0.0, n, st);
  launch_fill_double(v.fkz, 0.0, n, st);
  launch_fill_double(v.fbx, 0.0, n, st);
  launch_fill_double(v.fby, 0.0, n, st);
  launch_fill_double(v.fbz, 0.0, n, st);
  // thermostat parameters (ExecutionNVT::InitialCondition, ExecutionNVT.cpp:139-187)
  NhcParams& np = s->np;
  std::memset(&np, 0, sizeof(np));
  np.tdof = 3 * n - 3;
  if (p->fix_shake == RBMD_SHAKE_TRUE) np.tdof -= n;
  np.mode = p->temp_ctrl;
  np.inbuild = p->inbuild;
  np.fix_shake = p->fix_shake;
  np.n = n;
  np.nb = s->nb;
  np.na = s->na;
  np.boltz = s->u.boltz;
  np.kB = s->u.kB;
  np.nktv2p = s->u.nktv2p;
  np.qqr2e = s->u.qqr2e;
  np.t_start = p->temperature[0];
  np.t_stop = p->temperature[1];
  np.t_period = p->temperature[2];
  np.t_freq = 1 / np.t_period;
  np.dt = p->dt;
  np.dthalf = 0.5 * p->dt;
  np.dt4 = 0.25 * p->dt;
  np.dt8 = 0.125 * p->dt;
  np.tdrag = 1.0 - (p->dt * np.t_freq * 0.0 / 1);
  np.volume_range = v.volume;
  np.num_steps = p->num_steps;
  np.lj_only = v.lj_only;
  np.cvff = p->force_field == RBMD_FF_CVFF;
  np.shake_vir = p->fix_shake == RBMD_SHAKE_TRUE && s->na > 0;
  np.shake_iter = s->shake_iter;
  NhcState zero;
  std::memset(&zero, 0, sizeof(zero));
  hipError_t e = hipMemcpy(s->nhc, &zero, sizeof(zero), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    rbmd_sys_destroy(s);
    return fail(RBMD_ERR_HIP, "hipMemcpy: %s", hipGetErrorString(e));
  }
  rc = check_launch("sys_create");
  if (rc != RBMD_OK) {
    rbmd_sys_destroy(s);
    return rc;
  }
  *out = s;
  return RBMD_OK;
}

rbmd_status rbmd_sys_destroy(rbmd_sys* s) {
  if (!s) return RBMD_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  for (auto g : s->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (void* q : s->allocs) (void)hipFree(q);
  for (int h = 0; h < 2; ++h) {
    if (s->stage[h]) (void)hipHostFree(s->stage[h]);
    if (s->stage_ev[h]) (void)hipEventDestroy(s->stage_ev[h]);
  }
  (void)hipEventDestroy(s->ev_fork);
  (void)hipEventDestroy(s->ev_join);
  delete s;
  return RBMD_OK;
}

rbmd_status rbmd_sys_footprint(const rbmd_sys* s, int64_t* bytes) {
  if (!s || !bytes) return fail(RBMD_ERR_ARG, "null argument");
  *bytes = s->bytes;
  return RBMD_OK;
}

rbmd_status rbmd_sys_upload_state(rbmd_sys* s, const double* x, const double* v, const int32_t* image) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  int n = s->n;
  std::vector<double> a(n), b(n), c(n);
  hipStream_t st = s->ctx->stream;
  auto put = [&](const double* src, double* dx, double* dy, double* dz) -> rbmd_status {
    for (int i = 0; i < n; ++i) {
      a[i] = src[3 * i];
      b[i] = src[3 * i + 1];
      c[i] = src[3 * i + 2];
    }
    HIP_TRY(hipMemcpyAsync(dx, a.data(), n * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dy, b.data(), n * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dz, c.data(), n * 8, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    return RBMD_OK;
  };
  if (x) RC_TRY(put(x, s->v.x, s->v.y, s->v.z));
  if (v) RC_TRY(put(v, s->v.vx, s->v.vy, s->v.vz));
  if (image) {
    HIP_TRY(hipMemcpyAsync(s->v.img, image, (size_t)3 * n * 4, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  s->locator_built = s->nbr_built = false;
  s->skin_ready = false;  // a skin list is rebuilt on the next step
  return RBMD_OK;
}

rbmd_status rbmd_sys_download_state(rbmd_sys* s, double* x, double* v, double* f, int32_t* image) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  int n = s->n;
  hipStream_t st = s->ctx->stream;
  std::vector<double> a(n), b(n), c(n);
  auto get = [&](double* dst, const double* dx, const double* dy, const double* dz) -> rbmd_status {
    HIP_TRY(hipMemcpyAsync(a.data(), dx, n * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(b.data(), dy, n * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(c.data(), dz, n * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i) {
      dst[3 * i] = a[i];
      dst[3 * i + 1] = b[i];
      dst[3 * i + 2] = c[i];
    }
    return RBMD_OK;
  };
  if (x) RC_TRY(get(x, s->v.x, s->v.y, s->v.z));
  if (v) RC_TRY(get(v, s->v.vx, s->v.vy, s->v.vz));
  if (f) RC_TRY(get(f, s->v.fx, s->v.fy, s->v.fz));
  if (image) {
    HIP_TRY(hipMemcpyAsync(image, s->v.img, (size_t)3 * n * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  return RBMD_OK;
}

rbmd_status rbmd_locator_geometry(const rbmd_sys* s, int32_t dims[3], int32_t* nc, double dx[3]) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  for (int d = 0; d < 3; ++d) {  // the reference's locator (a skin list's cells are the MD step's own)
    if (dims) dims[d] = s->g_ref.dims[d];
    if (dx) dx[d] = s->g_ref.dxdydz[d];
  }
  if (nc) *nc = s->v.nc;
  return RBMD_OK;
}

rbmd_status rbmd_locator_build(rbmd_sys* s) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  RC_TRY(use_geom(s, false));  // the reference's cells; a skin list is rebuilt on the next step
  build_cells(s, true);
  s->locator_built = true;
  s->nbr_built = false;
  return check_launch("locator_build");
}

rbmd_status rbmd_locator_export(rbmd_sys* s, int32_t* cell_of, int32_t* point_ids, int32_t* lower, int32_t* upper) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  if (!s->locator_built) return fail(RBMD_ERR_STATE, "locator not built");
  hipStream_t st = s->ctx->stream;
  int n = s->n, C = s->v.ncell;
  std::vector<int> start(C + 1);
  if (cell_of) HIP_TRY(hipMemcpyAsync(cell_of, s->v.cell_of, n * 4, hipMemcpyDeviceToHost, st));
  if (point_ids) HIP_TRY(hipMemcpyAsync(point_ids, s->v.point_ids, n * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(start.data(), s->v.cell_start, (C + 1) * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int c = 0; c < C; ++c) {
    if (lower) lower[c] = start[c];
    if (upper) upper[c] = start[c + 1];
  }
  int stt = 0;
  RC_TRY(read_status(s, &stt));
  return status_to_rc(stt);
}

rbmd_status rbmd_neighbors_build(rbmd_sys* s, int32_t cap) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  if (!s->locator_built || s->skin_active) RC_TRY(rbmd_locator_build(s));
  RC_TRY(ensure_nbr(s, cap));
  s->ctx->run("nbr_build", [&] { launch_nbr_build(s->v, s->ctx->stream); });
  s->nbr_built = true;
  return check_launch("neighbors_build");
}

rbmd_status rbmd_neighbors_export(rbmd_sys* s, int32_t cap, int32_t* ids, double* offsets, int32_t* num) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  if (!s->nbr_built) return fail(RBMD_ERR_STATE, "neighbour lists not built");
  if (cap != s->export_cap) return fail(RBMD_ERR_ARG, "cap %d differs from the built width %d", cap, s->export_cap);
  int n = s->n;
  hipStream_t st = s->ctx->stream;
  std::vector<int> nbr((size_t)s->v.cap * n), cnt(9 * (size_t)n), pid(n), rank(n);
  HIP_TRY(hipMemcpyAsync(nbr.data(), s->v.nbr, nbr.size() * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(cnt.data(), s->v.nbr_num, cnt.size() * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(pid.data(), s->v.point_ids, n * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(rank.data(), s->v.rank, n * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  int stt = 0;
  RC_TRY(read_status(s, &stt));
  bool over = false;
  for (int a = 0; a < n; ++a) {
    int i = rank[a];
    if (ids)
      for (int k = 0; k < cap; ++k) ids[(size_t)a * cap + k] = -1;
    if (offsets)
      for (int k = 0; k < 3 * cap; ++k) offsets[(size_t)a * cap * 3 + k] = 0.0;
    auto emit = [&](int out_slot, int slot) {
      if (out_slot >= cap) {
        over = true;
        return;
      }
      int e = nbr[nbr_index(slot, i, s->v.cap)];
      int js = e >> kShiftBits;
      int code = e & ((1 << kShiftBits) - 1);
      int sh[3] = {(code & 7) - 3, ((code >> 3) & 7) - 3, ((code >> 6) & 7) - 3};
      if (ids) ids[(size_t)a * cap + out_slot] = pid[js];
      if (offsets)
        for (int d = 0; d < 3; ++d)
          offsets[((size_t)a * cap + out_slot) * 3 + d] = s->v.dxdydz[d] * (double)(sh[d] * s->v.dims[d]);
    };
    if (!s->v.rbl) {
      int out = 0;
      for (int g = 0; g < s->v.nseg; ++g)
        for (int k = 0; k < cnt[(size_t)g * n + i]; ++k) emit(out++, g * s->v.seg_cap + k);
      if (num) num[a] = out;
    } else {
      for (int k = 0; k < cnt[2 * i]; ++k) emit(k, k);
      for (int k = 0; k < cnt[2 * i + 1]; ++k) emit(s->v.rs_num + k, s->v.rs_num + k);
      if (num) {
        num[2 * a] = cnt[2 * i];
        num[2 * a + 1] = cnt[2 * i + 1];
      }
    }
  }
  if (over) return fail(RBMD_ERR_CAPACITY, "a list is longer than the export width %d", cap);
  return status_to_rc(stt);
}

rbmd_status rbmd_neighbors_capacity(const rbmd_sys* s, int32_t* cap) {
  if (!s || !cap) return fail(RBMD_ERR_ARG, "bad arguments");
  *cap = s->export_cap;
  return RBMD_OK;
}

rbmd_status rbmd_neighbors_count(rbmd_sys* s, int64_t* entries) {
  if (!s || !entries) return fail(RBMD_ERR_ARG, "bad arguments");
  if (!s->nbr_built) return fail(RBMD_ERR_STATE, "neighbour lists not built");
  int m = s->v.rbl ? 2 * s->n : s->v.nseg * s->n;
  std::vector<int> cnt(m);
  HIP_TRY(hipMemcpyAsync(cnt.data(), s->v.nbr_num, (size_t)m * 4, hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  int64_t t = 0;
  for (int c : cnt) t += c;
  *entries = t;
  return RBMD_OK;
}

rbmd_status rbmd_force_real(rbmd_sys* s, double* force, double* vir, double* elj, double* ecoul) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  RC_TRY(use_geom(s, false));  // the reference's per-step list
  if (s->fused) {
    if (!s->locator_built) RC_TRY(rbmd_locator_build(s));
  } else if (!s->nbr_built) {
    RC_TRY(rbmd_neighbors_build(s, s->v.cap));
  }
  hipStream_t st = s->ctx->stream;
  launch_real(s, 1, 1);
  RC_TRY(check_launch("force_real"));
  const int n = s->n, ns = s->v.nslab;
  std::vector<double> fx((size_t)ns * n), fy((size_t)ns * n), fz((size_t)ns * n), vv(6 * (size_t)ns * n),
      el((size_t)ns * n), ec((size_t)ns * n);
  std::vector<int> rank(n);
  HIP_TRY(hipMemcpyAsync(fx.data(), s->v.fnx, fx.size() * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(fy.data(), s->v.fny, fy.size() * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(fz.data(), s->v.fnz, fz.size() * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(vv.data(), s->v.vir_near, vv.size() * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(el.data(), s->v.e_lj, el.size() * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(ec.data(), s->v.e_coul, ec.size() * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(rank.data(), s->v.rank, n * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  // slab partials added in slab order, as the integrator does
  auto slab_sum = [&](const std::vector<double>& a, size_t i, int stride, int k) {
    double t = a[i * stride + k];
    for (int sl = 1; sl < ns; ++sl) t += a[((size_t)sl * n + i) * stride + k];
    return t;
  };
  for (int a = 0; a < n; ++a) {
    const size_t i = rank[a];
    if (force) {
      force[3 * a] = slab_sum(fx, i, 1, 0);
      force[3 * a + 1] = slab_sum(fy, i, 1, 0);
      force[3 * a + 2] = slab_sum(fz, i, 1, 0);
    }
    if (vir)
      for (int k = 0; k < 6; ++k) vir[6 * a + k] = slab_sum(vv, i, 6, k);
    if (elj) elj[a] = slab_sum(el, i, 1, 0);
    if (ecoul) ecoul[a] = slab_sum(ec, i, 1, 0);
  }
  int stt = 0;
  RC_TRY(read_status(s, &stt));
  return status_to_rc(stt);
}

rbmd_status rbmd_rbe_S(double alpha, const double box[3], double* S) {
  if (!box || !S || !(alpha > 0)) return fail(RBMD_ERR_ARG, "bad arguments");
  *S = rbmd::rbe_S(alpha, box);
  return RBMD_OK;
}

rbmd_status rbmd_rbe_sample_host(double alpha, const double box[3], int32_t p, uint64_t seed, uint64_t stream,
                                 double* m_out) {
  if (!box || !m_out || p <= 0 || !(alpha > 0)) return fail(RBMD_ERR_ARG, "bad arguments");
  KSampler ks(alpha, box, p);
  ks.draw(seed, stream, m_out);
  return RBMD_OK;
}

rbmd_status rbmd_rbe_set_kvectors(rbmd_sys* s, int32_t p, const double* m) {
  if (!s || !m || p <= 0) return fail(RBMD_ERR_ARG, "bad arguments");
  if (s->p.coulomb_type != RBMD_COULOMB_RBE) return fail(RBMD_ERR_STATE, "coulomb type is not RBE");
  if (p != s->P) {
    s->P = p;
    RC_TRY(s->dalloc(&s->kcur, 3 * (size_t)p));
    RC_TRY(ensure_rho(s, p));
    s->kgraph = nullptr;
  }
  s->graph_ok = false;
  std::vector<double> k(3 * (size_t)p);
  for (int l = 0; l < p; ++l)
    for (int d = 0; d < 3; ++d) k[3 * l + d] = 2 * kPi * m[3 * l + d] / s->v.box[d];
  HIP_TRY(hipMemcpy(s->kcur, k.data(), k.size() * 8, hipMemcpyHostToDevice));
  s->injected = true;
  return RBMD_OK;
}

rbmd_status rbmd_rbe_sample(rbmd_sys* s, uint64_t seed, uint64_t stream, double* m_out) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  if (s->p.coulomb_type != RBMD_COULOMB_RBE) return fail(RBMD_ERR_STATE, "coulomb type is not RBE");
  std::vector<double> m(3 * (size_t)s->sampler.P());
  s->sampler.draw(seed, stream, m.data());
  if (m_out) std::copy(m.begin(), m.end(), m_out);
  return rbmd_rbe_set_kvectors(s, s->sampler.P(), m.data());
}

rbmd_status rbmd_rbe_structure_factor(rbmd_sys* s, double* rho_out) {
  if (!s || !rho_out) return fail(RBMD_ERR_ARG, "bad arguments");
  if (!has_k(s) || !s->kcur) return fail(RBMD_ERR_STATE, "no k vectors");
  KView k = s->kv;
  k.P = s->P;
  k.kvec = s->kcur;
  k.f32 = s->v.mixed;
  hipStream_t st = s->ctx->stream;
  s->ctx->run("rho_partial", [&] { launch_rho_partial(s->v, k, st); });
  s->ctx->run("rho_reduce", [&] { launch_rho_reduce(s->v, k, st); });
  RC_TRY(check_launch("structure_factor"));
  HIP_TRY(hipMemcpyAsync(rho_out, s->rho, 2 * (size_t)s->P * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return RBMD_OK;
}

static rbmd_status kforce_out(rbmd_sys* s, int ewald, double* force, double* vir) {
  KView k = s->kv;
  k.P = s->P;
  k.kvec = s->kcur;
  k.S_over_P = ewald ? 1.0 : s->S / (double)s->P;
  k.f32 = ewald ? 0 : s->v.mixed;  // the converged Ewald sum is always FP64
  k.ewald = ewald;
  hipStream_t st = s->ctx->stream;
  s->ctx->run("rho_partial", [&] { launch_rho_partial(s->v, k, st); });
  s->ctx->run("rho_reduce", [&] { launch_rho_reduce(s->v, k, st); });
  s->ctx->run("kforce", [&] {
    launch_kforce(s->v, k, st, 1, 1, s->seg_rows(SEG_K), &s->seg_used[SEG_K]);
  });
  RC_TRY(check_launch("kforce"));
  int n = s->n;
  std::vector<double> fx(n), fy(n), fz(n);
  HIP_TRY(hipMemcpyAsync(fx.data(), s->v.fkx, n * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(fy.data(), s->v.fky, n * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(fz.data(), s->v.fkz, n * 8, hipMemcpyDeviceToHost, st));
  if (vir) HIP_TRY(hipMemcpyAsync(vir, s->v.vir_k, 6 * (size_t)n * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (force)
    for (int i = 0; i < n; ++i) {
      force[3 * i] = fx[i];
      force[3 * i + 1] = fy[i];
      force[3 * i + 2] = fz[i];
    }
  return RBMD_OK;
}

rbmd_status rbmd_rbe_force(rbmd_sys* s, double* force, double* vir) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  if (s->p.coulomb_type != RBMD_COULOMB_RBE || !s->kcur) return fail(RBMD_ERR_STATE, "no RBE k vectors");
  return kforce_out(s, 0, force, vir);
}

rbmd_status rbmd_ewald_force(rbmd_sys* s, const int32_t kmax[3], double* force, double* vir, double* energy_total) {
  if (!s || !kmax) return fail(RBMD_ERR_ARG, "bad arguments");
  if (s->v.lj_only) return fail(RBMD_ERR_STATE, "LJ/CUT has no k space");
  std::vector<double> kg;
  for (int i = -kmax[0]; i <= kmax[0]; ++i)
    for (int j = -kmax[1]; j <= kmax[1]; ++j)
      for (int k = -kmax[2]; k <= kmax[2]; ++k)
        if (i || j || k) {
          kg.push_back(2 * kPi * i / s->v.box[0]);
          kg.push_back(2 * kPi * j / s->v.box[1]);
          kg.push_back(2 * kPi * k / s->v.box[2]);
        }
  int K = (int)(kg.size() / 3);
  double* saved_k = s->kcur;
  int saved_P = s->P;
  double* kdev = nullptr;
  HIP_TRY(hipMalloc(&kdev, kg.size() * 8));
  HIP_TRY(hipMemcpy(kdev, kg.data(), kg.size() * 8, hipMemcpyHostToDevice));
  s->P = K;
  s->kcur = kdev;
  rbmd_status rc = ensure_rho(s, K);
  if (rc == RBMD_OK) rc = kforce_out(s, 1, force, vir);
  s->kcur = saved_k;
  s->P = saved_P;
  if (rc == RBMD_OK) rc = ensure_rho(s, saved_P);  // the step's block sizes again
  (void)hipFree(kdev);
  if (rc != RBMD_OK) return rc;
  if (energy_total) {
    double e_avg = 0, grid = 0;
    RC_TRY(long_energy(s, kmax, &e_avg, &grid));
    *energy_total = s->u.qqr2e * grid * (2 * kPi / s->v.volume);
  }
  return RBMD_OK;
}

rbmd_status rbmd_static_table_gnear(rbmd_ctx* c, int32_t n, const double* dis, double* g) {
  if (!c || n < 0 || (n > 0 && (!dis || !g))) return fail(RBMD_ERR_ARG, "bad arguments");
  if (n == 0) return RBMD_OK;
  HIP_TRY(hipSetDevice(c->device));
  double *dd = nullptr, *dg = nullptr;
  HIP_TRY(hipMalloc(&dd, (size_t)n * 8));
  hipError_t e = hipMalloc(&dg, (size_t)n * 8);
  if (e == hipSuccess) e = hipMemcpyAsync(dd, dis, (size_t)n * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    launch_static_table(dd, dg, n, c->stream);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(g, dg, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(dd);
  if (dg) (void)hipFree(dg);
  if (e != hipSuccess) return fail(RBMD_ERR_HIP, "static table: %s", hipGetErrorString(e));
  return RBMD_OK;
}

rbmd_status rbmd_force_bonded(rbmd_sys* s, double* force, double* eb, double* ea) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  hipStream_t st = s->ctx->stream;
  int nblk = 0;
  s->ctx->run("bonded", [&] { launch_bonded(s->v, s->iv, st, 1, s->seg_rows(SEG_BOND), &nblk); });
  s->seg_used[SEG_BOND] = nblk;
  RC_TRY(check_launch("bonded"));
  int n = s->n;
  std::vector<double> fx(n), fy(n), fz(n), rows((size_t)nblk * kRowBond);
  HIP_TRY(hipMemcpyAsync(fx.data(), s->v.fbx, n * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(fy.data(), s->v.fby, n * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(fz.data(), s->v.fbz, n * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(rows.data(), s->seg_rows(SEG_BOND), rows.size() * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (force)
    for (int i = 0; i < n; ++i) {
      force[3 * i] = fx[i];
      force[3 * i + 1] = fy[i];
      force[3 * i + 2] = fz[i];
    }
  double b = 0, a = 0;
  for (int r = 0; r < nblk; ++r) {
    b += rows[(size_t)r * kRowBond + (S_EBOND - S_VIR_BOND)];
    a += rows[(size_t)r * kRowBond + (S_EANGLE - S_VIR_BOND)];
  }
  if (eb) *eb = b;
  if (ea) *ea = a;
  return RBMD_OK;
}

// ExecutionNVT::Init: locator, k-set of step 0, dof/NHC SetUp, first force
rbmd_status rbmd_md_init(rbmd_sys* s) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  rbmd_ctx* c = s->ctx;
  hipStream_t st = c->stream;
  launch_fill_int(s->v.status, 0, 1, st);
  s->injected = false;
  s->graph_ok = false;
  NhcState zero;
  std::memset(&zero, 0, sizeof(zero));
  HIP_TRY(hipMemcpyAsync(s->nhc, &zero, sizeof(zero), hipMemcpyHostToDevice, st));
  if (s->skin) {  // build the skin list now (step counter 0)
    RC_TRY(use_geom(s, true));
    launch_skin_request(s->v, 0, st);
  }
  build_cells(s, true);
  gather_and_list(s);
  s->skin_ready = s->skin;
  const double* kv = nullptr;
  RC_TRY(kvec_for_step(s, 0, &kv));
  compute_force(s, kv, 1);
  c->run("second_half", [&] {
    launch_second_half(s->v, s->iv, st, 0, has_k(s), has_bonded(s), s->seg_rows(SEG_KE), &s->seg_used[SEG_KE]);
  });
  c->run("nhc", [&] { launch_nhc_rows(s->nhc, s->np, rows_desc(s, true), 3, st); });
  RC_TRY(check_launch("md_init"));
  s->step = 0;
  s->initialized = true;
  if (s->p.energy_every > 0 && has_k(s)) {
    RC_TRY(long_energy(s, s->p.kmax, &s->e_long, nullptr));
    s->energy_step = 0;
  }
  int stt = 0;
  RC_TRY(read_status(s, &stt));
  return status_to_rc(stt);
}

// k vectors come from the host-filled ring (not from injected or Ewald sets)
static bool kring_on(const rbmd_sys* s) {
  return has_k(s) && s->p.coulomb_type == RBMD_COULOMB_RBE && !s->injected;
}

// Enqueue one MD step (ExecutionNVT::PreSolve + Solve, ExecutionNVT.cpp:76-135)
// with the k vectors at kv.  from_ring: inside a step graph; the step's k
// vectors are copied from the ring slot the device step counter names.
// thermo: the step also forms the virial and energy sums behind pressure and
// the reported energies; without them it computes the same forces and the
// kinetic energy the thermostat needs, and the thermo state it leaves holds
// no valid pressure or energies (the last step of every call has thermo on).
// true when the thermostat's step-start phase can run in the previous step's
// end-of-step launch (phase 6): Nose-Hoover and plain velocity Verlet
static bool nhc_merges(const rbmd_sys* s) {
  return s->p.temp_ctrl == RBMD_TEMP_NOSE_HOOVER || s->p.temp_ctrl == RBMD_TEMP_NONE;
}

// the thermostat's step-start launch alone (the first step of a call)
static void enqueue_step_start(rbmd_sys* s) {
  rbmd_ctx* c = s->ctx;
  c->run("nhc", [&] { launch_nhc_rows(s->nhc, s->np, rows_desc(s, false), nhc_merges(s) ? 1 : 5, c->stream); });
}

// One MD step.  start: the step begins with its own thermostat launch (else
// the previous step's end ran it); merge_end: the end-of-step thermostat
// launch also starts the next step (every step of a call but the last);
// defer_rattle: when every atom is constrained, the step's RATTLE is left to
// the next step's first launch (k_rattle_first_half_shake), which then must
// follow in the same call (rattle_pending; flush_rattle launches it alone).
static void enqueue_step(rbmd_sys* s, const double* kv, bool from_ring, bool thermo, bool start, bool merge_end,
                         bool defer_rattle = false) {
  rbmd_ctx* c = s->ctx;
  hipStream_t st = c->stream;
  const int tc = s->p.temp_ctrl;
  const bool nh = tc == RBMD_TEMP_NOSE_HOOVER || tc == RBMD_TEMP_NONE;
  const bool shake = shake_on(s);
  if (start || !nh) enqueue_step_start(s);
  if (from_ring)  // the nhc phase above advanced the device step counter
    c->run("kvec_slot", [&] { launch_kvec_slot(s->kgraph, s->kring, s->ring_R, 3 * s->P, &s->nhc->step, st); });
  const double* scale = (tc == RBMD_TEMP_NOSE_HOOVER) ? &s->nhc->factor : nullptr;
  // also bins (per-step lists: the reference's locator bins before SHAKE)
  const int do_bin = s->skin_active ? 0 : 1;
  if (shake && s->nfree == 0) {  // every atom constrained: one launch for both
    if (s->rattle_pending) {  // and the previous step's RATTLE in front of them
      c->run("rattle_first_half_shake", [&] {
        launch_rattle_first_half_shake(s->v, s->iv, st, scale, do_bin, s->seg_rows(SEG_SHAKE), s->shake_iter);
      });
      s->rattle_pending = false;
    } else {
      c->run("first_half_shake", [&] {
        launch_first_half_shake(s->v, s->iv, st, scale, do_bin, s->seg_rows(SEG_SHAKE), s->shake_iter);
      });
    }
    s->seg_used[SEG_SHAKE] = (s->nmol + kBlock - 1) / kBlock;
  } else {
    c->run("first_half", [&] { launch_first_half(s->v, s->iv, st, scale, do_bin); });
    if (shake) {
      c->run("shake", [&] { launch_shake(s->v, s->iv, st, s->seg_rows(SEG_SHAKE), s->shake_iter); });
      s->seg_used[SEG_SHAKE] = (s->nmol + kBlock - 1) / kBlock;
      if (s->nfree) launch_unconstrained_copy(s->v, st, s->free_dev, s->nfree, 1);
    }
  }
  // skin list: the displacement test on the final positions, which also bins
  // them for a rebuild; the cell and list kernels below run in full only on
  // the steps it (or a host request) marks
  if (s->skin_active) c->run("skin_check", [&] { launch_skin_check(s->v, st); });
  build_cells(s, false);  // the cell sort also writes the sorted copies: after SHAKE
  // the positions are final for this step: the k-space and bonded forces
  // branch off to the side stream while the list is built and the pair
  // forces computed (kernel profiling keeps one stream, for its event timings).
  // Forking before the cell sort instead put the k-space kernels beside the
  // short cell kernels of the critical path and slowed the step.
  const bool fork = !c->prof && (has_k(s) || has_bonded(s));
  if (fork) {
    (void)hipEventRecord(s->ev_fork, st);
    (void)hipStreamWaitEvent(c->side, s->ev_fork, 0);
    launch_k_bonded(s, kv, thermo, c->side);
    (void)hipEventRecord(s->ev_join, c->side);
  }
  gather_and_list(s);
  if (fork) {
    launch_real(s, thermo, 0);
    (void)hipStreamWaitEvent(st, s->ev_join, 0);
  } else {
    compute_force(s, kv, thermo);
  }
  c->run("second_half", [&] {
    launch_second_half(s->v, s->iv, st, 1, has_k(s), has_bonded(s), s->seg_rows(SEG_KE), &s->seg_used[SEG_KE]);
  });
  if (nh) {
    c->run("nhc", [&] {
      launch_nhc_rows(s->nhc, s->np, rows_desc(s, true), merge_end ? 6 : 2, st, thermo ? 0 : 1);
    });
    if (tc == RBMD_TEMP_NOSE_HOOVER && !(shake && s->nfree == 0))
      c->run("scale_v", [&] { launch_scale_v(s->v, st, &s->nhc->factor_end); });
    if (shake && s->nfree == 0 && defer_rattle && merge_end) {
      s->rattle_pending = true;
    } else if (shake) {
      c->run("rattle", [&] { launch_rattle(s->v, s->iv, st); });
      if (s->nfree) launch_unconstrained_copy(s->v, st, s->free_dev, s->nfree, 0);
    }
  } else {
    if (shake) {
      c->run("rattle", [&] { launch_rattle(s->v, s->iv, st); });
      if (s->nfree) launch_unconstrained_copy(s->v, st, s->free_dev, s->nfree, 0);
      c->run("ke_rows", [&] { launch_ke_rows(s->v, s->iv, st, s->seg_rows(SEG_KE)); });
    }
    c->run("nhc", [&] { launch_nhc_rows(s->nhc, s->np, rows_desc(s, true), 4, st, thermo ? 0 : 1); });
    c->run("scale_v", [&] { launch_scale_v(s->v, st, &s->nhc->factor_end); });
  }
}

// the RATTLE a step left to the next one, alone (a call that stops early)
static void flush_rattle(rbmd_sys* s) {
  if (!s->rattle_pending) return;
  s->ctx->run("rattle", [&] { launch_rattle(s->v, s->iv, s->ctx->stream); });
  s->rattle_pending = false;
}

static rbmd_status md_one_step(rbmd_sys* s, bool first, bool last) {
  const int64_t step = s->step + 1;
  const double* kv = nullptr;
  RC_TRY(kvec_for_step(s, step, &kv));
  enqueue_step(s, kv, false, last, first, nhc_merges(s) && !last, !last);
  s->step = step;
  return RBMD_OK;
}

// Capture one step into a graph, without and with the thermo sums.  The
// captured launches carry this system's buffers and sizes, so the graphs are
// rebuilt after anything changes them.
static rbmd_status capture_step(rbmd_sys* s) {
  rbmd_ctx* c = s->ctx;
  hipStream_t st = c->stream;
  for (auto& g : s->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  const bool ring = kring_on(s);
  if (ring && !s->kgraph) RC_TRY(s->dalloc(&s->kgraph, 3 * (size_t)s->P));
  for (int thermo = 0; thermo < 2; ++thermo) {
    const bool prof = c->prof;
    c->prof = false;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    // graph 0: a step between others of its call; graph 1: a call's last
    // step.  Neither starts with the thermostat launch: the previous step's
    // end (or a launch before the first replay of the call) ran it.
    enqueue_step(s, ring ? s->kgraph : s->kcur, ring, thermo != 0, !nhc_merges(s), nhc_merges(s) && !thermo);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    c->prof = prof;
    if (e != hipSuccess) return fail(RBMD_ERR_HIP, "step capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&s->graph[thermo], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(RBMD_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  s->graph_ok = true;
  return RBMD_OK;
}

rbmd_status rbmd_md_step(rbmd_sys* s, int64_t n_steps) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  if (!s->initialized) return fail(RBMD_ERR_STATE, "rbmd_md_init has not run");
  RC_TRY(prepare_skin(s));
  for (int64_t k = 0; k < n_steps; ++k) {
    rbmd_status rc = md_one_step(s, k == 0, k == n_steps - 1);
    if (rc == RBMD_OK && s->p.energy_every > 0 && has_k(s) && s->step % s->p.energy_every == 0) {
      rc = long_energy(s, s->p.kmax, &s->e_long, nullptr);  // positions only: a pending RATTLE may wait
      if (rc == RBMD_OK) s->energy_step = s->step;
    }
    if (rc != RBMD_OK) {
      flush_rattle(s);
      return rc;
    }
  }
  flush_rattle(s);  // a no-op: the last step of a call runs its own RATTLE
  RC_TRY(check_launch("md_step"));
  int stt = 0;
  RC_TRY(read_status(s, &stt));
  return status_to_rc(stt);
}

// Same steps as rbmd_md_step, each replayed from one captured step graph: the
// host only tops up the k-vector ring (every R steps) and launches the graph.
static rbmd_status enqueue_graph_steps(rbmd_sys* s, int64_t n_steps) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  if (!s->initialized) return fail(RBMD_ERR_STATE, "rbmd_md_init has not run");
  RC_TRY(prepare_skin(s));
  if (!s->graph_ok) RC_TRY(capture_step(s));
  hipStream_t st = s->ctx->stream;
  const bool ring = kring_on(s);
  if (n_steps > 0 && nhc_merges(s)) enqueue_step_start(s);  // the graphs start after it
  for (int64_t k = 0; k < n_steps; ++k) {
    const int64_t step = s->step + 1;
    if (ring) {
      const double* unused = nullptr;
      RC_TRY(kvec_for_step(s, step, &unused));  // fills the ring half this step reads
    }
    HIP_TRY(hipGraphLaunch(s->graph[k == n_steps - 1 ? 1 : 0], st));
    s->step = step;
    if (s->p.energy_every > 0 && has_k(s) && s->step % s->p.energy_every == 0) {
      RC_TRY(long_energy(s, s->p.kmax, &s->e_long, nullptr));
      s->energy_step = s->step;
    }
  }
  return check_launch("md_run_graph");
}

rbmd_status rbmd_md_run_graph(rbmd_sys* s, int64_t n_steps) {
  RC_TRY(enqueue_graph_steps(s, n_steps));
  int stt = 0;
  RC_TRY(read_status(s, &stt));
  return status_to_rc(stt);
}

rbmd_status rbmd_md_enqueue(rbmd_sys* s, int64_t n_steps) { return enqueue_graph_steps(s, n_steps); }

rbmd_status rbmd_md_wait(rbmd_sys* s) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  int stt = 0;
  RC_TRY(read_status(s, &stt));
  return status_to_rc(stt);
}

rbmd_status rbmd_md_query(rbmd_sys* s, int32_t* idle) {
  if (!s || !idle) return fail(RBMD_ERR_ARG, "null argument");
  hipError_t e = hipStreamQuery(s->ctx->stream);
  if (e == hipSuccess) {
    *idle = 1;
  } else if (e == hipErrorNotReady) {
    *idle = 0;
  } else {
    return fail(RBMD_ERR_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
  }
  return RBMD_OK;
}

rbmd_status rbmd_md_thermo(rbmd_sys* s, rbmd_thermo* o) {
  if (!s || !o) return fail(RBMD_ERR_ARG, "bad arguments");
  NhcState S;
  int stt = 0, it = 0, mf = 0, lb = 0;
  HIP_TRY(hipMemcpyAsync(&mf, s->v.misfit_count, 4, hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipMemcpyAsync(&S, s->nhc, sizeof(S), hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipMemcpyAsync(&stt, s->v.status, 4, hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipMemcpyAsync(&it, s->shake_iter, 4, hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipMemcpyAsync(&lb, s->v.list_builds, 4, hipMemcpyDeviceToHost, s->ctx->stream));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  fill_thermo(s, S, stt, it, o);
  o->misfit_atoms = mf;
  o->list_builds = lb;
  return RBMD_OK;
}

rbmd_status rbmd_md_energy(rbmd_sys* s) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  if (!has_k(s)) return fail(RBMD_ERR_STATE, "no coulomb k space");
  RC_TRY(long_energy(s, s->p.kmax, &s->e_long, nullptr));
  s->energy_step = s->step;
  return RBMD_OK;
}

rbmd_status rbmd_md_energy_noalloc(rbmd_sys* s) {
  if (!s) return fail(RBMD_ERR_ARG, "null sys");
  if (!has_k(s)) return fail(RBMD_ERR_STATE, "no coulomb k space");
  void* scratch = scratch_address(s->ctx->device);
  if (!scratch) return fail(RBMD_ERR_STATE, "no scratch address on device %d", s->ctx->device);
  RC_TRY(long_energy(s, s->p.kmax, &s->e_long, nullptr, scratch, kScratchBytes));
  s->energy_step = s->step;
  return RBMD_OK;
}

rbmd_status rbmd_profile_enable(rbmd_ctx* c, int32_t on) {
  if (!c) return fail(RBMD_ERR_ARG, "null ctx");
  c->flush();
  c->prof = on != 0;
  return RBMD_OK;
}

rbmd_status rbmd_profile_reset(rbmd_ctx* c) {
  if (!c) return fail(RBMD_ERR_ARG, "null ctx");
  c->flush();
  for (auto& r : c->recs) {
    r.launches = 0;
    r.ms = 0;
  }
  return RBMD_OK;
}

rbmd_status rbmd_profile_get(rbmd_ctx* c, int32_t i, const char** name, int64_t* launches, double* ms) {
  if (!c) return fail(RBMD_ERR_ARG, "null ctx");
  c->flush();
  if (i < 0 || i >= (int)c->recs.size()) return fail(RBMD_ERR_ARG, "profile index %d past the end", i);
  if (name) *name = c->recs[i].name.c_str();
  if (launches) *launches = c->recs[i].launches;
  if (ms) *ms = c->recs[i].ms;
  return RBMD_OK;
}

// =================================================================== packs
// Many systems of the same shape stepped together: every kernel of the MD
// step is one launch over the blocks of all systems (PackLaunch), so small
// systems no longer pay a full chain of launches each.  Each system keeps its
// own buffers, thermostat state, k-vector ring and status word; its blocks
// run exactly the grid of its own launches, so its results are the same bits
// as rbmd_md_step / rbmd_md_run_graph on that system alone.

struct rbmd_pack {
  rbmd_ctx* ctx = nullptr;
  std::vector<rbmd_sys*> sys;
  PackSig sig{};
  int tc = 0;
  bool ring = false, nh = false, shake = false, scale_v = false, has_k = false, has_b = false;
  PackSys* dsys = nullptr;
  int* dmap = nullptr;
  int* dfirst = nullptr;
  int* dstatus = nullptr;
  int* hstatus = nullptr;  // pinned
  PackLaunch L[PK_COUNT];
  int nblk[PK_COUNT] = {};
  size_t lds_first = 0, lds_scatter = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // captured packed step without / with thermo sums
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;  // side-stream branch of a packed step
  std::vector<void*> allocs;
};

static const char* const kPackName[PK_COUNT] = {
    "pk_nhc",   "pk_kvec_slot", "pk_first_half", "pk_scan", "pk_scatter",     "pk_cell_sort",
    "pk_one",   "pk_shake",     "pk_nbr_build", "pk_pair",   "pk_rho_partial",
    "pk_rho_reduce", "pk_kforce", "pk_bonded",   "pk_second_half", "pk_scale_v", "pk_rattle"};

// the thermo-row descriptor a system's own step reduces (rows per producer =
// the grid of that producer's launch)
static RowsDesc pack_rows_desc(rbmd_sys* s, bool has_kk, bool has_bb, bool shk) {
  int used[NSEG];
  used[SEG_PAIR] = pair_blocks(s->v);
  used[SEG_K] = has_kk ? kforce_blocks(s->v) : 0;
  used[SEG_BOND] = has_bb ? (s->n + kBlock - 1) / kBlock : 0;
  used[SEG_KE] = (s->n + kBlock - 1) / kBlock;
  used[SEG_SHAKE] = shk ? (s->nmol + kBlock - 1) / kBlock : 0;
  RowsDesc d;
  d.nseg = NSEG;
  for (int g = 0; g < NSEG; ++g) d.seg[g] = RowSeg{s->seg_rows(g), used[g], kSegWidth[g], kSegSlot0[g], kSegGroups[g]};
  return d;
}

// thermo: as for enqueue_step (the last step of every call has it on)
static void pack_enqueue_step(rbmd_pack* p, bool thermo) {
  rbmd_ctx* c = p->ctx;
  hipStream_t st = c->stream;
  PackSig sig = p->sig;
  sig.thermo = thermo ? 1 : 0;
  auto integ = [&](int k, int arg, size_t lds, hipStream_t s) {
    c->run(kPackName[k], [&] { launch_pack_integrate(k, p->L[k], p->nblk[k], sig, arg, lds, nullptr, s); });
  };
  auto cells = [&](int k, size_t lds) {
    c->run(kPackName[k], [&] { launch_pack_cells(k, p->L[k], p->nblk[k], lds, st); });
  };
  auto ksp = [&](int k, hipStream_t s) {
    c->run(kPackName[k], [&] { launch_pack_kspace(k, p->L[k], p->nblk[k], sig, s); });
  };
  integ(PK_NHC, p->nh ? 1 : 5, 0, st);
  if (p->ring) ksp(PK_KSLOT, st);
  integ(PK_FIRST_HALF, p->tc == RBMD_TEMP_NOSE_HOOVER ? 1 : 0, p->lds_first, st);
  if (p->shake) integ(PK_SHAKE, 0, 0, st);
  cells(PK_SCAN, 0);
  cells(PK_SCATTER, p->lds_scatter);
  cells(PK_CELL_SORT, 0);  // also the sorted copies: after SHAKE
  // as in a single system's step: the k-space and bonded forces branch off to
  // the side stream once the positions are final (one stream when profiling)
  const bool fork = !c->prof && (p->has_k || p->has_b);
  const hipStream_t sk = fork ? c->side : st;
  if (fork) {
    (void)hipEventRecord(p->ev_fork, st);
    (void)hipStreamWaitEvent(sk, p->ev_fork, 0);
  }
  if (p->has_k) {
    ksp(PK_RHO, sk);
    ksp(PK_RHO_RED, sk);
    ksp(PK_KFORCE, sk);
  }
  if (p->has_b) integ(PK_BONDED, 0, 0, sk);
  if (fork) (void)hipEventRecord(p->ev_join, sk);
  cells(PK_NBR, 0);
  c->run(kPackName[PK_PAIR], [&] { launch_pack_pair(p->L[PK_PAIR], p->nblk[PK_PAIR], sig, st); });
  if (fork) (void)hipStreamWaitEvent(st, p->ev_join, 0);
  integ(PK_SECOND, 1, 0, st);
  integ(PK_NHC, 2 | 8, 0, st);
  if (p->scale_v) integ(PK_ATOMS, 0, 0, st);
  if (p->shake) integ(PK_RATTLE, 0, 0, st);
}

static rbmd_status pack_fail_free(rbmd_pack* p, rbmd_status rc) {
  for (void* q : p->allocs) (void)hipFree(q);
  if (p->hstatus) (void)hipHostFree(p->hstatus);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  delete p;
  return rc;
}

rbmd_status rbmd_pack_create(rbmd_ctx* ctx, rbmd_sys* const* systems, int32_t n, rbmd_pack** out) {
  if (!ctx || !systems || n <= 0 || !out) return fail(RBMD_ERR_ARG, "bad arguments");
  rbmd_sys* s0 = systems[0];
  if (!s0) return fail(RBMD_ERR_ARG, "null system 0");
  auto mode_of = [](const rbmd_sys* s) { return s->v.lj_only ? 2 : (s->v.weighted ? 1 : 0); };
  auto sp_of = [](const rbmd_sys* s) {
    return s->v.weighted && !s->v.lj_only ? (s->v.sp_max <= 2 ? 2 : (s->v.sp_max <= 4 ? 4 : 16)) : 0;
  };
  for (int k = 0; k < n; ++k) {
    const rbmd_sys* s = systems[k];
    if (!s) return fail(RBMD_ERR_ARG, "null system %d", k);
    if (s->ctx != ctx) return fail(RBMD_ERR_ARG, "system %d belongs to another context", k);
    if (!s->initialized) return fail(RBMD_ERR_STATE, "system %d: rbmd_md_init has not run", k);
    if (s->v.rbl || s->fused) return fail(RBMD_ERR_UNSUPPORTED, "system %d: packs run the VERLET list path", k);
    if (s->skin) return fail(RBMD_ERR_UNSUPPORTED, "system %d: packs build their lists every step (skin 0)", k);
    if (s->ntypes > kMaxTypes) return fail(RBMD_ERR_UNSUPPORTED, "system %d: more than %d types", k, kMaxTypes);
    const int tc = s->p.temp_ctrl;
    if (tc != RBMD_TEMP_NOSE_HOOVER && tc != RBMD_TEMP_NONE)
      return fail(RBMD_ERR_UNSUPPORTED, "system %d: packs run Nose-Hoover or NVE steps", k);
    if (s->p.energy_every > 0) return fail(RBMD_ERR_UNSUPPORTED, "system %d: energy_every must be 0 in a pack", k);
    if (shake_on(s) && s->nfree > 0) return fail(RBMD_ERR_UNSUPPORTED, "system %d: atoms outside SHAKE triplets", k);
    if (s->injected) return fail(RBMD_ERR_UNSUPPORTED, "system %d: injected k vectors", k);
    if (mode_of(s) != mode_of(s0) || sp_of(s) != sp_of(s0) || s->v.mixed != s0->v.mixed || has_k(s) != has_k(s0) ||
        s->p.coulomb_type != s0->p.coulomb_type || has_bonded(s) != has_bonded(s0) || shake_on(s) != shake_on(s0) ||
        tc != s0->p.temp_ctrl)
      return fail(RBMD_ERR_ARG, "system %d differs from system 0 in force field, precision, constraints or thermostat", k);
    if (s->v.k_lanes_lg != s0->v.k_lanes_lg || s->v.rho_kb != s0->v.rho_kb)
      return fail(RBMD_ERR_ARG, "system %d differs from system 0 in its k-space tile choices", k);
    for (int j = 0; j < k; ++j)
      if (systems[j] == s) return fail(RBMD_ERR_ARG, "system %d appears twice", k);
  }
  HIP_TRY(hipSetDevice(ctx->device));
  rbmd_pack* p = new rbmd_pack();
  p->ctx = ctx;
  p->sys.assign(systems, systems + n);
  p->tc = s0->p.temp_ctrl;
  p->has_k = has_k(s0);
  p->has_b = has_bonded(s0);
  p->shake = shake_on(s0);
  p->ring = p->has_k && s0->p.coulomb_type == RBMD_COULOMB_RBE;
  p->nh = true;
  p->scale_v = p->tc == RBMD_TEMP_NOSE_HOOVER && !p->shake;
  p->sig.pair_mode = mode_of(s0);
  p->sig.pair_sp = sp_of(s0);
  p->sig.f32 = s0->v.mixed;
  p->sig.has_k = p->has_k;
  p->sig.ewald = s0->p.coulomb_type == RBMD_COULOMB_EWALD;
  p->sig.has_b = p->has_b;
  p->sig.k_lanes_lg = s0->v.k_lanes_lg;
  p->sig.rho_kb = s0->v.rho_kb;
  p->sig.thermo = 1;
  auto dalloc = [&](void** q, size_t bytes) -> rbmd_status {
    hipError_t e = hipMalloc(q, std::max<size_t>(bytes, 4));
    if (e != hipSuccess) return fail(RBMD_ERR_OOM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    p->allocs.push_back(*q);
    return RBMD_OK;
  };
  std::vector<PackSys> tab(n);
  std::vector<int> first((size_t)PK_COUNT * (n + 1));
  std::vector<std::vector<int>> blocks(PK_COUNT, std::vector<int>(n, 0));
  for (int k = 0; k < n; ++k) {
    rbmd_sys* s = systems[k];
    if (p->ring && !s->kgraph) {
      rbmd_status rc = s->dalloc(&s->kgraph, 3 * (size_t)s->P);
      if (rc != RBMD_OK) return pack_fail_free(p, rc);
    }
    PackSys& e = tab[k];
    std::memset(&e, 0, sizeof(e));
    e.v = s->v;
    e.v.nslab = 1;
    e.iv = s->iv;
    e.kv = s->kv;
    e.kv.P = s->P;
    e.kv.kvec = p->ring ? s->kgraph : s->kcur;
    e.kv.S_over_P = s->p.coulomb_type == RBMD_COULOMB_RBE ? s->S / (double)s->P : 1.0;
    e.kv.f32 = s->v.mixed;
    e.kv.ewald = s->p.coulomb_type == RBMD_COULOMB_EWALD;
    e.nhc = s->nhc;
    e.np = s->np;
    e.rd = pack_rows_desc(s, p->has_k, p->has_b, p->shake);
    e.rows_pair = s->seg_rows(SEG_PAIR);
    e.rows_k = s->seg_rows(SEG_K);
    e.rows_bond = s->seg_rows(SEG_BOND);
    e.rows_ke = s->seg_rows(SEG_KE);
    e.rows_shake = s->seg_rows(SEG_SHAKE);
    e.shake_iter = s->shake_iter;
    e.kring = p->ring ? s->kring : nullptr;
    e.ring_R = s->ring_R;
    for (int q = 0; q < PK_COUNT; ++q) {
      bool on = true;
      if (q == PK_KSLOT) on = p->ring;
      if (q == PK_RHO || q == PK_RHO_RED || q == PK_KFORCE) on = p->has_k;
      if (q == PK_BONDED) on = p->has_b;
      if (q == PK_SHAKE || q == PK_RATTLE) on = p->shake;
      blocks[q][k] = on ? pack_blocks(s->v, s->iv, s->P, q) : 0;
    }
    if (s->v.lds_bins) {
      p->lds_first = std::max(p->lds_first, (size_t)s->v.ncell * sizeof(int));
      p->lds_scatter = std::max(p->lds_scatter, 2 * (size_t)s->v.ncell * sizeof(int));
    }
  }
  // block maps: kernel q's map starts at moff[q]
  std::vector<size_t> moff(PK_COUNT + 1, 0);
  for (int q = 0; q < PK_COUNT; ++q) {
    int64_t tot = 0;
    for (int k = 0; k < n; ++k) {
      first[(size_t)q * (n + 1) + k] = (int)tot;
      tot += blocks[q][k];
    }
    if (tot > (int64_t)0x7fffffff) return pack_fail_free(p, fail(RBMD_ERR_UNSUPPORTED, "pack grid too large"));
    first[(size_t)q * (n + 1) + n] = (int)tot;
    p->nblk[q] = (int)tot;
    moff[q + 1] = moff[q] + (size_t)tot;
  }
  std::vector<int> map(std::max<size_t>(moff[PK_COUNT], 1));
  for (int q = 0; q < PK_COUNT; ++q)
    for (int k = 0; k < n; ++k)
      for (int b = 0; b < blocks[q][k]; ++b) map[moff[q] + first[(size_t)q * (n + 1) + k] + b] = k;
  rbmd_status rc = RBMD_OK;
  if (rc == RBMD_OK) rc = dalloc((void**)&p->dsys, sizeof(PackSys) * n);
  if (rc == RBMD_OK) rc = dalloc((void**)&p->dmap, sizeof(int) * map.size());
  if (rc == RBMD_OK) rc = dalloc((void**)&p->dfirst, sizeof(int) * first.size());
  if (rc == RBMD_OK) rc = dalloc((void**)&p->dstatus, sizeof(int) * n);
  if (rc != RBMD_OK) return pack_fail_free(p, rc);
  if (hipHostMalloc((void**)&p->hstatus, sizeof(int) * n) != hipSuccess)
    return pack_fail_free(p, fail(RBMD_ERR_OOM, "pinned status array"));
  if (hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess)
    return pack_fail_free(p, fail(RBMD_ERR_HIP, "hipEventCreate"));
  hipStream_t st = ctx->stream;
  if (hipMemcpyAsync(p->dsys, tab.data(), sizeof(PackSys) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(p->dmap, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(p->dfirst, first.data(), sizeof(int) * first.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return pack_fail_free(p, fail(RBMD_ERR_HIP, "pack table upload"));
  for (int q = 0; q < PK_COUNT; ++q) p->L[q] = PackLaunch{p->dsys, p->dmap + moff[q], p->dfirst + (size_t)q * (n + 1)};
  *out = p;
  return RBMD_OK;
}

rbmd_status rbmd_pack_destroy(rbmd_pack* p) {
  if (!p) return RBMD_OK;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  for (auto g : p->graph)
    if (g) (void)hipGraphExecDestroy(g);
  return pack_fail_free(p, RBMD_OK);
}

// n packed steps; graph: replay one captured packed step
static rbmd_status pack_steps(rbmd_pack* p, int64_t n_steps, bool graph) {
  if (!p) return fail(RBMD_ERR_ARG, "null pack");
  hipStream_t st = p->ctx->stream;
  for (int thermo = 0; graph && thermo < 2; ++thermo) {
    if (p->graph[thermo]) continue;
    const bool prof = p->ctx->prof;
    p->ctx->prof = false;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    pack_enqueue_step(p, thermo != 0);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    p->ctx->prof = prof;
    if (e != hipSuccess) return fail(RBMD_ERR_HIP, "pack step capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&p->graph[thermo], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(RBMD_ERR_HIP, "pack graph instantiate: %s", hipGetErrorString(e));
  }
  std::vector<int64_t> next(p->sys.size());
  for (int64_t k = 0; k < n_steps; ++k) {
    for (size_t q = 0; q < p->sys.size(); ++q) next[q] = p->sys[q]->step + 1;
    RC_TRY(ring_fill(p->sys.data(), next.data(), (int)p->sys.size()));
    const bool thermo = k == n_steps - 1;
    if (graph)
      HIP_TRY(hipGraphLaunch(p->graph[thermo ? 1 : 0], st));
    else
      pack_enqueue_step(p, thermo);
    for (rbmd_sys* s : p->sys) s->step += 1;
  }
  return check_launch("pack steps");
}

rbmd_status rbmd_pack_step(rbmd_pack* p, int64_t n_steps) { return pack_steps(p, n_steps, false); }

rbmd_status rbmd_pack_enqueue(rbmd_pack* p, int64_t n_steps) { return pack_steps(p, n_steps, true); }

rbmd_status rbmd_pack_wait(rbmd_pack* p, int32_t* status_out) {
  if (!p) return fail(RBMD_ERR_ARG, "null pack");
  hipStream_t st = p->ctx->stream;
  launch_pack_integrate(PK_ONE, p->L[PK_ONE], p->nblk[PK_ONE], p->sig, 1, 0, p->dstatus, st);
  RC_TRY(check_launch("pack status"));
  const int n = (int)p->sys.size();
  HIP_TRY(hipMemcpyAsync(p->hstatus, p->dstatus, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  int first_bad = -1;
  for (int k = 0; k < n; ++k) {
    if (status_out) status_out[k] = p->hstatus[k];
    if (p->hstatus[k] && first_bad < 0) first_bad = k;
  }
  if (first_bad < 0) return RBMD_OK;
  rbmd_status rc = status_to_rc(p->hstatus[first_bad]);
  std::string msg = g_last_error;
  return fail(rc, "pack system %d: %s", first_bad, msg.c_str());
}

// sizes of the ABI structs, checked from Python against the ctypes mirror
int32_t rbmd_abi_sizeof(int32_t which) {
  switch (which) {
    case 0:
      return (int32_t)sizeof(rbmd_sys_desc);
    case 1:
      return (int32_t)sizeof(rbmd_params);
    case 2:
      return (int32_t)sizeof(rbmd_thermo);
    default:
      return -1;
  }
}

}  // extern "C"
