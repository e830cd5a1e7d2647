// k_integrate.hip -- velocity Verlet, PBC, SHAKE/RATTLE, bonded terms, force
// assembly, reductions and the Nose-Hoover chain, all FP64 and device resident.
//
//   UpdateVelocity        RunWorklet.cpp:2357-2375 (v += 0.5 F/m dt fmt2v)
//   UpdatePosition + PBC  RunWorklet.cpp:2446-2490, 2602-2636 (wrap on x < lo / x >= hi)
//   SHAKE                 MolecularWorklet.h:2367-2645 (NewConstraintAWaterBondAngleWorklet)
//   RATTLE                MolecularWorklet.h:2647-2832 (NewConstraintBWaterBondAngleWorklet)
//   bonds / angles        MolecularWorklet.h:58-397 (gathered per atom instead of keyed reduce)
//   ComputeTempe          ExecutionNVT.cpp:736-773
//   NHCTempIntegrate      ExecutionNVT.cpp:1172-1256 (run on the device by one thread)
//   Compute_Pressure      ExecutionNVT.cpp:1258-1276, ComputeVirial ExecutionMD.cpp:1263-1366
#include "dev.hpp"
#include "pack.hpp"
#include "reduce.hpp"

namespace rbmd {

// NHC scale (optional), kick, drift, PBC of atom i; returns its cell when
// do_bin (the binning of the drifted position), else -1
__device__ __forceinline__ int first_half_atom(const SysView& s, const IntView& iv, const double* scale, int do_bin,
                                               int i) {
#pragma clang fp contract(off)  // first_half_regs rounds the same way
  double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
  if (scale) {
    double f = *scale;
    vx = f * vx;
    vy = f * vy;
    vz = f * vz;
  }
  s.ovx[i] = vx;
  s.ovy[i] = vy;
  s.ovz[i] = vz;
  double m = s.mass[i];
  vx += 0.5 * s.fx[i] / m * iv.dt * iv.fmt2v;
  vy += 0.5 * s.fy[i] / m * iv.dt * iv.fmt2v;
  vz += 0.5 * s.fz[i] / m * iv.dt * iv.fmt2v;
  s.vx[i] = vx;
  s.vy[i] = vy;
  s.vz[i] = vz;
  double p[3] = {s.x[i], s.y[i], s.z[i]};
  s.ox[i] = p[0];
  s.oy[i] = p[1];
  s.oz[i] = p[2];
  p[0] += vx * iv.dt;
  p[1] += vy * iv.dt;
  p[2] += vz * iv.dt;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (p[d] < s.lo[d]) {
      p[d] += s.box[d];
      s.img[3 * i + d] -= 1;
    } else if (p[d] >= s.hi[d]) {
      p[d] -= s.box[d];
      s.img[3 * i + d] += 1;
    }
  }
  s.x[i] = p[0];
  s.y[i] = p[1];
  s.z[i] = p[2];
  if (!isfinite(p[0] + p[1] + p[2] + vx + vy + vz)) atomicOr(s.status, 1);
  if (!do_bin) return -1;
  int a = bin_axis(p[0], s.lo[0], s.dxdydz[0], s.dims[0]);
  int b = bin_axis(p[1], s.lo[1], s.dxdydz[1], s.dims[1]);
  int c = bin_axis(p[2], s.lo[2], s.dxdydz[2], s.dims[2]);
  const int cell = a + b * s.dims[0] + c * s.dims[0] * s.dims[1];
  s.cell_of[i] = cell;
  return cell;
}

// NHC scale (optional), kick, drift, PBC, binning
__device__ __forceinline__ void first_half_body(const SysView& s, const IntView& iv, const double* scale, int do_bin, const int bid) {
  extern __shared__ int lds_cells[];
  int i = bid * blockDim.x + threadIdx.x;
  int cell = -1;
  if (i < s.n) cell = first_half_atom(s, iv, scale, do_bin, i);
  if (do_bin) count_cells(s, cell, lds_cells);
}

__global__ __launch_bounds__(kBlock) void k_first_half(SysView s, IntView iv, const double* scale, int do_bin) {
  first_half_body(s, iv, scale, do_bin, (int)blockIdx.x);
}

void launch_first_half(const SysView& s, const IntView& iv, hipStream_t st, const double* scale, int do_bin) {
  size_t lds = (do_bin && s.lds_bins) ? (size_t)s.ncell * sizeof(int) : 0;
  hipLaunchKernelGGL(k_first_half, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), lds, st, s, iv, scale, do_bin);
}

struct D3 {
  double x, y, z;
};
// The vector helpers, SHAKE and RATTLE compile with FMA contraction off: a
// product and a sum written in separate helpers are then rounded separately
// in every kernel that inlines them, so the fused and the separate launches
// of these steps agree to the bit.
__device__ __forceinline__ D3 sub(D3 a, D3 b) {
#pragma clang fp contract(off)
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
__device__ __forceinline__ D3 add(D3 a, D3 b) {
#pragma clang fp contract(off)
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
__device__ __forceinline__ D3 mul(double k, D3 a) {
#pragma clang fp contract(off)
  return {k * a.x, k * a.y, k * a.z};
}
__device__ __forceinline__ D3 divs(D3 a, double k) {
#pragma clang fp contract(off)
  return {a.x / k, a.y / k, a.z / k};
}
__device__ __forceinline__ double dotp(D3 a, D3 b) {
#pragma clang fp contract(off)
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ D3 mimg(const SysView& s, D3 a, D3 b) {
#pragma clang fp contract(off)
  D3 v = sub(a, b);
  if (fabs(v.x) > s.box[0] * 0.5) v.x -= (v.x > 0 ? s.box[0] : -s.box[0]);
  if (fabs(v.y) > s.box[1] * 0.5) v.y -= (v.y > 0 ? s.box[1] : -s.box[1]);
  if (fabs(v.z) > s.box[2] * 0.5) v.z -= (v.z > 0 ? s.box[2] : -s.box[2]);
  return v;
}
__device__ __forceinline__ D3 ld(const double* x, const double* y, const double* z, int i) { return {x[i], y[i], z[i]}; }
__device__ __forceinline__ void st3(double* x, double* y, double* z, int i, D3 v) {
  x[i] = v.x;
  y[i] = v.y;
  z[i] = v.z;
}

// SHAKE on (old x, old v, F) of one H-O-H triplet; writes x, v, image flags
// SHAKE of one molecule from its atoms' old positions ox, old velocities ovv,
// forces ff and masses m; writes x, v, image flags; vir: the constraint virial
__device__ __forceinline__ void shake_mol(const SysView& s, const IntView& iv, const int (&ids)[3], const D3 (&ox)[3],
                                          const D3 (&ovv)[3], const D3 (&ff)[3], const double (&m)[3], double (&vir)[6],
                                          int& my_iter) {
#pragma clang fp contract(off)
  {
    const double dt = iv.dt, fmt2v = iv.fmt2v;
    D3 sp[3];
    for (int t = 0; t < 3; ++t)
      sp[t] = add(add(ox[t], mul(dt, ovv[t])), mul(fmt2v, divs(mul(0.5 * dt * dt, ff[t]), m[t])));
    const double bond1 = 1.0, bond2 = 1.0;
    const double bond12 = iv.shake_d12;  // sqrt(b1^2 + b2^2 - 2 b1 b2 cos(109.47 deg)), MolecularWorklet.h's geometry
    D3 r01 = mimg(s, ox[1], ox[0]);
    D3 r12 = mimg(s, ox[2], ox[1]);
    D3 r20 = mimg(s, ox[0], ox[2]);
    D3 s10 = mimg(s, sp[0], sp[1]);
    D3 s21 = mimg(s, sp[1], sp[2]);
    D3 s02 = mimg(s, sp[2], sp[0]);
    double r01sq = dotp(r01, r01), r02sq = dotp(r20, r20), r12sq = dotp(r12, r12);
    double s01sq = dotp(s10, s10), s02sq = dotp(s02, s02), s12sq = dotp(s21, s21);
    double im0 = 1.0 / m[0], im1 = 1.0 / m[1], im2 = 1.0 / m[2];
    double a11 = 2.0 * (im0 + im1) * dotp(s10, r01);
    double a12 = -2.0 * im1 * dotp(s10, r12);
    double a13 = -2.0 * im0 * dotp(s10, r20);
    double a21 = -2.0 * im1 * dotp(s21, r01);
    double a22 = 2.0 * (im1 + im2) * dotp(s21, r12);
    double a23 = -2.0 * im2 * dotp(s21, r20);
    double a31 = -2.0 * im0 * dotp(s02, r01);
    double a32 = -2.0 * im2 * dotp(s02, r12);
    double a33 = 2.0 * (im0 + im2) * dotp(s02, r20);
    double det = a11 * a22 * a33 + a12 * a23 * a31 + a13 * a21 * a32 - a11 * a23 * a32 - a12 * a21 * a33 - a13 * a22 * a31;
    bool bad = fabs(det) < 0.0001;
    double di = 1.0 / det;
    double b11 = di * (a22 * a33 - a23 * a32), b12 = -di * (a12 * a33 - a13 * a32), b13 = di * (a12 * a23 - a13 * a22);
    double b21 = -di * (a21 * a33 - a23 * a31), b22 = di * (a11 * a33 - a13 * a31), b23 = -di * (a11 * a23 - a13 * a21);
    double b31 = di * (a21 * a32 - a22 * a31), b32 = -di * (a11 * a32 - a12 * a31), b33 = di * (a11 * a22 - a12 * a21);
    double r0120 = dotp(r01, r20), r0112 = dotp(r01, r12), r2012 = dotp(r20, r12);
    double q1_0101 = (im0 + im1) * (im0 + im1) * r01sq, q1_1212 = im1 * im1 * r12sq, q1_2020 = im0 * im0 * r02sq;
    double q1_0120 = -2.0 * (im0 + im1) * im0 * r0120, q1_0112 = -2.0 * (im0 + im1) * im1 * r0112;
    double q1_2012 = 2.0 * im0 * im1 * r2012;
    double q2_0101 = im1 * im1 * r01sq, q2_1212 = (im1 + im2) * (im1 + im2) * r12sq, q2_2020 = im2 * im2 * r02sq;
    double q2_0120 = 2.0 * im1 * im2 * r0120, q2_0112 = -2.0 * (im1 + im2) * im1 * r0112;
    double q2_2012 = -2.0 * (im1 + im2) * im2 * r2012;
    double q3_0101 = im0 * im0 * r01sq, q3_1212 = im2 * im2 * r12sq, q3_2020 = (im0 + im2) * (im0 + im2) * r02sq;
    double q3_0120 = -2.0 * (im0 + im2) * im0 * r0120, q3_0112 = 2.0 * im0 * im2 * r0112;
    double q3_2012 = -2.0 * (im0 + im2) * im2 * r2012;
    const double tol = iv.shake_tol;  // the reference's 0.00001 unless set (MolecularWorklet.h:2507-2508)
    const int max_iter = 5000;
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
      if (fabs(n01 - l01) > tol) done = 0;
      if (fabs(n20 - l20) > tol) done = 0;
      if (fabs(n12 - l12) > tol) done = 0;
      l01 = n01;
      l20 = n20;
      l12 = n12;
      if (isnan(l01) || isnan(l20) || isnan(l12) || fabs(l01) > 1e20 || fabs(l20) > 1e20 || fabs(l12) > 1e20) {
        done = 1;
        ovf = 1;
      }
      niter++;
    }
    if (!done || ovf || bad) atomicOr(s.status, 2);
    my_iter = niter;
    D3 pc[3];
    pc[0] = sub(mul(im0, mul(l01, r01)), mul(im0, mul(l20, r20)));
    pc[1] = sub(mul(im1, mul(l12, r12)), mul(im1, mul(l01, r01)));
    pc[2] = sub(mul(im2, mul(l20, r20)), mul(im2, mul(l12, r12)));
    for (int t = 0; t < 3; ++t) {
      int id = ids[t];
      D3 sv = add(ovv[t], mul(fmt2v, divs(mul(0.5 * dt, ff[t]), m[t])));
      st3(s.vx, s.vy, s.vz, id, add(sv, divs(pc[t], dt)));
      D3 np = add(sp[t], pc[t]);
      double p[3] = {np.x, np.y, np.z};
      for (int d = 0; d < 3; ++d) {
        double rng = s.hi[d] - s.lo[d];
        if (p[d] < s.lo[d]) {
          p[d] += rng;
          s.img[3 * id + d] -= 1;
        }
        if (p[d] > s.hi[d]) {
          p[d] -= rng;
          s.img[3 * id + d] += 1;
        }
      }
      s.x[id] = p[0];
      s.y[id] = p[1];
      s.z[id] = p[2];
    }
    double dtfsq = 0.5 * dt * dt * fmt2v;
    l01 /= dtfsq;
    l20 /= dtfsq;
    l12 /= dtfsq;
    vir[0] = l01 * r01.x * r01.x + l20 * r20.x * r20.x + l12 * r12.x * r12.x;
    vir[1] = l01 * r01.y * r01.y + l20 * r20.y * r20.y + l12 * r12.y * r12.y;
    vir[2] = l01 * r01.z * r01.z + l20 * r20.z * r20.z + l12 * r12.z * r12.z;
    vir[3] = l01 * r01.x * r01.y + l20 * r20.x * r20.y + l12 * r12.x * r12.y;
    vir[4] = l01 * r01.x * r01.z + l20 * r20.x * r20.z + l12 * r12.x * r12.z;
    vir[5] = l01 * r01.y * r01.z + l20 * r20.y * r20.z + l12 * r12.y * r12.z;
  }
}

// the SHAKE rows of a block and the iteration maximum
__device__ __forceinline__ void shake_finish(double* rows, int* max_iter_out, const double (&vir)[6], int my_iter,
                                             const int bid) {
  if (max_iter_out) {  // the wavefront's maximum, one atomic per wavefront (not one per molecule)
    int mx = my_iter;
    for (int off = 1; off < kWave; off <<= 1) mx = max(mx, __shfl_xor(mx, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(max_iter_out, mx);
  }
  const double vals[kRowShake] = {vir[0], vir[1], vir[2], vir[3], vir[4], vir[5]};
  block_write_row<kRowShake>(vals, rows + (size_t)bid * kRowShake);
}

// SHAKE on (old x, old v, F) of one H-O-H triplet per thread
__device__ __forceinline__ void shake_body(const SysView& s, const IntView& iv, double* rows, int* max_iter_out, const int bid) {
  int a = bid * blockDim.x + threadIdx.x;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  int my_iter = 0;
  if (a < iv.nmol) {
    int ids[3] = {iv.angle_atoms[3 * a], iv.angle_atoms[3 * a + 1], iv.angle_atoms[3 * a + 2]};
    D3 ox[3], ovv[3], ff[3];
    double m[3];
    for (int t = 0; t < 3; ++t) {
      int id = ids[t];
      ox[t] = ld(s.ox, s.oy, s.oz, id);
      ovv[t] = ld(s.ovx, s.ovy, s.ovz, id);
      ff[t] = ld(s.fx, s.fy, s.fz, id);
      m[t] = s.mass[id];
    }
    shake_mol(s, iv, ids, ox, ovv, ff, m, vir, my_iter);
  }
  shake_finish(rows, max_iter_out, vir, my_iter, bid);
}

__global__ __launch_bounds__(kBlock) void k_shake(SysView s, IntView iv, double* rows, int* max_iter_out) {
  shake_body(s, iv, rows, max_iter_out, (int)blockIdx.x);
}

// first_half_atom on values already in registers (x: the old position, v:
// the velocity, scaled in place to the saved old velocity), storing only what
// SHAKE does not overwrite: the saved velocity, image flags, status, cell.
// The same operations as first_half_atom (both without FMA contraction, so
// the compiler cannot round them differently in the two kernels).
__device__ __forceinline__ int first_half_regs(const SysView& s, const IntView& iv, const double* scale, int do_bin,
                                               int i, const D3& x, D3& v, const D3& f, double m) {
#pragma clang fp contract(off)
  double vx = v.x, vy = v.y, vz = v.z;
  if (scale) {
    double fs = *scale;
    vx = fs * vx;
    vy = fs * vy;
    vz = fs * vz;
  }
  s.ovx[i] = vx;
  s.ovy[i] = vy;
  s.ovz[i] = vz;
  v = {vx, vy, vz};
  vx += 0.5 * f.x / m * iv.dt * iv.fmt2v;
  vy += 0.5 * f.y / m * iv.dt * iv.fmt2v;
  vz += 0.5 * f.z / m * iv.dt * iv.fmt2v;
  double p[3] = {x.x, x.y, x.z};
  p[0] += vx * iv.dt;
  p[1] += vy * iv.dt;
  p[2] += vz * iv.dt;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (p[d] < s.lo[d]) {
      p[d] += s.box[d];
      s.img[3 * i + d] -= 1;
    } else if (p[d] >= s.hi[d]) {
      p[d] -= s.box[d];
      s.img[3 * i + d] += 1;
    }
  }
  if (!isfinite(p[0] + p[1] + p[2] + vx + vy + vz)) atomicOr(s.status, 1);
  if (!do_bin) return -1;
  int ca = bin_axis(p[0], s.lo[0], s.dxdydz[0], s.dims[0]);
  int cb = bin_axis(p[1], s.lo[1], s.dxdydz[1], s.dims[1]);
  int cc = bin_axis(p[2], s.lo[2], s.dxdydz[2], s.dims[2]);
  const int cell = ca + cb * s.dims[0] + cc * s.dims[0] * s.dims[1];
  s.cell_of[i] = cell;
  return cell;
}

// The first half step and SHAKE in one launch, for systems whose every atom
// is in a SHAKE triplet: the thread of molecule a runs first_half_atom on its
// three atoms from registers, counts their cells, then runs SHAKE on the same
// registers.  Bit for bit the two launches: the same arithmetic, with the
// kicked velocity, drifted position and saved old position that SHAKE
// replaces (or that only it reads) left unstored; the cells are counted from
// the drifted positions, as the reference's locator bins before SHAKE.
__global__ __launch_bounds__(kBlock) void k_first_half_shake(SysView s, IntView iv, const double* scale, int do_bin,
                                                             double* rows, int* max_iter_out) {
  extern __shared__ int lds_cells[];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  int cell[3] = {-1, -1, -1};
  double vir[6] = {0, 0, 0, 0, 0, 0};
  int my_iter = 0;
  int ids[3];
  D3 ox[3], ovv[3], ff[3];
  double m[3];
  if (a < iv.nmol) {
    for (int t = 0; t < 3; ++t) {  // every load first, then the arithmetic
      ids[t] = iv.angle_atoms[3 * a + t];
      ox[t] = ld(s.x, s.y, s.z, ids[t]);
      ovv[t] = ld(s.vx, s.vy, s.vz, ids[t]);
      ff[t] = ld(s.fx, s.fy, s.fz, ids[t]);
      m[t] = s.mass[ids[t]];
    }
    for (int t = 0; t < 3; ++t)
      cell[t] = first_half_regs(s, iv, scale, do_bin, ids[t], ox[t], ovv[t], ff[t], m[t]);
  }
  if (do_bin) count_cells3(s, cell, lds_cells);
  if (a < iv.nmol) shake_mol(s, iv, ids, ox, ovv, ff, m, vir, my_iter);
  shake_finish(rows, max_iter_out, vir, my_iter, (int)blockIdx.x);
}

void launch_first_half_shake(const SysView& s, const IntView& iv, hipStream_t st, const double* scale, int do_bin,
                             double* rows, int* max_iter) {
  const int grid = (iv.nmol + kBlock - 1) / kBlock;
  const size_t lds = (do_bin && s.lds_bins) ? (size_t)s.ncell * sizeof(int) : 0;
  if (grid > 0)
    hipLaunchKernelGGL(k_first_half_shake, dim3(grid), dim3(kBlock), lds, st, s, iv, scale, do_bin, rows, max_iter);
}

// atoms outside every SHAKE triplet get x = x_old, v = v_old: the reference
// copies the whole old arrays back after the constraint worklets
// (ExecutionNVT.cpp:869-870, 896)
__global__ __launch_bounds__(kBlock) void k_unconstrained_copy(SysView s, const int* free_ids, int nfree, int pos) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nfree) return;
  int i = free_ids[k];
  if (pos) {
    s.x[i] = s.ox[i];
    s.y[i] = s.oy[i];
    s.z[i] = s.oz[i];
  }
  s.vx[i] = s.ovx[i];
  s.vy[i] = s.ovy[i];
  s.vz[i] = s.ovz[i];
}

void launch_shake(const SysView& s, const IntView& iv, hipStream_t st, double* rows, int* max_iter) {
  int grid = (iv.nmol + kBlock - 1) / kBlock;
  if (grid > 0) hipLaunchKernelGGL(k_shake, dim3(grid), dim3(kBlock), 0, st, s, iv, rows, max_iter);
}

void launch_unconstrained_copy(const SysView& s, hipStream_t st, const int* free_ids, int nfree, int pos) {
  if (nfree <= 0) return;
  hipLaunchKernelGGL(k_unconstrained_copy, dim3((nfree + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, free_ids,
                     nfree, pos);
}

// RATTLE of one H-O-H triplet on values in registers: the velocity starts
// again from the saved old velocity ov plus the half kick of the new force f,
// as the reference's RATTLE does (MolecularWorklet.h:2647-2832); x are the
// positions; the corrected velocities go to vout.
__device__ __forceinline__ void rattle_mol(const SysView& s, const IntView& iv, const D3 (&ov)[3], const D3 (&f)[3],
                                           const double (&m)[3], const D3 (&x)[3], D3 (&vout)[3]) {
#pragma clang fp contract(off)
  const double dt = iv.dt, fmt2v = iv.fmt2v;
  D3 sv[3];
  for (int t = 0; t < 3; ++t) sv[t] = add(ov[t], mul(fmt2v, divs(mul(0.5 * dt, f[t]), m[t])));
  const D3 p0 = x[0], p1 = x[1], p2 = x[2];
  D3 r01 = mimg(s, p1, p0), r12 = mimg(s, p2, p1), r20 = mimg(s, p0, p2);
  D3 sv10 = sub(sv[0], sv[1]), sv21 = sub(sv[1], sv[2]), sv02 = sub(sv[2], sv[0]);
  double im0 = 1.0 / m[0], im1 = 1.0 / m[1], im2 = 1.0 / m[2];
  double A[3][3];
  A[0][0] = (im1 + im0) * dotp(r01, r01);
  A[0][1] = -im1 * dotp(r01, r12);
  A[0][2] = (-im0) * dotp(r01, r20);
  A[1][0] = A[0][1];
  A[1][1] = (im1 + im2) * dotp(r12, r12);
  A[1][2] = -(im2)*dotp(r20, r12);
  A[2][0] = A[0][2];
  A[2][1] = A[1][2];
  A[2][2] = (im0 + im2) * dotp(r20, r20);
  double c[3] = {-dotp(sv10, r01), -dotp(sv21, r12), -dotp(sv02, r20)};
  double det = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] -
               A[0][0] * A[1][2] * A[2][1] - A[0][1] * A[1][0] * A[2][2] - A[0][2] * A[1][1] * A[2][0];
  if (fabs(det) < 0.0001) atomicOr(s.status, 2);
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
  D3 vc0 = sub(mul(im0, mul(l[0], r01)), mul(im0, mul(l[2], r20)));
  D3 vc1 = sub(mul(im1, mul(l[1], r12)), mul(im1, mul(l[0], r01)));
  D3 vc2 = sub(mul(im2, mul(l[2], r20)), mul(im2, mul(l[1], r12)));
  vout[0] = add(sv[0], vc0);
  vout[1] = add(sv[1], vc1);
  vout[2] = add(sv[2], vc2);
}

// RATTLE on (x, old v, F) of one H-O-H triplet per thread
__device__ __forceinline__ void rattle_body(const SysView& s, const IntView& iv, const int bid) {
  const int a = bid * blockDim.x + threadIdx.x;
  if (a >= iv.nmol) return;
  int ids[3];
  D3 ov[3], f[3], x[3], v[3];
  double m[3];
  for (int t = 0; t < 3; ++t) {
    ids[t] = iv.angle_atoms[3 * a + t];
    ov[t] = ld(s.ovx, s.ovy, s.ovz, ids[t]);
    f[t] = ld(s.fx, s.fy, s.fz, ids[t]);
    x[t] = ld(s.x, s.y, s.z, ids[t]);
    m[t] = s.mass[ids[t]];
  }
  rattle_mol(s, iv, ov, f, m, x, v);
  for (int t = 0; t < 3; ++t) st3(s.vx, s.vy, s.vz, ids[t], v[t]);
}

__global__ __launch_bounds__(kBlock) void k_rattle(SysView s, IntView iv) {
  rattle_body(s, iv, (int)blockIdx.x);
}

void launch_rattle(const SysView& s, const IntView& iv, hipStream_t st) {
  int grid = (iv.nmol + kBlock - 1) / kBlock;
  if (grid > 0) hipLaunchKernelGGL(k_rattle, dim3(grid), dim3(kBlock), 0, st, s, iv);
}

// The end of one step and the start of the next in one launch, for systems
// whose every atom is in a SHAKE triplet: the thread of molecule a runs RATTLE
// on its three atoms (the step's end), then, from the same registers, the
// next step's first half step, the cell count and SHAKE, as
// k_first_half_shake does.  Nothing between the two reads a velocity: the
// thermostat's step boundary has already run (it reads the kinetic energy
// before RATTLE), and RATTLE leaves the positions and forces as they are.
// Bit for bit the two launches k_rattle and k_first_half_shake.
__global__ __launch_bounds__(kBlock) void k_rattle_first_half_shake(SysView s, IntView iv, const double* scale,
                                                                    int do_bin, double* rows, int* max_iter_out) {
  extern __shared__ int lds_cells[];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  int cell[3] = {-1, -1, -1};
  double vir[6] = {0, 0, 0, 0, 0, 0};
  int my_iter = 0;
  int ids[3];
  D3 x[3], v[3], f[3];
  double m[3];
  if (a < iv.nmol) {
    D3 ov[3];
    for (int t = 0; t < 3; ++t) {  // every load first, then the arithmetic
      ids[t] = iv.angle_atoms[3 * a + t];
      ov[t] = ld(s.ovx, s.ovy, s.ovz, ids[t]);
      f[t] = ld(s.fx, s.fy, s.fz, ids[t]);
      x[t] = ld(s.x, s.y, s.z, ids[t]);
      m[t] = s.mass[ids[t]];
    }
    rattle_mol(s, iv, ov, f, m, x, v);
    for (int t = 0; t < 3; ++t) cell[t] = first_half_regs(s, iv, scale, do_bin, ids[t], x[t], v[t], f[t], m[t]);
  }
  if (do_bin) count_cells3(s, cell, lds_cells);
  if (a < iv.nmol) shake_mol(s, iv, ids, x, v, f, m, vir, my_iter);
  shake_finish(rows, max_iter_out, vir, my_iter, (int)blockIdx.x);
}

void launch_rattle_first_half_shake(const SysView& s, const IntView& iv, hipStream_t st, const double* scale,
                                    int do_bin, double* rows, int* max_iter) {
  const int grid = (iv.nmol + kBlock - 1) / kBlock;
  const size_t lds = (do_bin && s.lds_bins) ? (size_t)s.ncell * sizeof(int) : 0;
  if (grid > 0)
    hipLaunchKernelGGL(k_rattle_first_half_shake, dim3(grid), dim3(kBlock), lds, st, s, iv, scale, do_bin, rows,
                       max_iter);
}

// bond + angle forces gathered per atom over its incident terms; the energy
// and virial of a term are counted by its first atom.  With iv.spec_coul
// (CVFF + RBL + special bonds) also the intramolecular Coulomb correction of
// ComputeSpecialCoulGeneralWorklet (RunWorklet.cpp:2012-2084, added to the
// force by ExecutionNVT::ComputeAllForce, ExecutionNVT.cpp:212-214): for the
// other atoms j of i's molecule (molecules of 3 or more atoms),
//   F_i += (1 - w_ij) (-332.06371 q_i q_j / r^3) r,  r = MinDistanceVec(x_i, x_j),
// 332.06371 hard-coded as in the reference (hazard Q24), w_ij the last match of
// j in i's special list (else 1); its virial +1/2 r (x) F (ComputePairVirial_fix)
// goes into the bond virial slots (both enter the pressure once,
// ExecutionMD.cpp:1323-1358).  The random-batch real space has no exclusions,
// so this removes the intramolecular erfc Coulomb the pair term added.
template <bool THERMO>
__device__ __forceinline__ void bonded_body(const SysView& s, const IntView& iv, double* rows, const int bid) {
  int i = bid * blockDim.x + threadIdx.x;
  double f[3] = {0, 0, 0};
  double eb = 0, ea = 0;
  double vb[6] = {0, 0, 0, 0, 0, 0}, va[6] = {0, 0, 0, 0, 0, 0};
  if (i < s.n && iv.spec_coul && iv.grp_off[i + 1] - iv.grp_off[i] >= 3) {
    const D3 pi = ld(s.x, s.y, s.z, i);
    const double qi = s.q[i];
    for (int k = iv.grp_off[i]; k < iv.grp_off[i + 1]; ++k) {
      const int j = iv.grp_ids[k];
      if (j == i) continue;
      const D3 r = mimg(s, pi, ld(s.x, s.y, s.z, j));
      const double dis = sqrt(dotp(r, r));
      const double fcomp = -332.06371 * qi * s.q[j] / (dis * dis * dis);
      double w = 1.0;
      if (s.sp_off)
        for (int t = s.sp_off[i]; t < s.sp_off[i + 1]; ++t)
          if (s.sp_ids[t] == j) w = s.sp_w[t];
      const D3 cp = mul((1.0 - w) * fcomp, r);
      f[0] += cp.x;
      f[1] += cp.y;
      f[2] += cp.z;
      if (THERMO) {
        vb[0] += 0.5 * (r.x * cp.x);
        vb[1] += 0.5 * (r.y * cp.y);
        vb[2] += 0.5 * (r.z * cp.z);
        vb[3] += 0.5 * (r.x * cp.y);
        vb[4] += 0.5 * (r.x * cp.z);
        vb[5] += 0.5 * (r.y * cp.z);
      }
    }
  }
  if (i < s.n) {
    for (int k = iv.inc_b_off[i]; k < iv.inc_b_off[i + 1]; ++k) {
      int b = iv.inc_b[k];
      int a0 = iv.bond_atoms[2 * b], a1 = iv.bond_atoms[2 * b + 1];
      D3 r = mimg(s, ld(s.x, s.y, s.z, a0), ld(s.x, s.y, s.z, a1));
      double d = sqrt(dotp(r, r));
      double dr = d - iv.bond_r0[b];
      double rk = iv.bond_k[b] * dr;
      double fb = d > 0.01 ? -2.0 * rk / d : 0.0;
      D3 fi = mul(fb, r);
      double sg = (a0 == i) ? 1.0 : -1.0;
      f[0] += sg * fi.x;
      f[1] += sg * fi.y;
      f[2] += sg * fi.z;
      if (THERMO && a0 == i) {
        eb += rk * dr;
        vb[0] += r.x * fi.x;
        vb[1] += r.y * fi.y;
        vb[2] += r.z * fi.z;
        vb[3] += r.x * fi.y;
        vb[4] += r.x * fi.z;
        vb[5] += r.y * fi.z;
      }
    }
    for (int k = iv.inc_a_off[i]; k < iv.inc_a_off[i + 1]; ++k) {
      int a = iv.inc_a[k];
      int ai = iv.angle_all[3 * a], aj = iv.angle_all[3 * a + 1], ak = iv.angle_all[3 * a + 2];
      D3 pj = ld(s.x, s.y, s.z, aj);
      D3 rij = mimg(s, ld(s.x, s.y, s.z, ai), pj);
      D3 rkj = mimg(s, ld(s.x, s.y, s.z, ak), pj);
      double dij2 = dotp(rij, rij), dij = sqrt(dij2);
      double dkj2 = dotp(rkj, rkj), dkj = sqrt(dkj2);
      double c = dotp(rij, rkj);
      c /= dij * dkj;
      if (c > 1.0) c = 1.0;
      if (c < -1.0) c = -1.0;
      double sn = sqrt(1.0 - c * c);
      if (sn < 0.001) sn = 0.001;
      sn = 1.0 / sn;
      double dth = acos(c) - (iv.angle_t0[a] * kPi) / 180;
      double tk = iv.angle_k[a] * dth;
      double aa = -2.0 * tk * sn;
      double a11 = aa * c / dij2, a12 = -aa / (dij * dkj), a22 = aa * c / dkj2;
      D3 fi = add(mul(a11, rij), mul(a12, rkj));
      D3 fk = add(mul(a22, rkj), mul(a12, rij));
      D3 mine = (i == ai) ? fi : ((i == ak) ? fk : mul(-1.0, add(fi, fk)));
      f[0] += mine.x;
      f[1] += mine.y;
      f[2] += mine.z;
      if (THERMO && i == ai) {
        ea += tk * dth;
        va[0] += rij.x * fi.x + rkj.x * fk.x;
        va[1] += rij.y * fi.y + rkj.y * fk.y;
        va[2] += rij.z * fi.z + rkj.z * fk.z;
        va[3] += rij.x * fi.y + rkj.x * fk.y;
        va[4] += rij.x * fi.z + rkj.x * fk.z;
        va[5] += rij.y * fi.z + rkj.y * fk.z;
      }
    }
    s.fbx[i] = f[0];
    s.fby[i] = f[1];
    s.fbz[i] = f[2];
  }
  if (THERMO) {
    // slots S_VIR_BOND .. S_VIR_ANGLE + 5, S_EBOND, S_EANGLE
    const double vals[kRowBond] = {vb[0], vb[1], vb[2], vb[3], vb[4], vb[5], va[0],
                                   va[1], va[2], va[3], va[4], va[5], eb,    ea};
    block_write_row<kRowBond>(vals, rows + (size_t)bid * kRowBond);
  }
}

template <bool THERMO>
__global__ __launch_bounds__(kBlock) void k_bonded(SysView s, IntView iv, double* rows) {
  bonded_body<THERMO>(s, iv, rows, (int)blockIdx.x);
}

void launch_bonded(const SysView& s, const IntView& iv, hipStream_t st, int thermo, double* rows, int* nblocks) {
  int grid = (s.n + kBlock - 1) / kBlock;
  if (nblocks) *nblocks = grid;
  if (thermo)
    hipLaunchKernelGGL(k_bonded<true>, dim3(grid), dim3(kBlock), 0, st, s, iv, rows);
  else
    hipLaunchKernelGGL(k_bonded<false>, dim3(grid), dim3(kBlock), 0, st, s, iv, rows);
}

// F = F_real[rank] (slab partials in order) + F_k + F_bonded; optional second kick;
// kinetic energy row
__device__ __forceinline__ void second_half_body(const SysView& s, const IntView& iv, int kick, int has_k, int has_b, double* rows, const int bid) {
  int i = bid * blockDim.x + threadIdx.x;
  double ke = 0.0;
  if (i < s.n) {
    int r = s.rank[i];
    double fx = s.fnx[r], fy = s.fny[r], fz = s.fnz[r];
    for (int k = 1; k < s.nslab; ++k) {
      const size_t o = (size_t)k * s.n + r;
      fx += s.fnx[o];
      fy += s.fny[o];
      fz += s.fnz[o];
    }
    if (has_k) {
      fx = s.fkx[i] + fx;
      fy = s.fky[i] + fy;
      fz = s.fkz[i] + fz;
    }
    if (has_b) {
      fx += s.fbx[i];
      fy += s.fby[i];
      fz += s.fbz[i];
    }
    s.fx[i] = fx;
    s.fy[i] = fy;
    s.fz[i] = fz;
    double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
    double m = s.mass[i];
    if (kick) {
      s.ovx[i] = vx;
      s.ovy[i] = vy;
      s.ovz[i] = vz;
      vx += 0.5 * fx / m * iv.dt * iv.fmt2v;
      vy += 0.5 * fy / m * iv.dt * iv.fmt2v;
      vz += 0.5 * fz / m * iv.dt * iv.fmt2v;
      s.vx[i] = vx;
      s.vy[i] = vy;
      s.vz[i] = vz;
    }
    ke = m * (vx * vx + vy * vy + vz * vz) * iv.mvv2e;
    if (!isfinite(fx + fy + fz)) atomicOr(s.status, 1);
  }
  const double vals[kRowKe] = {ke};
  block_write_row<kRowKe>(vals, rows + (size_t)bid * kRowKe);
}

__global__ __launch_bounds__(kBlock) void k_second_half(SysView s, IntView iv, int kick, int has_k, int has_b, double* rows) {
  second_half_body(s, iv, kick, has_k, has_b, rows, (int)blockIdx.x);
}

void launch_second_half(const SysView& s, const IntView& iv, hipStream_t st, int kick, int has_k, int has_b,
                        double* rows, int* nblocks) {
  int grid = (s.n + kBlock - 1) / kBlock;
  if (nblocks) *nblocks = grid;
  hipLaunchKernelGGL(k_second_half, dim3(grid), dim3(kBlock), 0, st, s, iv, kick, has_k, has_b, rows);
}

__global__ __launch_bounds__(kBlock) void k_ke_rows(SysView s, IntView iv, double* rows) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ke = 0.0;
  if (i < s.n) ke = s.mass[i] * (s.vx[i] * s.vx[i] + s.vy[i] * s.vy[i] + s.vz[i] * s.vz[i]) * iv.mvv2e;
  const double vals[kRowKe] = {ke};
  block_write_row<kRowKe>(vals, rows + (size_t)blockIdx.x * kRowKe);
}

void launch_ke_rows(const SysView& s, const IntView& iv, hipStream_t st, double* rows) {
  hipLaunchKernelGGL(k_ke_rows, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, iv, rows);
}

__device__ __forceinline__ void scale_v_body(const SysView& s, const double* scale, const int bid) {
  int i = bid * blockDim.x + threadIdx.x;
  if (i >= s.n) return;
  double f = *scale;
  s.vx[i] = f * s.vx[i];
  s.vy[i] = f * s.vy[i];
  s.vz[i] = f * s.vz[i];
}

__global__ __launch_bounds__(kBlock) void k_scale_v(SysView s, const double* scale) {
  scale_v_body(s, scale, (int)blockIdx.x);
}

void launch_scale_v(const SysView& s, hipStream_t st, const double* scale) {
  hipLaunchKernelGGL(k_scale_v, dim3((s.n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, s, scale);
}

// ------------------------------------------------------------------ NHC


// e^x, with e^(+-0) = 1 exactly without the call (the last chain element's
// velocity eta_dot[3] stays 0, so two of the six factors of a half step are
// exactly 1)
__device__ __forceinline__ double exp_or_one(double x) { return x == 0.0 ? 1.0 : exp(x); }

__device__ void nhc_integrate(NhcState& S, const NhcParams& p) {
#pragma clang fp contract(off)
  double T = S.tempT;
  double kecur = p.tdof * p.boltz * T;
  S.eta_mass[0] = p.tdof * p.boltz * S.t_target / (p.t_freq * p.t_freq);
  for (int c = 1; c < 3; ++c) S.eta_mass[c] = p.boltz * S.t_target / (p.t_freq * p.t_freq);
  S.eta_dotdot[0] = S.eta_mass[0] > 0.0 ? (kecur - S.ke_target) / S.eta_mass[0] : 0.0;
  const double ncfac = 1.0;
  double expfac;
  for (int c = 2; c > 0; --c) {
    expfac = exp_or_one(-ncfac * p.dt8 * S.eta_dot[c + 1]);
    S.eta_dot[c] *= expfac;
    S.eta_dot[c] += S.eta_dotdot[c] * ncfac * p.dt4;
    S.eta_dot[c] *= p.tdrag;
    S.eta_dot[c] *= expfac;
  }
  expfac = exp(-ncfac * p.dt8 * S.eta_dot[1]);
  S.eta_dot[0] *= expfac;
  S.eta_dot[0] += S.eta_dotdot[0] * ncfac * p.dt4;
  S.eta_dot[0] *= p.tdrag;
  S.eta_dot[0] *= expfac;
  S.factor = exp(-ncfac * p.dthalf * S.eta_dot[0]);
  S.tempT = T * S.factor * S.factor;
  kecur = p.tdof * p.boltz * S.tempT;
  S.eta_dotdot[0] = S.eta_mass[0] > 0.0 ? (kecur - S.ke_target) / S.eta_mass[0] : 0.0;
  for (int c = 0; c < 3; ++c) S.eta[c] += ncfac * p.dthalf * S.eta_dot[c];
  S.eta_dot[0] *= expfac;
  S.eta_dot[0] += S.eta_dotdot[0] * ncfac * p.dt4;
  S.eta_dot[0] *= expfac;
  for (int c = 1; c < 3; ++c) {
    expfac = exp_or_one(-ncfac * p.dt8 * S.eta_dot[c + 1]);
    S.eta_dot[c] *= expfac;
    S.eta_dotdot[c] = (S.eta_mass[c - 1] * S.eta_dot[c - 1] * S.eta_dot[c - 1] - p.boltz * S.t_target) / S.eta_mass[c];
    S.eta_dot[c] += S.eta_dotdot[c] * ncfac * p.dt4;
    S.eta_dot[c] *= expfac;
  }
}

__device__ void temp_target(NhcState& S, const NhcParams& p) {
#pragma clang fp contract(off)
  if (p.t_stop == p.t_start) {
    S.t_target = p.t_start;
  } else {
    double delta = (double)S.step;
    if (delta != 0.0) delta = delta / (double)p.num_steps;
    S.t_target = p.t_start + delta * (p.t_stop - p.t_start);
  }
  S.ke_target = p.tdof * p.boltz * S.t_target;
}

__device__ void compute_tempe(NhcState& S, const NhcParams& p, double ke_sum) {
#pragma clang fp contract(off)
  S.tempT_sum = ke_sum;
  double n = p.n;
  if (p.inbuild)
    S.tempT = 0.5 * ke_sum / (3 * n / 2.0);
  else if (p.fix_shake == 0)
    S.tempT = 0.5 * ke_sum / ((3 * n - 3) * p.kB / 2.0);
  else if (p.fix_shake == 1)
    S.tempT = 0.5 * ke_sum / ((3 * n - n - 3) * p.kB / 2.0);
  else
    S.tempT = 0.5 * ke_sum / ((3 * n) * p.kB / 2.0);
}

__device__ void compute_pressure(NhcState& S, const NhcParams& p, const double* sums) {
#pragma clang fp contract(off)
  for (int k = 0; k < 6; ++k) {
    double w = sums[S_VIR_NEAR + k];
    if (!p.lj_only) w += sums[S_VIR_K + k] * p.qqr2e;
    if (p.cvff) {
      w += sums[S_VIR_BOND + k] + sums[S_VIR_ANGLE + k];
      if (p.shake_vir) w += sums[S_VIR_SHAKE + k];
    }
    S.virial[k] = w;
  }
  S.pressure = (p.tdof * p.boltz * S.tempT + S.virial[0] + S.virial[1] + S.virial[2]) / 3.0 * (1.0 / p.volume_range) *
               p.nktv2p;
  S.e_lj = sums[S_ELJ] / p.n;
  S.e_coul = sums[S_ECOUL] / p.n;
  S.e_bond = p.nb > 0 ? sums[S_EBOND] / p.nb : 0.0;
  S.e_angle = p.na > 0 ? sums[S_EANGLE] / p.na : 0.0;
}

// Thread 0's thermostat step from the reduced sums.  The thermostat
// functions compile with FMA contraction off, so every kernel they are
// inlined into (single-system and packed) evaluates the same separately
// rounded operations and the chain variables agree to the bit.
__device__ __forceinline__ void nhc_phase(NhcState& S, const NhcParams& p, const double* sums, int phase) {
#pragma clang fp contract(off)
  switch (phase) {
    case 3:  // SetUp: ComputeTempe, ComputeTempTarget, chain masses (ExecutionNVT.cpp:1106-1122)
      compute_tempe(S, p, sums[S_KE]);
      temp_target(S, p);
      S.eta_mass[0] = p.tdof * p.boltz * S.t_target / (p.t_freq * p.t_freq);
      for (int c = 1; c < 3; ++c) S.eta_mass[c] = p.boltz * S.t_target / (p.t_freq * p.t_freq);
      for (int c = 1; c < 3; ++c)
        S.eta_dotdot[c] = (S.eta_mass[c - 1] * S.eta_dot[c - 1] * S.eta_dot[c - 1] - p.boltz * S.t_target) / S.eta_mass[c];
      compute_pressure(S, p, sums);
      S.factor = 1.0;
      S.factor_end = 1.0;
      break;
    case 1:  // InitialIntegrate: TakeStep increments the step, then ComputeTempTarget + NHC
      S.step += 1;
      if (p.shake_iter) *p.shake_iter = 0;
      if (p.mode == 0) {
        temp_target(S, p);
        nhc_integrate(S, p);
      } else {
        S.factor = 1.0;
      }
      break;
    case 2:  // FinalIntegrate: ComputeTempe, Compute_Pressure_Scalar, NHC
      compute_tempe(S, p, sums[S_KE]);
      compute_pressure(S, p, sums);
      if (p.mode == 0)
        nhc_integrate(S, p);
      else
        S.factor = 1.0;
      S.factor_end = S.factor;
      break;
    case 4: {  // Berendsen / rescale: ComputeTempe, rescale, pressure (ExecutionNVT.cpp:90-120, 270-291)
      compute_tempe(S, p, sums[S_KE]);
      double coeff;
      if (p.mode == 2)
        coeff = sqrt(p.t_start / S.tempT);
      else
        coeff = sqrt(1.0 + (p.dt / p.t_period) * (p.t_start / S.tempT - 1.0));
      S.factor = coeff;
      S.factor_end = coeff;
      compute_pressure(S, p, sums);
      break;
    }
    case 5:  // Berendsen-path step start: only the step counter
      S.step += 1;
      if (p.shake_iter) *p.shake_iter = 0;
      S.factor = 1.0;
      break;
    default:
      break;
  }
}

// phase 6: FinalIntegrate of this step (2) and then InitialIntegrate of the
// next (1), in one launch between two steps of a call: phase 1 reads only
// the state phase 2 leaves (the end-of-step scale stays in factor_end for the
// velocity scaling that follows).  The same operations in the same order as
// the two launches.
__device__ __forceinline__ void nhc_update(NhcState* state, const NhcParams& p, const double* sums, int phase) {
  NhcState S = *state;
  if (phase == 6) {
    nhc_phase(S, p, sums, 2);
    nhc_phase(S, p, sums, 1);
  } else {
    nhc_phase(S, p, sums, phase);
  }
  *state = S;
}

// One workgroup: reduce the partial rows (each wavefront owns a set of slots,
// lanes stride the rows, fixed-order shuffle tree), then thread 0 advances the
// thermostat state.
// nseg: thermo row segments to reduce (rd.nseg, or 0 for no reduction)
__device__ __forceinline__ void nhc_body(NhcState* state, const NhcParams& p, const RowsDesc& rd, const int nseg,
                                         int phase, const int ke_only, const int bid) {
  // All producer segments at once: segment g owns threads [base, base +
  // groups * width); thread (group, k) strides rows group, group + groups, ...
  // of column k (consecutive threads read consecutive doubles), then thread k
  // of the segment adds the group partials in order.  Fixed order throughout.
  __shared__ double part[1024];
  __shared__ double sums[S_NSLOT];
  const int tid = threadIdx.x;
  if (tid < S_NSLOT) sums[tid] = 0.0;
  int base = 0, mine = -1, mbase = 0;
  for (int g = 0; g < nseg; ++g) {
    const int span = rd.seg[g].groups * rd.seg[g].width;
    if (tid >= base && tid < base + span) {
      mine = g;
      mbase = base;
    }
    base += span;
  }
  if (mine >= 0) {
    const RowSeg sg = rd.seg[mine];
    const int k = (tid - mbase) % sg.width, grp = (tid - mbase) / sg.width;
    const int nrows = (ke_only && sg.slot0 != S_KE) ? 0 : sg.nrows;  // skipped sums read as zero
    double acc = 0.0;
    for (int r = grp; r < nrows; r += sg.groups) acc += sg.rows[(size_t)r * sg.width + k];
    part[tid] = acc;
  }
  __syncthreads();
  if (mine >= 0 && tid - mbase < rd.seg[mine].width) {
    const RowSeg sg = rd.seg[mine];
    double acc = 0.0;
    for (int grp = 0; grp < sg.groups; ++grp) acc += part[mbase + grp * sg.width + (tid - mbase)];
    sums[sg.slot0 + (tid - mbase)] = acc;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  nhc_update(state, p, sums, phase);
}

__global__ __launch_bounds__(1024) void k_nhc(NhcState* state, NhcParams p, RowsDesc rd, int phase, int ke_only) {
  nhc_body(state, p, rd, rd.nseg, phase, ke_only, (int)blockIdx.x);
}

void launch_nhc_rows(NhcState* st_dev, const NhcParams& p, const RowsDesc& rd, int phase, hipStream_t st,
                     int ke_only) {
  hipLaunchKernelGGL(k_nhc, dim3(1), dim3(1024), 0, st, st_dev, p, rd, phase, ke_only);
}

// ------------------------------------------------------------------ packs
// The same bodies as the kernels above, one launch over every system of a
// pack; each block takes its system's arguments from the pack table.

__global__ __launch_bounds__(1024) void kp_nhc(PACK_PARAMS, int phase, int reduce, int ke_only) {
  PACK_ENTRY(e, bid)
  nhc_body(e.nhc, e.np, e.rd, reduce ? e.rd.nseg : 0, phase, ke_only, bid);
}

__global__ __launch_bounds__(kBlock) void kp_first_half(PACK_PARAMS, int use_scale) {
  PACK_ENTRY(e, bid)
  first_half_body(e.v, e.iv, use_scale ? &e.nhc->factor : nullptr, 1, bid);
}

// what 0: zero the SHAKE iteration counter; 1: copy the status word to out[k]
__global__ __launch_bounds__(kWave) void kp_one(PACK_PARAMS, int what, int* out) {
  PACK_ENTRY(e, bid)
  if (threadIdx.x != 0) return;
  if (what == 0) {
    if (e.shake_iter) *e.shake_iter = 0;
  } else {
    out[PACK_INDEX()] = *e.v.status;
  }
}

__global__ __launch_bounds__(kBlock) void kp_shake(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  shake_body(e.v, e.iv, e.rows_shake, e.shake_iter, bid);
}

template <bool THERMO>
__global__ __launch_bounds__(kBlock) void kp_bonded(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  bonded_body<THERMO>(e.v, e.iv, e.rows_bond, bid);
}

__global__ __launch_bounds__(kBlock) void kp_second_half(PACK_PARAMS, int kick, int has_k, int has_b) {
  PACK_ENTRY(e, bid)
  second_half_body(e.v, e.iv, kick, has_k, has_b, e.rows_ke, bid);
}

__global__ __launch_bounds__(kBlock) void kp_scale_v(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  scale_v_body(e.v, &e.nhc->factor_end, bid);
}

__global__ __launch_bounds__(kBlock) void kp_rattle(PACK_PARAMS) {
  PACK_ENTRY(e, bid)
  rattle_body(e.v, e.iv, bid);
}

void launch_pack_integrate(int kernel, const PackLaunch& L, int nblocks, const PackSig& g, int arg, size_t lds,
                           int* out, hipStream_t st) {
  if (nblocks <= 0) return;
  const dim3 grid(nblocks);
  switch (kernel) {
    case PK_NHC:
      hipLaunchKernelGGL(kp_nhc, grid, dim3(1024), 0, st, PACK_ARGS(L), arg & 7, (arg >> 3) & 1, g.thermo ? 0 : 1);
      break;
    case PK_FIRST_HALF:
      hipLaunchKernelGGL(kp_first_half, grid, dim3(kBlock), lds, st, PACK_ARGS(L), arg);
      break;
    case PK_ONE:
      hipLaunchKernelGGL(kp_one, grid, dim3(kWave), 0, st, PACK_ARGS(L), arg, out);
      break;
    case PK_SHAKE:
      hipLaunchKernelGGL(kp_shake, grid, dim3(kBlock), 0, st, PACK_ARGS(L));
      break;
    case PK_BONDED:
      if (g.thermo)
        hipLaunchKernelGGL(kp_bonded<true>, grid, dim3(kBlock), 0, st, PACK_ARGS(L));
      else
        hipLaunchKernelGGL(kp_bonded<false>, grid, dim3(kBlock), 0, st, PACK_ARGS(L));
      break;
    case PK_SECOND:
      hipLaunchKernelGGL(kp_second_half, grid, dim3(kBlock), 0, st, PACK_ARGS(L), arg, g.has_k, g.has_b);
      break;
    case PK_ATOMS:
      hipLaunchKernelGGL(kp_scale_v, grid, dim3(kBlock), 0, st, PACK_ARGS(L));
      break;
    case PK_RATTLE:
      hipLaunchKernelGGL(kp_rattle, grid, dim3(kBlock), 0, st, PACK_ARGS(L));
      break;
    default:
      break;
  }
}

}  // namespace rbmd
