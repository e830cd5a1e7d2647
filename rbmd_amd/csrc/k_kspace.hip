// k_kspace.hip -- random-batch Ewald structure factor and k-space force.
//
// rho(k_l) = sum_i q_i (cos k_l.r_i, sin k_l.r_i)
//   ComputePnumberChargeStructureFactorWorklet + ReduceByKey (RunWorklet.cpp:2209-2291,
//   ExecutionMD.cpp:469-501).  The reference writes a P x N scratch and reduces it by
//   key; here each thread folds 4 atoms, the workgroup reduces with wavefront
//   shuffles (no atomics), and a second pass sums the per-workgroup partials in
//   a fixed order.
// F_i = (S/P) sum_l -(4 pi q_i / (V k_l^2)) (cos(k.r) Im rho - sin(k.r) Re rho) k_l, x qqr2e
//   ComputeNewRBEForceWorklet (RunWorklet.cpp:2086-2141), ComputeRBEForceSum_fix
//   (ExecForceFunction.h:748-762), ComputeKspaceVirial (:397-410), UnitRescale
//   (ExecutionNVT.cpp:718).  With an Ewald prefactor e^{-k^2/4 alpha} and scale 1
//   the same kernel is ComputeNewFarElectrostaticsWorklet (RunWorklet.cpp:2492-2540).
#include "dev.hpp"
#include "reduce.hpp"

namespace rbmd {

constexpr int kRhoAtomsPerThread = 4;
constexpr int kRhoK = 8;  // k-vectors per workgroup in the rho pass

__global__ __launch_bounds__(kBlock) void k_rho_partial(SysView s, KView kv) {
  __shared__ double lds[kBlock / kWave][2 * kRhoK];
  const int l0 = blockIdx.y * kRhoK;
  double re[kRhoK], im[kRhoK];
#pragma unroll
  for (int t = 0; t < kRhoK; ++t) re[t] = im[t] = 0.0;
  double kx[kRhoK], ky[kRhoK], kz[kRhoK];
#pragma unroll
  for (int t = 0; t < kRhoK; ++t) {
    int l = l0 + t;
    bool ok = l < kv.P;
    kx[t] = ok ? kv.kvec[3 * l] : 0.0;
    ky[t] = ok ? kv.kvec[3 * l + 1] : 0.0;
    kz[t] = ok ? kv.kvec[3 * l + 2] : 0.0;
  }
  const int base = blockIdx.x * kBlock * kRhoAtomsPerThread;
#pragma unroll
  for (int a = 0; a < kRhoAtomsPerThread; ++a) {
    int i = base + a * kBlock + threadIdx.x;
    if (i < s.n) {
      double x = s.x[i], y = s.y[i], z = s.z[i], q = s.q[i];
#pragma unroll
      for (int t = 0; t < kRhoK; ++t) {
        double ph = kx[t] * x + ky[t] * y + kz[t] * z;
        double sn, cs;
        sincos(ph, &sn, &cs);
        re[t] += q * cs;
        im[t] += q * sn;
      }
    }
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
#pragma unroll
  for (int t = 0; t < kRhoK; ++t) {
    double a = wave_sum(re[t]);
    double b = wave_sum(im[t]);
    if (lane == 0) {
      lds[wave][2 * t] = a;
      lds[wave][2 * t + 1] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * kRhoK) {
    int t = threadIdx.x >> 1;
    int l = l0 + t;
    if (l < kv.P) {
      double acc = 0.0;
      for (int w = 0; w < kBlock / kWave; ++w) acc += lds[w][threadIdx.x];
      kv.rho_part[(size_t)blockIdx.x * 2 * kv.P + 2 * l + (threadIdx.x & 1)] = acc;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_rho_reduce(KView kv) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * kv.P) return;
  double acc = 0.0;
  for (int b = 0; b < kv.nblk_rho; ++b) acc += kv.rho_part[(size_t)b * 2 * kv.P + t];
  kv.rho[t] = acc;
}

void launch_rho_partial(const SysView& s, const KView& k, hipStream_t st) {
  dim3 grid(k.nblk_rho, (k.P + kRhoK - 1) / kRhoK);
  hipLaunchKernelGGL(k_rho_partial, grid, dim3(kBlock), 0, st, s, k);
}

void launch_rho_reduce(const SysView& s, const KView& k, hipStream_t st) {
  hipLaunchKernelGGL(k_rho_reduce, dim3((2 * k.P + kBlock - 1) / kBlock), dim3(kBlock), 0, st, k);
}

constexpr int kForceTile = 256;

// prefactor pre_l = -4 pi / (V k^2) [* e^{-k^2/4alpha} for Ewald]
template <bool THERMO, bool PERATOM>
__global__ __launch_bounds__(kBlock) void k_kforce(SysView s, KView kv, int ewald, double alpha, double* rows) {
  __shared__ double tk[kForceTile][4];
  __shared__ double tr[kForceTile][2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < s.n;
  double x = 0, y = 0, z = 0, q = 0;
  if (act) {
    x = s.x[i];
    y = s.y[i];
    z = s.z[i];
    q = s.q[i];
  }
  double fx = 0, fy = 0, fz = 0;
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int base = 0; base < kv.P; base += kForceTile) {
    __syncthreads();
    for (int t = threadIdx.x; t < kForceTile; t += blockDim.x) {
      int l = base + t;
      if (l < kv.P) {
        double kx = kv.kvec[3 * l], ky = kv.kvec[3 * l + 1], kz = kv.kvec[3 * l + 2];
        double k2 = kx * kx + ky * ky + kz * kz;
        tk[t][0] = kx;
        tk[t][1] = ky;
        tk[t][2] = kz;
        tk[t][3] = ewald ? exp(-k2 / (4 * alpha)) / (s.volume * k2) : 1.0 / (s.volume * k2);
        tr[t][0] = kv.rho[2 * l];
        tr[t][1] = kv.rho[2 * l + 1];
      }
    }
    __syncthreads();
    int cnt = kv.P - base < kForceTile ? kv.P - base : kForceTile;
    for (int t = 0; t < cnt; ++t) {
      double kx = tk[t][0], ky = tk[t][1], kz = tk[t][2];
      double ph = kx * x + ky * y + kz * z;
      double sn, cs;
      sincos(ph, &sn, &cs);
      double fs = (-4.0 * kPi * q) * tk[t][3] * (cs * tr[t][1] - sn * tr[t][0]);
      fx += fs * kx;
      fy += fs * ky;
      fz += fs * kz;
      if (THERMO || PERATOM) {
        v[0] += (x * kx + x * kx) * fs;
        v[1] += (y * ky + y * ky) * fs;
        v[2] += (z * kz + z * kz) * fs;
        v[3] += (x * ky + y * kx) * fs;
        v[4] += (x * kz + z * kx) * fs;
        v[5] += (y * kz + z * ky) * fs;
      }
    }
  }
  const double sc = kv.S_over_P;
  if (act) {
    s.fkx[i] = s.qqr2e * (fx * sc);
    s.fky[i] = s.qqr2e * (fy * sc);
    s.fkz[i] = s.qqr2e * (fz * sc);
    if (PERATOM)
      for (int k = 0; k < 6; ++k) s.vir_k[(size_t)6 * i + k] = v[k] * sc;
  }
  if (THERMO) {
    const double vals[6] = {v[0] * sc, v[1] * sc, v[2] * sc, v[3] * sc, v[4] * sc, v[5] * sc};
    const int slots[6] = {S_VIR_K, S_VIR_K + 1, S_VIR_K + 2, S_VIR_K + 3, S_VIR_K + 4, S_VIR_K + 5};
    block_write_row<6>(vals, slots, rows + (size_t)blockIdx.x * S_NSLOT);
  }
}

void launch_kforce(const SysView& s, const KView& k, hipStream_t st, int ewald, double alpha, int thermo, int peratom,
                   double* rows, int* nblocks) {
  int grid = (s.n + kBlock - 1) / kBlock;
  if (nblocks) *nblocks = grid;
  if (peratom)
    hipLaunchKernelGGL((k_kforce<true, true>), dim3(grid), dim3(kBlock), 0, st, s, k, ewald, alpha, rows);
  else if (thermo)
    hipLaunchKernelGGL((k_kforce<true, false>), dim3(grid), dim3(kBlock), 0, st, s, k, ewald, alpha, rows);
  else
    hipLaunchKernelGGL((k_kforce<false, false>), dim3(grid), dim3(kBlock), 0, st, s, k, ewald, alpha, rows);
}

}  // namespace rbmd
