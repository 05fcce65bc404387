// SMD time-step kernel shared by the FP32 and FP64 translation units.
// See smd.hip for the algorithm and its reference.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "device_common.h"
#include "hsmd_internal.h"

namespace hsmd {

struct LatDev {
  Coup cp;
  Bonds bd;
  const int32_t *nb;
  const void *Jt;
  const int32_t *bi, *bj;
  const uint8_t *bc;
};

struct SampleDev {
  int li;
  int N;
  long long spin_off;   // into spins / S0 (3 * Npad per sample)
  long long traj_off;   // into trajectory chunk (Tc * 3 * N per sample)
  long long y_off;      // into Y accumulators (3 * N * 2 * n_omega per sample)
};

constexpr int kSmdThreads = 1024;

template <typename T>
__device__ __forceinline__ void cross3(T ax, T ay, T az, T bx, T by, T bz, T &cx, T &cy, T &cz) {
  cx = ay * bz - az * by;
  cy = az * bx - ax * bz;
  cz = ax * by - ay * bx;
}

template <typename T>
__device__ __forceinline__ double bond_energy(const LatDev &L, const T *sx, const T *sy, const T *sz) {
  double e = 0.0;
  for (int b = threadIdx.x; b < L.bd.Nb; b += blockDim.x) {
    const int i = L.bi[b], j = L.bj[b], k = L.bc[b];
    e += L.bd.jxy[k] * ((double)sx[i] * (double)sx[j] + (double)sy[i] * (double)sy[j]) +
         L.bd.jz[k] * ((double)sz[i] * (double)sz[j]);
  }
  return e;
}

template <typename T>
__device__ __forceinline__ double field_energy_term(const LatDev &L, const T *sz) {
  if (!L.cp.h_on) return 0.0;
  double m = 0.0;
  for (int p = threadIdx.x; p < L.cp.N; p += blockDim.x) m += (double)sz[p];
  return -L.cp.h * m;
}

struct StepArgs {
  const LatDev *lats;
  const SampleDev *samples;
  void *spins;          // per sample [3][Npad]
  const void *s0;       // per sample [3][Npad], t = 0 spins
  void *traj;           // per sample [Tc][3][N]
  double *qss;          // per sample [T]
  double *eref;         // per sample, energy at the current time
  void *stage;          // per sample [3][Npad] stage buffer (global-memory form)
  double dt;
  int t0;               // global time index of the first point of this launch
  int n_points;         // time points produced by this launch
  int chunk_base;       // time index stored at trajectory row 0
  int T;                // total time points
  int lds_npad;         // LDS stride (max Npad) when LDS is used
  int sample_base;      // first sample of this launch
};

// One launch produces n_points time points t0 .. t0+n_points-1 of every
// sample; t = 0 is the input configuration, every later point is one RK4
// step plus the energy correction.
template <typename T, int SPT, bool LDS, int ZC>
__global__ __launch_bounds__(kSmdThreads) void smd_kernel(const StepArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double scratch[2 * 16];
  const int sidx = blockIdx.x + a.sample_base;
  const SampleDev sd = a.samples[sidx];
  const LatDev &L = a.lats[sd.li];
  const int N = sd.N, Npad = N + 1;
  T *gS = (T *)a.spins + sd.spin_off;
  const T *g0 = (const T *)a.s0 + sd.spin_off;
  T *S, *B;
  int stride;
  if (LDS) {
    stride = a.lds_npad;
    S = (T *)smem;
    B = S + 3 * stride;
    for (int i = threadIdx.x; i < 3 * Npad; i += blockDim.x) {
      const int d = i / Npad, p = i - d * Npad;
      S[d * stride + p] = gS[i];
      B[d * stride + p] = T(0);
    }
  } else {
    // global-memory form: spins in place, stage buffer beside them
    stride = Npad;
    S = gS;
    B = (T *)a.stage + sd.spin_off;
  }
  __syncthreads();
  T *Sx = S, *Sy = S + stride, *Sz = S + 2 * stride;
  T *Bx = B, *By = B + stride, *Bz = B + 2 * stride;
  const T dt = T(a.dt);
  double Eref = a.eref[sidx];
  const T *Jt = (const T *)L.Jt;
  T *traj = (T *)a.traj + sd.traj_off;
  for (int it = 0; it < a.n_points; it++) {
    const int t = a.t0 + it;
    if (t > 0) {
      T K1[SPT][3], K2[SPT][3], K3[SPT][3];
      // stage 1: K1 from S, B = S + K1/2
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = threadIdx.x + u * kSmdThreads;
        if (p < N) {
          T hx, hy, hz;
          local_field<T, ZC>(L.cp, L.nb, Jt, p, Sx, Sy, Sz, hx, hy, hz);
          const T x = Sx[p], y = Sy[p], z = Sz[p];
          T px, py, pz;
          cross3(hx, hy, hz, x, y, z, px, py, pz);
          K1[u][0] = dt * px; K1[u][1] = dt * py; K1[u][2] = dt * pz;
          Bx[p] = x + T(0.5) * K1[u][0];
          By[p] = y + T(0.5) * K1[u][1];
          Bz[p] = z + T(0.5) * K1[u][2];
        }
      }
      __syncthreads();
      // stage 2
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = threadIdx.x + u * kSmdThreads;
        if (p < N) {
          T hx, hy, hz;
          local_field<T, ZC>(L.cp, L.nb, Jt, p, Bx, By, Bz, hx, hy, hz);
          T px, py, pz;
          cross3(hx, hy, hz, Bx[p], By[p], Bz[p], px, py, pz);
          K2[u][0] = dt * px; K2[u][1] = dt * py; K2[u][2] = dt * pz;
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = threadIdx.x + u * kSmdThreads;
        if (p < N) {
          Bx[p] = Sx[p] + T(0.5) * K2[u][0];
          By[p] = Sy[p] + T(0.5) * K2[u][1];
          Bz[p] = Sz[p] + T(0.5) * K2[u][2];
        }
      }
      __syncthreads();
      // stage 3
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = threadIdx.x + u * kSmdThreads;
        if (p < N) {
          T hx, hy, hz;
          local_field<T, ZC>(L.cp, L.nb, Jt, p, Bx, By, Bz, hx, hy, hz);
          T px, py, pz;
          cross3(hx, hy, hz, Bx[p], By[p], Bz[p], px, py, pz);
          K3[u][0] = dt * px; K3[u][1] = dt * py; K3[u][2] = dt * pz;
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = threadIdx.x + u * kSmdThreads;
        if (p < N) {
          Bx[p] = Sx[p] + K3[u][0];
          By[p] = Sy[p] + K3[u][1];
          Bz[p] = Sz[p] + K3[u][2];
        }
      }
      __syncthreads();
      // stage 4 and combine (S is not read by any thread in this stage)
      T Sn[SPT][3];
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = threadIdx.x + u * kSmdThreads;
        if (p < N) {
          T hx, hy, hz;
          local_field<T, ZC>(L.cp, L.nb, Jt, p, Bx, By, Bz, hx, hy, hz);
          T px, py, pz;
          cross3(hx, hy, hz, Bx[p], By[p], Bz[p], px, py, pz);
          const T f1 = T(1.0 / 6.0), f2 = T(2.0 / 6.0);
          T v[3];
          const T k4[3] = {dt * px, dt * py, dt * pz};
          const T s[3] = {Sx[p], Sy[p], Sz[p]};
#pragma unroll
          for (int d = 0; d < 3; d++)
            v[d] = s[d] + (f1 * (K1[u][d] + k4[d]) + f2 * (K2[u][d] + K3[u][d]));
          const T inrm = T(1) / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
          Sn[u][0] = v[0] * inrm; Sn[u][1] = v[1] * inrm; Sn[u][2] = v[2] * inrm;
        }
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = threadIdx.x + u * kSmdThreads;
        if (p < N) { Sx[p] = Sn[u][0]; Sy[p] = Sn[u][1]; Sz[p] = Sn[u][2]; }
      }
      __syncthreads();
      // energy correction: sum |h x S|^2 and E(S) in FP64
      double D[SPT][3];
      double v2[2] = {0.0, 0.0};
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = threadIdx.x + u * kSmdThreads;
        if (p < N) {
          T hx, hy, hz;
          local_field<T, ZC>(L.cp, L.nb, Jt, p, Sx, Sy, Sz, hx, hy, hz);
          const double x = Sx[p], y = Sy[p], z = Sz[p];
          double px, py, pz;
          cross3<double>(hx, hy, hz, x, y, z, px, py, pz);
          v2[0] += px * px + py * py + pz * pz;
          cross3<double>(x, y, z, px, py, pz, D[u][0], D[u][1], D[u][2]);
        }
      }
      v2[1] = bond_energy<T>(L, Sx, Sy, Sz) + field_energy_term<T>(L, Sz);
      block_sum<2>(v2, scratch);
      const double cfac = (Eref - v2[1]) / v2[0];
      double v3[2] = {0.0, 0.0};
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = threadIdx.x + u * kSmdThreads;
        if (p < N) {
          double x = (double)Sx[p] + cfac * D[u][0];
          double y = (double)Sy[p] + cfac * D[u][1];
          double z = (double)Sz[p] + cfac * D[u][2];
          const double nr = sqrt(x * x + y * y + z * z);
          x /= nr; y /= nr; z /= nr;
          v3[1] += (double)g0[p] * x + (double)g0[Npad + p] * y + (double)g0[2 * Npad + p] * z;
          Sx[p] = (T)x; Sy[p] = (T)y; Sz[p] = (T)z;
        }
      }
      __syncthreads();
      v3[0] = bond_energy<T>(L, Sx, Sy, Sz) + field_energy_term<T>(L, Sz);
      block_sum<2>(v3, scratch);
      Eref = v3[0];
      if (threadIdx.x == 0) a.qss[(size_t)sidx * a.T + t] = v3[1] / (double)N;
    } else {
      if (threadIdx.x == 0) a.qss[(size_t)sidx * a.T + 0] = 1.0;
    }
    // record the configuration at time t
    if (a.traj) {
      T *row = traj + (size_t)(t - a.chunk_base) * 3 * N;
      for (int p = threadIdx.x; p < N; p += blockDim.x) {
        row[p] = Sx[p]; row[N + p] = Sy[p]; row[2 * N + p] = Sz[p];
      }
    }
  }
  if (threadIdx.x == 0) a.eref[sidx] = Eref;
  if (LDS) {
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * Npad; i += blockDim.x) {
      const int d = i / Npad, p = i - d * Npad;
      gS[i] = S[d * stride + p];
    }
  }
}

// Energy of every sample (FP64 reduction), used to seed E_ref.
template <typename T>
__global__ __launch_bounds__(kSmdThreads) void smd_energy_kernel(const LatDev *lats,
                                                                 const SampleDev *samples,
                                                                 const T *spins, double *E) {
  __shared__ double scratch[16];
  const SampleDev sd = samples[blockIdx.x];
  const LatDev &L = lats[sd.li];
  const int Npad = sd.N + 1;
  const T *s = spins + sd.spin_off;
  double v[1] = {bond_energy<T>(L, s, s + Npad, s + 2 * Npad) + field_energy_term<T>(L, s + 2 * Npad)};
  block_sum<1>(v, scratch);
  if (threadIdx.x == 0) E[blockIdx.x] = v[0];
}

struct SmdLaunch {
  int S;
  int spt;
  int maxZ;
  bool use_lds;
  size_t lds_bytes;
};

void smd_launch_f32(const SmdLaunch &c, StepArgs a);
void smd_launch_f64(const SmdLaunch &c, StepArgs a);

// Launches the time-step kernel for element type T: all samples in one
// launch with LDS-resident lattices, or (global-memory form) one launch per
// sample.
template <typename T, int SPT, int ZC>
void smd_launch_t(const SmdLaunch &c, StepArgs a) {
  if (c.use_lds) {
    const void *fn = (const void *)smd_kernel<T, SPT, true, ZC>;
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.lds_bytes));
    a.sample_base = 0;
    hipLaunchKernelGGL((smd_kernel<T, SPT, true, ZC>), dim3(c.S), dim3(kSmdThreads), c.lds_bytes, 0, a);
  } else {
    for (int s = 0; s < c.S; s++) {
      a.sample_base = s;
      hipLaunchKernelGGL((smd_kernel<T, SPT, false, ZC>), dim3(1), dim3(kSmdThreads), 0, 0, a);
    }
  }
}

template <typename T, int SPT>
void smd_launch_zc(const SmdLaunch &c, const StepArgs &a) {
  const int Z = c.maxZ;
  if (Z <= 8) smd_launch_t<T, SPT, 8>(c, a);
  else if (Z <= 16) smd_launch_t<T, SPT, 16>(c, a);
  else smd_launch_t<T, SPT, 32>(c, a);
}

}  // namespace hsmd
