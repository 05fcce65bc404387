// DSSF by direct spatial sums, for tabulated graphs (get_qct_SqData inside
// get_dynSpectrum, CF/subrouts.cpp:11906-11912 and 8460-8525) and for any
// list of wavevectors (a q-path).
//
// The time contraction Y(r, w) = sum_t w(t) S(r, t) e^{-2 pi i n t / T} is
// already in the engine's Y buffers; this file applies
//   X_s(q, w) = sum_r e^{-i q.r} Y_s(r, w)
// as a complex GEMM per (sample, component): M = q, N = frequency bins,
// K = sites, and sums |X_s|^2 / (N_sites T) over the samples of one lattice.
// The phase table P[q][r] = e^{-i q.r} is built once per call in FP64 (q.r
// can reach 10^3 rad, beyond FP32 sincos accuracy) and stored at the engine
// precision; vacant sites get P = 0.
//
// Tiling: 256 threads as 16 x 16, each thread a 4 (q) x 4 (bin) block of
// complex accumulators; K staged through LDS 16 sites at a time.  A
// workgroup loops over one group of samples; when the q list is short (a
// q-path) the samples are split into groups so the launch still fills the
// chip, each group writing its own partial sum, which a second kernel adds
// in a fixed order (no atomics, deterministic).
#include <math.h>

#include <algorithm>

#include "smd_kernel.h"

namespace hsmd {

constexpr int kDQ = 64, kDB = 64, kDK = 16;

template <typename T>
__global__ void phase_table_kernel(int nq, const double *q, int N, const double *pos,
                                   const uint8_t *mask, T *P) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)nq * N) return;
  const int iq = (int)(idx / N), r = (int)(idx - (size_t)iq * N);
  double s = 0.0, c = 0.0;
  if (!mask[r]) sincos(q[2 * iq] * pos[2 * r] + q[2 * iq + 1] * pos[2 * r + 1], &s, &c);
  P[2 * idx] = (T)c;
  P[2 * idx + 1] = (T)(-s);
}

template <typename T>
__global__ __launch_bounds__(256) void dssf_direct_kernel(const SampleDev *samples, const int *ids,
                                                          int n_samples, int per_group, int N,
                                                          int nq, int nop, const T *P, const T *Y,
                                                          double scale, double *part) {
  __shared__ T Ps[kDK][kDQ][2];
  __shared__ T Ys[kDK][kDB][2];
  const int q0 = blockIdx.x * kDQ, b0 = blockIdx.y * kDB;
  const int d = blockIdx.z % 3, grp = blockIdx.z / 3;
  const int s_begin = grp * per_group, s_end = min(n_samples, s_begin + per_group);
  double *dssf = part + (size_t)grp * 3 * nq * nop;
  const int tq = threadIdx.x >> 4, tb = threadIdx.x & 15;
  const int ldy = 2 * nop;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
  for (int sl = s_begin; sl < s_end; sl++) {
    const SampleDev sd = samples[ids[sl]];
    const T *Yb = Y + sd.y_off + (size_t)d * N * ldy;
    T xr[4][4], xi[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) { xr[i][j] = T(0); xi[i][j] = T(0); }
    for (int k0 = 0; k0 < N; k0 += kDK) {
      // P tile: kDQ q x kDK sites (complex), 8 values per thread
      for (int e = threadIdx.x; e < kDQ * kDK; e += 256) {
        const int qq = e / kDK, kk = e - qq * kDK;
        const int q = q0 + qq, r = k0 + kk;
        T pr = T(0), pi = T(0);
        if (q < nq && r < N) {
          pr = P[((size_t)q * N + r) * 2];
          pi = P[((size_t)q * N + r) * 2 + 1];
        }
        Ps[kk][qq][0] = pr;
        Ps[kk][qq][1] = pi;
      }
      // Y tile: kDK sites x kDB bins (re, im)
      for (int e = threadIdx.x; e < kDK * kDB; e += 256) {
        const int kk = e / kDB, bb = e - kk * kDB;
        const int r = k0 + kk, n = b0 + bb;
        T yr = T(0), yi = T(0);
        if (r < N && n < nop) {
          yr = Yb[(size_t)r * ldy + n];
          yi = Yb[(size_t)r * ldy + nop + n];
        }
        Ys[kk][bb][0] = yr;
        Ys[kk][bb][1] = yi;
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < kDK; kk++) {
        T pr[4], pi[4], yr[4], yi[4];
#pragma unroll
        for (int i = 0; i < 4; i++) { pr[i] = Ps[kk][tq * 4 + i][0]; pi[i] = Ps[kk][tq * 4 + i][1]; }
#pragma unroll
        for (int j = 0; j < 4; j++) { yr[j] = Ys[kk][tb * 4 + j][0]; yi[j] = Ys[kk][tb * 4 + j][1]; }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int j = 0; j < 4; j++) {
            xr[i][j] += pr[i] * yr[j] - pi[i] * yi[j];
            xi[i][j] += pr[i] * yi[j] + pi[i] * yr[j];
          }
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        acc[i][j] += (double)xr[i][j] * (double)xr[i][j] + (double)xi[i][j] * (double)xi[i][j];
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int q = q0 + tq * 4 + i;
    if (q >= nq) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int n = b0 + tb * 4 + j;
      if (n < nop) dssf[((size_t)d * nq + q) * nop + n] = scale * acc[i][j];
    }
  }
}

// out[i] += sum over groups g of part[g][i], groups in order
__global__ void dssf_group_sum_kernel(int groups, size_t n, const double *part, double *out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int g = 0; g < groups; g++) s += part[(size_t)g * n + i];
  out[i] += s;
}

// part: scratch of groups x 3 x nq x nop doubles (groups from
// dssf_direct_groups); dssf accumulates the sample sums.
int dssf_direct_groups(int n_samples, int nq, int nop) {
  const int tiles = ((nq + kDQ - 1) / kDQ) * ((nop + kDB - 1) / kDB) * 3;
  const int want = (2048 + tiles - 1) / tiles;
  return std::max(1, std::min(n_samples, std::min(want, 64)));
}

template <typename T>
void dssf_direct_launch(const SampleDev *samples, const int *ids, int n_samples, int N, int nq,
                        int nop, const double *q, const double *pos, const uint8_t *mask, T *P,
                        const T *Y, double scale, int groups, double *part, double *dssf) {
  const size_t np = (size_t)nq * N;
  hipLaunchKernelGGL(phase_table_kernel<T>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, 0, nq,
                     q, N, pos, mask, P);
  const int per_group = (n_samples + groups - 1) / groups;
  const int used = (n_samples + per_group - 1) / per_group;
  dim3 grid((nq + kDQ - 1) / kDQ, (nop + kDB - 1) / kDB, 3 * used);
  hipLaunchKernelGGL(dssf_direct_kernel<T>, grid, dim3(256), 0, 0, samples, ids, n_samples,
                     per_group, N, nq, nop, P, Y, scale, part);
  const size_t n = (size_t)3 * nq * nop;
  hipLaunchKernelGGL(dssf_group_sum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, used,
                     n, (const double *)part, dssf);
  HIP_CHECK(hipGetLastError());
}

template void dssf_direct_launch<float>(const SampleDev *, const int *, int, int, int, int,
                                        const double *, const double *, const uint8_t *, float *,
                                        const float *, double, int, double *, double *);
template void dssf_direct_launch<double>(const SampleDev *, const int *, int, int, int, int,
                                         const double *, const double *, const uint8_t *, double *,
                                         const double *, double, int, double *, double *);

}  // namespace hsmd
