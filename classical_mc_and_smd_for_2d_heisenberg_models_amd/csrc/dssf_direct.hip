// DSSF by direct spatial sums, for tabulated graphs (get_qct_SqData inside
// get_dynSpectrum, CF/subrouts.cpp:11906-11912 and 8460-8525) and for any
// list of wavevectors (a q-path): the FP64 engine's VALU kernel, and the FP32
// engine's split-FP16 MFMA form (second half of the file).
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

// ---- FP32 q-path on the matrix cores -------------------------------------
// For one sample s and component d, with Y_sd[r][n'] the time-contracted
// trajectory (n' < nop: Re, n' >= nop: Im) and the phases split as
// B[j][r] = Re e^{-i q_j.r} (j < QP) and Im e^{-i q_{j-QP}.r} (j >= QP),
//   C_sd[n'][j] = sum_r Y_sd[r][n'] B[j][r]
// is the time-contraction GEMM shape with the sites as the reduction index:
// Y_sd is the FP32 operand split in-kernel, B the pre-split one, so the
// 128 x 512 MFMA tile of the time contraction computes every (s, d) block in
// one launch.  Then X(q, n) = C[n][q] - C[nop+n][QP+q] + i (C[nop+n][q] +
// C[n][QP+q]) and the DSSF sums |X|^2 over the lattice's samples.

// Phase table, FP16 hi/lo, chunk-major [ldk / 8][2 QP][8] over sites k (the
// layout of the time contraction's twiddles); vacant sites, pad sites and
// pad wavevectors are zero.  Phases in FP64: q.r reaches 10^3 rad.
__global__ void phase_split_kernel(int nq, int QP, const double *q, int N, int ldk,
                                   const double *pos, const uint8_t *mask, _Float16 *Bh,
                                   _Float16 *Bl) {
  const int J = 2 * QP;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)ldk * J) return;
  const int k8 = (int)(e & 7);
  const size_t jc = e >> 3;
  const int j = (int)(jc % J), k = (int)(jc / J) * 8 + k8;
  const int iq = j < QP ? j : j - QP;
  float v = 0.f;
  if (iq < nq && k < N && !mask[k]) {
    double sn, cs;
    sincos(q[2 * iq] * pos[2 * k] + q[2 * iq + 1] * pos[2 * k + 1], &sn, &cs);
    v = (float)(j < QP ? cs : -sn);
  }
  const _Float16 h = (_Float16)v;
  Bh[e] = h;
  Bl[e] = (_Float16)(v - (float)h);
}

// D[d][q][n] += scale sum over the G samples of |X_sd(q, n)|^2, samples in
// order (FP64 sums); C: G x 3 blocks of [2 nop][2 QP] FP32
__global__ void qpath_power_kernel(int G, int nq, int QP, int nop, const float *C, double scale,
                                   double *D) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)3 * nop * nq) return;
  const int iq = (int)(e % nq);
  const size_t t = e / nq;
  const int n = (int)(t % nop), d = (int)(t / nop);
  const size_t J = 2 * (size_t)QP, blk = (size_t)2 * nop * J;
  double acc = 0.0;
  for (int g = 0; g < G; g++) {
    const float *Cb = C + (size_t)(3 * g + d) * blk;
    const float xr = Cb[(size_t)n * J + iq] - Cb[(size_t)(nop + n) * J + QP + iq];
    const float xi = Cb[(size_t)(nop + n) * J + iq] + Cb[(size_t)n * J + QP + iq];
    acc += (double)xr * (double)xr + (double)xi * (double)xi;
  }
  D[((size_t)d * nq + iq) * nop + n] += scale * acc;
}

void launch_time_gemm_mfma_w(const void *samples_dev, int S, int maxM, const float *A,
                             const _Float16 *B8h, const _Float16 *B8l, int k0, int K, int ldb,
                             float *Y, hipStream_t st);

// qpath_phase_launch fills Bh, Bl (2 QP x ldk halves each, ldk = N rounded up
// to 32) for nq <= QP wavevectors; qpath_mfma_launch then runs G samples:
// samples_dev holds G x 3 GEMM blocks (M = 2 nop rows of Y_sd, row stride
// 2 nop, outputs [2 nop][2 QP] consecutive in C) and D accumulates the powers.
void qpath_phase_launch(int nq, int QP, const double *q, int N, const double *pos,
                        const uint8_t *mask, _Float16 *Bh, _Float16 *Bl) {
  const int ldk = ((N + 31) / 32) * 32;
  const size_t nb = (size_t)ldk * 2 * QP;
  hipLaunchKernelGGL(phase_split_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, 0, nq, QP,
                     q, N, ldk, pos, mask, Bh, Bl);
  HIP_CHECK(hipGetLastError());
}

void qpath_mfma_launch(int nq, int QP, int N, const _Float16 *Bh, const _Float16 *Bl,
                       const void *samples_dev, int G, int nop, const float *Y, float *C,
                       double scale, double *D) {
  launch_time_gemm_mfma_w(samples_dev, 3 * G, 2 * nop, Y, Bh, Bl, 0, N, 2 * QP, C, 0);
  const size_t n = (size_t)3 * nop * nq;
  hipLaunchKernelGGL(qpath_power_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, G, nq,
                     QP, nop, (const float *)C, scale, D);
  HIP_CHECK(hipGetLastError());
}

template void dssf_direct_launch<double>(const SampleDev *, const int *, int, int, int, int,
                                         const double *, const double *, const uint8_t *, double *,
                                         const double *, double, int, double *, double *);

}  // namespace hsmd
