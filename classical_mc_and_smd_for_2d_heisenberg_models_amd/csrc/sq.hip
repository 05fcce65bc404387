// Static structure factor kernels.
//
// Grid form, get_SqData_iMKL (CF/subrouts.cpp:8330-8443): spins scattered on
// the Gsz x Gsz DFT grid through gridMap (empty points 0), the unscaled
// forward 2D DFT, times 1/sqrt(Nq); output index k = i * Gsz + j.  Here the
// transform is separable: a row pass and a column pass, each one workgroup
// per (configuration, component, group of lines) with the lines and a
// cos/sin table indexed by (j * q) mod Gsz staged in LDS.
//
// Direct form, get_qct_SqData (CF/subrouts.cpp:8460-8525): for q = dq (nx, ny),
// nx, ny in [0, Lsz/2) then [-Lsz/2, 0), S(q) = 1/sqrt(N) sum_k S_k
// exp(-i dq (nx x_k + ny y_k)) at index ny_index * Lsz + nx_index; one thread
// per (q, configuration), sites staged through LDS in tiles.
//
// Measurements (CF/mc_hm_code.cpp:5203-5233): |S| and arg S summed per q into
// SqField; spec = max(1e-16, |S|^2) (get_StaticSFac, CF/subrouts.cpp:8799),
// reordered to xy = (x+y)/2, yz = (y+z)/2, zz, tt = (x+y+z)/3
// (get_OrderedSpecArray2D, CF/subrouts.cpp:8868-8930) and read at the five
// key points of get_corrLen (CF/subrouts.cpp:9335-9482): the peak SQ1 and the
// correlation length sqrt(SQ1/SQ2 - 1)/dq (0 when SQ1 <= SQ2).
#include <math.h>

#include <algorithm>

#include "sq.h"

namespace hsmd {

constexpr int kSqThreads = 256;
constexpr int kSqMaxG = 1024;

template <SpinType TY>
__device__ __forceinline__ double spin_comp(const void *spins, int comps, int Npad, int c, int d,
                                            int p) {
  if (TY == SPIN_F32) return (double)((const float *)spins)[((size_t)c * comps + d) * Npad + p];
  if (TY == SPIN_F64) return ((const double *)spins)[((size_t)c * comps + d) * Npad + p];
  return d == 0 ? (double)((const int8_t *)spins)[(size_t)c * Npad + p] : 0.0;
}

// Both passes give each workgroup kSqLines lines (grid rows, then columns of
// the row transform) and one thread per output frequency q; the thread keeps
// the kSqLines sums in registers, so every twiddle read (from an LDS copy of
// the table) serves kSqLines lines.  Per output the sum runs over j (i) in
// ascending order, as a one-line-per-workgroup loop would.
constexpr int kSqLines = 8;

// rows[c][d][i][qj] = sum_j S(i, j) e^{-2 pi i j qj / G}
template <SpinType TY>
__global__ __launch_bounds__(kSqThreads) void sq_rows_kernel(const void *spins, int comps, int Npad,
                                                             const int32_t *gmap, int G,
                                                             const double *tw, double2 *rows) {
  extern __shared__ double dsm[];
  double *line = dsm;                    // [kSqLines][G]
  double *tws = dsm + kSqLines * G;      // cos [G], sin [G]
  const int i0 = blockIdx.x * kSqLines, d = blockIdx.y, c = blockIdx.z;
  for (int e = threadIdx.x; e < kSqLines * G; e += blockDim.x) {
    const int r = e / G, j = e - r * G, i = i0 + r;
    double v = 0.0;
    if (i < G) {
      const int p = gmap[i * G + j];
      if (p >= 0) v = spin_comp<TY>(spins, comps, Npad, c, d, p);
    }
    line[e] = v;
  }
  for (int m = threadIdx.x; m < 2 * G; m += blockDim.x) tws[m] = tw[m];
  __syncthreads();
  for (int q = threadIdx.x; q < G; q += blockDim.x) {
    double re[kSqLines], im[kSqLines];
#pragma unroll
    for (int r = 0; r < kSqLines; r++) { re[r] = 0.0; im[r] = 0.0; }
    int m = 0;
    for (int j = 0; j < G; j++) {
      const double wc = tws[m], ws = tws[G + m];
#pragma unroll
      for (int r = 0; r < kSqLines; r++) {
        re[r] += line[r * G + j] * wc;
        im[r] -= line[r * G + j] * ws;
      }
      m += q;
      if (m >= G) m -= G;
    }
#pragma unroll
    for (int r = 0; r < kSqLines; r++)
      if (i0 + r < G) rows[(((size_t)c * 3 + d) * G + i0 + r) * G + q] = make_double2(re[r], im[r]);
  }
}

// sq[c][d][qi * G + qj] = fc sum_i rows[c][d][i][qj] e^{-2 pi i i qi / G}
__global__ __launch_bounds__(kSqThreads) void sq_cols_kernel(const double2 *rows, int G,
                                                            const double *tw, double fc,
                                                            double2 *sq) {
  extern __shared__ double dsm[];
  double2 *line = (double2 *)dsm;              // [kSqLines][G]
  double *tws = dsm + 2 * kSqLines * G;        // cos [G], sin [G]
  const int q0 = blockIdx.x * kSqLines, d = blockIdx.y, c = blockIdx.z;
  const double2 *src = rows + ((size_t)c * 3 + d) * G * G;
  for (int e = threadIdx.x; e < kSqLines * G; e += blockDim.x) {
    const int i = e / kSqLines, r = e - i * kSqLines, qj = q0 + r;
    line[r * G + i] = qj < G ? src[(size_t)i * G + qj] : make_double2(0.0, 0.0);
  }
  for (int m = threadIdx.x; m < 2 * G; m += blockDim.x) tws[m] = tw[m];
  __syncthreads();
  double2 *dst = sq + ((size_t)c * 3 + d) * G * G;
  for (int q = threadIdx.x; q < G; q += blockDim.x) {
    double re[kSqLines], im[kSqLines];
#pragma unroll
    for (int r = 0; r < kSqLines; r++) { re[r] = 0.0; im[r] = 0.0; }
    int m = 0;
    for (int i = 0; i < G; i++) {
      const double wc = tws[m], ws = -tws[G + m];
#pragma unroll
      for (int r = 0; r < kSqLines; r++) {
        const double a = line[r * G + i].x, b = line[r * G + i].y;
        re[r] += a * wc - b * ws;
        im[r] += a * ws + b * wc;
      }
      m += q;
      if (m >= G) m -= G;
    }
#pragma unroll
    for (int r = 0; r < kSqLines; r++)
      if (q0 + r < G) dst[(size_t)q * G + q0 + r] = make_double2(fc * re[r], fc * im[r]);
  }
}

template <SpinType TY>
__global__ __launch_bounds__(kSqThreads) void sq_direct_kernel(const void *spins, int comps,
                                                              int Npad, int N, const double *pos,
                                                              int Lq, int Nq, double dq,
                                                              double fc, double2 *sq) {
  __shared__ double tx[kSqThreads], ty[kSqThreads], ts[3][kSqThreads];
  const int c = blockIdx.y;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int l2 = Lq / 2;
  const bool on = idx < Lq * Lq;
  const int j = on ? idx / Lq : 0, i = on ? idx - j * Lq : 0;
  const int nx = i < l2 ? i : i - Lq, ny = j < l2 ? j : j - Lq;
  double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
  for (int base = 0; base < N; base += kSqThreads) {
    const int p = base + threadIdx.x;
    if (p < N) {
      tx[threadIdx.x] = pos[2 * p];
      ty[threadIdx.x] = pos[2 * p + 1];
      for (int d = 0; d < 3; d++) ts[d][threadIdx.x] = spin_comp<TY>(spins, comps, Npad, c, d, p);
    }
    __syncthreads();
    const int n = min(kSqThreads, N - base);
    if (on) {
      for (int k = 0; k < n; k++) {
        double s, co;
        sincos(-dq * (nx * tx[k] + ny * ty[k]), &s, &co);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          re[d] += ts[d][k] * co;
          im[d] += ts[d][k] * s;
        }
      }
    }
    __syncthreads();
  }
  if (idx < Nq) {
    for (int d = 0; d < 3; d++)
      sq[((size_t)c * 3 + d) * Nq + idx] = on ? make_double2(fc * re[d], fc * im[d]) : make_double2(0.0, 0.0);
  }
}

__global__ __launch_bounds__(kSqThreads) void sq_field_kernel(const double2 *sq, int Nq,
                                                             const int32_t *config_of_slot,
                                                             double *sqfield) {
  const int k = blockIdx.y;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= Nq) return;
  const int c = config_of_slot[k];
  double *f = sqfield + ((size_t)k * Nq + q) * 6;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double2 z = sq[((size_t)c * 3 + d) * Nq + q];
    f[d] += hypot(z.x, z.y);
    f[3 + d] += atan2(z.y, z.x);
  }
}

// get_OrderedSpecArray2D's index map: ordered row i reads spectrum row mvec[i].
__device__ __forceinline__ int ordered_to_fft(int i, int G) {
  const int l2 = G / 2;
  const int n = 2 * (i / l2);
  return i - l2 * (n - 1);
}

__global__ void sq_keypoints_kernel(const double2 *sq, int G, int Nq, int c3, double dq, int R,
                                    const int32_t *config_of_slot, double *moments, int nmom,
                                    int mom_base) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= R) return;
  const int c = config_of_slot[k];
  const int l2 = G / 2, l3 = G / 3, l4 = G / 4;
  int i1[5], j1[5], i2[5], j2[5];
  if (c3) {
    // K (hexagon corner), M1, M2, Gamma, quadrant centre
    i1[0] = l3 + l2; j1[0] = -l3 + l2; i2[0] = l3 + l2; j2[0] = -l3 + l2 + 1;
    i1[1] = l2; j1[1] = 0; i2[1] = l2; j2[1] = 1;
    i1[2] = 0; j1[2] = l2; i2[2] = 1; j2[2] = l2;
  } else {
    // M (corner), X, Y
    i1[0] = 0; j1[0] = 0; i2[0] = 0; j2[0] = 1;
    i1[1] = 0; j1[1] = l2; i2[1] = 1; j2[1] = l2;
    i1[2] = l2; j1[2] = 0; i2[2] = l2; j2[2] = 1;
  }
  i1[3] = l2; j1[3] = l2; i2[3] = l2; j2[3] = l2 + 1;
  i1[4] = l4; j1[4] = l4; i2[4] = l4; j2[4] = l4 + 1;
  const double2 *s = sq + (size_t)c * 3 * Nq;
  auto spec = [&](int oi, int oj, double (&v)[4]) {
    const int q = ordered_to_fft(oi, G) * G + ordered_to_fft(oj, G);
    double p[3];
    for (int d = 0; d < 3; d++) {
      const double2 z = s[(size_t)d * Nq + q];
      p[d] = fmax(1.0e-16, z.x * z.x + z.y * z.y);
    }
    v[0] = 0.5 * (p[0] + p[1]);
    v[1] = 0.5 * (p[1] + p[2]);
    v[2] = p[2];
    v[3] = (1.0 / 3.0) * (p[0] + p[1] + p[2]);
  };
  double *m = moments + (size_t)k * nmom + mom_base;
  for (int pt = 0; pt < 5; pt++) {
    double a[4], b[4];
    spec(i1[pt], j1[pt], a);
    spec(i2[pt], j2[pt], b);
    for (int f = 0; f < 4; f++) {
      const double w = a[f] / b[f];
      m[f * 5 + pt] += a[f];
      m[20 + f * 5 + pt] += w > 1.0 ? (1.0 / dq) * sqrt(w - 1.0) : 0.0;
    }
  }
  m[40] += 1.0;
}

void sq_plan_init(SqPlan &p, const Lattice &L, int R) {
  p.direct = L.sq_direct();
  p.N = L.N;
  p.Npad = L.N + 1;
  p.G = L.Gsz;
  p.Lq = L.L;
  p.Nq = L.Nq;
  p.dq = L.sq_dq();
  p.c3 = L.c3sym;
  p.R = R;
  if (p.G < 2 || p.G % 2 != 0) throw Error(HSMD_ERR_ARG, "S(q) needs an even DFT grid side");
  if (p.direct) {
    if (L.rvec.size() != (size_t)2 * L.N)
      throw Error(HSMD_ERR_ARG, "S(q) direct sums need site positions");
    if (p.Lq * p.Lq > p.Nq) throw Error(HSMD_ERR_ARG, "S(q) q grid larger than the spectrum");
    p.fc = 1.0 / sqrt((double)L.N);
    std::vector<double> pos((size_t)2 * L.N);
    for (int q = 0; q < L.N; q++) {
      pos[2 * q] = L.rvec[2 * (size_t)L.perm[q]];
      pos[2 * q + 1] = L.rvec[2 * (size_t)L.perm[q] + 1];
    }
    p.pos.upload(pos.data(), pos.size());
  } else {
    if (p.G > kSqMaxG) throw Error(HSMD_ERR_ARG, "S(q) grid side above 1024");
    if (L.gridMap.size() != (size_t)p.G * p.G) throw Error(HSMD_ERR_ARG, "lattice has no DFT grid map");
    p.fc = 1.0 / sqrt((double)p.Nq);
    std::vector<int32_t> gm(L.gridMap.size());
    for (size_t n = 0; n < gm.size(); n++) gm[n] = L.gridMap[n] >= 0 ? L.inv[L.gridMap[n]] : -1;
    p.gmap.upload(gm.data(), gm.size());
    std::vector<double> tw((size_t)2 * p.G);
    const double pi2 = 2.0 * acos(-1.0);
    for (int m = 0; m < p.G; m++) {
      tw[m] = cos(pi2 * m / p.G);
      tw[p.G + m] = sin(pi2 * m / p.G);
    }
    p.tw.upload(tw.data(), tw.size());
    p.rows.alloc((size_t)R * 3 * p.G * p.G * 2);
    // the passes keep kSqLines lines (complex in the column pass) and the
    // twiddles in LDS
    const size_t cols_lds = sizeof(double) * (size_t)(2 * kSqLines + 2) * p.G;
    const size_t rows_lds = sizeof(double) * (size_t)(kSqLines + 2) * p.G;
    if (cols_lds > 65536)
      HIP_CHECK(hipFuncSetAttribute((const void *)sq_cols_kernel,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)cols_lds));
    if (rows_lds > 65536) {
      HIP_CHECK(hipFuncSetAttribute((const void *)sq_rows_kernel<SPIN_F32>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)rows_lds));
      HIP_CHECK(hipFuncSetAttribute((const void *)sq_rows_kernel<SPIN_F64>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)rows_lds));
      HIP_CHECK(hipFuncSetAttribute((const void *)sq_rows_kernel<SPIN_I8>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)rows_lds));
    }
  }
  p.sq.alloc((size_t)R * 3 * p.Nq * 2);
}

template <SpinType TY>
static void sq_compute_t(SqPlan &p, const void *spins, int R) {
  const int comps = TY == SPIN_I8 ? 1 : 3;
  if (p.direct) {
    dim3 grid((p.Nq + kSqThreads - 1) / kSqThreads, R);
    hipLaunchKernelGGL(sq_direct_kernel<TY>, grid, dim3(kSqThreads), 0, 0, spins, comps, p.Npad,
                       p.N, p.pos.p, p.Lq, p.Nq, p.dq, p.fc, (double2 *)p.sq.p);
  } else {
    // one thread per frequency (a whole number of waves), kSqLines lines per workgroup
    const int nt = std::min(kSqThreads, ((p.G + 63) / 64) * 64);
    const int nl = (p.G + kSqLines - 1) / kSqLines;
    const size_t rows_lds = sizeof(double) * (size_t)(kSqLines + 2) * p.G;
    const size_t cols_lds = sizeof(double) * (size_t)(2 * kSqLines + 2) * p.G;
    hipLaunchKernelGGL(sq_rows_kernel<TY>, dim3(nl, 3, R), dim3(nt), rows_lds, 0, spins, comps,
                       p.Npad, p.gmap.p, p.G, p.tw.p, (double2 *)p.rows.p);
    hipLaunchKernelGGL(sq_cols_kernel, dim3(nl, 3, R), dim3(nt), cols_lds, 0,
                       (const double2 *)p.rows.p, p.G, p.tw.p, p.fc, (double2 *)p.sq.p);
  }
  HIP_CHECK(hipGetLastError());
}

void sq_compute(SqPlan &p, const void *spins, SpinType type, int R) {
  if (R > p.R) throw Error(HSMD_ERR_ARG, "S(q) plan sized for fewer configurations");
  if (type == SPIN_F32) sq_compute_t<SPIN_F32>(p, spins, R);
  else if (type == SPIN_F64) sq_compute_t<SPIN_F64>(p, spins, R);
  else sq_compute_t<SPIN_I8>(p, spins, R);
}

void sq_accumulate(SqPlan &p, int R, const int32_t *config_of_slot, double *sqfield,
                   double *moments, int nmom, int mom_base) {
  dim3 grid((p.Nq + kSqThreads - 1) / kSqThreads, R);
  hipLaunchKernelGGL(sq_field_kernel, grid, dim3(kSqThreads), 0, 0, (const double2 *)p.sq.p, p.Nq,
                     config_of_slot, sqfield);
  hipLaunchKernelGGL(sq_keypoints_kernel, dim3((R + 63) / 64), dim3(64), 0, 0,
                     (const double2 *)p.sq.p, p.G, p.Nq, p.c3 ? 1 : 0, p.dq, R, config_of_slot,
                     moments, nmom, mom_base);
  HIP_CHECK(hipGetLastError());
}

void static_sq(const Lattice &L, int n_configs, const double *spins, double *out) {
  if (n_configs < 1) throw Error(HSMD_ERR_ARG, "n_configs must be >= 1");
  SqPlan p;
  sq_plan_init(p, L, n_configs);
  const int N = L.N, Npad = N + 1;
  std::vector<double> h((size_t)n_configs * 3 * Npad, 0.0);
  for (int c = 0; c < n_configs; c++)
    for (int q = 0; q < N; q++)
      for (int d = 0; d < 3; d++)
        h[((size_t)c * 3 + d) * Npad + q] = spins[((size_t)c * N + L.perm[q]) * 3 + d];
  DevBuf<double> dev;
  dev.upload(h.data(), h.size());
  sq_compute(p, dev.p, SPIN_F64, n_configs);
  HIP_CHECK(hipDeviceSynchronize());
  HIP_CHECK(hipMemcpy(out, p.sq.p, (size_t)n_configs * 3 * p.Nq * 2 * sizeof(double),
                      hipMemcpyDeviceToHost));
}

}  // namespace hsmd
