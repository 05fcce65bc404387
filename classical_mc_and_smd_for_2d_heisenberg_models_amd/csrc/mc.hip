// Monte Carlo sweep engine: heat-bath + overrelaxation (Heisenberg),
// single-spin-flip Metropolis (Ising), on-device replica exchange.
//
// Reference hot path (one MPI rank per temperature):
//   sweep loops            CF/mc_hm_code.cpp:2778-2962, 4836-4945
//   overrelaxation         reflect_aboutVec  CF/subrouts.cpp:4956-5004
//   heat bath              get_rotZ2VMat / get_hBathSample CF/subrouts.cpp:5095-5179
//   Metropolis             CF/mc_hm_code.cpp:2872-2898
//   energy                 get_energyValue   CF/subrouts.cpp:5181-5312
//   replica exchange       CF/mc_hm_code.cpp:3015-3130, 5253-5365
//   observables            CF/mc_hm_code.cpp:4951-5160, CF/subrouts.cpp:5362-5800
//
// Design: one workgroup per replica.  Sites are relabelled by colour class
// (no two sites of a class share an active coupling), so a class is updated
// in parallel and the classes are separated by workgroup barriers.  The
// replica's spins live in LDS as SoA x/y/z arrays plus a zero sentinel site
// that absorbs missing neighbours; the neighbour table (slot-major, shared by
// all replicas) streams from L2.  The sweep epilogue reduces the FP64 energy
// over the bond list and the order-parameter sums; a one-block kernel then
// accumulates the observables per temperature slot and performs the
// exchange by permuting slot labels (configurations never move).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <tuple>

#include "device_common.h"
#include "hsmd_internal.h"
#include "sq.h"

namespace hsmd {

constexpr int NOBS = 32;   // scratch width of the sweep kernels' reductions
constexpr int NOBSX = 40;  // per-configuration sums of obs_kernel: order
                           // parameters [0..31], stiffness C-sums [32..37]
constexpr int NMOM = HSMD_N_MOMENTS;
constexpr int kThreads = 1024;
// quasicrystal measurement: raw sums per configuration (qc_obs_kernel) and
// per-slot histograms (Neel 41, stripes 41 x 4, stripe vector 41 x 41)
constexpr int NQC = 56;
constexpr int kHst = 41;
constexpr int kQcHist = HSMD_QC_HIST;

struct SweepArgs {
  Coup cp;
  Bonds bd;
  int n_colours;
  int col_off[kCMAX + 1];
  int n_or, n_hb, n_or_b, n_hb_b;
  int metro_passes;        // Ising: Metropolis passes per sweep (1; 0 = energies only)
  uint64_t seed;
  uint32_t sweep;
  int config_base;
  int measure;
  int n_stag, has_labels, z2, c3;
  const int32_t *nb;
  const uint32_t *nb16;    // packed 16-bit neighbour pairs, (Z+1)/2 x N (fast path)
  const void *Jt;
  const int32_t *perm;
  const int32_t *bi, *bj;
  const uint8_t *bc;
  const int8_t *lab;       // 4 x N: stag, hs, vs, ds (new labels)
  const int32_t *z2nb;     // 3 x N new labels
  void *spins;
  const double *beta_slot;
  const int32_t *slot_of_config;
  const int32_t *boost_slot;
  double *energy;          // per configuration
  double *obs;             // per configuration, NOBSX
  const double *stiff_w;   // 6 x Nb: J C0^2, J C1^2, J C2^2, J C0, J C1, J C2
  const uint32_t *sweep_dev;   // graph replays: sweep index read here instead of `sweep`
  // quasicrystal measurement (qc != 0): reference states [5][3][N], hotspot
  // phases e^{-i q_i.r} [8][N], shell-1 slot count, sums out [NQC] per config
  int qc, qc_zn1;
  const double *qc_ref;
  const double2 *qc_phase;
  double *qc_obs;
  // fast kernels: energy from the neighbour table over the slots of emask,
  // weighted by escale (table_energy != 0), instead of the bond list
  int table_energy;
  uint32_t emask;
  double escale;
  // the same energy from a label table of the emask slots alone: nEw words
  // per site of packed 16-bit pairs (sentinel-padded), couplings ejxy / ejz
  // per emask slot in slot order (zero for the padding)
  const uint32_t *nbE;
  int nEw;
  double ejxy[16], ejz[16];
  // compact labels (sparse tables, fast kernels): each site's coupled
  // neighbours first, in slot order, as 14-bit labels with the 2-bit shell
  // (0..2) of their slot; wcnt[col * 16 + chunk] = label words that hold
  // neighbours of any site of 64-site chunk `chunk` of class col
  const uint8_t *wcnt;
  double shj[3], shjz[3];   // couplings per shell (FP32 copies below)
  float fshj[3], fshjz[3];
  // fast kernels, wave-level ordering instead of the class barrier (wdep !=
  // nullptr): wdep[col * 16 + w] = bit mask of the waves that own a
  // neighbour (any class, 64 sites of a class per wave) of a site that wave
  // w updates in class col; sync_err is set when a wait ran out
  const uint32_t *wdep;
  uint32_t *sync_err;
  // fast Ising kernel: the energy's integer bond counts from nbE when
  // itable != 0, icls[e] = bond class of emask slot e (slot order, -1 for
  // the padding)
  int itable;
  int icls[16];
};

__device__ __forceinline__ uint32_t sweep_of(const SweepArgs &a) {
  return a.sweep_dev ? *a.sweep_dev : a.sweep;
}

// ------------------------------------------------------------ site updates

// Overrelaxation: S' = 2 (S.h)/|h|^2 h - S; identity when |h|^2 <= 1e-12.
// FP32 takes the hardware reciprocal (v_rcp_f32, 1 ulp) instead of an IEEE
// division; FP64 keeps the division of the reference's formula.
template <typename T>
__device__ __forceinline__ T inv2(T sq) { return T(2) / sq; }
template <>
__device__ __forceinline__ float inv2<float>(float sq) { return 2.0f * __builtin_amdgcn_rcpf(sq); }
template <typename T>
__device__ __forceinline__ void overrelax(T hx, T hy, T hz, T &x, T &y, T &z) {
  const T sq = hx * hx + hy * hy + hz * hz;
  if (sq > T(1.0e-12)) {
    const T fac = inv2<T>(sq) * (x * hx + y * hy + z * hz);
    x = fac * hx - x;
    y = fac * hy - y;
    z = fac * hz - z;
  }
}

// FP64 heat bath, the reference's formulas: inverse-CDF cos(theta) about h,
// rotation z -> h of get_rotZ2VMat (identity when h is along +-z).
__device__ __forceinline__ void heat_bath_f64(double beta, double hx, double hy, double hz,
                                              u32x4 r, double &x, double &y, double &z) {
  const double u1 = u01_f64(r.x, r.y), u2 = u01_f64(r.z, r.w);
  const double tht = 2.0 * M_PI * u1;
  const double H = sqrt(hx * hx + hy * hy + hz * hz);
  double chi;
  if (fabs(H) > 1.0e-12) {
    const double x0 = 1.0 + u2 * (exp(-2.0 * beta * H) - 1.0);
    const double fc = -1.0 / (beta * H);
    chi = fc * log(x0) - 1.0;
  } else {
    chi = 2.0 * u2 - 1.0;
  }
  const double w = sqrt(fmax(0.0, 1.0 - chi * chi));
  double st, ct;
  sincos(tht, &st, &ct);
  const double s0 = w * ct, s1 = w * st, s2 = chi;
  const double V0 = hx / H, V1 = hy / H, V2 = hz / H;
  const double perp = V0 * V0 + V1 * V1;
  if (perp > 0.0) {
    const double sfac = (1.0 - V2) / perp;
    const double x11 = 1.0 - sfac * V0 * V0;
    const double x22 = 1.0 - sfac * V1 * V1;
    const double x33 = x11 + x22 - 1.0;
    const double x12 = V0 * V1 * (-sfac);
    x = x11 * s0 + x12 * s1 + V0 * s2;
    y = x12 * s0 + x22 * s1 + V1 * s2;
    z = -V0 * s0 - V1 * s1 + x33 * s2;
  } else {
    x = s0;
    y = s1;
    z = s2;
  }
}

// expm1(x) for x <= 0 and log1p(v) for v in (-1, 0] from the hardware's
// base-2 exponential and logarithm (v_exp_f32, v_log_f32) by Kahan's
// compensations (Goldberg, "What every computer scientist should know about
// floating-point arithmetic", 1991): (e^x - 1) x / log(e^x) and
// log(1 + v) v / ((1 + v) - 1) cancel the rounding of e^x and of 1 + v, so
// both stay accurate to a few ulps near 0, where the plain forms lose all
// their digits.
__device__ __forceinline__ float expm1_neg_f32(float x) {
  const float u = __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
  const float um1 = u - 1.0f;
  const float r = um1 * x * __builtin_amdgcn_rcpf(__builtin_amdgcn_logf(u) * 0.693147180559945309f);
  return u == 1.0f ? x : (um1 == -1.0f ? -1.0f : r);
}
__device__ __forceinline__ float log1p_neg_f32(float v) {
  const float w = 1.0f + v;
  const float r = __builtin_amdgcn_logf(w) * 0.693147180559945309f * v * __builtin_amdgcn_rcpf(w - 1.0f);
  return w == 1.0f ? v : r;
}

// FP32 heat bath: same distribution, written for single precision.
// t = 1 + cos(theta) = -log1p(u (e^{-2bH} - 1)) / (bH) keeps the pole region
// accurate, and the frame about h is the branch-free orthonormal basis of
// Duff et al. (JCGT 2017), equal to the reference rotation for h_z >= 0.
// Reciprocals, square roots and the azimuth's sine and cosine are the
// hardware's (v_rcp_f32, v_rsq_f32, v_sqrt_f32; v_sin_f32 and v_cos_f32 take
// the angle in turns, so 2 pi u1 needs no range reduction).  w0 and w1 are
// the two Philox words the uniforms come from.
__device__ __forceinline__ void heat_bath_f32(float beta, float hx, float hy, float hz,
                                              uint32_t w0, uint32_t w1, float &x, float &y,
                                              float &z) {
  const float u1 = u01_f32(w0), u2 = u01_f32(w1);
  const float H2 = hx * hx + hy * hy + hz * hz;
  float t;
  float ux = 0.f, uy = 0.f, uz = 1.f;
  bool frame = false;
  if (H2 > 1.0e-24f) {
    const float iH = __builtin_amdgcn_rsqf(H2);
    const float bh = beta * (H2 * iH);
    t = -log1p_neg_f32(u2 * expm1_neg_f32(-2.0f * bh)) * __builtin_amdgcn_rcpf(bh);
    ux = hx * iH; uy = hy * iH; uz = hz * iH;
    frame = true;
  } else {
    t = 2.0f * u2;
  }
  t = fminf(fmaxf(t, 0.0f), 2.0f);
  const float chi = t - 1.0f;
  const float w = __builtin_amdgcn_sqrtf(t * (2.0f - t));
  const float st = __builtin_amdgcn_sinf(u1), ct = __builtin_amdgcn_cosf(u1);
  const float s0 = w * ct, s1 = w * st;
  if (frame) {
    const float sg = copysignf(1.0f, uz);
    const float a = -__builtin_amdgcn_rcpf(sg + uz);
    const float b = ux * uy * a;
    // e1 = (1 + sg ux^2 a, sg b, -sg ux), e2 = (b, sg + uy^2 a, -uy)
    const float e1x = 1.0f + sg * ux * ux * a, e1y = sg * b, e1z = -sg * ux;
    const float e2x = b, e2y = sg + uy * uy * a, e2z = -uy;
    x = s0 * e1x + s1 * e2x + chi * ux;
    y = s0 * e1y + s1 * e2y + chi * uy;
    z = s0 * e1z + s1 * e2z + chi * uz;
  } else {
    x = s0; y = s1; z = chi;
  }
}

template <typename T>
__device__ __forceinline__ void heat_bath(double beta, T hx, T hy, T hz, u32x4 r, T &x, T &y, T &z);
template <>
__device__ __forceinline__ void heat_bath<double>(double beta, double hx, double hy, double hz,
                                                  u32x4 r, double &x, double &y, double &z) {
  heat_bath_f64(beta, hx, hy, hz, r, x, y, z);
}
template <>
__device__ __forceinline__ void heat_bath<float>(double beta, float hx, float hy, float hz,
                                                 u32x4 r, float &x, float &y, float &z) {
  heat_bath_f32((float)beta, hx, hy, hz, r.x, r.y, x, y, z);
}

// ------------------------------------------------------- sweep epilogues

// Adds (x, y, z) to the sublattice sums selected by the labels; every index
// is a compile-time constant so the sums stay in registers.
__device__ __forceinline__ void add_labels(double *v, int ls, int lh, int lv, int ld, double x,
                                           double y, double z) {
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const double w = (ls == q) ? 1.0 : 0.0;
    v[3 + 3 * q] += w * x; v[4 + 3 * q] += w * y; v[5 + 3 * q] += w * z;
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const double wh = (lh == q) ? 1.0 : 0.0, wv = (lv == q) ? 1.0 : 0.0, wd = (ld == q) ? 1.0 : 0.0;
    v[12 + 3 * q] += wh * x; v[13 + 3 * q] += wh * y; v[14 + 3 * q] += wh * z;
    v[18 + 3 * q] += wv * x; v[19 + 3 * q] += wv * y; v[20 + 3 * q] += wv * z;
    v[24 + 3 * q] += wd * x; v[25 + 3 * q] += wd * y; v[26 + 3 * q] += wd * z;
  }
}

// FP64 bond-list energy (get_energyValue) of one configuration; get(i, x,
// y, z) reads site i of the workgroup's copy.
template <typename G>
__device__ void heis_energy(const SweepArgs &a, int c, double *scratch, G get) {
  const int N = a.cp.N;
  double e = 0.0;
  for (int b = threadIdx.x; b < a.bd.Nb; b += blockDim.x) {
    const int i = a.bi[b], j = a.bj[b], k = a.bc[b];
    double xi, yi, zi, xj, yj, zj;
    get(i, xi, yi, zi);
    get(j, xj, yj, zj);
    e += a.bd.jxy[k] * (xi * xj + yi * yj) + a.bd.jz[k] * (zi * zj);
  }
  if (a.cp.h_on) {
    double mz = 0.0;
    for (int p = threadIdx.x; p < N; p += blockDim.x) {
      double x, y, z;
      get(p, x, y, z);
      mz += z;
    }
    e -= a.cp.h * mz;
  }
  double v1[1] = {e};
  block_sum<1>(v1, scratch);
  if (threadIdx.x == 0) a.energy[c] = v1[0];
}

template <typename T>
__device__ void heis_epilogue(const SweepArgs &a, int c, const T *sx, const T *sy, const T *sz,
                              double *scratch) {
  heis_energy(a, c, scratch, [&](int i, double &x, double &y, double &z) {
    x = (double)sx[i]; y = (double)sy[i]; z = (double)sz[i];
  });
}

// Order-parameter sums of one configuration per workgroup, read back from
// global memory after the sweep (measurement sweeps only): total M, the
// staggered / stripe sublattice sums and the Z2 plaquette sum.
template <typename T>
__device__ __forceinline__ void load_spin(const T *g, int Npad, int p, double &x, double &y,
                                          double &z) {
  x = (double)g[p]; y = (double)g[Npad + p]; z = (double)g[2 * Npad + p];
}
template <>
__device__ __forceinline__ void load_spin<int8_t>(const int8_t *g, int Npad, int p, double &x,
                                                  double &y, double &z) {
  x = (double)g[p]; y = 0.0; z = 0.0;
}

// Order-parameter sums and the spin-stiffness bond sums of get_spinStiff
// (CF/subrouts.cpp:6519-6620): dprod = J (x_i x_j + y_i y_j),
// vprod = J (x_i y_j - y_i x_j), weighted by C^2 and C for the three lattice
// directions (weights precomputed per bond).
// 512 threads: the 40 FP64 accumulators per thread stay in registers (at
// 1024 threads per workgroup the register limit of 128 forces them to
// scratch memory).
// With LDS the configuration is staged into LDS first, so the bond and
// plaquette gathers read LDS instead of L2.
constexpr int kObsThreads = 512;
template <typename T, bool LDS>
__global__ __launch_bounds__(kObsThreads) void obs_kernel(const SweepArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double scratch[NOBSX * 16];
  const int c = blockIdx.x;
  const int N = a.cp.N, Npad = a.cp.Npad;
  const int comps = sizeof(T) == 1 ? 1 : 3;
  const T *g = (const T *)a.spins + (size_t)c * comps * Npad;
  if constexpr (LDS) {
    T *sm = (T *)smem;
    for (int i = threadIdx.x; i < comps * Npad; i += blockDim.x) sm[i] = g[i];
    __syncthreads();
    g = sm;
  }
  double v[NOBSX];
#pragma unroll
  for (int k = 0; k < NOBSX; k++) v[k] = 0.0;
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    double x, y, z;
    load_spin<T>(g, Npad, p, x, y, z);
    v[0] += x; v[1] += y; v[2] += z;
    if (a.has_labels) {
      add_labels(v, a.lab[p], a.lab[N + p], a.lab[2 * N + p], a.lab[3 * N + p], x, y, z);
      if (a.z2) {
        double ax, ay, az, bx, by, bz, cx, cy, cz;
        load_spin<T>(g, Npad, a.z2nb[p], ax, ay, az);
        load_spin<T>(g, Npad, a.z2nb[N + p], bx, by, bz);
        load_spin<T>(g, Npad, a.z2nb[2 * N + p], cx, cy, cz);
        // IsingZ2Parameter: [S(i) - S(i+x+y)] . [S(i+x) - S(i+y)]
        const double d = (x - cx) * (ax - bx) + (y - cy) * (ay - by) + (z - cz) * (az - bz);
        const double fc = (fabs(d) > 1.0e-12) ? 1.0 / fabs(d) : 0.25;
        v[30] += fc * d;
      }
    }
  }
  const int Nb = a.bd.Nb;
#pragma unroll 4
  for (int b = threadIdx.x; b < Nb; b += blockDim.x) {
    double xi, yi, zi, xj, yj, zj;
    load_spin<T>(g, Npad, a.bi[b], xi, yi, zi);
    load_spin<T>(g, Npad, a.bj[b], xj, yj, zj);
    const double dp = xi * xj + yi * yj, vp = xi * yj - yi * xj;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[32 + k] += a.stiff_w[(size_t)k * Nb + b] * dp;
      v[35 + k] += a.stiff_w[(size_t)(3 + k) * Nb + b] * vp;
    }
  }
  block_sum<NOBSX>(v, scratch);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NOBSX; k++) a.obs[(size_t)c * NOBSX + k] = v[k];
  }
}

// Quasicrystal order parameters of one configuration (one workgroup), as
// raw site sums: projections on the Neel and four stripe reference states
// (get_qctNeelPrjt / get_qctStrpPrjt, CF/subrouts.cpp:6053-6180), the
// clock-ordered stripe vector (get_qctStripeMag, :6183-6262), the
// nearest-neighbour frustration parameter (get_qctNeel_N1fpar, :6011-6050)
// and the S(q) sums at the eight hotspots (get_qct_SqPks4, :8637-8700):
//   [0] sum S.neel, [1..4] sum S.stripe_i, [5, 6] sum rfac (rx, ry),
//   [7] sum (1 - S.<S_nb>) / 2, [8 + 6 i + 2 d] / [.. + 1] Re / Im of
//   sum S_d e^{-i q_i.r}.
// rx, ry: the site's spin against the differences of the clock-ordered
// shell-1 neighbours (slots 0..7, missing = 0) over their count (0 when the
// site has none; the reference divides by zero there), clipped to length 1.
template <typename T>
__global__ __launch_bounds__(256) void qc_obs_kernel(const SweepArgs a) {
  __shared__ double scratch[NQC * 16];
  const int c = blockIdx.x;
  const int N = a.cp.N, Npad = a.cp.Npad;
  const int comps = sizeof(T) == 1 ? 1 : 3;
  const T *g = (const T *)a.spins + (size_t)c * comps * Npad;
  const double *rf = a.qc_ref;
  double v[NQC];
#pragma unroll
  for (int k = 0; k < NQC; k++) v[k] = 0.0;
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    double x, y, z;
    load_spin<T>(g, Npad, p, x, y, z);
#pragma unroll
    for (int r = 0; r < 5; r++)
      v[r] += x * rf[(size_t)(3 * r) * N + p] + y * rf[(size_t)(3 * r + 1) * N + p] +
              z * rf[(size_t)(3 * r + 2) * N + p];
    double nx[8], ny[8], nz[8];
    double ax = 0.0, ay = 0.0, az = 0.0;
    int zc = 0;
#pragma unroll
    for (int s = 0; s < 8; s++) { nx[s] = 0.0; ny[s] = 0.0; nz[s] = 0.0; }
    for (int s = 0; s < a.qc_zn1; s++) {
      const int q = a.nb[(size_t)s * N + p];
      double bx = 0.0, by = 0.0, bz = 0.0;
      if (q < N) {
        load_spin<T>(g, Npad, q, bx, by, bz);
        zc++;
      }
      ax += bx; ay += by; az += bz;
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k == s) { nx[k] = bx; ny[k] = by; nz[k] = bz; }
    }
    const double fcz = zc > 0 ? 1.0 / zc : 0.0;
    const double v1x = (nx[0] + nx[1] + nx[4] + nx[5]) - (nx[2] + nx[3] + nx[6] + nx[7]);
    const double v1y = (ny[0] + ny[1] + ny[4] + ny[5]) - (ny[2] + ny[3] + ny[6] + ny[7]);
    const double v1z = (nz[0] + nz[1] + nz[4] + nz[5]) - (nz[2] + nz[3] + nz[6] + nz[7]);
    const double v2x = (nx[1] + nx[2] + nx[5] + nx[6]) - (nx[3] + nx[4] + nx[7] + nx[0]);
    const double v2y = (ny[1] + ny[2] + ny[5] + ny[6]) - (ny[3] + ny[4] + ny[7] + ny[0]);
    const double v2z = (nz[1] + nz[2] + nz[5] + nz[6]) - (nz[3] + nz[4] + nz[7] + nz[0]);
    const double rx = fcz * (x * v1x + y * v1y + z * v1z);
    const double ry = fcz * (x * v2x + y * v2y + z * v2z);
    const double rn = sqrt(rx * rx + ry * ry);
    const double rfac = rn > 1.0 ? 1.0 / rn : 1.0;
    v[5] += rfac * rx;
    v[6] += rfac * ry;
    const double z1 = 1.0 / (double)max(1, zc);
    v[7] += 0.5 * (1.0 - (x * z1 * ax + y * z1 * ay + z * z1 * az));
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const double2 e = a.qc_phase[(size_t)i * N + p];
      v[8 + 6 * i + 0] += x * e.x; v[8 + 6 * i + 1] += x * e.y;
      v[8 + 6 * i + 2] += y * e.x; v[8 + 6 * i + 3] += y * e.y;
      v[8 + 6 * i + 4] += z * e.x; v[8 + 6 * i + 5] += z * e.y;
    }
  }
  block_sum<NQC>(v, scratch);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NQC; k++) a.qc_obs[(size_t)c * NQC + k] = v[k];
  }
}

// ----------------------------------------------------------- Heisenberg

template <typename T, bool LDS, int ZC>
__global__ __launch_bounds__(kThreads) void heis_sweep_kernel(const SweepArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double scratch[NOBS * 16];
  const int c = blockIdx.x + a.config_base;
  const int N = a.cp.N, Npad = a.cp.Npad;
  T *g = (T *)a.spins + (size_t)c * 3 * Npad;
  T *s;
  if (LDS) {
    s = (T *)smem;
    for (int i = threadIdx.x; i < 3 * Npad; i += blockDim.x) s[i] = g[i];
    __syncthreads();
  } else {
    s = g;
  }
  T *sx = s, *sy = s + Npad, *sz = s + 2 * Npad;
  const int slot = a.slot_of_config[c];
  const double beta = a.beta_slot[slot];
  const bool boost = a.boost_slot[slot] != 0;
  const int nor = boost ? a.n_or_b : a.n_or;
  const int nhb = boost ? a.n_hb_b : a.n_hb;
  const T *Jt = (const T *)a.Jt;
  for (int pass = 0; pass < nor + nhb; pass++) {
    const bool hb = pass >= nor;
    const uint32_t step = sweep_of(a) * 16u + (uint32_t)pass;
    for (int col = 0; col < a.n_colours; col++) {
      for (int p = a.col_off[col] + threadIdx.x; p < a.col_off[col + 1]; p += blockDim.x) {
        T hx, hy, hz;
        local_field<T, ZC>(a.cp, a.nb, Jt, p, sx, sy, sz, hx, hy, hz);
        T x = sx[p], y = sy[p], z = sz[p];
        if (!hb) {
          overrelax<T>(hx, hy, hz, x, y, z);
        } else {
          const u32x4 r = draw4(a.seed, (uint32_t)a.perm[p], step, (uint32_t)c, STREAM_MC_UPDATE);
          heat_bath<T>(beta, hx, hy, hz, r, x, y, z);
        }
        sx[p] = x; sy[p] = y; sz[p] = z;
      }
      __syncthreads();
    }
  }
  heis_epilogue<T>(a, c, sx, sy, sz, scratch);
  if (LDS) {
    for (int i = threadIdx.x; i < 3 * Npad; i += blockDim.x) g[i] = s[i];
  }
}


// Fast path of the sweep: LDS-resident replica, uniform couplings, every
// colour class no larger than the workgroup (one site per thread per class),
// neighbour labels packed as 16-bit pairs.  The labels of the next class are
// fetched before the current class is updated, so their L2 latency overlaps
// the gathers, the update and the class barrier.
// AOS (FP32 only): the LDS copy holds each site as one float4 (x, y, z, 0),
// so a neighbour is one 16-byte LDS read at one address instead of three
// reads at three addresses; otherwise three component arrays.
template <typename T>
__device__ __forceinline__ T shell_j(const SweepArgs &a, int sh, bool z);
template <>
__device__ __forceinline__ float shell_j<float>(const SweepArgs &a, int sh, bool z) {
  // selects values, not pointers: a pointer into the kernel's argument
  // block makes the compiler copy the whole block to scratch memory
  const float j0 = z ? a.fshjz[0] : a.fshj[0];
  const float j1 = z ? a.fshjz[1] : a.fshj[1];
  const float j2 = z ? a.fshjz[2] : a.fshj[2];
  return sh == 0 ? j0 : (sh == 1 ? j1 : j2);
}
template <>
__device__ __forceinline__ double shell_j<double>(const SweepArgs &a, int sh, bool z) {
  const double j0 = z ? a.shjz[0] : a.shj[0];
  const double j1 = z ? a.shjz[1] : a.shj[1];
  const double j2 = z ? a.shjz[2] : a.shj[2];
  return sh == 0 ? j0 : (sh == 1 ? j1 : j2);
}

// Sites of the statically placed float4 replica (STAT kernels): 150 KB.
constexpr int kStatSites = 9600;

// three floats read from a 16-byte aligned float4 site (a 12-byte LDS read)
typedef float float3v __attribute__((ext_vector_type(3)));

// Byte offset 16 x label of the low and high 16-bit label of a packed word,
// one instruction each (the sub-dword shift reads either half directly).
__device__ __forceinline__ uint32_t lab_lo16(uint32_t w) {
  uint32_t r;
  asm("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0"
      : "=v"(r) : "v"(w));
  return r;
}
__device__ __forceinline__ uint32_t lab_hi16(uint32_t w) {
  uint32_t r;
  asm("v_lshlrev_b32_sdwa %0, 4, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1"
      : "=v"(r) : "v"(w));
  return r;
}

// The replica in LDS allows one workgroup per CU, at most 4 waves per SIMD:
// the compiler may then use up to 128 VGPRs, which lets it keep a site's
// neighbour reads in flight together rather than two at a time.
// STAT (FP32 float4 replicas of more than 4096 sites): the replica sits in a
// statically placed LDS array, whose address is a constant; a neighbour's
// LDS address is then its label times 16 from one sub-dword shift, folded
// into the read's immediate offset, instead of an unpack, a shift and a base
// add.
// FP32 heat bath: a Philox block drawn for a thread's site of an even colour
// class serves, with its words z and w, the same thread's site of the next
// (odd) class; a site of an odd class whose thread had none in the class
// before draws its own block.  Counters stay (site, sweep * 16 + pass,
// configuration, 0), so no two draws share one.
// EW > 0 (with the table energy): the energy from the emask-slot label
// table of EW words per site (nbE), every neighbour read issued before the
// products, instead of the full table with a branch per slot.
template <typename T, int ZC, bool AOS, bool CP = false, bool STAT = false, int EW = 0>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 4)))
void heis_sweep_fast(const SweepArgs a) {
  static_assert(!AOS || sizeof(T) == 4, "the float4 layout is FP32 only");
  static_assert(!STAT || (AOS && !CP), "static placement is for the float4 layout");
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double scratch[NOBS * 16];
  constexpr int ZH = (ZC + 1) / 2;
  const int c = blockIdx.x;
  const int N = a.cp.N, Npad = a.cp.Npad;
  T *g = (T *)a.spins + (size_t)c * 3 * Npad;
  T *s = (T *)smem;
  float4 *s4;
  if constexpr (STAT) {
    __shared__ __align__(16) float4 stat4[kStatSites];
    s4 = stat4;
  } else {
    s4 = (float4 *)smem;
  }
  // neighbour read from a packed label word: half 0 or 1.  The float4 sites
  // are read as three floats (one 12-byte LDS read of the 16-byte aligned
  // site): a quarter less LDS traffic than the whole float4
  auto nget = [&](uint32_t w, int half, T &x, T &y, T &z) {
    if constexpr (STAT) {
      const uint32_t off = half ? lab_hi16(w) : lab_lo16(w);
      const float3v v = *(const float3v *)((const char *)s4 + off);
      x = v.x; y = v.y; z = v.z;
    } else if constexpr (AOS) {
      const float3v v = *(const float3v *)&s4[half ? (int)(w >> 16) : (int)(w & 0xffffu)];
      x = v.x; y = v.y; z = v.z;
    } else {
      const int i = half ? (int)(w >> 16) : (int)(w & 0xffffu);
      x = s[i]; y = s[Npad + i]; z = s[2 * Npad + i];
    }
  };
  if constexpr (AOS) {
    for (int i = threadIdx.x; i < Npad; i += blockDim.x)
      s4[i] = make_float4(g[i], g[Npad + i], g[2 * Npad + i], 0.f);
  } else {
    for (int i = threadIdx.x; i < 3 * Npad; i += blockDim.x) s[i] = g[i];
  }
  T *sx = s, *sy = s + Npad, *sz = s + 2 * Npad;
  auto get = [&](int i, T &x, T &y, T &z) {
    if constexpr (AOS) {
      const float3v v = *(const float3v *)&s4[i];
      x = v.x; y = v.y; z = v.z;
    } else {
      x = sx[i]; y = sy[i]; z = sz[i];
    }
  };
  const int slot = a.slot_of_config[c];
  const double beta = a.beta_slot[slot];
  const float fbeta = (float)beta;
  const bool boost = a.boost_slot[slot] != 0;
  const int nor = boost ? a.n_or_b : a.n_or;
  const int nhb = boost ? a.n_hb_b : a.n_hb;
  const int nc = a.n_colours;
  const int total = (nor + nhb) * nc;
  const int tid = threadIdx.x;
  const T hz0 = a.cp.h_on ? T(-a.cp.h) : T(0);
  const uint32_t sweep16 = sweep_of(a) * 16u;   // read once, before the label loads
  // FP32 heat bath: Philox words z, w kept for the next (odd) class
  uint32_t keep_z = 0u, keep_w = 0u;
  bool keep = false;
  uint32_t qa[ZH], qb[ZH];
  int la = 0, lb = 0;
  // the packed labels are padded to ZC slots with the zero-spin sentinel
  // (coupling 0), so every slot is read without a branch.  The next class's
  // labels are loaded unconditionally (lanes past the class read row 0), so
  // the loads stay in flight across the current class's update
  // buffer loads: the lane's site is the per-lane offset, the word row the
  // scalar offset, so the loads need no per-lane address arithmetic
  const __amdgpu_buffer_rsrc_t nrsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)a.nb16, (short)0, 0x7fffffff, 0x00020000);
  const __amdgpu_buffer_rsrc_t prsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)a.perm, (short)0, 0x7fffffff, 0x00020000);
  // class bounds: the current class's (cb, ce) and the next one's (cbn, cen)
  // travel with the labels; the class after that is read one iteration
  // ahead, so no scalar load is waited for at the top of an iteration
  int cb = 0, ce = 0, cbn = 0, cen = 0;
  auto fetch = [&](uint32_t (&q)[ZH], int &lab, bool &v, int b0, int b1) __attribute__((always_inline)) {
    const int p = b0 + tid;
    v = p < b1;
    const int pc4 = 4 * (v ? p : 0);
#pragma unroll
    for (int k = 0; k < ZH; k++) q[k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(nrsrc, pc4, 4 * k * N, 0);
    lab = __builtin_amdgcn_raw_buffer_load_b32(prsrc, pc4, 0, 0);
  };
  if (total > 0) {
    bool va;
    cb = a.col_off[0];
    ce = a.col_off[1];
    fetch(qa, la, va, cb, ce);
    if (total > 1) {
      cbn = a.col_off[1 % nc];
      cen = a.col_off[1 % nc + 1];
    }
  }
  // wave-level ordering (wdep != nullptr): wprog[w] = class iterations wave
  // w has completed (spins written).  Before iteration it a wave waits until
  // every wave that owns a neighbour of its sites has completed it - 1: then
  // the neighbours it reads hold their values of iteration it - 1, and no
  // wave still reading one of its sites for an earlier iteration remains, so
  // each site reads exactly what it reads under a barrier per class.
  __shared__ int wprog[16];
  const bool wsync = a.wdep != nullptr;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  if (tid < 16) wprog[tid] = 0;
  // this wave's masks, lane k holding class k's (at most 64 classes), so the
  // class loop picks one with a lane read instead of a memory load
  const uint32_t dmv = (wsync && lane < nc) ? a.wdep[lane * 16 + wv] : 0u;
  __syncthreads();
  // one colour class; the label registers alternate between two sets (the
  // loop runs two classes per trip), so no copies follow the barrier.  Lanes
  // past a class hold whatever row 0's labels were loaded and do no update.
  // always inlined: called twice, a lambda that stays a call takes the label
  // arrays by reference in scratch memory
  auto iteration = [&](int it, const uint32_t (&qc)[ZH], int lc, uint32_t (&qn)[ZH],
                       int &ln) __attribute__((always_inline)) {
    const int pass = it / nc;
    const int col = it - pass * nc;
    bool vn = false;
    if (it + 1 < total) fetch(qn, ln, vn, cbn, cen);
    (void)vn;
    int cb2 = 0, ce2 = 0;
    if (it + 2 < total) {
      const int c2 = (it + 2) % nc;
      cb2 = a.col_off[c2];
      ce2 = a.col_off[c2 + 1];
    }
    const int p = cb + tid;
    // compact labels: the words any site of this wave's 64-site chunk uses
    const int nw = CP ? __builtin_amdgcn_readfirstlane((int)a.wcnt[col * 16 + (tid >> 6)]) : ZH;
    const bool act = p < ce;
    // FP32: the site's own spin (written only by this wave) and the
    // heat-bath draws do not depend on the neighbours, so they are issued
    // before the wait for the neighbours' owners and overlap its polls; the
    // own spin's latency hides behind them.  FP64 does both after the field
    // sums: held across them, they would push its 128 VGPRs into scratch.
    constexpr bool EARLY = sizeof(T) == 4;
    T x = T(0), y = T(0), z = T(0);
    const uint32_t step = sweep16 + (uint32_t)pass;
    uint32_t w0 = keep_z, w1 = keep_w;
    u32x4 r64 = {0u, 0u, 0u, 0u};
    auto draw_early = [&]() __attribute__((always_inline)) {
      if (pass < nor) return;
      if constexpr (sizeof(T) == 4) {
        if (!((col & 1) && keep)) {
          const u32x4 r = draw4(a.seed, (uint32_t)lc, step, (uint32_t)c, STREAM_MC_UPDATE);
          w0 = r.x; w1 = r.y;
          keep_z = r.z; keep_w = r.w;
        }
      } else {
        r64 = draw4(a.seed, (uint32_t)lc, step, (uint32_t)c, STREAM_MC_UPDATE);
      }
    };
    if constexpr (EARLY) {
      if (act) {
        get(p, x, y, z);
        draw_early();
      }
    }
    if (wsync) {
      const uint32_t dm = (uint32_t)__builtin_amdgcn_readlane((int)dmv, col);
      if (dm) {
        // bounded: a wait that runs out flags sync_err and goes on, so the
        // kernel always ends (the host reports the flag)
        uint32_t n = 0;
        while (true) {
          const int v = __hip_atomic_load(&wprog[lane & 15], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          const bool behind = ((dm >> (lane & 15)) & 1u) && v < it;
          if (__builtin_amdgcn_ballot_w64(behind) == 0) break;
          if (++n > (1u << 20)) {
            if (lane == 0) atomicOr(a.sync_err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    if (act) {
      T hx = T(0), hy = T(0), hz = hz0;
      if constexpr (!CP) {
        // the neighbour reads in groups (10 neighbours in FP32, 4 in FP64),
        // each group issued before its first product (the scheduling
        // barriers keep them together, so the group's reads are in flight at
        // once); the sums run in slot order as below
        constexpr int GWT = sizeof(T) == 8 ? 2 : 5;
        constexpr int GW = ZH < GWT ? ZH : GWT;   // label words per group
#pragma unroll
        for (int k0 = 0; k0 < ZH; k0 += GW) {
          T nx[2 * GW], ny[2 * GW], nz[2 * GW];
#pragma unroll
          for (int k = 0; k < GW; k++) {
            if (k0 + k < ZH) {
              nget(qc[k0 + k], 0, nx[2 * k], ny[2 * k], nz[2 * k]);
              if (2 * (k0 + k) + 1 < ZC) nget(qc[k0 + k], 1, nx[2 * k + 1], ny[2 * k + 1], nz[2 * k + 1]);
            }
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int u = 0; u < 2 * GW; u++) {
            const int sl = 2 * k0 + u;
            if (sl < ZC) {
              const T J = cj_xy<T>(a.cp, sl), Jz = cj_z<T>(a.cp, sl);
              hx += J * nx[u]; hy += J * ny[u]; hz += Jz * nz[u];
            }
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
#pragma unroll
      for (int k = 0; k < (CP ? ZH : 0); k++) {
        if (CP && k >= nw) break;
        int i0, i1;
        T J0, Jz0, J1, Jz1;
        if constexpr (CP) {
          const uint32_t w = qc[k];
          i0 = (int)(w & 0x3fffu);
          i1 = (int)((w >> 16) & 0x3fffu);
          const int s0 = (int)((w >> 14) & 3u), s1 = (int)(w >> 30);
          J0 = shell_j<T>(a, s0, false); Jz0 = shell_j<T>(a, s0, true);
          J1 = shell_j<T>(a, s1, false); Jz1 = shell_j<T>(a, s1, true);
        } else {
          i0 = (int)(qc[k] & 0xffffu);
          i1 = (int)(qc[k] >> 16);
          J0 = cj_xy<T>(a.cp, 2 * k); Jz0 = cj_z<T>(a.cp, 2 * k);
          J1 = cj_xy<T>(a.cp, 2 * k + 1); Jz1 = cj_z<T>(a.cp, 2 * k + 1);
        }
        T x0, y0, z0, x1, y1, z1;
        get(i0, x0, y0, z0);
        get(i1, x1, y1, z1);
        hx += J0 * x0; hy += J0 * y0; hz += Jz0 * z0;
        hx += J1 * x1; hy += J1 * y1; hz += Jz1 * z1;
      }
      if constexpr (!EARLY) {
        get(p, x, y, z);
        draw_early();
      }
      if (pass < nor) {
        overrelax<T>(hx, hy, hz, x, y, z);
      } else {
        if constexpr (sizeof(T) == 4) heat_bath_f32(fbeta, hx, hy, hz, w0, w1, x, y, z);
        else heat_bath<T>(beta, hx, hy, hz, r64, x, y, z);
      }
      if constexpr (AOS) s4[p] = make_float4(x, y, z, 0.f);
      else { sx[p] = x; sy[p] = y; sz[p] = z; }
    }
    if (pass >= nor) keep = !(col & 1) && p < ce;
    if (wsync) {
      // the release orders this wave's spin writes before the count
      if (lane == 0) __hip_atomic_store(&wprog[wv], it + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      __syncthreads();
    }
    cb = cbn;
    ce = cen;
    cbn = cb2;
    cen = ce2;
  };
  for (int it = 0; it < total; it += 2) {
    iteration(it, qa, la, qb, lb);
    if (it + 1 < total) iteration(it + 1, qb, lb, qa, la);
  }
  if (wsync) __syncthreads();
  bool written = false;
  if constexpr (EW > 0) {
    // E = escale sum_i S_i . sum_{e < 2 EW} (ej_e S_nb(i, e)) - h sum_i z_i over
    // the emask slots in slot order: the same FP64 products and sums as the
    // branch-per-slot form below (a padding slot adds J 0 times the
    // sentinel's zero spin); EB sites' labels per batch, all 2 EW
    // neighbour reads of a site ahead of its products.  Each site's spin goes
    // back to global memory here, where it is in registers anyway.
    constexpr int EB = 3;
    const __amdgpu_buffer_rsrc_t ersrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)a.nbE, (short)0, 0x7fffffff, 0x00020000);
    double e = 0.0, mz = 0.0;
    for (int p0 = tid; p0 < N; p0 += EB * blockDim.x) {
      uint32_t q[EB][EW];
#pragma unroll
      for (int b = 0; b < EB; b++) {
        const int pc = 4 * min(p0 + b * (int)blockDim.x, N - 1);
#pragma unroll
        for (int k = 0; k < EW; k++) q[b][k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(ersrc, pc, 4 * k * N, 0);
      }
#pragma unroll
      for (int b = 0; b < EB; b++) {
        const int p = p0 + b * (int)blockDim.x;
        if (p < N) {
          T u[2 * EW], v[2 * EW], w[2 * EW];
#pragma unroll
          for (int k = 0; k < EW; k++) {
            nget(q[b][k], 0, u[2 * k], v[2 * k], w[2 * k]);
            nget(q[b][k], 1, u[2 * k + 1], v[2 * k + 1], w[2 * k + 1]);
          }
          T x, y, z;
          get(p, x, y, z);
          double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
          for (int k = 0; k < 2 * EW; k++) {
            ax = fma(a.ejxy[k], (double)u[k], ax);
            ay = fma(a.ejxy[k], (double)v[k], ay);
            az = fma(a.ejz[k], (double)w[k], az);
          }
          e = fma((double)x, ax, e);
          e = fma((double)y, ay, e);
          e = fma((double)z, az, e);
          mz += (double)z;
          g[p] = x; g[Npad + p] = y; g[2 * Npad + p] = z;
        }
      }
    }
    written = true;   // the sentinel site N keeps its zero spin
    e *= a.escale;
    if (a.cp.h_on) e -= a.cp.h * mz;
    double v1[1] = {e};
    block_sum<1>(v1, scratch);
    if (threadIdx.x == 0) a.energy[c] = v1[0];
  } else if (a.table_energy) {
    // E = escale sum_i S_i . sum_{s in emask} (J_s S_nb(i,s)) - h sum_i z_i,
    // every product and sum in FP64; the labels are read again (coalesced,
    // from L2), EB sites' worth per batch so their latency is met once per
    // batch (one site in FP64, whose registers a batch would spill), and the
    // spins from LDS.  Each site's spin goes back to global memory here,
    // where it is in registers anyway.
    constexpr int EB = sizeof(T) == 4 ? 3 : 1;
    double e = 0.0, mz = 0.0;
    for (int p0 = tid; p0 < N; p0 += EB * blockDim.x) {
      uint32_t q[EB][ZH];
#pragma unroll
      for (int b = 0; b < EB; b++) {
        const int pc = 4 * min(p0 + b * (int)blockDim.x, N - 1);
#pragma unroll
        for (int k = 0; k < ZH; k++) q[b][k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(nrsrc, pc, 4 * k * N, 0);
      }
#pragma unroll
      for (int b = 0; b < EB; b++) {
        const int p = p0 + b * (int)blockDim.x;
        if (p < N) {
          double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
          for (int k = 0; k < ZH; k++)
#pragma unroll
            for (int h = 0; h < 2; h++) {
              const int sl = 2 * k + h;
              if (a.emask & (1u << sl)) {
                T u, v, w;
                nget(q[b][k], h, u, v, w);
                ax = fma(a.cp.jxy[sl], (double)u, ax);
                ay = fma(a.cp.jxy[sl], (double)v, ay);
                az = fma(a.cp.jz[sl], (double)w, az);
              }
            }
          T x, y, z;
          get(p, x, y, z);
          e = fma((double)x, ax, e);
          e = fma((double)y, ay, e);
          e = fma((double)z, az, e);
          mz += (double)z;
          g[p] = x; g[Npad + p] = y; g[2 * Npad + p] = z;
        }
      }
    }
    written = true;   // the sentinel site N keeps its zero spin
    e *= a.escale;
    if (a.cp.h_on) e -= a.cp.h * mz;
    double v1[1] = {e};
    block_sum<1>(v1, scratch);
    if (threadIdx.x == 0) a.energy[c] = v1[0];
  } else {
    heis_energy(a, c, scratch, [&](int i, double &x, double &y, double &z) {
      T u, v, w;
      get(i, u, v, w);
      x = (double)u; y = (double)v; z = (double)w;
    });
  }
  if (written) return;
  if constexpr (AOS) {
    for (int i = threadIdx.x; i < Npad; i += blockDim.x) {
      const float4 v = s4[i];
      g[i] = v.x; g[Npad + i] = v.y; g[2 * Npad + i] = v.z;
    }
  } else {
    for (int i = threadIdx.x; i < 3 * Npad; i += blockDim.x) g[i] = s[i];
  }
}

// --------------------------------------------------------------- Ising

// Metropolis with FP64 fields and the deterministic exponential; spins are
// int8 +-1 with a zero sentinel.  Energies use exact integer bond counts per
// coupling class, E = sum_c J_c n_c - h M, evaluated in class order.
template <bool LDS, int ZC>
__global__ __launch_bounds__(kThreads) void ising_sweep_kernel(const SweepArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double scratch[NOBS * 16];
  __shared__ long long iscr[(kBCMAX + 1) * 16];
  const int c = blockIdx.x + a.config_base;
  const int N = a.cp.N, Npad = a.cp.Npad;
  int8_t *g = (int8_t *)a.spins + (size_t)c * Npad;
  int8_t *s;
  if (LDS) {
    s = (int8_t *)smem;
    for (int i = threadIdx.x; i < Npad; i += blockDim.x) s[i] = g[i];
    __syncthreads();
  } else {
    s = g;
  }
  const int slot = a.slot_of_config[c];
  const double beta = a.beta_slot[slot];
  const double *Jt = (const double *)a.Jt;
  const uint32_t step = sweep_of(a) * 16u;
  for (int col = 0; col < (a.metro_passes > 0 ? a.n_colours : 0); col++) {
    for (int p = a.col_off[col] + threadIdx.x; p < a.col_off[col + 1]; p += blockDim.x) {
      double h = a.cp.h_on ? -a.cp.h : 0.0;
      int q[ZC];
#pragma unroll
      for (int k = 0; k < ZC; k++) q[k] = (k < a.cp.Z) ? a.nb[(size_t)k * N + p] : N;
#pragma unroll
      for (int k = 0; k < ZC; k++) {
        if (k < a.cp.Z) {
          const double J = a.cp.general ? Jt[(size_t)k * N + p] : a.cp.jxy[k];
          h += J * (double)s[q[k]];
        }
      }
      const double spin = (double)s[p];
      const double dE = 2.0 * spin * h;
      const u32x4 r = draw4(a.seed, (uint32_t)a.perm[p], step, (uint32_t)c, STREAM_MC_UPDATE);
      const double u = u01_f64(r.x, r.y);
      if (accept_exp(u, beta * dE)) s[p] = (int8_t)(-s[p]);
    }
    __syncthreads();
  }
  // exact energy: integer bond counts per class
  long long cnt[kBCMAX + 1];
#pragma unroll
  for (int k = 0; k <= kBCMAX; k++) cnt[k] = 0;
  for (int b = threadIdx.x; b < a.bd.Nb; b += blockDim.x) {
    const int k = a.bc[b];
    const long long pr = (long long)(s[a.bi[b]] * s[a.bj[b]]);
#pragma unroll
    for (int kk = 0; kk < kBCMAX; kk++)
      if (kk == k) cnt[kk] += pr;
  }
  for (int p = threadIdx.x; p < N; p += blockDim.x) cnt[kBCMAX] += s[p];
  {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k <= kBCMAX; k++) cnt[k] = wave_sum_ll(cnt[k]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k <= kBCMAX; k++) iscr[k * 16 + w] = cnt[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k <= kBCMAX; k++) {
      long long t = 0;
      for (int i = 0; i < nw; i++) t += iscr[k * 16 + i];
      cnt[k] = t;
    }
  }
  if (threadIdx.x == 0) {
    double E = 0.0;
#pragma unroll
    for (int k = 0; k < kBCMAX; k++)
      if (k < a.bd.n_class) E = fma(a.bd.jxy[k], (double)cnt[k], E);
    if (a.cp.h_on) E = fma(-a.cp.h, (double)cnt[kBCMAX], E);
    a.energy[c] = E;
  }
  if (LDS) {
    for (int i = threadIdx.x; i < Npad; i += blockDim.x) g[i] = s[i];
  }
}

// Fast Ising sweep (one Metropolis pass): the int8 replica in LDS, uniform
// couplings per slot, every colour class no larger than the workgroup (one
// site per thread per class).  The packed 16-bit labels and the Philox
// counter label of the next class are loaded into registers while the
// current class updates; classes are ordered per wave (wdep) or by a
// barrier.  The field (FP64, slot order), the draw and the deterministic
// acceptance are those of ising_sweep_kernel, so the chain is bit-identical
// to it.  The energy's exact integer bond counts per coupling class come from
// the half-shell label table (each bond once, its class per slot in icls)
// when the host has verified that table against the bond list, else from the
// bond list.
template <int ZC>
__global__ __launch_bounds__(kThreads) void ising_sweep_fast(const SweepArgs a) {
  constexpr int ZH = (ZC + 1) / 2;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ long long iscr[(kBCMAX + 1) * 16];
  __shared__ int wprog[16];
  const int c = blockIdx.x;
  const int N = a.cp.N, Npad = a.cp.Npad;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  int8_t *g = (int8_t *)a.spins + (size_t)c * Npad;
  int8_t *s = (int8_t *)smem;
  for (int i = tid; i < Npad; i += blockDim.x) s[i] = g[i];
  const int slot = a.slot_of_config[c];
  const double beta = a.beta_slot[slot];
  const uint32_t step = sweep_of(a) * 16u;
  const int nc = a.n_colours;
  const int total = a.metro_passes > 0 ? nc : 0;
  const int Z = a.cp.Z;
  const double h0 = a.cp.h_on ? -a.cp.h : 0.0;
  const bool wsync = a.wdep != nullptr;
  const __amdgpu_buffer_rsrc_t nrsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)a.nb16, (short)0, 0x7fffffff, 0x00020000);
  const __amdgpu_buffer_rsrc_t prsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)a.perm, (short)0, 0x7fffffff, 0x00020000);
  uint32_t q[ZH], qn[ZH];
  int lab = 0, labn = 0;
  // lanes past a class read row 0 and do no update
  auto fetch = [&](uint32_t (&d)[ZH], int &l, int b0, int b1) __attribute__((always_inline)) {
    const int p = b0 + tid;
    const int pc4 = 4 * (p < b1 ? p : 0);
#pragma unroll
    for (int k = 0; k < ZH; k++) d[k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(nrsrc, pc4, 4 * k * N, 0);
    l = __builtin_amdgcn_raw_buffer_load_b32(prsrc, pc4, 0, 0);
  };
#pragma unroll
  for (int k = 0; k < ZH; k++) q[k] = qn[k] = 0u;
  if (total > 0) fetch(q, lab, a.col_off[0], a.col_off[1]);
  if (tid < 16) wprog[tid] = 0;
  const uint32_t dmv = (wsync && lane < nc) ? a.wdep[lane * 16 + wv] : 0u;
  __syncthreads();
  for (int col = 0; col < total; col++) {
    const int cb = a.col_off[col], ce = a.col_off[col + 1];
    if (col + 1 < total) fetch(qn, labn, a.col_off[col + 1], a.col_off[col + 2]);
    const int p = cb + tid;
    // the own spin (written only by this wave) and the draw do not depend on
    // the neighbours: issued before the wait, they fill its polls
    int8_t sp = 0;
    double u = 0.0;
    if (p < ce) {
      sp = s[p];
      const u32x4 r = draw4(a.seed, (uint32_t)lab, step, (uint32_t)c, STREAM_MC_UPDATE);
      u = u01_f64(r.x, r.y);
    }
    if (wsync) {
      // the ordering of heis_sweep_fast: wait until every wave owning a
      // neighbour of this wave's sites has completed class col - 1
      const uint32_t dm = (uint32_t)__builtin_amdgcn_readlane((int)dmv, col);
      if (dm) {
        uint32_t n = 0;
        while (true) {
          const int v = __hip_atomic_load(&wprog[lane & 15], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          const bool behind = ((dm >> (lane & 15)) & 1u) && v < col;
          if (__builtin_amdgcn_ballot_w64(behind) == 0) break;
          if (++n > (1u << 20)) {
            if (lane == 0) atomicOr(a.sync_err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    if (p < ce) {
      // every neighbour read issued before the sums, which run in slot order
      int8_t nv[2 * ZH];
#pragma unroll
      for (int k = 0; k < ZH; k++) {
        nv[2 * k] = s[q[k] & 0xffffu];
        nv[2 * k + 1] = s[q[k] >> 16];
      }
      double h = h0;
#pragma unroll
      for (int k = 0; k < ZC; k++)
        if (k < Z) h += a.cp.jxy[k] * (double)nv[k];
      const double spin = (double)sp;
      const double dE = 2.0 * spin * h;
      if (accept_exp(u, beta * dE)) s[p] = (int8_t)(-sp);
    }
    if (wsync) {
      if (lane == 0) __hip_atomic_store(&wprog[wv], col + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < ZH; k++) q[k] = qn[k];
    lab = labn;
  }
  if (wsync) __syncthreads();
  // exact energy: integer bond counts per coupling class
  long long cnt[kBCMAX + 1];
#pragma unroll
  for (int k = 0; k <= kBCMAX; k++) cnt[k] = 0;
  if (a.itable) {
    const __amdgpu_buffer_rsrc_t ersrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)a.nbE, (short)0, 0x7fffffff, 0x00020000);
    const int nEw = a.nEw;
    int acc[16], m = 0;
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0;
    for (int p = tid; p < N; p += blockDim.x) {
      uint32_t w[8];
#pragma unroll
      for (int k = 0; k < 8; k++)
        w[k] = k < nEw ? (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(ersrc, 4 * p, 4 * k * N, 0)
                       : (uint32_t)N | ((uint32_t)N << 16);
      const int si = s[p];
      m += si;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        acc[2 * k] += si * (int)s[w[k] & 0xffffu];
        acc[2 * k + 1] += si * (int)s[w[k] >> 16];
      }
    }
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const int k = a.icls[e];
#pragma unroll
      for (int kk = 0; kk < kBCMAX; kk++)
        if (kk == k) cnt[kk] += acc[e];
    }
    cnt[kBCMAX] = m;
  } else {
    for (int b = tid; b < a.bd.Nb; b += blockDim.x) {
      const int k = a.bc[b];
      const long long pr = (long long)(s[a.bi[b]] * s[a.bj[b]]);
#pragma unroll
      for (int kk = 0; kk < kBCMAX; kk++)
        if (kk == k) cnt[kk] += pr;
    }
    for (int p = tid; p < N; p += blockDim.x) cnt[kBCMAX] += s[p];
  }
  {
    const int w = tid >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k <= kBCMAX; k++) cnt[k] = wave_sum_ll(cnt[k]);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k <= kBCMAX; k++) iscr[k * 16 + w] = cnt[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k <= kBCMAX; k++) {
      long long t = 0;
      for (int i = 0; i < nw; i++) t += iscr[k * 16 + i];
      cnt[k] = t;
    }
  }
  if (tid == 0) {
    double E = 0.0;
#pragma unroll
    for (int k = 0; k < kBCMAX; k++)
      if (k < a.bd.n_class) E = fma(a.bd.jxy[k], (double)cnt[k], E);
    if (a.cp.h_on) E = fma(-a.cp.h, (double)cnt[kBCMAX], E);
    a.energy[c] = E;
  }
  for (int i = tid; i < Npad; i += blockDim.x) g[i] = s[i];
}

// ------------------------------------------- measurement + replica exchange

struct PTArgs {
  int R, N;
  int parity;          // -1 none, 0/1 pair parity
  int cap50;
  int measure;
  int n_stag, has_labels, z2, c3, h_on, ising;
  uint64_t seed;
  uint32_t sweep;
  const double *energy;    // per configuration
  const double *obs;       // per configuration
  const double *beta_slot;
  int32_t *config_of_slot;
  int32_t *slot_of_config;
  int32_t *boost_slot;
  long long *pt_cnt, *pt_exc;
  double *moments;         // per slot, NMOM
  long long *n_meas;
  double *etrace;          // per slot energies of this sweep (nullable)
  uint32_t *sweep_dev;     // graph replays: sweep index (advanced by this kernel)
  int parity_alt;          // with sweep_dev: pair parity = sweep % 2
  // quasicrystal measurement: sums from qc_obs_kernel, 0.5 / Nq of
  // get_qct_SqPks4, per-slot histograms
  int qc;
  double qc_fc;
  const double *qc_obs;
  int32_t *qc_hist;
};

// get_HistIndex (CF/subrouts.cpp:795-812) with r0 = 1 and fc = (kHst - 1) / 2
__device__ __forceinline__ int hist_index(double r) {
  const int i = (int)round((r + 1.0) * (0.5 * (kHst - 1)));
  return i < 0 ? 0 : (i >= kHst ? kHst - 1 : i);
}

__device__ double stag_mag(const double *sub, int n0, double iN) {
  // absStaggMag (CF/subrouts.cpp:5382-5462)
  const double norm = (double)n0 * iN;
  double sx = 0, sy = 0, sz = 0, sq = 0;
  for (int n = 0; n < n0; n++) {
    const double mx = norm * sub[3 * n], my = norm * sub[3 * n + 1], mz = norm * sub[3 * n + 2];
    sx += mx; sy += my; sz += mz;
    sq += mx * mx + my * my + mz * mz;
  }
  const double r = sq - (sx * sx + sy * sy + sz * sz) / (double)n0;
  return (1.0 / sqrt((double)n0)) * sqrt(fmax(r, 0.0));
}

__global__ void measure_exchange_kernel(const PTArgs a) {
  const int t = threadIdx.x;
  const uint32_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
  const int parity = (a.sweep_dev && a.parity_alt) ? (int)(sweep % 2u) : a.parity;
  const double iN = 1.0 / (double)a.N;
  if (t < a.R && (a.measure || a.etrace)) {
    const int k = t;
    const int c = a.config_of_slot[k];
    const double E = a.energy[c];
    if (a.etrace) a.etrace[k] = E;
    if (a.measure) {
      const double *o = a.obs + (size_t)c * NOBSX;
      double *m = a.moments + (size_t)k * NMOM;
      const double F1 = iN * sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
      const double F2 = F1 * F1;
      double T1 = (a.ising ? 0.0 : o[2]) * iN;
      if (!a.h_on) T1 = fabs(T1);
      double Q1 = 0, SH = 0, SV = 0, SD = 0, X1 = 0;
      double S1q = -1.0;
      if (a.qc) {
        // quasicrystal forms (CF/mc_hm_code.cpp:4973-5056): Q = |Neel
        // projection|, S = |stripe projections|, X = |stripe vector|
        const double *q = a.qc_obs + (size_t)c * NQC;
        const double P0 = iN * q[0];
        double P[4];
        bool cut = false;
        int kmax = 0;
        for (int i = 0; i < 4; i++) {
          P[i] = iN * q[1 + i];
          if (fabs(fabs(P[i]) - 1.0) < 1.0e-12) { kmax = i; cut = true; }
        }
        if (cut)
          for (int i = 0; i < 4; i++)
            if (i != kmax) P[i] = 0.0;
        Q1 = fabs(P0);
        S1q = sqrt(P[0] * P[0] + P[1] * P[1] + P[2] * P[2] + P[3] * P[3]);
        const double qx = iN * q[5], qy = iN * q[6];
        X1 = sqrt(qx * qx + qy * qy);
        const double N1 = fabs(iN * q[7]);
        double pk[4];
        for (int j = 0; j < 4; j++) {
          double sv = 0.0;
          for (int h = 0; h < 2; h++) {
            const double *z = q + 8 + 6 * (2 * j + h);
            sv += (1.0 / 3.0) * (z[0] * z[0] + z[1] * z[1] + z[2] * z[2] + z[3] * z[3] +
                                 z[4] * z[4] + z[5] * z[5]);
          }
          pk[j] = sqrt(a.qc_fc * sv);
        }
        for (int i = 0; i < 4; i++) {
          const double v1 = fabs(P[i]), v2 = v1 * v1;
          m[66 + i] += v1; m[70 + i] += v2; m[74 + i] += v2 * v2;
          const double w2 = pk[i] * pk[i];
          m[81 + i] += pk[i]; m[85 + i] += w2; m[89 + i] += w2 * w2;
        }
        m[78] += N1; m[79] += N1 * N1; m[80] += N1 * N1 * N1 * N1;
        // histograms (CF/mc_hm_code.cpp:5004-5028)
        int32_t *hg = a.qc_hist + (size_t)k * kQcHist;
        hg[hist_index(P0)] += 1;
        for (int i = 0; i < 4; i++) hg[kHst + 4 * hist_index(P[i]) + i] += 1;
        const double pre = fabs(P[0]) - fabs(P[2]), pim = fabs(P[1]) - fabs(P[3]);
        hg[5 * kHst + kHst * hist_index(pim) + hist_index(pre)] += 1;
      } else if (a.has_labels) {
        Q1 = stag_mag(o + 3, a.n_stag, iN);
        SH = stag_mag(o + 12, 2, iN);
        SV = stag_mag(o + 18, 2, iN);
        SD = stag_mag(o + 24, 2, iN);
        if (a.c3) {
          // cplxC3Parameter (CF/subrouts.cpp:5730-5800), z = dMag component
          const double s0 = o[5], s1 = o[8], s2 = o[11];
          const double re = sqrt(6.0) * iN * (s0 + s1 - 2.0 * s2);
          const double im = -3.0 * sqrt(2.0) * iN * (s1 - s0);
          X1 = sqrt(re * re + im * im);
        } else if (a.z2) {
          X1 = iN * o[30];
        }
      }
      const double S1 = S1q >= 0.0 ? S1q : sqrt(SH * SH + SV * SV + (a.c3 ? SD * SD : 0.0));
      const double Q2 = Q1 * Q1, S2 = S1 * S1, X2 = X1 * X1;
      m[0] += E; m[1] += E * E;
      m[2] += F1; m[3] += F2; m[4] += F2 * F2;
      m[5] += T1; m[6] += T1 * T1;
      m[7] += Q1; m[8] += Q2; m[9] += Q2 * Q2;
      m[10] += S1; m[11] += S2; m[12] += S2 * S2;
      m[13] += SH; m[14] += SV; m[15] += SD;
      m[16] += (a.z2 && !a.c3) ? fabs(X1) : X1; m[17] += X2; m[18] += X2 * X2;
      // spin stiffness: rhoS1vec summed, rhoS2vec = squared C-sums
      m[19] += o[32]; m[20] += o[33]; m[21] += o[34];
      m[22] += o[35] * o[35]; m[23] += o[36] * o[36]; m[24] += o[37] * o[37];
      if (k == 0) a.n_meas[0] += 1;
    }
  }
  // the exchange: pairs (k1, k1 + 1) of one parity are disjoint, so every
  // pair is decided by its own thread
  __syncthreads();
  if (t < a.R) a.boost_slot[t] = 0;
  __syncthreads();
  if (parity >= 0) {
    const int k1 = parity + 2 * t;
    if (k1 + 1 < a.R) {
      const int k2 = k1 + 1;
      const long long cnt1 = a.pt_cnt[k1] + 1, cnt2 = a.pt_cnt[k2] + 1;
      a.pt_cnt[k1] = cnt1;
      a.pt_cnt[k2] = cnt2;
      const double ratio = (double)(a.pt_exc[k2] + 1) * (100.0 / (double)cnt2);
      const int c1 = a.config_of_slot[k1], c2 = a.config_of_slot[k2];
      const double dE = a.energy[c2] - a.energy[c1];
      const double dB = a.beta_slot[k2] - a.beta_slot[k1];
      const double fc = dB * dE;
      const u32x4 r = draw4(a.seed, (uint32_t)k1, sweep, 0u, STREAM_PT);
      const double u = u01_f64(r.x, r.y);
      if (accept_exp(u, fc) && (!a.cap50 || ratio <= 50.0)) {
        a.config_of_slot[k1] = c2;
        a.config_of_slot[k2] = c1;
        a.slot_of_config[c2] = k1;
        a.slot_of_config[c1] = k2;
        a.pt_exc[k1] += 1;
        a.pt_exc[k2] += 1;
        a.boost_slot[k1] = 1;
        a.boost_slot[k2] = 1;
      }
    }
  }
  if (a.sweep_dev) {
    __syncthreads();
    if (t == 0) *a.sweep_dev = sweep + 1u;
  }
}

// Sweep end in replayed graphs without an exchange: clears the boost flags
// (set by the last exchange, used by the sweep just done) and advances the
// device sweep counter.
__global__ void sweep_bump_kernel(uint32_t *sweep_dev, int32_t *boost_slot, int R) {
  for (int k = threadIdx.x; k < R; k += blockDim.x) boost_slot[k] = 0;
  if (threadIdx.x == 0) *sweep_dev += 1u;
}

// ------------------------------------------------------------------ engine

struct MCEngine {
  Lattice lat;
  hsmd_mc_spec spec{};
  int R = 0;
  bool ising = false;
  int prec = HSMD_BATCHED_F32;
  bool use_lds = false;
  size_t lds_bytes = 0;
  uint32_t sweep = 0;
  bool spins_set = false;
  std::vector<double> temps;
  DevBuf<int32_t> nb, perm, bi, bj, z2nb;
  DevBuf<uint32_t> nb16;
  DevBuf<uint32_t> nbE;   // emask-slot labels of the table energy
  bool fast = false;
  // launch shape of the batched sweeps: threads per replica workgroup; the
  // register-label kernel needs every colour class to fit one pass of them
  int threads = kThreads;
  bool fast_ok = false;
  bool ifast_ok = false, ifast = false;   // fast Ising kernel possible / in use
  bool compact = false;   // fast kernels read compact labels (see SweepArgs::wcnt)
  DevBuf<uint8_t> wcnt;
  DevBuf<uint32_t> wdep, sync_err;   // wave-level ordering of the fast kernels
  int maxcol = 0;
  bool aos = false;       // FP32 fast path with float4 sites in LDS
  bool stat = false;      // ... in the statically placed LDS array (large replicas)
  size_t aos_bytes = 0;
  DevBuf<uint8_t> bc;
  DevBuf<int8_t> lab;
  DevBuf<float> Jf;
  DevBuf<double> Jd;
  DevBuf<unsigned char> spins;
  DevBuf<double> beta, energy, obs, moments, etrace;
  DevBuf<int32_t> cfg_of_slot, slot_of_cfg, boost;
  DevBuf<long long> pt_cnt, pt_exc, n_meas;
  DevBuf<double> stiff_w;
  // quasicrystal measurement
  DevBuf<double> qc_ref, qc_obs;
  DevBuf<double2> qc_phase;
  DevBuf<int32_t> qc_hist;
  // static S(q) measurement
  int sq_every = 0;
  long long meas_count = 0;   // measured sweeps since the last moments reset
  std::unique_ptr<SqPlan> sq;
  DevBuf<double> sqfield;
  long long sq_count = 0;
  SweepArgs args{};
  // captured sweep loop: one sweep (+ observables + exchange) as a graph,
  // replayed per sweep with the sweep index kept on the device
  DevBuf<uint32_t> sweep_dev;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int gkey[3] = {-9, -9, -9};
  ~MCEngine() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static void fill_coup(const Lattice &L, Coup &cp, bool ising) {
  cp.N = L.N;
  cp.Npad = L.N + 1;
  cp.Z = L.Z;
  cp.general = L.general ? 1 : 0;
  cp.h_on = L.h_on ? 1 : 0;
  cp.h = L.c.h;
  for (int s = 0; s < kZMAX; s++) { cp.jxy[s] = cp.jz[s] = 0.0; cp.lam[s] = 1.0; }
  for (int s = 0; s < L.Z; s++) {
    const int sh = L.slot_shell[s];
    cp.jxy[s] = L.shellJ[sh - 1];
    cp.jz[s] = ising ? L.shellJ[sh - 1] : L.shellJ[sh - 1] * L.shellLam[sh - 1];
    cp.lam[s] = L.shellLam[sh - 1];
  }
  finish_coup(cp);
}

void upload_lattice_common(const Lattice &L, DevBuf<int32_t> &nb, DevBuf<int32_t> &perm,
                           DevBuf<int32_t> &bi, DevBuf<int32_t> &bj, DevBuf<uint8_t> &bc,
                           DevBuf<float> &Jf, DevBuf<double> &Jd, bool f32) {
  nb.upload(L.g_nb.data(), L.g_nb.size());
  perm.upload(L.g_perm.data(), L.g_perm.size());
  bi.upload(L.g_bi.data(), L.g_bi.size());
  bj.upload(L.g_bj.data(), L.g_bj.size());
  bc.upload(L.g_bc.data(), L.g_bc.size());
  if (L.general) {
    if (f32) {
      std::vector<float> t(L.g_J.begin(), L.g_J.end());
      Jf.upload(t.data(), t.size());
    } else {
      Jd.upload(L.g_J.data(), L.g_J.size());
    }
  }
}

void fill_bonds(const Lattice &L, Bonds &bd) {
  bd.Nb = (int)L.g_bi.size();
  bd.n_class = (int)L.bc_Jxy.size();
  for (int k = 0; k < kBCMAX; k++) { bd.jxy[k] = bd.jz[k] = 0.0; }
  for (int k = 0; k < bd.n_class; k++) { bd.jxy[k] = L.bc_Jxy[k]; bd.jz[k] = L.bc_Jz[k]; }
}

// Energy form of the fast sweep kernels.  The bond-list energy
// sum_b J_b (x_i x_j + y_i y_j) + Jz_b z_i z_j is a quadratic form in the
// spins; the table form w sum_i sum_{s in F} S_i . (J_s, J_s, Jz_s) S_nb(i,s)
// is another.  F = the first half of every shell's slots with w = 1 (each
// bond met once: centrally symmetric shells of a Bravais lattice in
// descending-angle order), else all slots with w = 1/2.  A form is taken
// only if it equals the bond list's, checked on a pseudo-random
// configuration (two different forms differ there by far more than the
// rounding tolerance); otherwise the kernels keep the bond list.  Sparse
// tables (irregular graphs, where most of the Z slots of a site hold the
// sentinel) also keep the bond list, which visits only real bonds.
static void pick_energy_form(const Lattice &L, const Coup &cp, SweepArgs &a) {
  a.table_energy = 0;
  a.emask = 0u;
  a.escale = 1.0;
  if (L.general || L.Z > 32) return;
  const int N = L.N;
  size_t filled = 0;
  for (size_t k = 0; k < (size_t)L.Z * N; k++) filled += L.g_nb[k] < N ? 1 : 0;
  if (2 * filled < (size_t)L.Z * N) return;
  std::vector<double> S((size_t)3 * (N + 1), 0.0);
  uint32_t st = 0x2545f491u;
  auto rnd = [&]() { st ^= st << 13; st ^= st >> 17; st ^= st << 5; return (st >> 8) * (1.0 / 16777216.0) - 0.5; };
  for (int i = 0; i < N; i++) {
    double v[3] = {rnd(), rnd(), rnd()};
    const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + 1e-12;
    for (int d = 0; d < 3; d++) S[3 * i + d] = v[d] / n;
  }
  double eb = 0.0;
  for (size_t b = 0; b < L.g_bi.size(); b++) {
    const int i = L.g_bi[b], j = L.g_bj[b], k = L.g_bc[b];
    eb += L.bc_Jxy[k] * (S[3 * i] * S[3 * j] + S[3 * i + 1] * S[3 * j + 1]) +
          L.bc_Jz[k] * (S[3 * i + 2] * S[3 * j + 2]);
  }
  auto table = [&](uint32_t mask, double w) {
    double e = 0.0;
    for (int p = 0; p < N; p++)
      for (int s = 0; s < L.Z; s++) {
        if (!(mask & (1u << s))) continue;
        const int q = L.g_nb[(size_t)s * N + p];
        e += cp.jxy[s] * (S[3 * p] * S[3 * q] + S[3 * p + 1] * S[3 * q + 1]) +
             cp.jz[s] * (S[3 * p + 2] * S[3 * q + 2]);
      }
    return w * e;
  };
  const double tol = 1e-9 * (1.0 + fabs(eb));
  int count[4] = {0, 0, 0, 0};
  for (int s = 0; s < L.Z; s++) count[L.slot_shell[s]]++;
  uint32_t half = 0u, all = 0u;
  bool even = true;
  for (int s = 0; s < L.Z; s++) {
    all |= 1u << s;
    const int n = count[L.slot_shell[s]];
    if (n % 2) even = false;
    if (L.slot_k[s] < n / 2) half |= 1u << s;
  }
  if (even && fabs(table(half, 1.0) - eb) <= tol) {
    a.table_energy = 1; a.emask = half; a.escale = 1.0;
  } else if (fabs(table(all, 0.5) - eb) <= tol) {
    a.table_energy = 1; a.emask = all; a.escale = 0.5;
  }
}

static size_t lds_limit() { return 160 * 1024 - NOBS * 16 * 8 - (kBCMAX + 1) * 16 * 8 - 1024; }

template <int ZC>
static void set_lds_attr(MCEngine &e);

// Wave dependencies of the fast kernels: wave w updates sites [64 w, 64 w +
// 64) of each colour class, so the owner of site j is its position in its
// class over 64.  For each (class, wave) the mask holds the owners of every
// label the kernel reads for that wave's sites (the packed table, compact or
// full; the sentinel N belongs to no wave), without the wave itself.  The
// fast kernels then order the classes per wave instead of per workgroup.
// Used unless HSMD_MC_WAVE_SYNC = 0: generated lattices in site order give
// masks of the waves either side, and even the broad masks of the
// Poisson-disk graph (whose classes are sorted by degree: 14 of 16 waves)
// sweep faster than with a barrier per class.
static void build_wave_deps(MCEngine &e, const std::vector<uint32_t> &pk, int words) {
  const Lattice &lat = e.lat;
  const int N = lat.N, nc = lat.n_colours;
  SweepArgs &a = e.args;
  a.wdep = nullptr;
  a.sync_err = nullptr;
  std::vector<int> owner(N + 1, -1);
  for (int col = 0; col < nc; col++)
    for (int p = lat.col_off[col]; p < lat.col_off[col + 1]; p++) owner[p] = (p - lat.col_off[col]) >> 6;
  const uint32_t lmask = e.compact ? 0x3fffu : 0xffffu;
  std::vector<uint32_t> dep((size_t)nc * 16, 0u);
  for (int col = 0; col < nc; col++)
    for (int p = lat.col_off[col]; p < lat.col_off[col + 1]; p++) {
      const int w = owner[p];
      if (w < 0 || w > 15) return;   // a class beyond 1024 sites: no fast kernel
      uint32_t &m = dep[(size_t)col * 16 + w];
      for (int k = 0; k < words; k++) {
        const uint32_t word = pk[(size_t)k * N + p];
        const uint32_t l0 = word & lmask, l1 = (word >> 16) & lmask;
        if ((int)l0 < N && owner[l0] >= 0) m |= 1u << owner[l0];
        if ((int)l1 < N && owner[l1] >= 0) m |= 1u << owner[l1];
      }
      m &= ~(1u << w);
    }
  bool on = N > 0;
  if (const char *ev = getenv("HSMD_MC_WAVE_SYNC")) on = atoi(ev) != 0;
  if (!on) return;
  e.wdep.upload(dep.data(), dep.size());
  e.sync_err.alloc(1);
  e.sync_err.zero();
  a.wdep = e.wdep.p;
  a.sync_err = e.sync_err.p;
}

// The fast kernels' neighbour labels: (ZC + 1) / 2 words per site for the
// slot class ZC >= Z, two 16-bit labels per word, slots past Z (and holes)
// holding the sentinel N.
static std::vector<uint32_t> pack_labels16(const Lattice &lat) {
  const int N = lat.N, ZH = (lat.Z + 1) / 2, ZHC = (zc_for(lat.Z) + 1) / 2;
  std::vector<uint32_t> pk((size_t)ZHC * N, (uint32_t)N | ((uint32_t)N << 16));
  for (int k = 0; k < ZH; k++)
    for (int p = 0; p < N; p++) {
      const uint32_t lo = (uint32_t)lat.g_nb[(size_t)(2 * k) * N + p];
      const uint32_t hi = (2 * k + 1 < lat.Z) ? (uint32_t)lat.g_nb[(size_t)(2 * k + 1) * N + p] : (uint32_t)N;
      pk[(size_t)k * N + p] = lo | (hi << 16);
    }
  return pk;
}

// Energy table of the fast Ising kernel: the half-shell slots (the first half
// of each shell's slots), each slot's bond class found by its couplings (bond
// classes are keyed by (Jxy, Jz)).  Used only when the table's bonds, with
// those classes, are exactly the bond list's as a multiset: then its integer
// counts per class equal the bond list's and so does the energy.
static void ising_energy_table(MCEngine &e) {
  const Lattice &lat = e.lat;
  SweepArgs &a = e.args;
  a.itable = 0;
  for (int k = 0; k < 16; k++) a.icls[k] = -1;
  const int N = lat.N;
  int count[4] = {0, 0, 0, 0};
  for (int s = 0; s < lat.Z; s++) {
    if (lat.slot_shell[s] < 1 || lat.slot_shell[s] > 3) return;
    count[lat.slot_shell[s]]++;
  }
  std::vector<int> sl, cls;
  for (int s = 0; s < lat.Z; s++) {
    const int sh = lat.slot_shell[s], n = count[sh];
    if (n % 2) return;
    if (lat.slot_k[s] >= n / 2) continue;
    const double jxy = lat.shellJ[sh - 1], jz = lat.shellJ[sh - 1] * lat.shellLam[sh - 1];
    int k = -1;
    for (size_t c = 0; c < lat.bc_Jxy.size(); c++)
      if (lat.bc_Jxy[c] == jxy && lat.bc_Jz[c] == jz) k = (int)c;
    if (k < 0) return;
    sl.push_back(s);
    cls.push_back(k);
  }
  if (sl.empty() || sl.size() > 16) return;
  typedef std::tuple<int, int, int> Bond;
  std::vector<Bond> tb, tt;
  tb.reserve(lat.g_bi.size());
  for (size_t b = 0; b < lat.g_bi.size(); b++) {
    const int i = lat.g_bi[b], j = lat.g_bj[b];
    tb.emplace_back(std::min(i, j), std::max(i, j), (int)lat.g_bc[b]);
  }
  for (size_t k = 0; k < sl.size(); k++)
    for (int p = 0; p < N; p++) {
      const int q = lat.g_nb[(size_t)sl[k] * N + p];
      if (q < N) tt.emplace_back(std::min(p, q), std::max(p, q), cls[k]);
    }
  std::sort(tb.begin(), tb.end());
  std::sort(tt.begin(), tt.end());
  if (tb != tt) return;
  const int nw = ((int)sl.size() + 1) / 2;
  std::vector<uint32_t> pe((size_t)nw * N, (uint32_t)N | ((uint32_t)N << 16));
  for (int k = 0; k < nw; k++)
    for (int p = 0; p < N; p++) {
      const uint32_t lo = (uint32_t)lat.g_nb[(size_t)sl[2 * k] * N + p];
      const uint32_t hi = 2 * k + 1 < (int)sl.size() ? (uint32_t)lat.g_nb[(size_t)sl[2 * k + 1] * N + p]
                                                      : (uint32_t)N;
      pe[(size_t)k * N + p] = lo | (hi << 16);
    }
  for (size_t k = 0; k < cls.size(); k++) a.icls[k] = cls[k];
  e.nbE.upload(pe.data(), pe.size());
  a.nbE = e.nbE.p;
  a.nEw = nw;
  a.itable = 1;
}

MCEngine *mc_create(const Lattice &L, int R, const double *temps, const hsmd_mc_spec &sp) {
  if (R < 1 || R > kThreads) throw Error(HSMD_ERR_ARG, "replica count must be in [1, 1024]");
  if (sp.model != HSMD_HEISENBERG && sp.model != HSMD_ISING) throw Error(HSMD_ERR_ARG, "unknown model");
  if (sp.precision < 0 || sp.precision > 2) throw Error(HSMD_ERR_ARG, "unknown precision form");
  if (sp.n_or < 0 || sp.n_hb < 0 || sp.n_or_boost < 0 || sp.n_hb_boost < 0 ||
      sp.n_or + sp.n_hb > 16 || sp.n_or_boost + sp.n_hb_boost > 16)
    throw Error(HSMD_ERR_ARG, "passes per sweep must be in [0, 16]");
  for (int k = 0; k < R; k++)
    if (!(temps[k] > 0.0)) throw Error(HSMD_ERR_ARG, "temperatures must be positive");
  if (L.n_colours > kCMAX) throw Error(HSMD_ERR_ARG, "too many colours");
  std::unique_ptr<MCEngine> e(new MCEngine());
  e->lat = L;
  e->spec = sp;
  e->R = R;
  e->ising = sp.model == HSMD_ISING;
  e->prec = sp.precision;
  e->temps.assign(temps, temps + R);
  const Lattice &lat = e->lat;
  const bool f32 = (e->prec == HSMD_BATCHED_F32) && !e->ising;
  upload_lattice_common(lat, e->nb, e->perm, e->bi, e->bj, e->bc, e->Jf, e->Jd, f32);
  if (e->ising && lat.general) e->Jd.upload(lat.g_J.data(), lat.g_J.size());
  // labels in new order
  const int N = lat.N;
  if (lat.labels.present) {
    std::vector<int8_t> lb((size_t)4 * N);
    for (int p = 0; p < N; p++) {
      const int i = lat.perm[p];
      lb[p] = lat.labels.stag[i];
      lb[N + p] = lat.labels.hs[i];
      lb[2 * N + p] = lat.labels.vs[i];
      lb[3 * N + p] = lat.labels.ds[i];
    }
    e->lab.upload(lb.data(), lb.size());
    if (lat.labels.z2) {
      std::vector<int32_t> z((size_t)3 * N);
      for (int p = 0; p < N; p++) {
        const int i = lat.perm[p];
        for (int r = 0; r < 3; r++) {
          const int q = lat.labels.z2nb[(size_t)r * N + i];
          z[(size_t)r * N + p] = q >= 0 ? lat.inv[q] : N;
        }
      }
      e->z2nb.upload(z.data(), z.size());
    }
  }
  const size_t Npad = (size_t)N + 1;
  size_t elem = e->ising ? 1 : (f32 ? 3 * sizeof(float) : 3 * sizeof(double));
  e->spins.alloc(Npad * elem * R);
  e->spins.zero();
  e->lds_bytes = Npad * elem;
  e->use_lds = (e->prec != HSMD_PLAIN_F64) && e->lds_bytes <= lds_limit();
  std::vector<double> b(R);
  for (int k = 0; k < R; k++) b[k] = 1.0 / temps[k];
  e->beta.upload(b.data(), R);
  e->energy.alloc(R); e->energy.zero();
  e->obs.alloc((size_t)R * NOBSX); e->obs.zero();
  {
    // stiffness weights per bond (new labels, reference bond order)
    const double c60 = 0.5, s60 = sqrt(3.0) / 2.0;
    const size_t Nb = lat.bi.size();
    std::vector<double> w(6 * std::max<size_t>(Nb, 1), 0.0);
    if (lat.rvec.size() == (size_t)2 * N) {
      for (size_t b = 0; b < Nb; b++) {
        const int i = lat.bi[b], j = lat.bj[b];
        const double dx = lat.rvec[2 * i] - lat.rvec[2 * j];
        const double dy = lat.rvec[2 * i + 1] - lat.rvec[2 * j + 1];
        const double C[3] = {dx, s60 * dy + c60 * dx, s60 * dy - c60 * dx};
        const double J = lat.stiffJ(lat.bshell[b], i, j);
        for (int k = 0; k < 3; k++) {
          w[k * Nb + b] = J * C[k] * C[k];
          w[(3 + k) * Nb + b] = J * C[k];
        }
      }
    }
    e->stiff_w.upload(w.data(), w.size());
  }
  e->moments.alloc((size_t)R * NMOM); e->moments.zero();
  if (lat.qc_on()) {
    // reference states and hotspot phases in the engine's labels
    std::vector<double> ref((size_t)15 * N);
    std::vector<double2> ph((size_t)8 * N);
    const double dq = lat.sq_dq();
    for (int p = 0; p < N; p++) {
      const int i = lat.perm[p];
      for (int r = 0; r < 5; r++)
        for (int d = 0; d < 3; d++) ref[(size_t)(3 * r + d) * N + p] = lat.qc_ref[((size_t)r * N + i) * 3 + d];
      for (int q = 0; q < 8; q++) {
        const double th = -dq * (lat.qc_nx[q] * lat.rvec[2 * (size_t)i] + lat.qc_ny[q] * lat.rvec[2 * (size_t)i + 1]);
        ph[(size_t)q * N + p] = make_double2(cos(th), sin(th));
      }
    }
    e->qc_ref.upload(ref.data(), ref.size());
    e->qc_phase.upload(ph.data(), ph.size());
    e->qc_obs.alloc((size_t)R * NQC);
    e->qc_obs.zero();
    e->qc_hist.alloc((size_t)R * kQcHist);
    e->qc_hist.zero();
  }
  std::vector<int32_t> id(R);
  for (int k = 0; k < R; k++) id[k] = k;
  e->cfg_of_slot.upload(id.data(), R);
  e->slot_of_cfg.upload(id.data(), R);
  e->boost.alloc(R); e->boost.zero();
  e->pt_cnt.alloc(R); e->pt_cnt.zero();
  e->pt_exc.alloc(R); e->pt_exc.zero();
  e->n_meas.alloc(1); e->n_meas.zero();
  // kernel arguments
  SweepArgs &a = e->args;
  fill_coup(lat, a.cp, e->ising);
  fill_bonds(lat, a.bd);
  a.n_colours = lat.n_colours;
  for (int k = 0; k <= lat.n_colours; k++) a.col_off[k] = lat.col_off[k];
  a.n_or = sp.n_or; a.n_hb = sp.n_hb; a.n_or_b = sp.n_or_boost; a.n_hb_b = sp.n_hb_boost;
  a.metro_passes = 1;
  a.seed = sp.seed;
  a.n_stag = lat.labels.n_stag;
  a.has_labels = lat.labels.present ? 1 : 0;
  a.z2 = lat.labels.z2 ? 1 : 0;
  a.c3 = lat.labels.c3 ? 1 : 0;
  a.nb = e->nb.p;
  a.Jt = f32 ? (const void *)e->Jf.p : (const void *)e->Jd.p;
  a.perm = e->perm.p;
  a.bi = e->bi.p; a.bj = e->bj.p; a.bc = e->bc.p;
  a.lab = e->lab.p;
  a.z2nb = e->z2nb.p;
  a.spins = e->spins.p;
  a.beta_slot = e->beta.p;
  a.slot_of_config = e->slot_of_cfg.p;
  a.boost_slot = e->boost.p;
  a.energy = e->energy.p;
  a.obs = e->obs.p;
  a.stiff_w = e->stiff_w.p;
  a.qc = lat.qc_on() ? 1 : 0;
  a.qc_zn1 = lat.zn1;
  a.qc_ref = e->qc_ref.p;
  a.qc_phase = e->qc_phase.p;
  a.qc_obs = e->qc_obs.p;
  {
    int maxcol = 0;
    for (int k = 0; k < lat.n_colours; k++) maxcol = std::max(maxcol, lat.col_off[k + 1] - lat.col_off[k]);
    e->maxcol = maxcol;
    e->fast_ok = e->use_lds && !e->ising && !lat.general && (lat.N + 1) < 65535 && maxcol <= kThreads;
    e->fast = e->fast_ok;
    e->aos_bytes = (size_t)(N + 1) * sizeof(float4);
    e->aos = e->fast_ok && f32 && e->aos_bytes <= lds_limit();
    e->stat = e->aos && N + 1 <= kStatSites && e->aos_bytes > 64 * 1024;
    if (e->fast_ok) {
      // (ZC + 1) / 2 words per site for the kernel's slot class ZC >= Z;
      // slots past Z hold the sentinel N on both halves
      std::vector<uint32_t> pk = pack_labels16(lat);
      // sparse tables (irregular graphs, where most slots hold the
      // sentinel): compact labels, so a wave visits only as many label words
      // as its sites' largest neighbour count needs
      size_t filled = 0;
      for (size_t k = 0; k < (size_t)lat.Z * N; k++) filled += lat.g_nb[k] < N ? 1 : 0;
      bool cp = !lat.general && N + 1 <= 16384 && 2 * filled < (size_t)lat.Z * N;
      double shj[3] = {0, 0, 0}, shjz[3] = {0, 0, 0};
      bool seen[3] = {false, false, false};
      for (int sl = 0; sl < lat.Z && cp; sl++) {
        const int sh = lat.slot_shell[sl] - 1;
        if (sh < 0 || sh > 2) { cp = false; break; }
        if (!seen[sh]) {
          seen[sh] = true; shj[sh] = a.cp.jxy[sl]; shjz[sh] = a.cp.jz[sl];
        } else if (shj[sh] != a.cp.jxy[sl] || shjz[sh] != a.cp.jz[sl]) {
          cp = false;
        }
      }
      if (cp) {
        std::fill(pk.begin(), pk.end(), (uint32_t)N | ((uint32_t)N << 16));
        std::vector<int> deg(N, 0);
        for (int p = 0; p < N; p++) {
          int n = 0;
          for (int sl = 0; sl < lat.Z; sl++) {
            const int q = lat.g_nb[(size_t)sl * N + p];
            if (q >= N) continue;
            const uint32_t ent = (uint32_t)q | ((uint32_t)(lat.slot_shell[sl] - 1) << 14);
            uint32_t &w = pk[(size_t)(n / 2) * N + p];
            w = (n % 2) ? ((w & 0xffffu) | (ent << 16)) : ((w & 0xffff0000u) | ent);
            n++;
          }
          deg[p] = n;
        }
        std::vector<uint8_t> wc((size_t)lat.n_colours * 16, 0);
        for (int col = 0; col < lat.n_colours; col++)
          for (int p = lat.col_off[col]; p < lat.col_off[col + 1]; p++) {
            const int ch = (p - lat.col_off[col]) / 64;
            uint8_t &w = wc[(size_t)col * 16 + ch];
            w = std::max<uint8_t>(w, (uint8_t)((deg[p] + 1) / 2));
          }
        e->wcnt.upload(wc.data(), wc.size());
        a.wcnt = e->wcnt.p;
        for (int k = 0; k < 3; k++) {
          a.shj[k] = shj[k]; a.shjz[k] = shjz[k];
          a.fshj[k] = (float)shj[k]; a.fshjz[k] = (float)shjz[k];
        }
        e->compact = true;
      }
      e->nb16.upload(pk.data(), pk.size());
      a.nb16 = e->nb16.p;
      build_wave_deps(*e, pk, (zc_for(lat.Z) + 1) / 2);
      if (!e->compact) pick_energy_form(lat, a.cp, a);
      a.nbE = nullptr;
      a.nEw = 0;
      for (int k = 0; k < 16; k++) a.ejxy[k] = a.ejz[k] = 0.0;
      if (a.table_energy && __builtin_popcount(a.emask) <= 16) {
        // the emask slots' labels in slot order, two per word, sentinel-padded
        std::vector<int> sl;
        for (int s = 0; s < lat.Z; s++)
          if (a.emask & (1u << s)) sl.push_back(s);
        const int nw = ((int)sl.size() + 1) / 2;
        std::vector<uint32_t> pe((size_t)nw * N, (uint32_t)N | ((uint32_t)N << 16));
        for (int k = 0; k < nw; k++)
          for (int p = 0; p < N; p++) {
            const uint32_t lo = (uint32_t)lat.g_nb[(size_t)sl[2 * k] * N + p];
            const uint32_t hi = 2 * k + 1 < (int)sl.size() ? (uint32_t)lat.g_nb[(size_t)sl[2 * k + 1] * N + p]
                                                            : (uint32_t)N;
            pe[(size_t)k * N + p] = lo | (hi << 16);
          }
        for (size_t k = 0; k < sl.size(); k++) {
          a.ejxy[k] = a.cp.jxy[sl[k]];
          a.ejz[k] = a.cp.jz[sl[k]];
        }
        e->nbE.upload(pe.data(), pe.size());
        a.nbE = e->nbE.p;
        a.nEw = nw;
      }
    }
  }
  {
    // fast Ising kernel: register labels, wave ordering, table energy
    e->ifast_ok = e->use_lds && e->ising && !lat.general && (lat.N + 1) < 65535 && e->maxcol <= kThreads;
    e->ifast = e->ifast_ok;
    a.itable = 0;
    if (e->ifast_ok) {
      std::vector<uint32_t> pk = pack_labels16(lat);
      e->nb16.upload(pk.data(), pk.size());
      a.nb16 = e->nb16.p;
      build_wave_deps(*e, pk, (zc_for(lat.Z) + 1) / 2);
      ising_energy_table(*e);
    }
  }
  if (e->use_lds) HSMD_ZC_DISPATCH(lat.Z, set_lds_attr, *e);
  return e.release();
}

void mc_set_spins(MCEngine &e, const double *spins) {
  const Lattice &L = e.lat;
  const int N = L.N, R = e.R;
  const size_t Npad = (size_t)N + 1;
  const bool f32 = e.prec == HSMD_BATCHED_F32;
  std::vector<int32_t> id(R);
  for (int k = 0; k < R; k++) id[k] = k;
  HIP_CHECK(hipMemcpy(e.cfg_of_slot.p, id.data(), R * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(e.slot_of_cfg.p, id.data(), R * sizeof(int32_t), hipMemcpyHostToDevice));
  if (e.ising) {
    std::vector<int8_t> h(Npad * R, 0);
    for (int r = 0; r < R; r++)
      for (int i = 0; i < N; i++)
        h[r * Npad + L.inv[i]] = spins[((size_t)r * N + i) * 3] < 0.0 ? -1 : 1;
    HIP_CHECK(hipMemcpy(e.spins.p, h.data(), h.size(), hipMemcpyHostToDevice));
  } else if (f32) {
    std::vector<float> h(3 * Npad * R, 0.f);
    for (int r = 0; r < R; r++)
      for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++)
          h[(r * 3 + d) * Npad + L.inv[i]] = (float)spins[((size_t)r * N + i) * 3 + d];
    HIP_CHECK(hipMemcpy(e.spins.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  } else {
    std::vector<double> h(3 * Npad * R, 0.0);
    for (int r = 0; r < R; r++)
      for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++)
          h[(r * 3 + d) * Npad + L.inv[i]] = spins[((size_t)r * N + i) * 3 + d];
    HIP_CHECK(hipMemcpy(e.spins.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  e.spins_set = true;
}

// a fast-kernel wave wait that ran out (never expected: the waits end
// once every wave has written its classes) makes the results invalid
static void check_wave_sync(MCEngine &e) {
  if (!e.args.sync_err) return;
  uint32_t f = 0;
  HIP_CHECK(hipMemcpy(&f, e.sync_err.p, sizeof f, hipMemcpyDeviceToHost));
  if (f) throw Error(HSMD_ERR_DEVICE, "sweep kernel: a wave dependency wait ran out");
}

void mc_get_spins(MCEngine &e, double *spins) {
  check_wave_sync(e);
  const Lattice &L = e.lat;
  const int N = L.N, R = e.R;
  const size_t Npad = (size_t)N + 1;
  const bool f32 = e.prec == HSMD_BATCHED_F32;
  std::vector<int32_t> cos(R);
  HIP_CHECK(hipMemcpy(cos.data(), e.cfg_of_slot.p, R * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (e.ising) {
    std::vector<int8_t> h(Npad * R);
    HIP_CHECK(hipMemcpy(h.data(), e.spins.p, h.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < R; k++)
      for (int i = 0; i < N; i++) {
        double *o = spins + ((size_t)k * N + i) * 3;
        o[0] = (double)h[cos[k] * Npad + L.inv[i]];
        o[1] = o[2] = 0.0;
      }
  } else if (f32) {
    std::vector<float> h(3 * Npad * R);
    HIP_CHECK(hipMemcpy(h.data(), e.spins.p, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int k = 0; k < R; k++)
      for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++)
          spins[((size_t)k * N + i) * 3 + d] = h[(cos[k] * 3 + d) * Npad + L.inv[i]];
  } else {
    std::vector<double> h(3 * Npad * R);
    HIP_CHECK(hipMemcpy(h.data(), e.spins.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < R; k++)
      for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++)
          spins[((size_t)k * N + i) * 3 + d] = h[(cos[k] * 3 + d) * Npad + L.inv[i]];
  }
}

template <int ZC>
static void launch_sweep_zc(MCEngine &e, hipStream_t st) {
  SweepArgs &a = e.args;
  const bool f32 = e.prec == HSMD_BATCHED_F32;
  if (e.prec == HSMD_PLAIN_F64) {
    // one instance per launch, FP64, spins in global memory
    for (int c = 0; c < e.R; c++) {
      a.config_base = c;
      if (e.ising) hipLaunchKernelGGL((ising_sweep_kernel<false, ZC>), dim3(1), dim3(kThreads), 0, st, a);
      else hipLaunchKernelGGL((heis_sweep_kernel<double, false, ZC>), dim3(1), dim3(kThreads), 0, st, a);
    }
    a.config_base = 0;
    return;
  }
  a.config_base = 0;
  const size_t sh = e.use_lds ? e.lds_bytes : 0;
  if (e.ising) {
    if (e.ifast) hipLaunchKernelGGL((ising_sweep_fast<ZC>), dim3(e.R), dim3(e.threads), sh, st, a);
    else if (e.use_lds) hipLaunchKernelGGL((ising_sweep_kernel<true, ZC>), dim3(e.R), dim3(e.threads), sh, st, a);
    else hipLaunchKernelGGL((ising_sweep_kernel<false, ZC>), dim3(e.R), dim3(e.threads), 0, st, a);
  } else if (e.fast && e.compact && a.n_or + a.n_hb + a.n_or_b + a.n_hb_b > 0) {
    if (f32 && e.aos)
      hipLaunchKernelGGL((heis_sweep_fast<float, ZC, true, true>), dim3(e.R), dim3(e.threads), e.aos_bytes, st, a);
    else if (f32) hipLaunchKernelGGL((heis_sweep_fast<float, ZC, false, true>), dim3(e.R), dim3(e.threads), sh, st, a);
    else hipLaunchKernelGGL((heis_sweep_fast<double, ZC, false, true>), dim3(e.R), dim3(e.threads), sh, st, a);
  } else if (e.fast && a.n_or + a.n_hb + a.n_or_b + a.n_hb_b > 0) {
    if (f32 && e.aos && e.stat)
      hipLaunchKernelGGL((heis_sweep_fast<float, ZC, true, false, true>), dim3(e.R), dim3(e.threads), 0, st, a);
    else if (f32 && e.aos)
      hipLaunchKernelGGL((heis_sweep_fast<float, ZC, true>), dim3(e.R), dim3(e.threads), e.aos_bytes, st, a);
    else if (f32) hipLaunchKernelGGL((heis_sweep_fast<float, ZC, false>), dim3(e.R), dim3(e.threads), sh, st, a);
    else hipLaunchKernelGGL((heis_sweep_fast<double, ZC, false>), dim3(e.R), dim3(e.threads), sh, st, a);
  } else if (f32) {
    if (e.use_lds) hipLaunchKernelGGL((heis_sweep_kernel<float, true, ZC>), dim3(e.R), dim3(e.threads), sh, st, a);
    else hipLaunchKernelGGL((heis_sweep_kernel<float, false, ZC>), dim3(e.R), dim3(e.threads), 0, st, a);
  } else {
    if (e.use_lds) hipLaunchKernelGGL((heis_sweep_kernel<double, true, ZC>), dim3(e.R), dim3(e.threads), sh, st, a);
    else hipLaunchKernelGGL((heis_sweep_kernel<double, false, ZC>), dim3(e.R), dim3(e.threads), 0, st, a);
  }
}

template <int ZC>
static void set_lds_attr(MCEngine &e) {
  const bool f32 = e.prec == HSMD_BATCHED_F32;
  const void *fn;
  if (e.ising) fn = (const void *)ising_sweep_kernel<true, ZC>;
  else if (f32) fn = (const void *)heis_sweep_kernel<float, true, ZC>;
  else fn = (const void *)heis_sweep_kernel<double, true, ZC>;
  HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.lds_bytes));
  if (e.ising) fn = (const void *)obs_kernel<int8_t, true>;
  else if (f32) fn = (const void *)obs_kernel<float, true>;
  else fn = (const void *)obs_kernel<double, true>;
  HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.lds_bytes));
  if (e.fast_ok) {
    fn = f32 ? (const void *)heis_sweep_fast<float, ZC, false> : (const void *)heis_sweep_fast<double, ZC, false>;
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.lds_bytes));
    fn = f32 ? (const void *)heis_sweep_fast<float, ZC, false, true>
             : (const void *)heis_sweep_fast<double, ZC, false, true>;
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.lds_bytes));
  }
  if (e.aos) {
    fn = (const void *)heis_sweep_fast<float, ZC, true>;
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.aos_bytes));
    fn = (const void *)heis_sweep_fast<float, ZC, true, true>;
    HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.aos_bytes));
  }
}

static void launch_sweep(MCEngine &e, hipStream_t st) {
  const SweepArgs &a = e.args;
  if (e.stat && e.fast && !e.compact && a.n_or + a.n_hb + a.n_or_b + a.n_hb_b > 0 &&
      e.lat.Z > 16 && e.lat.Z <= 18) {
    // eighteen coupled slots (triangular lattice with three shells): the
    // statically placed kernel with exactly nine label words per site
    e.args.config_base = 0;
    if (a.table_energy && a.nEw == 5)   // nine emask slots: the half-shell table energy
      hipLaunchKernelGGL((heis_sweep_fast<float, 18, true, false, true, 5>), dim3(e.R), dim3(e.threads), 0,
                         st, e.args);
    else
      hipLaunchKernelGGL((heis_sweep_fast<float, 18, true, false, true>), dim3(e.R), dim3(e.threads), 0,
                         st, e.args);
    return;
  }
  HSMD_ZC_DISPATCH(e.lat.Z, launch_sweep_zc, e, st);
}

// The sweep kernel launch_sweep starts for the engine's current form, as
// "<kernel><template arguments> x <threads>", and " wave-ordered" when the
// fast kernel orders its classes per wave instead of per workgroup.
std::string mc_sweep_kernel(const MCEngine &e) {
  const SweepArgs &a = e.args;
  const bool f32 = e.prec == HSMD_BATCHED_F32;
  const int ZC = zc_for(e.lat.Z);
  const bool passes = a.n_or + a.n_hb + a.n_or_b + a.n_hb_b > 0;
  const char *T = f32 ? "float" : "double";
  char b[160];
  if (e.prec == HSMD_PLAIN_F64) {
    snprintf(b, sizeof b, "%s<%s%d> x %d, one replica per launch",
             e.ising ? "ising_sweep_kernel" : "heis_sweep_kernel", e.ising ? "false," : "double,false,", ZC,
             kThreads);
    return b;
  }
  bool fast = false;
  if (e.ising && e.ifast) {
    snprintf(b, sizeof b, "ising_sweep_fast<%d> x %d%s", ZC, e.threads,
             a.itable ? ", table energy" : ", bond-list energy");
    fast = true;
  } else if (e.ising) {
    snprintf(b, sizeof b, "ising_sweep_kernel<%s,%d> x %d", e.use_lds ? "true" : "false", ZC, e.threads);
  } else if (e.stat && e.fast && !e.compact && passes && e.lat.Z > 16 && e.lat.Z <= 18) {
    snprintf(b, sizeof b, "heis_sweep_fast<float,18,true,false,true%s> x %d",
             a.table_energy && a.nEw == 5 ? ",5" : "", e.threads);
    fast = true;
  } else if (e.fast && passes) {
    snprintf(b, sizeof b, "heis_sweep_fast<%s,%d,%s,%s> x %d", T, ZC, f32 && e.aos ? "true" : "false",
             e.compact ? "true" : "false", e.threads);
    fast = true;
  } else {
    snprintf(b, sizeof b, "heis_sweep_kernel<%s,%s,%d> x %d", T, e.use_lds ? "true" : "false", ZC, e.threads);
  }
  std::string r = b;
  if (fast && a.wdep) r += " wave-ordered";
  return r;
}

static void launch_obs(MCEngine &e, hipStream_t st) {
  const size_t sh = e.use_lds ? e.lds_bytes : 0;
  if (e.ising) {
    if (e.use_lds) hipLaunchKernelGGL((obs_kernel<int8_t, true>), dim3(e.R), dim3(kObsThreads), sh, st, e.args);
    else hipLaunchKernelGGL((obs_kernel<int8_t, false>), dim3(e.R), dim3(kObsThreads), 0, st, e.args);
  } else if (e.prec == HSMD_BATCHED_F32) {
    if (e.use_lds) hipLaunchKernelGGL((obs_kernel<float, true>), dim3(e.R), dim3(kObsThreads), sh, st, e.args);
    else hipLaunchKernelGGL((obs_kernel<float, false>), dim3(e.R), dim3(kObsThreads), 0, st, e.args);
  } else {
    if (e.use_lds) hipLaunchKernelGGL((obs_kernel<double, true>), dim3(e.R), dim3(kObsThreads), sh, st, e.args);
    else hipLaunchKernelGGL((obs_kernel<double, false>), dim3(e.R), dim3(kObsThreads), 0, st, e.args);
  }
  if (e.args.qc) {
    if (e.ising) hipLaunchKernelGGL(qc_obs_kernel<int8_t>, dim3(e.R), dim3(256), 0, st, e.args);
    else if (e.prec == HSMD_BATCHED_F32)
      hipLaunchKernelGGL(qc_obs_kernel<float>, dim3(e.R), dim3(256), 0, st, e.args);
    else hipLaunchKernelGGL(qc_obs_kernel<double>, dim3(e.R), dim3(256), 0, st, e.args);
  }
}

// Sweep loop as a replayed hipGraph: the graph of one sweep (sweep kernel,
// observables when measuring, exchange or a counter bump) is captured once
// per (measure, parity mode, exchange) combination and launched n times; the
// kernels read the sweep index (Philox counter, pair parity) from the device
// counter the exchange kernel advances.  Same arithmetic and counters as the
// launch-by-launch loop, which remains for energy traces, S(q) measurement
// and the plain form.  No memset is recorded in the graph.
static void run_graph(MCEngine &e, PTArgs p, int n, int pt_parity, int measure, bool post,
                      int pthreads) {
  if (!e.stream) HIP_CHECK(hipStreamCreate(&e.stream));
  if (e.sweep_dev.n == 0) e.sweep_dev.alloc(1);
  HIP_CHECK(hipMemcpy(e.sweep_dev.p, &e.sweep, sizeof(uint32_t), hipMemcpyHostToDevice));
  const int key[3] = {measure, pt_parity, post ? 1 : 0};
  if (!e.gexec || memcmp(key, e.gkey, sizeof key) != 0) {
    if (e.gexec) { HIP_CHECK(hipGraphExecDestroy(e.gexec)); e.gexec = nullptr; }
    if (e.graph) { HIP_CHECK(hipGraphDestroy(e.graph)); e.graph = nullptr; }
    const SweepArgs saved = e.args;
    e.args.sweep_dev = e.sweep_dev.p;
    e.args.measure = measure;
    HIP_CHECK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
    launch_sweep(e, e.stream);
    if (measure) launch_obs(e, e.stream);
    if (post) {
      p.sweep_dev = e.sweep_dev.p;
      p.parity_alt = pt_parity == 2 ? 1 : 0;
      p.parity = pt_parity == 2 ? 0 : pt_parity;
      p.etrace = nullptr;
      hipLaunchKernelGGL(measure_exchange_kernel, dim3(1), dim3(pthreads), 0, e.stream, p);
    } else {
      hipLaunchKernelGGL(sweep_bump_kernel, dim3(1), dim3(256), 0, e.stream, e.sweep_dev.p,
                       e.boost.p, e.R);
    }
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e.stream, &g);
    e.args = saved;
    HIP_CHECK(ce);
    e.graph = g;
    HIP_CHECK(hipGraphInstantiate(&e.gexec, g, nullptr, nullptr, 0));
    memcpy(e.gkey, key, sizeof key);
  }
  for (int s = 0; s < n; s++) HIP_CHECK(hipGraphLaunch(e.gexec, e.stream));
  HIP_CHECK(hipStreamSynchronize(e.stream));
  HIP_CHECK(hipGetLastError());
  e.sweep += (uint32_t)n;
  if (measure) e.meas_count += n;
}

void mc_run(MCEngine &e, int n_sweeps, int pt_parity, int measure, double *etrace_out) {
  if (!e.spins_set) throw Error(HSMD_ERR_STATE, "hsmd_mc_run before hsmd_mc_set_spins");
  if (n_sweeps < 0) throw Error(HSMD_ERR_ARG, "n_sweeps must be >= 0");
  if (pt_parity < -1 || pt_parity > 2) throw Error(HSMD_ERR_ARG, "pt_parity must be -1, 0, 1 or 2");
  if (etrace_out && n_sweeps > 0) e.etrace.alloc((size_t)n_sweeps * e.R);
  PTArgs p{};
  p.R = e.R; p.N = e.lat.N;
  p.cap50 = e.spec.pt_cap50;
  p.measure = measure;
  p.n_stag = e.args.n_stag; p.has_labels = e.args.has_labels; p.z2 = e.args.z2; p.c3 = e.args.c3;
  p.h_on = e.lat.h_on ? 1 : 0;
  p.ising = e.ising ? 1 : 0;
  p.seed = e.spec.seed;
  p.energy = e.energy.p; p.obs = e.obs.p; p.beta_slot = e.beta.p;
  p.config_of_slot = e.cfg_of_slot.p; p.slot_of_config = e.slot_of_cfg.p;
  p.boost_slot = e.boost.p; p.pt_cnt = e.pt_cnt.p; p.pt_exc = e.pt_exc.p;
  p.moments = e.moments.p; p.n_meas = e.n_meas.p;
  p.qc = e.args.qc;
  p.qc_fc = 0.5 / (double)e.lat.Nq;
  p.qc_obs = e.qc_obs.p;
  p.qc_hist = e.qc_hist.p;
  e.args.measure = measure;
  const int pthreads = ((e.R + 63) / 64) * 64;
  // the measurement / exchange kernel is skipped when it has nothing to do
  const bool post = pt_parity >= 0 || measure || etrace_out;
  if (e.prec != HSMD_PLAIN_F64 && !etrace_out && !(measure && e.sq_every > 0) && n_sweeps > 1) {
    run_graph(e, p, n_sweeps, pt_parity, measure, post, pthreads);
    return;
  }
  for (int s = 0; s < n_sweeps; s++) {
    e.args.sweep = e.sweep;
    launch_sweep(e, 0);
    if (measure) launch_obs(e, 0);
    if (measure && e.sq_every > 0 && e.meas_count % e.sq_every == 0) {
      const SpinType ty = e.ising ? SPIN_I8 : (e.prec == HSMD_BATCHED_F32 ? SPIN_F32 : SPIN_F64);
      sq_compute(*e.sq, e.spins.p, ty, e.R);
      sq_accumulate(*e.sq, e.R, e.cfg_of_slot.p, e.sqfield.p, e.moments.p, NMOM, 25);
      e.sq_count++;
    }
    if (measure) e.meas_count++;
    if (post) {
      p.sweep = e.sweep;
      p.parity = pt_parity == 2 ? (int)(e.sweep % 2u) : pt_parity;
      p.etrace = etrace_out ? e.etrace.p + (size_t)s * e.R : nullptr;
      hipLaunchKernelGGL(measure_exchange_kernel, dim3(1), dim3(pthreads), 0, 0, p);
    } else if (s == 0) {
      e.boost.zero();
    }
    e.sweep++;
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  if (etrace_out && n_sweeps > 0)
    HIP_CHECK(hipMemcpy(etrace_out, e.etrace.p, (size_t)n_sweeps * e.R * sizeof(double),
                        hipMemcpyDeviceToHost));
}

void mc_energies(MCEngine &e, double *E) {
  check_wave_sync(e);
  std::vector<double> ec(e.R);
  std::vector<int32_t> cos(e.R);
  HIP_CHECK(hipMemcpy(ec.data(), e.energy.p, e.R * sizeof(double), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(cos.data(), e.cfg_of_slot.p, e.R * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int k = 0; k < e.R; k++) E[k] = ec[cos[k]];
}

void mc_refresh_energies(MCEngine &e) {
  // zero-pass sweep: recompute energies of the current configurations
  SweepArgs a = e.args;
  a.n_or = a.n_hb = a.n_or_b = a.n_hb_b = 0;
  a.metro_passes = 0;
  a.measure = 0;
  std::swap(a, e.args);
  try {
    launch_sweep(e, 0);
  } catch (...) {
    std::swap(a, e.args);
    throw;
  }
  std::swap(a, e.args);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
}

void mc_pt_counters(MCEngine &e, int64_t *att, int64_t *acc) {
  std::vector<long long> a(e.R), b(e.R);
  HIP_CHECK(hipMemcpy(a.data(), e.pt_cnt.p, e.R * sizeof(long long), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(b.data(), e.pt_exc.p, e.R * sizeof(long long), hipMemcpyDeviceToHost));
  for (int k = 0; k < e.R; k++) { att[k] = a[k]; acc[k] = b[k]; }
}

void mc_reset_counters(MCEngine &e) {
  e.pt_cnt.zero();
  e.pt_exc.zero();
  e.boost.zero();
  e.sweep = 0;
}

void mc_moments(MCEngine &e, double *out, int64_t *n, int reset) {
  HIP_CHECK(hipMemcpy(out, e.moments.p, (size_t)e.R * NMOM * sizeof(double), hipMemcpyDeviceToHost));
  long long nm = 0;
  HIP_CHECK(hipMemcpy(&nm, e.n_meas.p, sizeof(long long), hipMemcpyDeviceToHost));
  if (n) *n = nm;
  if (reset) { e.moments.zero(); e.n_meas.zero(); e.meas_count = 0; }
}

void mc_set_sq(MCEngine &e, int every) {
  if (every < 0) throw Error(HSMD_ERR_ARG, "S(q) interval must be >= 0");
  if (every > 0 && !e.sq) {
    std::unique_ptr<SqPlan> p(new SqPlan());
    sq_plan_init(*p, e.lat, e.R);
    e.sqfield.alloc((size_t)e.R * p->Nq * 6);
    e.sqfield.zero();
    e.sq = std::move(p);
    e.sq_count = 0;
  }
  e.sq_every = every;
}

void mc_sq_field(MCEngine &e, double *out, int64_t *count, int reset) {
  if (!e.sq) throw Error(HSMD_ERR_STATE, "S(q) measurement was never enabled");
  HIP_CHECK(hipDeviceSynchronize());
  if (out)
    HIP_CHECK(hipMemcpy(out, e.sqfield.p, (size_t)e.R * e.sq->Nq * 6 * sizeof(double),
                        hipMemcpyDeviceToHost));
  if (count) *count = e.sq_count;
  if (reset) { e.sqfield.zero(); e.sq_count = 0; }
}

void mc_qc_histograms(MCEngine &e, int32_t *out, int reset) {
  if (!e.args.qc) throw Error(HSMD_ERR_STATE, "the lattice has no quasicrystal reference states");
  HIP_CHECK(hipMemcpy(out, e.qc_hist.p, (size_t)e.R * kQcHist * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (reset) e.qc_hist.zero();
}

McHandoff mc_handoff(MCEngine &e) {
  if (!e.spins_set) throw Error(HSMD_ERR_STATE, "the MC engine has no spins yet");
  McHandoff h;
  h.spins = e.spins.p;
  h.type = e.ising ? 0 : (e.prec == HSMD_BATCHED_F32 ? 1 : 2);
  h.R = e.R;
  h.cfg_of_slot = e.cfg_of_slot.p;
  h.lat = &e.lat;
  return h;
}

void mc_slot_map(MCEngine &e, int32_t *m) {
  HIP_CHECK(hipMemcpy(m, e.cfg_of_slot.p, e.R * sizeof(int32_t), hipMemcpyDeviceToHost));
}

void mc_destroy(MCEngine *e) { delete e; }

McBondsDev mc_bonds_dev(const MCEngine &e) {
  McBondsDev b;
  b.bi = e.bi.p;
  b.bj = e.bj.p;
  b.bc = e.bc.p;
  b.perm = e.perm.p;
  b.N = e.lat.N;
  b.bd = e.args.bd;
  b.h_on = e.lat.h_on ? 1 : 0;
  b.h = e.lat.c.h;
  return b;
}

// Sweep-kernel timing hook for the benchmark: launches n sweeps of every
// replica back to back on stream st (no exchange), without synchronising.
void mc_launch_sweeps(MCEngine &e, int n, hipStream_t st) {
  for (int s = 0; s < n; s++) {
    e.args.sweep = e.sweep++;
    e.args.measure = 0;
    launch_sweep(e, st);
  }
}

// Threads per replica workgroup of the batched sweeps (a multiple of 64 in
// [64, 1024]).  Per-site arithmetic does not depend on it; block reductions
// (energies) group their terms by wave, so FP64 energies can differ in the
// last bits between shapes.
void mc_set_threads(MCEngine &e, int threads) {
  if (threads < 64 || threads > kThreads || threads % 64 != 0)
    throw Error(HSMD_ERR_ARG, "threads per workgroup must be a multiple of 64 in [64, 1024]");
  e.threads = threads;
  e.fast = e.fast_ok && e.maxcol <= threads;
  e.ifast = e.ifast_ok && e.maxcol <= threads;
  if (e.gexec) { HIP_CHECK(hipGraphExecDestroy(e.gexec)); e.gexec = nullptr; }
  if (e.graph) { HIP_CHECK(hipGraphDestroy(e.graph)); e.graph = nullptr; }
  e.gkey[0] = e.gkey[1] = e.gkey[2] = -9;
}

// Autotuner of the batched sweep: colour schedule x workgroup shape.
//
// Colour schedules are first-fit colourings with a cap on the sites per
// colour: the library default, no cap, and caps of k x (64, 256, 512, 1024)
// sites; a pass over a colour class of n sites takes ceil(n / threads)
// rounds separated by barriers, so the best cap depends on the workgroup
// width and on how the graph packs.  Workgroup shapes are 256, 512 and 1024
// threads (the plain form has one shape).  Every distinct (colouring, shape)
// pair is timed on the device over n_sweeps sweeps of R replicas of random
// spins after two warm-up sweeps; the fastest wins.  Results are keyed by
// the lattice signature, replica count, model, precision, pass composition
// and device, and kept in the tuning cache when one is configured.
void mc_autotune(const Lattice &L0, int R, const double *temps, const hsmd_mc_spec &sp, int n_sweeps,
                 hsmd_mc_tuning *out) {
  if (n_sweeps < 1) throw Error(HSMD_ERR_ARG, "n_sweeps must be >= 1");
  char key[256];
  snprintf(key, sizeof key, "mc_%016llx_R%d_m%d_p%d_%d.%d.%d.%d_%s",
           (unsigned long long)lattice_signature(L0), R, sp.model, sp.precision, sp.n_or, sp.n_hb,
           sp.n_or_boost, sp.n_hb_boost, device_tag().c_str());
  std::vector<double> cv;
  if (tune_cache_get(key, cv) && cv.size() >= 4) {
    out->colour_cap = (int32_t)cv[0];
    out->threads = (int32_t)cv[1];
    out->ms_per_sweep = cv[2];
    out->default_ms = cv[3];
    out->n_candidates = 0;
    out->cached = 1;
    return;
  }
  const int N = L0.N;
  std::vector<int> caps = {0, N};
  for (int w : {64, 256, 512, 1024})
    for (int k = 1; k <= 16; k++)
      if (k * w < N && k * w * kCMAX >= N) caps.push_back(k * w);
  std::vector<int> shapes = {1024};
  if (sp.precision != HSMD_PLAIN_F64) shapes = {1024, 512, 256};
  // random unit (or +-1) spins
  std::vector<double> spins((size_t)R * N * 3, 0.0);
  {
    uint64_t st = 0x9e3779b97f4a7c15ull ^ (uint64_t)N;
    auto u01 = [&]() {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    for (size_t i = 0; i < (size_t)R * N; i++) {
      if (sp.model == HSMD_ISING) {
        spins[3 * i] = u01() < 0.5 ? -1.0 : 1.0;
      } else {
        const double ph = 2.0 * M_PI * u01(), ct = 1.0 - 2.0 * u01(), stt = sqrt(1.0 - ct * ct);
        spins[3 * i] = stt * cos(ph);
        spins[3 * i + 1] = stt * sin(ph);
        spins[3 * i + 2] = ct;
      }
    }
  }
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  std::vector<std::vector<int32_t>> seen;
  double best = 1e300, def_ms = -1.0;
  int best_cap = 0, best_thr = kThreads, ncand = 0;
  try {
    for (int cap : caps) {
      Lattice Lc = L0;
      Lc.col_cap = cap;
      try {
        Lc.rebuild();
      } catch (const Error &) {
        continue;   // the cap needs more colours than a sweep supports
      }
      if (Lc.n_colours > kCMAX) continue;
      bool dup = false;
      for (auto &c : seen) dup = dup || c == Lc.col_off;
      if (dup && cap != 0) continue;
      seen.push_back(Lc.col_off);
      for (int thr : shapes) {
        std::unique_ptr<MCEngine> e(mc_create(Lc, R, temps, sp));
        if (sp.precision != HSMD_PLAIN_F64) mc_set_threads(*e, thr);
        mc_set_spins(*e, spins.data());
        mc_launch_sweeps(*e, 2, 0);
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipEventRecord(ev0, 0));
        mc_launch_sweeps(*e, n_sweeps, 0);
        HIP_CHECK(hipEventRecord(ev1, 0));
        HIP_CHECK(hipEventSynchronize(ev1));
        HIP_CHECK(hipGetLastError());
        float ms = 0.f;
        HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
        const double per = (double)ms / n_sweeps;
        if (getenv("HSMD_TUNE_VERBOSE"))
          fprintf(stderr, "[hsmd-tune] mc cap %d (%d colours) x %d threads: %.4f ms/sweep\n", cap,
                  Lc.n_colours, thr, per);
        ncand++;
        if (cap == 0 && thr == kThreads) def_ms = per;
        if (per < best) { best = per; best_cap = cap; best_thr = thr; }
      }
    }
  } catch (...) {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    throw;
  }
  HIP_CHECK(hipEventDestroy(ev0));
  HIP_CHECK(hipEventDestroy(ev1));
  out->colour_cap = best_cap;
  out->threads = best_thr;
  out->ms_per_sweep = best;
  out->default_ms = def_ms;
  out->n_candidates = ncand;
  out->cached = 0;
  tune_cache_put(key, {(double)best_cap, (double)best_thr, best, def_ms});
}

// ------------------------------------------------- fields / energies (ABI)

template <typename T, int ZC>
__global__ void fields_kernel(const Coup cp, const int32_t *nb, const T *Jt, const T *spins,
                              T *out) {
  const int c = blockIdx.y;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= cp.N) return;
  const T *s = spins + (size_t)c * 3 * cp.Npad;
  T hx, hy, hz;
  local_field<T, ZC>(cp, nb, Jt, p, s, s + cp.Npad, s + 2 * cp.Npad, hx, hy, hz);
  T *o = out + (size_t)c * 3 * cp.N;
  o[p] = hx; o[cp.N + p] = hy; o[2 * cp.N + p] = hz;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void energy_kernel(const Coup cp, const Bonds bd,
                                                          const int32_t *bi, const int32_t *bj,
                                                          const uint8_t *bc, const T *spins,
                                                          double *E) {
  __shared__ double scratch[16];
  const int c = blockIdx.x;
  const T *s = spins + (size_t)c * 3 * cp.Npad;
  const T *sx = s, *sy = s + cp.Npad, *sz = s + 2 * cp.Npad;
  double e = 0.0, mz = 0.0;
  for (int b = threadIdx.x; b < bd.Nb; b += blockDim.x) {
    const int i = bi[b], j = bj[b], k = bc[b];
    e += bd.jxy[k] * ((double)sx[i] * (double)sx[j] + (double)sy[i] * (double)sy[j]) +
         bd.jz[k] * ((double)sz[i] * (double)sz[j]);
  }
  for (int p = threadIdx.x; p < cp.N; p += blockDim.x) mz += (double)sz[p];
  double v[1] = {e - (cp.h_on ? cp.h * mz : 0.0)};
  block_sum<1>(v, scratch);
  if (threadIdx.x == 0) E[c] = v[0];
}

template <int ZC, typename T>
static void launch_fields(dim3 grid, const Coup &cp, const int32_t *nb, const T *Jt, const T *s,
                          T *out) {
  hipLaunchKernelGGL((fields_kernel<T, ZC>), grid, dim3(256), 0, 0, cp, nb, Jt, s, out);
}

template <typename T>
static void fields_energies_impl(const Lattice &L, int R, const double *spins, double *fields,
                                 double *energies) {
  const int N = L.N;
  const size_t Npad = (size_t)N + 1;
  DevBuf<int32_t> nb, perm, bi, bj;
  DevBuf<uint8_t> bc;
  DevBuf<float> Jf;
  DevBuf<double> Jd;
  const bool f32 = sizeof(T) == 4;
  upload_lattice_common(L, nb, perm, bi, bj, bc, Jf, Jd, f32);
  Coup cp;
  fill_coup(L, cp, false);
  Bonds bd;
  fill_bonds(L, bd);
  std::vector<T> h(3 * Npad * R, T(0));
  for (int r = 0; r < R; r++)
    for (int i = 0; i < N; i++)
      for (int d = 0; d < 3; d++) h[(r * 3 + d) * Npad + L.inv[i]] = (T)spins[((size_t)r * N + i) * 3 + d];
  DevBuf<T> ds;
  ds.upload(h.data(), h.size());
  const T *Jt = f32 ? (const T *)Jf.p : (const T *)Jd.p;
  if (fields) {
    DevBuf<T> out;
    out.alloc((size_t)3 * N * R);
    HSMD_ZC_DISPATCH(L.Z, launch_fields, dim3((N + 255) / 256, R), cp, nb.p, Jt, ds.p, out.p);
    HIP_CHECK(hipGetLastError());
    std::vector<T> o((size_t)3 * N * R);
    HIP_CHECK(hipMemcpy(o.data(), out.p, o.size() * sizeof(T), hipMemcpyDeviceToHost));
    for (int r = 0; r < R; r++)
      for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++)
          fields[((size_t)r * N + i) * 3 + d] = (double)o[((size_t)r * 3 + d) * N + L.inv[i]];
  }
  if (energies) {
    DevBuf<double> E;
    E.alloc(R);
    hipLaunchKernelGGL(energy_kernel<T>, dim3(R), dim3(kThreads), 0, 0, cp, bd, bi.p, bj.p, bc.p,
                       ds.p, E.p);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(energies, E.p, R * sizeof(double), hipMemcpyDeviceToHost));
  }
  HIP_CHECK(hipDeviceSynchronize());
}

void fields_energies(const Lattice &L, int precision, int R, const double *spins, double *fields,
                     double *energies) {
  if (R < 1) throw Error(HSMD_ERR_ARG, "R must be >= 1");
  if (precision == HSMD_BATCHED_F32) fields_energies_impl<float>(L, R, spins, fields, energies);
  else fields_energies_impl<double>(L, R, spins, fields, energies);
}

// ---------------------------------------------------------- self checks

__global__ void philox_kernel(uint64_t seed, uint32_t step, uint32_t inst, uint32_t stream, int n,
                              uint32_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const u32x4 r = draw4(seed, (uint32_t)i, step, inst, stream);
  out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
}

__global__ void exp_kernel(int n, const double *x, double *y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = det_exp_neg(x[i]);
}

void debug_philox(uint64_t seed, uint32_t step, uint32_t inst, uint32_t stream, int n, uint32_t *out) {
  if (n <= 0) return;
  DevBuf<uint32_t> d;
  d.alloc((size_t)4 * n);
  hipLaunchKernelGGL(philox_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, seed, step, inst, stream, n, d.p);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(out, d.p, (size_t)4 * n * sizeof(uint32_t), hipMemcpyDeviceToHost));
}

void debug_exp(int n, const double *x, double *y) {
  if (n <= 0) return;
  DevBuf<double> dx, dy;
  dx.upload(x, n);
  dy.alloc(n);
  hipLaunchKernelGGL(exp_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, n, dx.p, dy.p);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpy(y, dy.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace hsmd
