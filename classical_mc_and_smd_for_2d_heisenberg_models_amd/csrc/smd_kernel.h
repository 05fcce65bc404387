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
  const uint32_t *nb16;   // packed 16-bit neighbour pairs, kZMAX/2 x N (sentinel-padded)
  const void *Jt;
  const int32_t *bi, *bj;
  const uint8_t *bc;
  int field_energy;       // 1: bond energy == 1/2 sum S.h - 1/2 h sum Sz (symmetric tables)
  // compact labels (sparse tables, read by the global-memory form): each
  // site's coupled neighbours first, in slot order, as 14-bit labels with the
  // 2-bit shell of their slot, kZMAX/2 x N; wcnt[p / 64] = label words any
  // site of 64-site chunk p / 64 uses
  int compact;
  const uint32_t *nb16c;
  const uint8_t *wcnt;
  double shj[3], shjz[3];  // couplings per shell (J, J lambda)
  // LDS-resident form with field energies: the RK stages ordered per wave
  // (wsync != 0) instead of by workgroup barriers; wdep[w] = bit mask of the
  // waves owning a neighbour of a site wave w integrates (site p belongs to
  // thread p mod 1024 at any sites-per-thread count)
  int wsync;
  uint32_t wdep[16];
};

// Load through a global-address-space pointer (a plain global load rather
// than a flat one that must also consider LDS).
template <typename T>
__device__ __forceinline__ T gload(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}

struct SampleDev {
  int li;
  int N;
  long long spin_off;   // into spins / S0 (3 * Npad per sample)
  long long traj_off;   // into trajectory chunk (Tc * ldm per sample)
  long long y_off;      // into Y accumulators (3 * N * 2 * n_omega per sample)
  int ldm;              // trajectory row stride: 3 N rounded up to a multiple of 4
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

// Stores the RK4 combination v / |v|, normalised in FP64 whatever T is.
// Normalised in FP32 (squares, square root, division and products each
// rounded), unit vectors come out about 0.3 ulp too long on average
// (<|S| - 1> = 1.7e-8); the energy correction, evaluated on those spins,
// restores the energy of spins that are too long, and the correction's own
// FP64 normalisation then shortens them, so E rose by 2 <|S| - 1> |E| every
// step (1.7e-4 per step on 4096 sites).  Rounding an FP64 unit vector to
// FP32 is unbiased.
// 1 / sqrt(n2) in FP64.  FP32 spins: from the hardware's FP64 reciprocal
// square-root estimate and one Newton step (relative error ~1e-13, far below
// the FP32 rounding of the stored spin, so the rounding stays unbiased)
// instead of a correctly rounded FP64 square root and division; FP64 spins
// keep the correctly rounded form.
template <typename T>
__device__ __forceinline__ double inv_norm(double n2) {
  if constexpr (sizeof(T) == 4) {
    const double r = __builtin_amdgcn_rsq(n2);
    return r * (1.5 - 0.5 * n2 * r * r);
  } else {
    return 1.0 / sqrt(n2);
  }
}

template <typename T>
__device__ __forceinline__ void unit_store(const T (&v)[3], T &x, T &y, T &z) {
  const double a = v[0], b = v[1], c = v[2];
  const double inr = inv_norm<T>(a * a + b * b + c * c);
  x = (T)(a * inr);
  y = (T)(b * inr);
  z = (T)(c * inr);
}

struct StepArgs {
  const LatDev *lats;
  const SampleDev *samples;
  void *spins;          // per sample [3][Npad]
  const void *s0;       // per sample [3][Npad], t = 0 spins
  void *traj;           // per sample [Tc][3][N]
  double *qss;          // per sample [T]
  double *eref;         // per sample, energy at the current time
  void *stage;          // per sample [3][Npad] stage buffers B1, B2 (global-memory form)
  long long stage_stride;   // elements from B1 to B2
  void *ks;             // wide form: per sample [6][Npad] K1 and K2+K3
  double *dd;           // wide form: per sample [3][Npad] correction directions
  double dt;
  int t0;               // global time index of the first point of this launch
  int n_points;         // time points produced by this launch
  int chunk_base;       // time index stored at trajectory row 0
  int T;                // total time points
  int lds_npad;         // LDS component stride, SPT x 1024 + 1 (LDS form)
  int sample_base;      // first sample of this launch
  uint32_t *sync_err;   // set when a wave-ordering wait ran out
};

// Per-lattice constants the time loop reads.  The kernel copies them out of
// the lattice table into registers before it stores anything, which lets the
// compiler keep them in scalar registers; read through the table after a
// store, every coupling would be a vector load waited on in the inner loop.
template <typename T, int ZC>
struct FieldCtx {
  int Z, N;
  bool general;
  T hz0;           // -h (field along z) or 0
  const T *Jt;     // per-slot J table (general lattices)
  T jxy[ZC];       // uniform couplings per slot
  T jz[ZC];        // J * lambda per slot, or lambda alone for general lattices
  bool compact;    // compact labels (LatDev::compact)
  const uint8_t *wcnt;
  // couplings per shell as separate scalars: an array indexed through a
  // select would be placed in scratch memory
  T sj0, sj1, sj2, sz0, sz1, sz2;
};

template <typename T, int ZC>
__device__ __forceinline__ FieldCtx<T, ZC> field_ctx(const Coup &cp, const T *Jt) {
  FieldCtx<T, ZC> c;
  c.Z = cp.Z;
  c.N = cp.N;
  c.general = cp.general != 0;
  c.hz0 = cp.h_on ? T(-cp.h) : T(0);
  c.Jt = Jt;
#pragma unroll
  for (int s = 0; s < ZC; s++) {
    c.jxy[s] = cj_xy<T>(cp, s);
    c.jz[s] = c.general ? cj_lam<T>(cp, s) : cj_z<T>(cp, s);
  }
  c.compact = false;
  c.wcnt = nullptr;
  return c;
}

template <typename T, int ZC>
__device__ __forceinline__ FieldCtx<T, ZC> field_ctx(const LatDev &L) {
  FieldCtx<T, ZC> c = field_ctx<T, ZC>(L.cp, (const T *)L.Jt);
  c.compact = L.compact != 0;
  c.wcnt = L.wcnt;
  c.sj0 = (T)L.shj[0]; c.sj1 = (T)L.shj[1]; c.sj2 = (T)L.shj[2];
  c.sz0 = (T)L.shjz[0]; c.sz1 = (T)L.shjz[1]; c.sz2 = (T)L.shjz[2];
  return c;
}

// Local field of site p from its packed 16-bit neighbour labels.  Each label
// word passes through an empty asm statement so that the compiler unpacks it
// here instead of hoisting Z x 6 loop-invariant LDS addresses out of the time
// loop (which spills them to scratch).  Without per-slot J tables (GEN false)
// all ZC slots are summed with no branch: slots beyond the lattice's Z hold
// the zero sentinel site and zero couplings, and the LDS reads of every slot
// issue back to back.  GEN lattices (disorder, Lieb J2 asymmetry) read J per
// site and slot and test the slot count at run time.
#ifndef SMD_GN
#define SMD_GN 12   // neighbours whose reads are grouped in field16
#endif
template <typename T, int ZC, int ZH, bool GEN, bool CP = false>
__device__ __forceinline__ void field16(const FieldCtx<T, ZC> &c, const uint32_t (&q)[ZH], int p,
                                        const T *sx, const T *sy, const T *sz, T &hx, T &hy,
                                        T &hz) {
  hx = T(0);
  hy = T(0);
  hz = c.hz0;
  if constexpr (CP && !GEN) {
    // compact labels: the words this wave's 64-site chunk needs, couplings
    // selected by the shell bits; the same sums in the same slot order as
    // the padded form (which adds zero for every sentinel slot)
    const int nw = __builtin_amdgcn_readfirstlane((int)c.wcnt[p >> 6]);
    // the shell couplings as values (selects between fields of c would be
    // turned into a load from a selected address, forcing c into scratch)
    T a0 = c.sj0, a1 = c.sj1, a2 = c.sj2, b0 = c.sz0, b1 = c.sz1, b2 = c.sz2;
    asm volatile("" : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(b0), "+v"(b1), "+v"(b2));
#pragma unroll
    for (int k = 0; k < ZH; k++) {
      if (k < nw) {
        uint32_t w = q[k];
        asm volatile("" : "+v"(w));
        const int i0 = (int)(w & 0x3fffu), i1 = (int)((w >> 16) & 0x3fffu);
        const int s0 = (int)((w >> 14) & 3u), s1 = (int)(w >> 30);
        const T j0 = s0 == 0 ? a0 : (s0 == 1 ? a1 : a2);
        const T z0 = s0 == 0 ? b0 : (s0 == 1 ? b1 : b2);
        hx += j0 * sx[i0]; hy += j0 * sy[i0]; hz += z0 * sz[i0];
        if (2 * k + 1 < ZC) {
          const T j1 = s1 == 0 ? a0 : (s1 == 1 ? a1 : a2);
          const T z1 = s1 == 0 ? b0 : (s1 == 1 ? b1 : b2);
          hx += j1 * sx[i1]; hy += j1 * sy[i1]; hz += z1 * sz[i1];
        }
      }
    }
  } else if constexpr (!GEN) {
    // the reads of SMD_GN neighbours are issued before their products (the
    // scheduling barriers keep each group together, so its reads are in
    // flight at once); the sums run in slot order
    constexpr int GN = SMD_GN;
#pragma unroll
    for (int s0 = 0; s0 < ZC; s0 += GN) {
      T vx[GN], vy[GN], vz[GN];
#pragma unroll
      for (int u = 0; u < GN; u++) {
        const int sl = s0 + u;
        if (sl < ZC) {
          uint32_t w = q[sl / 2];
          asm volatile("" : "+v"(w));
          const int i = (sl & 1) ? (int)(w >> 16) : (int)(w & 0xffffu);
          vx[u] = sx[i]; vy[u] = sy[i]; vz[u] = sz[i];
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < GN; u++) {
        const int sl = s0 + u;
        if (sl < ZC) {
          hx += c.jxy[sl] * vx[u]; hy += c.jxy[sl] * vy[u]; hz += c.jz[sl] * vz[u];
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    const T *jr = c.Jt;
    if (c.general) {
      int pp = p;
      asm volatile("" : "+v"(pp));
      jr = c.Jt + pp;
    }
#pragma unroll
    for (int k = 0; k < ZH; k++) {
      uint32_t w = q[k];
      asm volatile("" : "+v"(w));
      const int i0 = (int)(w & 0xffffu), i1 = (int)(w >> 16);
      if (2 * k < c.Z) {
        const T J = c.general ? jr[(size_t)(2 * k) * c.N] : c.jxy[2 * k];
        const T Jz = c.general ? J * c.jz[2 * k] : c.jz[2 * k];
        hx += J * sx[i0]; hy += J * sy[i0]; hz += Jz * sz[i0];
      }
      if (2 * k + 1 < ZC && 2 * k + 1 < c.Z) {
        const T J = c.general ? jr[(size_t)(2 * k + 1) * c.N] : c.jxy[2 * k + 1];
        const T Jz = c.general ? J * c.jz[2 * k + 1] : c.jz[2 * k + 1];
        hx += J * sx[i1]; hy += J * sy[i1]; hz += Jz * sz[i1];
      }
    }
  }
}

// One launch produces n_points time points t0 .. t0+n_points-1 of every
// sample; t = 0 is the input configuration, every later point is one RK4
// step plus the energy correction.  Stage configurations alternate between
// two buffers (B1, B2), so each stage writes its successor's input without a
// second barrier, and stage 4 writes S in place (it reads only B1 and the
// site's own S).  Neighbour labels stay in registers when SPT x ZH <= 16
// words and the t = 0 spins (for qSS) when SPT <= 2; otherwise they are
// re-read where used.
// When the tables are symmetric (field_energy), E_ref and E(S) come from the
// fields the step computes anyway, E = 1/2 sum_i S_i.h_i - 1/2 h sum_i S_z,i;
// otherwise from the bond list.
// Wave ordering (LDS form with field energies, lattice flag wsync): each
// of the five neighbour-reading phases of a step (stages 1-4 and the
// correction's fields) and the next step's stage 1 start once the waves that
// own the phase's neighbours have published the phase before (an LDS count
// per wave, released after the wave's stores).  A wave's dependencies also
// cover every reader of the sites it writes (the neighbour relation is
// symmetric), so no stage buffer is overwritten while a slower wave still
// reads it; the per-step FP64 sums keep one workgroup barrier, with the
// reduction scratch alternating between two halves.  Each site reads the
// same values as under a barrier per phase.
template <typename T, int SPT, bool LDS, int ZC, bool GEN>
__device__ __forceinline__ void smd_body(const StepArgs &a, unsigned char *smem, double *scratch,
                                         int *wprog) {
  constexpr int ZH = (ZC + 1) / 2;
  constexpr bool QREG = SPT * ZH <= 16;   // labels in registers
  constexpr bool GREG = SPT <= 2;         // t = 0 spins in registers
  const int sidx = blockIdx.x + a.sample_base;
  const SampleDev sd = a.samples[sidx];
  const LatDev &L = a.lats[sd.li];
  const int N = sd.N, Npad = N + 1;
  const FieldCtx<T, ZC> fc = field_ctx<T, ZC>(L);
  const uint32_t *nb16 = L.nb16;
  const bool fe = L.field_energy != 0;
  const double hext = L.cp.h_on ? L.cp.h : 0.0;
  const bool ws = LDS && fe && L.wsync != 0;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const uint32_t dm = ws ? L.wdep[wv] : 0u;
  T *gS = (T *)a.spins + sd.spin_off;
  const T *g0 = (const T *)a.s0 + sd.spin_off;
  T *S, *B1, *B2;
  int stride;
  if (LDS) {
    // a compile-time component stride: the y and z reads of a neighbour then
    // take the x address plus an immediate LDS offset
    stride = SPT * kSmdThreads + 64;
    S = (T *)smem;
    B1 = S + 3 * stride;
    B2 = B1 + 3 * stride;
    for (int i = threadIdx.x; i < 3 * Npad; i += blockDim.x) {
      const int d = i / Npad, p = i - d * Npad;
      S[d * stride + p] = gS[i];
      B1[d * stride + p] = T(0);
      B2[d * stride + p] = T(0);
    }
  } else {
    // global-memory form: spins in place, two zeroed stage buffers beside them
    stride = Npad;
    S = gS;
    B1 = (T *)a.stage + sd.spin_off;
    B2 = (T *)a.stage + a.stage_stride + sd.spin_off;
  }
  const int Z = fc.Z;
  const uint32_t pad = (uint32_t)N | ((uint32_t)N << 16);
  uint32_t q[QREG ? SPT : 1][ZH];
  T G0[GREG ? SPT : 1][3];
#pragma unroll
  for (int u = 0; u < SPT; u++) {
    const int p = threadIdx.x + u * kSmdThreads;
    if constexpr (QREG) {
#pragma unroll
      for (int k = 0; k < ZH; k++) q[u][k] = p < N ? gload(nb16 + (size_t)k * N + p) : pad;
    }
    if constexpr (GREG) {
#pragma unroll
      for (int d = 0; d < 3; d++) G0[u][d] = p < N ? g0[d * Npad + p] : T(0);
    }
  }
  if (threadIdx.x < 16) wprog[threadIdx.x] = 0;
  __syncthreads();
  int wcount = 0;   // phases this wave has published
  auto sync = [&]() __attribute__((always_inline)) {
    if (ws) {
      wcount++;
      if (lane == 0) __hip_atomic_store(&wprog[wv], wcount, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (dm) {
        // bounded: a wait that runs out flags sync_err and goes on
        uint32_t n = 0;
        while (true) {
          const int v = __hip_atomic_load(&wprog[lane & 15], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          const bool behind = ((dm >> (lane & 15)) & 1u) && v < wcount;
          if (__builtin_amdgcn_ballot_w64(behind) == 0) break;
          if (++n > (1u << 20)) {
            if (lane == 0) atomicOr(a.sync_err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
      __syncthreads();
    }
  };
  int half = 0;   // reduction scratch half of the wave-ordered form
  T *Sx = S, *Sy = S + stride, *Sz = S + 2 * stride;
  T *Px = B1, *Py = B1 + stride, *Pz = B1 + 2 * stride;
  T *Qx = B2, *Qy = B2 + stride, *Qz = B2 + 2 * stride;
  const T dt = T(a.dt);
  double Eref = a.eref[sidx];
  T *traj = (T *)a.traj + sd.traj_off;
  // field of site p (slot u) from configuration (fx, fy, fz)
  auto field = [&](int u, int p, const T *fx, const T *fy, const T *fz, T &hx, T &hy, T &hz) {
    uint32_t w[ZH];
    if constexpr (QREG) {
#pragma unroll
      for (int k = 0; k < ZH; k++) w[k] = q[u][k];
    } else {
      int pp = p;
      asm volatile("" : "+v"(pp));
#pragma unroll
      for (int k = 0; k < ZH; k++) w[k] = gload(nb16 + (size_t)k * N + pp);
    }
    field16<T, ZC, ZH, GEN>(fc, w, p, fx, fy, fz, hx, hy, hz);
  };
  auto spin0 = [&](int u, int p, int d) -> double {
    if constexpr (GREG) return (double)G0[u][d];
    else return (double)g0[d * Npad + p];
  };
  // own-site index, opaque so that per-site LDS addresses are formed where
  // they are used rather than held across the time loop
  auto site = [&](int u) {
    int p = threadIdx.x + u * kSmdThreads;
    asm volatile("" : "+v"(p));
    return p;
  };
  double qss_prev = 0.0;   // qSS partial of the previous step, reduced with this step's sums
  int t_prev = -1;
  for (int it = 0; it < a.n_points; it++) {
    const int t = a.t0 + it;
    if (t > 0) {
      T K1[SPT][3], A2[SPT][3];
      double eref_part = 0.0;
      // stage 1: K1 from S, B1 = S + K1/2 (and the energy of S from its fields)
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = site(u);
        if (p < N) {
          T hx, hy, hz;
          field(u, p, Sx, Sy, Sz, hx, hy, hz);
          const T x = Sx[p], y = Sy[p], z = Sz[p];
          if (fe) eref_part += 0.5 * ((double)x * hx + (double)y * hy + (double)z * hz) - 0.5 * hext * z;
          T px, py, pz;
          cross3(hx, hy, hz, x, y, z, px, py, pz);
          K1[u][0] = dt * px; K1[u][1] = dt * py; K1[u][2] = dt * pz;
          Px[p] = x + T(0.5) * K1[u][0];
          Py[p] = y + T(0.5) * K1[u][1];
          Pz[p] = z + T(0.5) * K1[u][2];
        }
      }
      sync();
      // stage 2: K2 from B1, B2 = S + K2/2
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = site(u);
        if (p < N) {
          T hx, hy, hz;
          field(u, p, Px, Py, Pz, hx, hy, hz);
          T px, py, pz;
          cross3(hx, hy, hz, Px[p], Py[p], Pz[p], px, py, pz);
          A2[u][0] = dt * px; A2[u][1] = dt * py; A2[u][2] = dt * pz;
          Qx[p] = Sx[p] + T(0.5) * A2[u][0];
          Qy[p] = Sy[p] + T(0.5) * A2[u][1];
          Qz[p] = Sz[p] + T(0.5) * A2[u][2];
        }
      }
      sync();
      // stage 3: K3 from B2, B1 = S + K3, A2 = K2 + K3
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = site(u);
        if (p < N) {
          T hx, hy, hz;
          field(u, p, Qx, Qy, Qz, hx, hy, hz);
          T px, py, pz;
          cross3(hx, hy, hz, Qx[p], Qy[p], Qz[p], px, py, pz);
          const T k3[3] = {dt * px, dt * py, dt * pz};
          Px[p] = Sx[p] + k3[0];
          Py[p] = Sy[p] + k3[1];
          Pz[p] = Sz[p] + k3[2];
#pragma unroll
          for (int d = 0; d < 3; d++) A2[u][d] += k3[d];
        }
      }
      sync();
      // stage 4 and combine, S written in place
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = site(u);
        if (p < N) {
          T hx, hy, hz;
          field(u, p, Px, Py, Pz, hx, hy, hz);
          T px, py, pz;
          cross3(hx, hy, hz, Px[p], Py[p], Pz[p], px, py, pz);
          const T f1 = T(1.0 / 6.0), f2 = T(2.0 / 6.0);
          T v[3];
          const T k4[3] = {dt * px, dt * py, dt * pz};
          const T s[3] = {Sx[p], Sy[p], Sz[p]};
#pragma unroll
          for (int d = 0; d < 3; d++) v[d] = s[d] + (f1 * (K1[u][d] + k4[d]) + f2 * A2[u][d]);
          unit_store<T>(v, Sx[p], Sy[p], Sz[p]);
        }
      }
      sync();
      // energy correction: sum |h x S|^2 and E(S) in FP64
      double D[SPT][3];
      double v4[4] = {0.0, 0.0, eref_part, qss_prev};
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = site(u);
        if (p < N) {
          T hx, hy, hz;
          field(u, p, Sx, Sy, Sz, hx, hy, hz);
          const double x = Sx[p], y = Sy[p], z = Sz[p];
          if (fe) v4[1] += 0.5 * (x * hx + y * hy + z * hz) - 0.5 * hext * z;
          double px, py, pz;
          cross3<double>(hx, hy, hz, x, y, z, px, py, pz);
          v4[0] += px * px + py * py + pz * pz;
          cross3<double>(x, y, z, px, py, pz, D[u][0], D[u][1], D[u][2]);
        }
      }
      if (!fe) v4[1] = bond_energy<T>(L, Sx, Sy, Sz) + field_energy_term<T>(L, Sz);
      if (ws) {
        block_sum_once<4>(v4, scratch + half * 4 * 16);   // its barrier: S may be overwritten below
        half ^= 1;
      } else {
        block_sum<4>(v4, scratch);   // ends with a barrier: S may be overwritten below
      }
      if (fe) Eref = v4[2];
      if (t_prev >= 0 && threadIdx.x == 0) a.qss[(size_t)sidx * a.T + t_prev] = v4[3] / (double)N;
      const double cfac = (Eref - v4[1]) / v4[0];
      T *row = a.traj ? traj + (size_t)(t - a.chunk_base) * sd.ldm : nullptr;
      double qs = 0.0;
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = site(u);
        if (p < N) {
          double x = (double)Sx[p] + cfac * D[u][0];
          double y = (double)Sy[p] + cfac * D[u][1];
          double z = (double)Sz[p] + cfac * D[u][2];
          const double inr = inv_norm<T>(x * x + y * y + z * z);
          x *= inr; y *= inr; z *= inr;
          qs += spin0(u, p, 0) * x + spin0(u, p, 1) * y + spin0(u, p, 2) * z;
          const T tx = (T)x, ty = (T)y, tz = (T)z;
          Sx[p] = tx; Sy[p] = ty; Sz[p] = tz;
          if (row) { row[p] = tx; row[N + p] = ty; row[2 * N + p] = tz; }
        }
      }
      sync();
      qss_prev = qs;
      t_prev = t;
      if (!fe) {
        double v1[1] = {bond_energy<T>(L, Sx, Sy, Sz) + field_energy_term<T>(L, Sz)};
        block_sum<1>(v1, scratch);
        Eref = v1[0];
      }
    } else {
      if (threadIdx.x == 0) a.qss[(size_t)sidx * a.T + 0] = 1.0;
      if (a.traj) {
        T *row = traj + (size_t)(t - a.chunk_base) * sd.ldm;
        for (int p = threadIdx.x; p < N; p += blockDim.x) {
          row[p] = Sx[p]; row[N + p] = Sy[p]; row[2 * N + p] = Sz[p];
        }
      }
    }
  }
  // flush the last step's qSS and, with field energies, the energy of the
  // final configuration (the next launch's E_ref)
  {
    double e_part = 0.0;
    if (fe) {
#pragma unroll
      for (int u = 0; u < SPT; u++) {
        const int p = site(u);
        if (p < N) {
          T hx, hy, hz;
          field(u, p, Sx, Sy, Sz, hx, hy, hz);
          const double x = Sx[p], y = Sy[p], z = Sz[p];
          e_part += 0.5 * (x * hx + y * hy + z * hz) - 0.5 * hext * z;
        }
      }
    }
    double v2[2] = {e_part, qss_prev};
    block_sum<2>(v2, scratch);
    if (fe) Eref = v2[0];
    if (t_prev >= 0 && threadIdx.x == 0) a.qss[(size_t)sidx * a.T + t_prev] = v2[1] / (double)N;
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

template <typename T, int SPT, bool LDS, int ZC, bool GEN>
__global__ __launch_bounds__(kSmdThreads) void smd_kernel(const StepArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double scratch[2 * 4 * 16];
  __shared__ int wprog[16];
  smd_body<T, SPT, LDS, ZC, GEN>(a, smem, scratch, wprog);
}

// Mixed batches whose lattices fall in two slot classes: each workgroup runs
// the body of its own lattice's class, so a lattice with 8 coupled slots is
// not summed over the 12 slots another lattice of the batch needs.
template <typename T, int SPT, int ZLO, int ZHI>
__global__ __launch_bounds__(kSmdThreads) void smd_kernel_two(const StepArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double scratch[2 * 4 * 16];
  __shared__ int wprog[16];
  const SampleDev sd = a.samples[blockIdx.x + a.sample_base];
  if (a.lats[sd.li].cp.Z <= ZLO) smd_body<T, SPT, true, ZLO, false>(a, smem, scratch, wprog);
  else smd_body<T, SPT, true, ZHI, false>(a, smem, scratch, wprog);
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

// Wide form for lattices beyond the LDS: one workgroup per sample, every
// array (S, B1, B2, the RK increments K1 and K2 + K3, the correction
// directions) in global memory, sites visited with a workgroup-stride loop;
// the same arithmetic, in the same order, as smd_kernel.  All samples run in
// one launch; the working set (tens of MB at N ~ 10^4) stays in the L2 /
// Infinity Cache across the time loop.
template <typename T, int ZC, bool GEN>
__global__ __launch_bounds__(kSmdThreads) void smd_kernel_wide(const StepArgs a) {
  __shared__ double scratch[4 * 16];
  constexpr int ZH = (ZC + 1) / 2;
  const int sidx = blockIdx.x + a.sample_base;
  const SampleDev sd = a.samples[sidx];
  const LatDev &L = a.lats[sd.li];
  const int N = sd.N, Npad = N + 1;
  const FieldCtx<T, ZC> fc = field_ctx<T, ZC>(L);
  const uint32_t *nb16 = L.nb16;
  const bool fe = L.field_energy != 0;
  const double hext = L.cp.h_on ? L.cp.h : 0.0;
  const int Z = fc.Z;
  const uint32_t pad = (uint32_t)N | ((uint32_t)N << 16);
  T *S = (T *)a.spins + sd.spin_off;
  const T *g0 = (const T *)a.s0 + sd.spin_off;
  T *B1 = (T *)a.stage + sd.spin_off;
  T *B2 = (T *)a.stage + a.stage_stride + sd.spin_off;
  T *K = (T *)a.ks + 2 * sd.spin_off;
  double *D = a.dd + sd.spin_off;
  T *Sx = S, *Sy = S + Npad, *Sz = S + 2 * Npad;
  T *Px = B1, *Py = B1 + Npad, *Pz = B1 + 2 * Npad;
  T *Qx = B2, *Qy = B2 + Npad, *Qz = B2 + 2 * Npad;
  const T dt = T(a.dt);
  double Eref = a.eref[sidx];
  T *traj = (T *)a.traj + sd.traj_off;
  const bool cpl = !GEN && fc.compact;
  const uint32_t *nbw = cpl ? L.nb16c : nb16;
  auto field = [&](int p, const T *fx, const T *fy, const T *fz, T &hx, T &hy, T &hz) {
    uint32_t w[ZH];
    int pp = p;
    asm volatile("" : "+v"(pp));
#pragma unroll
    for (int k = 0; k < ZH; k++) w[k] = gload(nbw + (size_t)k * N + pp);
    if (cpl) field16<T, ZC, ZH, GEN, true>(fc, w, p, fx, fy, fz, hx, hy, hz);
    else field16<T, ZC, ZH, GEN>(fc, w, p, fx, fy, fz, hx, hy, hz);
  };
  double qss_prev = 0.0;
  int t_prev = -1;
  for (int it = 0; it < a.n_points; it++) {
    const int t = a.t0 + it;
    if (t > 0) {
      double eref_part = 0.0;
      for (int p = threadIdx.x; p < N; p += kSmdThreads) {
        T hx, hy, hz;
        field(p, Sx, Sy, Sz, hx, hy, hz);
        const T x = Sx[p], y = Sy[p], z = Sz[p];
        if (fe) eref_part += 0.5 * ((double)x * hx + (double)y * hy + (double)z * hz) - 0.5 * hext * z;
        T px, py, pz;
        cross3(hx, hy, hz, x, y, z, px, py, pz);
        const T k1[3] = {dt * px, dt * py, dt * pz};
        K[p] = k1[0]; K[Npad + p] = k1[1]; K[2 * Npad + p] = k1[2];
        Px[p] = x + T(0.5) * k1[0];
        Py[p] = y + T(0.5) * k1[1];
        Pz[p] = z + T(0.5) * k1[2];
      }
      __syncthreads();
      for (int p = threadIdx.x; p < N; p += kSmdThreads) {
        T hx, hy, hz;
        field(p, Px, Py, Pz, hx, hy, hz);
        T px, py, pz;
        cross3(hx, hy, hz, Px[p], Py[p], Pz[p], px, py, pz);
        const T k2[3] = {dt * px, dt * py, dt * pz};
        K[3 * Npad + p] = k2[0]; K[4 * Npad + p] = k2[1]; K[5 * Npad + p] = k2[2];
        Qx[p] = Sx[p] + T(0.5) * k2[0];
        Qy[p] = Sy[p] + T(0.5) * k2[1];
        Qz[p] = Sz[p] + T(0.5) * k2[2];
      }
      __syncthreads();
      for (int p = threadIdx.x; p < N; p += kSmdThreads) {
        T hx, hy, hz;
        field(p, Qx, Qy, Qz, hx, hy, hz);
        T px, py, pz;
        cross3(hx, hy, hz, Qx[p], Qy[p], Qz[p], px, py, pz);
        const T k3[3] = {dt * px, dt * py, dt * pz};
        Px[p] = Sx[p] + k3[0];
        Py[p] = Sy[p] + k3[1];
        Pz[p] = Sz[p] + k3[2];
        K[3 * Npad + p] += k3[0]; K[4 * Npad + p] += k3[1]; K[5 * Npad + p] += k3[2];
      }
      __syncthreads();
      for (int p = threadIdx.x; p < N; p += kSmdThreads) {
        T hx, hy, hz;
        field(p, Px, Py, Pz, hx, hy, hz);
        T px, py, pz;
        cross3(hx, hy, hz, Px[p], Py[p], Pz[p], px, py, pz);
        const T f1 = T(1.0 / 6.0), f2 = T(2.0 / 6.0);
        const T k4[3] = {dt * px, dt * py, dt * pz};
        const T s[3] = {Sx[p], Sy[p], Sz[p]};
        T v[3];
#pragma unroll
        for (int d = 0; d < 3; d++) v[d] = s[d] + (f1 * (K[d * Npad + p] + k4[d]) + f2 * K[(3 + d) * Npad + p]);
        unit_store<T>(v, Sx[p], Sy[p], Sz[p]);
      }
      __syncthreads();
      double v4[4] = {0.0, 0.0, eref_part, qss_prev};
      for (int p = threadIdx.x; p < N; p += kSmdThreads) {
        T hx, hy, hz;
        field(p, Sx, Sy, Sz, hx, hy, hz);
        const double x = Sx[p], y = Sy[p], z = Sz[p];
        if (fe) v4[1] += 0.5 * (x * hx + y * hy + z * hz) - 0.5 * hext * z;
        double px, py, pz, dx, dy, dz;
        cross3<double>(hx, hy, hz, x, y, z, px, py, pz);
        v4[0] += px * px + py * py + pz * pz;
        cross3<double>(x, y, z, px, py, pz, dx, dy, dz);
        D[p] = dx; D[Npad + p] = dy; D[2 * Npad + p] = dz;
      }
      if (!fe) v4[1] = bond_energy<T>(L, Sx, Sy, Sz) + field_energy_term<T>(L, Sz);
      block_sum<4>(v4, scratch);
      if (fe) Eref = v4[2];
      if (t_prev >= 0 && threadIdx.x == 0) a.qss[(size_t)sidx * a.T + t_prev] = v4[3] / (double)N;
      const double cfac = (Eref - v4[1]) / v4[0];
      T *row = a.traj ? traj + (size_t)(t - a.chunk_base) * sd.ldm : nullptr;
      double qs = 0.0;
      for (int p = threadIdx.x; p < N; p += kSmdThreads) {
        double x = (double)Sx[p] + cfac * D[p];
        double y = (double)Sy[p] + cfac * D[Npad + p];
        double z = (double)Sz[p] + cfac * D[2 * Npad + p];
        const double inr = inv_norm<T>(x * x + y * y + z * z);
        x *= inr; y *= inr; z *= inr;
        qs += (double)g0[p] * x + (double)g0[Npad + p] * y + (double)g0[2 * Npad + p] * z;
        const T tx = (T)x, ty = (T)y, tz = (T)z;
        Sx[p] = tx; Sy[p] = ty; Sz[p] = tz;
        if (row) { row[p] = tx; row[N + p] = ty; row[2 * N + p] = tz; }
      }
      __syncthreads();
      qss_prev = qs;
      t_prev = t;
      if (!fe) {
        double v1[1] = {bond_energy<T>(L, Sx, Sy, Sz) + field_energy_term<T>(L, Sz)};
        block_sum<1>(v1, scratch);
        Eref = v1[0];
      }
    } else {
      if (threadIdx.x == 0) a.qss[(size_t)sidx * a.T + 0] = 1.0;
      if (a.traj) {
        T *row = traj + (size_t)(t - a.chunk_base) * sd.ldm;
        for (int p = threadIdx.x; p < N; p += blockDim.x) {
          row[p] = Sx[p]; row[N + p] = Sy[p]; row[2 * N + p] = Sz[p];
        }
      }
    }
  }
  double e_part = 0.0;
  if (fe) {
    for (int p = threadIdx.x; p < N; p += kSmdThreads) {
      T hx, hy, hz;
      field(p, Sx, Sy, Sz, hx, hy, hz);
      const double x = Sx[p], y = Sy[p], z = Sz[p];
      e_part += 0.5 * (x * hx + y * hy + z * hz) - 0.5 * hext * z;
    }
  }
  double v2[2] = {e_part, qss_prev};
  block_sum<2>(v2, scratch);
  if (fe) Eref = v2[0];
  if (t_prev >= 0 && threadIdx.x == 0) a.qss[(size_t)sidx * a.T + t_prev] = v2[1] / (double)N;
  if (threadIdx.x == 0) a.eref[sidx] = Eref;
}

struct SmdLaunch {
  int S;
  int spt;
  int maxZ;
  int minZ;
  bool general;   // some lattice needs per-slot J tables
  bool use_lds;
  bool wide;      // global-memory wide form
  bool plain;     // one sample per launch (the reference's one instance per rank)
  size_t lds_bytes;
  bool one_class; // LDS form: every workgroup uses the largest slot class
};

template <typename T, int ZC, bool GEN>
void smd_launch_wide_t(const SmdLaunch &c, StepArgs a) {
  if (c.plain) {
    for (int s = 0; s < c.S; s++) {
      a.sample_base = s;
      hipLaunchKernelGGL((smd_kernel_wide<T, ZC, GEN>), dim3(1), dim3(kSmdThreads), 0, 0, a);
    }
  } else {
    a.sample_base = 0;
    hipLaunchKernelGGL((smd_kernel_wide<T, ZC, GEN>), dim3(c.S), dim3(kSmdThreads), 0, 0, a);
  }
}

template <typename T>
void smd_launch_wide(const SmdLaunch &c, const StepArgs &a) {
  const int Z = c.maxZ;
  if (c.general) smd_launch_wide_t<T, 32, true>(c, a);
  else if (Z <= 4) smd_launch_wide_t<T, 4, false>(c, a);
  else if (Z <= 8) smd_launch_wide_t<T, 8, false>(c, a);
  else if (Z <= 12) smd_launch_wide_t<T, 12, false>(c, a);
  else if (Z <= 20) smd_launch_wide_t<T, 20, false>(c, a);
  else smd_launch_wide_t<T, 32, false>(c, a);
}

// LDS-resident form: S, B1, B2 of one sample fit 150 KB only for N <= 4096
// (FP32) or N <= 2048 (FP64), i.e. SPT x sizeof(T) <= 16.
template <typename T, int SPT, int ZC, bool GEN>
void smd_launch_t(const SmdLaunch &c, StepArgs a) {
  if (c.use_lds) {
    if constexpr (SPT * sizeof(T) <= 16) {
      const void *fn = (const void *)smd_kernel<T, SPT, true, ZC, GEN>;
      HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.lds_bytes));
      a.sample_base = 0;
      hipLaunchKernelGGL((smd_kernel<T, SPT, true, ZC, GEN>), dim3(c.S), dim3(kSmdThreads), c.lds_bytes,
                         0, a);
    } else {
      throw Error(HSMD_ERR_ARG, "LDS-resident SMD form requested for a lattice that does not fit");
    }
  } else {
    for (int s = 0; s < c.S; s++) {
      a.sample_base = s;
      hipLaunchKernelGGL((smd_kernel<T, SPT, false, ZC, GEN>), dim3(1), dim3(kSmdThreads), 0, 0, a);
    }
  }
}

// Slot-count classes: 4 (square, kagome, hexagonal and Lieb J1), 8, 12, 20
// (triangular and kagome up to the third shell) and 32; lattices with per-slot
// J tables take the 32-slot form with run-time slot tests.
template <typename T, int SPT, int ZLO, int ZHI>
void smd_launch_two(const SmdLaunch &c, StepArgs a) {
  const void *fn = (const void *)smd_kernel_two<T, SPT, ZLO, ZHI>;
  HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.lds_bytes));
  a.sample_base = 0;
  hipLaunchKernelGGL((smd_kernel_two<T, SPT, ZLO, ZHI>), dim3(c.S), dim3(kSmdThreads), c.lds_bytes, 0,
                     a);
}

template <typename T, int SPT>
void smd_launch_zc(const SmdLaunch &c, const StepArgs &a) {
  const int Z = c.maxZ;
  // two slot classes in one LDS-resident batch (FP32 only)
  if constexpr (sizeof(T) == 4 && SPT <= 4) {
    // a largest class of 9 or 10 slots (the hexagonal lattice with J1 and
    // J2) is summed over 10 slots, not 12
    if (c.use_lds && !c.one_class && !c.general && c.minZ <= 8 && Z > 8 && Z <= 10) {
      if (c.minZ <= 4) smd_launch_two<T, SPT, 4, 10>(c, a);
      else smd_launch_two<T, SPT, 8, 10>(c, a);
      return;
    }
    if (c.use_lds && !c.one_class && !c.general && c.minZ <= 8 && Z > 8 && Z <= 12) {
      if (c.minZ <= 4) smd_launch_two<T, SPT, 4, 12>(c, a);
      else smd_launch_two<T, SPT, 8, 12>(c, a);
      return;
    }
    if (c.use_lds && !c.one_class && !c.general && c.minZ <= 4 && Z > 4 && Z <= 8) {
      smd_launch_two<T, SPT, 4, 8>(c, a);
      return;
    }
  }
  if (c.general) smd_launch_t<T, SPT, 32, true>(c, a);
  else if (Z <= 4) smd_launch_t<T, SPT, 4, false>(c, a);
  else if (Z <= 8) smd_launch_t<T, SPT, 8, false>(c, a);
  else if (Z <= 12) smd_launch_t<T, SPT, 12, false>(c, a);
  else if (Z <= 20) smd_launch_t<T, SPT, 20, false>(c, a);
  else smd_launch_t<T, SPT, 32, false>(c, a);
}

void smd_launch_f32(const SmdLaunch &c, StepArgs a);   // smd_f32.hip
void smd_launch_f64(const SmdLaunch &c, StepArgs a);   // smd_f64.hip

}  // namespace hsmd
