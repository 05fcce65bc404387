// Device-side building blocks shared by the MC and SMD kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hsmd_common.h"
#include "hsmd_internal.h"

namespace hsmd {

constexpr int kZMAX = 32;
constexpr int kCMAX = 64;
constexpr int kBCMAX = 16;

#define HIP_CHECK(x)                                                          \
  do {                                                                        \
    hipError_t e_ = (x);                                                      \
    if (e_ != hipSuccess)                                                     \
      throw ::hsmd::Error(HSMD_ERR_DEVICE, std::string(#x) + ": " +          \
                                               hipGetErrorString(e_));        \
  } while (0)

// Owning device buffer (movable, not copyable).
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
  }
  void upload(const T *h, size_t count) {
    alloc(count);
    if (count) HIP_CHECK(hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void zero() { if (n) HIP_CHECK(hipMemset(p, 0, n * sizeof(T))); }
};

// Couplings of the active neighbour slots of a lattice (new labels).
// Slot s of site p reads neighbour nb[s*N + p]; the sentinel N holds a zero
// spin, so missing neighbours add nothing.
struct Coup {
  int N, Npad, Z;
  int general;               // 1: per-slot J table (disorder / Lieb)
  int h_on;
  double h;                  // external field (z for Heisenberg, Ising axis for Ising)
  double jxy[kZMAX];         // uniform couplings per slot
  double jz[kZMAX];          // uniform z couplings per slot (J * lambda)
  double lam[kZMAX];         // lambda per slot (general mode)
  float fjxy[kZMAX];         // single-precision copies for the FP32 kernels
  float fjz[kZMAX];
  float flam[kZMAX];
};

template <typename T>
__device__ __forceinline__ T cj_xy(const Coup &cp, int s);
template <>
__device__ __forceinline__ double cj_xy<double>(const Coup &cp, int s) { return cp.jxy[s]; }
template <>
__device__ __forceinline__ float cj_xy<float>(const Coup &cp, int s) { return cp.fjxy[s]; }
template <typename T>
__device__ __forceinline__ T cj_z(const Coup &cp, int s);
template <>
__device__ __forceinline__ double cj_z<double>(const Coup &cp, int s) { return cp.jz[s]; }
template <>
__device__ __forceinline__ float cj_z<float>(const Coup &cp, int s) { return cp.fjz[s]; }
template <typename T>
__device__ __forceinline__ T cj_lam(const Coup &cp, int s);
template <>
__device__ __forceinline__ double cj_lam<double>(const Coup &cp, int s) { return cp.lam[s]; }
template <>
__device__ __forceinline__ float cj_lam<float>(const Coup &cp, int s) { return cp.flam[s]; }

// Fills the single-precision copies of the slot couplings.
inline void finish_coup(Coup &cp) {
  for (int s = 0; s < kZMAX; s++) {
    cp.fjxy[s] = (float)cp.jxy[s];
    cp.fjz[s] = (float)cp.jz[s];
    cp.flam[s] = (float)cp.lam[s];
  }
}

struct Bonds {
  int Nb;
  int n_class;
  double jxy[kBCMAX];
  double jz[kBCMAX];
};

// Device-resident bond list of an MC engine (new labels) with the map back
// to the ABI's site order, for entry points on caller-owned device buffers.
struct McBondsDev {
  const int32_t *bi, *bj;
  const uint8_t *bc;
  const int32_t *perm;   // new label -> site
  int N;
  Bonds bd;
  int h_on;
  double h;
};

// h_i of the reference's get_localField (CF/subrouts.cpp:4465-4575):
// start from -h on z, then add shells in slot order.  ZC is a compile-time
// bound on the slot count, so every neighbour index is loaded up front and
// the gathers overlap instead of waiting on one index at a time.
template <typename T, int ZC>
__device__ __forceinline__ void local_field(const Coup &cp, const int32_t *__restrict__ nb,
                                            const T *__restrict__ Jt, int p,
                                            const T *sx, const T *sy, const T *sz,
                                            T &hx, T &hy, T &hz) {
  const int N = cp.N, Z = cp.Z;
  int q[ZC];
#pragma unroll
  for (int s = 0; s < ZC; s++) q[s] = (s < Z) ? nb[(size_t)s * N + p] : N;
  hx = T(0);
  hy = T(0);
  hz = cp.h_on ? T(-cp.h) : T(0);
  if (!cp.general) {
#pragma unroll
    for (int s = 0; s < ZC; s++) {
      if (s < Z) {
        const T J = cj_xy<T>(cp, s), Jz = cj_z<T>(cp, s);
        hx += J * sx[q[s]];
        hy += J * sy[q[s]];
        hz += Jz * sz[q[s]];
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < ZC; s++) {
      if (s < Z) {
        const T J = Jt[(size_t)s * N + p];
        const T Jz = J * cj_lam<T>(cp, s);
        hx += J * sx[q[s]];
        hy += J * sy[q[s]];
        hz += Jz * sz[q[s]];
      }
    }
  }
}

// Compile-time slot bound for a lattice with Z active slots.
inline int zc_for(int Z) { return Z <= 8 ? 8 : Z <= 16 ? 16 : Z <= 20 ? 20 : 32; }

#define HSMD_ZC_DISPATCH(Z, FN, ...)             \
  do {                                           \
    switch (::hsmd::zc_for(Z)) {                 \
      case 8: FN<8>(__VA_ARGS__); break;         \
      case 16: FN<16>(__VA_ARGS__); break;       \
      case 20: FN<20>(__VA_ARGS__); break;       \
      default: FN<32>(__VA_ARGS__); break;       \
    }                                            \
  } while (0)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-wide FP64 sums of NV values; result valid in every thread.
// scratch: at least NV * 16 doubles of shared memory.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *scratch) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) scratch[k * 16 + w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
    for (int i = 0; i < nw; i++) s += scratch[k * 16 + i];
    v[k] = s;
  }
  __syncthreads();
}

// block_sum with one barrier: for callers that alternate between two scratch
// halves (a half is written again only after every wave has passed the
// barrier of the call in between, so no wave still reads it) and need no
// barrier after the sums.
template <int NV>
__device__ __forceinline__ void block_sum_once(double (&v)[NV], double *scratch) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) scratch[k * 16 + w] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
    for (int i = 0; i < nw; i++) s += scratch[k * 16 + i];
    v[k] = s;
  }
}

}  // namespace hsmd
