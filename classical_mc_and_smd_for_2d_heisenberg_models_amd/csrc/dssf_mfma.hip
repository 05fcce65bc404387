// DSSF time contraction on the matrix cores: split-precision FP16 MFMA with
// FP32 accumulation.
//
// The SMD engine accumulates, per sample, Y[m][j] += sum_k A[k][m] B[k][j]
// where A is a chunk of the trajectory (time k, spin component x site m,
// FP32) and B the windowed DFT twiddles (time k, frequency bin j with cos in
// the first n_omega columns and -sin in the rest).  Each FP32 operand is split
// into two FP16 halves, x = x_hi + x_lo with x_lo = fp16(x - x_hi), and the
// product is formed as
//   hi*hi + hi*lo + lo*hi     (the lo*lo term is below FP32 rounding)
// with three MFMAs per block and FP32 accumulation, which keeps about 21
// significant bits -- FP32-level accuracy for operands in [-1, 1].  (Emulated
// on the CPU for a 2381-point trajectory of the square L=48 / kagome L=32
// batch, on the bins above 1e-2 of their peak: one FP16 product leaves a
// relative error of 3.7e-3, hi*hi + lo*hi with FP16 twiddles 4e-4, the three
// products 8e-7.)  This
// replaces the reference's per-q MKL FFT over time (CF/subrouts.cpp:12062-12105);
// the DFT is exact for every bin kept.
//
// Kernels: time_gemm_mfma_kernel, a 64 x 64 tile with on-the-fly splitting
// (x_hi = fp16(x), v_mfma_f32_16x16x16_f16), kept for bin counts that are not
// a multiple of 128; the 128-row tiles (time_gemm_mfma_q / q256 / db /
// w kernels: x_hi = x cut to 11 significant bits, v_mfma_f32_16x16x32_f16,
// twiddles pre-split on the host), each described where it is defined.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "hsmd_internal.h"

namespace hsmd {

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

struct GemmSample {
  int M;                 // rows of Y (3 N)
  int lda;               // row stride of A: M rounded up to a multiple of 4
  long long a_off;       // into A (chunk, [K][lda])
  long long y_off;       // into Y ([M][ldb])
};

constexpr int TM = 64, TJ = 64, TK = 16, KP = TK + 4;  // KP: padded k stride (halves)

// Y update of one accumulator: the first chunk of a run (k0 == 0) stores, later
// chunks add with a no-return global atomic.  Each element of Y has exactly one
// owning lane per launch, so the sum is the same as a read-add-write, but the
// wave issues its 128 updates without waiting for reads of Y (a read-modify-
// write epilogue left the matrix cores idle for tens of microseconds per tile).
__device__ __forceinline__ void y_update(float *p, float v, bool first) {
  if (first) *p = v;
  else unsafeAtomicAdd(p, v);
}

__global__ __launch_bounds__(256) void time_gemm_mfma_kernel(const GemmSample *samples,
                                                             const float *__restrict__ A,
                                                             const float *__restrict__ Bm, int ldb,
                                                             int k0, int K, float *Y) {
  __shared__ __align__(16) _Float16 Ah[TM * KP], Al[TM * KP];
  __shared__ __align__(16) _Float16 Bh[TJ * KP], Bl[TJ * KP];
  const GemmSample sd = samples[blockIdx.z];
  const int M = sd.M;
  const int m0 = blockIdx.x * TM, j0 = blockIdx.y * TJ;
  if (m0 >= M) return;
  const float *Ab = A + sd.a_off;
  float *Yb = Y + sd.y_off;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wj = (wave & 1) * 32;   // wave's quarter
  const int li = lane & 15, lk = (lane >> 4) * 4;           // fragment coordinates
  float4_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = (float4_t){0.f, 0.f, 0.f, 0.f};
  for (int kk = 0; kk < K; kk += TK) {
    // stage A (16 x 64) and B (16 x 64): 1024 elements each, 4 per thread
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int e = tid + r * 256;
      const int k = e >> 6, col = e & 63;
      const int kg = kk + k;
      const int m = m0 + col;
      const float av = (kg < K && m < M) ? Ab[(size_t)kg * sd.lda + m] : 0.f;
      const float bv = (kg < K) ? Bm[(size_t)(k0 + kg) * ldb + j0 + col] : 0.f;
      const _Float16 ah = (_Float16)av;
      const _Float16 bh = (_Float16)bv;
      Ah[col * KP + k] = ah;
      Al[col * KP + k] = (_Float16)(av - (float)ah);
      Bh[col * KP + k] = bh;
      Bl[col * KP + k] = (_Float16)(bv - (float)bh);
    }
    __syncthreads();
    half4_t ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int a = 0; a < 2; a++) {
      const int row = wm + a * 16 + li;
      ah[a] = *(const half4_t *)&Ah[row * KP + lk];
      al[a] = *(const half4_t *)&Al[row * KP + lk];
    }
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int col = wj + b * 16 + li;
      bh[b] = *(const half4_t *)&Bh[col * KP + lk];
      bl[b] = *(const half4_t *)&Bl[col * KP + lk];
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16f16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
    __syncthreads();
  }
  // D layout of the 16x16 block: lane holds rows 4*(lane/16) + r, column lane % 16
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int j = j0 + wj + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + a * 16 + (lane >> 4) * 4 + r;
        if (m < M) y_update(&Yb[(size_t)m * ldb + j], acc[a][b][r], k0 == 0);
      }
    }
}

// Host launcher: samples[S] on the device, Y accumulated in place.
void launch_time_gemm_mfma(const void *samples_dev, int S, int maxM, const float *A, const float *B,
                           int ldb, int k0, int K, float *Y, hipStream_t st) {
  if (ldb % TJ != 0) throw Error(HSMD_ERR_ARG, "2 * n_omega must be a multiple of 64");
  dim3 grid((maxM + TM - 1) / TM, ldb / TJ, S);
  hipLaunchKernelGGL(time_gemm_mfma_kernel, grid, dim3(256), 0, st,
                     (const GemmSample *)samples_dev, A, B, ldb, k0, K, Y);
}


typedef _Float16 half8_t __attribute__((ext_vector_type(8)));

constexpr int WK = 32, WKP = WK + 8;

// Main variant (2 n_omega a multiple of 128): a workgroup of 4 waves owns a
// 128 (m) x 128 (j) tile of Y, each wave a 64 x 64 sub-tile of 4 x 4 MFMA
// blocks.  The twiddles arrive pre-split (FP16 hi/lo, k contiguous per
// column) and the gfx950 v_mfma_f32_16x16x32_f16 consumes 32 time points per
// instruction.  The global loads of step k+1 (sixteen trajectory values and
// four 16-byte twiddle vectors per thread) are issued into registers before
// the MFMAs of step k; the single 40 KB LDS stage lets three workgroups share
// a CU (3 waves per SIMD), so one workgroup's LDS and barrier phases overlap
// another's matrix work.  Staging roles are chosen so that the 16 lanes of
// each LDS write phase touch 16 consecutive rows (20-dword row stride:
// conflict-free 16-byte writes).  Workgroup ids are XCD-aware: workgroups are
// dispatched round-robin over the 8 XCDs, so the frequency tiles of one row
// block get ids that land on the same XCD one after another and share the
// trajectory tile in that XCD's L2 (row-block count padded to a multiple of
// 8).  Measured on MI355X: 47 ms per 1024-point chunk of the 512-sample
// batch, 652 TFLOP/s counting the three split products.
constexpr int QM = 128, QT = 256;

// TJ = 128 (the main tile) or 256 (a wider tile: each trajectory value is
// split and staged half as often and the chunk is read from memory half as
// often, at 128 accumulator registers per lane and two workgroups per CU).
template <int TJ>
__device__ __forceinline__ void time_gemm_q_body(const GemmSample *samples, const float *__restrict__ A,
                                                 const _Float16 *__restrict__ Bth,
                                                 const _Float16 *__restrict__ Btl, int ldt, int k0,
                                                 int K, int ldb, float *Y, _Float16 *Ah, _Float16 *Al,
                                                 _Float16 *Bh, _Float16 *Bl) {
  constexpr int NB = TJ / 32;               // 16-column blocks per wave
  constexpr int BK = WK * TJ / QT;          // time points each thread stages for B
  const GemmSample sd = samples[blockIdx.z];
  const int M = sd.M, lda = sd.lda;
  const int nj = ldb / TJ;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int m0 = ((slot / nj) * 8 + xcd) * QM, j0 = (slot % nj) * TJ;
  if (m0 >= M) return;
  const float *Ab = A + sd.a_off;
  float *Yb = Y + sd.y_off;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 64, wj = (wave & 1) * (TJ / 2);
  const int li = lane & 15, lk = (lane >> 4) * 8;
  // staging: thread t owns row am = t % 128 (time points 16 kh .. + 15) and
  // column bc = t % TJ (BK time points from kb); 16 lanes write 16 rows
  const int am = tid % QM;
  const int kh = __builtin_amdgcn_readfirstlane(tid / QM);   // uniform per wave
  const bool arow = m0 + am < M;
  // trajectory loads of full time steps go through a buffer resource: the
  // column (this thread's row of Y, clamped into the row stride; rows past M
  // are computed and never stored) is the per-lane offset, the time row the
  // per-wave scalar offset, so the sixteen loads need no address arithmetic
  // and no bounds branches
  const __amdgpu_buffer_rsrc_t arsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)Ab, (short)0, 0x7fffffff, 0x00020000);
  const int acol = 4 * min(m0 + am, lda - 1);
  const int lda4 = 4 * lda;
  const int bc = tid % TJ, kb = (tid / TJ) * BK;
  const size_t bsrc = (size_t)(j0 + bc) * ldt + (size_t)(k0 + kb);
  float4_t acc[4][NB];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < NB; b++) acc[a][b] = (float4_t){0.f, 0.f, 0.f, 0.f};
  static_assert(BK == 16 || BK == 32, "B staging covers 16 or 32 time points per thread");
  float ra[16];
  uint4 rbh0, rbh1, rbl0, rbl1, rbh2, rbh3, rbl2, rbl3;
  auto fetch = [&](int kk) {
    if (kk + WK <= K) {
      const int r0 = (kk + 16 * kh) * lda4;
#pragma unroll
      for (int e = 0; e < 16; e++)
        ra[e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(arsrc, acol, r0 + e * lda4, 0));
    } else {
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int ka = kk + 16 * kh + e;
        ra[e] = (arow && ka < K) ? Ab[(size_t)ka * lda + m0 + am] : 0.f;
      }
    }
    const uint4 *h = (const uint4 *)(Bth + bsrc + kk);
    const uint4 *l = (const uint4 *)(Btl + bsrc + kk);
    rbh0 = h[0]; rbh1 = h[1];
    rbl0 = l[0]; rbl1 = l[1];
    if constexpr (BK == 32) {
      rbh2 = h[2]; rbh3 = h[3];
      rbl2 = l[2]; rbl3 = l[3];
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int g = 0; g < 2; g++) {
      half8_t hv, lv;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        // hi: the value cut to FP16's 11 significant bits (exact in FP16 for
        // |x| >= 2^-14); lo: the exact FP32 remainder, rounded to FP16
        const float x = ra[8 * g + e];
        const float h = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) & 0xffffe000u);
        hv[e] = (_Float16)h;
        lv[e] = (_Float16)(x - h);
      }
      *(half8_t *)&Ah[am * WKP + 16 * kh + 8 * g] = hv;
      *(half8_t *)&Al[am * WKP + 16 * kh + 8 * g] = lv;
    }
    uint4 *dh = (uint4 *)&Bh[bc * WKP + kb];
    uint4 *dl = (uint4 *)&Bl[bc * WKP + kb];
    dh[0] = rbh0; dh[1] = rbh1;
    dl[0] = rbl0; dl[1] = rbl1;
    if constexpr (BK == 32) {
      dh[2] = rbh2; dh[3] = rbh3;
      dl[2] = rbl2; dl[3] = rbl3;
    }
  };
  fetch(0);
  for (int kk = 0; kk < K; kk += WK) {
    stage();
    __syncthreads();
    if (kk + WK < K) fetch(kk + WK);
    half8_t ah[4], al[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int row = wm + a * 16 + li;
      ah[a] = *(const half8_t *)&Ah[row * WKP + lk];
      al[a] = *(const half8_t *)&Al[row * WKP + lk];
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int col = wj + b * 16 + li;
      const half8_t bh = *(const half8_t *)&Bh[col * WKP + lk];
      const half8_t bl = *(const half8_t *)&Bl[col * WKP + lk];
#pragma unroll
      for (int a = 0; a < 4; a++) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int j = j0 + wj + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + a * 16 + (lane >> 4) * 4 + r;
        if (m < M) y_update(&Yb[(size_t)m * ldb + j], acc[a][b][r], k0 == 0);
      }
    }
}

__global__ __launch_bounds__(QT) __attribute__((amdgpu_waves_per_eu(3, 3)))
void time_gemm_mfma_q_kernel(const GemmSample *samples, const float *__restrict__ A,
                             const _Float16 *__restrict__ Bth, const _Float16 *__restrict__ Btl,
                             int ldt, int k0, int K, int ldb, float *Y) {
  __shared__ __align__(16) _Float16 Ah[QM * WKP], Al[QM * WKP];
  __shared__ __align__(16) _Float16 Bh[128 * WKP], Bl[128 * WKP];
  time_gemm_q_body<128>(samples, A, Bth, Btl, ldt, k0, K, ldb, Y, Ah, Al, Bh, Bl);
}

__global__ __launch_bounds__(QT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void time_gemm_mfma_q256_kernel(const GemmSample *samples, const float *__restrict__ A,
                                const _Float16 *__restrict__ Bth, const _Float16 *__restrict__ Btl,
                                int ldt, int k0, int K, int ldb, float *Y) {
  __shared__ __align__(16) _Float16 Ah[QM * WKP], Al[QM * WKP];
  __shared__ __align__(16) _Float16 Bh[256 * WKP], Bl[256 * WKP];
  time_gemm_q_body<256>(samples, A, Bth, Btl, ldt, k0, K, ldb, Y, Ah, Al, Bh, Bl);
}

// Double-buffered form of the 128 x 128 tile: two LDS stages (80 KB, two
// workgroups per CU).  Step k's MFMAs read one stage while the same waves
// convert and store step k+1's tiles into the other, so the FP16 split and
// the LDS writes issue between matrix instructions, and one barrier per step
// separates the stages.
__global__ __launch_bounds__(QT) __attribute__((amdgpu_waves_per_eu(2, 2)))
void time_gemm_mfma_db_kernel(const GemmSample *samples, const float *__restrict__ A,
                              const _Float16 *__restrict__ Bth, const _Float16 *__restrict__ Btl,
                              int ldt, int k0, int K, int ldb, float *Y) {
  constexpr int TJ = 128, NB = TJ / 32, BK = WK * TJ / QT;
  __shared__ __align__(16) _Float16 Ah[2][QM * WKP], Al[2][QM * WKP];
  __shared__ __align__(16) _Float16 Bh[2][TJ * WKP], Bl[2][TJ * WKP];
  const GemmSample sd = samples[blockIdx.z];
  const int M = sd.M, lda = sd.lda;
  const int nj = ldb / TJ;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int m0 = ((slot / nj) * 8 + xcd) * QM, j0 = (slot % nj) * TJ;
  if (m0 >= M) return;
  const float *Ab = A + sd.a_off;
  float *Yb = Y + sd.y_off;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 64, wj = (wave & 1) * (TJ / 2);
  const int li = lane & 15, lk = (lane >> 4) * 8;
  const int am = tid % QM;
  const int kh = __builtin_amdgcn_readfirstlane(tid / QM);
  const bool arow = m0 + am < M;
  const __amdgpu_buffer_rsrc_t arsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)Ab, (short)0, 0x7fffffff, 0x00020000);
  const int acol = 4 * min(m0 + am, lda - 1);
  const int lda4 = 4 * lda;
  const int bc = tid % TJ, kb = (tid / TJ) * BK;
  const size_t bsrc = (size_t)(j0 + bc) * ldt + (size_t)(k0 + kb);
  float4_t acc[4][NB];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < NB; b++) acc[a][b] = (float4_t){0.f, 0.f, 0.f, 0.f};
  float ra[16];
  uint4 rbh0, rbh1, rbl0, rbl1;
  auto fetch = [&](int kk) {
    if (kk + WK <= K) {
      const int r0 = (kk + 16 * kh) * lda4;
#pragma unroll
      for (int e = 0; e < 16; e++)
        ra[e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(arsrc, acol, r0 + e * lda4, 0));
    } else {
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int ka = kk + 16 * kh + e;
        ra[e] = (arow && ka < K) ? Ab[(size_t)ka * lda + m0 + am] : 0.f;
      }
    }
    const uint4 *h = (const uint4 *)(Bth + bsrc + kk);
    const uint4 *l = (const uint4 *)(Btl + bsrc + kk);
    rbh0 = h[0]; rbh1 = h[1];
    rbl0 = l[0]; rbl1 = l[1];
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int g = 0; g < 2; g++) {
      half8_t hv, lv;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const float x = ra[8 * g + e];
        const float h = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) & 0xffffe000u);
        hv[e] = (_Float16)h;
        lv[e] = (_Float16)(x - h);
      }
      *(half8_t *)&Ah[buf][am * WKP + 16 * kh + 8 * g] = hv;
      *(half8_t *)&Al[buf][am * WKP + 16 * kh + 8 * g] = lv;
    }
    uint4 *dh = (uint4 *)&Bh[buf][bc * WKP + kb];
    uint4 *dl = (uint4 *)&Bl[buf][bc * WKP + kb];
    dh[0] = rbh0; dh[1] = rbh1;
    dl[0] = rbl0; dl[1] = rbl1;
  };
  fetch(0);
  stage(0);
  __syncthreads();
  int cur = 0;
  for (int kk = 0; kk < K; kk += WK) {
    const bool more = kk + WK < K;
    if (more) fetch(kk + WK);
    half8_t ah[4], al[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int row = wm + a * 16 + li;
      ah[a] = *(const half8_t *)&Ah[cur][row * WKP + lk];
      al[a] = *(const half8_t *)&Al[cur][row * WKP + lk];
    }
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int col = wj + b * 16 + li;
      const half8_t bh = *(const half8_t *)&Bh[cur][col * WKP + lk];
      const half8_t bl = *(const half8_t *)&Bl[cur][col * WKP + lk];
#pragma unroll
      for (int a = 0; a < 4; a++) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh, acc[a][b], 0, 0, 0);
      }
    }
    if (more) stage(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < NB; b++) {
      const int j = j0 + wj + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + a * 16 + (lane >> 4) * 4 + r;
        if (m < M) y_update(&Yb[(size_t)m * ldb + j], acc[a][b][r], k0 == 0);
      }
    }
}

// Wide variant (2 n_omega a multiple of 512): a workgroup of 8 waves owns a
// 128 (m) x 512 (j) tile of Y, each wave a 64 x 128 sub-tile of 4 x 8 MFMA
// blocks (128 accumulator registers), so at 512 bins the trajectory is read
// twice per chunk instead of eight times, and each trajectory value is split
// into its FP16 halves twice instead of eight times.  The twiddles come in a
// chunk-major layout, [t / 8][2 n_omega][8] FP16 hi and lo: the eight time
// points of one column are 16 bytes, and a wave's 64 lanes take 64
// consecutive columns of one time chunk -- 1 KB contiguous in memory and in
// LDS -- so the loads are coalesced, the LDS writes conflict-free, and the
// MFMA B fragments (16 columns x 8 time points per 16-lane group) read 256
// contiguous bytes per lane group.  Step k + 1's trajectory tile (128 rows x
// 32 time points, FP32) and twiddles are loaded into registers while step
// k's MFMAs run, then split / stored to the single LDS stage (20 KB +
// 64 KB).  One workgroup per CU, 2 waves per SIMD.
constexpr int W_TJ = 512, W_THREADS = 512;
constexpr int W_BSTAGE = 4 * W_TJ * 8;   // halves per (hi or lo) twiddle stage: 4 chunks x 512 columns x 8

__global__ __launch_bounds__(W_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
void time_gemm_mfma_w_kernel(const GemmSample *samples, const float *__restrict__ A,
                             const _Float16 *__restrict__ B8h, const _Float16 *__restrict__ B8l,
                             int k0, int K, int ldb, float *Y) {
  __shared__ __align__(16) _Float16 Ah[QM * WKP], Al[QM * WKP];
  __shared__ __align__(16) _Float16 Bs[2][W_BSTAGE];   // [hi, lo][chunk][col][8]
  const GemmSample sd = samples[blockIdx.z];
  const int M = sd.M, lda = sd.lda;
  const int nj = ldb / W_TJ;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int m0 = ((slot / nj) * 8 + xcd) * QM, j0 = (slot % nj) * W_TJ;
  if (m0 >= M) return;
  const float *Ab = A + sd.a_off;
  float *Yb = Y + sd.y_off;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (wave >> 2) * 64, wj = (wave & 3) * 128;
  const int li = lane & 15, lg = lane >> 4;
  // trajectory staging: thread t owns row am = t % 128, time points 8 kh .. 8 kh + 7
  const int am = tid % QM;
  const int kh = __builtin_amdgcn_readfirstlane(tid / QM);   // uniform per wave
  const bool arow = m0 + am < M;
  const __amdgpu_buffer_rsrc_t arsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)Ab, (short)0, 0x7fffffff, 0x00020000);
  const int acol = 4 * min(m0 + am, lda - 1);
  const int lda4 = 4 * lda;
  // twiddle staging: wave w copies region w of the stage (time chunk w % 4 of
  // the hi table for w < 4, of the lo table otherwise), eight 1 KB pieces
  const int bchunk = wave & 3;
  const uint4 *bsrc = (const uint4 *)((wave < 4 ? B8h : B8l) + (size_t)bchunk * ldb * 8) + j0 + lane;
  uint4 *bdst = (uint4 *)&Bs[wave >> 2][bchunk * W_TJ * 8] + lane;
  float4_t acc[4][8];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 8; b++) acc[a][b] = (float4_t){0.f, 0.f, 0.f, 0.f};
  float ra[8];
  // named registers, not an array: an array captured by the lambdas below
  // would be placed in LDS by the compiler, 128 bytes per thread
  uint4 rb0, rb1, rb2, rb3, rb4, rb5, rb6, rb7;
  auto fetch = [&](int kk) {
    if (kk + WK <= K) {
      const int r0 = (kk + 8 * kh) * lda4;
#pragma unroll
      for (int e = 0; e < 8; e++)
        ra[e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(arsrc, acol, r0 + e * lda4, 0));
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int ka = kk + 8 * kh + e;
        ra[e] = (arow && ka < K) ? Ab[(size_t)ka * lda + m0 + am] : 0.f;
      }
    }
    // time chunk (k0 + kk) / 8 + bchunk, columns j0 .. j0 + 511
    const uint4 *g = bsrc + (size_t)((k0 + kk) / 8) * ldb;
    rb0 = g[0]; rb1 = g[64]; rb2 = g[128]; rb3 = g[192];
    rb4 = g[256]; rb5 = g[320]; rb6 = g[384]; rb7 = g[448];
  };
  auto stage = [&]() {
    half8_t hv, lv;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      // hi: the value cut to FP16's 11 significant bits (exact in FP16 for
      // |x| >= 2^-14); lo: the exact FP32 remainder, rounded to FP16
      const float x = ra[e];
      const float h = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) & 0xffffe000u);
      hv[e] = (_Float16)h;
      lv[e] = (_Float16)(x - h);
    }
    *(half8_t *)&Ah[am * WKP + 8 * kh] = hv;
    *(half8_t *)&Al[am * WKP + 8 * kh] = lv;
    bdst[0] = rb0; bdst[64] = rb1; bdst[128] = rb2; bdst[192] = rb3;
    bdst[256] = rb4; bdst[320] = rb5; bdst[384] = rb6; bdst[448] = rb7;
  };
  fetch(0);
  for (int kk = 0; kk < K; kk += WK) {
    stage();
    __syncthreads();
    if (kk + WK < K) fetch(kk + WK);
    half8_t ah[4], al[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int row = wm + a * 16 + li;
      ah[a] = *(const half8_t *)&Ah[row * WKP + 8 * lg];
      al[a] = *(const half8_t *)&Al[row * WKP + 8 * lg];
    }
    const _Float16 *bh0 = &Bs[0][lg * W_TJ * 8];
    const _Float16 *bl0 = &Bs[1][lg * W_TJ * 8];
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const int col = wj + b * 16 + li;
      const half8_t bh = *(const half8_t *)&bh0[col * 8];
      const half8_t bl = *(const half8_t *)&bl0[col * 8];
#pragma unroll
      for (int a = 0; a < 4; a++) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const int j = j0 + wj + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + a * 16 + lg * 4 + r;
        if (m < M) y_update(&Yb[(size_t)m * ldb + j], acc[a][b][r], k0 == 0);
      }
    }
}

// Pipelined form of the 128 x 512 tile: four waves, one per SIMD, each
// owning all 128 rows and 128 of the columns (8 x 8 MFMA blocks, 256
// accumulators).  The twiddles a wave multiplies are its own (no other wave
// uses its columns), so they skip LDS: each wave loads its FP16 hi / lo
// fragments straight from the chunk-major table into registers (16-byte
// loads, 256 contiguous bytes per 16-lane group), one 32-point step ahead.
// Only the trajectory tile, which all four waves share, passes through LDS,
// in a ring of three 16 KB stages: during step k the waves multiply step k's
// fragments (in registers), refill each row block's fragments from stage
// k + 1 as soon as its products are issued, and split / store step k + 2's
// trajectory values (loaded one step earlier) into the third stage; one
// barrier per step.  LDS traffic per step and CU is 64 KB of reads and 16 KB
// of writes instead of the 192 KB + 80 KB of the single-stage tile.  The
// products, their order per accumulator and the Y update are those of the
// 128 x 512 kernel above, so the two give identical sums.  Measured on
// MI355X: 50.3 ms per 2048-point chunk of the 512-sample batch against
// 49.8 ms for the single-stage tile (a copy with the twiddles staged in LDS
// by direct-to-LDS loads over two 80 KB stages: 52 ms).  In a probe of the
// same instruction stream the matrix instructions alone ran at 2.14 PFLOP/s,
// and 28 % faster than the full kernel when every load hit the cache: with
// one wave per SIMD a one-step-ahead prefetch does not cover the loads'
// latency.  The autotuner times it beside the other tiles.
constexpr int P_THREADS = 256;
constexpr int P_AST = 2 * 4 * QM * 8;     // halves of a trajectory stage [hi, lo][4 chunks][128 rows][8]

__global__ __launch_bounds__(P_THREADS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void time_gemm_mfma_p_kernel(const GemmSample *samples, const float *__restrict__ A,
                             const _Float16 *__restrict__ B8h, const _Float16 *__restrict__ B8l,
                             int k0, int K, int ldb, float *Y) {
  __shared__ __align__(16) _Float16 As[3 * P_AST];
  const GemmSample sd = samples[blockIdx.z];
  const int M = sd.M, lda = sd.lda;
  const int nj = ldb / W_TJ;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int m0 = ((slot / nj) * 8 + xcd) * QM, j0 = (slot % nj) * W_TJ;
  if (m0 >= M) return;
  const float *Ab = A + sd.a_off;
  float *Yb = Y + sd.y_off;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wj = wave * 128;
  const int li = lane & 15, lg = lane >> 4;
  // trajectory staging: thread t owns row am = t % 128, time points 16 kh .. 16 kh + 15
  const int am = tid & 127;
  const int kh = __builtin_amdgcn_readfirstlane(tid >> 7);
  const __amdgpu_buffer_rsrc_t arsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)Ab, (short)0, 0x7fffffff, 0x00020000);
  const int acol = 4 * min(m0 + am, lda - 1);
  const int lda4 = 4 * lda;
  // twiddle fragments: lane (li, lg) of column block b reads table row
  // (k0 + 32 st) / 8 + lg, column j0 + wj + 16 b + li: the per-lane part of
  // the byte offset is fixed, the step's row the scalar part, the block the
  // instruction's immediate offset
  const __amdgpu_buffer_rsrc_t hrsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)B8h, (short)0, 0x7fffffff, 0x00020000);
  const __amdgpu_buffer_rsrc_t lrsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)B8l, (short)0, 0x7fffffff, 0x00020000);
  const int bvoff = lg * ldb * 16 + (j0 + wj + li) * 16;
  const int nsteps = (K + WK - 1) / WK;
  float4_t acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int b = 0; b < 8; b++) acc[a][b] = (float4_t){0.f, 0.f, 0.f, 0.f};

  // every step issues the same loads (so the compiler's wait counts are
  // exact): time rows past the chunk are clamped to its last row, and their
  // values are zeroed when split (lim: the number of valid rows); the time
  // index is uniform over the wave, so the condition is scalar (rows past M
  // are clamped into the row stride, computed and never stored)
  auto load_a = [&](float (&ra)[16], int &lim, int st) __attribute__((always_inline)) {
    const int k1 = st * WK + 16 * kh;
    lim = K - k1;
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const int r = min(k1 + e, K - 1) * lda4;
      ra[e] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(arsrc, acol, r, 0));
    }
  };
  auto store_a = [&](const float (&ra)[16], int lim, _Float16 *as) __attribute__((always_inline)) {
#pragma unroll
    for (int g = 0; g < 2; g++) {
      half8_t hv, lv;
#pragma unroll
      for (int e = 0; e < 8; e++) {
        // hi: the value cut to FP16's 11 significant bits (exact in FP16 for
        // |x| >= 2^-14); lo: the exact FP32 remainder, rounded to FP16
        const float x = 8 * g + e < lim ? ra[8 * g + e] : 0.f;
        const float h = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, x) & 0xffffe000u);
        hv[e] = (_Float16)h;
        lv[e] = (_Float16)(x - h);
      }
      *(half8_t *)&as[((2 * kh + g) * QM + am) * 8] = hv;
      *(half8_t *)&as[((4 + 2 * kh + g) * QM + am) * 8] = lv;
    }
  };
  auto load_b = [&](half8_t (&bh)[8], half8_t (&bl)[8], int st) __attribute__((always_inline)) {
    const int soff = ((k0 + st * WK) / 8) * ldb * 16;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      bh[b] = __builtin_bit_cast(half8_t, __builtin_amdgcn_raw_buffer_load_b128(hrsrc, bvoff, soff + b * 256, 0));
      bl[b] = __builtin_bit_cast(half8_t, __builtin_amdgcn_raw_buffer_load_b128(lrsrc, bvoff, soff + b * 256, 0));
    }
  };
  auto read_a = [&](half8_t &ah, half8_t &al, const _Float16 *as, int a) __attribute__((always_inline)) {
    const int row = a * 16 + li;
    al = *(const half8_t *)&as[((4 + lg) * QM + row) * 8];
    ah = *(const half8_t *)&as[(lg * QM + row) * 8];
  };
  auto barrier = [&]() __attribute__((always_inline)) {
    // LDS stores done, then the workgroup barrier; the memory clobber keeps
    // the compiler from moving LDS accesses across it
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  };

  float ra[16];
  int lim;
  half8_t bh[8], bl[8], nbh[8], nbl[8], ah[8], al[8];
  // prologue: stages 0 and 1, step 0's twiddles, step 2's trajectory values
  load_b(bh, bl, 0);
  load_a(ra, lim, 0);
  store_a(ra, lim, As);
  if (nsteps > 1) {
    load_a(ra, lim, 1);
    store_a(ra, lim, As + P_AST);
  }
  load_a(ra, lim, min(2, nsteps - 1));
  barrier();
#pragma unroll
  for (int a = 0; a < 8; a++) read_a(ah[a], al[a], As, a);
  int cur = 0;
#pragma unroll 1
  for (int st = 0; st < nsteps; st++) {
    const int n1 = cur == 2 ? 0 : cur + 1, n2 = n1 == 2 ? 0 : n1 + 1;
    const bool more = st + 1 < nsteps;
    load_b(nbh, nbl, min(st + 1, nsteps - 1));
#pragma unroll
    for (int a = 0; a < 8; a++) {
#pragma unroll
      for (int b = 0; b < 8; b++) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[a], bh[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bl[b], acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[a], bh[b], acc[a][b], 0, 0, 0);
      }
      // this row block's products are issued: its registers take step
      // st + 1's fragments (stage n1, complete since the last barrier)
      if (more) read_a(ah[a], al[a], As + n1 * P_AST, a);
    }
    if (st + 2 < nsteps) store_a(ra, lim, As + n2 * P_AST);
    load_a(ra, lim, min(st + 3, nsteps - 1));
#pragma unroll
    for (int b = 0; b < 8; b++) { bh[b] = nbh[b]; bl[b] = nbl[b]; }
    barrier();
    cur = n1;
  }
  // Y update: one row pointer per (a, r), the eight column blocks at
  // constant offsets from it
  float *yrow = Yb + (size_t)(m0 + lg * 4) * ldb + j0 + wj + li;
  const bool first = k0 == 0;
#pragma unroll
  for (int a = 0; a < 8; a++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (m0 + a * 16 + lg * 4 + r < M) {
        float *yp = yrow + (size_t)(a * 16 + r) * ldb;
        if (first) {
#pragma unroll
          for (int b = 0; b < 8; b++) yp[b * 16] = acc[a][b][r];
        } else {
#pragma unroll
          for (int b = 0; b < 8; b++) unsafeAtomicAdd(yp + b * 16, acc[a][b][r]);
        }
      }
    }
}

// B8h / B8l: twiddles in the chunk-major layout [T' / 8][ldb][8] (T' a multiple of 32)
void launch_time_gemm_mfma_w(const void *samples_dev, int S, int maxM, const float *A,
                             const _Float16 *B8h, const _Float16 *B8l, int k0, int K, int ldb,
                             float *Y, hipStream_t st) {
  if (ldb % W_TJ != 0) throw Error(HSMD_ERR_ARG, "2 * n_omega must be a multiple of 512");
  if (k0 % WK != 0) throw Error(HSMD_ERR_ARG, "trajectory chunks start at multiples of 32 points");
  const int nm = (((maxM + QM - 1) / QM) + 7) / 8 * 8;
  dim3 grid(nm * (ldb / W_TJ), 1, S);
  hipLaunchKernelGGL(time_gemm_mfma_w_kernel, grid, dim3(W_THREADS), 0, st,
                     (const GemmSample *)samples_dev, A, B8h, B8l, k0, K, ldb, Y);
}

// the pipelined 128 x 512 tile, same operands and preconditions as above
void launch_time_gemm_mfma_p(const void *samples_dev, int S, int maxM, const float *A,
                             const _Float16 *B8h, const _Float16 *B8l, int k0, int K, int ldb,
                             float *Y, hipStream_t st) {
  if (ldb % W_TJ != 0) throw Error(HSMD_ERR_ARG, "2 * n_omega must be a multiple of 512");
  if (k0 % WK != 0) throw Error(HSMD_ERR_ARG, "trajectory chunks start at multiples of 32 points");
  if ((long long)((k0 + K + WK - 1) / 8 + 4) * ldb * 16 >= (1LL << 31))
    throw Error(HSMD_ERR_ARG, "the twiddle table exceeds 2 GiB");
  const int nm = (((maxM + QM - 1) / QM) + 7) / 8 * 8;
  dim3 grid(nm * (ldb / W_TJ), 1, S);
  hipLaunchKernelGGL(time_gemm_mfma_p_kernel, grid, dim3(P_THREADS), 0, st,
                     (const GemmSample *)samples_dev, A, B8h, B8l, k0, K, ldb, Y);
}

// tj: 128 or 256 columns per workgroup tile; db: the double-buffered 128 tile
void launch_time_gemm_mfma_wide(const void *samples_dev, int S, int maxM, const float *A,
                                const _Float16 *Bth, const _Float16 *Btl, int ldt, int k0, int K,
                                int ldb, float *Y, hipStream_t st, int tj, bool db) {
  if (tj != 128 && tj != 256) throw Error(HSMD_ERR_ARG, "contraction tile must be 128 or 256 columns");
  if (ldb % tj != 0) throw Error(HSMD_ERR_ARG, "2 * n_omega must be a multiple of the tile width");
  const int nm = (((maxM + QM - 1) / QM) + 7) / 8 * 8;
  dim3 grid(nm * (ldb / tj), 1, S);
  if (db && tj == 128)
    hipLaunchKernelGGL(time_gemm_mfma_db_kernel, grid, dim3(QT), 0, st,
                       (const GemmSample *)samples_dev, A, Bth, Btl, ldt, k0, K, ldb, Y);
  else if (tj == 128)
    hipLaunchKernelGGL(time_gemm_mfma_q_kernel, grid, dim3(QT), 0, st,
                       (const GemmSample *)samples_dev, A, Bth, Btl, ldt, k0, K, ldb, Y);
  else
    hipLaunchKernelGGL(time_gemm_mfma_q256_kernel, grid, dim3(QT), 0, st,
                       (const GemmSample *)samples_dev, A, Bth, Btl, ldt, k0, K, ldb, Y);
}

}  // namespace hsmd
