// DSSF time transform as a pruned FFT (FP32 trajectories).
//
// The SMD engine needs, per sample, the lowest n_omega bins of the T-point
// DFT of every windowed trajectory row,
//   Y[m][n] + i Y[m][nop + n] = sum_t w_t x_t[m] e^{-2 pi i n t / T},
// the same sums the matrix-core contraction forms as a GEMM with the
// twiddle table (and the reference forms with a per-q FFT over time,
// CF/subrouts.cpp:12062-12105; the transform is linear, so taking it per
// site before the spatial sum gives the same spectrum).  When T is a
// multiple of 512 and n_omega <= 512 the sums factor: with P = T / 512 and
// t = P t1 + t2,
//   Y[m][n] = sum_{t2 < P} e^{-2 pi i n t2 / T} F_{t2}[n],
//   F_{t2}[k] = sum_{t1 < 512} w_t x_t e^{-2 pi i k t1 / 512},
// i.e. P 512-point FFTs of the decimated rows and one complex multiply-add
// per bin and t2: about 150 operations per row and time point against 2
// n_omega (2048 at 512 bins) for the GEMM.  A trajectory chunk [k0, k0 + K)
// enters as the rows with the other time points zero; each chunk adds its
// part, like the contraction's chunks.
//
// One workgroup (512 threads, 80 KB of LDS: two per CU, so that one
// works while the other waits at a barrier) owns 16 consecutive rows of one
// sample.  The 16 real rows go in pairs into 8 complex series z = x_c + i
// x_{c+8}.  Per t2: the workgroup writes the windowed decimated rows (loaded
// one t2 ahead, 64 bytes per time point; the workgroups of the two halves of
// a 128-byte line run on one XCD, one after the other in dispatch order, and
// share it through L2) into LDS; a barrier; each wave runs the 512-point FFT
// of one series (Stockham radix-8, three passes, natural order out) waiting
// only for its own LDS writes between passes, and ends with bins fj + 64 s
// of its series in registers; it adds w_n Z[n] and zeta^t2 w_n Z[n] into
// registers (the second sum is what the mirror bin 512 - n needs to separate
// the pair, X_c = (Z[n] + conj Z[-n]) / 2, X_{c+8} = -i (Z[n] - conj Z[-n]) /
// 2; the lanes exchange it once, after the loop); a barrier.  Twiddles come
// from tables rounded from FP64 on the host; the bin twiddles of a t2 pass
// through LDS, the pass twiddles are laid out so that the lanes of a wave
// read consecutive entries.
//
// Measured on MI355X for one 8192-point chunk of the configuration-4 batch
// (4,866,048 rows, 512 bins): 55 ms against 210 ms for the 128 x 512
// matrix-core tile, within 6e-7 of the FP64 transform's peak (the tile:
// 2e-6); VALU issue 0.63 of the SIMDs' cycles, no LDS bank conflicts, the
// chunk read from HBM once.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "hsmd_internal.h"

namespace hsmd {

namespace {

constexpr int F_N = 512;          // FFT length
constexpr int F_MB = 16;          // rows per workgroup (8 complex series)
constexpr int F_THREADS = 512;
// series stride in LDS (complex values): 2 F_SS = 4 (mod 32), so that the
// input writes of one wave (8 series x 2 parts x 4 consecutive points) fall
// on 32 different banks
constexpr int F_SS = 514;

struct FftSample {
  int M;
  int lda;
  long long a_off;
  long long y_off;
};

// LDS position of point i of a series: i with its low 4 bits XORed with
// ((i >> 4) ^ 9 (i >> 6)) & 15.  Complex (8-byte) LDS accesses are served
// 16 lanes at a time over 16 bank pairs; under this map every access of the
// FFT -- the passes' reads (points fj + 64 r) and writes (8 fj + s,
// 64 (fj / 8) + fj % 8 + 8 s, fj + 64 s) and the bin reads -- puts the 16
// lanes of a group on 16 different bank pairs (the mirror reads, points
// 512 - n, on 15).  Each access's position is a per-thread invariant XOR a
// compile-time constant.
__device__ __forceinline__ int fswz(int i) { return i ^ (((i >> 4) ^ (9 * (i >> 6))) & 15); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// 8-point forward DFT in registers (three radix-2 decimation-in-frequency
// stages); the result of bin s is left in v[rev(s)], rev = 3-bit reversal
__device__ __forceinline__ void dft8(float2 *v) {
  const float r2 = 0.70710678118654752f;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float2 a = v[r], b = v[r + 4];
    v[r] = make_float2(a.x + b.x, a.y + b.y);
    const float2 d = make_float2(a.x - b.x, a.y - b.y);
    if (r == 0) v[4] = d;
    else if (r == 1) v[5] = make_float2(r2 * (d.x + d.y), r2 * (d.y - d.x));     // d e^{-i pi/4}
    else if (r == 2) v[6] = make_float2(d.y, -d.x);                             // d (-i)
    else v[7] = make_float2(r2 * (d.y - d.x), -r2 * (d.x + d.y));                // d e^{-3 i pi/4}
  }
#pragma unroll
  for (int h = 0; h < 8; h += 4)
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const float2 a = v[h + r], b = v[h + r + 2];
      v[h + r] = make_float2(a.x + b.x, a.y + b.y);
      const float2 d = make_float2(a.x - b.x, a.y - b.y);
      v[h + r + 2] = r == 0 ? d : make_float2(d.y, -d.x);
    }
#pragma unroll
  for (int h = 0; h < 8; h += 2) {
    const float2 a = v[h], b = v[h + 1];
    v[h] = make_float2(a.x + b.x, a.y + b.y);
    v[h + 1] = make_float2(a.x - b.x, a.y - b.y);
  }
}

// tw: the pass twiddles, [7][8] e^{-2 pi i 8 r k / 512} (pass 1) and [7][64]
// e^{-2 pi i r k / 512} (pass 2), r = 1..7, padded to 512 entries; then
// [P][512] w_n = e^{-2 pi i n t2 / T} and [P][512] zeta^t2 w_n.  win: [T] the
// window.
__global__ __launch_bounds__(F_THREADS) __attribute__((amdgpu_waves_per_eu(4, 4)))
void time_dft_fft_kernel(const FftSample *samples, const float *__restrict__ A,
                         const float *__restrict__ win, const float2 *__restrict__ tw, int T,
                         int k0, int K, int nop, int nkeep, float *Y) {
  __shared__ float2 Zs[2][8 * F_SS];
  __shared__ float2 Ws[7 * 8 + 7 * 64];
  __shared__ float Wn[F_N];   // the window at the decimated row's points
  __shared__ float2 Cs[2][F_N];   // this t2's bin twiddles w_n = e^{-2 pi i n t2 / T} and zeta^t2 w_n
  const FftSample sd = samples[blockIdx.z];
  const int M = sd.M, lda = sd.lda;
  // XCD-aware row blocks: workgroups are dispatched round-robin over the 8
  // XCDs, so ids b and b + 8 (same XCD, consecutive there) take the two row
  // blocks of one 32-row pair
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int m0 = (2 * ((slot >> 1) * 8 + xcd) + (slot & 1)) * F_MB;
  if (m0 >= M) return;
  float *Yb = Y + sd.y_off;
  const int tid = threadIdx.x;
  const int P = T / F_N;
  for (int k = tid; k < 7 * 8 + 7 * 64; k += F_THREADS) Ws[k] = tw[k];
  const float2 *ctw = tw + F_N;
  // the sample's chunk, [K][lda] FP32: byte offsets stay below 2^31 (the
  // engine sizes chunks so), and a read past its end returns zero
  const __amdgpu_buffer_rsrc_t arsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)(A + sd.a_off), (short)0, K * lda * 4, 0x00020000);
  constexpr int kOut = 0x7ffffff0;
  // loads: row li of the workgroup, points t1 = lr + 32 j of the decimated row
  const int li = tid & 15, lr = tid >> 4;
  const bool lrow = m0 + li < M;
  const int lcol4 = 4 * (m0 + li);
  // fswz(lr + 32 j) = 16 (lr / 16) + 32 j + (a_in ^ c_j), c_j = (2 j ^ 9 (j / 2)) & 15
  const int a_in = (lr & 15) ^ (lr >> 4);
  float *zin = (float *)&Zs[0][0] + 2 * (li & 7) * F_SS + (li >> 3) + 32 * (lr >> 4);
  // FFT: series fc (one per wave), butterfly fj; fswz(fj + 64 r) = 64 r +
  // (b_fj ^ d_r), fswz(8 fj + s) = b0 ^ s, fswz(64 (fj / 8) + fj % 8 + 8 s) =
  // b1 ^ f_s
  const int fc = tid >> 6, fj = tid & 63;
  const int b_fj = 16 * (fj >> 4) + ((fj & 15) ^ (fj >> 4));
  const int b0 = fswz(8 * fj);
  const int b1 = 64 * (fj >> 3) + (((fj & 7) ^ ((4 * (fj >> 3)) ^ (9 * (fj >> 3)))) & 15);
  // combine: after pass 2 the thread holds bins n = fj + 64 s of series fc,
  // Z[n] = X_c[n] + i X_{c+8}[n] with X_c[n] = (Z[n] + conj Z[-n]) / 2 and
  // X_{c+8}[n] = -i (Z[n] - conj Z[-n]) / 2.  It keeps
  //   acc1[s] = sum_t2 w_n Z[n],   accz[s] = sum_t2 zeta^t2 w_n Z[n],
  // zeta = e^{2 pi i / P}: since conj(w_{512 - n}) = zeta^t2 w_n, the sum
  // sum_t2 w_m conj(Z[-m]) that bin m = 512 - n needs is conj(accz[s]) of
  // this lane, exchanged once after the loop.
  float2 acc1[8], accz[8];
#pragma unroll
  for (int h = 0; h < 8; h++) acc1[h] = accz[h] = make_float2(0.f, 0.f);
  float xv[16], wn;
  float2 cwn, czn;
  // the window over the chunk's points only: a point outside reads zero
  const __amdgpu_buffer_rsrc_t wrsrc =
      __builtin_amdgcn_make_buffer_rsrc((void *)(win + k0), (short)0, K * 4, 0x00020000);
  // the next t2 (from `from` on) whose decimated row meets the chunk
  auto next_t2 = [&](int from) __attribute__((always_inline)) {
    for (int u = from; u < P; u++) {
      const int lo = k0 - u <= 0 ? 0 : (k0 - u + P - 1) / P;
      const int hi = k0 + K - u <= 0 ? 0 : min(F_N, (k0 + K - u + P - 1) / P);
      if (lo < hi) return u;
    }
    return P;
  };
  // loads complete in issue order for the wait counter: the window value
  // and the bin twiddle (stored after pass 2) go out before the 16
  // trajectory values (needed at the top of the next trip).  Every trip
  // issues all 18 loads, the last one with out-of-range trajectory offsets
  // (no memory traffic), so that the counter is the same on every path.
  auto fetch = [&](int t2, bool live) __attribute__((always_inline)) {
    const int tc = live ? t2 : P - 1;
    const int tw_ = P * tid + tc;   // the window of point t1 = tid
    wn = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                       wrsrc, live ? 4 * (tw_ - k0) : kOut, 0, 0));
    cwn = ctw[(size_t)tc * F_N + tid];
    czn = ctw[(size_t)(P + tc) * F_N + tid];
    __builtin_amdgcn_sched_barrier(0);   // keep the issue order
    const int lo = k0 - tc <= 0 ? 0 : (k0 - tc + P - 1) / P;
    const int hi = k0 + K - tc <= 0 ? 0 : min(F_N, (k0 + K - tc + P - 1) / P);
    if (!live || (lo == 0 && hi == F_N)) {
      // the chunk holds the whole decimated row: offsets in steps of 32 P rows
      const bool on = live && lrow;
      const int step = on ? 32 * P * 4 * lda : 0;
      const int base = on ? (P * lr + tc - k0) * (4 * lda) + lcol4 : kOut;
#pragma unroll
      for (int j = 0; j < 16; j++)
        xv[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(arsrc, base + j * step, 0, 0));
    } else {
#pragma unroll
      for (int j = 0; j < 16; j++) {
        const int tau = P * (lr + 32 * j) + tc - k0;
        const bool ok = lrow && tau >= 0 && tau < K;
        xv[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                              arsrc, ok ? tau * (4 * lda) + lcol4 : kOut, 0, 0));
      }
    }
  };
  int t2 = next_t2(0);
  if (t2 < P) {
    fetch(t2, true);
    Wn[tid] = wn;
  }
  __syncthreads();
  while (t2 < P) {
    // the positions derived from these bases are recomputed every trip (a
    // VALU XOR each) instead of being kept in registers across the loop
    int ai = a_in, bf = b_fj, c0 = b0, c1 = b1;
    asm volatile("" : "+v"(ai), "+v"(bf), "+v"(c0), "+v"(c1));
    // windowed decimated rows into buffer 0 (real / imaginary parts of the pairs)
#pragma unroll
    for (int j = 0; j < 16; j++)
      zin[2 * (32 * j + (ai ^ (((2 * j) ^ (9 * (j >> 1))) & 15)))] = xv[j] * Wn[lr + 32 * j];
    // this t2's bin twiddles (read in the combine, behind the barrier below;
    // the last trip's reads were behind the barrier that ended it)
    Cs[0][tid] = cwn;
    Cs[1][tid] = czn;
    const int t2n = next_t2(t2 + 1);
    __syncthreads();
    fetch(t2n, t2n < P);
    // three radix-8 Stockham passes over the wave's own series, 0 -> 1 -> 0
    // -> registers; between passes the wave waits for its own LDS writes
    float2 v[8];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const float2 *src = &Zs[q & 1][fc * F_SS];
      float2 *dst = &Zs[(q & 1) ^ 1][fc * F_SS];
      const int Ns = q == 0 ? 1 : (q == 1 ? 8 : 64);
      const int k = fj & (Ns - 1);
#pragma unroll
      for (int r = 0; r < 8; r++) v[r] = src[64 * r + (bf ^ (((4 * (r & 3)) ^ (9 * r)) & 15))];
      if (q > 0) {
        // v_r times e^{-2 pi i r k / (8 Ns)}
        const float2 *wt = q == 1 ? &Ws[k] : &Ws[7 * 8 + k];
        const int wst = q == 1 ? 8 : 64;
#pragma unroll
        for (int r = 1; r < 8; r++) v[r] = cmul(v[r], wt[(r - 1) * wst]);
      }
      dft8(v);
      if (q < 2) {
        // bin s of butterfly fj goes to point (fj - k) 8 + k + s Ns
        constexpr int rev[8] = {0, 4, 2, 6, 1, 5, 3, 7};
#pragma unroll
        for (int s = 0; s < 8; s++) {
          const int pos = q == 0 ? c0 ^ s : c1 ^ ((16 * (s >> 1)) | ((8 * (s & 1)) ^ (s >> 1)));
          dst[pos] = v[rev[s]];
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      }
    }
    // the next t2's window, read in the next trip behind the barrier that
    // ends this one (its last reads were before this trip's first barrier)
    Wn[tid] = wn;
    // bins n = fj + 64 s are v[rev(s)]
    constexpr int rev[8] = {0, 4, 2, 6, 1, 5, 3, 7};
#pragma unroll
    for (int s = 0; s < 8; s++) {
      const float2 w = Cs[0][fj + 64 * s], wz = Cs[1][fj + 64 * s];
      const float2 z = v[rev[s]];
      acc1[s].x += w.x * z.x - w.y * z.y;
      acc1[s].y += w.x * z.y + w.y * z.x;
      accz[s].x += wz.x * z.x - wz.y * z.y;
      accz[s].y += wz.x * z.y + wz.y * z.x;
    }
    t2 = t2n;
    // every wave's pass-2 reads of buffer 0 come before the next input writes
    __syncthreads();
  }
  // acc2[s] = sum_t2 w_n conj(Z[-n]) = conj(accz) of the bin -n: lane
  // 64 - fj, register 7 - s (lane 0: its own register (8 - s) % 8, and for
  // bin 0, whose weights are all 1, conj(acc1[0]))
  const int mlane4 = 4 * ((64 - fj) & 63);
  float2 acc2[8];
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const float2 mr = accz[7 - s];
    float2 m;
    m.x = __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(mlane4, __builtin_bit_cast(int, mr.x)));
    m.y = __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(mlane4, __builtin_bit_cast(int, mr.y)));
    if (fj == 0) m = s == 0 ? acc1[0] : accz[8 - s];
    acc2[s] = make_float2(m.x, -m.y);
  }
  // Y_c = (acc1 + acc2) / 2, Y_{c+8} = -i (acc1 - acc2) / 2
  const bool first = k0 == 0;
  const int ldb = 2 * nop;
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const int nn = fj + 64 * s;
    if (nn >= nop) continue;
    const bool keep = nn < nkeep;
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int m = m0 + fc + 8 * u;
      if (m >= M) continue;
      const float2 a = u == 0 ? make_float2(0.5f * (acc1[s].x + acc2[s].x), 0.5f * (acc1[s].y + acc2[s].y))
                              : make_float2(0.5f * (acc1[s].y - acc2[s].y), 0.5f * (acc2[s].x - acc1[s].x));
      float *y = Yb + (size_t)m * ldb;
      if (first) {
        y[nn] = keep ? a.x : 0.f;
        y[nop + nn] = keep ? a.y : 0.f;
      } else if (keep) {
        unsafeAtomicAdd(&y[nn], a.x);
        unsafeAtomicAdd(&y[nop + nn], a.y);
      }
    }
  }
}

}  // namespace

// The host tables of the FFT form, as (cos, -sin) pairs rounded from FP64:
// [7][8] e^{-2 pi i 8 r k / 512} and [7][64] e^{-2 pi i r k / 512} for
// r = 1..7 (the radix-8 passes' twiddles, 504 of 512 entries), then
// [T / 512][512] e^{-2 pi i n t2 / T}, then [T / 512][512]
// e^{-2 pi i (n - 512) t2 / T}.
std::vector<float> time_dft_fft_tables(int T) {
  const int P = T / F_N;
  std::vector<float> t((size_t)2 * F_N * (1 + 2 * P), 0.f);
  const double tp = 2.0 * acos(-1.0);
  auto put = [&](size_t e, long long ph, long long period) {
    t[2 * e] = (float)cos(tp * (double)ph / (double)period);
    t[2 * e + 1] = (float)-sin(tp * (double)ph / (double)period);
  };
  for (int r = 1; r < 8; r++)
    for (int k = 0; k < 8; k++) put((size_t)(r - 1) * 8 + k, (8LL * r * k) % F_N, F_N);
  for (int r = 1; r < 8; r++)
    for (int k = 0; k < 64; k++) put(56 + (size_t)(r - 1) * 64 + k, (long long)r * k, F_N);
  for (int u = 0; u < P; u++)
    for (int k = 0; k < F_N; k++) {
      put((size_t)F_N * (1 + u) + k, ((long long)k * u) % T, T);
      // zeta^u w_k = e^{-2 pi i (k - 512) u / T}
      put((size_t)F_N * (1 + P + u) + k, (((long long)(k - F_N) * u) % T + T) % T, T);
    }
  return t;
}

bool time_dft_fft_ok(int T, int nop) { return T >= F_N && T % F_N == 0 && nop <= F_N; }

// samples_dev: the engine's per-sample (M, lda, a_off, y_off) records;
// tables: time_dft_fft_tables(T) on the device; win: the window, FP32
void launch_time_dft_fft(const void *samples_dev, int S, int maxM, const float *A, const float *win,
                         const float *tables, int T, int k0, int K, int nop, int nkeep, float *Y,
                         hipStream_t st) {
  if (!time_dft_fft_ok(T, nop))
    throw Error(HSMD_ERR_ARG, "the FFT form needs n_steps a multiple of 512 and n_omega <= 512");
  if (k0 < 0 || K < 0 || k0 + K > T) throw Error(HSMD_ERR_ARG, "chunk outside the trajectory");
  if (K == 0) return;
  // row blocks in pairs over the 8 XCDs (see the kernel): the grid is a
  // multiple of 16 workgroups
  const int nb = (maxM + F_MB - 1) / F_MB;
  dim3 grid((nb + 15) / 16 * 16, 1, S);
  hipLaunchKernelGGL(time_dft_fft_kernel, grid, dim3(F_THREADS), 0, st, (const FftSample *)samples_dev,
                     A, win, (const float2 *)tables, T, k0, K, nop, nkeep, Y);
}

}  // namespace hsmd
