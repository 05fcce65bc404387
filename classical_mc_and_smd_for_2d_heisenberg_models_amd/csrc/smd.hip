// Semiclassical spin dynamics (SMD) engine: RK4 on the undamped LLG
// equations, the energy-correction step, and DSSF accumulation.
//
// Reference: get_dynSpectrum CF/subrouts.cpp:11547-12149, driven by the RK
// loop CF/smd_code.cpp:3079-3312 (one MPI rank per slice of samples) and the
// root reduction CF/smd_code.cpp:3330-3396.
//
// Per time step and sample (reference order):
//   K1..K4 = dt (h(S_k) x S_k) with S_1 = S, S_2 = S + K1/2, S_3 = S + K2/2,
//   S_4 = S + K3 (stages unnormalised), S <- norm(S + (K1+K4)/6 + (K2+K3)/3);
//   c = (E_ref - E(S)) / sum_i |h_i x S_i|^2, S_i <- norm(S_i + c S_i x (h_i x S_i));
//   qSS_t = (1/N) sum_i S_i(0).S_i(t);  S(q,t) = DFT2D(S)/sqrt(Nq).
// After the loop: X(q,w) = DFT_t(w(t) S(q,t))/sqrt(T), DSSF = |X|^2.
//
// Design: one workgroup per sample, all samples (of any mix of lattices) in
// one launch that integrates a chunk of time steps with the lattice resident
// in LDS; FP64 block reductions for E_ref, E(S), sum |h x S|^2 and qSS; the
// correction is applied in FP64 (Jacobi order: every site uses the fields of
// the post-RK configuration).  Each step's configuration is written to a
// trajectory chunk; the DSSF is linear in the spins, so the time transform is
// contracted first, Y(r,w) += sum_t A(t, r) B(t, w) with B = window x
// twiddles, as a GEMM per chunk, and the spatial transform of Y and |.|^2
// follow once at the end.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <memory>
#include <string>

#include "device_common.h"
#include "hsmd_internal.h"
#include "smd_kernel.h"

namespace hsmd {

// ------------------------------------------------------------- DSSF GEMM

// Y[s][m][j] += sum_k A[s][k][m] * B[k0+k][j], m < 3 N_s, j < 2 n_omega.
// 64 x 64 output tile per workgroup of 256 threads, 4 x 4 per thread.
template <typename T>
__global__ __launch_bounds__(256) void time_gemm_kernel(const SampleDev *samples, const T *A,
                                                        const T *Bm, int ldb, int k0, int K,
                                                        T *Y) {
  __shared__ T As[16][64];
  __shared__ T Bs[16][64];
  const SampleDev sd = samples[blockIdx.z];
  const int M = 3 * sd.N;
  const int m0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  if (m0 >= M) return;
  const T *Ab = A + sd.traj_off;
  const int lda = sd.ldm;
  T *Yb = Y + sd.y_off;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = T(0);
  for (int kk = 0; kk < K; kk += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int r = e >> 6, cidx = e & 63;
      const int k = kk + r;
      const int m = m0 + cidx;
      As[r][cidx] = (k < K && m < M) ? Ab[(size_t)k * lda + m] : T(0);
      Bs[r][cidx] = (k < K) ? Bm[(size_t)(k0 + k) * ldb + j0 + cidx] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) {
      T av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; i++) av[i] = As[r][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; j++) bv[j] = Bs[r][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bv[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int m = m0 + ty * 4 + i;
    if (m >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; j++) Yb[(size_t)m * ldb + j0 + tx * 4 + j] += acc[i][j];
  }
}

// Wave dependencies of the LDS-resident RK4 form: site p is integrated by
// thread p mod 1024, so by wave (p mod 1024) / 64 at any sites-per-thread
// count; wdep[w] = the waves owning a neighbour (any coupled or padded slot
// the kernel reads) of a site of wave w, without w.  The stages are ordered
// per wave when the masks are local (on average at most eight waves; lattices
// in the reference's site order: the waves either side), or as
// HSMD_SMD_WAVE_SYNC = 1 / 0 forces.
static void set_wave_deps(const Lattice &L, LatDev &d) {
  const int N = L.N;
  for (int w = 0; w < 16; w++) d.wdep[w] = 0u;
  bool used[16] = {false};
  for (int p = 0; p < N; p++) {
    const int w = (p & 1023) >> 6;
    used[w] = true;
    for (int k = 0; k < L.Z; k++) {
      const int q = L.g_nb[(size_t)k * N + p];
      if (q >= 0 && q < N) d.wdep[w] |= 1u << ((q & 1023) >> 6);
    }
  }
  size_t bits = 0, nw = 0;
  for (int w = 0; w < 16; w++) {
    d.wdep[w] &= ~(1u << w);
    if (used[w]) {
      bits += __builtin_popcount(d.wdep[w]);
      nw++;
    }
  }
  bool on = nw > 0 && bits <= 8 * nw;
  if (const char *ev = getenv("HSMD_SMD_WAVE_SYNC")) on = atoi(ev) != 0;
  d.wsync = on ? 1 : 0;
}

// --------------------------------------------------- spatial transform

// Row pass over the DFT grid of one sample: for grid row a and component d,
// R[a][b'][n] = sum_b Yg[a][b][n] e^{-2 pi i b b'/G}, Yg the site values
// scattered onto the grid (0 where no site).  One workgroup per (row, d, s).
template <typename T>
__global__ void dft_rows_kernel(const SampleDev *samples, const int *sample_ids,
                                const int32_t *gmap, int G, int n_omega, const T *Y,
                                const T *cs, const T *sn, T *Rout) {
  const int a = blockIdx.x, d = blockIdx.y, sl = blockIdx.z;
  const SampleDev sd = samples[sample_ids[sl]];
  const int N = sd.N;
  const int ldy = 2 * n_omega;
  const T *Yb = Y + sd.y_off;
  T *Rb = Rout + ((size_t)(sl * 3 + d) * G + a) * (size_t)G * ldy;
  for (int e = threadIdx.x; e < G * n_omega; e += blockDim.x) {
    const int bp = e / n_omega, n = e - bp * n_omega;
    T re = T(0), im = T(0);
    for (int b = 0; b < G; b++) {
      const int site = gmap[a * G + b];
      if (site < 0) continue;
      const T *y = Yb + (size_t)(d * N + site) * ldy;
      const T yr = y[n], yi = y[n_omega + n];
      const int ph = (b * bp) % G;
      const T c = cs[ph], s = sn[ph];   // e^{-i th} = c - i s
      re += yr * c + yi * s;
      im += yi * c - yr * s;
    }
    T *o = Rb + (size_t)bp * ldy;
    o[n] = re;
    o[n_omega + n] = im;
  }
}

// Column pass: X[a'][b'][n] = sum_a R[a][b'][n] e^{-2 pi i a a'/G}; the
// squared moduli, scaled by 1/(Nq T), are summed over the samples of the
// batch into dssf[d][q = a' G + b'][n].
template <typename T>
__global__ void dft_cols_kernel(int n_batch, int G, int n_omega, const T *R, const T *cs,
                                const T *sn, double scale, double *dssf) {
  const int bp = blockIdx.x, d = blockIdx.y;
  const int ldy = 2 * n_omega;
  const int Nq = G * G;
  for (int e = threadIdx.x; e < G * n_omega; e += blockDim.x) {
    const int ap = e / n_omega, n = e - ap * n_omega;
    double acc = 0.0;
    for (int sl = 0; sl < n_batch; sl++) {
      const T *Rb = R + (size_t)(sl * 3 + d) * G * (size_t)G * ldy;
      T re = T(0), im = T(0);
      for (int a = 0; a < G; a++) {
        const T *r = Rb + ((size_t)a * G + bp) * ldy;
        const T rr = r[n], ri = r[n_omega + n];
        const int ph = (a * ap) % G;
        const T c = cs[ph], s = sn[ph];
        re += rr * c + ri * s;
        im += ri * c - rr * s;
      }
      acc += (double)re * (double)re + (double)im * (double)im;
    }
    dssf[((size_t)d * Nq + (size_t)ap * G + bp) * n_omega + n] += scale * acc;
  }
}

// ---------------------------------------------------------------- engine

void launch_time_gemm_mfma(const void *samples_dev, int S, int maxM, const float *A, const float *B,
                           int ldb, int k0, int K, float *Y, hipStream_t st);

void launch_time_gemm_mfma_w(const void *samples_dev, int S, int maxM, const float *A,
                             const _Float16 *B8h, const _Float16 *B8l, int k0, int K, int ldb,
                             float *Y, hipStream_t st);
void launch_time_gemm_mfma_p(const void *samples_dev, int S, int maxM, const float *A,
                             const _Float16 *B8h, const _Float16 *B8l, int k0, int K, int ldb,
                             float *Y, hipStream_t st);
void launch_time_gemm_mfma_wide(const void *samples_dev, int S, int maxM, const float *A,
                                const _Float16 *Bth, const _Float16 *Btl, int ldt, int k0, int K,
                                int ldb, float *Y, hipStream_t st, int tj, bool db);
std::vector<float> time_dft_fft_tables(int T);
bool time_dft_fft_ok(int T, int nop);
void launch_time_dft_fft(const void *samples_dev, int S, int maxM, const float *A, const float *win,
                         const float *tables, int T, int k0, int K, int nop, int nkeep, float *Y,
                         hipStream_t st);

struct GemmSampleHost {
  int M;
  int lda;
  long long a_off;
  long long y_off;
};

struct SMDEngine {
  std::vector<Lattice> lats;
  hsmd_smd_spec spec{};
  int S = 0;
  bool f32 = false;
  bool use_lds = false;
  bool wide = false;
  int spt = 1;
  int maxZ = 0;
  int minZ = 1 << 30;
  bool general = false;
  size_t lds_bytes = 0;
  int lds_npad = 0;
  int Tc = 0;
  int t_next = 0;        // next time index to produce
  int chunk_base = 0;
  int chunk_fill = 0;
  bool dssf_done = false;
  std::vector<int32_t> sample_lat;
  std::vector<SampleDev> hs;
  std::vector<double> window;
  // device
  std::vector<DevBuf<int32_t>> nb, bi, bj, gmap;
  std::vector<DevBuf<uint32_t>> nb16;
  std::vector<DevBuf<uint32_t>> nb16c;   // compact labels (global-memory form)
  std::vector<DevBuf<uint8_t>> wcnt;     // compact labels: label words per 64-site chunk
  std::vector<DevBuf<uint8_t>> bc;
  std::vector<DevBuf<float>> Jf;
  std::vector<DevBuf<double>> Jd;
  DevBuf<LatDev> dlats;
  DevBuf<SampleDev> dsamples;
  DevBuf<unsigned char> spins, s0, stage, traj, ybuf, twid, dftcs;
  DevBuf<double> qss, eref, dssf, dd;
  DevBuf<uint32_t> sync_err;   // wave-ordering waits of the LDS form that ran out
  std::vector<char> lat_ws;     // per lattice: stages ordered per wave in the LDS form
  DevBuf<unsigned char> ks;
  DevBuf<int> sample_ids;
  DevBuf<GemmSampleHost> gemm_samples;
  DevBuf<_Float16> tw_hi, tw_lo;   // twiddles split into FP16 hi/lo, [2 nop][ldt]
  DevBuf<_Float16> tw8_hi, tw8_lo; // the same, chunk-major [ldt / 8][2 nop][8] (2 nop % 512 == 0)
  DevBuf<float> fft_tab, win32;    // FFT form: its twiddle tables and the FP32 window
  int nop = 0;                     // frequency bins held internally (n_omega padded)
  int ldt = 0;
  bool set = false;
  // time points appended by smd_feed instead of integrated, and the first
  // point fed (the qSS reference), in the ABI's site order
  bool fed = false;
  std::vector<double> feed0;
  // launch shape (hsmd_smd_set_shape): LDS-resident form possible, one slot
  // class for every workgroup, and the contraction tile (0: 128 x 128 with
  // pre-split twiddles, 1: 64 x 64, 2: 128 x 256, 3: 128 x 128 with two
  // LDS stages, 4: 128 x 512 with chunk-major twiddles, 5: 4 pipelined,
  // 6: the pruned FFT)
  bool lds_ok = false;
  bool one_class = false;
  int gemm = 0;
};

template <typename T>
static void upload_twiddles(SMDEngine &e) {
  // bins n >= n_omega of the padded table are zero columns
  const int T_ = e.spec.n_steps, no = e.nop, nkeep = e.spec.n_omega;
  std::vector<T> B((size_t)T_ * 2 * no, T(0));
  const double tp = 2.0 * acos(-1.0);
  for (int t = 0; t < T_; t++) {
    const double w = e.window[t];
    for (int n = 0; n < nkeep; n++) {
      const long long ph = ((long long)n * t) % T_;
      const double th = tp * (double)ph / (double)T_;
      B[(size_t)t * 2 * no + n] = (T)(w * cos(th));
      B[(size_t)t * 2 * no + no + n] = (T)(-w * sin(th));
    }
  }
  e.twid.alloc(B.size() * sizeof(T));
  HIP_CHECK(hipMemcpy(e.twid.p, B.data(), B.size() * sizeof(T), hipMemcpyHostToDevice));
  if (sizeof(T) == 4 && (2 * no) % 256 == 0) {
    // transposed FP16 split of the twiddles for the wide MFMA contraction;
    // rows padded to a multiple of 32 time points with zeros
    e.ldt = ((T_ + 31) / 32) * 32;
    std::vector<_Float16> hi((size_t)2 * no * e.ldt, (_Float16)0.f), lo(hi.size(), (_Float16)0.f);
    for (int t = 0; t < T_; t++)
      for (int j = 0; j < 2 * no; j++) {
        const float v = (float)B[(size_t)t * 2 * no + j];
        const _Float16 h = (_Float16)v;
        hi[(size_t)j * e.ldt + t] = h;
        lo[(size_t)j * e.ldt + t] = (_Float16)(v - (float)h);
      }
    e.tw_hi.upload(hi.data(), hi.size());
    e.tw_lo.upload(lo.data(), lo.size());
    if ((2 * no) % 512 == 0) {
      // chunk-major copy: the eight time points of one column are 16 bytes,
      // consecutive columns of one 8-point chunk are consecutive
      std::vector<_Float16> h8(hi.size()), l8(lo.size());
      for (int j = 0; j < 2 * no; j++)
        for (int t = 0; t < e.ldt; t++) {
          const size_t dst = ((size_t)(t / 8) * 2 * no + j) * 8 + (t % 8);
          h8[dst] = hi[(size_t)j * e.ldt + t];
          l8[dst] = lo[(size_t)j * e.ldt + t];
        }
      e.tw8_hi.upload(h8.data(), h8.size());
      e.tw8_lo.upload(l8.data(), l8.size());
    }
  }
  if (sizeof(T) == 4 && time_dft_fft_ok(T_, no)) {
    const std::vector<float> tab = time_dft_fft_tables(T_);
    e.fft_tab.upload(tab.data(), tab.size());
    std::vector<float> w32(T_);
    for (int t = 0; t < T_; t++) w32[t] = (float)e.window[t];
    e.win32.upload(w32.data(), w32.size());
  }
}

// The SMD kernels do not use the MC colour order, which scatters the
// neighbours of consecutive sites.  Regular lattices keep the reference's
// site order: the neighbour gathers of consecutive sites (i +- 1, i +- L,
// ...) fall on consecutive LDS banks and cache lines.  Tabulated graphs
// (point sets, quasicrystal tables) come in whatever order their generator
// produced; they are put in Morton (Z-curve) order of their positions, so
// that neighbours in space are neighbours in memory.  Sums over sites and
// bonds are order-independent up to rounding; per-site results are not
// affected at all.  Returns L with its device tables re-expressed in the
// chosen order (perm: new label -> site, inv: site -> new label).
static uint32_t spread16(uint32_t v) {
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

static Lattice natural_order(const Lattice &L) {
  Lattice o = L;
  const int N = L.N, Z = L.Z;
  std::vector<int> ord(N);
  for (int i = 0; i < N; i++) ord[i] = i;
  if (L.tabulated && L.rvec.size() == (size_t)2 * N) {
    double lo[2] = {L.rvec[0], L.rvec[1]}, hi[2] = {L.rvec[0], L.rvec[1]};
    for (int i = 0; i < N; i++)
      for (int d = 0; d < 2; d++) {
        lo[d] = std::min(lo[d], L.rvec[2 * (size_t)i + d]);
        hi[d] = std::max(hi[d], L.rvec[2 * (size_t)i + d]);
      }
    std::vector<uint32_t> key(N);
    for (int i = 0; i < N; i++) {
      uint32_t c[2];
      for (int d = 0; d < 2; d++) {
        const double w = hi[d] > lo[d] ? (L.rvec[2 * (size_t)i + d] - lo[d]) / (hi[d] - lo[d]) : 0.0;
        c[d] = (uint32_t)std::min(65535.0, std::max(0.0, w * 65535.0));
      }
      key[i] = spread16(c[0]) | (spread16(c[1]) << 1);
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
  }
  for (int p = 0; p < N; p++) { o.perm[p] = ord[p]; o.inv[ord[p]] = p; }
  for (int s = 0; s < Z; s++)
    for (int p = 0; p < N; p++) {
      const int q = L.g_nb[(size_t)s * N + L.inv[ord[p]]];
      o.g_nb[(size_t)s * N + p] = q < N ? o.inv[L.perm[q]] : N;
    }
  if (!L.g_J.empty())
    for (int s = 0; s < Z; s++)
      for (int p = 0; p < N; p++) o.g_J[(size_t)s * N + p] = L.g_J[(size_t)s * N + L.inv[ord[p]]];
  o.g_bi.resize(L.bi.size());
  o.g_bj.resize(L.bj.size());
  for (size_t b = 0; b < L.bi.size(); b++) {
    o.g_bi[b] = o.inv[L.bi[b]];
    o.g_bj[b] = o.inv[L.bj[b]];
  }
  o.g_perm = o.perm;
  return o;
}

// max_chunk > 0 caps the FP32 trajectory chunk (the autotuner's scratch engine)
SMDEngine *smd_create(const Lattice *const *lats, int nl, const int32_t *sample_lattice, int S,
                      const hsmd_smd_spec &sp, const double *window, int max_chunk) {
  if (S < 1) throw Error(HSMD_ERR_ARG, "sample count must be >= 1");
  if (!(sp.dt > 0.0)) throw Error(HSMD_ERR_ARG, "dt must be positive");
  if (sp.n_steps < 2) throw Error(HSMD_ERR_ARG, "n_steps must be >= 2");
  if (sp.precision < 0 || sp.precision > 2) throw Error(HSMD_ERR_ARG, "unknown precision form");
  if (sp.gauss_seidel) throw Error(HSMD_ERR_ARG, "in-place (Gauss-Seidel) correction is CPU-only");
  if (sp.dssf && (sp.n_omega < 1 || sp.n_omega > sp.n_steps))
    throw Error(HSMD_ERR_ARG, "n_omega must be in [1, n_steps]");
  std::unique_ptr<SMDEngine> e(new SMDEngine());
  e->spec = sp;
  e->S = S;
  e->f32 = sp.precision == HSMD_BATCHED_F32;
  const size_t es = e->f32 ? sizeof(float) : sizeof(double);
  for (int i = 0; i < nl; i++) e->lats.push_back(natural_order(*lats[i]));
  e->sample_lat.assign(sample_lattice, sample_lattice + S);
  int maxN = 0;
  for (int s = 0; s < S; s++) {
    if (sample_lattice[s] < 0 || sample_lattice[s] >= nl) throw Error(HSMD_ERR_ARG, "bad sample lattice index");
    maxN = std::max(maxN, e->lats[sample_lattice[s]].N);
  }
  e->window.assign(sp.n_steps, 1.0);
  if (window) e->window.assign(window, window + sp.n_steps);
  // per-lattice device tables
  e->nb.resize(nl); e->bi.resize(nl); e->bj.resize(nl); e->bc.resize(nl); e->nb16.resize(nl);
  e->wcnt.resize(nl); e->nb16c.resize(nl);
  e->Jf.resize(nl); e->Jd.resize(nl); e->gmap.resize(nl);
  std::vector<LatDev> hl(nl);
  for (int i = 0; i < nl; i++) {
    const Lattice &L = e->lats[i];
    if (L.Z > kZMAX) throw Error(HSMD_ERR_ARG, "too many neighbour slots");
    e->maxZ = std::max(e->maxZ, L.Z);
    e->minZ = std::min(e->minZ, L.Z);
    e->general = e->general || L.general;
    e->nb[i].upload(L.g_nb.data(), L.g_nb.size());
    e->bi[i].upload(L.g_bi.data(), L.g_bi.size());
    e->bj[i].upload(L.g_bj.data(), L.g_bj.size());
    e->bc[i].upload(L.g_bc.data(), L.g_bc.size());
    LatDev &d = hl[i];
    Coup &cp = d.cp;
    cp.N = L.N; cp.Npad = L.N + 1; cp.Z = L.Z; cp.general = L.general; cp.h_on = L.h_on; cp.h = L.c.h;
    for (int s = 0; s < kZMAX; s++) { cp.jxy[s] = cp.jz[s] = 0; cp.lam[s] = 1; }
    for (int s = 0; s < L.Z; s++) {
      const int sh = L.slot_shell[s];
      cp.jxy[s] = L.shellJ[sh - 1];
      cp.jz[s] = L.shellJ[sh - 1] * L.shellLam[sh - 1];
      cp.lam[s] = L.shellLam[sh - 1];
    }
    finish_coup(cp);
    d.bd.Nb = (int)L.g_bi.size();
    d.bd.n_class = (int)L.bc_Jxy.size();
    for (int k = 0; k < kBCMAX; k++) d.bd.jxy[k] = d.bd.jz[k] = 0;
    for (int k = 0; k < d.bd.n_class; k++) { d.bd.jxy[k] = L.bc_Jxy[k]; d.bd.jz[k] = L.bc_Jz[k]; }
    d.Jt = nullptr;
    if (L.general) {
      if (e->f32) {
        std::vector<float> t(L.g_J.begin(), L.g_J.end());
        e->Jf[i].upload(t.data(), t.size());
        d.Jt = e->Jf[i].p;
      } else {
        e->Jd[i].upload(L.g_J.data(), L.g_J.size());
        d.Jt = e->Jd[i].p;
      }
    }
    d.nb = e->nb[i].p; d.bi = e->bi[i].p; d.bj = e->bj[i].p; d.bc = e->bc[i].p;
    // 16-bit neighbour labels, two slots per word, kept in registers by the kernel
    if (L.N + 1 >= 65535) throw Error(HSMD_ERR_ARG, "SMD lattices are limited to 65534 sites");
    {
      // padded to kZMAX / 2 words per site with the sentinel, so a kernel of
      // any slot class reads every word of its class without a bounds test
      const int ZH = kZMAX / 2;
      std::vector<uint32_t> pk((size_t)ZH * L.N, (uint32_t)L.N | ((uint32_t)L.N << 16));
      for (int k = 0; 2 * k < L.Z; k++)
        for (int p = 0; p < L.N; p++) {
          const uint32_t lo = (uint32_t)L.g_nb[(size_t)(2 * k) * L.N + p];
          const uint32_t hi = (2 * k + 1 < L.Z) ? (uint32_t)L.g_nb[(size_t)(2 * k + 1) * L.N + p] : (uint32_t)L.N;
          pk[(size_t)k * L.N + p] = lo | (hi << 16);
        }
      // sparse tables (irregular graphs): compact labels, so the field loop
      // of a wave stops after the label words its 64 sites use
      size_t filled = 0;
      for (size_t k = 0; k < (size_t)L.Z * L.N; k++) filled += L.g_nb[k] < L.N ? 1 : 0;
      d.compact = 0;
      d.nb16c = nullptr;
      d.wcnt = nullptr;
      for (int k = 0; k < 3; k++) {
        d.shj[k] = L.shellJ[k];
        d.shjz[k] = L.shellJ[k] * L.shellLam[k];
      }
      e->nb16[i].upload(pk.data(), pk.size());
      d.nb16 = e->nb16[i].p;
      if (!L.general && L.N + 1 <= 16384 && 2 * filled < (size_t)L.Z * L.N) {
        std::fill(pk.begin(), pk.end(), (uint32_t)L.N | ((uint32_t)L.N << 16));
        std::vector<uint8_t> wc((L.N + 63) / 64, 0);
        for (int p = 0; p < L.N; p++) {
          int n = 0;
          for (int sl = 0; sl < L.Z; sl++) {
            const int q = L.g_nb[(size_t)sl * L.N + p];
            if (q >= L.N) continue;
            const uint32_t ent = (uint32_t)q | ((uint32_t)(L.slot_shell[sl] - 1) << 14);
            uint32_t &w = pk[(size_t)(n / 2) * L.N + p];
            w = (n % 2) ? ((w & 0xffffu) | (ent << 16)) : ((w & 0xffff0000u) | ent);
            n++;
          }
          wc[p / 64] = std::max<uint8_t>(wc[p / 64], (uint8_t)((n + 1) / 2));
        }
        e->wcnt[i].upload(wc.data(), wc.size());
        e->nb16c[i].upload(pk.data(), pk.size());
        d.wcnt = e->wcnt[i].p;
        d.nb16c = e->nb16c[i].p;
        d.compact = 1;
      }
    }
    d.field_energy = L.field_energy_ok() ? 1 : 0;
    set_wave_deps(L, d);
    e->lat_ws.push_back((char)(d.wsync && d.field_energy));
    // DFT grid in new labels; vacant sites are masked out of the transform
    if (!L.gridMap.empty()) {
      std::vector<int32_t> g(L.gridMap.size());
      for (size_t k = 0; k < g.size(); k++) {
        const int site = L.gridMap[k];
        g[k] = (site < 0 || L.vacant[site]) ? -1 : L.inv[site];
      }
      e->gmap[i].upload(g.data(), g.size());
    } else if (sp.dssf && L.rvec.size() != (size_t)2 * L.N) {
      throw Error(HSMD_ERR_ARG, "DSSF on a tabulated graph needs its site positions");
    }
  }
  e->dlats.upload(hl.data(), hl.size());
  // sizes and layout
  const int T_ = sp.n_steps;
  // trajectory chunk: the first contraction of a run stores Y and every
  // later one adds to it (a read and a write of all of Y, at 512 bins as
  // large as a 1024-point chunk), and each contraction launch pays its tile
  // prologues and epilogues once, so the FP32 chunk holds up to 8192 points
  // (the 512-sample batch of 1.6M sites: 160 GB, one contraction of 210 ms
  // per 8192-step run).  The chunk shrinks to the next lower power of two
  // (down to 256 points) while it, Y and the spin buffers would not fit in
  // 90 % of the free device memory; HSMD_SMD_MAX_CHUNK caps it (a power of
  // two >= 256).
  e->nop = e->f32 ? ((sp.n_omega + 127) / 128) * 128 : ((sp.n_omega + 31) / 32) * 32;
  const int no = e->nop;
  int cap = e->f32 ? 8192 : 256;
  if (const char *mc = getenv("HSMD_SMD_MAX_CHUNK")) {
    const int v = atoi(mc);
    if (v < 256 || (v & (v - 1)) != 0)
      throw Error(HSMD_ERR_ARG, "HSMD_SMD_MAX_CHUNK must be a power of two >= 256");
    cap = std::min(cap, v);
  }
  if (max_chunk > 0) cap = std::min(cap, max_chunk);
  e->Tc = std::min(T_, cap);
  if (sp.dssf) {
    size_t freeb = 0, totb = 0;
    HIP_CHECK(hipMemGetInfo(&freeb, &totb));
    long long rows = 0, ysz = 0, sp3 = 0;
    for (int s = 0; s < S; s++) {
      const int N = e->lats[sample_lattice[s]].N;
      rows += ((3LL * N + 3) / 4) * 4;
      ysz += 3LL * N * 2 * no;
      sp3 += 3LL * (N + 1);
    }
    // spins, t = 0 spins, stages and the wide form's increments: at most 11 copies
    const auto need = [&](int tc) { return (double)es * ((double)tc * rows + ysz + 11.0 * sp3); };
    // next lower power of two: every chunk but the first starts at a multiple of 32 points
    const auto lower = [](int tc) { int p = 256; while (2 * p < tc) p *= 2; return p; };
    while (e->Tc > 256 && need(e->Tc) > 0.9 * (double)freeb) e->Tc = lower(e->Tc);
    if (need(e->Tc) > 0.9 * (double)freeb)
      throw Error(HSMD_ERR_DEVICE, "the DSSF buffers of this batch exceed the free device memory");
    // the contraction addresses a sample's trajectory chunk with 32-bit byte
    // offsets (buffer loads): the chunk shrinks until its Tc x ldm FP32 values
    // stay below 2^31 B for every sample
    if (e->f32) {
      long long ldmax = 0;
      for (int s = 0; s < S; s++)
        ldmax = std::max(ldmax, ((3LL * e->lats[sample_lattice[s]].N + 3) / 4) * 4);
      while (e->Tc > 256 && (long long)e->Tc * ldmax * 4 >= (1LL << 31)) e->Tc = lower(e->Tc);
    }
  }
  if (e->f32 && sp.dssf)
    for (int s = 0; s < S; s++) {
      const long long ldm = ((3LL * e->lats[sample_lattice[s]].N + 3) / 4) * 4;
      if ((long long)e->Tc * ldm * 4 >= (1LL << 31))
        throw Error(HSMD_ERR_ARG, "a sample's trajectory chunk exceeds 2 GiB");
    }
  long long so = 0, to = 0, yo = 0;
  e->hs.resize(S);
  for (int s = 0; s < S; s++) {
    const int N = e->lats[sample_lattice[s]].N;
    const int ldm = ((3 * N + 3) / 4) * 4;
    e->hs[s] = {sample_lattice[s], N, so, to, yo, ldm};
    so += 3LL * (N + 1);
    to += (long long)e->Tc * ldm;
    yo += 3LL * N * 2 * no;
  }
  e->dsamples.upload(e->hs.data(), S);
  {
    std::vector<GemmSampleHost> gs(S);
    for (int s = 0; s < S; s++)
      gs[s] = {3 * e->hs[s].N, e->hs[s].ldm, e->hs[s].traj_off, e->hs[s].y_off};
    e->gemm_samples.upload(gs.data(), S);
  }
  e->spins.alloc(so * es);
  e->spins.zero();
  e->s0.alloc(so * es);
  e->qss.alloc((size_t)S * T_);
  e->qss.zero();
  e->eref.alloc(S);
  e->sync_err.alloc(1);
  e->sync_err.zero();
  if (sp.dssf) {
    e->traj.alloc(to * es);
    e->traj.zero();   // the pad columns of each row stay zero
    e->ybuf.alloc(yo * es);
    e->ybuf.zero();
    if (e->f32) upload_twiddles<float>(*e); else upload_twiddles<double>(*e);
  }
  // launch shape
  e->spt = 1;
  while (e->spt * kSmdThreads < maxN) e->spt *= 2;
  e->lds_npad = e->spt * kSmdThreads + 64;   // the kernels' compile-time component stride
  e->lds_bytes = (size_t)3 * 3 * e->lds_npad * es;   // S, B1, B2
  e->use_lds = (sp.precision != HSMD_PLAIN_F64) && e->lds_bytes <= 150 * 1024 &&
               e->spt * (int)es <= 16;
  e->lds_ok = e->use_lds;
  e->gemm = e->tw8_hi.n > 0 ? 4 : (e->f32 && e->ldt > 0) ? 0 : 1;
  // the FFT form costs one whole transform per chunk: the default while the
  // run needs at most two chunks (a third of the 128 x 512 tile's time for
  // one 8192-point chunk of the configuration-4 batch)
  if (e->fft_tab.n > 0 && (T_ + e->Tc - 1) / e->Tc <= 2) e->gemm = 6;
  // register-resident global form only for the plain form's small lattices
  e->wide = !e->use_lds && (sp.precision != HSMD_PLAIN_F64 || e->spt > (e->f32 ? 8 : 4));
  if (!e->use_lds) {
    e->stage.alloc(2 * so * es);
    e->stage.zero();
  }
  if (e->wide) {
    e->ks.alloc(6 * so * es);
    e->dd.alloc(3 * so);
  }
  return e.release();
}

void smd_set_samples(SMDEngine &e, const double *spins) {
  std::vector<double> h64;
  std::vector<float> h32;
  const size_t total = e.spins.n / (e.f32 ? sizeof(float) : sizeof(double));
  if (e.f32) h32.assign(total, 0.f); else h64.assign(total, 0.0);
  size_t in = 0;
  for (int s = 0; s < e.S; s++) {
    const Lattice &L = e.lats[e.hs[s].li];
    const int N = L.N, Npad = N + 1;
    const size_t off = (size_t)e.hs[s].spin_off;
    for (int i = 0; i < N; i++)
      for (int d = 0; d < 3; d++) {
        const double v = spins[in + (size_t)i * 3 + d];
        if (e.f32) h32[off + (size_t)d * Npad + L.inv[i]] = (float)v;
        else h64[off + (size_t)d * Npad + L.inv[i]] = v;
      }
    in += (size_t)N * 3;
  }
  if (e.f32) HIP_CHECK(hipMemcpy(e.spins.p, h32.data(), total * 4, hipMemcpyHostToDevice));
  else HIP_CHECK(hipMemcpy(e.spins.p, h64.data(), total * 8, hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(e.s0.p, e.spins.p, e.spins.n, hipMemcpyDeviceToDevice));
  e.t_next = 0;
  e.chunk_base = 0;
  e.chunk_fill = 0;
  e.fed = false;
  e.dssf_done = false;
  e.qss.zero();
  if (e.spec.dssf) e.ybuf.zero();
  if (e.f32)
    hipLaunchKernelGGL(smd_energy_kernel<float>, dim3(e.S), dim3(kSmdThreads), 0, 0, e.dlats.p,
                       e.dsamples.p, (const float *)e.spins.p, e.eref.p);
  else
    hipLaunchKernelGGL(smd_energy_kernel<double>, dim3(e.S), dim3(kSmdThreads), 0, 0, e.dlats.p,
                       e.dsamples.p, (const double *)e.spins.p, e.eref.p);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  e.set = true;
}

// Sample s <- the configuration at temperature slot slots[s] of an MC
// engine, copied on the device (MC labels -> SMD labels, MC precision -> SMD
// precision).  map[p] = MC label of the site at SMD label p.
template <typename TD, typename TS>
__global__ void handoff_kernel(const TS *src, int comps, int N, const int32_t *cfg_of_slot,
                               const int32_t *slots, const int32_t *map, const SampleDev *samples,
                               TD *dst) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (p >= N) return;
  const int c = cfg_of_slot[slots[s]];
  const int q = map[p];
  const TS *g = src + (size_t)c * comps * (N + 1);
  TD *o = dst + samples[s].spin_off;
  for (int d = 0; d < 3; d++) o[(size_t)d * (N + 1) + p] = d < comps ? (TD)g[(size_t)d * (N + 1) + q] : TD(0);
}

void smd_set_samples_from_mc(SMDEngine &e, const McHandoff &h, const int32_t *slots) {
  const Lattice &ML = *h.lat;
  const uint64_t sig = lattice_signature(ML);
  for (int s = 0; s < e.S; s++) {
    if (slots[s] < 0 || slots[s] >= h.R) throw Error(HSMD_ERR_ARG, "slot index out of range");
    const Lattice &L = e.lats[e.hs[s].li];
    if (L.N != ML.N || lattice_signature(L) != sig)
      throw Error(HSMD_ERR_ARG, "every sample must be on the MC engine's lattice");
  }
  const int N = ML.N;
  std::vector<int32_t> mcinv(N);
  for (int p = 0; p < N; p++) mcinv[ML.perm[p]] = p;
  // one label map per lattice of the batch (equal lattices given as
  // separate entries are handled separately)
  DevBuf<int32_t> dmap;
  std::vector<int> done(e.lats.size(), 0);
  for (int s = 0; s < e.S; s++) {
    const int li = e.hs[s].li;
    if (done[li]) continue;
    done[li] = 1;
    const Lattice &L = e.lats[li];
    std::vector<int32_t> map(N);
    for (int p = 0; p < N; p++) map[p] = mcinv[L.perm[p]];
    dmap.upload(map.data(), N);
    std::vector<int32_t> ids;
    for (int t = 0; t < e.S; t++)
      if (e.hs[t].li == li) ids.push_back(t);
    // one launch per lattice: the samples of this lattice, by sample index
    std::vector<int32_t> sl(ids.size());
    std::vector<SampleDev> sd(ids.size());
    for (size_t k = 0; k < ids.size(); k++) { sl[k] = slots[ids[k]]; sd[k] = e.hs[ids[k]]; }
    DevBuf<int32_t> dsl;
    DevBuf<SampleDev> dsd;
    dsl.upload(sl.data(), sl.size());
    dsd.upload(sd.data(), sd.size());
    const dim3 grid((N + 255) / 256, (unsigned)ids.size());
    const int comps = h.type == 0 ? 1 : 3;
#define HSMD_HANDOFF(TD, TS)                                                                     \
  hipLaunchKernelGGL((handoff_kernel<TD, TS>), grid, dim3(256), 0, 0, (const TS *)h.spins, comps, N, \
                     h.cfg_of_slot, dsl.p, dmap.p, dsd.p, (TD *)e.spins.p)
    if (e.f32) {
      if (h.type == 0) HSMD_HANDOFF(float, int8_t);
      else if (h.type == 1) HSMD_HANDOFF(float, float);
      else HSMD_HANDOFF(float, double);
    } else {
      if (h.type == 0) HSMD_HANDOFF(double, int8_t);
      else if (h.type == 1) HSMD_HANDOFF(double, float);
      else HSMD_HANDOFF(double, double);
    }
#undef HSMD_HANDOFF
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipDeviceSynchronize());
  }
  HIP_CHECK(hipMemcpy(e.s0.p, e.spins.p, e.spins.n, hipMemcpyDeviceToDevice));
  e.t_next = 0;
  e.chunk_base = 0;
  e.chunk_fill = 0;
  e.fed = false;
  e.dssf_done = false;
  e.qss.zero();
  if (e.spec.dssf) e.ybuf.zero();
  if (e.f32)
    hipLaunchKernelGGL(smd_energy_kernel<float>, dim3(e.S), dim3(kSmdThreads), 0, 0, e.dlats.p,
                       e.dsamples.p, (const float *)e.spins.p, e.eref.p);
  else
    hipLaunchKernelGGL(smd_energy_kernel<double>, dim3(e.S), dim3(kSmdThreads), 0, 0, e.dlats.p,
                       e.dsamples.p, (const double *)e.spins.p, e.eref.p);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  e.set = true;
}

void smd_get_samples(SMDEngine &e, double *spins) {
  const size_t total = e.spins.n / (e.f32 ? sizeof(float) : sizeof(double));
  std::vector<double> h64;
  std::vector<float> h32;
  if (e.f32) { h32.resize(total); HIP_CHECK(hipMemcpy(h32.data(), e.spins.p, total * 4, hipMemcpyDeviceToHost)); }
  else { h64.resize(total); HIP_CHECK(hipMemcpy(h64.data(), e.spins.p, total * 8, hipMemcpyDeviceToHost)); }
  size_t o = 0;
  for (int s = 0; s < e.S; s++) {
    const Lattice &L = e.lats[e.hs[s].li];
    const int N = L.N, Npad = N + 1;
    const size_t off = (size_t)e.hs[s].spin_off;
    for (int i = 0; i < N; i++)
      for (int d = 0; d < 3; d++)
        spins[o + (size_t)i * 3 + d] = e.f32 ? (double)h32[off + (size_t)d * Npad + L.inv[i]]
                                             : h64[off + (size_t)d * Npad + L.inv[i]];
    o += (size_t)N * 3;
  }
}

static void flush_chunk(SMDEngine &e) {
  if (!e.spec.dssf || e.chunk_fill == 0) return;
  const int no = e.nop;
  int maxM = 0;
  for (auto &s : e.hs) maxM = std::max(maxM, 3 * s.N);
  const int ldb = 2 * no;
  dim3 grid((maxM + 63) / 64, (ldb + 63) / 64, e.S);
  if (ldb % 64 != 0) throw Error(HSMD_ERR_ARG, "n_omega must be a multiple of 32");
  if (e.gemm == 6)
    launch_time_dft_fft(e.gemm_samples.p, e.S, maxM, (const float *)e.traj.p, e.win32.p, e.fft_tab.p,
                        e.spec.n_steps, e.chunk_base, e.chunk_fill, no, e.spec.n_omega,
                        (float *)e.ybuf.p, 0);
  else if (e.gemm == 5)
    launch_time_gemm_mfma_p(e.gemm_samples.p, e.S, maxM, (const float *)e.traj.p, e.tw8_hi.p,
                            e.tw8_lo.p, e.chunk_base, e.chunk_fill, ldb, (float *)e.ybuf.p, 0);
  else if (e.gemm == 4)
    launch_time_gemm_mfma_w(e.gemm_samples.p, e.S, maxM, (const float *)e.traj.p, e.tw8_hi.p,
                            e.tw8_lo.p, e.chunk_base, e.chunk_fill, ldb, (float *)e.ybuf.p, 0);
  else if (e.f32 && e.ldt > 0 && e.gemm != 1)
    launch_time_gemm_mfma_wide(e.gemm_samples.p, e.S, maxM, (const float *)e.traj.p, e.tw_hi.p,
                               e.tw_lo.p, e.ldt, e.chunk_base, e.chunk_fill, ldb,
                               (float *)e.ybuf.p, 0, e.gemm == 2 ? 256 : 128, e.gemm == 3);
  else if (e.f32)
    launch_time_gemm_mfma(e.gemm_samples.p, e.S, maxM, (const float *)e.traj.p,
                          (const float *)e.twid.p, ldb, e.chunk_base, e.chunk_fill,
                          (float *)e.ybuf.p, 0);
  else
    hipLaunchKernelGGL(time_gemm_kernel<double>, grid, dim3(256), 0, 0, e.dsamples.p,
                       (const double *)e.traj.p, (const double *)e.twid.p, ldb, e.chunk_base,
                       e.chunk_fill, (double *)e.ybuf.p);
  HIP_CHECK(hipGetLastError());
  e.chunk_base += e.chunk_fill;
  e.chunk_fill = 0;
}

void smd_step(SMDEngine &e, int n) {
  if (!e.set) throw Error(HSMD_ERR_STATE, "hsmd_smd_step before hsmd_smd_set_samples");
  if (e.fed) throw Error(HSMD_ERR_STATE, "the trajectory holds fed time points");
  if (n < 0) throw Error(HSMD_ERR_ARG, "n_steps must be >= 0");
  // the first call also records t = 0
  int points = n + (e.t_next == 0 ? 1 : 0);
  if (e.t_next + points > e.spec.n_steps)
    throw Error(HSMD_ERR_ARG, "more steps than n_steps - 1");
  while (points > 0) {
    const int room = e.Tc - e.chunk_fill;
    const int np = std::min(points, room);
    StepArgs a{};
    a.lats = e.dlats.p;
    a.samples = e.dsamples.p;
    a.spins = e.spins.p;
    a.s0 = e.s0.p;
    a.traj = e.spec.dssf ? (void *)e.traj.p : nullptr;
    a.qss = e.qss.p;
    a.eref = e.eref.p;
    a.dt = e.spec.dt;
    a.t0 = e.t_next;
    a.n_points = np;
    a.chunk_base = e.chunk_base;
    a.T = e.spec.n_steps;
    a.lds_npad = e.lds_npad;
    a.stage = e.stage.p;
    a.stage_stride = (long long)(e.stage.n / (2 * (e.f32 ? sizeof(float) : sizeof(double))));
    a.ks = e.ks.p;
    a.dd = e.dd.p;
    a.sync_err = e.sync_err.p;
    const SmdLaunch lc = {e.S, e.spt, e.maxZ, e.minZ, e.general, e.use_lds, e.wide,
                          e.spec.precision == HSMD_PLAIN_F64, e.lds_bytes, e.one_class};
    if (e.f32) smd_launch_f32(lc, a); else smd_launch_f64(lc, a);
    HIP_CHECK(hipGetLastError());
    e.t_next += np;
    e.chunk_fill += np;
    points -= np;
    if (e.chunk_fill == e.Tc) flush_chunk(e);
  }
  HIP_CHECK(hipDeviceSynchronize());
  uint32_t f = 0;
  HIP_CHECK(hipMemcpy(&f, e.sync_err.p, sizeof f, hipMemcpyDeviceToHost));
  if (f) throw Error(HSMD_ERR_DEVICE, "RK4 kernel: a wave-ordering wait ran out");
}

// Appends n time points of given configurations to every sample's
// trajectory instead of integrating them: spins holds n blocks laid out as
// hsmd_smd_set_samples.  The points go through the same trajectory chunks
// and contractions as integrated ones (so the DSSF of an externally
// computed trajectory, or the contraction's own error on a known one, can be
// taken); qSS is formed on the host against the first point fed.
void smd_feed(SMDEngine &e, int n, const double *spins) {
  if (!e.spec.dssf) throw Error(HSMD_ERR_STATE, "feeding time points needs the DSSF on");
  if (n < 0) throw Error(HSMD_ERR_ARG, "n_points must be >= 0");
  if (e.t_next > 0 && !e.fed)
    throw Error(HSMD_ERR_STATE, "time points were already integrated by hsmd_smd_step");
  if (e.t_next + n > e.spec.n_steps) throw Error(HSMD_ERR_ARG, "more points than n_steps");
  size_t tot = 0;
  for (auto &d : e.hs) tot += (size_t)d.N;
  const int T_ = e.spec.n_steps;
  std::vector<double> q((size_t)e.S * T_, 0.0);
  if (e.t_next == 0) {
    e.feed0.assign(spins, spins + 3 * tot);
    e.qss.zero();
  } else {
    HIP_CHECK(hipMemcpy(q.data(), e.qss.p, q.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  const size_t es = e.f32 ? sizeof(float) : sizeof(double);
  int done = 0;
  while (done < n) {
    const int np = std::min(n - done, e.Tc - e.chunk_fill);
    for (int s = 0; s < e.S; s++) {
      const SampleDev &sd = e.hs[s];
      const Lattice &L = e.lats[sd.li];
      const int N = sd.N;
      size_t in0 = 0;
      for (int r = 0; r < s; r++) in0 += (size_t)3 * e.hs[r].N;
      std::vector<unsigned char> rows((size_t)np * sd.ldm * es, 0);
      for (int k = 0; k < np; k++) {
        const double *src = spins + (size_t)(done + k) * 3 * tot + in0;
        double acc = 0.0;
        for (int i = 0; i < N; i++) {
          const int p = L.inv[i];
          for (int d = 0; d < 3; d++) {
            const double v = src[3 * (size_t)i + d];
            const size_t at = (size_t)k * sd.ldm + (size_t)d * N + p;
            if (e.f32) ((float *)rows.data())[at] = (float)v;
            else ((double *)rows.data())[at] = v;
            acc += e.feed0[in0 + 3 * (size_t)i + d] * v;
          }
        }
        q[(size_t)s * T_ + e.t_next + k] = acc / (double)N;
      }
      unsigned char *dst = e.traj.p + ((size_t)sd.traj_off + (size_t)e.chunk_fill * sd.ldm) * es;
      HIP_CHECK(hipMemcpy(dst, rows.data(), rows.size(), hipMemcpyHostToDevice));
    }
    e.t_next += np;
    e.chunk_fill += np;
    done += np;
    if (e.chunk_fill == e.Tc) flush_chunk(e);
  }
  HIP_CHECK(hipMemcpy(e.qss.p, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipDeviceSynchronize());
  e.fed = true;
}

int dssf_direct_groups(int n_samples, int nq, int nop);
template <typename T>
void dssf_direct_launch(const SampleDev *samples, const int *ids, int n_samples, int N, int nq,
                        int nop, const double *q, const double *pos, const uint8_t *mask, T *P,
                        const T *Y, double scale, int groups, double *part, double *dssf);

// DSSF at wavevectors q (nq x 2) by direct spatial sums over the sites of
// lattice li (vacant sites excluded): out[3][nq][n_omega], sample-summed
// |X|^2 / (N T).
void qpath_phase_launch(int nq, int QP, const double *q, int N, const double *pos,
                        const uint8_t *mask, _Float16 *Bh, _Float16 *Bl);
void qpath_mfma_launch(int nq, int QP, int N, const _Float16 *Bh, const _Float16 *Bl,
                       const void *samples_dev, int G, int nop, const float *Y, float *C,
                       double scale, double *D);

// FP32 q-path DSSF on the matrix cores: wavevectors in chunks of 256 (the
// phase table's 2 x 256 rows are the contraction tile's 512 columns), samples
// in groups whose GEMM outputs fit in 4 GB.  Same sums as the direct kernel:
// X_s(q, w) = sum_r e^{-i q.r} Y_s(r, w), D = sum_s |X_s|^2 / (N T).
static void dssf_qpath_mfma(SMDEngine &e, int li, int nq, const double *q, double *out) {
  const Lattice &L = e.lats[li];
  if (L.rvec.size() != (size_t)2 * L.N) throw Error(HSMD_ERR_ARG, "lattice has no site positions");
  const int N = L.N, no = e.nop, nkeep = e.spec.n_omega;
  constexpr int QP = 256;
  std::vector<int> ids;
  for (int s = 0; s < e.S; s++) if (e.sample_lat[s] == li) ids.push_back(s);
  std::vector<double> pos((size_t)2 * N);
  std::vector<uint8_t> mask(N);
  for (int p = 0; p < N; p++) {
    const int i = L.perm[p];
    pos[2 * p] = L.rvec[2 * (size_t)i];
    pos[2 * p + 1] = L.rvec[2 * (size_t)i + 1];
    mask[p] = L.vacant.empty() ? 0 : (uint8_t)(L.vacant[i] != 0);
  }
  DevBuf<double> dpos, dq, d;
  DevBuf<uint8_t> dmask;
  dpos.upload(pos.data(), pos.size());
  dmask.upload(mask.data(), mask.size());
  const int ldk = ((N + 31) / 32) * 32;
  DevBuf<_Float16> bh, bl;
  bh.alloc((size_t)ldk * 2 * QP);
  bl.alloc((size_t)ldk * 2 * QP);
  const size_t blk = (size_t)2 * no * 2 * QP;   // one (sample, component) output block
  const int nids = (int)ids.size();
  const int G = std::max(1, std::min(std::max(1, nids), (int)((1ull << 32) / (3 * blk * sizeof(float)))));
  DevBuf<float> C;
  C.alloc((size_t)3 * G * blk);
  DevBuf<GemmSampleHost> gsd;
  gsd.alloc((size_t)3 * G);
  d.alloc((size_t)3 * QP * no);
  const double scale = 1.0 / ((double)N * (double)e.spec.n_steps);
  const long long ldy = 2LL * no;
  std::vector<double> h((size_t)3 * QP * no);
  std::vector<GemmSampleHost> gs((size_t)3 * G);
  for (int q0 = 0; q0 < nq; q0 += QP) {
    const int nc = std::min(QP, nq - q0);
    d.zero();
    dq.upload(q + 2 * (size_t)q0, (size_t)2 * nc);
    qpath_phase_launch(nc, QP, dq.p, N, dpos.p, dmask.p, bh.p, bl.p);
    for (int g0 = 0; g0 < nids; g0 += G) {
      const int ng = std::min(G, nids - g0);
      for (int g = 0; g < ng; g++)
        for (int c = 0; c < 3; c++)
          gs[3 * g + c] = {2 * no, 2 * no, e.hs[ids[g0 + g]].y_off + (long long)c * N * ldy,
                           (long long)(3 * g + c) * (long long)blk};
      HIP_CHECK(hipMemcpy(gsd.p, gs.data(), (size_t)3 * ng * sizeof(GemmSampleHost),
                          hipMemcpyHostToDevice));
      qpath_mfma_launch(nc, QP, N, bh.p, bl.p, gsd.p, ng, no, (const float *)e.ybuf.p, C.p, scale,
                        d.p);
    }
    HIP_CHECK(hipMemcpy(h.data(), d.p, (size_t)3 * nc * no * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; c++)
      for (int k = 0; k < nc; k++)
        for (int n = 0; n < nkeep; n++)
          out[((size_t)c * nq + q0 + k) * nkeep + n] = h[((size_t)c * nc + k) * no + n];
  }
}

template <typename T>
static void dssf_direct(SMDEngine &e, int li, int nq, const double *q, double *out) {
  const Lattice &L = e.lats[li];
  if (L.rvec.size() != (size_t)2 * L.N) throw Error(HSMD_ERR_ARG, "lattice has no site positions");
  const int N = L.N, no = e.nop, nkeep = e.spec.n_omega;
  std::vector<int> ids;
  for (int s = 0; s < e.S; s++) if (e.sample_lat[s] == li) ids.push_back(s);
  std::vector<double> pos((size_t)2 * N);
  std::vector<uint8_t> mask(N);
  for (int p = 0; p < N; p++) {
    const int i = L.perm[p];
    pos[2 * p] = L.rvec[2 * (size_t)i];
    pos[2 * p + 1] = L.rvec[2 * (size_t)i + 1];
    mask[p] = L.vacant.empty() ? 0 : (uint8_t)(L.vacant[i] != 0);
  }
  DevBuf<double> dpos, dq, d;
  DevBuf<uint8_t> dmask;
  DevBuf<int> dids;
  dpos.upload(pos.data(), pos.size());
  dmask.upload(mask.data(), mask.size());
  if (!ids.empty()) dids.upload(ids.data(), ids.size());
  // the phase table is built per chunk of wavevectors, at most 2^28 entries
  const int qc = std::max(1, std::min(nq, (int)((1ll << 28) / std::max(1, N))));
  DevBuf<T> P;
  P.alloc((size_t)qc * N * 2);
  d.alloc((size_t)3 * qc * no);
  DevBuf<double> part;   // per-group partial sums, sized per chunk
  const double scale = 1.0 / ((double)N * (double)e.spec.n_steps);
  std::vector<double> h((size_t)3 * qc * no);
  for (int q0 = 0; q0 < nq; q0 += qc) {
    const int nc = std::min(qc, nq - q0);
    d.zero();
    if (!ids.empty()) {
      dq.upload(q + 2 * (size_t)q0, (size_t)2 * nc);
      const int groups = dssf_direct_groups((int)ids.size(), nc, no);
      if (part.n < (size_t)groups * 3 * nc * no) part.alloc((size_t)groups * 3 * nc * no);
      dssf_direct_launch<T>(e.dsamples.p, dids.p, (int)ids.size(), N, nc, no, dq.p, dpos.p,
                            dmask.p, P.p, (const T *)e.ybuf.p, scale, groups, part.p, d.p);
    }
    HIP_CHECK(hipMemcpy(h.data(), d.p, (size_t)3 * nc * no * sizeof(double), hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; c++)
      for (int k = 0; k < nc; k++)
        for (int n = 0; n < nkeep; n++)
          out[((size_t)c * nq + q0 + k) * nkeep + n] = h[((size_t)c * nc + k) * no + n];
  }
}

// The reference's wavevector grid of tabulated graphs: q = dq (nx, ny),
// nx, ny in [0, L/2) then [-L/2, 0), at index ny_index * L + nx_index
// (get_qct_SqData, CF/subrouts.cpp:8460-8525); dq = 4 (2 pi) / Gsz.
static std::vector<double> tabulated_q_grid(const Lattice &L) {
  const int Lq = L.L, l2 = Lq / 2;
  const double dq = L.sq_dq();
  std::vector<double> q((size_t)2 * L.Nq, 0.0);
  for (int j = 0; j < Lq; j++)
    for (int i = 0; i < Lq; i++) {
      const int nx = i < l2 ? i : i - Lq, ny = j < l2 ? j : j - Lq;
      q[2 * ((size_t)j * Lq + i)] = dq * nx;
      q[2 * ((size_t)j * Lq + i) + 1] = dq * ny;
    }
  return q;
}

template <typename T>
static void dssf_space(SMDEngine &e, int li, double *out) {
  const Lattice &L = e.lats[li];
  const int G = L.Gsz, no = e.nop, nkeep = e.spec.n_omega, Nq = G * G;
  std::vector<int> ids;
  for (int s = 0; s < e.S; s++) if (e.sample_lat[s] == li) ids.push_back(s);
  DevBuf<double> d;
  d.alloc((size_t)3 * Nq * no);
  d.zero();
  std::vector<T> cs(G), sn(G);
  const double tp = 2.0 * acos(-1.0);
  for (int k = 0; k < G; k++) { cs[k] = (T)cos(tp * k / G); sn[k] = (T)sin(tp * k / G); }
  DevBuf<T> dcs, dsn;
  dcs.upload(cs.data(), G);
  dsn.upload(sn.data(), G);
  const int batch = std::max(1, std::min((int)ids.size(), 16));
  DevBuf<T> R;
  R.alloc((size_t)batch * 3 * Nq * 2 * no);
  DevBuf<int> dids;
  const double scale = 1.0 / ((double)Nq * (double)e.spec.n_steps);
  for (size_t b0 = 0; b0 < ids.size(); b0 += batch) {
    const int nb = std::min(batch, (int)(ids.size() - b0));
    dids.upload(ids.data() + b0, nb);
    hipLaunchKernelGGL(dft_rows_kernel<T>, dim3(G, 3, nb), dim3(256), 0, 0, e.dsamples.p, dids.p,
                       e.gmap[li].p, G, no, (const T *)e.ybuf.p, dcs.p, dsn.p, R.p);
    hipLaunchKernelGGL(dft_cols_kernel<T>, dim3(G, 3), dim3(256), 0, 0, nb, G, no, R.p, dcs.p,
                       dsn.p, scale, d.p);
    HIP_CHECK(hipGetLastError());
  }
  std::vector<double> h((size_t)3 * Nq * no);
  HIP_CHECK(hipMemcpy(h.data(), d.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t dq = 0; dq < (size_t)3 * Nq; dq++)
    for (int n = 0; n < nkeep; n++) out[dq * nkeep + n] = h[dq * no + n];
}

void smd_dssf(SMDEngine &e, int li, double *out, double *qss) {
  if (li < 0 || li >= (int)e.lats.size()) throw Error(HSMD_ERR_ARG, "bad lattice index");
  if (!e.spec.dssf) throw Error(HSMD_ERR_STATE, "DSSF accumulation is off");
  if (e.t_next != e.spec.n_steps) throw Error(HSMD_ERR_STATE, "DSSF needs all n_steps time points");
  flush_chunk(e);
  const Lattice &L = e.lats[li];
  if (L.tabulated) {
    // get_dynSpectrum takes get_qct_SqData for tabulated graphs (CF/subrouts.cpp:11906-11908)
    const std::vector<double> q = tabulated_q_grid(L);
    if (e.f32) dssf_qpath_mfma(e, li, L.Nq, q.data(), out);
    else dssf_direct<double>(e, li, L.Nq, q.data(), out);
  } else {
    if (e.f32) dssf_space<float>(e, li, out); else dssf_space<double>(e, li, out);
  }
  if (qss) {
    const int T_ = e.spec.n_steps;
    std::vector<double> q((size_t)e.S * T_);
    HIP_CHECK(hipMemcpy(q.data(), e.qss.p, q.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int t = 0; t < T_; t++) qss[t] = 0.0;
    for (int s = 0; s < e.S; s++)
      if (e.sample_lat[s] == li)
        for (int t = 0; t < T_; t++) qss[t] += q[(size_t)s * T_ + t];
  }
  HIP_CHECK(hipDeviceSynchronize());
}

void smd_dssf_q(SMDEngine &e, int li, int nq, const double *q, double *out) {
  if (li < 0 || li >= (int)e.lats.size()) throw Error(HSMD_ERR_ARG, "bad lattice index");
  if (nq < 1) throw Error(HSMD_ERR_ARG, "nq must be >= 1");
  if (!e.spec.dssf) throw Error(HSMD_ERR_STATE, "DSSF accumulation is off");
  if (e.t_next != e.spec.n_steps) throw Error(HSMD_ERR_STATE, "DSSF needs all n_steps time points");
  flush_chunk(e);
  if (e.f32) dssf_qpath_mfma(e, li, nq, q, out); else dssf_direct<double>(e, li, nq, q, out);
  HIP_CHECK(hipDeviceSynchronize());
}

void smd_energies(SMDEngine &e, double *E) {
  HIP_CHECK(hipMemcpy(E, e.eref.p, e.S * sizeof(double), hipMemcpyDeviceToHost));
}

void smd_destroy(SMDEngine *e) { delete e; }

// True when the FP32 LDS batch would run per-workgroup slot classes.
static bool two_class(const SMDEngine &e) {
  if (!e.f32 || !e.lds_ok || e.general || e.spt > 4) return false;
  return (e.minZ <= 8 && e.maxZ > 8 && e.maxZ <= 12) || (e.minZ <= 4 && e.maxZ > 4 && e.maxZ <= 8);
}

// The RK4 kernel smd_step launches for the engine's current form, as
// "<kernel><template arguments> x 1024", with ", wave-ordered on k of n
// lattices" when the LDS-resident form orders its stages per wave.
std::string smd_step_kernel(const SMDEngine &e) {
  const char *T = e.f32 ? "float" : "double";
  const bool plain = e.spec.precision == HSMD_PLAIN_F64;
  const int Z = e.maxZ;
  const int zc = e.general ? 32 : Z <= 4 ? 4 : Z <= 8 ? 8 : Z <= 12 ? 12 : Z <= 20 ? 20 : 32;
  const int spt = e.spt <= 4 ? e.spt : 8;
  const char *gen = e.general ? "true" : "false";
  char b[200];
  if (e.wide) {
    snprintf(b, sizeof b, "smd_kernel_wide<%s,%d,%s> x 1024%s", T, zc, gen, plain ? ", one sample per launch" : "");
    return b;
  }
  if (!e.use_lds) {
    snprintf(b, sizeof b, "smd_kernel<%s,%d,false,%d,%s> x 1024, one sample per launch", T, spt, zc, gen);
    return b;
  }
  int lo = 0, hi = 0;
  if (e.f32 && spt <= 4 && !e.one_class && !e.general) {
    if (e.minZ <= 8 && Z > 8 && Z <= 10) { lo = e.minZ <= 4 ? 4 : 8; hi = 10; }
    else if (e.minZ <= 8 && Z > 8 && Z <= 12) { lo = e.minZ <= 4 ? 4 : 8; hi = 12; }
    else if (e.minZ <= 4 && Z > 4 && Z <= 8) { lo = 4; hi = 8; }
  }
  if (hi) snprintf(b, sizeof b, "smd_kernel_two<float,%d,%d,%d> x 1024", spt, lo, hi);
  else snprintf(b, sizeof b, "smd_kernel<%s,%d,true,%d,%s> x 1024", T, spt, zc, gen);
  std::string r = b;
  int k = 0;
  for (char c : e.lat_ws) k += c ? 1 : 0;
  if (k) r += ", wave-ordered on " + std::to_string(k) + " of " + std::to_string(e.lat_ws.size()) + " lattices";
  return r;
}

static int current_form(const SMDEngine &e) {
  if (e.use_lds) return (two_class(e) && !e.one_class) ? 2 : 1;
  return e.wide ? 3 : 0;
}

// Launch shape of the batched forms.  form: 0 keeps the current one;
// 1 = LDS-resident, every workgroup with the batch's largest slot class;
// 2 = LDS-resident, each workgroup with its own lattice's slot class (FP32
// batches mixing two classes); 3 = global-memory form (one workgroup per
// sample, stages in HBM / Infinity Cache).  gemm: 0 = 128 x 128 contraction
// tile with pre-split twiddles (FP32, 2 n_omega a multiple of 256), 1 = 64 x 64,
// 2 = 128 x 256 with pre-split twiddles, 3 = 128 x 128 with two LDS stages,
// 4 = 128 x 512 with chunk-major twiddles (2 n_omega a multiple of 512),
// 5 = the same tile pipelined (twiddles in registers, trajectory through a
// three-stage LDS ring; same sums as 4), 6 = the pruned FFT (FP32, n_steps a
// multiple of 512, n_omega <= 512; its cost per chunk is that of a whole
// run's transform once the chunk holds n_steps / 512 points).
// Per-site arithmetic is the same in every shape; FP64 reductions may group
// terms differently.
void smd_set_shape(SMDEngine &e, int form, int gemm) {
  if (e.spec.precision == HSMD_PLAIN_F64) throw Error(HSMD_ERR_ARG, "the plain form has one launch shape");
  if (form < 0 || form > 3) throw Error(HSMD_ERR_ARG, "form must be 0..3");
  if (gemm < 0 || gemm > 6) throw Error(HSMD_ERR_ARG, "gemm tile must be 0..6");
  if ((form == 1 || form == 2) && !e.lds_ok)
    throw Error(HSMD_ERR_ARG, "the LDS-resident form does not fit these lattices");
  if (form == 2 && !two_class(e)) throw Error(HSMD_ERR_ARG, "the batch has a single slot class");
  if (gemm == 6 && e.fft_tab.n == 0)
    throw Error(HSMD_ERR_ARG, "the FFT form needs FP32, n_steps a multiple of 512 and n_omega <= 512");
  if (gemm != 1 && gemm != 6 && !(e.f32 && e.ldt > 0))
    throw Error(HSMD_ERR_ARG, "the 128-row MFMA tiles need FP32 and 2 n_omega a multiple of 256");
  if ((gemm == 4 || gemm == 5) && e.tw8_hi.n == 0)
    throw Error(HSMD_ERR_ARG, "the 128 x 512 MFMA tile needs FP32 and 2 n_omega a multiple of 512");
  if (form == 1 || form == 2) {
    e.use_lds = true;
    e.wide = false;
    e.one_class = form == 1;
  } else if (form == 3) {
    const size_t es = e.f32 ? sizeof(float) : sizeof(double);
    const size_t so = e.spins.n / es;
    if (e.stage.n == 0) { e.stage.alloc(2 * so * es); e.stage.zero(); }
    if (e.ks.n == 0) e.ks.alloc(6 * so * es);
    if (e.dd.n == 0) e.dd.alloc(3 * so);
    e.use_lds = false;
    e.wide = true;
    e.one_class = false;
  }
  e.gemm = gemm;
}

// Autotuner of the batched SMD step: workgroup form (LDS-resident with one
// or two slot classes, global-memory) and the tile of the DSSF contraction.
// A scratch engine with the caller's lattices, batch and spec runs n_steps
// RK4 steps of random unit spins per form (no chunk flush inside) and one
// full-chunk contraction per tile, timed with HIP events; the fastest of
// each wins.  Cached like the MC tuner.
void smd_autotune(const Lattice *const *lats, int nl, const int32_t *sample_lattice, int S,
                  const hsmd_smd_spec &sp, const double *window, int n_steps, hsmd_smd_tuning *out) {
  if (n_steps < 1) throw Error(HSMD_ERR_ARG, "n_steps must be >= 1");
  if (sp.precision == HSMD_PLAIN_F64) throw Error(HSMD_ERR_ARG, "the plain form has one launch shape");
  uint64_t h = 0xcbf29ce484222325ull;
  for (int i = 0; i < nl; i++) {
    const uint64_t ls = lattice_signature(*lats[i]);
    h = fnv1a(&ls, sizeof ls, h);
  }
  h = fnv1a(sample_lattice, (size_t)S * sizeof(int32_t), h);
  char key[256];
  snprintf(key, sizeof key, "smd_%016llx_S%d_p%d_T%d_w%d_d%d_%s", (unsigned long long)h, S, sp.precision,
           sp.n_steps, sp.n_omega, sp.dssf, device_tag().c_str());
  std::vector<double> cv;
  if (tune_cache_get(key, cv) && cv.size() >= 6) {
    out->form = (int32_t)cv[0];
    out->gemm = (int32_t)cv[1];
    out->ms_per_step = cv[2];
    out->default_ms_per_step = cv[3];
    out->gemm_ms = cv[4];
    out->default_gemm_ms = cv[5];
    out->n_candidates = 0;
    out->cached = 1;
    return;
  }
  std::unique_ptr<SMDEngine> e(smd_create(lats, nl, sample_lattice, S, sp, window, 2048));
  std::vector<double> spins;
  {
    uint64_t st = 0x2545f4914f6cdd1dull;
    auto u01 = [&]() {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    size_t tot = 0;
    for (auto &d : e->hs) tot += (size_t)d.N;
    spins.resize(3 * tot);
    for (size_t i = 0; i < tot; i++) {
      const double ph = 2.0 * M_PI * u01(), ct = 1.0 - 2.0 * u01(), stt = sqrt(1.0 - ct * ct);
      spins[3 * i] = stt * cos(ph);
      spins[3 * i + 1] = stt * sin(ph);
      spins[3 * i + 2] = ct;
    }
  }
  const int def_form = current_form(*e), def_gemm = e->gemm;
  std::vector<int> forms;
  if (e->lds_ok) forms.push_back(1);
  if (two_class(*e)) forms.push_back(2);
  forms.push_back(3);
  const int k = std::max(1, std::min(n_steps, std::min(e->Tc, sp.n_steps) - 3));
  hipEvent_t ev0, ev1;
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  auto timed = [&](auto &&fn) {
    HIP_CHECK(hipDeviceSynchronize());
    HIP_CHECK(hipEventRecord(ev0, 0));
    fn();
    HIP_CHECK(hipEventRecord(ev1, 0));
    HIP_CHECK(hipEventSynchronize(ev1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    return (double)ms;
  };
  double best = 1e300, def_ms = -1.0, gbest = -1.0, gdef = -1.0;
  int best_form = def_form, best_gemm = def_gemm, ncand = 0;
  try {
    for (int f : forms) {
      smd_set_shape(*e, f, def_gemm);
      smd_set_samples(*e, spins.data());
      smd_step(*e, 1);
      const double per = timed([&] { smd_step(*e, k); }) / k;
      if (getenv("HSMD_TUNE_VERBOSE")) fprintf(stderr, "[hsmd-tune] smd form %d: %.4f ms/step\n", f, per);
      ncand++;
      if (f == def_form) def_ms = per;
      if (per < best) { best = per; best_form = f; }
    }
    if (sp.dssf && e->f32) {
      std::vector<int> tiles = {1};
      if (e->ldt > 0) tiles = {0, 3, 2, 1};
      if (e->tw8_hi.n > 0) tiles.insert(tiles.begin(), {5, 4});
      if (e->fft_tab.n > 0) tiles.insert(tiles.begin(), 6);
      gbest = 1e300;
      for (int g : tiles) {
        smd_set_shape(*e, best_form, g);
        // a contraction of up to 2048 points ranks the tiles (the cost per
        // point is flat beyond a few hundred)
        const int kt = std::min(std::min(e->Tc, sp.n_steps), 2048);
        e->chunk_base = 0;
        e->chunk_fill = kt;
        flush_chunk(*e);   // warm-up
        e->chunk_base = 0;
        e->chunk_fill = kt;
        double ms = timed([&] { flush_chunk(*e); });
        if (g == 6) {
          // the FFT form costs one whole transform per chunk (of at least
          // n_steps / 512 points): a run of n_steps points in chunks of up to
          // 8192, scaled to the kt points the other tiles are timed on
          const int P = sp.n_steps / 512;
          const double full = ms * (double)P / std::min(P, kt);
          ms = full * ((sp.n_steps + 8191) / 8192) * kt / sp.n_steps;
        }
        if (getenv("HSMD_TUNE_VERBOSE")) fprintf(stderr, "[hsmd-tune] smd contraction tile %d: %.3f ms\n", g, ms);
        ncand++;
        if (g == def_gemm) gdef = ms;
        if (ms < gbest) { gbest = ms; best_gemm = g; }
      }
    }
  } catch (...) {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    throw;
  }
  HIP_CHECK(hipEventDestroy(ev0));
  HIP_CHECK(hipEventDestroy(ev1));
  out->form = best_form;
  out->gemm = best_gemm;
  out->ms_per_step = best;
  out->default_ms_per_step = def_ms;
  out->gemm_ms = gbest;
  out->default_gemm_ms = gdef;
  out->n_candidates = ncand;
  out->cached = 0;
  tune_cache_put(key, {(double)best_form, (double)best_gemm, best, def_ms, gbest, gdef});
}

}  // namespace hsmd
