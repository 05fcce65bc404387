// cpu_ref -- CPU FP64 restatement of the reference's hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the GPU library is
// checked against; it is linked or executed only by tests/, by the smoke
// check in __graft_entry__.py and by the cpu_baseline leg of bench.py.  The
// product (libhsmd.so) never calls it.
//
// Each function restates one reference function (paths relative to the
// reference root, CF = Codes_Folder) from the source read as text; the
// reference itself cannot be compiled here (GSL, MKL and a working MPI
// compiler are absent), so there is no oracle/_ref build.
//
// Modes:
//  * reference order  -- random-permutation sequential sweeps and the
//    reference's replica-exchange rule; the statistical oracle and the CPU
//    baseline (one OpenMP thread per replica / sample, like one MPI rank each);
//  * colour replay    -- the GPU's colour order and counter-based stream,
//    for the bit-exact Ising check and the FP64 one-sweep Heisenberg check.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double PI = acos(-1.0);

// ----------------------------------------------------------------- RNG
// Philox4x32-10 (Salmon et al., SC'11), checked against the Random123
// known-answer vectors in tests/test_oracle.py.
struct W4 { uint32_t v[4]; };

W4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t key) {
  uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
  uint32_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
    const uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
    const uint32_t y1 = (uint32_t)p1;
    const uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
    const uint32_t y3 = (uint32_t)p0;
    x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  W4 w = {{x0, x1, x2, x3}};
  return w;
}

double u53(uint32_t hi, uint32_t lo) {
  return (double)((((uint64_t)hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0);
}

// exp(x) for x <= 0 from fma/rint/ldexp only (Cody-Waite split of ln 2 as
// in fdlibm, degree-13 Taylor polynomial); identical on every IEEE machine.
double dexp(double x) {
  if (!(x <= 0.0)) return 1.0;
  if (x < -700.0) return 0.0;
  const double k = rint(x * 1.44269504088896338700e+00);
  double r = fma(-k, 6.93147180369123816490e-01, x);
  r = fma(-k, 1.90821492927058770002e-10, r);
  static const double inv_fact[14] = {
      1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0, 1.0 / 5040.0,
      1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0, 1.0 / 39916800.0, 1.0 / 479001600.0,
      1.0 / 6227020800.0};
  double p = inv_fact[13];
  for (int n = 12; n >= 0; n--) p = fma(p, r, inv_fact[n]);
  return ldexp(p, (int)k);
}

// ------------------------------------------------------------- model
struct Model {
  int N = 0, zn1 = 0, zn2 = 0, znx = 0;
  std::vector<int32_t> nb;        // N x Zt, -1 = missing
  std::vector<int32_t> imp;       // N
  double J1 = 0, J2 = 0, JX = 0, J2d = 0, l1 = 1, l2 = 1, h = 0;
  double J1ic = 0, J2ic = 0, JXic = 0, J1ii = 0, J2ii = 0, JXii = 0;
  bool j2_on = false, jx_on = false, xxz_on = false, h_on = false, dis = false, j2ab = false;
  // bond list (make_bondList order)
  std::vector<int32_t> bi, bj, bs;
  int Zt() const { return zn1 + zn2 + znx; }
};

bool nz(double v) { return fabs(v) > 1.0e-12; }

// Flags of convert_inputs2data (CF/subrouts.cpp:1251-1400, 1900-2127).
void set_flags(Model &m, double imp_ratio, bool lieb) {
  m.dis = imp_ratio > 0.0;
  if (!m.dis) m.J1ic = m.J1ii = m.J2ic = m.J2ii = m.JXic = m.JXii = 0.0;
  m.j2_on = nz(m.J2);
  m.jx_on = nz(m.JX);
  m.j2ab = false;
  if (nz(m.J2d) && lieb) { m.j2_on = true; m.j2ab = true; }
  if (m.dis) {
    if (nz(m.J2ic) || nz(m.J2ii)) m.j2_on = true;
    if (nz(m.JXic) || nz(m.JXii)) m.jx_on = true;
  }
  m.xxz_on = (m.l1 != 1.0) || (m.l2 != 1.0);
  m.h_on = nz(m.h);
}

double jmat(const Model &m, int shell, int a1, int a2) {
  if (shell == 1) return a1 == a2 ? (a1 ? m.J1ii : m.J1) : m.J1ic;
  if (shell == 2) return a1 == a2 ? (a1 ? m.J2ii : m.J2) : m.J2ic;
  return a1 == a2 ? (a1 ? m.JXii : m.JX) : m.JXic;
}

// make_bondList (CF/subrouts.cpp:3426-3560): an unordered pair is kept once,
// shells visited 1, 2 (if on), X (if on), sites and slots in table order.
void make_bonds(Model &m) {
  m.bi.clear(); m.bj.clear(); m.bs.clear();
  std::vector<std::vector<int32_t>> seen(m.N);
  const int Zt = m.Zt();
  auto shell = [&](int s, int off, int zs) {
    for (int i = 0; i < m.N; i++)
      for (int k = 0; k < zs; k++) {
        const int nb = m.nb[(size_t)i * Zt + off + k];
        if (nb < 0) continue;
        const int a = std::min(i, nb), b = std::max(i, nb);
        if (std::find(seen[a].begin(), seen[a].end(), b) != seen[a].end()) continue;
        seen[a].push_back(b);
        m.bi.push_back(i); m.bj.push_back(nb); m.bs.push_back(s);
      }
  };
  shell(1, 0, m.zn1);
  if (m.j2_on) shell(2, m.zn1, m.zn2);
  if (m.jx_on) shell(3, m.zn1 + m.zn2, m.znx);
}

// get_localField (CF/subrouts.cpp:4465-4575).
void local_field(const Model &m, const double *S, int i, double out[3]) {
  const double lam1[3] = {1, 1, m.xxz_on ? m.l1 : 1.0};
  const double lam2[3] = {1, 1, m.xxz_on ? m.l2 : 1.0};
  const int Zt = m.Zt();
  out[0] = out[1] = out[2] = 0.0;
  if (m.h_on) out[2] = -m.h;
  for (int k = 0; k < m.zn1; k++) {
    const int nb = m.nb[(size_t)i * Zt + k];
    if (nb < 0) continue;
    const double J = m.dis ? jmat(m, 1, m.imp[i], m.imp[nb]) : m.J1;
    for (int n = 0; n < 3; n++) out[n] += (J * lam1[n]) * S[3 * nb + n];
  }
  if (m.j2_on) {
    for (int k = 0; k < m.zn2; k++) {
      const int nb = m.nb[(size_t)i * Zt + m.zn1 + k];
      if (nb < 0) continue;
      double J = m.dis ? jmat(m, 2, m.imp[i], m.imp[nb]) : m.J2;
      if (m.j2ab) J += ((i & 1) ? -1 : 1) * ((k & 1) ? -1 : 1) * m.J2d;
      for (int n = 0; n < 3; n++) out[n] += (J * lam2[n]) * S[3 * nb + n];
    }
  }
  if (m.jx_on) {
    for (int k = 0; k < m.znx; k++) {
      const int nb = m.nb[(size_t)i * Zt + m.zn1 + m.zn2 + k];
      if (nb < 0) continue;
      const double J = m.dis ? jmat(m, 3, m.imp[i], m.imp[nb]) : m.JX;
      for (int n = 0; n < 3; n++) out[n] += (J * lam2[n]) * S[3 * nb + n];
    }
  }
}

// get_localFieldX (CF/subrouts.cpp:4577-4668), spin in component 0.
double local_field_x(const Model &m, const double *S, int i) {
  const int Zt = m.Zt();
  double h = m.h_on ? -m.h : 0.0;
  for (int k = 0; k < m.zn1; k++) {
    const int nb = m.nb[(size_t)i * Zt + k];
    if (nb < 0) continue;
    h += (m.dis ? jmat(m, 1, m.imp[i], m.imp[nb]) : m.J1) * S[3 * nb];
  }
  if (m.j2_on)
    for (int k = 0; k < m.zn2; k++) {
      const int nb = m.nb[(size_t)i * Zt + m.zn1 + k];
      if (nb < 0) continue;
      double J = m.dis ? jmat(m, 2, m.imp[i], m.imp[nb]) : m.J2;
      if (m.j2ab) J += ((i & 1) ? -1 : 1) * ((k & 1) ? -1 : 1) * m.J2d;
      h += J * S[3 * nb];
    }
  if (m.jx_on)
    for (int k = 0; k < m.znx; k++) {
      const int nb = m.nb[(size_t)i * Zt + m.zn1 + m.zn2 + k];
      if (nb < 0) continue;
      h += (m.dis ? jmat(m, 3, m.imp[i], m.imp[nb]) : m.JX) * S[3 * nb];
    }
  return h;
}

double bond_J(const Model &m, int b, double &Jz) {
  const int i = m.bi[b], j = m.bj[b], s = m.bs[b];
  double J = m.dis ? jmat(m, s, m.imp[i], m.imp[j]) : (s == 1 ? m.J1 : s == 2 ? m.J2 : m.JX);
  if (s == 2 && m.j2ab) J += ((i & 1) ? -1 : 1) * ((j & 1) ? -1 : 1) * m.J2d;
  const double lz = m.xxz_on ? (s == 1 ? m.l1 : m.l2) : 1.0;
  Jz = J * lz;
  return J;
}

// get_energyValue (CF/subrouts.cpp:5181-5312): sequential sum over the bond
// list, then -h sum S_z.
double energy(const Model &m, const double *S) {
  double E = 0.0;
  for (size_t b = 0; b < m.bi.size(); b++) {
    double Jz;
    const double J = bond_J(m, (int)b, Jz);
    const int i = m.bi[b], j = m.bj[b];
    E += J * S[3 * i] * S[3 * j] + J * S[3 * i + 1] * S[3 * j + 1] + Jz * S[3 * i + 2] * S[3 * j + 2];
  }
  if (m.h_on)
    for (int i = 0; i < m.N; i++) E += (-m.h) * S[3 * i + 2];
  return E;
}

// Ising energy as the GPU defines it for exact replay: integer bond counts
// per coupling class (distinct (J, J lambda) pairs in order of first
// appearance in the bond list), E = sum_c fma(J_c, n_c, E), then fma(-h, M, E).
double energy_ising_classes(const Model &m, const double *S) {
  std::vector<double> cj, cz;
  std::vector<long long> cnt;
  for (size_t b = 0; b < m.bi.size(); b++) {
    double Jz;
    const double J = bond_J(m, (int)b, Jz);
    int c = -1;
    for (size_t k = 0; k < cj.size(); k++) if (cj[k] == J && cz[k] == Jz) { c = (int)k; break; }
    if (c < 0) { cj.push_back(J); cz.push_back(Jz); cnt.push_back(0); c = (int)cj.size() - 1; }
    cnt[c] += (long long)(S[3 * m.bi[b]] * S[3 * m.bj[b]]);
  }
  double E = 0.0;
  for (size_t k = 0; k < cj.size(); k++) E = fma(cj[k], (double)cnt[k], E);
  if (m.h_on) {
    long long M = 0;
    for (int i = 0; i < m.N; i++) M += (long long)S[3 * i];
    E = fma(-m.h, (double)M, E);
  }
  return E;
}

// reflect_aboutVec (CF/subrouts.cpp:4956-5004).
void reflect(const double s[3], const double h[3], double o[3]) {
  const double sq = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
  if (sq > 1.0e-12) {
    const double fac = (2.0 / sq) * (s[0] * h[0] + s[1] * h[1] + s[2] * h[2]);
    for (int n = 0; n < 3; n++) o[n] = fac * h[n] - s[n];
  } else {
    for (int n = 0; n < 3; n++) o[n] = s[n];
  }
}

// Heat-bath sample: normVec3d, get_rotZ2VMat (CF/subrouts.cpp:5095-5157),
// get_hBathSample (:5159-5179), MxVecProduct; inline block at
// CF/mc_hm_code.cpp:2919-2956.
void heat_bath(double beta, const double h[3], double u1, double u2, double o[3]) {
  const double H = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
  double z;
  if (nz(H)) {
    const double x0 = 1.0 + u2 * (exp(-2.0 * beta * H) - 1.0);
    z = (-1.0 / (beta * H)) * log(x0) - 1.0;
  } else {
    z = 2.0 * u2 - 1.0;
  }
  const double tht = 2.0 * PI * u1;
  const double w = sqrt(std::max(0.0, 1.0 - z * z));
  const double v[3] = {w * cos(tht), w * sin(tht), z};
  const double V[3] = {h[0] / H, h[1] / H, h[2] / H};
  double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double perp = V[0] * V[0] + V[1] * V[1];
  if (perp > 0) {
    const double sf = (1.0 - V[2]) / perp;
    const double x11 = 1.0 - sf * V[0] * V[0], x22 = 1.0 - sf * V[1] * V[1];
    const double x33 = x11 + x22 - 1.0, x12 = V[0] * V[1] * (-sf);
    const double w9[9] = {x11, x12, V[0], x12, x22, V[1], -V[0], -V[1], x33};
    memcpy(R, w9, sizeof(R));
  }
  for (int r = 0; r < 3; r++) o[r] = R[3 * r] * v[0] + R[3 * r + 1] * v[1] + R[3 * r + 2] * v[2];
}

// Replica exchange over slot pairs (CF/mc_hm_code.cpp:3015-3130): the
// receiver k2 accepts with min(1, exp((b2-b1)(E2-E1))) and, with the cap,
// only while (accepts+1)*100/attempts <= 50.  Configurations stay put; the
// slot -> configuration map is permuted.
void exchange(int R, int parity, bool cap, uint64_t seed, uint32_t sweep, const double *Ecfg,
              const double *beta, int32_t *cos, long long *cnt, long long *exc, int32_t *boost) {
  for (int k = 0; k < R; k++) boost[k] = 0;
  if (parity < 0) return;
  for (int k1 = parity; k1 + 1 < R; k1 += 2) {
    const int k2 = k1 + 1;
    cnt[k1]++; cnt[k2]++;
    const double ratio = (double)(exc[k2] + 1) * (100.0 / (double)cnt[k2]);
    const int c1 = cos[k1], c2 = cos[k2];
    const double fc = (beta[k2] - beta[k1]) * (Ecfg[c2] - Ecfg[c1]);
    const W4 r = philox((uint32_t)k1, sweep, 0u, 1u, seed);
    const double u = u53(r.v[0], r.v[1]);
    if (u < dexp(fc) && (!cap || ratio <= 50.0)) {
      cos[k1] = c2; cos[k2] = c1;
      exc[k1]++; exc[k2]++;
      boost[k1] = boost[k2] = 1;
    }
  }
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

// --- lattice generation: make_rvecList + make_nborsTable, reference scan ---
// Returns 0 on success, 1 unknown geometry, 2 neighbour-count mismatch,
// 3 DFT grid does not cover the sites.
int orc_lattice_dims(const char *geom, int L, int *N, int *zn1, int *zn2, int *znx, int *Gsz) {
  std::string g(geom);
  int ns, gf;
  if (g == "square") { ns = 1; gf = 1; *zn1 = 4; *zn2 = 4; *znx = 4; }
  else if (g == "triang") { ns = 1; gf = 1; *zn1 = 6; *zn2 = 6; *znx = 6; }
  else if (g == "hexagn") { ns = 2; gf = 3; *zn1 = 3; *zn2 = 6; *znx = 3; }
  else if (g == "kagome") { ns = 3; gf = 2; *zn1 = 4; *zn2 = 4; *znx = 6; }
  else if (g == "lieb") { ns = 2; gf = 2; *zn1 = 4; *zn2 = 4; *znx = 4; }
  else return 1;
  *N = ns * L * L;
  *Gsz = gf * L;
  return 0;
}

int orc_lattice_generate(const char *geom, int L, int pbc, int32_t *nbors, double *rvec,
                         int32_t *gridmap) {
  int N, zn1, zn2, znx, G;
  if (orc_lattice_dims(geom, L, &N, &zn1, &zn2, &znx, &G)) return 1;
  const std::string g(geom);
  const bool sq = (g == "square" || g == "lieb");
  const double sq2 = sqrt(2.0), sq3 = sqrt(3.0);
  const double c60 = cos(PI / 3.0), s60 = sin(PI / 3.0);
  const double szFac = g == "hexagn" ? sq3 : g == "kagome" ? 2.0 : g == "lieb" ? sq2 : 1.0;
  const int gridFac = G / L;
  // make_rvecList (CF/subrouts.cpp:2714-2905)
  int k = 0;
  for (int i = 0; i < L; i++)
    for (int j = 0; j < L; j++) {
      double r0, r1;
      if (sq) { r0 = i * szFac; r1 = j * szFac; }
      else { r0 = szFac * (i * 1.0 + j * c60); r1 = szFac * (i * 0.0 + j * s60); }
      if (g == "square" || g == "triang") { rvec[2 * k] = r0; rvec[2 * k + 1] = r1; k++; }
      else if (g == "hexagn") {
        rvec[2 * k] = r0 + 0.5 * sq3; rvec[2 * k + 1] = r1 + 0.5 * 1.0; k++;
        rvec[2 * k] = r0 + 1.0 * sq3; rvec[2 * k + 1] = r1 + 1.0 * 1.0; k++;
      } else if (g == "kagome") {
        for (int n = 0; n < 3; n++) {
          const double a = PI - n * (PI / 3.0);
          rvec[2 * k] = r0 + cos(a); rvec[2 * k + 1] = r1 + sin(a); k++;
        }
      } else {
        for (int n = 0; n < 2; n++) {
          const double a = PI - n * (PI / 2.0);
          rvec[2 * k] = r0 + 0.5 * szFac * cos(a); rvec[2 * k + 1] = r1 + 0.5 * szFac * sin(a); k++;
        }
      }
    }
  // DFT grid and index map (CF/subrouts.cpp:2907-2975), brute-force match
  if (gridmap) {
    const double fc = szFac / gridFac;
    const int i0 = (g == "lieb" || g == "kagome") ? 1 : 0;
    int cnt = 0;
    for (int i = 0; i < G; i++)
      for (int j = 0; j < G; j++) {
        double x, y;
        if (sq) { x = (i - i0) * fc; y = j * fc; }
        else { x = fc * ((i - i0) * 1.0 + j * c60); y = fc * ((i - i0) * 0.0 + j * s60); }
        const int n = j + i * G;
        gridmap[n] = -1;
        for (int s = 0; s < N; s++)
          if (fabs(x - rvec[2 * s]) < 1.0e-8 && fabs(y - rvec[2 * s + 1]) < 1.0e-8) {
            gridmap[n] = s; cnt++;
          }
      }
    if (cnt != N) return 3;
  }
  // make_nborsTable (CF/subrouts.cpp:2983-3316): integer coordinates, 8
  // periodic images appended per site, full scan, angle sort, rotations.
  const double spc = 500.0;
  struct P { int x, y, z; };
  std::vector<P> list;
  for (int s = 0; s < N; s++)
    list.push_back({(int)round(spc * rvec[2 * s]), (int)round(spc * rvec[2 * s + 1]), s});
  int nx1, nx2, ny;
  if (sq) { nx1 = nx2 = ny = (int)round(L * szFac * spc); }
  else {
    nx1 = (int)round(L * szFac * (spc * 1.0));
    nx2 = (int)round(L * szFac * (spc * c60));
    ny = (int)round(L * szFac * (spc * 0.0 + spc * s60)) - 1;
  }
  int sh[8][2] = {{nx1, 0}, {-nx1, 0}, {nx2, ny}, {nx2, -ny}, {-nx2, ny}, {-nx2, -ny}, {0, ny}, {0, -ny}};
  if (!sq) { sh[6][0] = nx2 + nx1; sh[6][1] = ny; sh[7][0] = -nx2 - nx1; sh[7][1] = -ny; }
  for (int s = 0; s < N; s++)
    for (int i = 0; i < 8; i++) list.push_back({sh[i][0] + list[s].x, sh[i][1] + list[s].y, s});
  const int rA1 = (int)floor(spc * 1.10), rA3 = (int)floor(2.0 * spc * 1.05);
  const int rA2 = sq ? (int)floor(sq2 * spc * 1.05) : (int)floor(sq3 * spc * 1.05);
  const int Zt = zn1 + zn2 + znx;
  for (int n = 0; n < N; n++) {
    const int nx = list[n].x, nyy = list[n].y;
    std::vector<std::pair<int, double>> a1, a2, a3;
    double angleSum = 0.0;
    for (size_t m = 0; m < list.size(); m++) {
      if (list[m].z == n) continue;
      const int dist = (int)floor(sqrt(pow(list[m].x - nx, 2) + pow(list[m].y - nyy, 2)));
      const double ang = atan2((double)(list[m].y - nyy), (double)(list[m].x - nx));
      int tag = list[m].z;
      if ((int)m >= N && !pbc) tag = -1;
      if (dist < rA1) { angleSum += ang; a1.push_back({tag, ang}); }
      else if (dist > rA1 && dist < rA2) a2.push_back({tag, ang});
      else if (dist > rA2 && dist < rA3) a3.push_back({tag, ang});
    }
    if ((int)a1.size() != zn1 || (int)a2.size() != zn2 || (int)a3.size() != znx) return 2;
    auto cmp = [](const std::pair<int, double> &x, const std::pair<int, double> &y) { return x.second > y.second; };
    std::sort(a1.begin(), a1.end(), cmp);
    std::sort(a2.begin(), a2.end(), cmp);
    std::sort(a3.begin(), a3.end(), cmp);
    if (g == "hexagn") {
      std::rotate(a2.begin(), a2.begin() + (n % 2 == 0 ? 1 : 2), a2.end());
    } else if (g == "kagome") {
      int ic1 = 1, ic2 = 1;
      const double avg = angleSum * (1.0 / zn1);
      if (avg > 0.0) { if (avg > 1.0) ic2 -= 1; else { ic1 -= 1; ic2 -= 1; } }
      std::rotate(a1.begin(), a1.begin() + ic1, a1.end());
      std::rotate(a2.begin(), a2.begin() + ic2, a2.end());
    }
    for (int q = 0; q < zn1; q++) nbors[(size_t)n * Zt + q] = a1[q].first;
    for (int q = 0; q < zn2; q++) nbors[(size_t)n * Zt + zn1 + q] = a2[q].first;
    for (int q = 0; q < znx; q++) nbors[(size_t)n * Zt + zn1 + zn2 + q] = a3[q].first;
  }
  return 0;
}

// --- model ---------------------------------------------------------------
// params[14]: J1 J2 JX J2delta lambda1 lambda2 h imp_ratio J1ic J2ic JXic J1ii J2ii JXii
void *orc_model_create(int N, int zn1, int zn2, int znx, const int32_t *nbors, const int32_t *imp,
                       const double *params, int lieb) {
  Model *m = new Model();
  m->N = N; m->zn1 = zn1; m->zn2 = zn2; m->znx = znx;
  m->nb.assign(nbors, nbors + (size_t)N * (zn1 + zn2 + znx));
  if (imp) m->imp.assign(imp, imp + N); else m->imp.assign(N, 0);
  m->J1 = params[0]; m->J2 = params[1]; m->JX = params[2]; m->J2d = params[3];
  m->l1 = params[4]; m->l2 = params[5]; m->h = params[6];
  m->J1ic = params[8]; m->J2ic = params[9]; m->JXic = params[10];
  m->J1ii = params[11]; m->J2ii = params[12]; m->JXii = params[13];
  set_flags(*m, params[7], lieb != 0);
  make_bonds(*m);
  return m;
}

void orc_model_destroy(void *m) { delete (Model *)m; }

int orc_bond_count(void *mp) { return (int)((Model *)mp)->bi.size(); }

void orc_bonds(void *mp, int32_t *pairs, int32_t *shell) {
  Model &m = *(Model *)mp;
  for (size_t b = 0; b < m.bi.size(); b++) {
    pairs[2 * b] = m.bi[b]; pairs[2 * b + 1] = m.bj[b]; shell[b] = m.bs[b];
  }
}

void orc_local_fields(void *mp, int R, const double *spins, double *out) {
  Model &m = *(Model *)mp;
  for (int r = 0; r < R; r++)
    for (int i = 0; i < m.N; i++)
      local_field(m, spins + (size_t)r * m.N * 3, i, out + ((size_t)r * m.N + i) * 3);
}

void orc_local_fields_x(void *mp, const double *spins, double *out) {
  Model &m = *(Model *)mp;
  for (int i = 0; i < m.N; i++) out[i] = local_field_x(m, spins, i);
}

void orc_energies(void *mp, int R, const double *spins, double *E) {
  Model &m = *(Model *)mp;
  for (int r = 0; r < R; r++) E[r] = energy(m, spins + (size_t)r * m.N * 3);
}

void orc_energies_ising(void *mp, int R, const double *spins, double *E) {
  Model &m = *(Model *)mp;
  for (int r = 0; r < R; r++) E[r] = energy_ising_classes(m, spins + (size_t)r * m.N * 3);
}

void orc_reflect(const double *s, const double *h, double *o) { reflect(s, h, o); }
void orc_heat_bath(double beta, const double *h, double u1, double u2, double *o) {
  heat_bath(beta, h, u1, u2, o);
}

void orc_philox(uint64_t seed, const uint32_t *ctr, uint32_t *out) {
  const W4 w = philox(ctr[0], ctr[1], ctr[2], ctr[3], seed);
  memcpy(out, w.v, 16);
}

double orc_det_exp(double x) { return dexp(x); }

// One replica-exchange round over the slot pairs of one parity (the rule of
// CF/mc_hm_code.cpp:3015-3130 as restated in exchange() above); Ecfg[c] is
// the energy of configuration c, cos/cnt/exc/boost are updated in place.
void orc_exchange(int R, int parity, int cap50, uint64_t seed, uint32_t sweep, const double *Ecfg,
                  const double *beta, int32_t *cos, long long *cnt, long long *exc, int32_t *boost) {
  exchange(R, parity, cap50 != 0, seed, sweep, Ecfg, beta, cos, cnt, exc, boost);
}

// --- colour replay of the GPU sweep --------------------------------------
// colour[N]: colour class of each site; the GPU visits class 0, 1, ... in
// turn, each class in parallel, which equals any sequential order inside a
// class.  spins: R configurations x N x 3 (configuration c = initial slot c).
// cos[R]: slot -> configuration map (in/out), cnt/exc[R]: exchange counters,
// boost[R] (in/out), sweep0: first sweep counter.  pt_parity: -1 none,
// 2 alternate by sweep counter, 0/1 fixed.  model: 0 Heisenberg, 1 Ising.
// etrace (nullable): n_sweeps x R slot energies after each sweep.
void orc_mc_replay(void *mp, int model, int R, const double *temps, const int32_t *colour,
                   int n_or, int n_hb, int n_or_b, int n_hb_b, int cap50, uint64_t seed,
                   uint32_t sweep0, int n_sweeps, int pt_parity, double *spins, int32_t *cos,
                   long long *cnt, long long *exc, int32_t *boost, double *etrace) {
  Model &m = *(Model *)mp;
  const int N = m.N;
  int nc = 0;
  for (int i = 0; i < N; i++) nc = std::max(nc, colour[i] + 1);
  std::vector<std::vector<int>> cls(nc);
  for (int i = 0; i < N; i++) cls[colour[i]].push_back(i);
  std::vector<double> beta(R), E(R);
  for (int k = 0; k < R; k++) beta[k] = 1.0 / temps[k];
  std::vector<int32_t> soc(R);
  for (int s = 0; s < n_sweeps; s++) {
    const uint32_t sweep = sweep0 + (uint32_t)s;
    for (int k = 0; k < R; k++) soc[cos[k]] = k;
    for (int c = 0; c < R; c++) {
      double *S = spins + (size_t)c * N * 3;
      const int slot = soc[c];
      const double b = beta[slot];
      if (model == 1) {
        for (int col = 0; col < nc; col++)
          for (int i : cls[col]) {
            const double h = local_field_x(m, S, i);
            const double dE = 2.0 * S[3 * i] * h;
            const W4 r = philox((uint32_t)i, sweep * 16u, (uint32_t)c, 0u, seed);
            if (u53(r.v[0], r.v[1]) < dexp(b * dE)) S[3 * i] = -S[3 * i];
          }
        E[c] = energy_ising_classes(m, S);
      } else {
        const int nor = boost[slot] ? n_or_b : n_or;
        const int nhb = boost[slot] ? n_hb_b : n_hb;
        for (int pass = 0; pass < nor + nhb; pass++) {
          const uint32_t step = sweep * 16u + (uint32_t)pass;
          for (int col = 0; col < nc; col++)
            for (int i : cls[col]) {
              double h[3], o[3];
              local_field(m, S, i, h);
              if (pass < nor) {
                reflect(S + 3 * i, h, o);
              } else {
                const W4 r = philox((uint32_t)i, step, (uint32_t)c, 0u, seed);
                heat_bath(b, h, u53(r.v[0], r.v[1]), u53(r.v[2], r.v[3]), o);
              }
              S[3 * i] = o[0]; S[3 * i + 1] = o[1]; S[3 * i + 2] = o[2];
            }
        }
        E[c] = energy(m, S);
      }
    }
    if (etrace)
      for (int k = 0; k < R; k++) etrace[(size_t)s * R + k] = E[cos[k]];
    const int par = pt_parity == 2 ? (int)(sweep % 2u) : pt_parity;
    exchange(R, par, cap50 != 0, seed, sweep, E.data(), beta.data(), cos, cnt, exc, boost);
  }
}

}  // extern "C"

namespace {

// --- heat-bath replay of the batched FP32 sweep ----------------------------
// The FP32 kernel's heat bath evaluated in FP64, for a check of that kernel's
// indexing and formulas at full size (the reference's own formula is
// heat_bath() above; both sample the same distribution).  Uniforms: 24-bit,
// u = (w >> 8) 2^-24 of Philox words; t = 1 + cos(theta) =
// -log1p(u2 expm1(-2 b H)) / (b H) clamped to [0, 2]; the orthonormal frame
// about h of Duff et al. (JCGT 2017).
void heat_bath_f32form(double beta, const double h[3], double u1, double u2, double o[3]) {
  const double H2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
  double t, ux = 0, uy = 0, uz = 1;
  bool frame = false;
  if (H2 > 1.0e-24) {
    const double H = sqrt(H2), bh = beta * H;
    t = -log1p(u2 * expm1(-2.0 * bh)) / bh;
    ux = h[0] / H; uy = h[1] / H; uz = h[2] / H;
    frame = true;
  } else {
    t = 2.0 * u2;
  }
  t = std::min(std::max(t, 0.0), 2.0);
  const double chi = t - 1.0, w = sqrt(t * (2.0 - t));
  const double s0 = w * cos(2.0 * PI * u1), s1 = w * sin(2.0 * PI * u1);
  if (frame) {
    const double sg = uz >= 0.0 ? 1.0 : -1.0;
    const double a = -1.0 / (sg + uz), b = ux * uy * a;
    const double e1[3] = {1.0 + sg * ux * ux * a, sg * b, -sg * ux};
    const double e2[3] = {b, sg + uy * uy * a, -uy};
    for (int d = 0; d < 3; d++) o[d] = s0 * e1[d] + s1 * e2[d] + chi * (d == 0 ? ux : d == 1 ? uy : uz);
  } else {
    o[0] = s0; o[1] = s1; o[2] = chi;
  }
}

double u24(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0); }

}  // namespace

extern "C" {
// n_hb heat-bath passes of sweep `sweep` (no overrelaxation, no exchange) of
// R configurations (configuration c at temperature temps[c]) in the FP32
// kernel's order: colour classes c = 0 .. nc - 1 in turn, thread t of class
// c updating site order[col_off[c] + t].  Philox counter (site, sweep * 16 +
// pass, c, 0): a thread draws a block for its site of an even class and uses
// words x, y; words z, w serve its site of the next (odd) class; a site of an
// odd class whose thread had no site in the class before draws its own block.
void orc_mc_replay_hb32(void *mp, int R, const double *temps, const int32_t *order,
                        const int32_t *col_off, int nc, int n_hb, uint64_t seed, uint32_t sweep,
                        double *spins) {
  Model &m = *(Model *)mp;
  const int N = m.N;
  for (int c = 0; c < R; c++) {
    double *S = spins + (size_t)c * N * 3;
    const double b = 1.0 / temps[c];
    for (int pass = 0; pass < n_hb; pass++) {
      const uint32_t step = sweep * 16u + (uint32_t)pass;
      std::vector<uint32_t> keep_z, keep_w;
      std::vector<char> keep;
      for (int col = 0; col < nc; col++) {
        const int n = col_off[col + 1] - col_off[col];
        std::vector<uint32_t> nz(n), nw(n);
        std::vector<char> nk(n, 0);
        for (int t = 0; t < n; t++) {
          const int i = order[col_off[col] + t];
          double h[3], o[3];
          local_field(m, S, i, h);
          uint32_t r0, r1;
          if ((col & 1) && t < (int)keep.size() && keep[t]) {
            r0 = keep_z[t]; r1 = keep_w[t];
          } else {
            const W4 r = philox((uint32_t)i, step, (uint32_t)c, 0u, seed);
            r0 = r.v[0]; r1 = r.v[1];
            if (!(col & 1)) { nz[t] = r.v[2]; nw[t] = r.v[3]; nk[t] = 1; }
          }
          heat_bath_f32form(b, h, u24(r0), u24(r1), o);
          S[3 * i] = o[0]; S[3 * i + 1] = o[1]; S[3 * i + 2] = o[2];
        }
        if (!(col & 1)) { keep_z = nz; keep_w = nw; keep = nk; }
        else keep.clear();
      }
    }
  }
}

// --- reference-order MC (statistical oracle and CPU baseline) --------------
// Per sweep and replica: n_or overrelaxation passes then n_hb heat-bath
// passes (Heisenberg) or one Metropolis pass (Ising), each over a fresh
// Fisher-Yates permutation (get_shuffledList CF/subrouts.cpp:2666-2699), the
// bond-list energy, then the replica exchange with the reference's rule.
// Replicas run on OpenMP threads (one MPI rank each in the reference).
// acc (nullable): per slot accumulators [R][2] of E and E^2 over the sweeps.
// cos (slot -> configuration) is read and updated; cnt, exc (exchange
// counters) and boost (pass boost per slot) are read and updated when not
// null, so a chain can be continued over several calls; sweep0 is the index
// of the first sweep (random streams, pair parity s % 2, exchange after
// pt_start).
void orc_mc_reference(void *mp, int model, int R, const double *temps, int n_or, int n_hb,
                      int n_or_b, int n_hb_b, int cap50, uint64_t seed, int n_sweeps,
                      int pt_start, double *spins, int32_t *cos, double *acc, int threads,
                      long long *cnt_io, long long *exc_io, int32_t *boost_io, int sweep0) {
  Model &m = *(Model *)mp;
  const int N = m.N;
  std::vector<double> beta(R), E(R);
  for (int k = 0; k < R; k++) beta[k] = 1.0 / temps[k];
  std::vector<long long> cnt(R, 0), exc(R, 0);
  std::vector<int32_t> boost(R, 0), soc(R);
  if (cnt_io) cnt.assign(cnt_io, cnt_io + R);
  if (exc_io) exc.assign(exc_io, exc_io + R);
  if (boost_io) boost.assign(boost_io, boost_io + R);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  for (int s = sweep0; s < sweep0 + n_sweeps; s++) {
    for (int k = 0; k < R; k++) soc[cos[k]] = k;
    // small chains run on the calling thread: a parallel region per sweep
    // costs more than a sweep of a few dozen sites
#pragma omp parallel for schedule(static) if ((long long)R * N >= 2048)
    for (int c = 0; c < R; c++) {
      double *S = spins + (size_t)c * N * 3;
      const int slot = soc[c];
      const double b = beta[slot];
      std::vector<int> perm(N);
      auto shuffle = [&](uint32_t pass) {
        for (int i = 0; i < N; i++) perm[i] = i;
        for (int i = N - 1; i > 0; --i) {
          const W4 r = philox((uint32_t)i, (uint32_t)s * 16u + pass, (uint32_t)c, 5u, seed);
          const int j = (int)(r.v[0] % (uint32_t)(i + 1));
          std::swap(perm[i], perm[j]);
        }
      };
      if (model == 1) {
        shuffle(0);
        for (int n = 0; n < N; n++) {
          const int i = perm[n];
          const double h = local_field_x(m, S, i);
          const double dE = 2.0 * S[3 * i] * h;
          const W4 r = philox((uint32_t)i, (uint32_t)s, (uint32_t)c, 6u, seed);
          if (u53(r.v[0], r.v[1]) < std::min(1.0, exp(b * dE))) S[3 * i] = -S[3 * i];
        }
        E[c] = energy(m, S);
      } else {
        const int nor = boost[slot] ? n_or_b : n_or;
        const int nhb = boost[slot] ? n_hb_b : n_hb;
        for (int pass = 0; pass < nor + nhb; pass++) {
          shuffle((uint32_t)pass);
          for (int n = 0; n < N; n++) {
            const int i = perm[n];
            double h[3], o[3];
            local_field(m, S, i, h);
            if (pass < nor) {
              reflect(S + 3 * i, h, o);
            } else {
              const W4 r = philox((uint32_t)i, (uint32_t)s * 16u + (uint32_t)pass, (uint32_t)c, 6u, seed);
              heat_bath(b, h, u53(r.v[0], r.v[1]), u53(r.v[2], r.v[3]), o);
            }
            S[3 * i] = o[0]; S[3 * i + 1] = o[1]; S[3 * i + 2] = o[2];
          }
        }
        E[c] = energy(m, S);
      }
    }
    if (acc)
      for (int k = 0; k < R; k++) {
        acc[2 * k] += E[cos[k]];
        acc[2 * k + 1] += E[cos[k]] * E[cos[k]];
      }
    if (s >= pt_start)
      exchange(R, s % 2, cap50 != 0, seed ^ 0x5151ULL, (uint32_t)s, E.data(), beta.data(), cos,
               cnt.data(), exc.data(), boost.data());
    else
      for (int k = 0; k < R; k++) boost[k] = 0;
  }
  if (cnt_io) std::copy(cnt.begin(), cnt.end(), cnt_io);
  if (exc_io) std::copy(exc.begin(), exc.end(), exc_io);
  if (boost_io) std::copy(boost.begin(), boost.end(), boost_io);
}

// --- SMD: get_dynSpectrum time loop (CF/subrouts.cpp:11734-11927) ----------
// spins: N x 3 in/out.  n_points time points (t = 0 included); the step loop
// runs n_points - 1 times.  gauss_seidel = 1 applies the correction in
// place in site order (reference); 0 uses the post-RK fields for every site.
// traj (nullable): n_points x N x 3; qss: n_points; energies (nullable):
// n_points energies after each step.
void orc_smd_run(void *mp, double *spins, double dt, int n_points, int gauss_seidel, double *traj,
                 double *qss, double *energies) {
  Model &m = *(Model *)mp;
  const int N = m.N;
  const double f1d6 = 1.0 / 6.0, f2d6 = 2.0 / 6.0;
  std::vector<double> S0(spins, spins + 3 * N), prev(3 * N), next(3 * N);
  std::vector<double> K1(3 * N), K2(3 * N), K3(3 * N), K4(3 * N);
  if (traj) memcpy(traj, spins, sizeof(double) * 3 * N);
  if (qss) qss[0] = 1.0;
  if (energies) energies[0] = energy(m, spins);
  double *S = spins;
  auto kstage = [&](const std::vector<double> &in, std::vector<double> &K) {
    for (int i = 0; i < N; i++) {
      double h[3];
      local_field(m, in.data(), i, h);
      const double *s = &in[3 * i];
      const double P[3] = {h[1] * s[2] - h[2] * s[1], h[2] * s[0] - h[0] * s[2], h[0] * s[1] - h[1] * s[0]};
      for (int n = 0; n < 3; n++) K[3 * i + n] = dt * P[n];  // damping dpc = 0
    }
  };
  for (int t = 1; t < n_points; t++) {
    const double ERef = energy(m, S);
    prev.assign(S, S + 3 * N);
    kstage(prev, K1);
    for (int q = 0; q < 3 * N; q++) next[q] = S[q] + 0.5 * K1[q];
    kstage(next, K2);
    for (int q = 0; q < 3 * N; q++) next[q] = S[q] + 0.5 * K2[q];
    kstage(next, K3);
    for (int q = 0; q < 3 * N; q++) next[q] = S[q] + K3[q];
    kstage(next, K4);
    for (int i = 0; i < N; i++) {
      double v[3];
      for (int n = 0; n < 3; n++)
        v[n] = S[3 * i + n] + (f1d6 * (K1[3 * i + n] + K4[3 * i + n]) + f2d6 * (K2[3 * i + n] + K3[3 * i + n]));
      const double nr = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int n = 0; n < 3; n++) S[3 * i + n] = v[n] / nr;
    }
    // energy correction (CF/subrouts.cpp:11860-11900)
    double sf = 0.0;
    std::vector<double> D(3 * N);
    for (int i = 0; i < N; i++) {
      double h[3];
      local_field(m, S, i, h);
      const double *s = S + 3 * i;
      const double P[3] = {h[1] * s[2] - h[2] * s[1], h[2] * s[0] - h[0] * s[2], h[0] * s[1] - h[1] * s[0]};
      sf += P[0] * P[0] + P[1] * P[1] + P[2] * P[2];
      D[3 * i] = s[1] * P[2] - s[2] * P[1];
      D[3 * i + 1] = s[2] * P[0] - s[0] * P[2];
      D[3 * i + 2] = s[0] * P[1] - s[1] * P[0];
    }
    const double E1 = energy(m, S);
    sf = (ERef - E1) / sf;
    double proj = 0.0;
    for (int i = 0; i < N; i++) {
      double *s = S + 3 * i;
      double d[3] = {D[3 * i], D[3 * i + 1], D[3 * i + 2]};
      if (gauss_seidel) {
        double h[3];
        local_field(m, S, i, h);
        const double P[3] = {h[1] * s[2] - h[2] * s[1], h[2] * s[0] - h[0] * s[2], h[0] * s[1] - h[1] * s[0]};
        d[0] = s[1] * P[2] - s[2] * P[1];
        d[1] = s[2] * P[0] - s[0] * P[2];
        d[2] = s[0] * P[1] - s[1] * P[0];
      }
      double v[3] = {s[0] + sf * d[0], s[1] + sf * d[1], s[2] + sf * d[2]};
      const double nr = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int n = 0; n < 3; n++) s[n] = v[n] / nr;
      proj += S0[3 * i] * s[0] + S0[3 * i + 1] * s[1] + S0[3 * i + 2] * s[2];
    }
    if (qss) qss[t] = proj / N;
    if (energies) energies[t] = energy(m, S);
    if (traj) memcpy(traj + (size_t)t * 3 * N, S, sizeof(double) * 3 * N);
  }
}

// --- DSSF from a trajectory (CF/subrouts.cpp:8330-8443, 12062-12105) -------
// traj: T x N x 3; gridmap[G*G] -> site or -1; window[T].  For every step the
// spins are scattered on the G x G grid, S(q,t) = DFT2D / sqrt(Nq) (textbook
// forward DFT, q = k1*G + k2), then X(q,n) = sum_t w(t) S(q,t)
// e^{-2 pi i n t / T} / sqrt(T) for n < n_omega, out[d][q][n] = |X|^2.
void orc_dssf(int N, int G, const int32_t *gridmap, int T, const double *traj,
              const double *window, int n_omega, double *out) {
  typedef std::complex<double> C;
  const int Nq = G * G;
  std::vector<C> hist((size_t)3 * Nq * T);
  std::vector<C> tw(G);
  for (int k = 0; k < G; k++) tw[k] = std::polar(1.0, -2.0 * PI * k / G);
  std::vector<C> grid(Nq), tmp(Nq);
  const double fc = sqrt(1.0 / Nq);
  for (int t = 0; t < T; t++)
    for (int d = 0; d < 3; d++) {
      for (int q = 0; q < Nq; q++) {
        const int s = gridmap[q];
        grid[q] = s < 0 ? C(0, 0) : C(traj[((size_t)t * N + s) * 3 + d], 0.0);
      }
      for (int a = 0; a < G; a++)
        for (int b2 = 0; b2 < G; b2++) {
          C acc(0, 0);
          for (int b = 0; b < G; b++) acc += grid[a * G + b] * tw[(b * b2) % G];
          tmp[a * G + b2] = acc;
        }
      for (int a2 = 0; a2 < G; a2++)
        for (int b2 = 0; b2 < G; b2++) {
          C acc(0, 0);
          for (int a = 0; a < G; a++) acc += tmp[a * G + b2] * tw[(a * a2) % G];
          hist[((size_t)d * Nq + a2 * G + b2) * T + t] = fc * acc;
        }
    }
  std::vector<C> twt(T);
  for (int k = 0; k < T; k++) twt[k] = std::polar(1.0, -2.0 * PI * k / T);
  const double sfc = sqrt(1.0 / T);
  for (int d = 0; d < 3; d++)
    for (int q = 0; q < Nq; q++) {
      const C *h = &hist[((size_t)d * Nq + q) * T];
      for (int n = 0; n < n_omega; n++) {
        C acc(0, 0);
        for (int t = 0; t < T; t++) acc += window[t] * h[t] * twt[((long long)n * t) % T];
        const C x = sfc * acc;
        out[((size_t)d * Nq + q) * n_omega + n] = std::real(x * std::conj(x));
      }
    }
}

// --- DSSF by direct spatial sums (get_dynSpectrum with get_qct_SqData,
// CF/subrouts.cpp:11906-11912, 8460-8525, 12062-12105): S(q, t) =
// (1/sqrt(N)) sum_r e^{-i q.r} S(r, t) at the given wavevectors (sites with
// mask[r] != 0 left out), then the windowed time DFT, |X|^2 / T.
void orc_dssf_direct(int N, const double *rvec, const int32_t *mask, int nq, const double *qv,
                     int T, const double *traj, const double *window, int n_omega, double *out) {
  typedef std::complex<double> C;
  const double fc = sqrt(1.0 / N), sfc = sqrt(1.0 / T);
  std::vector<C> hist((size_t)3 * T), ph(N), twt(T);
  for (int k = 0; k < T; k++) twt[k] = std::polar(1.0, -2.0 * PI * k / T);
  for (int q = 0; q < nq; q++) {
    for (int r = 0; r < N; r++)
      ph[r] = (mask && mask[r]) ? C(0, 0)
                                : std::exp(C(0.0, -1.0) * (qv[2 * q] * rvec[2 * r] + qv[2 * q + 1] * rvec[2 * r + 1]));
    for (int t = 0; t < T; t++)
      for (int d = 0; d < 3; d++) {
        C acc(0, 0);
        for (int r = 0; r < N; r++) acc += traj[((size_t)t * N + r) * 3 + d] * ph[r];
        hist[(size_t)d * T + t] = fc * acc;
      }
    for (int d = 0; d < 3; d++)
      for (int n = 0; n < n_omega; n++) {
        C acc(0, 0);
        for (int t = 0; t < T; t++) acc += window[t] * hist[(size_t)d * T + t] * twt[((long long)n * t) % T];
        const C x = sfc * acc;
        out[((size_t)d * nq + q) * n_omega + n] = std::real(x * std::conj(x));
      }
  }
}


// --- order parameters (CF/subrouts.cpp:5362-5800) ---------------------------
// geom: "square", "triang", "hexagn", "kagome", "lieb"; spins N x 3.
// out[8]: |M|/N, M_z/N (abs without a field), absStaggMag, absStripeMag
// Horz / Vert / Diag, SMag1 = sqrt(H^2 + V^2 + D^2 [C3]), and the
// cplxC3Parameter (C3 geometries) or IsingZ2Parameter (square).
void orc_observables(const char *geom, int L, int N, const int32_t *nbors, int zn1, int zn2,
                     int znx, int h_on, const double *S, double *out) {
  const std::string g(geom);
  const bool c3 = (g != "square" && g != "lieb");
  const int Zt = zn1 + zn2 + znx;
  const double iNs = 1.0 / N;
  double M[3] = {0, 0, 0};
  for (int k = 0; k < N; k++) for (int d = 0; d < 3; d++) M[d] += S[3 * k + d];
  out[0] = iNs * sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]);
  out[1] = M[2] * iNs;
  if (!h_on) out[1] = fabs(out[1]);
  auto finish = [&](double sub[][3], int n0, double norm) {
    double sv[3] = {0, 0, 0}, sq = 0;
    for (int n = 0; n < n0; n++) {
      double mv[3];
      for (int d = 0; d < 3; d++) { mv[d] = norm * sub[n][d]; sv[d] += mv[d]; }
      sq += mv[0] * mv[0] + mv[1] * mv[1] + mv[2] * mv[2];
    }
    return (1.0 / sqrt((double)n0)) * sqrt(sq - (sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]) / n0);
  };
  // absStaggMag
  {
    const int n0 = c3 ? 3 : 2, ifc = c3 ? 1 : 0;
    double sub[3][3] = {{0}};
    if (g != "lieb") {
      int k = 0;
      for (int i = 0; i < L; i++) {
        for (int j = 0; j < L; j++) {
          const int st = i * L + j, n = (i + j + ifc * k) % n0;
          for (int d = 0; d < 3; d++) sub[n][d] += S[3 * st + d];
        }
        k += 1;
      }
    } else {
      for (int k = 0; k < N; k++) for (int d = 0; d < 3; d++) sub[k % n0][d] += S[3 * k + d];
    }
    out[2] = finish(sub, n0, n0 * iNs);
  }
  // absStripeMag
  const char *types[3] = {"Horz", "Vert", "Diag"};
  for (int t = 0; t < 3; t++) {
    double sub[2][3] = {{0}};
    const std::string ty(types[t]);
    if (g != "lieb") {
      for (int i = 0; i < L; i++)
        for (int j = 0; j < L; j++) {
          const int st = i * L + j;
          const int n = ty == "Horz" ? j % 2 : ty == "Vert" ? i % 2 : (i + j) % 2;
          for (int d = 0; d < 3; d++) sub[n][d] += S[3 * st + d];
        }
    } else if (ty == "Horz" || ty == "Vert") {
      int n = 0;
      for (int i = 0; i < L; i++)
        for (int j = 0; j < L; j++) {
          const int n1 = (i + j) % 2;
          for (int d = 0; d < 3; d++) {
            if (ty == "Horz") sub[(i + j) & 1][d] += S[3 * n + d] + S[3 * (n + 1) + d];
            else { sub[n1][d] += S[3 * n + d]; sub[1 - n1][d] += S[3 * (n + 1) + d]; }
          }
          n += 2;
        }
    } else {
      for (int k = 0; k < N; k++) for (int d = 0; d < 3; d++) sub[k % 2][d] += S[3 * k + d];
    }
    out[3 + t] = finish(sub, 2, 2.0 * iNs);
  }
  out[6] = sqrt(out[3] * out[3] + out[4] * out[4] + (c3 ? out[5] * out[5] : 0.0));
  out[7] = 0.0;
  if (c3) {
    // cplxC3Parameter, z component (dMag = 2)
    double sm[3] = {0, 0, 0};
    int k = 0;
    for (int i = 0; i < L; i++) {
      for (int j = 0; j < L; j++) sm[(i + j + k) % 3] += S[3 * (i * L + j) + 2];
      k += 1;
    }
    const double re = sqrt(6.0) * iNs * (sm[0] + sm[1] - 2.0 * sm[2]);
    const double im = -3.0 * sqrt(2.0) * iNs * (sm[1] - sm[0]);
    out[7] = sqrt(re * re + im * im);
  } else if (g == "square") {
    // IsingZ2Parameter: [S(i) - S(i+x+y)] . [S(i+x) - S(i+y)], sign per plaquette
    double sum = 0.0;
    for (int i = 0; i < N; i++) {
      const int nxx = nbors[(size_t)i * Zt + 0], nyy = nbors[(size_t)i * Zt + 1];
      const int nxy = nbors[(size_t)i * Zt + zn1 + 0];
      double dp = 0.0;
      for (int d = 0; d < 3; d++)
        dp += (S[3 * i + d] - S[3 * nxy + d]) * (S[3 * nxx + d] - S[3 * nyy + d]);
      const double fc = fabs(dp) > 1.0e-12 ? 1.0 / fabs(dp) : 0.25;
      sum += fc * dp;
    }
    out[7] = iNs * sum;
  }
}

// --- windows (CF/aux_code.cpp:249-435, selection CF/smd_code.cpp:2825-2895) -
int orc_window(const char *name, int N, double *w) {
  const std::string s(name);
  const double tp = 2.0 * PI;
  auto sinc = [](double x) { return x == 0.0 ? 1.0 : sin(x) / x; };
  if (s == "nowin") { for (int i = 0; i < N; i++) w[i] = 1.0; return 0; }
  if (s == "gauss") {
    const double dx = 2.0 / (N - 1.0);
    for (int i = 0; i < N; i++) { const double x = i * dx - 1.0; w[i] = exp(-80.0 * pow(x, 2)); }
    return 0;
  }
  if (s == "lancz" || s == "sincw") {
    const double fc = s == "lancz" ? 0.5 : 1.0;
    const double dx = 2.0 / (N - 1.0);
    for (int i = 0; i < N; i++) {
      const double x = (i * dx - 1.0) * fc;
      w[i] = (s == "lancz" && fabs(x) > 0.5) ? 0.0 : sinc(tp * x);
    }
    return 0;
  }
  if (s == "hann1" || s == "hann2" || s == "hann3") {
    const int pw = s[4] - '0';
    const double dx = 1.0 / (N - 1.0);
    for (int i = 0; i < N; i++) w[i] = pow(0.5 * (1.0 - cos(tp * i * dx)), pw);
    return 0;
  }
  const double dx = 1.0 / (N - 1.0);
  if (s == "ntall" || s == "black") {
    const double *a;
    static const double nut[4] = {0.3635819, 0.4891775, 0.1365995, 0.0106411};
    static const double har[4] = {0.35875, 0.48829, 0.14128, 0.01168};
    a = s == "ntall" ? nut : har;
    for (int i = 0; i < N; i++) {
      const double x = i * dx;
      w[i] = a[0] + (-a[1] * cos(1.0 * tp * x) + a[2] * cos(2.0 * tp * x) - a[3] * cos(3.0 * tp * x));
    }
    return 0;
  }
  if (s == "flatw") {
    const double a[5] = {0.215578950, 0.416631580, 0.277263158, 0.083578947, 0.006947368};
    for (int i = 0; i < N; i++) {
      const double x = i * dx;
      w[i] = a[0] + (-a[1] * cos(1.0 * tp * x) + a[2] * cos(2.0 * tp * x) -
                     a[3] * cos(3.0 * tp * x) + a[4] * cos(4.0 * tp * x));
    }
    return 0;
  }
  return 1;
}

// --- spin stiffness (get_spinStiff, CF/subrouts.cpp:6519-6620) ------------
// Per configuration: rhoS1vec (sum J C^2 (x_i x_j + y_i y_j)) for the three
// directions C0 = dx, C1 = c60 dx + s60 dy, C2 = s60 dy - c60 dx, with raw
// position differences, then the C-sums of J (x_i y_j - y_i x_j) (before the
// reference squares them).  J: shell coupling or disorder matrix entry.
void orc_spin_stiffness(void *mp, const double *rvec, int R, const double *spins, double *out) {
  const Model &m = *(const Model *)mp;
  const double c60 = cos(PI / 3.0), s60 = sin(PI / 3.0);
  for (int r = 0; r < R; r++) {
    const double *S = spins + (size_t)r * m.N * 3;
    double a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
    for (size_t b = 0; b < m.bi.size(); b++) {
      const int i = m.bi[b], j = m.bj[b], sh = m.bs[b];
      const double dx = rvec[2 * i] - rvec[2 * j], dy = rvec[2 * i + 1] - rvec[2 * j + 1];
      const double C[3] = {dx, c60 * dx + s60 * dy, s60 * dy - c60 * dx};
      double J;
      if (m.dis) J = jmat(m, sh, m.imp[i], m.imp[j]);
      else J = sh == 1 ? m.J1 : (sh == 2 ? m.J2 : m.JX);
      const double dp = J * (S[3 * i] * S[3 * j] + S[3 * i + 1] * S[3 * j + 1]);
      const double vp = J * (S[3 * i] * S[3 * j + 1] - S[3 * i + 1] * S[3 * j]);
      for (int k = 0; k < 3; k++) {
        a1[k] += C[k] * C[k] * dp;
        a2[k] += C[k] * vp;
      }
    }
    for (int k = 0; k < 3; k++) {
      out[(size_t)r * 6 + k] = a1[k];
      out[(size_t)r * 6 + 3 + k] = a2[k];
    }
  }
}

// --- static S(q) on the DFT grid (get_SqData_iMKL, CF/subrouts.cpp:8330-8443)
// Textbook 2D DFT of the gridMap-scattered spins (MKL forward, unscaled),
// times 1/sqrt(Nq); out[R][3][G*G][2], index i * G + j.
void orc_sq_grid(int N, int G, const int32_t *gridmap, int R, const double *spins, double *out) {
  const int Nq = G * G;
  const double fc = 1.0 / sqrt((double)Nq);
  std::vector<double> cs(G), sn(G);
  for (int m = 0; m < G; m++) { cs[m] = cos(2.0 * PI * m / G); sn[m] = sin(2.0 * PI * m / G); }
  for (int r = 0; r < R; r++)
    for (int d = 0; d < 3; d++)
      for (int qi = 0; qi < G; qi++)
        for (int qj = 0; qj < G; qj++) {
          double re = 0.0, im = 0.0;
          for (int i = 0; i < G; i++)
            for (int j = 0; j < G; j++) {
              const int k = gridmap[i * G + j];
              if (k < 0) continue;
              const double v = spins[((size_t)r * N + k) * 3 + d];
              const int mm = (int)(((long long)i * qi + (long long)j * qj) % G);
              re += v * cs[mm];
              im -= v * sn[mm];
            }
          double *o = out + (((size_t)r * 3 + d) * Nq + (size_t)qi * G + qj) * 2;
          o[0] = fc * re;
          o[1] = fc * im;
        }
}

// --- static S(q) by direct sums (get_qct_SqData, CF/subrouts.cpp:8460-8525)
// q = dq (nx, ny), nx = i or i - L (i >= L/2), likewise ny from j; output
// index j * L + i of an Nq-long array (entries beyond L^2 stay 0).
void orc_sq_direct(int N, int L, int Nq, double dq, const double *rvec, int R,
                   const double *spins, double *out) {
  const int l2 = L / 2;
  const double fc = 1.0 / sqrt((double)N);
  memset(out, 0, sizeof(double) * (size_t)R * 3 * Nq * 2);
  for (int r = 0; r < R; r++)
    for (int j = 0; j < L; j++) {
      const int ny = j < l2 ? j : j - L;
      for (int i = 0; i < L; i++) {
        const int nx = i < l2 ? i : i - L;
        std::complex<double> S[3];
        for (int k = 0; k < N; k++) {
          const std::complex<double> f =
              std::exp(std::complex<double>(0.0, -dq) * (nx * rvec[2 * k] + ny * rvec[2 * k + 1]));
          for (int d = 0; d < 3; d++) S[d] += spins[((size_t)r * N + k) * 3 + d] * f;
        }
        for (int d = 0; d < 3; d++) {
          double *o = out + (((size_t)r * 3 + d) * Nq + (size_t)j * L + i) * 2;
          o[0] = fc * S[d].real();
          o[1] = fc * S[d].imag();
        }
      }
    }
}

// --- key points of the ordered spectrum (get_StaticSFac CF/subrouts.cpp:8799,
// get_OrderedSpecArray2D :8868, get_corrLen :9335).  sq: [3][G*G][2] of one
// configuration.  peaks[4][5] (xy, yz, zz, tt) = SQ1 at the five points,
// clen[4][5] = sqrt(SQ1/SQ2 - 1)/dq or 0.
void orc_corr_len(int G, int c3, double dq, const double *sq, double *peaks, double *clen) {
  const int Nq = G * G, l2 = G / 2, l3 = G / 3, l4 = G / 4;
  std::vector<int> mv(G);
  for (int i = 0; i < G; i++) { const int n = 2 * (i / l2); mv[i] = i - l2 * (n - 1); }
  // ordered arrays xy, yz, zz, tt
  std::vector<double> ord((size_t)4 * Nq);
  for (int i = 0; i < G; i++)
    for (int j = 0; j < G; j++) {
      const int k = mv[i] * G + mv[j];
      double p[3];
      for (int d = 0; d < 3; d++) {
        const double re = sq[((size_t)d * Nq + k) * 2], im = sq[((size_t)d * Nq + k) * 2 + 1];
        p[d] = std::max(1.0e-16, re * re + im * im);
      }
      ord[(size_t)0 * Nq + i * G + j] = 0.5 * (p[0] + p[1]);
      ord[(size_t)1 * Nq + i * G + j] = 0.5 * (p[1] + p[2]);
      ord[(size_t)2 * Nq + i * G + j] = p[2];
      ord[(size_t)3 * Nq + i * G + j] = (1.0 / 3.0) * (p[0] + p[1] + p[2]);
    }
  int i1[5], j1[5], i2[5], j2[5];
  if (c3) {
    i1[0] = l3 + l2; j1[0] = -l3 + l2; i2[0] = l3 + l2; j2[0] = -l3 + l2 + 1;
    i1[1] = l2; j1[1] = 0; i2[1] = l2; j2[1] = 1;
    i1[2] = 0; j1[2] = l2; i2[2] = 1; j2[2] = l2;
  } else {
    i1[0] = 0; j1[0] = 0; i2[0] = 0; j2[0] = 1;
    i1[1] = 0; j1[1] = l2; i2[1] = 1; j2[1] = l2;
    i1[2] = l2; j1[2] = 0; i2[2] = l2; j2[2] = 1;
  }
  i1[3] = l2; j1[3] = l2; i2[3] = l2; j2[3] = l2 + 1;
  i1[4] = l4; j1[4] = l4; i2[4] = l4; j2[4] = l4 + 1;
  for (int f = 0; f < 4; f++)
    for (int k = 0; k < 5; k++) {
      const double a = ord[(size_t)f * Nq + i1[k] * G + j1[k]];
      const double b = ord[(size_t)f * Nq + i2[k] * G + j2[k]];
      const double w = a / b;
      peaks[f * 5 + k] = a;
      clen[f * 5 + k] = w > 1.0 ? (1.0 / dq) * sqrt(w - 1.0) : 0.0;
    }
}

// Quasicrystal order parameters of R configurations (per configuration:
// [0] NeelPrjt, [1..4] StrpPrjt after the cut to the saturated state,
// [5] QctMag, [6] the signed Neel_N1fpar w0, [7..10] SqPeaks).
//  get_qctNeelPrjt    CF/subrouts.cpp:6053-6084
//  get_qctStrpPrjt    CF/subrouts.cpp:6086-6180 (cut: |P_k| within 1e-12 of 1)
//  get_qctStripeMag   CF/subrouts.cpp:6183-6262 (fcz = 1/znum1; 0 here when a
//                     site has no shell-1 neighbour, where the reference divides by 0)
//  get_qctNeel_N1fpar CF/subrouts.cpp:6011-6050
//  get_qct_SqPks4     CF/subrouts.cpp:8637-8700
// ref: 5 x N x 3 (Neel, stripes 1-4); pk: 8 x (nx, ny); nbors: N x Zt with
// shell 1 in columns 0..zn1-1 (-1 = missing).
void orc_qc_values(int N, const int32_t *nbors, int Zt, int zn1, int R, const double *spins,
                   const double *ref, const double *rvec, int Nq, double dq, const int32_t *pk,
                   double *out) {
  const double iNs = 1.0 / N;
  for (int c = 0; c < R; c++) {
    const double *S = spins + (size_t)c * N * 3;
    double *o = out + (size_t)c * 11;
    double pr[5] = {0, 0, 0, 0, 0};
    for (int r = 0; r < 5; r++) {
      double acc = 0.0;
      for (int k = 0; k < N; k++) {
        const double *rs = ref + ((size_t)r * N + k) * 3;
        acc += S[3 * k] * rs[0] + S[3 * k + 1] * rs[1] + S[3 * k + 2] * rs[2];
      }
      pr[r] = iNs * acc;
    }
    o[0] = pr[0];
    bool cut = false;
    int kmax = 0;
    for (int k = 0; k < 4; k++)
      if (std::abs(std::abs(pr[1 + k]) - 1.0) < 1.0e-12) { kmax = k; cut = true; }
    for (int k = 0; k < 4; k++) o[1 + k] = (cut && k != kmax) ? 0.0 : pr[1 + k];
    // stripe vector
    double m1 = 0.0, m2 = 0.0, fpar = 0.0;
    for (int i = 0; i < N; i++) {
      double set[8][3];
      int zc = 0;
      double sum[3] = {0, 0, 0};
      for (int j = 0; j < zn1; j++) {
        const int nb = nbors[(size_t)i * Zt + j];
        double v[3] = {0, 0, 0};
        if (nb >= 0) {
          zc++;
          for (int d = 0; d < 3; d++) v[d] = S[3 * nb + d];
        }
        for (int d = 0; d < 3; d++) sum[d] += v[d];
        if (j < 8)
          for (int d = 0; d < 3; d++) set[j][d] = v[d];
      }
      for (int j = zn1; j < 8; j++)
        for (int d = 0; d < 3; d++) set[j][d] = 0.0;
      double rx = 0.0, ry = 0.0;
      const double fcz = zc > 0 ? 1.0 / zc : 0.0;
      for (int d = 0; d < 3; d++) {
        const double v1 = (set[0][d] + set[1][d] + set[4][d] + set[5][d]) -
                          (set[2][d] + set[3][d] + set[6][d] + set[7][d]);
        const double v2 = (set[1][d] + set[2][d] + set[5][d] + set[6][d]) -
                          (set[3][d] + set[4][d] + set[7][d] + set[0][d]);
        rx += S[3 * i + d] * v1;
        ry += S[3 * i + d] * v2;
      }
      rx *= fcz;
      ry *= fcz;
      const double rn = sqrt(rx * rx + ry * ry);
      const double rf = rn > 1.0 ? 1.0 / rn : 1.0;
      m1 += rf * rx;
      m2 += rf * ry;
      const double z1 = 1.0 / std::max(1, zc);
      double dot = 0.0;
      for (int d = 0; d < 3; d++) dot += S[3 * i + d] * (z1 * sum[d]);
      fpar += 0.5 * (1.0 - dot);
    }
    o[5] = sqrt((iNs * m1) * (iNs * m1) + (iNs * m2) * (iNs * m2));
    o[6] = iNs * fpar;
    // S(q) at the eight hotspots
    double spec[8];
    for (int h = 0; h < 8; h++) {
      std::complex<double> sq[3];
      for (int k = 0; k < N; k++) {
        const double ph = -dq * (pk[2 * h] * rvec[2 * k] + pk[2 * h + 1] * rvec[2 * k + 1]);
        const std::complex<double> e(cos(ph), sin(ph));
        for (int d = 0; d < 3; d++) sq[d] += S[3 * k + d] * e;
      }
      spec[h] = (1.0 / 3.0) * (std::norm(sq[0]) + std::norm(sq[1]) + std::norm(sq[2]));
    }
    const double fc = 0.5 / Nq;
    for (int j = 0; j < 4; j++) o[7 + j] = sqrt(fc * (spec[2 * j] + spec[2 * j + 1]));
  }
}

// get_qct_SqPkLoc (CF/subrouts.cpp:8531-8632): the hotspots of a reference
// state on the Lsz x Lsz direct-sum grid; pk = 8 x (nx, ny).
void orc_qc_peaks(int N, const double *state, const double *rvec, int Lsz, double dq, int32_t *pk) {
  const int l2 = Lsz / 2;
  double best = 0.0, qxm = 0.0, qym = 0.0;
  for (int j = 0; j < Lsz; j++) {
    const int ny = j < l2 ? j : j - Lsz;
    for (int i = 0; i < Lsz; i++) {
      const int nx = i < l2 ? i : i - Lsz;
      const double qx = nx * dq, qy = ny * dq;
      const double th = atan2(qy, qx), qr = sqrt(qx * qx + qy * qy);
      if (!(qr > 2.0 && qr < 5.0 && th > 0.0 && th < PI / 4.0)) continue;
      std::complex<double> sq[3];
      for (int k = 0; k < N; k++) {
        const std::complex<double> e = std::exp(std::complex<double>(0.0, -dq) *
                                                (double)(nx * rvec[2 * k] + ny * rvec[2 * k + 1]));
        for (int d = 0; d < 3; d++) sq[d] += state[3 * k + d] * e;
      }
      const double v = std::norm(sq[0]) + std::norm(sq[1]) + std::norm(sq[2]);
      if (v > best) { best = v; qxm = qx; qym = qy; }
    }
  }
  const double qx[8] = {qxm, -qxm, qym, -qym, -qxm, qym, qxm, -qym};
  const double qy[8] = {qym, -qym, qxm, -qxm, qym, -qxm, -qym, qxm};
  for (int i = 0; i < 8; i++) {
    pk[2 * i] = (int)std::round(qx[i] / dq);
    pk[2 * i + 1] = (int)std::round(qy[i] / dq);
  }
}

}  // extern "C"
