// Quasicrystal measurement set-up: the reference states the projection
// order parameters are taken against, and the eight S(q) hotspots.
//
// The reference reads the Neel and four stripe states from
// stateNeel0/stateStrp1..4_AppxtSz(N).bin (CF/mc_hm_code.cpp:1747-1860) and
// locates the hotspots once from the Neel state or the first stripe state
// (qct_NeelPhase, CF/mc_hm_code.cpp:4451-4457) with get_qct_SqPkLoc
// (CF/subrouts.cpp:8531-8632): the largest sum_d |S_d(q)|^2 over the grid
// q = dq (nx, ny), nx, ny in [-Lsz/2, Lsz/2), restricted to 2 < |q| < 5 and
// 0 < arg q < pi/4, then the eight images of that q under the square's
// rotations and reflections.  (When no grid point lies in that region the
// reference leaves the peak unset; here it is q = 0.)
#include <math.h>

#include <complex>
#include <vector>

#include "hsmd_internal.h"

namespace hsmd {

void lattice_set_qc(Lattice &L, const double *neel, const double *strp, bool neel_phase) {
  const int N = L.N;
  if (L.rvec.size() != (size_t)2 * N) throw Error(HSMD_ERR_ARG, "quasicrystal states need site positions");
  L.qc_ref.assign((size_t)5 * N * 3, 0.0);
  std::copy(neel, neel + (size_t)N * 3, L.qc_ref.begin());
  std::copy(strp, strp + (size_t)4 * N * 3, L.qc_ref.begin() + (size_t)N * 3);
  const double *f = neel_phase ? neel : strp;
  const int Lsz = L.L, l2 = Lsz / 2;
  const double dq = L.sq_dq();
  const double a45 = acos(-1.0) / 4.0;
  double best = 0.0, qxm = 0.0, qym = 0.0;
  for (int j = 0; j < Lsz; j++) {
    const int ny = j < l2 ? j : j - Lsz;
    const double qy = ny * dq;
    for (int i = 0; i < Lsz; i++) {
      const int nx = i < l2 ? i : i - Lsz;
      const double qx = nx * dq;
      const double th = atan2(qy, qx), qr = sqrt(qx * qx + qy * qy);
      if (!(qr > 2.0 && qr < 5.0 && th > 0.0 && th < a45)) continue;
      std::complex<double> s[3];
      for (int k = 0; k < N; k++) {
        const double ph = -dq * (nx * L.rvec[2 * (size_t)k] + ny * L.rvec[2 * (size_t)k + 1]);
        const std::complex<double> e(cos(ph), sin(ph));
        for (int d = 0; d < 3; d++) s[d] += f[(size_t)k * 3 + d] * e;
      }
      const double v = std::norm(s[0]) + std::norm(s[1]) + std::norm(s[2]);
      if (v > best) {
        best = v;
        qxm = qx;
        qym = qy;
      }
    }
  }
  const double qx[8] = {qxm, -qxm, qym, -qym, -qxm, qym, qxm, -qym};
  const double qy[8] = {qym, -qym, qxm, -qxm, qym, -qxm, -qym, qxm};
  for (int i = 0; i < 8; i++) {
    L.qc_nx[i] = (int)lround(qx[i] / dq);
    L.qc_ny[i] = (int)lround(qy[i] / dq);
  }
}

}  // namespace hsmd
