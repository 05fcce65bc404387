// Static structure factor S(q) of many configurations at once, and the
// measurements the reference derives from it every dftDelay1 MC sweeps
// (CF/mc_hm_code.cpp:5185-5245).
#pragma once
#include "device_common.h"

namespace hsmd {

enum SpinType { SPIN_F32 = 0, SPIN_F64 = 1, SPIN_I8 = 2 };

struct SqPlan {
  bool direct = false;   // direct sums (tabulated / open) or the grid DFT
  int N = 0, Npad = 0;
  int G = 0;             // Gsz: side of the ordered spectrum arrays
  int Lq = 0;            // direct sums: side of the q grid (Lsz)
  int Nq = 0;            // outputs per component (Gsz^2)
  double fc = 1.0;       // 1/sqrt(Nq) (grid) or 1/sqrt(N) (direct)
  double dq = 0.0;
  bool c3 = false;       // key points of get_corrLen: hexagonal zone or square
  int R = 0;             // configurations the scratch is sized for
  DevBuf<int32_t> gmap;  // grid point -> new site label or -1
  DevBuf<double> tw;     // cos / sin of 2 pi m / G, m < G
  DevBuf<double> pos;    // x, y of every new label (direct sums)
  DevBuf<double> rows;   // grid DFT scratch, R x 3 x G x G complex
  DevBuf<double> sq;     // R x 3 x Nq complex
};

// Geometry of S(q) for lattice L with scratch for R configurations.
void sq_plan_init(SqPlan &p, const Lattice &L, int R);
// S(q) of configurations [0, R) of `spins` (per configuration comps x Npad,
// new labels) into p.sq, on the null stream.
void sq_compute(SqPlan &p, const void *spins, SpinType type, int R);
// Adds |S| and arg S of every slot's configuration to sqfield (R x Nq x 6)
// and the key-point peaks, correlation lengths and the measurement count to
// moments (R x nmom, entries mom_base .. mom_base + 40).
void sq_accumulate(SqPlan &p, int R, const int32_t *config_of_slot, double *sqfield,
                   double *moments, int nmom, int mom_base);

// One-off S(q) of host configurations (reference site order, FP64).
void static_sq(const Lattice &L, int n_configs, const double *spins, double *out);

}  // namespace hsmd
