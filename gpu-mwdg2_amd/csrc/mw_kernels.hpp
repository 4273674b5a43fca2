// Device-side data structures and kernel declarations of the MI355X step.
#pragma once

#include "mw_math.hpp"

namespace mw {

constexpr int WG = 256;          // work-group size (4 wave64s)
constexpr int EMIT_ITEMS = 4;    // level-(L-1) nodes per thread in leaf emission
constexpr int EMIT_TILE = WG * EMIT_ITEMS;
constexpr int MAX_SRC = 4096;
constexpr int RED_SLOTS = 4096;  // max work-groups of a leaf kernel (one reduction slot each)

// Scalars that live on the device for the whole run, so that a step needs no
// host round trip (the reference reads dt, N_a and the maxima back every step).
template <typename real>
struct DevState {
  real time, dt, sim_time;
  real mx[5];                     // eta, h, qx, qy(=qx), z used by this step's encode
  int n_leaves;
  int done;                       // set once time >= sim_time at the top of a step
  int first;                      // 1 until the first step's prologue ran
  int steps;
  real inlet[4];                  // N, E, S, W boundary inlet values
  int row[4];                     // time-series cursors
  int n_src;
  unsigned long long leaf_sum;    // leaves updated, summed over steps
  unsigned long long wet_sum[2];  // leaves of the wet lists of stage 1 and 2, summed over steps
};

// Time series of the boundaries (entries 0..3) and point sources (4..).
template <typename real>
struct Series {
  const int* off;
  const int* len;
  const real* t;
  const real* v;
};

template <typename real>
struct Params {
  int L, N, D, H, mesh_dim, xsz, ysz;
  int solver;
  real eps, dx, g, tol_h, tol_q, tol_s, manning, initial_tstep, wall_height, CFL, tol_Krivo;
  real eps_extra_factor;          // 2^(M_BAR+1), evaluated on the host
  int limitslopes;
  int btype[4];                   // N, E, S, W boundary type
  int bstart[4], bend[4];
};

template <typename real>
struct Hier {                     // flow scale coefficients, eta form
  real* f[9];                     // eta0 eta1x eta1y qx0 qx1x qx1y qy0 qy1x qy1y
};

template <typename real>
struct ZHier {
  real* z[3];                     // z0 z1x z1y
};

template <typename real>
struct Sources {
  int n;
  const int* code;
  const int* type;
  real* value;                    // updated in place by the prologue for HVAR/QVAR
  int* row;
};

}  // namespace mw
