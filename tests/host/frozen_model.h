// Host (CPU) model of one island solve built from the library's templated step functions
// (box2d-rs_amd/csrc/b2g_step_t.h), for checks on the CPU. Test infrastructure only.
//
// solve_step runs B2world::solve's island solve (src/private/dynamics/b2_island_private.rs:129-328)
// for all bodies at once, visiting the constraints in a given order; islands (connected
// components through non-static bodies) keep the reference's per-island early exit of the
// position iterations. With T = double it is compared with the oracle's f64 island solver.
#pragma once
#include <vector>

#include "../../box2d-rs_amd/csrc/b2g_step_t.h"

namespace fm {
using namespace b2g;

struct StepParams {
  double dt, dt_ratio, gx, gy;
  int vel_iters, pos_iters;
  bool warm, block;
};

inline int uf_find(std::vector<int>& parent, int x) {
  while (parent[x] != x) {
    parent[x] = parent[parent[x]];
    x = parent[x];
  }
  return x;
}

// island id per contact (root of the union-find over non-static bodies)
inline std::vector<int> contact_islands(int nb, const int* btype, int nc, const int* cb) {
  std::vector<int> parent(nb);
  for (int i = 0; i < nb; ++i) parent[i] = i;
  for (int k = 0; k < nc; ++k) {
    int a = cb[2 * k], b = cb[2 * k + 1];
    if (btype[a] == 0 || btype[b] == 0) continue;
    int ra = uf_find(parent, a), rb = uf_find(parent, b);
    if (ra != rb) parent[ra > rb ? ra : rb] = ra < rb ? ra : rb;
  }
  std::vector<int> isl(nc);
  for (int k = 0; k < nc; ++k) {
    int a = cb[2 * k];
    isl[k] = uf_find(parent, btype[a] != 0 ? a : cb[2 * k + 1]);
  }
  return isl;
}

// One solve. lam holds 4 impulses per contact (normal0, tangent0, normal1, tangent1): on entry
// the manifold's carried impulses, on return the stored ones. pos_runs (optional) receives the
// position iterations each contact took part in.
template <class T>
void solve_step(int nb, const int* btype, const TBodyMass<double>* bm, const double* force, TPos<T>* pos, TVel<T>* vel,
                int nc, const int* cb, const TContactConst<double>* kc, const TGeom<T>* geom, const T* friction, T* lam,
                const int* order, const StepParams& sp, int* pos_runs = nullptr) {
  for (int b = 0; b < nb; ++b)
    if (btype[b] == 2)
      integrate_velocity_t(vel[b], bm[b], force[3 * b], force[3 * b + 1], force[3 * b + 2], sp.gx, sp.gy, sp.dt);
  std::vector<TVelC<T>> vc(nc);
  std::vector<T> l(4 * nc, T(0.0));
  for (int i = 0; i < nc; ++i) {
    int k = order[i];
    int a = cb[2 * k], b = cb[2 * k + 1];
    vc[k] = init_velocity_t(kc[k], geom[k], friction[k], pos[a], pos[b], vel[a], vel[b], sp.block);
    if (sp.warm) {
      l[4 * k] = T(sp.dt_ratio) * lam[4 * k];
      l[4 * k + 1] = T(sp.dt_ratio) * lam[4 * k + 1];
      if (kc[k].count > 1) {
        l[4 * k + 2] = T(sp.dt_ratio) * lam[4 * k + 2];
        l[4 * k + 3] = T(sp.dt_ratio) * lam[4 * k + 3];
      }
    }
  }
  auto apply = [&](int k, auto&& fn) {
    int a = cb[2 * k], b = cb[2 * k + 1];
    TVel<T> A = vel[a], B = vel[b];
    fn(A, B);
    if (btype[a] == 2) vel[a] = A;
    if (btype[b] == 2) vel[b] = B;
  };
  if (sp.warm)
    for (int i = 0; i < nc; ++i) {
      int k = order[i];
      apply(k, [&](TVel<T>& A, TVel<T>& B) { warm_start_t(vc[k], kc[k], &l[4 * k], A, B); });
    }
  for (int it = 0; it < sp.vel_iters; ++it)
    for (int i = 0; i < nc; ++i) {
      int k = order[i];
      apply(k, [&](TVel<T>& A, TVel<T>& B) { solve_velocity_t(vc[k], kc[k], &l[4 * k], A, B, sp.block); });
    }
  for (int k = 0; k < nc; ++k) {
    lam[4 * k] = l[4 * k];
    lam[4 * k + 1] = l[4 * k + 1];
    if (vc[k].vcount > 1) {
      lam[4 * k + 2] = l[4 * k + 2];
      lam[4 * k + 3] = l[4 * k + 3];
    }
  }
  for (int b = 0; b < nb; ++b)
    if (btype[b] != 0) integrate_position_t(pos[b], vel[b], sp.dt);
  std::vector<int> isl = contact_islands(nb, btype, nc, cb);
  std::vector<char> done(nb, 0);
  std::vector<double> minsep(nb, 0.0);
  if (pos_runs)
    for (int k = 0; k < nc; ++k) pos_runs[k] = 0;
  for (int it = 0; it < sp.pos_iters; ++it) {
    for (int i = 0; i < nc; ++i) {
      int k = order[i];
      if (done[isl[k]]) continue;
      int a = cb[2 * k], b = cb[2 * k + 1];
      TPos<T> A = pos[a], B = pos[b];
      T ms = solve_position_t(kc[k], geom[k], A, B);
      double m = double(primal(ms));
      if (m < minsep[isl[k]]) minsep[isl[k]] = m;
      if (btype[a] != 0) pos[a] = A;
      if (btype[b] != 0) pos[b] = B;
      if (pos_runs) ++pos_runs[k];
    }
    // the early-exit test in the model's own precision (f32: as the kernels, f64: as the oracle)
    const double thr = sizeof(primal(T(0.0))) == sizeof(float) ? double(-3.0f * kLinearSlop) : -3.0 * kLinearSlopD;
    for (int k = 0; k < nc; ++k) {
      int r = isl[k];
      if (!done[r] && minsep[r] >= thr) done[r] = 1;
    }
    for (int k = 0; k < nc; ++k) minsep[isl[k]] = 0.0;
  }
}

}  // namespace fm
