// Host f64 model of a recorded run of frozen steps (the GPU tape's structure) and its reverse
// pass, built from the library's templated step functions and vector-Jacobian products
// (box2d-rs_amd/csrc/b2g_step_t.h, b2g_adjoint_t.h) instantiated in double. Used on the CPU for
// central differences of the frozen model and as the f64 reference of the GPU adjoint. Test
// infrastructure only. Constraints are visited sequentially in slot order (the recorded solve
// order), so no colouring is needed here.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../box2d-rs_amd/csrc/b2g_adjoint_t.h"
#include "shape_rec.h"

using namespace b2g;

namespace {

struct Step {
  int n = 0;
  long long off = 0;
  double dt = 0, dt_ratio = 0;
  int vel_iters = 0, pos_iters = 0;
  bool warm = true, block = true;
};

struct Model {
  int nb = 0, nf = 0;
  std::vector<int> btype;
  std::vector<TBodyMass<double>> mass;
  std::vector<double> force;  // 3 per body
  std::vector<ShapeV> shape;
  std::vector<double> mat;    // friction, restitution, threshold per fixture
  double gx = 0, gy = -10;
  std::vector<Step> steps;
  long long total = 0;
  std::vector<int> slot;       // 10 per slot (b2g_tape_export layout)
  std::vector<double> lam0;    // 4 per slot
  std::vector<double> store;   // 4 per slot, written by the forward model
  int mismatches = 0;
  std::vector<int> mismatch_body;  // per body: 1 if a slot of it mismatched in the last pass
};

TContactConst<double> contact_const(const Model& m, const int* sl) {
  TContactConst<double> k;
  k.type = sl[4] & 0xff;
  k.count = sl[4] >> 8;
  const double* ma = &m.mat[3 * sl[2]];
  const double* mb = &m.mat[3 * sl[3]];
  k.restitution = ma[1] > mb[1] ? ma[1] : mb[1];
  k.threshold = ma[2] < mb[2] ? ma[2] : mb[2];
  k.rA = m.shape[sl[2]].radius;
  k.rB = m.shape[sl[3]].radius;
  const TBodyMass<double>& A = m.mass[sl[0]];
  const TBodyMass<double>& B = m.mass[sl[1]];
  k.mA = A.inv_m; k.iA = A.inv_i; k.mB = B.inv_m; k.iB = B.inv_i;
  k.lcA[0] = A.lcx; k.lcA[1] = A.lcy; k.lcB[0] = B.lcx; k.lcB[1] = B.lcy;
  return k;
}

// one step's working data, kept for the reverse pass
struct StepTape {
  std::vector<double> q0, v0, vstar;  // 3 per body
  std::vector<double> geom;           // 8 per slot
  std::vector<double> lam_init, lam0eff, mu;
  std::vector<double> ws;             // 6 per slot
  std::vector<double> vel;            // iteration-major: 10 per slot
  std::vector<double> ip;             // 6 per body
  std::vector<double> pos;            // iteration-major: 6 per slot
  std::vector<int> vcount;
};

void vel_inputs(const Model& m, const StepTape& t, int i, const int* sl, const double* vA, const double* vB,
                const double* lam, double* x) {
  for (int j = 0; j < 3; ++j) {
    x[j] = vA[j];
    x[3 + j] = vB[j];
    x[10 + j] = t.q0[3 * sl[0] + j];
    x[13 + j] = t.q0[3 * sl[1] + j];
    x[25 + j] = t.vstar[3 * sl[0] + j];
    x[28 + j] = t.vstar[3 * sl[1] + j];
  }
  for (int j = 0; j < 4; ++j) x[6 + j] = lam[j];
  for (int j = 0; j < 8; ++j) x[16 + j] = t.geom[8 * i + j];
  x[24] = t.mu[i];
}

// forward of one step on q, v (3 per body); fills the tape
void forward_step(Model& m, const Step& st, std::vector<double>& q, std::vector<double>& v, StepTape& t) {
  const int n = st.n;
  t.q0 = q;
  t.v0 = v;
  t.geom.assign(8 * size_t(n), 0.0);
  t.mu.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    const int* sl = &m.slot[10 * (st.off + i)];
    const double la[2] = {m.mass[sl[0]].lcx, m.mass[sl[0]].lcy}, lb[2] = {m.mass[sl[1]].lcx, m.mass[sl[1]].lcy};
    int hdr = 0;
    uint32_t id[2] = {0, 0};
    GeomFn<double> f{&m.shape[sl[2]], &m.shape[sl[3]], la, lb, &hdr, id};
    double x[6] = {q[3 * sl[0]], q[3 * sl[0] + 1], q[3 * sl[0] + 2], q[3 * sl[1]], q[3 * sl[1] + 1], q[3 * sl[1] + 2]};
    f(x, &t.geom[8 * i]);
    if (hdr != sl[4] || id[0] != uint32_t(sl[8]) || id[1] != uint32_t(sl[9])) {
      ++m.mismatches;
      m.mismatch_body[sl[0]] = m.mismatch_body[sl[1]] = 1;
    }
    t.mu[i] = std::sqrt(m.mat[3 * sl[2]] * m.mat[3 * sl[3]]);
  }
  for (int b = 0; b < m.nb; ++b)
    if (m.btype[b] == 2) {
      IntVelFn<double> fn{m.mass[b], m.force[3 * b], m.force[3 * b + 1], m.force[3 * b + 2], m.gx, m.gy, st.dt};
      double y[3];
      fn(&v[3 * b], y);
      for (int j = 0; j < 3; ++j) v[3 * b + j] = y[j];
    }
  t.vstar = v;
  t.lam_init.assign(4 * size_t(n), 0.0);
  t.lam0eff.assign(4 * size_t(n), 0.0);
  std::vector<double> lam(4 * size_t(n), 0.0);
  for (int i = 0; i < n; ++i) {
    const long long s = st.off + i;
    const int* sl = &m.slot[10 * s];
    const int count = sl[4] >> 8;
    for (int p = 0; p < 2; ++p) {
      const int src = sl[5 + p];
      for (int j = 0; j < 2; ++j)
        t.lam0eff[4 * i + 2 * p + j] =
            src >= 0 ? m.store[4 * size_t(src >> 1) + 2 * (src & 1) + j] : m.lam0[4 * s + 2 * p + j];
    }
    if (st.warm)
      for (int j = 0; j < 4; ++j)
        if (j < 2 || count > 1) t.lam_init[4 * i + j] = st.dt_ratio * t.lam0eff[4 * i + j];
  }
  lam = t.lam_init;
  auto write_v = [&](const int* sl, const double* y) {
    for (int j = 0; j < 3; ++j) {
      if (m.btype[sl[0]] == 2) v[3 * sl[0] + j] = y[j];
      if (m.btype[sl[1]] == 2) v[3 * sl[1] + j] = y[3 + j];
    }
  };
  t.ws.assign(6 * size_t(n), 0.0);
  if (st.warm)
    for (int i = 0; i < n; ++i) {
      const int* sl = &m.slot[10 * (st.off + i)];
      double x[kVelIn], y[6];
      for (int j = 0; j < 3; ++j) {
        t.ws[6 * i + j] = v[3 * sl[0] + j];
        t.ws[6 * i + 3 + j] = v[3 * sl[1] + j];
      }
      vel_inputs(m, t, i, sl, &v[3 * sl[0]], &v[3 * sl[1]], &lam[4 * i], x);
      WarmFn<double>{contact_const(m, sl), st.block}(x, y);
      write_v(sl, y);
    }
  t.vel.assign(10 * size_t(n) * st.vel_iters, 0.0);
  for (int it = 0; it < st.vel_iters; ++it)
    for (int i = 0; i < n; ++i) {
      const int* sl = &m.slot[10 * (st.off + i)];
      double* rec = &t.vel[10 * (size_t(it) * n + i)];
      for (int j = 0; j < 3; ++j) {
        rec[j] = v[3 * sl[0] + j];
        rec[3 + j] = v[3 * sl[1] + j];
      }
      for (int j = 0; j < 4; ++j) rec[6 + j] = lam[4 * i + j];
      double x[kVelIn], y[10];
      vel_inputs(m, t, i, sl, &v[3 * sl[0]], &v[3 * sl[1]], &lam[4 * i], x);
      VelFn<double>{contact_const(m, sl), st.block}(x, y);
      write_v(sl, y);
      for (int j = 0; j < 4; ++j) lam[4 * i + j] = y[6 + j];
    }
  t.vcount.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    const long long s = st.off + i;
    const int* sl = &m.slot[10 * s];
    double x[kVelIn];
    const double z[4] = {0, 0, 0, 0};
    vel_inputs(m, t, i, sl, z, z, z, x);
    const TContactConst<double> k = contact_const(m, sl);
    const int vc = velc_of(x, k, st.block).vcount;
    t.vcount[i] = vc;
    double* o = &m.store[4 * s];
    o[0] = lam[4 * i];
    o[1] = lam[4 * i + 1];
    o[2] = vc > 1 ? lam[4 * i + 2] : (k.count > 1 ? t.lam0eff[4 * i + 2] : 0.0);
    o[3] = vc > 1 ? lam[4 * i + 3] : (k.count > 1 ? t.lam0eff[4 * i + 3] : 0.0);
  }
  t.ip.assign(6 * size_t(m.nb), 0.0);
  for (int b = 0; b < m.nb; ++b) {
    if (m.btype[b] == 0) continue;
    double x[6] = {q[3 * b], q[3 * b + 1], q[3 * b + 2], v[3 * b], v[3 * b + 1], v[3 * b + 2]}, y[6];
    for (int j = 0; j < 6; ++j) t.ip[6 * b + j] = x[j];
    IntPosFn<double>{st.dt}(x, y);
    for (int j = 0; j < 3; ++j) {
      q[3 * b + j] = y[j];
      v[3 * b + j] = y[3 + j];
    }
  }
  t.pos.assign(6 * size_t(n) * st.pos_iters, 0.0);
  for (int it = 0; it < st.pos_iters; ++it)
    for (int i = 0; i < n; ++i) {
      const int* sl = &m.slot[10 * (st.off + i)];
      if (sl[7] <= it) continue;
      double x[14], y[6];
      for (int j = 0; j < 3; ++j) {
        x[j] = q[3 * sl[0] + j];
        x[3 + j] = q[3 * sl[1] + j];
      }
      for (int j = 0; j < 8; ++j) x[6 + j] = t.geom[8 * i + j];
      for (int j = 0; j < 6; ++j) t.pos[6 * (size_t(it) * n + i) + j] = x[j];
      PosFn<double>{contact_const(m, sl)}(x, y);
      for (int j = 0; j < 3; ++j) {
        if (m.btype[sl[0]] == 2) q[3 * sl[0] + j] = y[j];
        if (m.btype[sl[1]] == 2) q[3 * sl[1] + j] = y[3 + j];
      }
    }
}

// reverse of one step: qb, vb (3 per body) and store_bar (4 per slot, all steps) in/out
void reverse_step(Model& m, const Step& st, const StepTape& t, std::vector<double>& qb, std::vector<double>& vb,
                  std::vector<double>& store_bar, std::vector<double>& fric_grad) {
  const int n = st.n;
  std::vector<double> gbar(8 * size_t(n), 0.0), qnbar(6 * size_t(n), 0.0), vsbar(6 * size_t(n), 0.0),
      fricbar(n, 0.0), lambar(4 * size_t(n), 0.0), lam0bar(4 * size_t(n), 0.0);
  auto dyn = [&](int b) { return m.btype[b] == 2; };
  for (int it = st.pos_iters - 1; it >= 0; --it)
    for (int i = n - 1; i >= 0; --i) {
      const int* sl = &m.slot[10 * (st.off + i)];
      if (sl[7] <= it) continue;
      double x[14], yb[6], xb[14] = {0};
      for (int j = 0; j < 6; ++j) x[j] = t.pos[6 * (size_t(it) * n + i) + j];
      for (int j = 0; j < 8; ++j) x[6 + j] = t.geom[8 * i + j];
      for (int j = 0; j < 3; ++j) {
        yb[j] = dyn(sl[0]) ? qb[3 * sl[0] + j] : 0.0;
        yb[3 + j] = dyn(sl[1]) ? qb[3 * sl[1] + j] : 0.0;
      }
      vjp<14, 6, 7>(x, yb, xb, PosFn<double>{contact_const(m, sl)});
      for (int j = 0; j < 3; ++j) {
        if (dyn(sl[0])) qb[3 * sl[0] + j] = xb[j];
        if (dyn(sl[1])) qb[3 * sl[1] + j] = xb[3 + j];
      }
      for (int j = 0; j < 8; ++j) gbar[8 * i + j] += xb[6 + j];
    }
  for (int b = 0; b < m.nb; ++b) {
    if (!dyn(b)) continue;
    double yb[6] = {qb[3 * b], qb[3 * b + 1], qb[3 * b + 2], vb[3 * b], vb[3 * b + 1], vb[3 * b + 2]}, xb[6] = {0};
    vjp<6, 6, 6>(&t.ip[6 * b], yb, xb, IntPosFn<double>{st.dt});
    for (int j = 0; j < 3; ++j) {
      qb[3 * b + j] = xb[j];
      vb[3 * b + j] = xb[3 + j];
    }
  }
  for (int i = 0; i < n; ++i) {
    const long long s = st.off + i;
    const int count = m.slot[10 * s + 4] >> 8;
    const double* g = &store_bar[4 * s];
    lambar[4 * i] = g[0];
    lambar[4 * i + 1] = g[1];
    if (t.vcount[i] > 1) {
      lambar[4 * i + 2] = g[2];
      lambar[4 * i + 3] = g[3];
    } else if (count > 1) {
      lam0bar[4 * i + 2] = g[2];
      lam0bar[4 * i + 3] = g[3];
    }
  }
  auto accumulate = [&](int i, const double* xb) {
    for (int j = 0; j < 6; ++j) qnbar[6 * i + j] += xb[10 + j];
    for (int j = 0; j < 8; ++j) gbar[8 * i + j] += xb[16 + j];
    fricbar[i] += xb[24];
    for (int j = 0; j < 6; ++j) vsbar[6 * i + j] += xb[25 + j];
  };
  for (int it = st.vel_iters - 1; it >= 0; --it)
    for (int i = n - 1; i >= 0; --i) {
      const int* sl = &m.slot[10 * (st.off + i)];
      const double* rec = &t.vel[10 * (size_t(it) * n + i)];
      double x[kVelIn], yb[10], xb[kVelIn] = {0};
      vel_inputs(m, t, i, sl, rec, rec + 3, rec + 6, x);
      for (int j = 0; j < 3; ++j) {
        yb[j] = dyn(sl[0]) ? vb[3 * sl[0] + j] : 0.0;
        yb[3 + j] = dyn(sl[1]) ? vb[3 * sl[1] + j] : 0.0;
      }
      for (int j = 0; j < 4; ++j) yb[6 + j] = lambar[4 * i + j];
      vjp<kVelIn, 10, 8>(x, yb, xb, VelFn<double>{contact_const(m, sl), st.block});
      for (int j = 0; j < 3; ++j) {
        if (dyn(sl[0])) vb[3 * sl[0] + j] = xb[j];
        if (dyn(sl[1])) vb[3 * sl[1] + j] = xb[3 + j];
      }
      for (int j = 0; j < 4; ++j) lambar[4 * i + j] = xb[6 + j];
      accumulate(i, xb);
    }
  if (st.warm)
    for (int i = n - 1; i >= 0; --i) {
      const int* sl = &m.slot[10 * (st.off + i)];
      double x[kVelIn], yb[6], xb[kVelIn] = {0};
      vel_inputs(m, t, i, sl, &t.ws[6 * i], &t.ws[6 * i + 3], &t.lam_init[4 * i], x);
      for (int j = 0; j < 3; ++j) {
        yb[j] = dyn(sl[0]) ? vb[3 * sl[0] + j] : 0.0;
        yb[3 + j] = dyn(sl[1]) ? vb[3 * sl[1] + j] : 0.0;
      }
      vjp<kVelIn, 6, 8>(x, yb, xb, WarmFn<double>{contact_const(m, sl), st.block});
      for (int j = 0; j < 3; ++j) {
        if (dyn(sl[0])) vb[3 * sl[0] + j] = xb[j];
        if (dyn(sl[1])) vb[3 * sl[1] + j] = xb[3 + j];
      }
      for (int j = 0; j < 4; ++j) lambar[4 * i + j] += xb[6 + j];
      accumulate(i, xb);
    }
  for (int i = 0; i < n; ++i) {
    const long long s = st.off + i;
    const int* sl = &m.slot[10 * s];
    const int count = sl[4] >> 8;
    double eb[4];
    for (int j = 0; j < 4; ++j) {
      eb[j] = lam0bar[4 * i + j];
      if (st.warm && (j < 2 || count > 1)) eb[j] += st.dt_ratio * lambar[4 * i + j];
    }
    for (int p = 0; p < 2; ++p) {
      const int src = sl[5 + p];
      if (src < 0) continue;
      store_bar[4 * size_t(src >> 1) + 2 * (src & 1)] += eb[2 * p];
      store_bar[4 * size_t(src >> 1) + 2 * (src & 1) + 1] += eb[2 * p + 1];
    }
    const double fa = m.mat[3 * sl[2]], fb = m.mat[3 * sl[3]];
    const double mu = std::sqrt(fa * fb);
    if (mu > 0) {
      fric_grad[sl[2]] += fricbar[i] * 0.5 * fb / mu;
      fric_grad[sl[3]] += fricbar[i] * 0.5 * fa / mu;
    }
    const double la[2] = {m.mass[sl[0]].lcx, m.mass[sl[0]].lcy}, lb[2] = {m.mass[sl[1]].lcx, m.mass[sl[1]].lcy};
    double x[6], xb[6] = {0};
    for (int j = 0; j < 3; ++j) {
      x[j] = t.q0[3 * sl[0] + j];
      x[3 + j] = t.q0[3 * sl[1] + j];
    }
    vjp<6, 8, 6>(x, &gbar[8 * i], xb, GeomFn<double>{&m.shape[sl[2]], &m.shape[sl[3]], la, lb, nullptr, nullptr});
    for (int j = 0; j < 3; ++j) {
      if (dyn(sl[0])) {
        qb[3 * sl[0] + j] += qnbar[6 * i + j] + xb[j];
        vb[3 * sl[0] + j] += vsbar[6 * i + j];
      }
      if (dyn(sl[1])) {
        qb[3 * sl[1] + j] += qnbar[6 * i + 3 + j] + xb[3 + j];
        vb[3 * sl[1] + j] += vsbar[6 * i + 3 + j];
      }
    }
  }
  for (int b = 0; b < m.nb; ++b) {
    if (!dyn(b)) continue;
    double xb[3] = {0};
    vjp<3, 3, 3>(&t.v0[3 * b], &vb[3 * b], xb,
                 IntVelFn<double>{m.mass[b], m.force[3 * b], m.force[3 * b + 1], m.force[3 * b + 2], m.gx, m.gy, st.dt});
    for (int j = 0; j < 3; ++j) vb[3 * b + j] = xb[j];
  }
}

void split(const double* s6, int nb, std::vector<double>& q, std::vector<double>& v) {
  q.resize(3 * size_t(nb));
  v.resize(3 * size_t(nb));
  for (int b = 0; b < nb; ++b)
    for (int j = 0; j < 3; ++j) {
      q[3 * b + j] = s6[6 * b + j];
      v[3 * b + j] = s6[6 * b + 3 + j];
    }
}
void join(double* s6, int nb, const std::vector<double>& q, const std::vector<double>& v) {
  for (int b = 0; b < nb; ++b)
    for (int j = 0; j < 3; ++j) {
      s6[6 * b + j] = q[3 * b + j];
      s6[6 * b + 3 + j] = v[3 * b + j];
    }
}

}  // namespace

extern "C" {

// bodies: type, mass data (8 doubles: inv_m, inv_I, lc.x, lc.y, mass, gravity scale, linear and
// angular damping), force (3 doubles); fixtures: shape records, material (3 doubles: friction,
// restitution, threshold)
void* ac_new(int nb, const int* btype, const double* mass, const double* force, int nf, const shape_rec::Shape* shapes,
             const double* mat, double gx, double gy) {
  Model* m = new Model();
  m->nb = nb;
  m->nf = nf;
  m->btype.assign(btype, btype + nb);
  m->mass.resize(nb);
  for (int b = 0; b < nb; ++b) {
    const double* p = mass + 8 * b;
    m->mass[b] = TBodyMass<double>{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
  }
  m->force.assign(force, force + 3 * nb);
  m->shape.resize(nf);
  for (int f = 0; f < nf; ++f) m->shape[f] = shape_rec::view(shapes[f]);
  m->mat.assign(mat, mat + 3 * nf);
  m->gx = gx;
  m->gy = gy;
  return m;
}

void ac_free(void* h) { delete static_cast<Model*>(h); }

// one recorded step: slots (10 ints each, b2g_tape_export layout), carried impulses (4 doubles),
// parameters (dt, dt_ratio, velocity iterations, position iterations, warm, block)
void ac_add_step(void* h, int n, const int* slots, const double* lam0, const double* params) {
  Model* m = static_cast<Model*>(h);
  Step s;
  s.n = n;
  s.off = m->total;
  s.dt = params[0];
  s.dt_ratio = params[1];
  s.vel_iters = int(params[2]);
  s.pos_iters = int(params[3]);
  s.warm = params[4] != 0.0;
  s.block = params[5] != 0.0;
  m->steps.push_back(s);
  m->slot.insert(m->slot.end(), slots, slots + 10 * size_t(n));
  m->lam0.insert(m->lam0.end(), lam0, lam0 + 4 * size_t(n));
  m->total += n;
  m->store.assign(4 * size_t(m->total), 0.0);
}

// frozen forward from state0 (6 doubles per body) with fixture frictions fric; returns the
// number of recomputed manifolds whose type, count or features differ from the recording
int ac_forward(void* h, const double* state0, const double* fric, double* final_state) {
  Model* m = static_cast<Model*>(h);
  for (int f = 0; f < m->nf; ++f) m->mat[3 * f] = fric[f];
  m->mismatches = 0;
  m->mismatch_body.assign(m->nb, 0);
  std::vector<double> q, v;
  split(state0, m->nb, q, v);
  StepTape t;
  for (const Step& st : m->steps) forward_step(*m, st, q, v, t);
  join(final_state, m->nb, q, v);
  return m->mismatches;
}

// reverse pass: cot = dL/d(final state) -> dL/d(state0), dL/d(friction)
int ac_backward(void* h, const double* state0, const double* fric, const double* cot, double* grad0,
                double* grad_fric) {
  Model* m = static_cast<Model*>(h);
  for (int f = 0; f < m->nf; ++f) m->mat[3 * f] = fric[f];
  m->mismatches = 0;
  m->mismatch_body.assign(m->nb, 0);
  const int N = int(m->steps.size());
  std::vector<std::vector<double>> qs(N), vs(N);
  std::vector<double> q, v;
  split(state0, m->nb, q, v);
  StepTape t;
  for (int k = 0; k < N; ++k) {
    qs[k] = q;
    vs[k] = v;
    forward_step(*m, m->steps[k], q, v, t);
  }
  std::vector<double> qb, vb, store_bar(4 * size_t(std::max<long long>(m->total, 1)), 0.0), fg(m->nf, 0.0);
  split(cot, m->nb, qb, vb);
  for (int k = N - 1; k >= 0; --k) {
    q = qs[k];
    v = vs[k];
    forward_step(*m, m->steps[k], q, v, t);
    reverse_step(*m, m->steps[k], t, qb, vb, store_bar, fg);
  }
  join(grad0, m->nb, qb, vb);
  for (int f = 0; f < m->nf; ++f) grad_fric[f] = fg[f];
  return m->mismatches;
}

// bodies of the slots whose recomputed manifold differed from the recording in the last
// forward or backward pass (flags: one int per body)
void ac_mismatch_bodies(void* h, int* flags) {
  const Model* m = static_cast<const Model*>(h);
  for (int b = 0; b < m->nb; ++b) flags[b] = b < int(m->mismatch_body.size()) ? m->mismatch_body[b] : 0;
}

}  // extern "C"
