// CPU oracle, part 2 implementation. TEST INFRASTRUCTURE ONLY (see b2o_core.h).
#include "b2o_world.h"
#include "b2o_joints.h"
#include "b2o_toi.h"

#include <chrono>

namespace b2o {

// ---- dynamic tree ------------------------------------------------------------------------
// src/private/collision/b2_dynamic_tree.rs:34-66
int DynamicTree::allocate_node() {
  if (free_list == kNullNode) {
    node_capacity *= 2;
    nodes.resize(node_capacity);
    for (int i = node_count; i < node_capacity - 1; ++i) { nodes[i].parent = i + 1; nodes[i].height = -1; }
    nodes[node_capacity - 1].parent = kNullNode;
    nodes[node_capacity - 1].height = -1;
    free_list = node_count;
  }
  int id = free_list;
  free_list = nodes[id].parent;
  nodes[id].parent = kNullNode;
  nodes[id].child1 = kNullNode;
  nodes[id].child2 = kNullNode;
  nodes[id].height = 0;
  nodes[id].user = -1;
  nodes[id].moved = false;
  ++node_count;
  return id;
}

// src/private/collision/b2_dynamic_tree.rs:69-76
void DynamicTree::free_node(int id) {
  nodes[id].parent = free_list;
  nodes[id].height = -1;
  free_list = id;
  --node_count;
}

// src/private/collision/b2_dynamic_tree.rs:81-99
int DynamicTree::create_proxy(const AABB& aabb, int user) {
  int id = allocate_node();
  Vec2 r(kAabbExtension, kAabbExtension);
  nodes[id].aabb.lower = aabb.lower - r;
  nodes[id].aabb.upper = aabb.upper + r;
  nodes[id].user = user;
  nodes[id].height = 0;
  nodes[id].moved = true;
  insert_leaf(id);
  return id;
}

// src/private/collision/b2_dynamic_tree.rs:109-168
bool DynamicTree::move_proxy(int id, const AABB& aabb, Vec2 displacement) {
  Vec2 r(kAabbExtension, kAabbExtension);
  AABB fat;
  fat.lower = aabb.lower - r;
  fat.upper = aabb.upper + r;
  Vec2 d = kAabbMultiplier * displacement;
  if (d.x < 0) fat.lower.x += d.x; else fat.upper.x += d.x;
  if (d.y < 0) fat.lower.y += d.y; else fat.upper.y += d.y;
  const AABB treeAABB = nodes[id].aabb;
  if (treeAABB.contains(aabb)) {
    AABB huge;
    huge.lower = fat.lower - real(4.0) * r;
    huge.upper = fat.upper + real(4.0) * r;
    if (huge.contains(treeAABB)) return false;
  }
  remove_leaf(id);
  nodes[id].aabb = fat;
  insert_leaf(id);
  nodes[id].moved = true;
  return true;
}

// src/private/collision/b2_dynamic_tree.rs:170-298
void DynamicTree::insert_leaf(int leaf) {
  ++insertion_count;
  if (root == kNullNode) {
    root = leaf;
    nodes[root].parent = kNullNode;
    return;
  }
  AABB leafAABB = nodes[leaf].aabb;
  int index = root;
  while (!nodes[index].is_leaf()) {
    int child1 = nodes[index].child1, child2 = nodes[index].child2;
    real area = nodes[index].aabb.perimeter();
    AABB combined;
    combined.combine(nodes[index].aabb, leafAABB);
    real combinedArea = combined.perimeter();
    real cost = real(2.0) * combinedArea;
    real inheritanceCost = real(2.0) * (combinedArea - area);
    real cost1, cost2;
    if (nodes[child1].is_leaf()) {
      AABB a; a.combine(leafAABB, nodes[child1].aabb);
      cost1 = a.perimeter() + inheritanceCost;
    } else {
      AABB a; a.combine(leafAABB, nodes[child1].aabb);
      real oldArea = nodes[child1].aabb.perimeter();
      real newArea = a.perimeter();
      cost1 = (newArea - oldArea) + inheritanceCost;
    }
    if (nodes[child2].is_leaf()) {
      AABB a; a.combine(leafAABB, nodes[child2].aabb);
      cost2 = a.perimeter() + inheritanceCost;
    } else {
      AABB a; a.combine(leafAABB, nodes[child2].aabb);
      real oldArea = nodes[child2].aabb.perimeter();
      real newArea = a.perimeter();
      cost2 = newArea - oldArea + inheritanceCost;
    }
    if (cost < cost1 && cost < cost2) break;
    index = cost1 < cost2 ? child1 : child2;
  }
  int sibling = index;
  int oldParent = nodes[sibling].parent;
  int newParent = allocate_node();
  nodes[newParent].parent = oldParent;
  nodes[newParent].user = -1;
  nodes[newParent].aabb.combine(leafAABB, nodes[sibling].aabb);
  nodes[newParent].height = nodes[sibling].height + 1;
  if (oldParent != kNullNode) {
    if (nodes[oldParent].child1 == sibling) nodes[oldParent].child1 = newParent;
    else nodes[oldParent].child2 = newParent;
    nodes[newParent].child1 = sibling;
    nodes[newParent].child2 = leaf;
    nodes[sibling].parent = newParent;
    nodes[leaf].parent = newParent;
  } else {
    nodes[newParent].child1 = sibling;
    nodes[newParent].child2 = leaf;
    nodes[sibling].parent = newParent;
    nodes[leaf].parent = newParent;
    root = newParent;
  }
  index = nodes[leaf].parent;
  while (index != kNullNode) {
    index = balance(index);
    int child1 = nodes[index].child1, child2 = nodes[index].child2;
    nodes[index].height = 1 + bmax(nodes[child1].height, nodes[child2].height);
    AABB a1 = nodes[child1].aabb, a2 = nodes[child2].aabb;
    nodes[index].aabb.combine(a1, a2);
    index = nodes[index].parent;
  }
}

// src/private/collision/b2_dynamic_tree.rs:300-353
void DynamicTree::remove_leaf(int leaf) {
  if (leaf == root) { root = kNullNode; return; }
  int parent = nodes[leaf].parent;
  int grand = nodes[parent].parent;
  int sibling = nodes[parent].child1 == leaf ? nodes[parent].child2 : nodes[parent].child1;
  if (grand != kNullNode) {
    if (nodes[grand].child1 == parent) nodes[grand].child1 = sibling;
    else nodes[grand].child2 = sibling;
    nodes[sibling].parent = grand;
    free_node(parent);
    int index = grand;
    while (index != kNullNode) {
      index = balance(index);
      int child1 = nodes[index].child1, child2 = nodes[index].child2;
      AABB a1 = nodes[child1].aabb, a2 = nodes[child2].aabb;
      nodes[index].aabb.combine(a1, a2);
      nodes[index].height = 1 + bmax(nodes[child1].height, nodes[child2].height);
      index = nodes[index].parent;
    }
  } else {
    root = sibling;
    nodes[sibling].parent = kNullNode;
    free_node(parent);
  }
}

// src/private/collision/b2_dynamic_tree.rs:357-489
int DynamicTree::balance(int iA) {
  TreeNode A = nodes[iA];
  if (A.is_leaf() || A.height < 2) return iA;
  int iB = A.child1, iC = A.child2;
  TreeNode B = nodes[iB], C = nodes[iC];
  int bal = C.height - B.height;
  if (bal > 1) {
    int iF = C.child1, iG = C.child2;
    TreeNode F = nodes[iF], G = nodes[iG];
    C.child1 = iA;
    C.parent = A.parent;
    A.parent = iC;
    if (C.parent != kNullNode) {
      if (nodes[C.parent].child1 == iA) nodes[C.parent].child1 = iC;
      else nodes[C.parent].child2 = iC;
    } else {
      root = iC;
    }
    if (F.height > G.height) {
      C.child2 = iF; A.child2 = iG; G.parent = iA;
      A.aabb.combine(B.aabb, G.aabb);
      C.aabb.combine(A.aabb, F.aabb);
      A.height = 1 + bmax(B.height, G.height);
      C.height = 1 + bmax(A.height, F.height);
    } else {
      C.child2 = iG; A.child2 = iF; F.parent = iA;
      A.aabb.combine(B.aabb, F.aabb);
      C.aabb.combine(A.aabb, G.aabb);
      A.height = 1 + bmax(B.height, F.height);
      C.height = 1 + bmax(A.height, G.height);
    }
    nodes[iA] = A; nodes[iC] = C; nodes[iF] = F; nodes[iG] = G;
    return iC;
  }
  if (bal < -1) {
    int iD = B.child1, iE = B.child2;
    TreeNode D = nodes[iD], E = nodes[iE];
    B.child1 = iA;
    B.parent = A.parent;
    A.parent = iB;
    if (B.parent != kNullNode) {
      if (nodes[B.parent].child1 == iA) nodes[B.parent].child1 = iB;
      else nodes[B.parent].child2 = iB;
    } else {
      root = iB;
    }
    if (D.height > E.height) {
      B.child2 = iD; A.child1 = iE; E.parent = iA;
      A.aabb.combine(C.aabb, E.aabb);
      B.aabb.combine(A.aabb, D.aabb);
      A.height = 1 + bmax(C.height, E.height);
      B.height = 1 + bmax(A.height, D.height);
    } else {
      B.child2 = iE; A.child1 = iD; D.parent = iA;
      A.aabb.combine(C.aabb, D.aabb);
      B.aabb.combine(A.aabb, E.aabb);
      A.height = 1 + bmax(C.height, D.height);
      B.height = 1 + bmax(A.height, E.height);
    }
    nodes[iA] = A; nodes[iB] = B; nodes[iE] = E; nodes[iD] = D;
    return iB;
  }
  return iA;
}

// ---- body ---------------------------------------------------------------------------------
// src/b2_body.rs:783-801
void Body::set_awake(bool flag) {
  if (type == kStatic) return;
  if (flag) {
    awake = true;
    sleep_time = 0;
  } else {
    awake = false;
    sleep_time = 0;
    v.set_zero();
    w = 0;
    force.set_zero();
    torque = 0;
  }
}

World::~World() {
  for (Contact* c = contact_list; c;) { Contact* n = c->next; delete c; c = n; }
  for (Joint* j : joints) delete j;
  for (Fixture* f : fixtures) delete f;
  for (Body* b : bodies) delete b;
}

// src/private/dynamics/b2_body.rs:15-89 and src/private/dynamics/b2_world.rs:83-98
Body* World::create_body(int type, Vec2 pos, real angle, Vec2 v, real w, real lin_damp, real ang_damp, real gscale,
                         bool allow_sleep_, bool awake, bool fixed_rotation) {
  Body* b = new Body();
  b->index = int(bodies.size());
  b->type = type;
  b->fixed_rotation = fixed_rotation;
  b->auto_sleep = allow_sleep_;
  b->awake = awake && type != kStatic;
  b->enabled = true;
  b->xf = Transform(pos, Rot(angle));
  b->sweep.local_center.set_zero();
  b->sweep.c0 = b->xf.p;
  b->sweep.c = b->xf.p;
  b->sweep.a0 = angle;
  b->sweep.a = angle;
  b->sweep.alpha0 = 0;
  b->v = v;
  b->w = w;
  b->linear_damping = lin_damp;
  b->angular_damping = ang_damp;
  b->gravity_scale = gscale;
  bodies.push_back(b);
  return b;
}

// src/private/dynamics/b2_body.rs:292-349
void World::reset_mass_data(Body* b) {
  b->mass = 0; b->inv_mass = 0; b->I = 0; b->inv_I = 0;
  b->sweep.local_center.set_zero();
  if (b->type == kStatic || b->type == kKinematic) {
    b->sweep.c0 = b->xf.p;
    b->sweep.c = b->xf.p;
    b->sweep.a0 = b->sweep.a;
    return;
  }
  Vec2 localCenter;
  for (Fixture* f : b->fixtures) {
    if (f->density == 0) continue;
    MassData md;
    compute_mass(f->shape, &md, f->density);
    b->mass += md.mass;
    localCenter += md.mass * md.center;
    b->I += md.I;
  }
  if (b->mass > 0) {
    b->inv_mass = real(1.0) / b->mass;
    localCenter *= b->inv_mass;
  }
  if (b->I > 0 && !b->fixed_rotation) {
    b->I -= b->mass * dot(localCenter, localCenter);
    b->inv_I = real(1.0) / b->I;
  } else {
    b->I = 0;
    b->inv_I = 0;
  }
  Vec2 oldCenter = b->sweep.c;
  b->sweep.local_center = localCenter;
  b->sweep.c0 = mul(b->xf, b->sweep.local_center);
  b->sweep.c = b->sweep.c0;
  b->v += cross(b->w, b->sweep.c - oldCenter);
}

// src/private/dynamics/b2_body.rs:155-200, src/private/dynamics/b2_fixture.rs:114-133
Fixture* World::create_fixture(Body* b, const Shape& s, real density, real friction, real restitution, real threshold,
                               Filter filter, bool sensor) {
  Fixture* f = new Fixture();
  f->index = int(fixtures.size());
  f->body = b;
  f->shape = s;
  f->density = density;
  f->friction = friction;
  f->restitution = restitution;
  f->threshold = threshold;
  f->filter = filter;
  f->sensor = sensor;
  if (b->enabled) {
    compute_aabb(f->shape, &f->aabb, b->xf);
    f->proxy_id = bp.create_proxy(f->aabb, f->index);
  }
  fixtures.push_back(f);
  b->fixtures.insert(b->fixtures.begin(), f);
  if (density > 0) reset_mass_data(b);
  new_contacts = true;
  return f;
}

// ---- joints: src/private/dynamics/b2_world.rs:156-275 ----------------------------------------
Joint* World::create_joint(const Joint& def) {
  Joint* j = new Joint(def);
  j->index = int(joints.size());
  joints.push_back(j);
  // the bodies' joint lists (push_front)
  j->edge_a.joint = j;
  j->edge_a.other = j->b;
  j->edge_a.prev = nullptr;
  j->edge_a.next = j->a->joint_list;
  if (j->a->joint_list) j->a->joint_list->prev = &j->edge_a;
  j->a->joint_list = &j->edge_a;
  j->edge_b.joint = j;
  j->edge_b.other = j->a;
  j->edge_b.prev = nullptr;
  j->edge_b.next = j->b->joint_list;
  if (j->b->joint_list) j->b->joint_list->prev = &j->edge_b;
  j->b->joint_list = &j->edge_b;
  // a joint that prevents collisions flags the contacts between its bodies for filtering
  if (!j->collide_connected)
    for (ContactEdge* e = j->b->contact_list; e; e = e->next)
      if (e->other == j->a) e->contact->filter = true;
  // (creating a joint does not wake the bodies)
  return j;
}

// ---- contact manager ----------------------------------------------------------------------
static bool should_collide_bodies(const Body* b, const Body* other) {
  // src/private/dynamics/b2_body.rs:391-416
  if (b->type != kDynamic && other->type != kDynamic) return false;
  for (const JointEdge* je = b->joint_list; je; je = je->next)
    if (je->other == other && !je->joint->collide_connected) return false;
  return true;
}
// src/private/dynamics/b2_world_callbacks.rs:6-18
static bool should_collide_fixtures(const Fixture* a, const Fixture* b) {
  if (a->filter.group == b->filter.group && a->filter.group != 0) return a->filter.group > 0;
  return (a->filter.mask & b->filter.category) != 0 && (a->filter.category & b->filter.mask) != 0;
}

static void list_push_front(ContactEdge** head, ContactEdge* e) {
  e->prev = nullptr;
  e->next = *head;
  if (*head) (*head)->prev = e;
  *head = e;
}
static void list_remove(ContactEdge** head, ContactEdge* e) {
  if (e->prev) e->prev->next = e->next;
  if (e->next) e->next->prev = e->prev;
  if (e == *head) *head = e->next;
  e->prev = e->next = nullptr;
}

// src/private/dynamics/b2_contact_manager.rs:178-300 with b2_contact_create (b2_contact.rs:11-31)
void World::add_pair(int fia, int fib) {
  Fixture* fa = fixtures[fia];
  Fixture* fb = fixtures[fib];
  Body* ba = fa->body;
  Body* bb = fb->body;
  if (ba == bb) return;
  for (ContactEdge* e = bb->contact_list; e; e = e->next) {
    if (e->other == ba) {
      Fixture* cfa = e->contact->fa;
      Fixture* cfb = e->contact->fb;
      if (cfa == fa && cfb == fb) return;
      if (cfa == fb && cfb == fa) return;
    }
  }
  if (!should_collide_bodies(bb, ba)) return;
  if (!should_collide_fixtures(fa, fb)) return;
  bool swap = false;
  if (!registry_lookup(fa->shape.type, fb->shape.type, &swap)) return;
  if (swap) { Fixture* t = fa; fa = fb; fb = t; }
  Contact* c = new Contact();
  c->fa = fa;
  c->fb = fb;
  c->enabled = true;
  // src/b2_contact.rs:29-46
  c->friction = std::sqrt(fa->friction * fb->friction);
  c->restitution = fa->restitution > fb->restitution ? fa->restitution : fb->restitution;
  c->threshold = fa->threshold < fb->threshold ? fa->threshold : fb->threshold;
  c->tangent_speed = 0;
  ba = fa->body;
  bb = fb->body;
  c->prev = nullptr;
  c->next = contact_list;
  if (contact_list) contact_list->prev = c;
  contact_list = c;
  c->node_a.contact = c;
  c->node_a.other = bb;
  list_push_front(&ba->contact_list, &c->node_a);
  c->node_b.contact = c;
  c->node_b.other = ba;
  list_push_front(&bb->contact_list, &c->node_b);
  ++contact_count;
}

// src/private/dynamics/b2_contact_manager.rs:24-78
void World::destroy_contact(Contact* c) {
  Body* ba = c->fa->body;
  Body* bb = c->fb->body;
  if (c->prev) c->prev->next = c->next;
  if (c->next) c->next->prev = c->prev;
  if (c == contact_list) contact_list = c->next;
  list_remove(&ba->contact_list, &c->node_a);
  list_remove(&bb->contact_list, &c->node_b);
  // b2_contact_destroy: wake both bodies if the contact was touching (src/private/dynamics/b2_contact.rs:33-45)
  if (c->manifold.point_count > 0 && !c->fa->sensor && !c->fb->sensor) {
    ba->set_awake(true);
    bb->set_awake(true);
  }
  delete c;
  --contact_count;
}

// src/private/dynamics/b2_contact_manager.rs:83-171 with B2contact::update (b2_contact.rs:103-217)
void World::collide() {
  std::vector<Contact*> to_destroy;
  for (Contact* c = contact_list; c; c = c->next) {
    Fixture* fa = c->fa;
    Fixture* fb = c->fb;
    Body* ba = fa->body;
    Body* bb = fb->body;
    // flagged for filtering (src/private/dynamics/b2_contact_manager.rs:100-127)
    if (c->filter) {
      if (!should_collide_bodies(bb, ba) || !should_collide_fixtures(fa, fb)) {
        to_destroy.push_back(c);
        continue;
      }
      c->filter = false;
    }
    bool activeA = ba->awake && ba->type != kStatic;
    bool activeB = bb->awake && bb->type != kStatic;
    if (!activeA && !activeB) continue;
    bool overlap = test_overlap(bp.tree.nodes[fa->proxy_id].aabb, bp.tree.nodes[fb->proxy_id].aabb);
    if (!overlap) { to_destroy.push_back(c); continue; }
    update_contact(c);
  }
  for (Contact* c : to_destroy) destroy_contact(c);
}

// b2_test_overlap (src/private/collision/b2_collision.rs:216-245): GJK distance with radii from
// an empty simplex cache, below ten epsilons
static bool test_overlap(const Shape& a, const Transform& xa, const Shape& b, const Transform& xb) {
  DistanceProxy pa, pb;
  pa.set(a);
  pb.set(b);
  SimplexCache cache;
  DistanceOutput out;
  distance(&out, &cache, pa, xa, pb, xb, true);
  return out.distance < real(10.0) * kEpsilon;
}

// B2contact::update (src/private/dynamics/b2_contact.rs:103-217)
void World::update_contact(Contact* c) {
  {
    Fixture* fa = c->fa;
    Fixture* fb = c->fb;
    Body* ba = fa->body;
    Body* bb = fb->body;
    c->enabled = true;
    Manifold old = c->manifold;
    bool was_touching = c->touching;
    bool sensor = fa->sensor || fb->sensor;
    bool touching;
    if (sensor) {
      // sensors get no manifold and wake nobody (src/private/dynamics/b2_contact.rs:148-162)
      touching = test_overlap(fa->shape, ba->xf, fb->shape, bb->xf);
      c->manifold.point_count = 0;
    } else {
      Manifold nm;
      evaluate(&nm, fa->shape, ba->xf, fb->shape, bb->xf);
      c->manifold = nm;
      touching = c->manifold.point_count > 0;
      for (int i = 0; i < c->manifold.point_count; ++i) {
        ManifoldPoint* mp2 = &c->manifold.points[i];
        mp2->normal_impulse = 0;
        mp2->tangent_impulse = 0;
        for (int j = 0; j < old.point_count; ++j) {
          if (old.points[j].id == mp2->id) {
            mp2->normal_impulse = old.points[j].normal_impulse;
            mp2->tangent_impulse = old.points[j].tangent_impulse;
            break;
          }
        }
      }
      if (touching != was_touching) { ba->set_awake(true); bb->set_awake(true); }
    }
    c->touching = touching;
  }
}

void World::find_new_contacts() {
  bp.update_pairs([this](int ua, int ub) { add_pair(ua, ub); });
}

// src/private/dynamics/b2_body.rs:455-475 and src/private/dynamics/b2_fixture.rs:147-173
void World::synchronize_fixtures(Body* b) {
  Transform xf1;
  if (b->awake) {
    xf1.q.set(b->sweep.a0);
    xf1.p = b->sweep.c0 - mul(xf1.q, b->sweep.local_center);
  } else {
    xf1 = b->xf;
  }
  const Transform& xf2 = b->awake ? b->xf : b->xf;
  for (Fixture* f : b->fixtures) {
    if (f->proxy_id < 0) continue;
    AABB a1, a2;
    compute_aabb(f->shape, &a1, xf1);
    compute_aabb(f->shape, &a2, xf2);
    f->aabb.combine(a1, a2);
    Vec2 disp = a2.center() - a1.center();
    bp.move_proxy(f->proxy_id, f->aabb, disp);
  }
}

// ---- contact solver -----------------------------------------------------------------------
// src/private/dynamics/b2_contact_solver_private.rs:20-110
void ContactSolver::init(const std::vector<SolverContact>& contacts, const std::vector<SolverBody>& bodies) {
  size_t n = contacts.size();
  vcs.assign(n, VelocityConstraint());
  pcs.assign(n, PositionConstraint());
  for (size_t i = 0; i < n; ++i) {
    const SolverContact& sc = contacts[i];
    const Manifold& m = *sc.manifold;
    const SolverBody& A = bodies[sc.indexA];
    const SolverBody& B = bodies[sc.indexB];
    VelocityConstraint& vc = vcs[i];
    vc.friction = sc.friction;
    vc.restitution = sc.restitution;
    vc.threshold = sc.threshold;
    vc.tangent_speed = sc.tangent_speed;
    vc.indexA = sc.indexA;
    vc.indexB = sc.indexB;
    vc.invMassA = A.inv_mass;
    vc.invMassB = B.inv_mass;
    vc.invIA = A.inv_I;
    vc.invIB = B.inv_I;
    vc.contact_index = int(i);
    vc.point_count = m.point_count;
    vc.K.set_zero();
    vc.normal_mass.set_zero();
    PositionConstraint& pc = pcs[i];
    pc.indexA = sc.indexA;
    pc.indexB = sc.indexB;
    pc.invMassA = A.inv_mass;
    pc.invMassB = B.inv_mass;
    pc.localCenterA = A.local_center;
    pc.localCenterB = B.local_center;
    pc.invIA = A.inv_I;
    pc.invIB = B.inv_I;
    pc.local_normal = m.local_normal;
    pc.local_point = m.local_point;
    pc.point_count = m.point_count;
    pc.radiusA = sc.radiusA;
    pc.radiusB = sc.radiusB;
    pc.type = m.type;
    for (int j = 0; j < m.point_count; ++j) {
      const ManifoldPoint& cp = m.points[j];
      VelocityConstraintPoint& vcp = vc.points[j];
      if (step.warm_starting) {
        vcp.normal_impulse = step.dt_ratio * cp.normal_impulse;
        vcp.tangent_impulse = step.dt_ratio * cp.tangent_impulse;
      } else {
        vcp.normal_impulse = 0;
        vcp.tangent_impulse = 0;
      }
      vcp.rA.set_zero();
      vcp.rB.set_zero();
      vcp.normal_mass = 0;
      vcp.tangent_mass = 0;
      vcp.velocity_bias = 0;
      pc.local_points[j] = cp.local_point;
    }
  }
}

// src/private/dynamics/b2_contact_solver_private.rs:113-226
void ContactSolver::initialize_velocity_constraints(const std::vector<SolverContact>& contacts, const Position* pos,
                                                    const Velocity* vel) {
  for (size_t i = 0; i < vcs.size(); ++i) {
    VelocityConstraint& vc = vcs[i];
    PositionConstraint& pc = pcs[i];
    real radiusA = pc.radiusA, radiusB = pc.radiusB;
    const Manifold& m = *contacts[vc.contact_index].manifold;
    int iA = vc.indexA, iB = vc.indexB;
    real mA = vc.invMassA, mB = vc.invMassB, IA = vc.invIA, IB = vc.invIB;
    Vec2 lcA = pc.localCenterA, lcB = pc.localCenterB;
    Vec2 cA = pos[iA].c; real aA = pos[iA].a;
    Vec2 vA = vel[iA].v; real wA = vel[iA].w;
    Vec2 cB = pos[iB].c; real aB = pos[iB].a;
    Vec2 vB = vel[iB].v; real wB = vel[iB].w;
    Transform xfA, xfB;
    xfA.q.set(aA);
    xfB.q.set(aB);
    xfA.p = cA - mul(xfA.q, lcA);
    xfB.p = cB - mul(xfB.q, lcB);
    WorldManifold wm;
    world_manifold_initialize(&wm, m, xfA, radiusA, xfB, radiusB);
    vc.normal = wm.normal;
    int pcount = vc.point_count;
    for (int j = 0; j < pcount; ++j) {
      VelocityConstraintPoint& vcp = vc.points[j];
      vcp.rA = wm.points[j] - cA;
      vcp.rB = wm.points[j] - cB;
      real rnA = cross(vcp.rA, vc.normal);
      real rnB = cross(vcp.rB, vc.normal);
      real kNormal = mA + mB + IA * rnA * rnA + IB * rnB * rnB;
      vcp.normal_mass = kNormal > 0 ? real(1.0) / kNormal : real(0);
      Vec2 tangent = cross(vc.normal, real(1.0));
      real rtA = cross(vcp.rA, tangent);
      real rtB = cross(vcp.rB, tangent);
      real kTangent = mA + mB + IA * rtA * rtA + IB * rtB * rtB;
      vcp.tangent_mass = kTangent > 0 ? real(1.0) / kTangent : real(0);
      vcp.velocity_bias = 0;
      real vRel = dot(vc.normal, ((vB + cross(wB, vcp.rB)) - vA) - cross(wA, vcp.rA));
      if (vRel < -vc.threshold) vcp.velocity_bias = -vc.restitution * vRel;
    }
    if (vc.point_count == 2 && block_solve) {
      const VelocityConstraintPoint& p1 = vc.points[0];
      const VelocityConstraintPoint& p2 = vc.points[1];
      real rn1A = cross(p1.rA, vc.normal), rn1B = cross(p1.rB, vc.normal);
      real rn2A = cross(p2.rA, vc.normal), rn2B = cross(p2.rB, vc.normal);
      real k11 = mA + mB + IA * rn1A * rn1A + IB * rn1B * rn1B;
      real k22 = mA + mB + IA * rn2A * rn2A + IB * rn2B * rn2B;
      real k12 = mA + mB + IA * rn1A * rn2A + IB * rn1B * rn2B;
      const real kMaxCond = real(1000.0);
      if (k11 * k11 < kMaxCond * (k11 * k22 - k12 * k12)) {
        vc.K.ex.set(k11, k12);
        vc.K.ey.set(k12, k22);
        vc.normal_mass = vc.K.get_inverse();
      } else {
        vc.point_count = 1;
      }
    }
  }
}

// src/private/dynamics/b2_contact_solver_private.rs:228-266
void ContactSolver::warm_start(Velocity* vel) {
  for (size_t i = 0; i < vcs.size(); ++i) {
    const VelocityConstraint& vc = vcs[i];
    int iA = vc.indexA, iB = vc.indexB;
    real mA = vc.invMassA, IA = vc.invIA, mB = vc.invMassB, IB = vc.invIB;
    Vec2 vA = vel[iA].v; real wA = vel[iA].w;
    Vec2 vB = vel[iB].v; real wB = vel[iB].w;
    Vec2 normal = vc.normal;
    Vec2 tangent = cross(normal, real(1.0));
    for (int j = 0; j < vc.point_count; ++j) {
      const VelocityConstraintPoint& vcp = vc.points[j];
      Vec2 P = vcp.normal_impulse * normal + vcp.tangent_impulse * tangent;
      wA -= IA * cross(vcp.rA, P);
      vA -= mA * P;
      wB += IB * cross(vcp.rB, P);
      vB += mB * P;
    }
    vel[iA].v = vA; vel[iA].w = wA;
    vel[iB].v = vB; vel[iB].w = wB;
  }
}

// src/private/dynamics/b2_contact_solver_private.rs:268-583
void ContactSolver::solve_velocity_constraints(Velocity* vel) {
  for (size_t i = 0; i < vcs.size(); ++i) {
    VelocityConstraint& vc = vcs[i];
    int iA = vc.indexA, iB = vc.indexB;
    real mA = vc.invMassA, IA = vc.invIA, mB = vc.invMassB, IB = vc.invIB;
    int pcount = vc.point_count;
    Vec2 vA = vel[iA].v; real wA = vel[iA].w;
    Vec2 vB = vel[iB].v; real wB = vel[iB].w;
    Vec2 normal = vc.normal;
    Vec2 tangent = cross(normal, real(1.0));
    real friction = vc.friction;
    for (int j = 0; j < pcount; ++j) {
      VelocityConstraintPoint& vcp = vc.points[j];
      Vec2 dv = ((vB + cross(wB, vcp.rB)) - vA) - cross(wA, vcp.rA);
      real vt = dot(dv, tangent) - vc.tangent_speed;
      real lambda = vcp.tangent_mass * (-vt);
      real maxFriction = friction * vcp.normal_impulse;
      real newImpulse = bclamp(vcp.tangent_impulse + lambda, -maxFriction, maxFriction);
      lambda = newImpulse - vcp.tangent_impulse;
      vcp.tangent_impulse = newImpulse;
      Vec2 P = lambda * tangent;
      vA -= mA * P;
      wA -= IA * cross(vcp.rA, P);
      vB += mB * P;
      wB += IB * cross(vcp.rB, P);
    }
    if (pcount == 1 || !block_solve) {
      for (int j = 0; j < pcount; ++j) {
        VelocityConstraintPoint& vcp = vc.points[j];
        Vec2 dv = ((vB + cross(wB, vcp.rB)) - vA) - cross(wA, vcp.rA);
        real vn = dot(dv, normal);
        real lambda = -vcp.normal_mass * (vn - vcp.velocity_bias);
        real newImpulse = bmax(vcp.normal_impulse + lambda, real(0));
        lambda = newImpulse - vcp.normal_impulse;
        vcp.normal_impulse = newImpulse;
        Vec2 P = lambda * normal;
        vA -= mA * P;
        wA -= IA * cross(vcp.rA, P);
        vB += mB * P;
        wB += IB * cross(vcp.rB, P);
      }
    } else {
      VelocityConstraintPoint& cp1 = vc.points[0];
      VelocityConstraintPoint& cp2 = vc.points[1];
      Vec2 a(cp1.normal_impulse, cp2.normal_impulse);
      Vec2 dv1 = ((vB + cross(wB, cp1.rB)) - vA) - cross(wA, cp1.rA);
      Vec2 dv2 = ((vB + cross(wB, cp2.rB)) - vA) - cross(wA, cp2.rA);
      real vn1 = dot(dv1, normal);
      real vn2 = dot(dv2, normal);
      Vec2 b(vn1 - cp1.velocity_bias, vn2 - cp2.velocity_bias);
      b -= mul(vc.K, a);
      for (;;) {
        // Case 1: vn = 0
        Vec2 x = -mul(vc.normal_mass, b);
        if (x.x >= 0 && x.y >= 0) {
          Vec2 d = x - a;
          Vec2 P1 = d.x * normal, P2 = d.y * normal;
          vA -= mA * (P1 + P2);
          wA -= IA * (cross(cp1.rA, P1) + cross(cp2.rA, P2));
          vB += mB * (P1 + P2);
          wB += IB * (cross(cp1.rB, P1) + cross(cp2.rB, P2));
          cp1.normal_impulse = x.x;
          cp2.normal_impulse = x.y;
          break;
        }
        // Case 2: vn1 = 0 and x2 = 0
        x.x = -cp1.normal_mass * b.x;
        x.y = 0;
        vn1 = 0;
        vn2 = vc.K.ex.y * x.x + b.y;
        if (x.x >= 0 && vn2 >= 0) {
          Vec2 d = x - a;
          Vec2 P1 = d.x * normal, P2 = d.y * normal;
          vA -= mA * (P1 + P2);
          wA -= IA * (cross(cp1.rA, P1) + cross(cp2.rA, P2));
          vB += mB * (P1 + P2);
          wB += IB * (cross(cp1.rB, P1) + cross(cp2.rB, P2));
          cp1.normal_impulse = x.x;
          cp2.normal_impulse = x.y;
          break;
        }
        // Case 3: vn2 = 0 and x1 = 0
        x.x = 0;
        x.y = -cp2.normal_mass * b.y;
        vn1 = vc.K.ey.x * x.y + b.x;
        vn2 = 0;
        if (x.y >= 0 && vn1 >= 0) {
          Vec2 d = x - a;
          Vec2 P1 = d.x * normal, P2 = d.y * normal;
          vA -= mA * (P1 + P2);
          wA -= IA * (cross(cp1.rA, P1) + cross(cp2.rA, P2));
          vB += mB * (P1 + P2);
          wB += IB * (cross(cp1.rB, P1) + cross(cp2.rB, P2));
          cp1.normal_impulse = x.x;
          cp2.normal_impulse = x.y;
          break;
        }
        // Case 4: x1 = 0 and x2 = 0
        x.x = 0;
        x.y = 0;
        vn1 = b.x;
        vn2 = b.y;
        if (vn1 >= 0 && vn2 >= 0) {
          Vec2 d = x - a;
          Vec2 P1 = d.x * normal, P2 = d.y * normal;
          vA -= mA * (P1 + P2);
          wA -= IA * (cross(cp1.rA, P1) + cross(cp2.rA, P2));
          vB += mB * (P1 + P2);
          wB += IB * (cross(cp1.rB, P1) + cross(cp2.rB, P2));
          cp1.normal_impulse = x.x;
          cp2.normal_impulse = x.y;
          break;
        }
        // No solution, give up.
        break;
      }
    }
    vel[iA].v = vA; vel[iA].w = wA;
    vel[iB].v = vB; vel[iB].w = wB;
  }
}

// src/private/dynamics/b2_contact_solver_private.rs:585-598
void ContactSolver::store_impulses(const std::vector<SolverContact>& contacts) {
  for (const VelocityConstraint& vc : vcs) {
    Manifold* m = contacts[vc.contact_index].manifold;
    for (int j = 0; j < vc.point_count; ++j) {
      m->points[j].normal_impulse = vc.points[j].normal_impulse;
      m->points[j].tangent_impulse = vc.points[j].tangent_impulse;
    }
  }
}

// src/private/dynamics/b2_contact_solver_private.rs:600-642
struct PositionSolverManifold {
  Vec2 normal, point;
  real separation = 0;
  void initialize(const PositionConstraint& pc, const Transform& xfA, const Transform& xfB, int index) {
    if (pc.type == kCircles) {
      Vec2 pA = mul(xfA, pc.local_point);
      Vec2 pB = mul(xfB, pc.local_points[0]);
      normal = pB - pA;
      normal.normalize();
      point = real(0.5) * (pA + pB);
      separation = dot(pB - pA, normal) - pc.radiusA - pc.radiusB;
    } else if (pc.type == kFaceA) {
      normal = mul(xfA.q, pc.local_normal);
      Vec2 plane = mul(xfA, pc.local_point);
      Vec2 clip = mul(xfB, pc.local_points[index]);
      separation = dot(clip - plane, normal) - pc.radiusA - pc.radiusB;
      point = clip;
    } else {
      normal = mul(xfB.q, pc.local_normal);
      Vec2 plane = mul(xfB, pc.local_point);
      Vec2 clip = mul(xfA, pc.local_points[index]);
      separation = dot(clip - plane, normal) - pc.radiusA - pc.radiusB;
      point = clip;
      normal = -normal;
    }
  }
};

// src/private/dynamics/b2_contact_solver_private.rs:653-730
bool ContactSolver::solve_position_constraints(Position* pos, real* min_sep_out) {
  real minSeparation = 0;
  for (size_t i = 0; i < pcs.size(); ++i) {
    const PositionConstraint& pc = pcs[i];
    int iA = pc.indexA, iB = pc.indexB;
    Vec2 lcA = pc.localCenterA, lcB = pc.localCenterB;
    real mA = pc.invMassA, IA = pc.invIA, mB = pc.invMassB, IB = pc.invIB;
    Vec2 cA = pos[iA].c; real aA = pos[iA].a;
    Vec2 cB = pos[iB].c; real aB = pos[iB].a;
    for (int j = 0; j < pc.point_count; ++j) {
      Transform xfA, xfB;
      xfA.q.set(aA);
      xfB.q.set(aB);
      xfA.p = cA - mul(xfA.q, lcA);
      xfB.p = cB - mul(xfB.q, lcB);
      PositionSolverManifold psm;
      psm.initialize(pc, xfA, xfB, j);
      Vec2 normal = psm.normal, point = psm.point;
      real separation = psm.separation;
      Vec2 rA = point - cA, rB = point - cB;
      minSeparation = bmin(minSeparation, separation);
      real C = bclamp(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, real(0));
      real rnA = cross(rA, normal), rnB = cross(rB, normal);
      real K = mA + mB + IA * rnA * rnA + IB * rnB * rnB;
      real impulse = K > 0 ? -C / K : real(0);
      Vec2 P = impulse * normal;
      cA -= mA * P;
      aA -= IA * cross(rA, P);
      cB += mB * P;
      aB += IB * cross(rB, P);
    }
    pos[iA].c = cA; pos[iA].a = aA;
    pos[iB].c = cB; pos[iB].a = aB;
  }
  if (min_sep_out) *min_sep_out = minSeparation;
  return minSeparation >= real(-3.0) * kLinearSlop;
}

// src/private/dynamics/b2_contact_solver_private.rs:733-822: only the two TOI bodies move
bool ContactSolver::solve_toi_position_constraints(Position* pos, int toi_index_a, int toi_index_b) {
  real minSeparation = 0;
  for (size_t i = 0; i < pcs.size(); ++i) {
    const PositionConstraint& pc = pcs[i];
    int iA = pc.indexA, iB = pc.indexB;
    Vec2 lcA = pc.localCenterA, lcB = pc.localCenterB;
    real mA = 0, IA = 0, mB = 0, IB = 0;
    if (iA == toi_index_a || iA == toi_index_b) { mA = pc.invMassA; IA = pc.invIA; }
    if (iB == toi_index_a || iB == toi_index_b) { mB = pc.invMassB; IB = pc.invIB; }
    Vec2 cA = pos[iA].c; real aA = pos[iA].a;
    Vec2 cB = pos[iB].c; real aB = pos[iB].a;
    for (int j = 0; j < pc.point_count; ++j) {
      Transform xfA, xfB;
      xfA.q.set(aA);
      xfB.q.set(aB);
      xfA.p = cA - mul(xfA.q, lcA);
      xfB.p = cB - mul(xfB.q, lcB);
      PositionSolverManifold psm;
      psm.initialize(pc, xfA, xfB, j);
      Vec2 normal = psm.normal, point = psm.point;
      real separation = psm.separation;
      Vec2 rA = point - cA, rB = point - cB;
      minSeparation = bmin(minSeparation, separation);
      real C = bclamp(kToiBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, real(0));
      real rnA = cross(rA, normal), rnB = cross(rB, normal);
      real K = mA + mB + IA * rnA * rnA + IB * rnB * rnB;
      real impulse = K > 0 ? -C / K : real(0);
      Vec2 P = impulse * normal;
      cA -= mA * P;
      aA -= IA * cross(rA, P);
      cB += mB * P;
      aB += IB * cross(rB, P);
    }
    pos[iA].c = cA; pos[iA].a = aA;
    pos[iB].c = cB; pos[iB].a = aB;
  }
  return minSeparation >= real(-1.5) * kLinearSlop;
}

void Body::advance(real alpha) {
  sweep_advance(&sweep, alpha);
  sweep.c = sweep.c0;
  sweep.a = sweep.a0;
  xf.q.set(sweep.a);
  xf.p = sweep.c - mul(xf.q, sweep.local_center);
}

// B2island::solve_toi (src/private/dynamics/b2_island_private.rs:330-458)
static void solve_toi_island(World* w, IslandIO& island, const TimeStep& sub, int toi_a, int toi_b) {
  size_t nb = island.bodies.size();
  std::vector<Position> positions(nb);
  std::vector<Velocity> velocities(nb);
  std::vector<SolverBody> sbodies(nb);
  for (size_t i = 0; i < nb; ++i) {
    Body* b = island.bodies[i];
    positions[i].c = b->sweep.c;
    positions[i].a = b->sweep.a;
    velocities[i].v = b->v;
    velocities[i].w = b->w;
    sbodies[i].inv_mass = b->inv_mass;
    sbodies[i].inv_I = b->inv_I;
    sbodies[i].local_center = b->sweep.local_center;
  }
  std::vector<SolverContact> scs(island.contacts.size());
  for (size_t i = 0; i < island.contacts.size(); ++i) {
    Contact* c = island.contacts[i];
    scs[i] = SolverContact{&c->manifold, c->fa->body->island_index, c->fb->body->island_index, c->fa->shape.radius,
                           c->fb->shape.radius, c->friction, c->restitution, c->threshold, c->tangent_speed};
  }
  ContactSolver cs;
  cs.step = sub;
  cs.block_solve = w->block_solve;
  cs.init(scs, sbodies);
  for (int i = 0; i < sub.position_iterations; ++i)
    if (cs.solve_toi_position_constraints(positions.data(), toi_a, toi_b)) break;
  island.bodies[toi_a]->sweep.c0 = positions[toi_a].c;
  island.bodies[toi_a]->sweep.a0 = positions[toi_a].a;
  island.bodies[toi_b]->sweep.c0 = positions[toi_b].c;
  island.bodies[toi_b]->sweep.a0 = positions[toi_b].a;
  cs.initialize_velocity_constraints(scs, positions.data(), velocities.data());
  for (int i = 0; i < sub.velocity_iterations; ++i) cs.solve_velocity_constraints(velocities.data());
  const real h = sub.dt;
  for (size_t i = 0; i < nb; ++i) {
    Vec2 c = positions[i].c;
    real a = positions[i].a;
    Vec2 v = velocities[i].v;
    real wv = velocities[i].w;
    Vec2 translation = h * v;
    if (dot(translation, translation) > kMaxTranslationSquared) {
      real ratio = kMaxTranslation / translation.length();
      v *= ratio;
    }
    real rotation = h * wv;
    if (rotation * rotation > kMaxRotationSquared) {
      real ratio = kMaxRotation / babs(rotation);
      wv *= ratio;
    }
    c += h * v;
    a += h * wv;
    Body* b = island.bodies[i];
    b->sweep.c = c;
    b->sweep.a = a;
    b->v = v;
    b->w = wv;
    b->synchronize_transform();
  }
}

// src/private/dynamics/b2_world.rs:534-901 (no sub-stepping): a contact gets a time of impact when one of
// its bodies is static, kinematic or a bullet
void World::solve_toi(const TimeStep& step) {
  toi_events = 0;
  if (step_complete) {
    for (Body* b : bodies) {
      b->island_flag = false;
      b->sweep.alpha0 = 0;
    }
    for (Contact* c = contact_list; c; c = c->next) {
      c->toi_flag = false;
      c->island_flag = false;
      c->toi_count = 0;
      c->toi = 1;
    }
  }
  IslandIO island;
  for (;;) {
    Contact* min_contact = nullptr;
    real min_alpha = 1;
    for (Contact* c = contact_list; c; c = c->next) {
      if (!c->enabled) continue;
      if (c->toi_count > kMaxSubSteps) continue;
      real alpha;
      if (c->toi_flag) {
        alpha = c->toi;
      } else {
        Fixture* fa = c->fa;
        Fixture* fb = c->fb;
        if (fa->sensor || fb->sensor) continue;
        Body* ba = fa->body;
        Body* bb = fb->body;
        bool activeA = ba->awake && ba->type != kStatic;
        bool activeB = bb->awake && bb->type != kStatic;
        if (!activeA && !activeB) continue;
        bool collideA = ba->bullet || ba->type != kDynamic;
        bool collideB = bb->bullet || bb->type != kDynamic;
        if (!collideA && !collideB) continue;
        real alpha0 = ba->sweep.alpha0;
        if (ba->sweep.alpha0 < bb->sweep.alpha0) {
          alpha0 = bb->sweep.alpha0;
          sweep_advance(&ba->sweep, alpha0);
        } else if (bb->sweep.alpha0 < ba->sweep.alpha0) {
          alpha0 = ba->sweep.alpha0;
          sweep_advance(&bb->sweep, alpha0);
        }
        DistanceProxy pa, pb;
        pa.set(fa->shape);
        pb.set(fb->shape);
        ToiOutput out;
        time_of_impact(&out, pa, pb, ba->sweep, bb->sweep, real(1.0));
        real beta = out.t;
        if (out.state == kToiTouching) alpha = bmin(alpha0 + (real(1.0) - alpha0) * beta, real(1.0));
        else alpha = 1;
        c->toi = alpha;
        c->toi_flag = true;
      }
      if (alpha < min_alpha) {
        min_contact = c;
        min_alpha = alpha;
      }
    }
    if (!min_contact || real(1.0) - real(10.0) * kEpsilon < min_alpha) {
      step_complete = true;
      break;
    }

    Body* ba = min_contact->fa->body;
    Body* bb = min_contact->fb->body;
    Sweep backup1 = ba->sweep, backup2 = bb->sweep;
    ba->advance(min_alpha);
    bb->advance(min_alpha);
    update_contact(min_contact);
    min_contact->toi_flag = false;
    ++min_contact->toi_count;
    if (!min_contact->enabled || !min_contact->touching) {
      min_contact->enabled = false;
      ba->sweep = backup1;
      bb->sweep = backup2;
      ba->synchronize_transform();
      bb->synchronize_transform();
      continue;
    }
    ++toi_events;
    ba->set_awake(true);
    bb->set_awake(true);
    island.bodies.clear();
    island.contacts.clear();
    island.joints.clear();
    ba->island_index = 0;
    island.bodies.push_back(ba);
    bb->island_index = 1;
    island.bodies.push_back(bb);
    island.contacts.push_back(min_contact);
    ba->island_flag = true;
    bb->island_flag = true;
    min_contact->island_flag = true;
    Body* pair[2] = {ba, bb};
    for (Body* body : pair) {
      if (body->type != kDynamic) continue;
      for (ContactEdge* ce = body->contact_list; ce; ce = ce->next) {
        Contact* contact = ce->contact;
        if (contact->island_flag) continue;
        Body* other = ce->other;
        // only static, kinematic or bullet bodies join (src/private/dynamics/b2_world.rs:770-777)
        if (other->type == kDynamic && !body->bullet && !other->bullet) continue;
        if (contact->fa->sensor || contact->fb->sensor) continue;
        Sweep backup = other->sweep;
        if (!other->island_flag) other->advance(min_alpha);
        update_contact(contact);
        if (!contact->enabled || !contact->touching) {
          other->sweep = backup;
          other->synchronize_transform();
          continue;
        }
        contact->island_flag = true;
        island.contacts.push_back(contact);
        if (other->island_flag) continue;
        other->island_flag = true;
        if (other->type != kStatic) other->set_awake(true);
        other->island_index = int(island.bodies.size());
        island.bodies.push_back(other);
      }
    }
    TimeStep sub;
    sub.dt = (real(1.0) - min_alpha) * step.dt;
    sub.inv_dt = real(1.0) / sub.dt;
    sub.dt_ratio = 1;
    sub.position_iterations = 20;
    sub.velocity_iterations = step.velocity_iterations;
    sub.warm_starting = false;
    solve_toi_island(this, island, sub, 0, 1);
    for (Body* body : island.bodies) {
      body->island_flag = false;
      if (body->type != kDynamic) continue;
      synchronize_fixtures(body);
      for (ContactEdge* ce = body->contact_list; ce; ce = ce->next) {
        ce->contact->toi_flag = false;
        ce->contact->island_flag = false;
      }
    }
    find_new_contacts();
    if (sub_stepping) {
      step_complete = false;
      break;
    }
  }
}

// ---- island solve: src/private/dynamics/b2_island_private.rs:129-328 ---------------------
bool solve_island(World* w, IslandIO& island, const TimeStep& step) {
  const real h = step.dt;
  size_t nb = island.bodies.size();
  std::vector<Position> positions(nb);
  std::vector<Velocity> velocities(nb);
  std::vector<SolverBody> sbodies(nb);
  for (size_t i = 0; i < nb; ++i) {
    Body* b = island.bodies[i];
    Vec2 c = b->sweep.c;
    real a = b->sweep.a;
    Vec2 v = b->v;
    real wv = b->w;
    b->sweep.c0 = b->sweep.c;
    b->sweep.a0 = b->sweep.a;
    if (b->type == kDynamic) {
      v += (h * b->inv_mass) * ((b->gravity_scale * b->mass) * w->gravity + b->force);
      wv += h * b->inv_I * b->torque;
      v *= real(1.0) / (real(1.0) + h * b->linear_damping);
      wv *= real(1.0) / (real(1.0) + h * b->angular_damping);
    }
    positions[i].c = c;
    positions[i].a = a;
    velocities[i].v = v;
    velocities[i].w = wv;
    sbodies[i].inv_mass = b->inv_mass;
    sbodies[i].inv_I = b->inv_I;
    sbodies[i].local_center = b->sweep.local_center;
  }
  std::vector<SolverContact> scs(island.contacts.size());
  for (size_t i = 0; i < island.contacts.size(); ++i) {
    Contact* c = island.contacts[i];
    scs[i] = SolverContact{&c->manifold, c->fa->body->island_index, c->fb->body->island_index, c->fa->shape.radius,
                           c->fb->shape.radius, c->friction, c->restitution, c->threshold, c->tangent_speed};
  }
  ContactSolver cs;
  cs.step = step;
  cs.block_solve = w->block_solve;
  cs.init(scs, sbodies);
  cs.initialize_velocity_constraints(scs, positions.data(), velocities.data());
  if (step.warm_starting) cs.warm_start(velocities.data());
  const JointSolverData jd{step, positions.data(), velocities.data()};
  for (Joint* j : island.joints) joint_init_velocity(j, jd);
  for (int i = 0; i < step.velocity_iterations; ++i) {
    for (Joint* j : island.joints) joint_solve_velocity(j, jd);
    cs.solve_velocity_constraints(velocities.data());
  }
  cs.store_impulses(scs);
  for (size_t i = 0; i < nb; ++i) {
    Vec2 c = positions[i].c;
    real a = positions[i].a;
    Vec2 v = velocities[i].v;
    real wv = velocities[i].w;
    Vec2 translation = h * v;
    if (dot(translation, translation) > kMaxTranslationSquared) {
      real ratio = kMaxTranslation / translation.length();
      v *= ratio;
    }
    real rotation = h * wv;
    if (rotation * rotation > kMaxRotationSquared) {
      real ratio = kMaxRotation / babs(rotation);
      wv *= ratio;
    }
    c += h * v;
    a += h * wv;
    positions[i].c = c;
    positions[i].a = a;
    velocities[i].v = v;
    velocities[i].w = wv;
  }
  bool position_solved = false;
  for (int i = 0; i < step.position_iterations; ++i) {
    bool contacts_okay = cs.solve_position_constraints(positions.data());
    bool joints_okay = true;
    for (Joint* j : island.joints) {
      const bool joint_okay = joint_solve_position(j, jd);
      joints_okay = joints_okay && joint_okay;
    }
    if (contacts_okay && joints_okay) { position_solved = true; break; }
  }
  for (size_t i = 0; i < nb; ++i) {
    Body* b = island.bodies[i];
    b->sweep.c = positions[i].c;
    b->sweep.a = positions[i].a;
    b->v = velocities[i].v;
    b->w = velocities[i].w;
    b->synchronize_transform();
  }
  if (w->allow_sleep) {
    real minSleep = kMaxFloat;
    const real linTolSqr = kLinearSleepTolerance * kLinearSleepTolerance;
    const real angTolSqr = kAngularSleepTolerance * kAngularSleepTolerance;
    for (Body* b : island.bodies) {
      if (b->type == kStatic) continue;
      if (!b->auto_sleep || b->w * b->w > angTolSqr || dot(b->v, b->v) > linTolSqr) {
        b->sleep_time = 0;
        minSleep = 0;
      } else {
        b->sleep_time += h;
        minSleep = bmin(minSleep, b->sleep_time);
      }
    }
    if (minSleep >= kTimeToSleep && position_solved)
      for (Body* b : island.bodies) b->set_awake(false);
  }
  return position_solved;
}

// ---- world solve: src/private/dynamics/b2_world.rs:356-531 -------------------------------
void World::solve(const TimeStep& step) {
  for (Body* b : bodies) b->island_flag = false;
  for (Contact* c = contact_list; c; c = c->next) c->island_flag = false;
  for (Joint* j : joints) j->island_flag = false;
  IslandIO island;
  std::vector<Body*> stack;
  stack.reserve(bodies.size());
  for (auto it = bodies.rbegin(); it != bodies.rend(); ++it) {
    Body* seed = *it;
    if (seed->island_flag) continue;
    if (!seed->awake || !seed->enabled) continue;
    if (seed->type == kStatic) continue;
    island.bodies.clear();
    island.contacts.clear();
    island.joints.clear();
    stack.clear();
    stack.push_back(seed);
    seed->island_flag = true;
    while (!stack.empty()) {
      Body* b = stack.back();
      stack.pop_back();
      b->island_index = int(island.bodies.size());
      island.bodies.push_back(b);
      if (b->type == kStatic) continue;
      b->awake = true;
      for (ContactEdge* ce = b->contact_list; ce; ce = ce->next) {
        Contact* c = ce->contact;
        if (c->island_flag) continue;
        if (!c->enabled || !c->touching) continue;
        if (c->fa->sensor || c->fb->sensor) continue;
        island.contacts.push_back(c);
        c->island_flag = true;
        Body* other = ce->other;
        if (other->island_flag) continue;
        stack.push_back(other);
        other->island_flag = true;
      }
      for (JointEdge* je = b->joint_list; je; je = je->next) {
        Joint* j = je->joint;
        if (j->island_flag) continue;
        Body* other = je->other;
        if (!other->enabled) continue;
        island.joints.push_back(j);
        j->island_flag = true;
        if (other->island_flag) continue;
        stack.push_back(other);
        other->island_flag = true;
      }
    }
    solve_island(this, island, step);
    for (Body* b : island.bodies)
      if (b->type == kStatic) b->island_flag = false;
  }
  auto t0 = std::chrono::steady_clock::now();
  for (auto it = bodies.rbegin(); it != bodies.rend(); ++it) {
    Body* b = *it;
    if (!b->island_flag) continue;
    if (b->type == kStatic) continue;
    synchronize_fixtures(b);
  }
  find_new_contacts();
  profile.broadphase = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// src/private/dynamics/b2_world.rs:903-959 (solve_toi when continuous physics is on; sub-stepping is off)
void World::step(real dt, int vel_iters, int pos_iters) {
  auto t0 = std::chrono::steady_clock::now();
  if (new_contacts) {
    find_new_contacts();
    new_contacts = false;
  }
  TimeStep st;
  st.dt = dt;
  st.velocity_iterations = vel_iters;
  st.position_iterations = pos_iters;
  st.inv_dt = dt > 0 ? real(1.0) / dt : real(0);
  st.dt_ratio = inv_dt0 * dt;
  st.warm_starting = warm_starting;
  auto t1 = std::chrono::steady_clock::now();
  collide();
  auto t2 = std::chrono::steady_clock::now();
  profile.collide = std::chrono::duration<double, std::milli>(t2 - t1).count();
  if (step_complete && st.dt > 0) solve(st);
  auto t3 = std::chrono::steady_clock::now();
  profile.solve = std::chrono::duration<double, std::milli>(t3 - t2).count();
  if (continuous && st.dt > 0) solve_toi(st);
  if (st.dt > 0) inv_dt0 = st.inv_dt;
  for (Body* b : bodies) { b->force.set_zero(); b->torque = 0; }
  profile.step = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace b2o
