// oracle_capi.cpp -- TEST INFRASTRUCTURE ONLY.
//
// C entry points of the CPU oracle for ctypes: load one scene of an
// avbd_world, step it (FP32 or FP64, natural or colour-major visit order),
// save it back, and expose the unit-level functions (box-box collision, LDLT,
// FEM vertex terms, SVD, rotation helpers) so they can be pinned against
// closed forms and FP64 references.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

#include "avbd_oracle.hpp"

using namespace orc;

namespace {

struct Handle {
    bool fp64 = false;
    World<float> wf;
    World<double> wd;
};

// Solver::RemoveBody (solver.cpp:104-116) for the bodies at the listed
// indices (indices before the call), in the listed order.
// Mesh::isDragged / addedDragVelocity of the listed bodies (velocity: 3 per body).
template <class R>
int set_drag(World<R>& w, const int32_t* idx, const float* vel, const int32_t* dragged, int n) {
    for (int k = 0; k < n; ++k)
        if (idx[k] < 0 || idx[k] >= int(w.bodies.size())) return -1;
    for (int k = 0; k < n; ++k) {
        Body<R>* b = w.bodies[idx[k]].get();
        b->isDragged = dragged[k] != 0;
        b->addedDragVelocity = V3<R>(R(vel[3 * k]), R(vel[3 * k + 1]), R(vel[3 * k + 2]));
    }
    return 0;
}
template <class R>
int remove_listed(World<R>& w, const int32_t* idx, int n) {
    std::vector<Body<R>*> victims;
    for (int k = 0; k < n; ++k) {
        if (idx[k] < 0 || idx[k] >= int(w.bodies.size())) return -1;
        victims.push_back(w.bodies[idx[k]].get());
    }
    for (Body<R>* b : victims) w.remove_body(b);
    return 0;
}

template <class R>
V3<R> v3(const float* p) { return V3<R>(R(p[0]), R(p[1]), R(p[2])); }
template <class R>
Q<R> q4(const float* p) { return Q<R>(R(p[0]), R(p[1]), R(p[2]), R(p[3])); }
template <class R>
void put3(float* o, const V3<R>& v) { o[0] = float(v.x); o[1] = float(v.y); o[2] = float(v.z); }
template <class R>
void put4(float* o, const Q<R>& q) { o[0] = float(q.w); o[1] = float(q.x); o[2] = float(q.y); o[3] = float(q.z); }

template <class R>
int load(World<R>& W, const avbd_world* w, int scene) {
    W.bodies.clear();
    W.forces.clear();
    W.energies.clear();
    if (scene < 0 || scene >= w->n_scenes) return -1;
    const avbd_scene& sc = w->scenes[scene];
    W.dt = R(sc.dt);
    W.iterations = sc.iterations;
    W.alpha = R(sc.alpha);
    W.beta = R(sc.beta);
    W.gamma = R(sc.gamma);
    W.onPenetrationPenalty = R(sc.on_penetration_penalty);
    W.postStabilization = sc.post_stabilization != 0;
    W.projectionMode = sc.projection_mode;
    W.gravity = v3<R>(sc.gravity);
    W.boundsMin = v3<R>(sc.bounds_min);
    W.boundsMax = v3<R>(sc.bounds_max);
    W.emergencyStop = sc.emergency_stop != 0;
    W.nRemoved = sc.n_removed;
    W.options = sc.options;
    W.trustRegionRho = R(sc.trust_region_rho);
    const bool sphereCollision = (sc.options & AVBD_SCENE_SPHERE_COLLISION) != 0;
    std::vector<int> local(w->n_bodies, -1);
    for (int i = 0; i < w->n_bodies; ++i) {
        const avbd_body& b = w->bodies[i];
        if (b.scene != scene) continue;
        auto B = std::make_unique<Body<R>>();
        B->index = int(W.bodies.size());
        local[i] = B->index;
        B->pos = v3<R>(b.pos);
        B->rot = q4<R>(b.rot);
        B->scale = v3<R>(b.scale);
        B->vel = v3<R>(b.vel);
        B->angVel = v3<R>(b.ang_vel);
        B->prevVel = v3<R>(b.prev_vel);
        B->prevAngVel = v3<R>(b.prev_ang_vel);
        B->mass = R(b.mass);
        B->inertia = v3<R>(b.inertia);
        B->friction = R(b.friction);
        B->radius = R(b.radius);
        B->isStatic = (b.flags & AVBD_BODY_STATIC) != 0;
        B->isParticle = (b.flags & AVBD_BODY_PARTICLE) != 0;
        B->isSphere = (b.flags & AVBD_BODY_SPHERE) != 0;
        B->collidesAsSphere = B->isSphere && sphereCollision;
        W.bodies.push_back(std::move(B));
    }
    auto body = [&](int g) -> Body<R>* {
        if (g < 0 || g >= w->n_bodies || local[g] < 0) return nullptr;
        return W.bodies[local[g]].get();
    };
    // forces in serial order: (serial, kind, index)
    struct Ref { int64_t serial; int kind; int idx; };
    std::vector<Ref> refs;
    for (int i = 0; i < w->n_joints; ++i)
        if (body(w->joints[i].body_b)) refs.push_back({w->joints[i].serial, 0, i});
    for (int i = 0; i < w->n_springs; ++i)
        if (body(w->springs[i].body_b)) refs.push_back({w->springs[i].serial, 1, i});
    for (int i = 0; i < w->n_manifolds; ++i)
        if (body(w->manifolds[i].body_b)) refs.push_back({w->manifolds[i].serial, 2, i});
    std::stable_sort(refs.begin(), refs.end(), [](const Ref& a, const Ref& b) { return a.serial < b.serial; });
    int64_t maxSerial = -1;
    for (const Ref& r : refs) {
        maxSerial = std::max(maxSerial, r.serial);
        if (r.kind == 0) {
            const avbd_joint& j = w->joints[r.idx];
            auto J = std::make_unique<Joint<R>>(body(j.body_a), v3<R>(j.r_a), body(j.body_b), v3<R>(j.r_b));
            J->serial = j.serial;
            J->restRot = q4<R>(j.rest_rot);
            J->torqueArm = R(j.torque_arm);
            J->disabled = (j.flags & AVBD_FORCE_DISABLED) != 0;
            for (int k = 0; k < 6; ++k) {
                Row<R>& row = J->rows[k];
                row.stiffness = R(j.stiffness[k]);
                row.fracture = R(j.fracture[k]);
                row.fmin = R(j.fmin[k]);
                row.fmax = R(j.fmax[k]);
                row.lambda = R(j.lambda[k]);
                row.penalty = R(j.penalty[k]);
            }
            W.add_force(std::move(J));
        } else if (r.kind == 1) {
            const avbd_spring& s = w->springs[r.idx];
            auto S = std::make_unique<Spring<R>>(body(s.body_a), v3<R>(s.r_a), body(s.body_b), v3<R>(s.r_b));
            S->serial = s.serial;
            S->restLength = R(s.rest_length);
            S->disabled = (s.flags & AVBD_FORCE_DISABLED) != 0;
            Row<R>& row = S->rows[0];
            row.stiffness = R(s.stiffness);
            row.fracture = R(s.fracture);
            row.fmin = R(s.fmin);
            row.fmax = R(s.fmax);
            row.lambda = R(s.lambda);
            row.penalty = R(s.penalty);
            W.add_force(std::move(S));
        } else {
            const avbd_manifold& m = w->manifolds[r.idx];
            Body<R>* a = body(m.body_a);
            if (!a) return -2;
            auto M = std::make_unique<Manifold<R>>(a, body(m.body_b), W.onPenetrationPenalty);
            M->serial = m.serial;
            M->n = std::min(8, std::max(0, int(m.n_contacts)));
            M->friction = R(m.friction);
            for (int c = 0; c < M->n; ++c) {
                const avbd_contact& ct = m.contacts[c];
                M->contacts[c].id = ct.id;
                M->contacts[c].rA = v3<R>(ct.r_a);
                M->contacts[c].rB = v3<R>(ct.r_b);
                M->contacts[c].normal = v3<R>(ct.normal);
                M->contacts[c].penetration = R(ct.penetration);
                M->infos[c].stick = ct.stick != 0;
                for (int k = 0; k < 3; ++k) {
                    M->rows[c * 3 + k].lambda = R(ct.lambda[k]);
                    M->rows[c * 3 + k].penalty = R(ct.penalty[k]);
                }
            }
            W.add_force(std::move(M));
        }
    }
    W.nextSerial = maxSerial + 1;
    struct ERef { int32_t serial; int kind; int idx; };
    std::vector<ERef> erefs;
    for (int i = 0; i < w->n_tets; ++i)
        if (body(w->tets[i].v[0])) erefs.push_back({w->tets[i].serial, 0, i});
    for (int i = 0; i < w->n_tris; ++i)
        if (body(w->tris[i].v[0])) erefs.push_back({w->tris[i].serial, 1, i});
    std::stable_sort(erefs.begin(), erefs.end(), [](const ERef& a, const ERef& b) { return a.serial < b.serial; });
    for (const ERef& r : erefs) {
        if (r.kind == 0) {
            const avbd_tet& t = w->tets[r.idx];
            auto T = std::make_unique<NeoHookeanTet<R>>(body(t.v[0]), body(t.v[1]), body(t.v[2]), body(t.v[3]));
            T->serial = t.serial;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) T->DmInv.m[i][j] = R(t.dm_inv[i * 3 + j]);
            T->V0 = R(t.rest_volume);
            T->mu = R(t.mu);
            T->lam = R(t.lambda);
            T->index = int(W.energies.size());
            W.energies.push_back(std::move(T));
        } else {
            const avbd_tri& t = w->tris[r.idx];
            auto T = std::make_unique<StvkTri<R>>(body(t.v[0]), body(t.v[1]), body(t.v[2]));
            T->serial = t.serial;
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j) T->DmInv[i][j] = R(t.dm_inv[i * 2 + j]);
            T->area = R(t.rest_area);
            T->mu = R(t.mu);
            T->lam = R(t.lambda);
            T->index = int(W.energies.size());
            W.energies.push_back(std::move(T));
        }
    }
    return 0;
}

template <class R>
void counts(const World<R>& W, avbd_world* out) {
    out->n_scenes = 1;
    out->n_bodies = int(W.bodies.size());
    int nj = 0, ns = 0, nm = 0, nt = 0, nr = 0;
    for (auto& f : W.forces) {
        if (f->kind == ForceKind::Joint) ++nj;
        else if (f->kind == ForceKind::Spring) ++ns;
        else ++nm;
    }
    for (auto& e : W.energies) {
        if (dynamic_cast<NeoHookeanTet<R>*>(e.get())) ++nt; else ++nr;
    }
    out->n_joints = nj;
    out->n_springs = ns;
    out->n_manifolds = nm;
    out->n_tets = nt;
    out->n_tris = nr;
}

template <class R>
int save(const World<R>& W, avbd_world* out) {
    avbd_world c;
    counts(W, &c);
    if (out->n_bodies < c.n_bodies || out->n_joints < c.n_joints || out->n_springs < c.n_springs ||
        out->n_manifolds < c.n_manifolds || out->n_tets < c.n_tets || out->n_tris < c.n_tris || out->n_scenes < 1)
        return -1;
    avbd_scene& sc = out->scenes[0];
    sc.dt = float(W.dt);
    sc.iterations = W.iterations;
    sc.alpha = float(W.alpha);
    sc.beta = float(W.beta);
    sc.gamma = float(W.gamma);
    sc.on_penetration_penalty = float(W.onPenetrationPenalty);
    sc.post_stabilization = W.postStabilization ? 1 : 0;
    sc.projection_mode = W.projectionMode;
    put3(sc.gravity, W.gravity);
    put3(sc.bounds_min, W.boundsMin);
    put3(sc.bounds_max, W.boundsMax);
    sc.emergency_stop = W.emergencyStop ? 1 : 0;
    sc.n_removed = W.nRemoved;
    sc.options = W.options;
    sc.trust_region_rho = float(W.trustRegionRho);
    for (size_t i = 0; i < W.bodies.size(); ++i) {
        const Body<R>& B = *W.bodies[i];
        avbd_body& b = out->bodies[i];
        std::memset(&b, 0, sizeof(b));
        put3(b.pos, B.pos);
        put4(b.rot, B.rot);
        put3(b.scale, B.scale);
        put3(b.vel, B.vel);
        put3(b.ang_vel, B.angVel);
        put3(b.prev_vel, B.prevVel);
        put3(b.prev_ang_vel, B.prevAngVel);
        b.mass = float(B.mass);
        put3(b.inertia, B.inertia);
        b.friction = float(B.friction);
        b.radius = float(B.radius);
        b.flags = (B.isStatic ? AVBD_BODY_STATIC : 0u) | (B.isParticle ? AVBD_BODY_PARTICLE : 0u) |
                  (B.isSphere ? AVBD_BODY_SPHERE : 0u);
        b.scene = 0;
    }
    std::vector<const Force<R>*> fs;
    for (auto& f : W.forces) fs.push_back(f.get());
    std::stable_sort(fs.begin(), fs.end(), [](const Force<R>* a, const Force<R>* b) { return a->serial < b->serial; });
    int nj = 0, ns = 0, nm = 0;
    for (const Force<R>* f : fs) {
        if (f->kind == ForceKind::Joint) {
            const Joint<R>* J = static_cast<const Joint<R>*>(f);
            avbd_joint& j = out->joints[nj++];
            std::memset(&j, 0, sizeof(j));
            j.body_a = J->A ? J->A->index : -1;
            j.body_b = J->B->index;
            j.serial = J->serial;
            j.flags = J->disabled ? AVBD_FORCE_DISABLED : 0u;
            put3(j.r_a, J->rA);
            put3(j.r_b, J->rB);
            put4(j.rest_rot, J->restRot);
            j.torque_arm = float(J->torqueArm);
            for (int k = 0; k < 6; ++k) {
                j.stiffness[k] = float(J->rows[k].stiffness);
                j.fracture[k] = float(J->rows[k].fracture);
                j.fmin[k] = float(J->rows[k].fmin);
                j.fmax[k] = float(J->rows[k].fmax);
                j.lambda[k] = float(J->rows[k].lambda);
                j.penalty[k] = float(J->rows[k].penalty);
            }
        } else if (f->kind == ForceKind::Spring) {
            const Spring<R>* S = static_cast<const Spring<R>*>(f);
            avbd_spring& s = out->springs[ns++];
            std::memset(&s, 0, sizeof(s));
            s.body_a = S->A ? S->A->index : -1;
            s.body_b = S->B->index;
            s.serial = S->serial;
            s.flags = S->disabled ? AVBD_FORCE_DISABLED : 0u;
            put3(s.r_a, S->rA);
            put3(s.r_b, S->rB);
            s.rest_length = float(S->restLength);
            s.stiffness = float(S->rows[0].stiffness);
            s.fracture = float(S->rows[0].fracture);
            s.fmin = float(S->rows[0].fmin);
            s.fmax = float(S->rows[0].fmax);
            s.lambda = float(S->rows[0].lambda);
            s.penalty = float(S->rows[0].penalty);
        } else {
            const Manifold<R>* M = static_cast<const Manifold<R>*>(f);
            avbd_manifold& m = out->manifolds[nm++];
            std::memset(&m, 0, sizeof(m));
            m.body_a = M->A->index;
            m.body_b = M->B->index;
            m.serial = M->serial;
            m.n_contacts = M->n;
            m.friction = float(M->friction);
            for (int c = 0; c < M->n; ++c) {
                avbd_contact& ct = m.contacts[c];
                ct.id = M->contacts[c].id;
                ct.stick = M->infos[c].stick ? 1 : 0;
                put3(ct.r_a, M->contacts[c].rA);
                put3(ct.r_b, M->contacts[c].rB);
                put3(ct.normal, M->contacts[c].normal);
                ct.penetration = float(M->contacts[c].penetration);
                for (int k = 0; k < 3; ++k) {
                    ct.lambda[k] = float(M->rows[c * 3 + k].lambda);
                    ct.penalty[k] = float(M->rows[c * 3 + k].penalty);
                }
            }
        }
    }
    int nt = 0, nr = 0;
    std::vector<const Energy<R>*> es;
    for (auto& e : W.energies) es.push_back(e.get());
    std::stable_sort(es.begin(), es.end(), [](const Energy<R>* a, const Energy<R>* b) { return a->serial < b->serial; });
    for (const Energy<R>* e : es) {
        if (auto T = dynamic_cast<const NeoHookeanTet<R>*>(e)) {
            avbd_tet& t = out->tets[nt++];
            for (int k = 0; k < 4; ++k) t.v[k] = T->v[k]->index;
            t.serial = T->serial;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) t.dm_inv[i * 3 + j] = float(T->DmInv.m[i][j]);
            t.rest_volume = float(T->V0);
            t.mu = float(T->mu);
            t.lambda = float(T->lam);
        } else {
            auto S = static_cast<const StvkTri<R>*>(e);
            avbd_tri& t = out->tris[nr++];
            for (int k = 0; k < 3; ++k) t.v[k] = S->v[k]->index;
            t.serial = S->serial;
            for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j) t.dm_inv[i * 2 + j] = float(S->DmInv[i][j]);
            t.rest_area = float(S->area);
            t.mu = float(S->mu);
            t.lambda = float(S->lam);
        }
    }
    out->n_scenes = 1;
    out->n_bodies = c.n_bodies;
    out->n_joints = nj;
    out->n_springs = ns;
    out->n_manifolds = nm;
    out->n_tets = nt;
    out->n_tris = nr;
    return 0;
}

template <class R>
int box_collide(const float* a, const float* b, float* out, uint32_t* ids) {
    CollisionResult<R> r = collide_boxes(v3<R>(a), q4<R>(a + 3), v3<R>(a + 7), v3<R>(b), q4<R>(b + 3), v3<R>(b + 7));
    for (int i = 0; i < r.n; ++i) {
        float* o = out + i * 13;
        put3(o, r.c[i].position);
        put3(o + 3, r.c[i].normal);
        o[6] = float(r.c[i].penetration);
        put3(o + 7, r.c[i].rA);
        put3(o + 10, r.c[i].rB);
        ids[i] = r.c[i].id;
    }
    return r.n;
}

// collide_bodies on two bodies given as (pos 3, rot 4, scale 3) with their
// sphere flags (collide_boxes when neither is a sphere)
template <class R>
int body_collide(const float* a, const float* b, int sphereA, int sphereB, float* out, uint32_t* ids) {
    Body<R> A, B;
    A.pos = v3<R>(a); A.rot = q4<R>(a + 3); A.scale = v3<R>(a + 7); A.isSphere = A.collidesAsSphere = sphereA != 0;
    B.pos = v3<R>(b); B.rot = q4<R>(b + 3); B.scale = v3<R>(b + 7); B.isSphere = B.collidesAsSphere = sphereB != 0;
    CollisionResult<R> r = collide_bodies(&A, &B);
    for (int i = 0; i < r.n; ++i) {
        float* o = out + i * 13;
        put3(o, r.c[i].position);
        put3(o + 3, r.c[i].normal);
        o[6] = float(r.c[i].penetration);
        put3(o + 7, r.c[i].rA);
        put3(o + 10, r.c[i].rB);
        ids[i] = r.c[i].id;
    }
    return r.n;
}

template <class R>
static double nh_energy_change(const double* p, const double* dm, double V0, double mu, double lam, int vtx, const double* dx) {
    V3<R> P[4];
    for (int i = 0; i < 4; ++i) P[i] = V3<R>(R(p[i * 3]), R(p[i * 3 + 1]), R(p[i * 3 + 2]));
    M3<R> D;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) D.m[i][j] = R(dm[i * 3 + j]);
    return double(neohookean_energy_change(P, D, R(V0), R(mu), R(lam), vtx, V3<R>(R(dx[0]), R(dx[1]), R(dx[2]))));
}
template <class R>
static double stvk_energy_change_(const double* p, const double* dm, double area, double mu, double lam, int vtx, const double* dx) {
    V3<R> P[3];
    for (int i = 0; i < 3; ++i) P[i] = V3<R>(R(p[i * 3]), R(p[i * 3 + 1]), R(p[i * 3 + 2]));
    R D[2][2] = {{R(dm[0]), R(dm[1])}, {R(dm[2]), R(dm[3])}};
    return double(stvk_energy_change(P, D, R(area), R(mu), R(lam), vtx, V3<R>(R(dx[0]), R(dx[1]), R(dx[2]))));
}

template <class R>
static void nh_terms(const double* p, const double* dm, double V0, double mu, double lam, int vtx, int proj, double* g,
                     double* H) {
    V3<R> P[4];
    for (int i = 0; i < 4; ++i) P[i] = V3<R>(R(p[i * 3]), R(p[i * 3 + 1]), R(p[i * 3 + 2]));
    M3<R> D;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) D.m[i][j] = R(dm[i * 3 + j]);
    V3<R> gg; M3<R> HH;
    neohookean_vertex_terms(P, D, R(V0), R(mu), R(lam), vtx, proj, R(0), gg, HH);
    for (int i = 0; i < 3; ++i) { g[i] = double(gg[i]); for (int j = 0; j < 3; ++j) H[i * 3 + j] = double(HH.m[i][j]); }
}
template <class R>
static void stvk_terms(const double* p, const double* dm, double area, double mu, double lam, int vtx, int proj,
                       double* g, double* H) {
    V3<R> P[3];
    for (int i = 0; i < 3; ++i) P[i] = V3<R>(R(p[i * 3]), R(p[i * 3 + 1]), R(p[i * 3 + 2]));
    R D[2][2] = {{R(dm[0]), R(dm[1])}, {R(dm[2]), R(dm[3])}};
    V3<R> gg; M3<R> HH;
    stvk_vertex_terms(P, D, R(area), R(mu), R(lam), vtx, proj, R(0), gg, HH);
    for (int i = 0; i < 3; ++i) { g[i] = double(gg[i]); for (int j = 0; j < 3; ++j) H[i * 3 + j] = double(HH.m[i][j]); }
}
}  // namespace

extern "C" {

void* orc_create(int fp64) {
    Handle* h = new Handle();
    h->fp64 = fp64 != 0;
    return h;
}
void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

int orc_load(void* hp, const avbd_world* w, int scene) {
    Handle* h = static_cast<Handle*>(hp);
    return h->fp64 ? load(h->wd, w, scene) : load(h->wf, w, scene);
}
int orc_set_order(void* hp, int order) {
    Handle* h = static_cast<Handle*>(hp);
    h->wf.order = h->wd.order = order ? VisitOrder::ColourMajor : VisitOrder::Natural;
    return 0;
}
int orc_step(void* hp, int n) {
    Handle* h = static_cast<Handle*>(hp);
    for (int i = 0; i < n; ++i) {
        if (h->fp64) h->wd.step(); else h->wf.step();
    }
    return 0;
}
int orc_set_drag(void* hp, const int32_t* idx, const float* vel, const int32_t* dragged, int n) {
    Handle* h = static_cast<Handle*>(hp);
    return h->fp64 ? set_drag(h->wd, idx, vel, dragged, n) : set_drag(h->wf, idx, vel, dragged, n);
}
int orc_remove_bodies(void* hp, const int32_t* idx, int n) {
    Handle* h = static_cast<Handle*>(hp);
    return h->fp64 ? remove_listed(h->wd, idx, n) : remove_listed(h->wf, idx, n);
}
// Line / point data of the last step: copies at most cap records (10 / 8
// floats each), writes the full counts.
int orc_debug(void* hp, float* lines, int64_t lcap, float* points, int64_t pcap, int64_t* nl, int64_t* np) {
    Handle* h = static_cast<Handle*>(hp);
    const std::vector<float>& L = h->fp64 ? h->wd.lineData : h->wf.lineData;
    const std::vector<float>& P = h->fp64 ? h->wd.pointData : h->wf.pointData;
    *nl = (int64_t)L.size() / 10;
    *np = (int64_t)P.size() / 8;
    if (lines) std::copy(L.begin(), L.begin() + std::min<int64_t>(lcap, *nl) * 10, lines);
    if (points) std::copy(P.begin(), P.begin() + std::min<int64_t>(pcap, *np) * 8, points);
    return 0;
}
int orc_counts(void* hp, avbd_world* out) {
    Handle* h = static_cast<Handle*>(hp);
    if (h->fp64) counts(h->wd, out); else counts(h->wf, out);
    return 0;
}
int orc_save(void* hp, avbd_world* out) {
    Handle* h = static_cast<Handle*>(hp);
    return h->fp64 ? save(h->wd, out) : save(h->wf, out);
}
// Greedy colouring of the current topology (fills n entries, -1 = static).
int orc_colours(void* hp, int32_t* out, int n) {
    Handle* h = static_cast<Handle*>(hp);
    std::vector<int> col;
    int nc;
    if (h->fp64) { h->wd.compute_colouring(); col = h->wd.colour; nc = h->wd.nColours; }
    else { h->wf.compute_colouring(); col = h->wf.colour; nc = h->wf.nColours; }
    for (int i = 0; i < n && i < int(col.size()); ++i) out[i] = col[i];
    return nc;
}
int orc_timings(void* hp, double* out) {
    Handle* h = static_cast<Handle*>(hp);
    const Timings& t = h->fp64 ? h->wd.timings : h->wf.timings;
    out[0] = t.broad; out[1] = t.warm; out[2] = t.predict; out[3] = t.primalDual;
    out[4] = t.velocity; out[5] = t.postStab; out[6] = t.total;
    return 0;
}

// Box-box collision. a, b: pos[3], rot[4] (w,x,y,z), scale[3].  out: up to 8
// records of 13 floats {position[3], normal[3], penetration, rA[3], rB[3]}.
int orc_collide(const float* a, const float* b, int fp64, float* out, uint32_t* ids) {
    return fp64 ? box_collide<double>(a, b, out, ids) : box_collide<float>(a, b, out, ids);
}

// Contacts of two bodies, either possibly a sphere (AVBD_SCENE_SPHERE_COLLISION).
int orc_collide_bodies(const float* a, const float* b, int sphereA, int sphereB, int fp64, float* out, uint32_t* ids) {
    return fp64 ? body_collide<double>(a, b, sphereA, sphereB, out, ids) : body_collide<float>(a, b, sphereA, sphereB, out, ids);
}

// E(x - dx) - E(x) of one element for a move of its vertex vtx.
double orc_nh_energy_change(const double* p, const double* dm, double V0, double mu, double lam, int vtx,
                            const double* dx, int fp64) {
    return fp64 ? nh_energy_change<double>(p, dm, V0, mu, lam, vtx, dx) : nh_energy_change<float>(p, dm, V0, mu, lam, vtx, dx);
}
double orc_stvk_energy_change(const double* p, const double* dm, double area, double mu, double lam, int vtx,
                              const double* dx, int fp64) {
    return fp64 ? stvk_energy_change_<double>(p, dm, area, mu, lam, vtx, dx)
                : stvk_energy_change_<float>(p, dm, area, mu, lam, vtx, dx);
}

// LDLT solve of a symmetric 6x6 (row-major, lower triangle read).
int orc_ldlt_solve(const double* A, const double* b, double* x, int fp64) {
    if (fp64) {
        M6<double> M; V6<double> r;
        for (int i = 0; i < 6; ++i) { r[i] = b[i]; for (int j = 0; j < 6; ++j) M.m[i][j] = A[i * 6 + j]; }
        LDLT6<double> f; f.factor(M);
        V6<double> s = f.solve(r);
        for (int i = 0; i < 6; ++i) x[i] = s[i];
    } else {
        M6<float> M; V6<float> r;
        for (int i = 0; i < 6; ++i) { r[i] = float(b[i]); for (int j = 0; j < 6; ++j) M.m[i][j] = float(A[i * 6 + j]); }
        LDLT6<float> f; f.factor(M);
        V6<float> s = f.solve(r);
        for (int i = 0; i < 6; ++i) x[i] = s[i];
    }
    return 0;
}

int orc_nh_terms(const double* p, const double* dm, double V0, double mu, double lam, int vtx, int proj, int fp64,
                 double* g, double* H) {
    if (fp64) nh_terms<double>(p, dm, V0, mu, lam, vtx, proj, g, H);
    else nh_terms<float>(p, dm, V0, mu, lam, vtx, proj, g, H);
    return 0;
}

int orc_stvk_terms(const double* p, const double* dm, double area, double mu, double lam, int vtx, int proj, int fp64,
                   double* g, double* H) {
    if (fp64) stvk_terms<double>(p, dm, area, mu, lam, vtx, proj, g, H);
    else stvk_terms<float>(p, dm, area, mu, lam, vtx, proj, g, H);
    return 0;
}

int orc_svd3(const double* Fp, int fp64, double* Up, double* Sp, double* Vp) {
    if (fp64) {
        M3<double> F, U, V; V3<double> S;
        for (int i = 0; i < 9; ++i) F.m[i / 3][i % 3] = Fp[i];
        svd3(F, U, S, V);
        for (int i = 0; i < 9; ++i) { Up[i] = U.m[i / 3][i % 3]; Vp[i] = V.m[i / 3][i % 3]; }
        for (int i = 0; i < 3; ++i) Sp[i] = S[i];
    } else {
        M3<float> F, U, V; V3<float> S;
        for (int i = 0; i < 9; ++i) F.m[i / 3][i % 3] = float(Fp[i]);
        svd3(F, U, S, V);
        for (int i = 0; i < 9; ++i) { Up[i] = U.m[i / 3][i % 3]; Vp[i] = V.m[i / 3][i % 3]; }
        for (int i = 0; i < 3; ++i) Sp[i] = S[i];
    }
    return 0;
}

int orc_sym_eig3(const double* Ap, double* w, double* Vp) {
    M3<double> A, V; V3<double> ww;
    for (int i = 0; i < 9; ++i) A.m[i / 3][i % 3] = Ap[i];
    sym_eig3(A, ww, V);
    for (int i = 0; i < 3; ++i) w[i] = ww[i];
    for (int i = 0; i < 9; ++i) Vp[i] = V.m[i / 3][i % 3];
    return 0;
}

// Rotation helpers (FP32): op 0 = RotationDifference(q, p) -> 3,
// 1 = QuaternionFromDifference(v = a[0..2], sign = a[3]) -> 4,
// 2 = QuaternionToRotVec(q) -> 3, 3 = rotation matrix of q -> 9.
int orc_rot_helper(int op, const float* a, const float* b, float* out) {
    if (op == 0) { put3(out, rotation_difference(q4<float>(a), q4<float>(b))); return 3; }
    if (op == 1) { put4(out, quaternion_from_difference(v3<float>(a), a[3])); return 4; }
    if (op == 2) { put3(out, quaternion_to_rotvec(q4<float>(a))); return 3; }
    if (op == 3) { M3<float> R = rotmat(q4<float>(a)); for (int i = 0; i < 9; ++i) out[i] = R.m[i / 3][i % 3]; return 9; }
    return -1;
}

}  // extern "C"
