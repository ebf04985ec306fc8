// solver.cpp -- Solver (src/physics/solver.hpp, solver.cpp:18-164,199-627)
// over the AVBD step's C ABI.  The host objects are converted to the ABI's
// records on upload and updated from them after every step; see solver.hpp
// for when each transfer happens.
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>


namespace {

void check(int rc, const char* what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + avbd_last_error());
}

template <class V>
void put3(float* d, const V& v) {
    d[0] = v[0];
    d[1] = v[1];
    d[2] = v[2];
}

Vector3f get3(const float* s) { return Vector3f(s[0], s[1], s[2]); }

// The body fields a step changes or a caller may write between steps.
bool same_state(const avbd_body& a, const avbd_body& b) {
    return std::memcmp(a.pos, b.pos, sizeof(a.pos)) == 0 && std::memcmp(a.rot, b.rot, sizeof(a.rot)) == 0 &&
           std::memcmp(a.vel, b.vel, sizeof(a.vel)) == 0 && std::memcmp(a.ang_vel, b.ang_vel, sizeof(a.ang_vel)) == 0 &&
           std::memcmp(a.prev_vel, b.prev_vel, sizeof(a.prev_vel)) == 0 &&
           std::memcmp(a.prev_ang_vel, b.prev_ang_vel, sizeof(a.prev_ang_vel)) == 0;
}

// Everything else: a difference needs a full upload.
bool same_structure(const avbd_body& a, const avbd_body& b) {
    return std::memcmp(a.scale, b.scale, sizeof(a.scale)) == 0 && a.mass == b.mass &&
           std::memcmp(a.inertia, b.inertia, sizeof(a.inertia)) == 0 && a.friction == b.friction &&
           a.radius == b.radius && a.flags == b.flags && a.scene == b.scene;
}

}  // namespace

Solver::Solver() {}

Solver::~Solver() {
    Clear();
    if (ctx_) avbd_destroy(ctx_);
}

void Solver::Start() { Clear(); }

void Solver::Clear() {
    mirroring_ = true;  // the whole world goes: no per-object bookkeeping
    solverForces.clear();
    solverBodies.clear();
    solverEnergies.clear();
    mirroring_ = false;
    forcePtrs.clear();
    bodyPtrs.clear();
    energyPtrs.clear();
    prevTotalEnergy = 0.f;
    trustRegionRho = 0.f;
    emergencyStop = false;
    forcesStale_ = false;
    structureDirty_ = true;
    pendingAdds_.clear();
    timingAccumulator.clear();
}

void Solver::RebuildPtrCaches() {
    bodyPtrs.clear();
    bodyPtrs.reserve(solverBodies.size());
    for (auto& b : solverBodies) bodyPtrs.push_back(b.get());
    forcePtrs.clear();
    forcePtrs.reserve(solverForces.size());
    for (auto& f : solverForces) forcePtrs.push_back(f.get());
    energyPtrs.clear();
    energyPtrs.reserve(solverEnergies.size());
    for (auto& e : solverEnergies) energyPtrs.push_back(e.get());
}

// ---------------------------------------------------------------- mutation
void Solver::MarkStructureChanged() {
    if (mirroring_ || removingOnDevice_) return;
    // the host forces must be current before the world changes under them
    if (forcesStale_) SyncForces();
    structureDirty_ = true;
}

int64_t Solver::TakeForceSerial() {
    if (mirroring_) return -1;  // a mirrored contact takes the device's serial
    MarkStructureChanged();
    return nextForceSerial_++;
}

int32_t Solver::TakeEnergySerial() {
    if (mirroring_) return -1;
    MarkStructureChanged();
    return nextEnergySerial_++;
}

Mesh* Solver::AddBody(ModelType modelType, float density, float friction, const Vector3f& position,
                      const Vector3f& scale, const Vector3f& velocity, const Quaternionf rotation,
                      const Vector3f& angularVelocity, bool isStatic, const Vector3f& color) {
    // With the device world current, the next Step adds the body on the
    // device (avbd_add_bodies, into a slot a removal freed) if it can.
    const bool on_device = ctx_ && uploaded_ && !structureDirty_ && !mirroring_;
    if (on_device) {
        if (forcesStale_) SyncForces();
    } else {
        MarkStructureChanged();
    }
    auto mesh = std::make_unique<Mesh>(this, modelType, color);
    Mesh* raw = mesh.get();
    raw->density = density;
    raw->friction = friction;
    raw->velocity = velocity;
    raw->angularVelocity = angularVelocity;
    raw->isStatic = isStatic;
    raw->transform.SetPosition(position);
    raw->transform.SetRotation(rotation);
    raw->transform.SetScale(scale);
    const float sx = scale[0], sy = scale[1], sz = scale[2];
    raw->detectionRadius = std::sqrt((sx * sx + sy * sy) + sz * sz) * 0.5f;
    if (isStatic) {
        raw->inertiaTensorBody = Matrix3f::Zero();
        raw->inertiaTensorBodyInv = Matrix3f::Zero();
    } else {
        const float sx2 = sx * sx, sy2 = sy * sy, sz2 = sz * sz;
        float I[3];
        if (modelType == ModelType_Sphere) {
            const float r = sx * 0.5f;
            raw->mass = float(double(density * (4.f / 3.f)) * 3.141592653589793 * double(r) * double(r) * double(r));
            I[0] = I[1] = I[2] = (2.f / 5.f) * raw->mass * r * r;
        } else {
            // cubes; the reference leaves a pyramid's inertia undefined
            // (solver.cpp:77-90), here it takes the box's
            raw->mass = ((density * sx) * sy) * sz;
            const float k = raw->mass / 12.f;
            I[0] = k * (sy2 + sz2);
            I[1] = k * (sx2 + sz2);
            I[2] = k * (sx2 + sy2);
        }
        raw->inertiaTensorBody = Matrix3f::Zero();
        raw->inertiaTensorBodyInv = Matrix3f::Zero();
        for (int i = 0; i < 3; ++i) {
            raw->inertiaTensorBody(i, i) = I[i];
            raw->inertiaTensorBodyInv(i, i) = I[i] != 0.f ? 1.f / I[i] : 0.f;
        }
    }
    raw->solverIndex = static_cast<int>(solverBodies.size());
    solverBodies.push_back(std::move(mesh));
    if (on_device) pendingAdds_.push_back(raw);
    RebuildPtrCaches();
    return raw;
}

Mesh* Solver::AddParticle(float mass, float friction, const Vector3f& position, const Vector3f& velocity,
                          bool isStatic, const Vector3f& color) {
    const float s = 0.1f;
    Mesh* p = AddBody(ModelType_Cube, mass / ((s * s) * s), friction, position, Vector3f(s, s, s), velocity,
                      Quaternionf::Identity(), Vector3f::Zero(), isStatic, color);
    p->isParticle = true;
    p->inertiaTensorBody = Matrix3f::Zero();
    p->inertiaTensorBodyInv = Matrix3f::Zero();
    p->angularVelocity = Vector3f::Zero();
    return p;
}

Force* Solver::AddForce(std::unique_ptr<Force> force) {
    Force* raw = force.get();
    raw->solverIndex = static_cast<int>(solverForces.size());
    solverForces.push_back(std::move(force));
    return raw;
}

void Solver::RemoveForce(Force* force) {
    const int idx = force->solverIndex;
    if (idx < 0) return;
    MarkStructureChanged();
    const int last = static_cast<int>(solverForces.size()) - 1;
    if (idx != last) {
        std::swap(solverForces[idx], solverForces[last]);
        solverForces[idx]->solverIndex = idx;
    }
    solverForces.pop_back();  // ~Force leaves the bodies' lists
}

Energy* Solver::AddEnergy(std::unique_ptr<Energy> energy) {
    Energy* raw = energy.get();
    raw->solverIndex = static_cast<int>(solverEnergies.size());
    solverEnergies.push_back(std::move(energy));
    return raw;
}

void Solver::RemoveEnergy(Energy* energy) {
    const int idx = energy->solverIndex;
    if (idx < 0) return;
    MarkStructureChanged();
    const int last = static_cast<int>(solverEnergies.size()) - 1;
    if (idx != last) {
        std::swap(solverEnergies[idx], solverEnergies[last]);
        solverEnergies[idx]->solverIndex = idx;
    }
    solverEnergies.pop_back();
}

void Solver::RemoveBody(Mesh* body) {
    const int idx = body->solverIndex;
    const int last = static_cast<int>(solverBodies.size()) - 1;
    // With the device world current, the device removes the body itself
    // (avbd_remove_bodies) and the next Step needs no re-upload.
    const bool on_device = ctx_ && uploaded_ && !structureDirty_ && !mirroring_ && pendingAdds_.empty();
    if (on_device) {
        if (forcesStale_) SyncForces();
        const int32_t i = idx;
        check(avbd_remove_bodies(ctx_, &i, 1), "avbd_remove_bodies");
        avbd_body* rec = reinterpret_cast<avbd_body*>(lastBodies_.data());
        const size_t n = lastBodies_.size() / sizeof(avbd_body);
        if (n == solverBodies.size()) {
            if (idx != last) rec[idx] = rec[last];
            lastBodies_.resize((n - 1) * sizeof(avbd_body));
        } else {
            structureDirty_ = true;  // out of step: upload at the next Step
        }
        removingOnDevice_ = true;
    } else {
        MarkStructureChanged();
    }
    if (idx != last) {
        std::swap(solverBodies[idx], solverBodies[last]);
        solverBodies[idx]->solverIndex = idx;
    }
    solverBodies.pop_back();  // ~Mesh removes the body's forces and energies
    removingOnDevice_ = false;
    RebuildPtrCaches();
}

// ---------------------------------------------------------------- records
void Solver::SceneRecord(std::vector<unsigned char>& out) const {
    avbd_scene s{};
    s.dt = stepValue;
    s.iterations = numIterations;
    s.alpha = alpha;
    s.beta = beta;
    s.gamma = gamma;
    s.on_penetration_penalty = onPenetrationPenalty;
    s.post_stabilization = postStabilization ? 1 : 0;
    s.projection_mode = static_cast<int32_t>(projectionMode);
    put3(s.gravity, gravity);
    put3(s.bounds_min, boundsMin);
    put3(s.bounds_max, boundsMax);
    s.options = (trustRegionUpdate ? AVBD_SCENE_TRUST_REGION : 0u) | (sphereCollision ? AVBD_SCENE_SPHERE_COLLISION : 0u);
    s.trust_region_rho = trustRegionRho;
    out.assign(reinterpret_cast<const unsigned char*>(&s), reinterpret_cast<const unsigned char*>(&s) + sizeof(s));
}

void Solver::BodyRecords(std::vector<avbd_body>& out) const {
    out.assign(solverBodies.size(), avbd_body{});
    for (size_t i = 0; i < solverBodies.size(); ++i) {
        const Mesh& m = *solverBodies[i];
        avbd_body& b = out[i];
        put3(b.pos, m.transform.GetPosition());
        const Quaternionf q = m.transform.GetRotation();
        b.rot[0] = q.w(); b.rot[1] = q.x(); b.rot[2] = q.y(); b.rot[3] = q.z();
        put3(b.scale, m.transform.GetScale());
        put3(b.vel, m.velocity);
        put3(b.ang_vel, m.angularVelocity);
        put3(b.prev_vel, m.prevVelocity);
        put3(b.prev_ang_vel, m.prevAngularVelocity);
        b.mass = m.isStatic ? 0.f : m.mass;
        for (int k = 0; k < 3; ++k) b.inertia[k] = m.isStatic ? 0.f : m.inertiaTensorBody(k, k);
        b.friction = m.friction;
        b.radius = m.detectionRadius;
        b.flags = (m.isStatic ? AVBD_BODY_STATIC : 0u) | (m.isParticle ? AVBD_BODY_PARTICLE : 0u) |
                  (m.modelType == ModelType_Sphere ? AVBD_BODY_SPHERE : 0u);
        b.scene = 0;
    }
}

void Solver::EnsureContext() {
    if (!ctx_) check(avbd_create(device, &ctx_), "avbd_create");
}

avbd_world WorldRecords::view() {
    avbd_world w{};
    w.n_scenes = 1;
    w.n_bodies = (int32_t)bodies.size();
    w.n_joints = (int32_t)joints.size();
    w.n_springs = (int32_t)springs.size();
    w.n_manifolds = (int32_t)manifolds.size();
    w.n_tets = (int32_t)tets.size();
    w.n_tris = (int32_t)tris.size();
    w.scenes = &scene;
    w.bodies = bodies.data();
    w.joints = joints.data();
    w.springs = springs.data();
    w.manifolds = manifolds.data();
    w.tets = tets.data();
    w.tris = tris.data();
    return w;
}

avbd_ctx* Solver::DeviceContext() {
    EnsureContext();
    return ctx_;
}

namespace {

// A manifold's contact points and row state from its record.
void set_contacts(Manifold* m, const avbd_manifold& r) {
    m->numContactPoints = r.n_contacts;
    m->friction = r.friction;
    for (int c = 0; c < 8; ++c) {
        const avbd_contact& ct = r.contacts[c];
        ContactPoint& cp = m->contactPoints[c];
        cp.id = ct.id;
        cp.rA = get3(ct.r_a);
        cp.rB = get3(ct.r_b);
        cp.normal = get3(ct.normal);
        cp.penetration = ct.penetration;
        m->contactInfos[c].stick = ct.stick != 0;
        for (int k = 0; k < 3; ++k) {
            m->constraintPoints[3 * c + k].lambda = c < r.n_contacts ? ct.lambda[k] : 0.f;
            m->constraintPoints[3 * c + k].penalty = c < r.n_contacts ? ct.penalty[k] : 0.f;
        }
    }
}

void set_row(ConstraintPointProperties& p, float stiffness, float fracture, float fmin, float fmax, float lambda,
             float penalty) {
    p.stiffness = stiffness;
    p.fracture = fracture;
    p.fminMagnitude = fmin;
    p.fmaxMagnitude = fmax;
    p.lambda = lambda;
    p.penalty = penalty;
}

}  // namespace

void Solver::SaveSnapshot(const std::string& path) {
    WorldRecords r;
    ExportWorld(r);
    avbd_world w = r.view();
    if (avbd_snapshot_write(path.c_str(), &w) != 0) throw std::runtime_error(avbd_last_error());
}

void Solver::LoadSnapshot(const std::string& path) {
    avbd_world cnt{};
    if (avbd_snapshot_counts(path.c_str(), &cnt) != 0) throw std::runtime_error(avbd_last_error());
    if (cnt.n_scenes != 1)
        throw std::runtime_error("Solver::LoadSnapshot: a solver holds one scene, the snapshot has " +
                                 std::to_string(cnt.n_scenes));
    WorldRecords r;
    r.bodies.resize(cnt.n_bodies);
    r.joints.resize(cnt.n_joints);
    r.springs.resize(cnt.n_springs);
    r.manifolds.resize(cnt.n_manifolds);
    r.tets.resize(cnt.n_tets);
    r.tris.resize(cnt.n_tris);
    avbd_world w = r.view();
    if (avbd_snapshot_read(path.c_str(), &w) != 0) throw std::runtime_error(avbd_last_error());
    // Every check before the world is replaced: a bad file leaves it as it was.
    const int nb = cnt.n_bodies;
    auto in_range = [&](int32_t i, bool may_be_world) { return (i == -1 && may_be_world) || (i >= 0 && i < nb); };
    bool ok = true;
    for (const avbd_body& b : r.bodies) ok = ok && b.scene == 0;
    for (const avbd_joint& j : r.joints) ok = ok && in_range(j.body_a, true) && in_range(j.body_b, false);
    for (const avbd_spring& sp : r.springs) ok = ok && in_range(sp.body_a, true) && in_range(sp.body_b, false);
    for (const avbd_manifold& m : r.manifolds)
        ok = ok && in_range(m.body_a, false) && in_range(m.body_b, false) && m.n_contacts >= 0 && m.n_contacts <= 8;
    for (const avbd_tet& t : r.tets)
        for (int k = 0; k < 4; ++k) ok = ok && in_range(t.v[k], false);
    for (const avbd_tri& t : r.tris)
        for (int k = 0; k < 3; ++k) ok = ok && in_range(t.v[k], false);
    if (!ok) throw std::runtime_error("Solver::LoadSnapshot: a record names a body out of range or of another scene");
    auto body = [&](int32_t i) -> Mesh* { return i == -1 ? nullptr : solverBodies[i].get(); };

    Clear();
    const avbd_scene& sc = r.scene;
    stepValue = sc.dt;
    numIterations = sc.iterations;
    alpha = sc.alpha;
    beta = sc.beta;
    gamma = sc.gamma;
    onPenetrationPenalty = sc.on_penetration_penalty;
    postStabilization = sc.post_stabilization != 0;
    projectionMode = static_cast<EigenProjectionMode>(sc.projection_mode);
    gravity = get3(sc.gravity);
    boundsMin = get3(sc.bounds_min);
    boundsMax = get3(sc.bounds_max);
    emergencyStop = sc.emergency_stop != 0;
    trustRegionRho = sc.trust_region_rho;
    trustRegionUpdate = (sc.options & AVBD_SCENE_TRUST_REGION) != 0;
    sphereCollision = (sc.options & AVBD_SCENE_SPHERE_COLLISION) != 0;

    // The objects take the records' serials (the constructors' serial
    // requests are answered with -1 while mirroring_ is set).
    mirroring_ = true;
    for (const avbd_body& b : r.bodies) {
        auto mesh = std::make_unique<Mesh>(this, (b.flags & AVBD_BODY_SPHERE) ? ModelType_Sphere : ModelType_Cube);
        Mesh& m = *mesh;
        m.transform.SetPosition(get3(b.pos));
        m.transform.SetRotation(Quaternionf(b.rot[0], b.rot[1], b.rot[2], b.rot[3]));
        m.transform.SetScale(get3(b.scale));
        m.velocity = get3(b.vel);
        m.angularVelocity = get3(b.ang_vel);
        m.prevVelocity = get3(b.prev_vel);
        m.prevAngularVelocity = get3(b.prev_ang_vel);
        m.isStatic = (b.flags & AVBD_BODY_STATIC) != 0;
        m.isParticle = (b.flags & AVBD_BODY_PARTICLE) != 0;
        m.mass = b.mass;
        const float vol = (b.scale[0] * b.scale[1]) * b.scale[2];
        m.density = vol > 0.f ? b.mass / vol : 0.f;
        m.friction = b.friction;
        m.detectionRadius = b.radius;
        m.inertiaTensorBody = Matrix3f::Zero();
        m.inertiaTensorBodyInv = Matrix3f::Zero();
        for (int k = 0; k < 3; ++k) {
            m.inertiaTensorBody(k, k) = b.inertia[k];
            m.inertiaTensorBodyInv(k, k) = b.inertia[k] != 0.f ? 1.f / b.inertia[k] : 0.f;
        }
        m.solverIndex = static_cast<int>(solverBodies.size());
        solverBodies.push_back(std::move(mesh));
    }
    // Forces in creation order, so that each body's list is in serial order
    // as it would be had they been created one by one.
    struct Item {
        int64_t serial;
        int kind, i;
    };
    std::vector<Item> order;
    for (int i = 0; i < (int)r.joints.size(); ++i) order.push_back({r.joints[i].serial, 0, i});
    for (int i = 0; i < (int)r.springs.size(); ++i) order.push_back({r.springs[i].serial, 1, i});
    for (int i = 0; i < (int)r.manifolds.size(); ++i) order.push_back({r.manifolds[i].serial, 2, i});
    std::sort(order.begin(), order.end(), [](const Item& a, const Item& b) { return a.serial < b.serial; });
    int64_t max_force = -1;
    for (const Item& it : order) {
        Force* f = nullptr;
        if (it.kind == 0) {
            const avbd_joint& j = r.joints[it.i];
            auto owned = std::make_unique<Joint>(this, body(j.body_a), get3(j.r_a), body(j.body_b),
                                                 get3(j.r_b));
            owned->restRotation = Quaternionf(j.rest_rot[0], j.rest_rot[1], j.rest_rot[2], j.rest_rot[3]);
            owned->torqueArm = j.torque_arm;
            for (int k = 0; k < 6; ++k)
                set_row(owned->constraintPoints[k], j.stiffness[k], j.fracture[k], j.fmin[k], j.fmax[k], j.lambda[k],
                        j.penalty[k]);
            owned->disabled = (j.flags & AVBD_FORCE_DISABLED) != 0;
            f = AddForce(std::move(owned));
        } else if (it.kind == 1) {
            const avbd_spring& s = r.springs[it.i];
            auto owned = std::make_unique<Spring>(this, body(s.body_a), get3(s.r_a), body(s.body_b),
                                                  get3(s.r_b), s.stiffness, s.rest_length, s.fmin == 0.f);
            set_row(owned->constraintPoints[0], s.stiffness, s.fracture, s.fmin, s.fmax, s.lambda, s.penalty);
            owned->disabled = (s.flags & AVBD_FORCE_DISABLED) != 0;
            f = AddForce(std::move(owned));
        } else {
            const avbd_manifold& mr = r.manifolds[it.i];
            auto owned = std::make_unique<Manifold>(this, body(mr.body_a), body(mr.body_b));
            set_contacts(owned.get(), mr);
            f = AddForce(std::move(owned));
        }
        f->serial = it.serial;
        max_force = std::max<int64_t>(max_force, it.serial);
    }
    // Energies likewise: the constructors derive the rest state from the
    // current positions, the records' rest state replaces it.
    struct EItem {
        int32_t serial;
        int kind, i;
    };
    std::vector<EItem> eorder;
    for (int i = 0; i < (int)r.tets.size(); ++i) eorder.push_back({r.tets[i].serial, 0, i});
    for (int i = 0; i < (int)r.tris.size(); ++i) eorder.push_back({r.tris[i].serial, 1, i});
    std::sort(eorder.begin(), eorder.end(), [](const EItem& a, const EItem& b) { return a.serial < b.serial; });
    // Young's modulus and Poisson's ratio back from the Lame parameters
    auto young = [](float mu, float lam) { return mu + lam != 0.f ? mu * (3.f * lam + 2.f * mu) / (lam + mu) : 0.f; };
    auto poisson = [](float mu, float lam) { return mu + lam != 0.f ? lam / (2.f * (lam + mu)) : 0.f; };
    int32_t max_energy = -1;
    for (const EItem& it : eorder) {
        Energy* e = nullptr;
        if (it.kind == 0) {
            const avbd_tet& t = r.tets[it.i];
            auto owned = std::make_unique<NeoHookeanFEM>(this, body(t.v[0]), body(t.v[1]),
                                                         body(t.v[2]), body(t.v[3]),
                                                         young(t.mu, t.lambda), poisson(t.mu, t.lambda));
            std::memcpy(owned->DmInv, t.dm_inv, sizeof(owned->DmInv));
            owned->restVolume = t.rest_volume;
            owned->lameMu = t.mu;
            owned->lameLambda = t.lambda;
            e = AddEnergy(std::move(owned));
        } else {
            const avbd_tri& t = r.tris[it.i];
            auto owned = std::make_unique<STVKFEM>(this, body(t.v[0]), body(t.v[1]), body(t.v[2]),
                                                   young(t.mu, t.lambda), poisson(t.mu, t.lambda));
            std::memcpy(owned->DmInv2D, t.dm_inv, sizeof(owned->DmInv2D));
            owned->restArea = t.rest_area;
            owned->lameMu = t.mu;
            owned->lameLambda = t.lambda;
            e = AddEnergy(std::move(owned));
        }
        e->serial = it.serial;
        max_energy = std::max(max_energy, it.serial);
    }
    mirroring_ = false;
    nextForceSerial_ = max_force + 1;
    nextEnergySerial_ = max_energy + 1;
    structureDirty_ = true;
    RebuildPtrCaches();
}

void Solver::ExportWorld(WorldRecords& out) {
    RebuildPtrCaches();
    std::vector<unsigned char> sc;
    SceneRecord(sc);
    std::memcpy(&out.scene, sc.data(), sizeof(out.scene));
    BodyRecords(out.bodies);
    out.joints.clear();
    out.springs.clear();
    out.manifolds.clear();
    out.tets.clear();
    out.tris.clear();
    std::vector<std::pair<int64_t, const Manifold*>> mans;
    auto index = [](const Mesh* m) { return m ? m->solverIndex : -1; };
    for (const auto& fp : solverForces) {
        const Force* f = fp.get();
        if (const Joint* j = dynamic_cast<const Joint*>(f)) {
            avbd_joint r{};
            r.body_a = index(j->bodyA);
            r.body_b = index(j->bodyB);
            r.serial = j->serial;
            r.flags = j->disabled ? AVBD_FORCE_DISABLED : 0u;
            put3(r.r_a, j->rA);
            put3(r.r_b, j->rB);
            r.rest_rot[0] = j->restRotation.w(); r.rest_rot[1] = j->restRotation.x();
            r.rest_rot[2] = j->restRotation.y(); r.rest_rot[3] = j->restRotation.z();
            r.torque_arm = j->torqueArm;
            for (int k = 0; k < 6; ++k) {
                const ConstraintPointProperties& p = j->constraintPoints[k];
                r.stiffness[k] = p.stiffness; r.fracture[k] = p.fracture; r.fmin[k] = p.fminMagnitude;
                r.fmax[k] = p.fmaxMagnitude; r.lambda[k] = p.lambda; r.penalty[k] = p.penalty;
            }
            out.joints.push_back(r);
        } else if (const Spring* s = dynamic_cast<const Spring*>(f)) {
            avbd_spring r{};
            r.body_a = index(s->bodyA);
            r.body_b = index(s->bodyB);
            r.serial = s->serial;
            r.flags = s->disabled ? AVBD_FORCE_DISABLED : 0u;
            put3(r.r_a, s->rA);
            put3(r.r_b, s->rB);
            r.rest_length = s->restLength;
            const ConstraintPointProperties& p = s->constraintPoints[0];
            r.stiffness = p.stiffness; r.fracture = p.fracture; r.fmin = p.fminMagnitude; r.fmax = p.fmaxMagnitude;
            r.lambda = p.lambda; r.penalty = p.penalty;
            out.springs.push_back(r);
        } else if (const Manifold* m = dynamic_cast<const Manifold*>(f)) {
            mans.emplace_back(m->serial, m);
        } else {
            throw std::runtime_error("Solver: the device step supports Joint, Spring and Manifold forces only");
        }
    }
    // forces in creation order (solverForces is swap-with-last ordered)
    std::sort(out.joints.begin(), out.joints.end(), [](const avbd_joint& a, const avbd_joint& b) { return a.serial < b.serial; });
    std::sort(out.springs.begin(), out.springs.end(), [](const avbd_spring& a, const avbd_spring& b) { return a.serial < b.serial; });
    std::sort(mans.begin(), mans.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (const auto& pm : mans) {
        const Manifold* m = pm.second;
        avbd_manifold r{};
        r.body_a = index(m->bodyA);
        r.body_b = index(m->bodyB);
        r.serial = m->serial;
        r.n_contacts = m->numContactPoints;
        r.friction = m->friction;
        for (int c = 0; c < m->numContactPoints && c < 8; ++c) {
            const ContactPoint& cp = m->contactPoints[c];
            avbd_contact& ct = r.contacts[c];
            ct.id = cp.id;
            ct.stick = m->contactInfos[c].stick ? 1 : 0;
            put3(ct.r_a, cp.rA);
            put3(ct.r_b, cp.rB);
            put3(ct.normal, cp.normal);
            ct.penetration = cp.penetration;
            for (int k = 0; k < 3; ++k) {
                ct.lambda[k] = m->constraintPoints[3 * c + k].lambda;
                ct.penalty[k] = m->constraintPoints[3 * c + k].penalty;
            }
        }
        out.manifolds.push_back(r);
    }
    for (const auto& ep : solverEnergies) {
        const Energy* e = ep.get();
        if (const NeoHookeanFEM* t = dynamic_cast<const NeoHookeanFEM*>(e)) {
            avbd_tet r{};
            r.v[0] = t->body0->solverIndex; r.v[1] = t->body1->solverIndex;
            r.v[2] = t->body2->solverIndex; r.v[3] = t->body3->solverIndex;
            r.serial = t->serial;
            std::memcpy(r.dm_inv, t->DmInv, sizeof(r.dm_inv));
            r.rest_volume = t->restVolume;
            r.mu = t->lameMu;
            r.lambda = t->lameLambda;
            out.tets.push_back(r);
        } else if (const STVKFEM* t = dynamic_cast<const STVKFEM*>(e)) {
            avbd_tri r{};
            r.v[0] = t->body0->solverIndex; r.v[1] = t->body1->solverIndex; r.v[2] = t->body2->solverIndex;
            r.serial = t->serial;
            std::memcpy(r.dm_inv, t->DmInv2D, sizeof(r.dm_inv));
            r.rest_area = t->restArea;
            r.mu = t->lameMu;
            r.lambda = t->lameLambda;
            out.tris.push_back(r);
        } else {
            throw std::runtime_error("Solver: the device step supports NeoHookeanFEM and STVKFEM energies only");
        }
    }
    std::sort(out.tets.begin(), out.tets.end(), [](const avbd_tet& a, const avbd_tet& b) { return a.serial < b.serial; });
    std::sort(out.tris.begin(), out.tris.end(), [](const avbd_tri& a, const avbd_tri& b) { return a.serial < b.serial; });
}

void Solver::Upload() {
    WorldRecords rec;
    ExportWorld(rec);
    avbd_world w = rec.view();
    EnsureContext();
    check(avbd_set_option(ctx_, "spare_bodies", spareBodies), "avbd_set_option");
    check(avbd_upload(ctx_, &w, 0), "avbd_upload");
    ++uploads;
    pendingAdds_.clear();
    uploaded_ = true;
    structureDirty_ = false;
    lastScene_.assign(reinterpret_cast<const unsigned char*>(&rec.scene),
                      reinterpret_cast<const unsigned char*>(&rec.scene) + sizeof(rec.scene));
    lastBodies_.assign(reinterpret_cast<const unsigned char*>(rec.bodies.data()),
                       reinterpret_cast<const unsigned char*>(rec.bodies.data() + rec.bodies.size()));
}

// ---------------------------------------------------------------- step
void Solver::Step() {
    if (emergencyStop) return;
    RebuildPtrCaches();
    EnsureContext();
    std::vector<unsigned char> sc;
    SceneRecord(sc);
    if (sc != lastScene_) structureDirty_ = true;
    std::vector<avbd_body> bodies;
    BodyRecords(bodies);
    if (!structureDirty_ && !pendingAdds_.empty()) {
        // bodies added since the last step are the last ones
        const size_t n_old = lastBodies_.size() / sizeof(avbd_body);
        if (n_old + pendingAdds_.size() == bodies.size()) {
            const int rc = avbd_add_bodies(ctx_, bodies.data() + n_old, (int32_t)pendingAdds_.size());
            if (rc == AVBD_E_CAPACITY) {
                structureDirty_ = true;  // no free slot: upload the world
            } else {
                check(rc, "avbd_add_bodies");
                lastBodies_.insert(lastBodies_.end(), reinterpret_cast<const unsigned char*>(bodies.data() + n_old),
                                   reinterpret_cast<const unsigned char*>(bodies.data() + bodies.size()));
            }
        } else {
            structureDirty_ = true;
        }
    }
    pendingAdds_.clear();
    bool state_changed = false;
    if (!structureDirty_) {
        const avbd_body* prev = reinterpret_cast<const avbd_body*>(lastBodies_.data());
        if (lastBodies_.size() != bodies.size() * sizeof(avbd_body)) {
            structureDirty_ = true;
        } else {
            for (size_t i = 0; i < bodies.size(); ++i) {
                if (!same_structure(bodies[i], prev[i])) {
                    structureDirty_ = true;
                    break;
                }
                state_changed = state_changed || !same_state(bodies[i], prev[i]);
            }
        }
    }
    if (structureDirty_) {
        if (forcesStale_) SyncForces();
        Upload();
        dragOnDevice_ = false;  // a fresh upload starts undragged
    } else if (state_changed) {
        check(avbd_upload_bodies(ctx_, bodies.data(), (int32_t)bodies.size()), "avbd_upload_bodies");
    }
    // Mesh::isDragged / addedDragVelocity (geometry.hpp:193-194): the velocity
    // update adds a dragged body's added velocity once (solver.cpp:506-510)
    bool dragging = false;
    for (Mesh* m : bodyPtrs) dragging = dragging || m->isDragged;
    if (dragging || dragOnDevice_) {
        const int32_t n = (int32_t)bodyPtrs.size();
        std::vector<int32_t> idx(n), flag(n);
        std::vector<float> vel(3 * (size_t)n);
        for (int32_t i = 0; i < n; ++i) {
            const Mesh* m = bodyPtrs[i];
            idx[i] = i;
            flag[i] = m->isDragged ? 1 : 0;
            for (int k = 0; k < 3; ++k) vel[3 * i + k] = m->isDragged ? m->addedDragVelocity[k] : 0.f;
        }
        check(avbd_set_drag(ctx_, idx.data(), vel.data(), flag.data(), n), "avbd_set_drag");
        dragOnDevice_ = dragging;
    }
    check(avbd_set_option(ctx_, "debug_output", debugOutput ? 1 : 0), "avbd_set_option");
    if (debugOutput) {
        std::vector<float> rgb(3 * solverBodies.size());
        for (size_t i = 0; i < solverBodies.size(); ++i)
            for (int k = 0; k < 3; ++k) rgb[3 * i + k] = solverBodies[i]->color[k];
        check(avbd_set_body_colours(ctx_, rgb.data(), (int32_t)solverBodies.size()), "avbd_set_body_colours");
    }
    check(avbd_step(ctx_, 1), "avbd_step");
    if (dragging)  // the step's velocity update consumed the added velocities
        for (Mesh* m : bodyPtrs)
            if (m->isDragged && !m->isStatic) m->addedDragVelocity = Vector3f::Zero();
    lineData.clear();
    debugPointData.clear();
    if (debugOutput) {
        static_assert(sizeof(GPULineData) == sizeof(avbd_line), "line record layouts");
        static_assert(sizeof(GPUDebugPointData) == sizeof(avbd_debug_point), "point record layouts");
        int64_t nl = 0, np = 0;
        check(avbd_get_debug_lines(ctx_, nullptr, 0, &nl), "avbd_get_debug_lines");
        lineData.resize((size_t)nl);
        check(avbd_get_debug_lines(ctx_, reinterpret_cast<avbd_line*>(lineData.data()), nl, &nl),
              "avbd_get_debug_lines");
        check(avbd_get_debug_points(ctx_, nullptr, 0, &np), "avbd_get_debug_points");
        debugPointData.resize((size_t)np);
        check(avbd_get_debug_points(ctx_, reinterpret_cast<avbd_debug_point*>(debugPointData.data()), np, &np),
              "avbd_get_debug_points");
    }
    int32_t em = 0, removed = 0;
    check(avbd_get_status(ctx_, &em, &removed, 1), "avbd_get_status");
    if (trustRegionUpdate && projectionMode == EigenProjectionMode::ADAPTIVE)
        check(avbd_get_trust_region_rho(ctx_, &trustRegionRho, 1), "avbd_get_trust_region_rho");
    emergencyStop = em != 0;
    ApplyRemovals();
    ReadBack();
    if (syncForces) SyncForces();
    else forcesStale_ = true;
    // CheckExplosion returns from Step before the timing update
    // (solver.cpp:528-529): the stopping step is not accumulated
    if (!emergencyStop) AccumulateTimings();
}

// Bodies the step removed for leaving the bounds: the same RemoveBody calls
// in the same order (their forces and energies go with them).
void Solver::ApplyRemovals() {
    int32_t n = 0;
    check(avbd_last_removed(ctx_, nullptr, nullptr, 0, &n), "avbd_last_removed");
    if (n == 0) return;
    std::vector<int32_t> scene(n), index(n);
    check(avbd_last_removed(ctx_, scene.data(), index.data(), n, &n), "avbd_last_removed");
    mirroring_ = true;
    for (int k = 0; k < n; ++k) {
        if (index[k] < 0 || index[k] >= (int)bodyPtrs.size()) {
            mirroring_ = false;
            throw std::runtime_error("Solver::Step: removed body index out of range");
        }
        RemoveBody(bodyPtrs[index[k]]);
    }
    mirroring_ = false;
}

void Solver::ReadBack() {
    std::vector<avbd_body> out(solverBodies.size());
    check(avbd_download_bodies(ctx_, out.data(), (int32_t)out.size()), "avbd_download_bodies");
    for (size_t i = 0; i < out.size(); ++i) {
        Mesh& m = *solverBodies[i];
        const avbd_body& b = out[i];
        m.transform.SetPosition(get3(b.pos));
        m.transform.SetRotation(Quaternionf(b.rot[0], b.rot[1], b.rot[2], b.rot[3]));
        m.velocity = get3(b.vel);
        m.angularVelocity = get3(b.ang_vel);
        m.prevVelocity = get3(b.prev_vel);
        m.prevAngularVelocity = get3(b.prev_ang_vel);
    }
    std::vector<avbd_body> now;
    BodyRecords(now);
    lastBodies_.assign(reinterpret_cast<const unsigned char*>(now.data()),
                       reinterpret_cast<const unsigned char*>(now.data() + now.size()));
}

void Solver::SyncForces() {
    if (!ctx_ || !uploaded_) return;
    avbd_world cnt{};
    check(avbd_get_counts(ctx_, &cnt), "avbd_get_counts");
    std::vector<avbd_scene> sc(cnt.n_scenes);
    std::vector<avbd_body> bodies(cnt.n_bodies);
    std::vector<avbd_joint> joints(cnt.n_joints);
    std::vector<avbd_spring> springs(cnt.n_springs);
    std::vector<avbd_manifold> mans(cnt.n_manifolds);
    std::vector<avbd_tet> tets(cnt.n_tets);
    std::vector<avbd_tri> tris(cnt.n_tris);
    avbd_world w = cnt;
    w.scenes = sc.data(); w.bodies = bodies.data(); w.joints = joints.data(); w.springs = springs.data();
    w.manifolds = mans.data(); w.tets = tets.data(); w.tris = tris.data();
    check(avbd_download(ctx_, &w), "avbd_download");
    if (cnt.n_bodies != (int)solverBodies.size())
        throw std::runtime_error("Solver::SyncForces: device and host body counts differ");
    RebuildPtrCaches();
    std::map<int64_t, Force*> by_serial;
    for (auto& f : solverForces) by_serial[f->serial] = f.get();
    std::map<int64_t, bool> seen;
    int64_t max_serial = -1;
    for (const avbd_joint& r : joints) {
        max_serial = std::max<int64_t>(max_serial, r.serial);
        auto it = by_serial.find(r.serial);
        if (it == by_serial.end()) continue;
        seen[r.serial] = true;
        Force* f = it->second;
        f->disabled = (r.flags & AVBD_FORCE_DISABLED) != 0;
        for (int k = 0; k < 6; ++k) {
            f->constraintPoints[k].lambda = r.lambda[k];
            f->constraintPoints[k].penalty = r.penalty[k];
        }
    }
    for (const avbd_spring& r : springs) {
        max_serial = std::max<int64_t>(max_serial, r.serial);
        auto it = by_serial.find(r.serial);
        if (it == by_serial.end()) continue;
        seen[r.serial] = true;
        Force* f = it->second;
        f->disabled = (r.flags & AVBD_FORCE_DISABLED) != 0;
        f->constraintPoints[0].lambda = r.lambda;
        f->constraintPoints[0].penalty = r.penalty;
    }
    mirroring_ = true;
    for (const avbd_manifold& r : mans) {
        max_serial = std::max<int64_t>(max_serial, r.serial);
        Manifold* m = nullptr;
        auto it = by_serial.find(r.serial);
        if (it != by_serial.end()) {
            m = dynamic_cast<Manifold*>(it->second);
        } else {
            // a contact the device's broad phase created
            auto owned = std::make_unique<Manifold>(this, bodyPtrs.at(r.body_a), bodyPtrs.at(r.body_b));
            m = owned.get();
            m->serial = r.serial;
            AddForce(std::move(owned));
        }
        if (!m) continue;
        seen[r.serial] = true;
        set_contacts(m, r);
    }
    // forces the device retired (contacts without points, fractured springs
    // and joints, forces of removed bodies)
    std::vector<Force*> gone;
    for (auto& f : solverForces)
        if (!seen.count(f->serial)) gone.push_back(f.get());
    for (Force* f : gone) RemoveForce(f);
    mirroring_ = false;
    nextForceSerial_ = std::max<int64_t>(nextForceSerial_, max_serial + 1);
    forcesStale_ = false;
    RebuildPtrCaches();
}

void Solver::AccumulateTimings() {
    avbd_timings t{};
    check(avbd_get_timings(ctx_, &t), "avbd_get_timings");
    SolverTimings s;
    s.broadPhaseMs = t.broad_phase_ms;
    s.warmstartMs = t.warmstart_ms;
    s.predictionMs = t.prediction_ms;
    s.primalDualMs = t.primal_dual_ms;
    s.velocityUpdateMs = t.velocity_update_ms;
    s.postStabMs = t.post_stab_ms;
    s.totalSubstepMs = t.total_ms;
    // rolling mean over the last 60 steps (solver.cpp:599-626)
    timingAccumulator.push_back(s);
    if ((int)timingAccumulator.size() > TIMING_WINDOW) timingAccumulator.erase(timingAccumulator.begin());
    SolverTimings avg;
    for (const SolverTimings& x : timingAccumulator) {
        avg.broadPhaseMs += x.broadPhaseMs;
        avg.warmstartMs += x.warmstartMs;
        avg.predictionMs += x.predictionMs;
        avg.primalDualMs += x.primalDualMs;
        avg.velocityUpdateMs += x.velocityUpdateMs;
        avg.postStabMs += x.postStabMs;
        avg.totalSubstepMs += x.totalSubstepMs;
    }
    const float n = float(timingAccumulator.size());
    timings.broadPhaseMs = avg.broadPhaseMs / n;
    timings.warmstartMs = avg.warmstartMs / n;
    timings.predictionMs = avg.predictionMs / n;
    timings.primalDualMs = avg.primalDualMs / n;
    timings.velocityUpdateMs = avg.velocityUpdateMs / n;
    timings.postStabMs = avg.postStabMs / n;
    timings.totalSubstepMs = avg.totalSubstepMs / n;
    averageStepTime = timings.totalSubstepMs;
}
