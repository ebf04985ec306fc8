// elements.cpp -- bodies, forces and energies of the solver facade: their
// constructors with the reference's set-up arithmetic, and the
// registration of forces and energies in their bodies' lists.
//
//   Transform::TransformPoint   geometry.hpp:64-67
//   Mesh::~Mesh                 geometry.cpp:11-24
//   Force / ~Force              force.cpp:5-19
//   Spring                      spring.cpp:13-38
//   Joint                       joint.cpp:24-66
//   Manifold                    manifold.cpp:16-29
//   Energy / ~Energy            energy.cpp:5-19
//   NeoHookeanFEM               neohookeanfem.cpp:5-37
//   STVKFEM                     stvkfem.cpp:5-35
//
// Rest shapes are computed in double precision and rounded once to float
// (the reference computes them in float with Eigen; the records differ by at
// most an ulp or two and are the solver's input either way).
#include <algorithm>
#include <cassert>
#include <cmath>

#include "solver.hpp"

namespace {

// Rotation matrix of a unit quaternion (w, x, y, z) with the reference's
// toRotationMatrix arithmetic.
void rotation_matrix(const Quaternionf& q, float R[3][3]) {
    const float w = q.w(), x = q.x(), y = q.y(), z = q.z();
    const float tx = 2.f * x, ty = 2.f * y, tz = 2.f * z;
    const float twx = tx * w, twy = ty * w, twz = tz * w;
    const float txx = tx * x, txy = ty * x, txz = tz * x;
    const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
    R[0][0] = 1.f - (tyy + tzz); R[0][1] = txy - twz;         R[0][2] = txz + twy;
    R[1][0] = txy + twz;         R[1][1] = 1.f - (txx + tzz); R[1][2] = tyz - twx;
    R[2][0] = txz - twy;         R[2][1] = tyz + twx;         R[2][2] = 1.f - (txx + tyy);
}

float dot3(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

std::vector<Mesh*> non_null(std::initializer_list<Mesh*> bodies) {
    std::vector<Mesh*> r;
    for (Mesh* m : bodies)
        if (m) r.push_back(m);
    return r;
}

void lame(float E, float nu, float* mu, float* lambda) {
    *mu = E / (2.f * (1.f + nu));
    *lambda = E * nu / ((1.f + nu) * (1.f - 2.f * nu));
}

}  // namespace

Vector3f Transform::TransformPoint(const Vector3f& p) const {
    float R[3][3];
    rotation_matrix(rotation, R);
    const float sp[3] = {scale[0] * p[0], scale[1] * p[1], scale[2] * p[2]};
    return Vector3f(dot3(R[0], sp) + position[0], dot3(R[1], sp) + position[1], dot3(R[2], sp) + position[2]);
}

Mesh::Mesh(Solver* solver, ModelType modelType, const Vector3f& color)
    : modelType(modelType), color(color), solver(solver) {}

Mesh::~Mesh() {
    auto fs = forces;
    for (Force* f : fs) solver->RemoveForce(f);
    auto es = energies;
    for (Energy* e : es) solver->RemoveEnergy(e);
}

Force::Force(Solver* solver, std::vector<Mesh*> linkedBodies, int)
    : linkedBodies(std::move(linkedBodies)), solver(solver) {
    serial = solver->TakeForceSerial();
    for (Mesh* m : this->linkedBodies) m->forces.push_back(this);
}

Force::~Force() {
    for (Mesh* m : linkedBodies) {
        auto& f = m->forces;
        f.erase(std::remove(f.begin(), f.end(), this), f.end());
    }
}

bool isConstrainedTo(Mesh* mesh, Mesh* other) {
    for (Force* f : mesh->forces)
        if (std::find(f->linkedBodies.begin(), f->linkedBodies.end(), other) != f->linkedBodies.end()) return true;
    return false;
}

Spring::Spring(Solver* solver, Mesh* bodyA, const Vector3f& rA, Mesh* bodyB, const Vector3f& rB, float stiffness,
               float restLength, bool isRope)
    : Force(solver, non_null({bodyA, bodyB}), NUM_CONSTRAINTS), rA(rA), rB(rB), bodyA(bodyA), bodyB(bodyB),
      restLength(restLength), isRope(isRope) {
    assert(bodyB && "a spring must be connected to at least one body");
    constraintPoints[0].stiffness = stiffness;
    constraintPoints[0].fminMagnitude = isRope ? 0.f : -INFINITY;
    constraintPoints[0].fmaxMagnitude = INFINITY;
    if (this->restLength < 0.f) {
        const Vector3f a = bodyA ? bodyA->transform.TransformPoint(rA) : rA;
        const Vector3f b = bodyB->transform.TransformPoint(rB);
        const float d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        this->restLength = std::sqrt(dot3(d, d));
    }
}

Joint::Joint(Solver* solver, Mesh* bodyA, const Vector3f& rA, Mesh* bodyB, const Vector3f& rB,
             const Vector6f& stiffness, const Vector6f& fractures)
    : Force(solver, non_null({bodyA, bodyB}), NUM_CONSTRAINTS), rA(rA), rB(rB), bodyA(bodyA), bodyB(bodyB) {
    assert(bodyB && "a joint must connect to at least one body");
    includeHessian = true;
    for (int i = 0; i < NUM_CONSTRAINTS; ++i) {
        constraintPoints[i].stiffness = stiffness[i];
        constraintPoints[i].fracture = fractures[i];
        constraintPoints[i].fminMagnitude = -fractures[i];
        constraintPoints[i].fmaxMagnitude = fractures[i];
    }
    const Quaternionf qA = bodyA ? bodyA->transform.GetRotation() : Quaternionf::Identity();
    const Quaternionf qB = bodyB->transform.GetRotation();
    const Quaternionf ca(qA.w(), -qA.x(), -qA.y(), -qA.z());
    // Hamilton product conj(qA) * qB
    const float aw = ca.w(), ax = ca.x(), ay = ca.y(), az = ca.z();
    const float bw = qB.w(), bx = qB.x(), by = qB.y(), bz = qB.z();
    restRotation = Quaternionf(((aw * bw - ax * bx) - ay * by) - az * bz, ((aw * bx + ax * bw) + ay * bz) - az * by,
                               ((aw * by - ax * bz) + ay * bw) + az * bx, ((aw * bz + ax * by) - ay * bx) + az * bw);
    const Vector3f sA = bodyA ? bodyA->transform.GetScale() : Vector3f::Zero();
    const Vector3f sB = bodyB->transform.GetScale();
    const float s[3] = {sA[0] + sB[0], sA[1] + sB[1], sA[2] + sB[2]};
    torqueArm = dot3(s, s);
    if (torqueArm < 1e-6f) torqueArm = 1.f;
}

Manifold::Manifold(Solver* solver, Mesh* bodyA, Mesh* bodyB)
    : Force(solver, non_null({bodyA, bodyB}), NUM_CONSTRAINTS), bodyA(bodyA), bodyB(bodyB) {
    assert(bodyA && bodyB && "a manifold must be connected to two bodies");
    for (int i = 0; i < NUM_CONSTRAINTS; i += 3) {
        constraintPoints[i].fmaxMagnitude = 0.f;  // the normal row is compressive only
        constraintPoints[i].fminMagnitude = -INFINITY;
    }
    isManifold = true;
    includeHessian = false;
}

Energy::Energy(Solver* solver, std::vector<Mesh*> linkedBodies) : linkedBodies(std::move(linkedBodies)), solver(solver) {
    serial = solver->TakeEnergySerial();
    for (Mesh* m : this->linkedBodies) m->energies.push_back(this);
}

Energy::~Energy() {
    for (Mesh* m : linkedBodies) {
        auto& e = m->energies;
        e.erase(std::remove(e.begin(), e.end(), this), e.end());
    }
}

NeoHookeanFEM::NeoHookeanFEM(Solver* solver, Mesh* body0, Mesh* body1, Mesh* body2, Mesh* body3, float youngsModulus,
                             float poissonRatio)
    : Energy(solver, {body0, body1, body2, body3}), body0(body0), body1(body1), body2(body2), body3(body3),
      poissonRatio(poissonRatio), youngsModulus(youngsModulus) {
    lame(youngsModulus, poissonRatio, &lameMu, &lameLambda);
    const Vector3f p0 = body0->transform.GetPosition(), p1 = body1->transform.GetPosition(),
                   p2 = body2->transform.GetPosition(), p3 = body3->transform.GetPosition();
    double D[3][3];  // columns p1 - p0, p2 - p0, p3 - p0
    for (int r = 0; r < 3; ++r) {
        D[r][0] = double(p1[r]) - double(p0[r]);
        D[r][1] = double(p2[r]) - double(p0[r]);
        D[r][2] = double(p3[r]) - double(p0[r]);
    }
    const double c00 = D[1][1] * D[2][2] - D[1][2] * D[2][1];
    const double c01 = D[1][2] * D[2][0] - D[1][0] * D[2][2];
    const double c02 = D[1][0] * D[2][1] - D[1][1] * D[2][0];
    const double det = D[0][0] * c00 + D[0][1] * c01 + D[0][2] * c02;
    const double inv[3][3] = {
        {c00 / det, (D[0][2] * D[2][1] - D[0][1] * D[2][2]) / det, (D[0][1] * D[1][2] - D[0][2] * D[1][1]) / det},
        {c01 / det, (D[0][0] * D[2][2] - D[0][2] * D[2][0]) / det, (D[0][2] * D[1][0] - D[0][0] * D[1][2]) / det},
        {c02 / det, (D[0][1] * D[2][0] - D[0][0] * D[2][1]) / det, (D[0][0] * D[1][1] - D[0][1] * D[1][0]) / det},
    };
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) DmInv[r * 3 + c] = float(inv[r][c]);
    restVolume = float(std::fabs(det) / 6.0);
}

STVKFEM::STVKFEM(Solver* solver, Mesh* body0, Mesh* body1, Mesh* body2, float youngsModulus, float poissonRatio)
    : Energy(solver, {body0, body1, body2}), body0(body0), body1(body1), body2(body2), poissonRatio(poissonRatio),
      youngsModulus(youngsModulus) {
    lame(youngsModulus, poissonRatio, &lameMu, &lameLambda);
    const Vector3f p0 = body0->transform.GetPosition(), p1 = body1->transform.GetPosition(),
                   p2 = body2->transform.GetPosition();
    double e0[3], e1[3];
    for (int r = 0; r < 3; ++r) {
        e0[r] = double(p1[r]) - double(p0[r]);
        e1[r] = double(p2[r]) - double(p0[r]);
    }
    const double x1 = std::sqrt(e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2]);
    const double x2 = (e1[0] * e0[0] + e1[1] * e0[1] + e1[2] * e0[2]) / x1;
    const double y2 = std::sqrt(std::max(0.0, (e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]) - x2 * x2));
    // Dm = [[x1, x2], [0, y2]]
    DmInv2D[0] = float(1.0 / x1);
    DmInv2D[1] = float(-x2 / (x1 * y2));
    DmInv2D[2] = 0.f;
    DmInv2D[3] = float(1.0 / y2);
    restArea = float(0.5 * x1 * y2);
}
