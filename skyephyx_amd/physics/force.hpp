// force.hpp -- constraints of the solver facade: Force, Spring, Joint and
// the contact Manifold, with the reference's names, constructors and row
// state (src/physics/force.hpp:9-198).
//
// The row state a caller reads or seeds (stiffness, fracture, bounds,
// penalty, lambda) lives in constraintPoints as in the reference; the
// Jacobians, Hessians and C are computed on the device every iteration and
// are not mirrored.  The constraint evaluation itself (Initialize,
// ComputeConstraints, ComputeDerivatives) runs only on the device.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "geometry.hpp"

struct ConstraintPointProperties {
    float fminMagnitude = -INFINITY;
    float fmaxMagnitude = INFINITY;
    float stiffness = INFINITY;
    float fracture = INFINITY;
    float penalty = 0.f;
    float lambda = 0.f;
    float C = 0.f;
};

// Base of every force.  The constructor appends the force to each linked
// body's `forces` list (force.cpp:5-9) and takes the next creation serial;
// the destructor erases it from them (force.cpp:12-19).
struct Force {
    Force(Solver* solver, std::vector<Mesh*> linkedBodies, int numConstraintPoints);
    virtual ~Force();

    int solverIndex = -1;
    std::vector<Mesh*> linkedBodies;
    Solver* solver;
    ConstraintPointProperties constraintPoints[24];

    bool isManifold = false;
    bool includeHessian = true;
    bool markedForRemoval = false;
    bool disabled = false;

    // Creation order: a body visits its forces in increasing serial (the
    // append order of Mesh::forces).  Shared with the contacts the device
    // creates.
    int64_t serial = -1;

    virtual int numConstraints() const = 0;
    virtual int numBodies() const = 0;

    void Disable() {
        disabled = true;
        ResetValues();
    }
    void ResetValues() {
        for (int i = 0; i < numConstraints(); ++i) {
            constraintPoints[i].lambda = 0.f;
            constraintPoints[i].penalty = 0.f;
            constraintPoints[i].C = 0.f;
        }
    }
};

// True if a force links mesh and other (force.hpp:81-89).
bool isConstrainedTo(Mesh* mesh, Mesh* other);

// Distance spring between two anchor points (spring.cpp:13-38).  bodyA may
// be null (anchored in the world at rA); restLength < 0 takes the current
// distance; a rope pulls only (fmin = 0).
struct Spring : Force {
    static constexpr int NUM_CONSTRAINTS = 1;
    Spring(Solver* solver, Mesh* bodyA, const Vector3f& rA, Mesh* bodyB, const Vector3f& rB, float stiffness,
           float restLength = -1.f, bool isRope = false);
    Vector3f rA, rB;
    Mesh* bodyA = nullptr;
    Mesh* bodyB = nullptr;
    float restLength;
    bool isRope;
    int numConstraints() const override { return 1; }
    int numBodies() const override { return 2; }
};

// Six-row joint: three position rows, three angle rows scaled by the torque
// arm (joint.cpp:24-66).  bodyA may be null (world anchor).
struct Joint : Force {
    static constexpr int NUM_CONSTRAINTS = 6;
    Joint(Solver* solver, Mesh* bodyA, const Vector3f& rA, Mesh* bodyB, const Vector3f& rB,
          const Vector6f& stiffness = Vector6f::Constant(INFINITY),
          const Vector6f& fractures = Vector6f::Constant(INFINITY));
    Vector3f rA, rB;
    Mesh* bodyA = nullptr;
    Mesh* bodyB = nullptr;
    Quaternionf restRotation;
    float torqueArm = 1.f;
    int numConstraints() const override { return NUM_CONSTRAINTS; }
    int numBodies() const override { return 2; }
};

static constexpr float COLLISION_MARGIN = 0.0005f;
static constexpr float STICK_THRESHOLD = 0.01f;

// One contact point (collision.hpp:8-18): body-local unscaled lever arms,
// normal from A to B, penetration, feature id.
struct ContactPoint {
    Vector3f position;
    Vector3f normal;
    float penetration = 0.f;
    Vector3f rA;
    Vector3f rB;
    uint32_t id = 0;
};

struct ManifoldContactInfo {
    bool stick = false;
};

// Box-box contact between two bodies (manifold.cpp:16-29): three rows per
// contact point (normal, two tangents), the normal compressive only.
// Manifolds are created and retired by the device's broad phase and
// warm start; the facade mirrors them after each step.
struct Manifold : Force {
    static constexpr int NUM_CONSTRAINTS = 24;
    Manifold(Solver* solver, Mesh* bodyA, Mesh* bodyB);
    ContactPoint contactPoints[8];
    ManifoldContactInfo contactInfos[8];
    int numContactPoints = 0;
    float friction = 0.f;
    Mesh* bodyA = nullptr;
    Mesh* bodyB = nullptr;
    int numConstraints() const override { return numContactPoints * 3; }
    int numBodies() const override { return 2; }
};
