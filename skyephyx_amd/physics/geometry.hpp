// geometry.hpp -- rigid bodies and particles of the solver facade (Mesh,
// Transform), with the reference's names and meanings
// (src/helpers/geometry.hpp:17-22, 64-256).  Rendering state (vertex
// buffers, model matrices) is not part of the headless facade.
#pragma once

#include <string>
#include <vector>

#include "linalg.hpp"

class Solver;
struct Force;
struct Energy;

enum ModelType { ModelType_Cube = 0, ModelType_Sphere = 1, ModelType_Pyramid = 2 };

// Debug output records of the reference's renderer (geometry.hpp:48-60).
struct GPULineData {
    float start[3];
    float end[3];
    float color[4];
};
struct GPUDebugPointData {
    float position[3];
    float _pad;
    float color[4];
};

// Position, rotation and non-uniform scale of a body (geometry.hpp:64-145).
class Transform {
public:
    Transform() : position(Vector3f::Zero()), rotation(Quaternionf::Identity()), scale(Vector3f::Ones()) {}
    // R(q) (s . p) + x
    Vector3f TransformPoint(const Vector3f& p) const;
    void SetPosition(const Vector3f& p) { position = p; }
    void SetRotation(const Quaternionf& q) { rotation = q; }
    void SetScale(const Vector3f& s) { scale = s; }
    Vector3f GetPosition() const { return position; }
    Quaternionf GetRotation() const { return rotation; }
    Vector3f GetScale() const { return scale; }

private:
    Vector3f position;
    Quaternionf rotation;
    Vector3f scale;
};

// A rigid box or particle (geometry.hpp:148-256).  `forces` and `energies`
// list what acts on the body in creation order; the constructors of Force
// and Energy append to them and their destructors erase from them.
struct Mesh {
    Mesh(Solver* solver, ModelType modelType, const Vector3f& color = Vector3f(1.f, 1.f, 1.f));
    ~Mesh();  // removes the body's forces and energies from the solver

    int solverIndex = -1;
    Transform transform;
    std::vector<Force*> forces;
    std::vector<Energy*> energies;
    ModelType modelType;
    Vector3f color;
    Solver* solver;
    std::string name = "Mesh";

    float mass = 0.f;
    float density = 0.f;
    float friction = 0.f;
    bool isStatic = false;
    Matrix3f inertiaTensorBody = Matrix3f::Identity();  // diagonal in body frame
    Matrix3f inertiaTensorBodyInv = Matrix3f::Identity();

    Vector3f velocity = Vector3f::Zero();
    Vector3f prevVelocity = Vector3f::Zero();
    Vector3f angularVelocity = Vector3f::Zero();
    Vector3f prevAngularVelocity = Vector3f::Zero();

    float detectionRadius = 0.f;
    bool isDragged = false;
    Vector3f addedDragVelocity = Vector3f::Zero();
    bool isParticle = false;
};
