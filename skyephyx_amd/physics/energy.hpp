// energy.hpp -- elastic energies of the solver facade: Neo-Hookean
// tetrahedra and StVK membrane triangles over particles, with the
// reference's names and constructors (src/physics/energy.hpp:18-121,
// neohookeanfem.cpp:5-37, stvkfem.cpp:5-35).  The energy terms are
// evaluated only on the device.
#pragma once

#include <cstdint>
#include <vector>

#include "geometry.hpp"

enum class EigenProjectionMode { CLAMP = 0, ABSOLUTE = 1, ADAPTIVE = 2 };

// Base of every energy: the constructor appends it to each linked body's
// `energies` list (energy.cpp:5-9), the destructor erases it.
struct Energy {
    Energy(Solver* solver, std::vector<Mesh*> linkedBodies);
    virtual ~Energy();

    int solverIndex = -1;
    std::vector<Mesh*> linkedBodies;
    Solver* solver;
    int32_t serial = -1;  // creation order: a vertex visits its energies in increasing serial
    virtual int numBodies() const = 0;
};

// Linear tetrahedron: rest shape inverse Dm^-1, rest volume |det Dm| / 6,
// Lame parameters from Young's modulus and Poisson's ratio.
struct NeoHookeanFEM : Energy {
    NeoHookeanFEM(Solver* solver, Mesh* body0, Mesh* body1, Mesh* body2, Mesh* body3, float youngsModulus,
                  float poissonRatio);
    Mesh* body0;
    Mesh* body1;
    Mesh* body2;
    Mesh* body3;
    float restVolume = 0.f;
    float lameMu = 0.f;
    float lameLambda = 0.f;
    float poissonRatio;
    float youngsModulus;
    float DmInv[9] = {0.f};  // row-major
    int numBodies() const override { return 4; }
};

// Triangle membrane: 2-D rest frame (edge 0 along x), Dm^-1 of that frame,
// rest area.
struct STVKFEM : Energy {
    STVKFEM(Solver* solver, Mesh* body0, Mesh* body1, Mesh* body2, float youngsModulus, float poissonRatio);
    Mesh* body0;
    Mesh* body1;
    Mesh* body2;
    float restArea = 0.f;
    float lameMu = 0.f;
    float lameLambda = 0.f;
    float poissonRatio;
    float youngsModulus;
    float DmInv2D[4] = {0.f};  // row-major
    int numBodies() const override { return 3; }
};
