// levels.hpp -- scene builders over the solver facade: the reference's
// GeometryHelpers::makeTetVolume (src/helpers/geometry.cpp:30-94) and
// parameterised versions of its levels (src/levels.h) that the tests and
// benchmarks use.  Every builder only calls the public Solver API
// (AddBody, AddParticle, AddForce, AddEnergy), as a level does.
#pragma once

#include <functional>
#include <vector>

#include "solver.hpp"

namespace GeometryHelpers {

// nx*ny*nz particles on a grid of spacing cellSize from (ox, oy, oz), five
// Neo-Hookean tets per cell; pinMinX makes the i = 0 face static.  Returns
// the particles, index i + j*nx + k*nx*ny.
std::vector<Mesh*> makeTetVolume(Solver* solver, float ox, float oy, float oz, int nx, int ny, int nz, float cellSize,
                                 float E, float nu, const Vector3f& color, float particleMass = 1.f,
                                 float friction = 0.5f, bool pinMinX = false);
// The same with any set of static vertices: pin(i, j, k).
std::vector<Mesh*> makeTetVolume(Solver* solver, double ox, double oy, double oz, int nx, int ny, int nz,
                                 double cellSize, float E, float nu, float particleMass, float friction,
                                 const std::function<bool(int, int, int)>& pin);

}  // namespace GeometryHelpers

namespace Levels {

// Static ground box of every level (levels.h:30).
Mesh* ground(Solver* s);
// 2-D pyramid of unit boxes, rows, rows-1, ..., 1, `gap` apart, each row
// `lift` above the one below (alpha = gamma = 0.99).
void pyramid2D(Solver* s, int rows = 10, double gap = 0.1, double lift = 0.02);
// The JointPlayground pendulum (levels.h:203-233): ball-jointed links, the
// first anchored in the world.
void jointChain(Solver* s, int nLinks = 5);
// The rope bridge of JointPlayground (levels.h:235-300) with a weight;
// springs of link `fractureLink` break above 2000 N.
void springBridge(Solver* s, int nLinks = 6, int fractureLink = 3);
// StVK membrane as in ClothSimulation (levels.h:442-491), two corners pinned.
void clothSheet(Solver* s, int n, double spacing = 0.25, float E = 1e4f, float nu = 0.3f, float mass = 0.1f,
                double y = 5.0);
// Bodies that leave bounds of +-6 m in x and z in the first step: a fast
// box with the spring tying it to a resting box, the fast column of a small
// Neo-Hookean block, and a box resting on another.
void outOfBounds(Solver* s);

}  // namespace Levels
