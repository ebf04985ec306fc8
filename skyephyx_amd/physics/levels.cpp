// levels.cpp -- scene builders over the solver facade.  Coordinates are
// computed in double and rounded once to float.
#include "levels.hpp"

#include <cmath>
#include <memory>

namespace {

const Vector3f kWhite(1.f, 1.f, 1.f);

Vector3f v3(double x, double y, double z) { return Vector3f(float(x), float(y), float(z)); }

Mesh* box(Solver* s, double x, double y, double z, double sx, double sy, double sz,
          const Vector3f& vel = Vector3f::Zero()) {
    return s->AddBody(ModelType_Cube, 1.f, 0.5f, v3(x, y, z), v3(sx, sy, sz), vel, Quaternionf::Identity(),
                      Vector3f::Zero(), false);
}

}  // namespace

namespace GeometryHelpers {

std::vector<Mesh*> makeTetVolume(Solver* solver, double ox, double oy, double oz, int nx, int ny, int nz,
                                 double cellSize, float E, float nu, float particleMass, float friction,
                                 const std::function<bool(int, int, int)>& pin) {
    std::vector<Mesh*> parts((size_t)nx * ny * nz);
    auto idx = [&](int i, int j, int k) { return (size_t)i + (size_t)j * nx + (size_t)k * nx * ny; };
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                Mesh* p = solver->AddParticle(particleMass, friction,
                                              v3(ox + i * cellSize, oy + j * cellSize, oz + k * cellSize),
                                              Vector3f::Zero(), pin && pin(i, j, k));
                p->name = "V_" + std::to_string(i) + "_" + std::to_string(j) + "_" + std::to_string(k);
                parts[idx(i, j, k)] = p;
            }
    for (int ci = 0; ci < nx - 1; ++ci)
        for (int cj = 0; cj < ny - 1; ++cj)
            for (int ck = 0; ck < nz - 1; ++ck) {
                Mesh* v[8] = {parts[idx(ci, cj, ck)],         parts[idx(ci + 1, cj, ck)],
                              parts[idx(ci, cj + 1, ck)],     parts[idx(ci + 1, cj + 1, ck)],
                              parts[idx(ci, cj, ck + 1)],     parts[idx(ci + 1, cj, ck + 1)],
                              parts[idx(ci, cj + 1, ck + 1)], parts[idx(ci + 1, cj + 1, ck + 1)]};
                solver->AddEnergy(std::make_unique<NeoHookeanFEM>(solver, v[0], v[1], v[2], v[4], E, nu));
                solver->AddEnergy(std::make_unique<NeoHookeanFEM>(solver, v[3], v[2], v[1], v[7], E, nu));
                solver->AddEnergy(std::make_unique<NeoHookeanFEM>(solver, v[5], v[1], v[4], v[7], E, nu));
                solver->AddEnergy(std::make_unique<NeoHookeanFEM>(solver, v[6], v[4], v[2], v[7], E, nu));
                solver->AddEnergy(std::make_unique<NeoHookeanFEM>(solver, v[1], v[2], v[4], v[7], E, nu));
            }
    return parts;
}

std::vector<Mesh*> makeTetVolume(Solver* solver, float ox, float oy, float oz, int nx, int ny, int nz, float cellSize,
                                 float E, float nu, const Vector3f&, float particleMass, float friction,
                                 bool pinMinX) {
    return makeTetVolume(solver, double(ox), double(oy), double(oz), nx, ny, nz, double(cellSize), E, nu,
                         particleMass, friction, [pinMinX](int i, int, int) { return pinMinX && i == 0; });
}

}  // namespace GeometryHelpers

namespace Levels {

Mesh* ground(Solver* s) {
    return s->AddBody(ModelType_Cube, 1.f, 0.5f, Vector3f(0.f, -10.f, 0.f), Vector3f(20.f, 1.f, 20.f),
                      Vector3f::Zero(), Quaternionf::Identity(), Vector3f::Zero(), true);
}

void pyramid2D(Solver* s, int rows, double gap, double lift) {
    s->alpha = 0.99f;
    s->gamma = 0.99f;
    s->beta = 1e5f;
    ground(s);
    const double pitch = 1.0 + gap;
    for (int r = 0; r < rows; ++r) {
        const int n = rows - r;
        const double y = -9.5 + lift + 0.5 + r * (1.0 + lift);
        for (int k = 0; k < n; ++k) box(s, (k - (n - 1) * 0.5) * pitch, y, 0.0, 1.0, 1.0, 1.0);
    }
}

void jointChain(Solver* s, int nLinks) {
    Vector6f stiff = Vector6f::Constant(INFINITY);
    stiff[4] = 0.f;
    stiff[5] = 0.f;
    const double anchor[3] = {0.0, 5.0, 0.0};
    double pos[3] = {anchor[0] + 1.0, anchor[1], anchor[2]};
    Mesh* last = nullptr;
    for (int i = 0; i < nLinks; ++i) {
        Mesh* body = box(s, pos[0], pos[1], pos[2], 2.0, 0.7, 0.7);
        const Vector3f ra = last ? Vector3f(0.5f, 0.f, 0.f) : v3(anchor[0], anchor[1], anchor[2]);
        s->AddForce(std::make_unique<Joint>(s, last, ra, body, Vector3f(-0.5f, 0.f, 0.f), stiff));
        last = body;
        pos[0] += 2.0;
    }
}

void springBridge(Solver* s, int nLinks, int fractureLink) {
    const double start[3] = {-6.0, 5.0, -5.0};
    const double scale[3] = {2.0, 0.5, 3.0};
    const double spacing = 1.0;
    Mesh* last = nullptr;
    double px = start[0] + scale[0] / 2;
    const double left = start[0] - spacing;
    for (int i = 0; i < nLinks; ++i) {
        Mesh* body = box(s, px, start[1], start[2], scale[0], scale[1], scale[2]);
        for (int side = 0; side < 2; ++side) {
            const double zs = side == 0 ? 1.0 : -1.0;
            Vector3f ra, rb = v3(-0.5, 0.0, zs * 0.5);
            if (last) ra = v3(0.5, 0.0, zs * 0.5);
            else ra = v3(left, start[1], start[2] + zs * 0.5 * scale[2]);
            Force* f = s->AddForce(std::make_unique<Spring>(s, last, ra, body, rb, 5000.f));
            if (i == fractureLink) f->constraintPoints[0].fracture = 2000.f;
        }
        last = body;
        px += scale[0] + spacing;
    }
    const double end = start[0] + scale[0] * nLinks + (nLinks - 1) * spacing;
    const double right = end + spacing;
    for (int side = 0; side < 2; ++side) {
        const double zs = side == 0 ? 1.0 : -1.0;
        s->AddForce(std::make_unique<Spring>(s, nullptr, v3(right, start[1], start[2] + zs * 0.5 * scale[2]), last,
                                             v3(0.5, 0.0, zs * 0.5), INFINITY));
    }
    box(s, (start[0] + end) * 0.5, 8.0, start[2], 2.0, 2.0, 2.0);
}

void clothSheet(Solver* s, int n, double spacing, float E, float nu, float mass, double y) {
    const double sx = -(n - 1) * spacing * 0.5, sz = -(n - 1) * spacing * 0.5;
    std::vector<Mesh*> parts((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
            parts[i + (size_t)j * n] = s->AddParticle(mass, 0.5f, v3(sx + i * spacing, y, sz + j * spacing),
                                                      Vector3f::Zero(), j == 0 && (i == 0 || i == n - 1));
    for (int j = 0; j < n - 1; ++j)
        for (int i = 0; i < n - 1; ++i) {
            Mesh* v00 = parts[i + (size_t)j * n];
            Mesh* v10 = parts[i + 1 + (size_t)j * n];
            Mesh* v01 = parts[i + (size_t)(j + 1) * n];
            Mesh* v11 = parts[i + 1 + (size_t)(j + 1) * n];
            s->AddEnergy(std::make_unique<STVKFEM>(s, v00, v10, v11, E, nu));
            s->AddEnergy(std::make_unique<STVKFEM>(s, v00, v11, v01, E, nu));
        }
}

void outOfBounds(Solver* s) {
    s->boundsMin = Vector3f(-6.f, -20.f, -6.f);
    s->boundsMax = Vector3f(6.f, 20.f, 6.f);
    ground(s);
    box(s, 0.0, -9.0, 0.0, 1.0, 1.0, 1.0);
    Mesh* fast = box(s, 5.5, -5.0, 0.0, 1.0, 1.0, 1.0, Vector3f(60.f, 0.f, 0.f));
    fast->name = "fast";
    Mesh* anchor = box(s, -2.0, -9.0, 0.0, 1.0, 1.0, 1.0);
    s->AddForce(std::make_unique<Spring>(s, anchor, Vector3f::Zero(), fast, Vector3f::Zero(), 100.f));
    std::vector<Mesh*> block = GeometryHelpers::makeTetVolume(s, 5.0, -6.0, -0.5, 3, 2, 2, 0.5, 1e4f, 0.3f, 0.1f, 0.5f,
                                                              nullptr);
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j) block[2 + (size_t)j * 3 + (size_t)k * 6]->velocity = Vector3f(60.f, 0.f, 0.f);
    Mesh* top = box(s, 0.0, -7.98, 0.0, 1.0, 1.0, 1.0);
    top->name = "top";
}

}  // namespace Levels
