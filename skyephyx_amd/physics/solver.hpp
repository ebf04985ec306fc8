// solver.hpp -- the reference's Solver API (src/physics/solver.hpp:10-95)
// on top of the MI355X AVBD step (include/avbd_capi.h).
//
// A caller builds the world exactly as with the reference -- AddBody,
// AddParticle, `new` forces and energies handed to AddForce/AddEnergy,
// Remove* -- and calls Step() once per frame.  The host objects (Mesh,
// Force, Energy) stay the caller's view of the world; the step itself runs
// on the GPU:
//
//   * the first Step() after a structural change (a body, force or energy
//     added or removed, a tunable changed, a body's mass/inertia/scale/flags
//     written) uploads the whole world (avbd_upload);
//   * writes to a body's transform or velocities between steps are sent
//     with avbd_upload_bodies;
//   * after the step the body state is read back into every Mesh
//     (avbd_download_bodies), bodies the step removed for leaving the world
//     bounds are removed from the host world in the same order
//     (avbd_last_removed + RemoveBody), and -- unless syncForces is off --
//     forces, contact manifolds and their row state are mirrored into
//     solverForces (new manifolds appear, retired ones are removed).
//
// With syncForces off the host forces are brought up to date lazily, before
// the next structural change, and by SyncForces().
#pragma once

#include <memory>
#include <vector>

#include "avbd_capi.h"
#include "energy.hpp"
#include "force.hpp"
#include "geometry.hpp"

// The world as the C ABI's records (one scene): what Step() uploads, and a
// snapshot a caller can store or hand to avbd_upload / avbd_footprint.
struct WorldRecords {
    avbd_scene scene{};
    std::vector<avbd_body> bodies;
    std::vector<avbd_joint> joints;
    std::vector<avbd_spring> springs;
    std::vector<avbd_manifold> manifolds;
    std::vector<avbd_tet> tets;
    std::vector<avbd_tri> tris;
    avbd_world view();  // avbd_world pointing into the vectors
};

struct SolverTimings {
    float broadPhaseMs = 0.f;
    float warmstartMs = 0.f;
    float predictionMs = 0.f;
    float primalDualMs = 0.f;
    float velocityUpdateMs = 0.f;
    float postStabMs = 0.f;
    float totalSubstepMs = 0.f;
};

class Solver {
public:
    Solver();
    ~Solver();

    std::vector<std::unique_ptr<Mesh>> solverBodies;
    std::vector<std::unique_ptr<Force>> solverForces;
    std::vector<std::unique_ptr<Energy>> solverEnergies;

    std::vector<Mesh*> bodyPtrs;
    std::vector<Force*> forcePtrs;
    std::vector<Energy*> energyPtrs;

    // Debug geometry of the reference's renderer, filled after each Step from
    // the device when debugOutput is on (avbd_get_debug_lines / _points;
    // element edges take the bodies' Mesh::color), else left empty.
    std::vector<GPULineData> lineData;
    std::vector<GPUDebugPointData> debugPointData;
    bool debugOutput = false;
    int spareBodies = 0;  // free body slots per upload (option spare_bodies): AddBody between steps fills them

    void Start();
    void Clear();
    void Step();

    Mesh* AddBody(ModelType modelType, float density, float friction, const Vector3f& position, const Vector3f& scale,
                  const Vector3f& velocity, const Quaternionf rotation, const Vector3f& angularVelocity, bool isStatic,
                  const Vector3f& color = Vector3f(1.f, 1.f, 1.f));
    Force* AddForce(std::unique_ptr<Force> force);
    Energy* AddEnergy(std::unique_ptr<Energy> energy);
    Mesh* AddParticle(float mass, float friction, const Vector3f& position, const Vector3f& velocity, bool isStatic,
                      const Vector3f& color = Vector3f(1.f, 1.f, 1.f));
    void RemoveForce(Force* force);
    void RemoveEnergy(Energy* energy);
    void RemoveBody(Mesh* body);

    void RebuildPtrCaches();

    float averageStepTime = 0.f;
    SolverTimings timings;

    // Changing parameters (solver.hpp:69-84)
    bool postStabilization = true;
    int numIterations = 10;
    float alpha = 0.95f;
    float beta = 100000.f;
    float gamma = 0.99f;
    float onPenetrationPenalty = 0.f;
    float stepValue = 1.f / 60.f;
    EigenProjectionMode projectionMode = EigenProjectionMode::ABSOLUTE;
    float trustRegionRho = 0.f;
    // The two features the reference declares but leaves unfinished, off by
    // default (the reference's behaviour; not reference members):
    // trustRegionUpdate updates trustRegionRho after every primal sweep under
    // ADAPTIVE projection (the empty step 5.2, solver.cpp:451-455);
    // sphereCollision makes ModelType_Sphere bodies collide as spheres
    // (the reference collides every pair as boxes, manifold.cpp:57).
    bool trustRegionUpdate = false;
    bool sphereCollision = false;
    float prevTotalEnergy = 0.f;
    float trustRegionThreshold = 0.01f;

    bool emergencyStop = false;

    // ---- facade additions (no reference counterpart) ----
    // Constants of the reference made settable: gravity (solver.cpp:7) and
    // the world bounds (constants.hpp:11-12).
    Vector3f gravity = Vector3f(0.f, -9.81f, 0.f);
    Vector3f boundsMin = Vector3f(-50.f, -50.f, -50.f);
    Vector3f boundsMax = Vector3f(50.f, 50.f, 50.f);
    int device = 0;          // HIP device of the step
    bool syncForces = true;  // mirror forces and contacts into solverForces after every step
    int64_t uploads = 0;     // full world uploads so far (RemoveBody, and AddBody into slots removals
                             // freed, between steps need none)
    // Bring solverForces (joints, springs, manifolds and their row state) up
    // to date with the device now.
    void SyncForces();
    // The device context, created on first use (options such as the lane
    // widths can be set on it before the first Step()).
    avbd_ctx* DeviceContext();
    // The host world as ABI records (bodies in solverIndex order, forces and
    // energies with their creation serials, contacts in creation order).
    void ExportWorld(WorldRecords& out);
    // ExportWorld written as a world snapshot file (avbd_snapshot_write);
    // throws std::runtime_error with the library's message on failure.
    void SaveSnapshot(const std::string& path);
    // Replaces the world with a one-scene snapshot file (avbd_snapshot_read):
    // scene parameters, bodies (Mesh::modelType recorded as sphere or cube),
    // joints, springs, contact manifolds with their row state, and energies,
    // each keeping its creation serial, so that ExportWorld gives the file's
    // records back.  The next Step uploads the world.  Throws
    // std::runtime_error on a bad file or a snapshot of several scenes.
    void LoadSnapshot(const std::string& path);

    // Used by the Force / Energy constructors and the Mesh destructor.
    int64_t TakeForceSerial();
    int32_t TakeEnergySerial();
    void MarkStructureChanged();

private:
    static constexpr int TIMING_WINDOW = 60;
    std::vector<SolverTimings> timingAccumulator;

    avbd_ctx* ctx_ = nullptr;
    bool uploaded_ = false;
    bool structureDirty_ = true;
    bool forcesStale_ = false;
    bool mirroring_ = false;  // applying device changes to the host objects
    bool removingOnDevice_ = false;  // a RemoveBody the device applies itself
    std::vector<Mesh*> pendingAdds_;  // bodies added since the last Step, for avbd_add_bodies
    int64_t nextForceSerial_ = 0;
    int32_t nextEnergySerial_ = 0;
    std::vector<unsigned char> lastScene_;   // scene record of the last upload
    std::vector<unsigned char> lastBodies_;  // body records as last uploaded / read back
    bool dragOnDevice_ = false;  // some body's dragged flag is set on the device

    void EnsureContext();
    void Upload();
    void ReadBack();
    void ApplyRemovals();
    void BodyRecords(std::vector<avbd_body>& out) const;
    void SceneRecord(std::vector<unsigned char>& out) const;
    void AccumulateTimings();
};
