// DEMEngine: the STARDUST engine API (src/engine/engine.hpp:21-190,303-305;
// engine.cpp:601-609; simulation_kernel.cu:52-153) kept as the drop-in
// boundary. A program written against the reference's engine
// (sandbox/testDEM.cpp:80-128 is the model) builds and runs unchanged:
//
//     STARDUST::engineConfig config{STARDUST::SAP, STARDUST::GPU};
//     STARDUST::DEMEngine engine(config, 0.005f);
//     engine.add(STARDUST::DEMParticle(0, 1, 0.5f, 0.5f, 9.0f, pos, vel));
//     engine.prep();
//     engine.transfer();
//     for (int i = 0; i < 300; ++i) engine.update(0.0005f);
//
// transfer() picks one of three step paths of libstardust_hip.so:
//  - any DEMMesh or DEMHull added: the rigid-body path (sd_hull_*): every
//    single-sphere DEMParticle is a sphere body, every DEMHull a convex body,
//    every mesh triangle a static TRIANGLE body; sweep-and-prune, GJK/EPA
//    contacts (normal spring-dashpot plus Coulomb-capped viscous friction),
//    the semi-implicit integration; at most 65,536 bodies; multi-sphere
//    particles, runtime edits, checkpoints and VTK output are not available
//    on this path;
//  - every entity a single sphere (the reference driver's 1,000,000-sphere
//    scene, testDEM.cpp:80-128): a sphere context (sd_ctx_*): hash grid of
//    cell = the largest diameter, packed 64-bit keys re-sorted every step,
//    the sphere law with tangential history, semi-implicit Euler. When every
//    sphere lies inside the cube [0, domain]^3 (the constructor's domain, in
//    metres, as the reference's scene files state it), its six faces are
//    walls; otherwise the domain is treated as unbounded (the reference's
//    "TODO: Add option for unbounded", engine.cpp:73): no walls, a grid over
//    the spheres' extent (spheres leaving it are hashed to its border cells).
//    A sphere's orientation is not integrated (its quaternion stays the
//    identity); its spin is.
//  - otherwise: every entity is a rigid body made of its spheres, in a clump
//    context (sd_clump_*): sphere pose, LBVH broadphase over body boxes,
//    sphere-sphere contacts with the linear spring-dashpot law, per-body force
//    and torque gather, quaternion integration.
// update(dt) advances one step of that path. Differences from the reference,
// each one a correction of a documented defect:
//  - inverse inertia is R I^-1 R^T with I the member spheres' own 2/5 m r^2
//    plus the parallel-axis term (the reference inverts only det >= 1e-4 and
//    applies R^T I^-1 R; a single sphere had I = 0);
//  - the quaternion is renormalised every step;
//  - the contact law is the linear spring-dashpot law with a Coulomb-limited
//    tangential spring on the sphere radius (diameter / 2), set by
//    setContactLaw (the reference multiplied component-wise and used the
//    diameter as the radius);
//  - bodies with mass >= the static threshold (10, the reference's rule,
//    physics_update.cu:168) do not move; setStaticMassThreshold changes it.
// Errors raise std::runtime_error with the library's message (the reference
// called exit()).
#ifndef STARDUST_AMD_ENGINE_ENGINE_HPP
#define STARDUST_AMD_ENGINE_ENGINE_HPP

#include <string>
#include <vector>

#include "entity.hpp"
#include "entity_handler.hpp"

struct sd_clump_ctx;
struct sd_ctx;
struct sd_hull_ctx;

namespace STARDUST {

enum CollisionEngine { SAP, SHP };
enum Device { CPU, GPU };

struct engineConfig {
    CollisionEngine collision_engine;
    Device device;
};

class DEMEngine {
public:
    DEMEngine(engineConfig config, Scalar domain);
    ~DEMEngine();
    DEMEngine(const DEMEngine&) = delete;
    DEMEngine& operator=(const DEMEngine&) = delete;

    void add(DEMParticle particle);
    void add(DEMMesh mesh);  // static triangle walls (rigid-body path)
    void add(DEMHull hull);  // convex hull body (rigid-body path)
    void prep();
    void transfer();
    void update(float dt);

    // The reference's other names for the same steps (engine.hpp:172-185;
    // engine.cpp:110-130,243,601-612): addParticle/addMesh add, prepArrays
    // packs, transferDataToDevice transfers, step advances one step. The grid
    // is set up by transfer() (setUpGrid does nothing); cleanBuffers releases
    // the device state (a later transfer() uploads again).
    void addParticle(DEMParticle particle) { add(particle); }
    void addMesh(DEMMesh mesh) { add(mesh); }
    void setUpGrid() {}
    void prepArrays() { prep(); }
    void transferDataToDevice() { transfer(); }
    void step(Scalar dt) { update((float)dt); }
    void cleanBuffers();

    // Scene file of the reference's format (setup/star_test.json): entities of
    // type "generic" become DEMParticle(count + 1, 10, size, size, mass, position,
    // velocity) (the mapping of the commented loader, engine.cpp:64-108, with
    // the mass field); "domain" and "gravity" (magnitude, along -z) are read.
    void loadJSONSetup(const std::string& filepath);

    std::vector<DEMParticle> getEntities() const { return entity_handler.entities(); }
    int getEntityLength() const { return entity_handler.getNumberOfParticles(); }
    int getNumberOfSpheres() const { return entity_handler.getNumberOfSpheres(); }
    EntityHandler& getEntityHandler() { return entity_handler; }

    // Copies the current body and sphere state into the EntityHandler host arrays.
    void sync();

    // The step path transfer() chose, and (sphere path) its grid and walls.
    struct SphereGrid {
        float lo[3];
        float cell;
        int n[3];
        bool walls;
        float wall_lo[3], wall_hi[3];
    };
    bool sphereMode() const { return m_sphere_mode; }
    bool hullMode() const { return m_hull_mode; }
    const SphereGrid& sphereGrid() const { return m_sgrid; }
    long long stepsDone() const { return m_steps; }

    // Runtime add / remove (after transfer): the other entities keep their
    // state, contact history included (EntityHandler::addEntityRuntime and
    // removeEntityRuntime forward here). On the sphere path an added entity
    // must be a single sphere; indices above a removed entity shift down.
    void addRuntime(const DEMParticle& particle);
    void removeRuntime(int idx);

    // Checkpoint of the device state (step path, step count and the context's
    // state buffer) to a file; loadState restores it into an engine that has
    // transferred the same entities. Stepping on from a loaded state is
    // bit-identical to stepping on from the saved one.
    void saveState(const std::string& path);
    void loadState(const std::string& path);

    // State dump in place of the reference's GL interop (engine.cpp:36-62):
    // legacy VTK (binary) polydata, one point per sphere with its radius, its
    // body's velocity and its body index.
    void writeVTK(const std::string& path);

    // Contact law and options (defaults: kn = 2e6 N/m as the reference's
    // constant, damping ratio 0.3, mu = 0.5, kt = 2/7 kn, gravity 9.81 along -z).
    void setContactLaw(float kn, float damping_ratio, float mu, float kt_ratio);
    void setGravity(float gx, float gy, float gz);
    void setStaticMassThreshold(float threshold);

    bool is_first_step = true;

private:
    void ensure_context(float dt);
    void transfer_spheres();
    void transfer_clumps();
    void transfer_bodies();
    void bind_runtime();

    engineConfig engineConfiguration;
    Scalar m_domain;
    EntityHandler entity_handler;
    sd_clump_ctx* m_ctx = nullptr;
    sd_ctx* m_sctx = nullptr;
    sd_hull_ctx* m_hctx = nullptr;
    bool m_sphere_mode = false;
    bool m_hull_mode = false;
    SphereGrid m_sgrid{};
    bool m_uploaded = false;
    float m_dt = 0.0f;
    float m_kn = 2e6f, m_zeta = 0.3f, m_mu = 0.5f, m_kt_ratio = 2.0f / 7.0f;
    float m_g[3] = {0.0f, 0.0f, -9.81f};
    float m_static_mass = 10.0f;
    long long m_steps = 0;
    std::vector<float> up_pos4, up_vel4, up_quat4, up_angvel4, up_inertia6, up_local4;
    std::vector<int> up_first, up_count;
};

}  // namespace STARDUST

#endif
