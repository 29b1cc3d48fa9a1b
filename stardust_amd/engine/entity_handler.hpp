// EntityHandler of the STARDUST engine API (src/engine/entities/entity_handler.hpp:9-125).
//
// Holds the entities added to an engine and, after prepareArrays(), the host
// arrays the reference packs (entity_handler.cpp:8-98) under the same names:
// per body position / velocity / mass / quaternion / inertia tensor /
// momenta, per sphere position / velocity / mass / size / owner / relative
// position, and entity start / length ranges. The arrays are std::vector
// storage exposed through the reference's pointer members; the device-side
// d_* arrays of the reference are owned by libstardust_hip.so instead
// (DEMEngine::update forwards to it) and are not exposed.
#ifndef STARDUST_AMD_ENGINE_ENTITY_HANDLER_HPP
#define STARDUST_AMD_ENGINE_ENTITY_HANDLER_HPP

#include <functional>
#include <vector>

#include "entity.hpp"

namespace STARDUST {

class EntityHandler {
public:
    EntityHandler() = default;
    EntityHandler(const EntityHandler& o) { *this = o; }
    EntityHandler& operator=(const EntityHandler& o);

    int getNumberOfSpheres() const { return n_spheres; }
    int getNumberOfParticles() const { return (int)particles.size(); }
    int getNumberOfMeshes() const { return (int)meshes.size(); }
    int getNumberOfHulls() const { return (int)hulls.size(); }
    // spheres plus mesh triangles, as the reference counts primitives
    int getNumberOfPrimitives() const { return n_spheres + n_triangles; }

    void addEntity(const DEMParticle& particle);
    void addEntity(const DEMMesh& mesh);
    void addEntity(const DEMHull& hull);
    const std::vector<DEMMesh>& getMeshes() const { return meshes; }
    std::vector<DEMHull>& getHulls() { return hulls; }
    // Before transfer(): drops entity idx of the given type (0 = particle) from
    // the list (call prepareArrays again). The reference's is a stub
    // (entity_handler.cpp:118).
    void removeEntity(int type, int idx);
    // After transfer(): add or remove an entity in the running simulation
    // (forwarded to the DEMEngine that owns this handler; the reference's are
    // stubs, entity_handler.cpp:108-119). The other entities keep their state;
    // entity indices above a removed one shift down by one.
    void addEntityRuntime(const DEMParticle& particle);
    void removeEntityRuntime(int type, int idx);

    void prepareArrays();
    const std::vector<DEMParticle>& entities() const { return particles; }

    // Packed host arrays (valid after prepareArrays; refreshed by DEMEngine::sync).
    int* h_entity_start_ptr = nullptr;
    int* h_entity_length_ptr = nullptr;
    int* h_particle_idx_ptr = nullptr;
    int* h_particle_sorted_idx_ptr = nullptr;
    float* h_rigid_body_mass_ptr = nullptr;
    float4* h_rigid_body_position_ptr = nullptr;
    float4* h_rigid_body_velocity_ptr = nullptr;
    float4* h_rigid_body_angular_velocity_ptr = nullptr;
    float4* h_rigid_body_linear_momentum_ptr = nullptr;
    float4* h_rigid_body_angular_momentum_ptr = nullptr;
    float4* h_rigid_body_quaternion_ptr = nullptr;  // (w, x, y, z) in (x, y, z, w)
    float9* h_rigid_body_inertia_tensor_ptr = nullptr;  // reference formula: point masses (entity_handler.cpp:72-90)
    float4* h_rigid_body_forces_ptr = nullptr;
    float4* h_rigid_body_torques_ptr = nullptr;
    int* h_particle_to_rigid_idx_ptr = nullptr;
    float4* h_particle_position_ptr = nullptr;
    float4* h_particle_velocity_ptr = nullptr;
    float4* h_particle_forces_ptr = nullptr;
    float* h_particle_mass_ptr = nullptr;
    float* h_particle_size_ptr = nullptr;
    float4* h_particle_relative_position_ptr = nullptr;
    float4* h_particle_init_relative_position_ptr = nullptr;

private:
    void bind();

    // set by the owning DEMEngine once transferred
    std::function<void(const DEMParticle&)> runtime_add;
    std::function<void(int)> runtime_remove;

    std::vector<DEMParticle> particles;
    std::vector<DEMMesh> meshes;
    std::vector<DEMHull> hulls;
    int n_spheres = 0;
    int n_triangles = 0;
    bool prepared = false;

    std::vector<int> entity_start, entity_length, particle_idx, particle_sorted_idx, particle_to_rigid;
    std::vector<float> rb_mass, p_mass, p_size;
    std::vector<float4> rb_pos, rb_vel, rb_angvel, rb_linmom, rb_angmom, rb_quat, rb_force, rb_torque;
    std::vector<float9> rb_inertia;
    std::vector<float4> p_pos, p_vel, p_force, p_rel, p_init_rel;

    friend class DEMEngine;
};

}  // namespace STARDUST

#endif
