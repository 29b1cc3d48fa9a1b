// Host packing of entities (reference: src/engine/entities/entity_handler.cpp:8-121).
#include "entity_handler.hpp"

#include <stdexcept>
#include <string>

namespace STARDUST {

EntityHandler& EntityHandler::operator=(const EntityHandler& o) {
    if (this == &o) return *this;
    particles = o.particles;
    meshes = o.meshes;
    hulls = o.hulls;
    n_spheres = o.n_spheres;
    n_triangles = o.n_triangles;
    prepared = o.prepared;
    entity_start = o.entity_start;
    entity_length = o.entity_length;
    particle_idx = o.particle_idx;
    particle_sorted_idx = o.particle_sorted_idx;
    particle_to_rigid = o.particle_to_rigid;
    rb_mass = o.rb_mass;
    p_mass = o.p_mass;
    p_size = o.p_size;
    rb_pos = o.rb_pos;
    rb_vel = o.rb_vel;
    rb_angvel = o.rb_angvel;
    rb_linmom = o.rb_linmom;
    rb_angmom = o.rb_angmom;
    rb_quat = o.rb_quat;
    rb_force = o.rb_force;
    rb_torque = o.rb_torque;
    rb_inertia = o.rb_inertia;
    p_pos = o.p_pos;
    p_vel = o.p_vel;
    p_force = o.p_force;
    p_rel = o.p_rel;
    p_init_rel = o.p_init_rel;
    bind();
    return *this;
}

void EntityHandler::addEntity(const DEMParticle& particle) {
    particles.push_back(particle);
    n_spheres += (int)particle.getParticles().size();
    prepared = false;
}

void EntityHandler::addEntity(const DEMMesh& mesh) {
    meshes.push_back(mesh);
    n_triangles += mesh.n_triangles;
}

void EntityHandler::addEntity(const DEMHull& hull) { hulls.push_back(hull); }

void EntityHandler::removeEntity(int type, int idx) {
    if (type != 0) throw std::invalid_argument("EntityHandler::removeEntity: only particles (type 0) are held");
    if (idx < 0 || idx >= (int)particles.size())
        throw std::out_of_range("EntityHandler::removeEntity: no entity " + std::to_string(idx));
    n_spheres -= (int)particles[idx].getParticles().size();
    particles.erase(particles.begin() + idx);
    prepared = false;
}

void EntityHandler::addEntityRuntime(const DEMParticle& particle) {
    if (!runtime_add) throw std::runtime_error("EntityHandler::addEntityRuntime: not transferred to a DEMEngine yet");
    runtime_add(particle);
}

void EntityHandler::removeEntityRuntime(int type, int idx) {
    if (type != 0) throw std::invalid_argument("EntityHandler::removeEntityRuntime: only particles (type 0) are held");
    if (!runtime_remove)
        throw std::runtime_error("EntityHandler::removeEntityRuntime: not transferred to a DEMEngine yet");
    runtime_remove(idx);
}

void EntityHandler::bind() {
    auto p = [](auto& v) { return v.empty() ? nullptr : v.data(); };
    h_entity_start_ptr = p(entity_start);
    h_entity_length_ptr = p(entity_length);
    h_particle_idx_ptr = p(particle_idx);
    h_particle_sorted_idx_ptr = p(particle_sorted_idx);
    h_rigid_body_mass_ptr = p(rb_mass);
    h_rigid_body_position_ptr = p(rb_pos);
    h_rigid_body_velocity_ptr = p(rb_vel);
    h_rigid_body_angular_velocity_ptr = p(rb_angvel);
    h_rigid_body_linear_momentum_ptr = p(rb_linmom);
    h_rigid_body_angular_momentum_ptr = p(rb_angmom);
    h_rigid_body_quaternion_ptr = p(rb_quat);
    h_rigid_body_inertia_tensor_ptr = p(rb_inertia);
    h_rigid_body_forces_ptr = p(rb_force);
    h_rigid_body_torques_ptr = p(rb_torque);
    h_particle_to_rigid_idx_ptr = p(particle_to_rigid);
    h_particle_position_ptr = p(p_pos);
    h_particle_velocity_ptr = p(p_vel);
    h_particle_forces_ptr = p(p_force);
    h_particle_mass_ptr = p(p_mass);
    h_particle_size_ptr = p(p_size);
    h_particle_relative_position_ptr = p(p_rel);
    h_particle_init_relative_position_ptr = p(p_init_rel);
}

// Same packing as entity_handler.cpp:8-98: bodies in add order, their spheres
// contiguous; body position = entity position; relative sphere positions and
// the point-mass inertia tensor about the entity position.
void EntityHandler::prepareArrays() {
    const size_t nb = particles.size(), ns = (size_t)n_spheres;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    entity_start.assign(nb, 0);
    entity_length.assign(nb, 0);
    rb_mass.assign(nb, 0.0f);
    rb_pos.assign(nb, zero);
    rb_vel.assign(nb, zero);
    rb_angvel.assign(nb, zero);
    rb_linmom.assign(nb, zero);
    rb_angmom.assign(nb, zero);
    rb_quat.assign(nb, make_float4(1.0f, 0.0f, 0.0f, 0.0f));
    rb_force.assign(nb, zero);
    rb_torque.assign(nb, zero);
    rb_inertia.assign(nb, float9{});
    particle_idx.assign(ns, 0);
    particle_sorted_idx.assign(ns, 0);
    particle_to_rigid.assign(ns, 0);
    p_mass.assign(ns, 0.0f);
    p_size.assign(ns, 0.0f);
    p_pos.assign(ns, zero);
    p_vel.assign(ns, zero);
    p_force.assign(ns, zero);
    p_rel.assign(ns, zero);
    p_init_rel.assign(ns, zero);
    int offset = 0;
    for (size_t i = 0; i < nb; ++i) {
        const DEMParticle& e = particles[i];
        const std::vector<DEMSphere> sph = e.getParticles();
        rb_pos[i] = e.position;
        rb_vel[i] = e.velocity;
        rb_mass[i] = e.mass;
        rb_quat[i] = e.quaternion;
        rb_angvel[i] = e.angular_velocity;
        entity_start[i] = offset;
        entity_length[i] = (int)sph.size();
        float Ixx = 0.0f, Iyy = 0.0f, Izz = 0.0f, Ixy = 0.0f, Ixz = 0.0f, Iyz = 0.0f;
        for (size_t j = 0; j < sph.size(); ++j) {
            const DEMSphere& s = sph[j];
            const size_t k = (size_t)offset + j;
            particle_sorted_idx[k] = (int)k;
            particle_idx[k] = (int)k;
            p_pos[k] = s.position;
            p_vel[k] = e.velocity;
            p_mass[k] = s.mass;
            p_size[k] = s.size;
            particle_to_rigid[k] = (int)i;
            // world offset, and the body-frame offset (the same for a lattice,
            // which starts unrotated); the inertia is the body frame's
            p_rel[k] = s.position - e.position;
            const float4 r = e.bodyOffset(j);
            p_init_rel[k] = r;
            Ixx += s.mass * (r.y * r.y + r.z * r.z);
            Iyy += s.mass * (r.x * r.x + r.z * r.z);
            Izz += s.mass * (r.x * r.x + r.y * r.y);
            Ixy += -s.mass * r.x * r.y;
            Ixz += -s.mass * r.x * r.z;
            Iyz += -s.mass * r.y * r.z;
        }
        rb_inertia[i] = float9{{Ixx, Ixy, Ixz, Ixy, Iyy, Iyz, Ixz, Iyz, Izz}};
        offset += (int)sph.size();
    }
    prepared = true;
    bind();
}

}  // namespace STARDUST
