// DEMParticle sphere lattice (reference: src/engine/entities/entity.cpp:9-59).
#include "entity.hpp"

#include <cmath>
#include <stdexcept>
#include <string>

namespace STARDUST {

DEMParticle::DEMParticle(unsigned int id, int grid_resolution, Scalar size, Scalar diameter, Scalar mass_,
                         float4 position_, float4 velocity_)
    : position(position_), velocity(velocity_), COM(make_float4(0.0f, 0.0f, 0.0f, 0.0f)), mass(mass_),
      diameter(diameter), m_id(id), m_size(size), m_grid_resolution(grid_resolution) {
    if (grid_resolution < 1) throw std::invalid_argument("DEMParticle: grid_resolution must be >= 1");
    if (!(diameter > 0.0f) || !(size >= diameter))
        throw std::invalid_argument("DEMParticle: need diameter > 0 and size >= diameter");
    initParticles(m_grid_resolution);
    getCenterOfMass();
    setParticlesInWorldSpace();
}

// grid_resolution^3 spheres at (i, j, k) * dx, i, j, k = 1 .. grid_resolution,
// with dx = size / g and mass / g each, g = floor(size / diameter)^2 (entity.cpp:12-30).
void DEMParticle::initParticles(int grid_resolution) {
    const float per_side = std::floor(m_size / diameter);
    const float g = per_side * per_side;
    const float dx = m_size / g;
    particles.clear();
    for (int i = 1; i <= grid_resolution; ++i)
        for (int j = 1; j <= grid_resolution; ++j)
            for (int k = 1; k <= grid_resolution; ++k) {
                DEMSphere s;
                s.position = make_float4((float)i * dx, (float)j * dx, (float)k * dx, 0.0f);
                s.size = diameter;
                s.mass = mass / g;
                particles.push_back(s);
            }
}

// Centroid of the lattice (entity.cpp:34-40).
void DEMParticle::getCenterOfMass() {
    float4 c = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (const DEMSphere& s : particles) c += s.position;
    COM = c / (float)particles.size();
}

// world = position - (p - COM) (entity.cpp:42-57).
void DEMParticle::setParticlesInWorldSpace() {
    for (DEMSphere& s : particles) {
        const float4 rel = s.position - COM;
        s.position = position - rel;
    }
}

DEMMesh::DEMMesh(const char* filename, float4 position_, float4 quaternion) : position(position_), quat(quaternion) {
    throw std::runtime_error(std::string("DEMMesh: triangle meshes are not supported (") +
                             (filename ? filename : "null") + ")");
}

}  // namespace STARDUST
