// Entities of the STARDUST engine API (src/engine/entities/entity.hpp:9-93).
//
// DEMParticle keeps the reference constructor and its sphere layout exactly
// (entity.cpp:9-59): a grid_resolution^3 lattice of spheres of the given
// diameter with spacing dx = size / floor(size / diameter)^2, each carrying
// mass / floor(size / diameter)^2, placed in world space by mirroring about the
// lattice centroid (world = position - (p - centroid)). These are the
// reference's own formulas, kept so that scenes built through this API have
// the same spheres as in the reference.
#ifndef STARDUST_AMD_ENGINE_ENTITY_HPP
#define STARDUST_AMD_ENGINE_ENTITY_HPP

#include <vector>

#include "types.hpp"

namespace STARDUST {

struct DEMSphere {
    float4 position;
    Scalar size;  // diameter
    Scalar mass;
};

class DEMParticle {
public:
    DEMParticle(unsigned int id, int grid_resolution, Scalar size, Scalar diameter, Scalar mass, float4 position,
                float4 velocity);

    void initParticles(int grid_resolution);
    void setParticlesInWorldSpace();
    void getCenterOfMass();

    std::vector<DEMSphere> getParticles() const { return particles; }
    unsigned int getId() const { return m_id; }

    float4 position;  // world position of the entity (the centroid of its spheres)
    float4 velocity;
    float4 COM;       // lattice centroid in entity space
    float mass;
    float diameter;

private:
    unsigned int m_id;
    Scalar m_size;
    int m_grid_resolution;
    std::vector<DEMSphere> particles;
};

// Triangle-mesh walls (entity.hpp:64-93, mesh_loader.cpp) are not part of this
// implementation: the constructor throws std::runtime_error.
class DEMMesh {
public:
    DEMMesh(const char* filename, float4 position, float4 quaternion);
    int n_triangles = 0;
    float4 position;
    float4 quat;
};

}  // namespace STARDUST

#endif
