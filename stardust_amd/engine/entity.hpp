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

// A triangle mesh (entity.hpp:64-93): the triangles of an STL file (binary
// or ASCII), placed at position and rotated by quaternion (stored as the
// reference's float4 quaternions: x = w, y = i, z = j, w = k). Its triangles
// are static walls: DEMEngine steps scenes with meshes on the rigid-body path,
// where spheres and hulls collide with each triangle (GJK/EPA with the
// TRIANGLE shape). Members as in the reference: vertices (unique, mesh frame),
// indicies (3 per triangle), normals (per triangle, from the file), and the
// bounding sphere of every triangle (positions, radii; computeBoundingSpheres).
class DEMMesh {
public:
    DEMMesh(const char* filename, float4 position, float4 quaternion);

    // Per triangle: the smallest sphere through or around its corners - the
    // longest edge as diameter for a right or obtuse triangle, else the
    // circumscribed sphere (entity.cpp:75-120; that code's circumcentre
    // expression is not the circumcentre, here the circumcentre is exact).
    void computeBoundingSpheres();
    // Corners of triangle t in world space (the mesh pose applied).
    void worldTriangle(int t, float out[9]) const;

    int n_triangles = 0;
    float4 position;
    float4 quat;

    std::vector<float4> positions;  // bounding sphere centres (mesh frame)
    std::vector<float> radii;

    std::vector<float4> vertices;  // unique vertices, lexicographic order
    std::vector<float4> normals;   // one per triangle
    std::vector<int> indicies;     // 3 per triangle
};

// A convex hull (primitives.hpp:37-57, the sandbox's Hull with POLYHEDRA
// type; it has no entity in the reference's engine): vertices in the body
// frame about its centre, mass (>= the static threshold: static), and the
// pose. Inertia: equal point masses at the vertices. Stepped on the
// rigid-body path (GJK/EPA contacts).
class DEMHull {
public:
    DEMHull(unsigned int id, const std::vector<float4>& vertices, Scalar mass, float4 position, float4 velocity,
            float4 quaternion);
    unsigned int getId() const { return m_id; }

    std::vector<float4> vertices;
    Scalar mass;
    float4 position;
    float4 velocity;
    float4 quat;             // (w, x, y, z) in (x, y, z, w)
    float4 angular_momentum; // world, after DEMEngine::sync

private:
    unsigned int m_id;
};

}  // namespace STARDUST

#endif
