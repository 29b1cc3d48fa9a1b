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

#include <cstddef>
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
    // A rigid clump of arbitrary spheres (no counterpart in the reference,
    // whose DEMParticle is the lattice above; the bodies of configuration 3):
    // spheres[j].position is sphere j's centre relative to the body position
    // in the body frame (w unused), .size its diameter, .mass its mass. The
    // body sits at position with orientation quaternion (x = w, y = i, z = j,
    // w = k; unit) and angular velocity angular_velocity (world, rad/s); its
    // mass is the sum of the spheres' and its spheres' world positions are
    // position + R(quaternion) offset. (A scene of single spheres runs on the
    // sphere path, which takes no initial spin: transfer() rejects one.)
    DEMParticle(unsigned int id, const std::vector<DEMSphere>& spheres, float4 position, float4 velocity,
                float4 quaternion, float4 angular_velocity);

    void initParticles(int grid_resolution);
    void setParticlesInWorldSpace();
    void getCenterOfMass();

    std::vector<DEMSphere> getParticles() const { return particles; }
    unsigned int getId() const { return m_id; }

    // Body-frame offset of sphere j from the body position (the lattice: its
    // world offset, the body starting unrotated).
    float4 bodyOffset(std::size_t j) const;

    float4 position;  // world position of the entity (the centroid of its spheres)
    float4 velocity;
    float4 COM;       // lattice centroid in entity space
    float mass;
    float diameter;
    float4 quaternion = make_float4(1.0f, 0.0f, 0.0f, 0.0f);        // initial orientation (x = w, y = i, z = j, w = k)
    float4 angular_velocity = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // initial, world

private:
    unsigned int m_id;
    Scalar m_size;
    int m_grid_resolution;
    std::vector<DEMSphere> particles;
    std::vector<float4> m_body;  // body-frame offsets (general clumps only)
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
