// DEMParticle sphere lattice (reference: src/engine/entities/entity.cpp:9-59).
#include "entity.hpp"

#include "mesh_loader.hpp"

#include <algorithm>
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

DEMParticle::DEMParticle(unsigned int id, const std::vector<DEMSphere>& spheres, float4 position_, float4 velocity_,
                         float4 quaternion_, float4 angular_velocity_)
    : position(position_), velocity(velocity_), COM(make_float4(0.0f, 0.0f, 0.0f, 0.0f)), mass(0.0f),
      diameter(0.0f), quaternion(quaternion_), angular_velocity(angular_velocity_), m_id(id), m_size(0.0f),
      m_grid_resolution(0) {
    if (spheres.empty()) throw std::invalid_argument("DEMParticle: a clump needs at least one sphere");
    const double qw = quaternion.x, qx = quaternion.y, qy = quaternion.z, qz = quaternion.w;
    const double qn = qw * qw + qx * qx + qy * qy + qz * qz;
    if (!(std::fabs(qn - 1.0) < 1e-4)) throw std::invalid_argument("DEMParticle: the quaternion must be of unit length");
    // R(q) in double; each world coordinate rounded once
    const double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz),     2 * (qx * qz + qw * qy),
                         2 * (qx * qy + qw * qz),     1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx),
                         2 * (qx * qz - qw * qy),     2 * (qy * qz + qw * qx),     1 - 2 * (qx * qx + qy * qy)};
    for (const DEMSphere& b : spheres) {
        if (!(b.size > 0.0f) || !(b.mass > 0.0f))
            throw std::invalid_argument("DEMParticle: clump spheres need a positive diameter and mass");
        const double x = b.position.x, y = b.position.y, z = b.position.z;
        DEMSphere w = b;
        w.position = make_float4((float)(position.x + (R[0] * x + R[1] * y + R[2] * z)),
                                 (float)(position.y + (R[3] * x + R[4] * y + R[5] * z)),
                                 (float)(position.z + (R[6] * x + R[7] * y + R[8] * z)), 0.0f);
        particles.push_back(w);
        m_body.push_back(make_float4(b.position.x, b.position.y, b.position.z, 0.0f));
        mass += b.mass;
        diameter = std::max(diameter, b.size);
    }
}

float4 DEMParticle::bodyOffset(std::size_t j) const {
    if (!m_body.empty()) return m_body[j];
    return particles[j].position - position;
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
    loadAndConvertMesh(filename, vertices, indicies, normals);
    n_triangles = (int)normals.size();
    if ((int)indicies.size() != 3 * n_triangles) throw std::runtime_error("DEMMesh: inconsistent STL data");
    computeBoundingSpheres();
}

namespace {
struct D3 {
    double x, y, z;
};
D3 d3(const float4& v) { return D3{v.x, v.y, v.z}; }
D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 add(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 mul(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
}  // namespace

// In double precision: the longest edge opposite corner a; right or obtuse
// at a (B^2 + C^2 <= A^2): centre = the edge's midpoint, radius = A / 2;
// else the circumcentre c + ((|al|^2 be - |be|^2 al) x (al x be)) / (2 |al x be|^2)
// (al = a - c, be = b - c) and the circumradius.
void DEMMesh::computeBoundingSpheres() {
    positions.clear();
    radii.clear();
    for (int t = 0; t < n_triangles; ++t) {
        D3 p[3] = {d3(vertices[indicies[3 * t]]), d3(vertices[indicies[3 * t + 1]]), d3(vertices[indicies[3 * t + 2]])};
        // corner k is opposite the edge between the other two
        double e[3];
        for (int k = 0; k < 3; ++k) {
            const D3 d = sub(p[(k + 1) % 3], p[(k + 2) % 3]);
            e[k] = std::sqrt(dot(d, d));
        }
        int ia = 0;
        for (int k = 1; k < 3; ++k)
            if (e[k] > e[ia]) ia = k;
        const D3 a = p[ia], b = p[(ia + 1) % 3], c = p[(ia + 2) % 3];
        const double A = e[ia], B = e[(ia + 1) % 3], C = e[(ia + 2) % 3];
        D3 centre;
        double radius;
        if (B * B + C * C <= A * A) {
            centre = mul(add(b, c), 0.5);
            radius = 0.5 * A;
        } else {
            const D3 al = sub(a, c), be = sub(b, c), x = cross(al, be);
            centre = add(c, mul(cross(sub(mul(be, dot(al, al)), mul(al, dot(be, be))), x), 0.5 / dot(x, x)));
            const D3 r = sub(a, centre);
            radius = std::sqrt(dot(r, r));
        }
        positions.push_back(make_float4((float)centre.x, (float)centre.y, (float)centre.z, 0.0f));
        radii.push_back((float)radius);
    }
}

void DEMMesh::worldTriangle(int t, float out[9]) const {
    const double w = quat.x, x = quat.y, y = quat.z, z = quat.w;
    const double nq = std::sqrt(w * w + x * x + y * y + z * z);
    const double qw = w / nq, qx = x / nq, qy = y / nq, qz = z / nq;
    const double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy),
                         2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx),
                         2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)};
    for (int k = 0; k < 3; ++k) {
        const float4 v = vertices[indicies[3 * t + k]];
        const double pos[3] = {position.x, position.y, position.z};
        for (int r = 0; r < 3; ++r)
            out[3 * k + r] = (float)(R[3 * r] * v.x + R[3 * r + 1] * v.y + R[3 * r + 2] * v.z + pos[r]);
    }
}

DEMHull::DEMHull(unsigned int id, const std::vector<float4>& verts, Scalar mass_, float4 position_, float4 velocity_,
                 float4 quaternion)
    : vertices(verts), mass(mass_), position(position_), velocity(velocity_), quat(quaternion),
      angular_momentum(make_float4(0.0f, 0.0f, 0.0f, 0.0f)), m_id(id) {
    if (vertices.size() < 3) throw std::invalid_argument("DEMHull: at least three vertices");
    if (!(mass >= 0.0f)) throw std::invalid_argument("DEMHull: negative mass");
}

}  // namespace STARDUST
