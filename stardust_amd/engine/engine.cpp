// DEMEngine forwarding to libstardust_hip.so (reference: src/engine/engine.cpp:601-609,
// src/engine/cuda/simulation_kernel.cu:52-153).
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "json.hpp"
#include "sd_api.h"

namespace STARDUST {

namespace {

void check(sd_status st, const char* what) {
    if (st != SD_OK) throw std::runtime_error(std::string(what) + ": " + sd_last_error());
}

}  // namespace

DEMEngine::DEMEngine(engineConfig config, Scalar domain) : engineConfiguration(config), m_domain(domain) {
    if (config.collision_engine != SAP && config.collision_engine != SHP)
        throw std::invalid_argument("DEMEngine: unknown collision engine");
    if (config.device != GPU) throw std::invalid_argument("DEMEngine: this engine runs on the GPU only");
}

DEMEngine::~DEMEngine() {
    if (m_ctx) sd_clump_destroy(m_ctx);
    if (m_sctx) sd_ctx_destroy(m_sctx);
    if (m_hctx) sd_hull_destroy(m_hctx);
}

void DEMEngine::cleanBuffers() {
    if (m_ctx) sd_clump_destroy(m_ctx);
    if (m_sctx) sd_ctx_destroy(m_sctx);
    if (m_hctx) sd_hull_destroy(m_hctx);
    m_ctx = nullptr;
    m_sctx = nullptr;
    m_hctx = nullptr;
    m_uploaded = false;
}

void DEMEngine::add(DEMParticle particle) {
    if (m_uploaded) throw std::runtime_error("DEMEngine::add: entities cannot be added after transfer()");
    entity_handler.addEntity(particle);
}

void DEMEngine::add(DEMMesh mesh) {
    if (m_uploaded) throw std::runtime_error("DEMEngine::add: meshes cannot be added after transfer()");
    entity_handler.addEntity(mesh);
}

void DEMEngine::add(DEMHull hull) {
    if (m_uploaded) throw std::runtime_error("DEMEngine::add: hulls cannot be added after transfer()");
    entity_handler.addEntity(hull);
}

void DEMEngine::prep() { entity_handler.prepareArrays(); }

void DEMEngine::setContactLaw(float kn, float damping_ratio, float mu, float kt_ratio) {
    m_kn = kn;
    m_zeta = damping_ratio;
    m_mu = mu;
    m_kt_ratio = kt_ratio;
    m_dt = 0.0f;  // parameters are pushed on the next update
}

void DEMEngine::setGravity(float gx, float gy, float gz) {
    m_g[0] = gx;
    m_g[1] = gy;
    m_g[2] = gz;
    m_dt = 0.0f;
}

void DEMEngine::setStaticMassThreshold(float threshold) {
    if (m_uploaded) throw std::runtime_error("DEMEngine::setStaticMassThreshold: call before transfer()");
    m_static_mass = threshold;
}

namespace {

// Upload arrays of sd_clump_upload / sd_clump_add_bodies.
struct ClumpPack {
    std::vector<float> pos4, vel4, quat4, angvel4, inertia6, local4;
    std::vector<int> first, count;
};

// Packs the bodies of a prepared handler: centre = entity position, spheres
// relative to it with radius diameter / 2, inertia about the centre from the
// member spheres (own 2/5 m r^2 plus the parallel-axis term, in double),
// bodies at or above the static mass threshold uploaded with mass 0.
struct ClumpArrays {  // a handler's body and sphere arrays (read by DEMEngine members)
    const std::vector<int>& estart;
    const std::vector<int>& elen;
    const std::vector<float4>& rb_pos;
    const std::vector<float4>& rb_vel;
    const std::vector<float4>& rb_quat;
    const std::vector<float4>& rb_angvel;
    const std::vector<float>& rb_mass;
    const std::vector<float4>& p_init_rel;
    const std::vector<float>& p_size;
    const std::vector<float>& p_mass;
};

void pack_clumps(const EntityHandler& h, const ClumpArrays& a, float static_mass, ClumpPack& u) {
    const std::vector<int>& estart = a.estart;
    const std::vector<int>& elen = a.elen;
    const int nb = h.getNumberOfParticles(), ns = h.getNumberOfSpheres();
    u.pos4.assign(4 * (size_t)nb, 0.0f);
    u.vel4.assign(4 * (size_t)nb, 0.0f);
    u.quat4.assign(4 * (size_t)nb, 0.0f);
    u.angvel4.assign(4 * (size_t)nb, 0.0f);
    u.inertia6.assign(6 * (size_t)nb, 0.0f);
    u.first.assign(nb, 0);
    u.count.assign(nb, 0);
    u.local4.assign(4 * (size_t)ns, 0.0f);
    for (int b = 0; b < nb; ++b) {
        const float4 X = a.rb_pos[b], V = a.rb_vel[b], q = a.rb_quat[b], w = a.rb_angvel[b];
        const float m = a.rb_mass[b];
        const bool fixed = m >= static_mass;
        u.pos4[4 * b] = X.x;
        u.pos4[4 * b + 1] = X.y;
        u.pos4[4 * b + 2] = X.z;
        u.pos4[4 * b + 3] = fixed ? 0.0f : m;
        u.vel4[4 * b] = V.x;
        u.vel4[4 * b + 1] = V.y;
        u.vel4[4 * b + 2] = V.z;
        u.quat4[4 * b] = q.x;  // (w, x, y, z), as the reference's float4 stores it
        u.quat4[4 * b + 1] = q.y;
        u.quat4[4 * b + 2] = q.z;
        u.quat4[4 * b + 3] = q.w;
        u.angvel4[4 * b] = w.x;
        u.angvel4[4 * b + 1] = w.y;
        u.angvel4[4 * b + 2] = w.z;
        u.first[b] = estart[b];
        u.count[b] = elen[b];
        double I[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < elen[b]; ++j) {
            const int k = estart[b] + j;
            const float4 d = a.p_init_rel[k];  // body frame
            const float r = 0.5f * a.p_size[k];
            u.local4[4 * (size_t)k] = d.x;
            u.local4[4 * (size_t)k + 1] = d.y;
            u.local4[4 * (size_t)k + 2] = d.z;
            u.local4[4 * (size_t)k + 3] = r;
            const double mj = a.p_mass[k], dx = d.x, dy = d.y, dz = d.z;
            const double own = 0.4 * mj * (double)r * (double)r;
            I[0] += own + mj * (dy * dy + dz * dz);
            I[1] += own + mj * (dx * dx + dz * dz);
            I[2] += own + mj * (dx * dx + dy * dy);
            I[3] += -mj * dx * dy;
            I[4] += -mj * dx * dz;
            I[5] += -mj * dy * dz;
        }
        for (int c = 0; c < 6; ++c) u.inertia6[6 * (size_t)b + c] = (float)I[c];
    }
}

}  // namespace

void DEMEngine::transfer() {
    EntityHandler& h = entity_handler;
    if (!h.prepared) throw std::runtime_error("DEMEngine::transfer: call prep() first");
    const int nb = h.getNumberOfParticles();
    if (nb + h.getNumberOfHulls() + h.getNumberOfMeshes() == 0)
        throw std::runtime_error("DEMEngine::transfer: no entities");
    bool single = true;
    for (int b = 0; b < nb && single; ++b) single = h.entity_length[b] == 1;
    m_hull_mode = h.getNumberOfHulls() + h.getNumberOfMeshes() > 0;
    m_sphere_mode = single && !m_hull_mode;
    if (m_hull_mode) transfer_bodies();
    else if (single) transfer_spheres();
    else transfer_clumps();
    m_uploaded = true;
    m_dt = 0.0f;
    m_steps = 0;
    bind_runtime();
}

void DEMEngine::bind_runtime() {
    entity_handler.runtime_add = [this](const DEMParticle& e) { addRuntime(e); };
    entity_handler.runtime_remove = [this](int idx) { removeRuntime(idx); };
}

void DEMEngine::addRuntime(const DEMParticle& e) {
    if (m_hull_mode) throw std::runtime_error("DEMEngine::addRuntime: not available on the rigid-body path");
    if (!m_uploaded) throw std::runtime_error("DEMEngine::addRuntime: call transfer() first");
    EntityHandler one;
    one.addEntity(e);
    one.prepareArrays();
    if (m_sphere_mode) {
        if (one.getNumberOfSpheres() != 1)
            throw std::runtime_error("DEMEngine::addRuntime: this engine steps single spheres; the entity has " +
                                     std::to_string(one.getNumberOfSpheres()));
        const float4 X = one.rb_pos[0], V = one.rb_vel[0];
        const float r = 0.5f * one.p_size[0], mass = one.rb_mass[0] >= m_static_mass ? 0.0f : one.rb_mass[0];
        check(sd_ctx_add_spheres(m_sctx, 1, &X.x, &X.y, &X.z, &V.x, &V.y, &V.z, &r, &mass), "DEMEngine::addRuntime");
    } else {
        ClumpPack u;
        pack_clumps(one,
                    ClumpArrays{one.entity_start, one.entity_length, one.rb_pos, one.rb_vel, one.rb_quat, one.rb_angvel,
                                one.rb_mass, one.p_init_rel, one.p_size, one.p_mass},
                    m_static_mass, u);
        check(sd_clump_add_bodies(m_ctx, 1, u.pos4.data(), u.vel4.data(), u.quat4.data(), u.angvel4.data(),
                                  u.inertia6.data(), u.first.data(), u.count.data(), one.getNumberOfSpheres(),
                                  u.local4.data()),
              "DEMEngine::addRuntime");
    }
    entity_handler.addEntity(e);
    entity_handler.prepareArrays();
    sync();
}

void DEMEngine::removeRuntime(int idx) {
    if (m_hull_mode) throw std::runtime_error("DEMEngine::removeRuntime: not available on the rigid-body path");
    if (!m_uploaded) throw std::runtime_error("DEMEngine::removeRuntime: call transfer() first");
    if (idx < 0 || idx >= entity_handler.getNumberOfParticles())
        throw std::out_of_range("DEMEngine::removeRuntime: no entity " + std::to_string(idx));
    if (m_sphere_mode) check(sd_ctx_remove_spheres(m_sctx, &idx, 1), "DEMEngine::removeRuntime");
    else check(sd_clump_remove_bodies(m_ctx, &idx, 1), "DEMEngine::removeRuntime");
    entity_handler.removeEntity(0, idx);
    entity_handler.prepareArrays();
    sync();
}

namespace {
constexpr char ENGINE_MAGIC[8] = {'S', 'D', 'E', 'N', 'G', 'I', 'N', '1'};
}

void DEMEngine::saveState(const std::string& path) {
    if (m_hull_mode) throw std::runtime_error("DEMEngine::saveState: not available on the rigid-body path");
    if (!m_uploaded) throw std::runtime_error("DEMEngine::saveState: call transfer() first");
    size_t bytes = 0;
    if (m_sphere_mode) check(sd_ctx_state_bytes(m_sctx, &bytes), "DEMEngine::saveState");
    else check(sd_clump_state_bytes(m_ctx, &bytes), "DEMEngine::saveState");
    std::vector<char> buf(bytes);
    if (m_sphere_mode) check(sd_ctx_save(m_sctx, buf.data(), bytes), "DEMEngine::saveState");
    else check(sd_clump_save(m_ctx, buf.data(), bytes), "DEMEngine::saveState");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("DEMEngine::saveState: cannot open " + path);
    const int32_t mode = m_sphere_mode ? 1 : 0;
    const int64_t steps = m_steps, nbytes = (int64_t)bytes;
    f.write(ENGINE_MAGIC, 8);
    f.write((const char*)&mode, sizeof mode);
    f.write((const char*)&steps, sizeof steps);
    f.write((const char*)&nbytes, sizeof nbytes);
    f.write(buf.data(), (std::streamsize)bytes);
    if (!f) throw std::runtime_error("DEMEngine::saveState: write failed on " + path);
}

void DEMEngine::loadState(const std::string& path) {
    if (m_hull_mode) throw std::runtime_error("DEMEngine::loadState: not available on the rigid-body path");
    if (!m_uploaded) throw std::runtime_error("DEMEngine::loadState: transfer() the same entities first");
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("DEMEngine::loadState: cannot open " + path);
    char magic[8];
    int32_t mode = -1;
    int64_t steps = 0, nbytes = 0;
    f.read(magic, 8);
    f.read((char*)&mode, sizeof mode);
    f.read((char*)&steps, sizeof steps);
    f.read((char*)&nbytes, sizeof nbytes);
    if (!f || std::memcmp(magic, ENGINE_MAGIC, 8) != 0 || nbytes < 0)
        throw std::runtime_error("DEMEngine::loadState: " + path + " is not an engine checkpoint");
    if (mode != (m_sphere_mode ? 1 : 0))
        throw std::runtime_error("DEMEngine::loadState: the checkpoint is of the other step path");
    std::vector<char> buf((size_t)nbytes);
    f.read(buf.data(), (std::streamsize)nbytes);
    if (!f) throw std::runtime_error("DEMEngine::loadState: " + path + " is truncated");
    if (m_sphere_mode) check(sd_ctx_load(m_sctx, buf.data(), buf.size()), "DEMEngine::loadState");
    else check(sd_clump_load(m_ctx, buf.data(), buf.size()), "DEMEngine::loadState");
    m_steps = steps;
    is_first_step = steps == 0;
    sync();
}

namespace {

// Legacy VTK binary data is big-endian.
void put_be(std::ofstream& f, const void* p) {
    const unsigned char* b = (const unsigned char*)p;
    const unsigned char r[4] = {b[3], b[2], b[1], b[0]};
    f.write((const char*)r, 4);
}

}  // namespace

void DEMEngine::writeVTK(const std::string& path) {
    if (m_hull_mode) throw std::runtime_error("DEMEngine::writeVTK: not available on the rigid-body path");
    if (!m_uploaded) throw std::runtime_error("DEMEngine::writeVTK: call transfer() first");
    sync();
    EntityHandler& h = entity_handler;
    const int ns = h.getNumberOfSpheres();
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("DEMEngine::writeVTK: cannot open " + path);
    f << "# vtk DataFile Version 3.0\nstardust_amd spheres, step " << m_steps << "\nBINARY\nDATASET POLYDATA\n";
    f << "POINTS " << ns << " float\n";
    for (int k = 0; k < ns; ++k) {
        const float4 p = h.p_pos[k];
        put_be(f, &p.x);
        put_be(f, &p.y);
        put_be(f, &p.z);
    }
    f << "\nVERTICES " << ns << " " << 2 * (long long)ns << "\n";
    for (int k = 0; k < ns; ++k) {
        const int one = 1;
        put_be(f, &one);
        put_be(f, &k);
    }
    f << "\nPOINT_DATA " << ns << "\nSCALARS radius float 1\nLOOKUP_TABLE default\n";
    for (int k = 0; k < ns; ++k) {
        const float r = 0.5f * h.p_size[k];
        put_be(f, &r);
    }
    f << "\nVECTORS body_velocity float\n";
    for (int k = 0; k < ns; ++k) {
        const float4 v = h.rb_vel[h.particle_to_rigid[k]];
        put_be(f, &v.x);
        put_be(f, &v.y);
        put_be(f, &v.z);
    }
    f << "\nSCALARS body int 1\nLOOKUP_TABLE default\n";
    for (int k = 0; k < ns; ++k) put_be(f, &h.particle_to_rigid[k]);
    f << "\n";
    if (!f) throw std::runtime_error("DEMEngine::writeVTK: write failed on " + path);
}

namespace {

sd_contact_params sphere_law(float kn, float zeta, float mu, float kt_ratio, float dt, const float* g,
                             const DEMEngine::SphereGrid& sg) {
    sd_contact_params cp{};
    cp.kn = kn;
    cp.damping_ratio = zeta;
    cp.kt_ratio = kt_ratio;
    cp.mu = mu;
    cp.dt = dt;
    for (int a = 0; a < 3; ++a) cp.gravity[a] = g[a];
    cp.walls = sg.walls ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        cp.wall_lo[a] = sg.wall_lo[a];
        cp.wall_hi[a] = sg.wall_hi[a];
    }
    cp.max_contacts = 24;
    return cp;
}

}  // namespace

// Single-sphere entities to a sphere context: radius = diameter / 2, the
// entity's mass (0 = static when it reaches the static threshold), grid cell
// = the largest diameter; walls on the faces of [0, domain]^3 when every
// sphere lies inside it, else an unbounded grid over the spheres' extent.
void DEMEngine::transfer_spheres() {
    EntityHandler& h = entity_handler;
    const int n = h.getNumberOfParticles();
    for (int b = 0; b < n; ++b) {
        const float4 w = h.rb_angvel[b];
        if (w.x != 0.0f || w.y != 0.0f || w.z != 0.0f)
            throw std::invalid_argument("DEMEngine::transfer: entity " + std::to_string(b) +
                                        " has an initial angular velocity; single-sphere scenes (the sphere "
                                        "path) start every sphere without spin");
    }
    std::vector<float> x(n), y(n), z(n), vx(n), vy(n), vz(n), r(n), m(n);
    float rmax = 0.0f;
    bool inside = m_domain > 0.0f;
    float lo[3] = {3.4e38f, 3.4e38f, 3.4e38f}, hi[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
    for (int b = 0; b < n; ++b) {
        const int k = h.entity_start[b];
        const float4 X = h.rb_pos[b], V = h.rb_vel[b];
        const float mass = h.rb_mass[b];
        x[b] = X.x;
        y[b] = X.y;
        z[b] = X.z;
        vx[b] = V.x;
        vy[b] = V.y;
        vz[b] = V.z;
        r[b] = 0.5f * h.p_size[k];
        m[b] = mass >= m_static_mass ? 0.0f : mass;
        if (!(r[b] > 0.0f)) throw std::runtime_error("DEMEngine::transfer: sphere diameter must be positive");
        rmax = std::max(rmax, r[b]);
        const float p[3] = {X.x, X.y, X.z};
        for (int a = 0; a < 3; ++a) {
            inside = inside && p[a] >= r[b] && p[a] <= m_domain - r[b];
            lo[a] = std::min(lo[a], p[a] - r[b]);
            hi[a] = std::max(hi[a], p[a] + r[b]);
        }
    }
    SphereGrid& sg = m_sgrid;
    sg.cell = 2.0f * rmax;
    sg.walls = inside;
    for (int a = 0; a < 3; ++a) {
        if (inside) {
            sg.lo[a] = 0.0f;
            sg.wall_lo[a] = 0.0f;
            sg.wall_hi[a] = m_domain;
        } else {
            sg.lo[a] = lo[a] - 2.0f * sg.cell;
            hi[a] = hi[a] + 2.0f * sg.cell;
            sg.wall_lo[a] = sg.wall_hi[a] = 0.0f;
        }
    }
    const float ext[3] = {inside ? m_domain : hi[0] - sg.lo[0], inside ? m_domain : hi[1] - sg.lo[1],
                          inside ? m_domain : hi[2] - sg.lo[2]};
    for (;;) {  // at most 2^26 cells: a sparse scene gets coarser cells (any cell >= the largest diameter works)
        double cells = 1.0;
        for (int a = 0; a < 3; ++a) {
            sg.n[a] = std::max(1, (int)std::ceil(ext[a] / sg.cell));
            cells *= sg.n[a];
        }
        if (cells <= (double)(1 << 26)) break;
        sg.cell *= 2.0f;
    }
    sd_sphere_params p{};
    for (int a = 0; a < 3; ++a) {
        p.grid.lo[a] = sg.lo[a];
        p.grid.n[a] = sg.n[a];
    }
    p.grid.cell = sg.cell;
    p.contact = sphere_law(m_kn, m_zeta, m_mu, m_kt_ratio, 1e-5f, m_g, sg);  // dt of the first update replaces it
    p.key64 = 1;
    p.use_graph = 1;
    p.tile_items = 0;
    if (m_sctx) sd_ctx_destroy(m_sctx);
    m_sctx = nullptr;
    check(sd_ctx_create(&p, &m_sctx), "DEMEngine::transfer");
    check(sd_ctx_upload_spheres(m_sctx, n, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(), r.data(),
                                m.data()),
          "DEMEngine::transfer");
}


void DEMEngine::transfer_clumps() {
    EntityHandler& h = entity_handler;
    const int nb = h.getNumberOfParticles(), ns = h.getNumberOfSpheres();
    ClumpPack u;
    pack_clumps(h,
                ClumpArrays{h.entity_start, h.entity_length, h.rb_pos, h.rb_vel, h.rb_quat, h.rb_angvel, h.rb_mass,
                            h.p_init_rel, h.p_size, h.p_mass},
                m_static_mass, u);
    if (!m_ctx) {
        sd_clump_params p{};
        p.contact.kn = m_kn;
        p.contact.damping_ratio = m_zeta;
        p.contact.kt_ratio = m_kt_ratio;
        p.contact.mu = m_mu;
        p.contact.dt = 1e-5f;  // replaced by the dt of the first update
        for (int a = 0; a < 3; ++a) p.contact.gravity[a] = m_g[a];
        p.contact.max_contacts = 16;
        check(sd_clump_create(&p, &m_ctx), "DEMEngine::transfer");
    }
    check(sd_clump_upload(m_ctx, nb, u.pos4.data(), u.vel4.data(), u.quat4.data(), u.angvel4.data(),
                          u.inertia6.data(), u.first.data(), u.count.data(), ns, u.local4.data()),
          "DEMEngine::transfer");
}

namespace {

sd_hull_params hull_law(float kn, float zeta, float mu, float dt, const float* g) {
    sd_hull_params p{};
    p.contact.kn = kn;
    p.contact.damping_ratio = zeta;
    p.contact.mu = mu;
    p.contact.dt = dt;
    for (int a = 0; a < 3; ++a) p.contact.gravity[a] = g[a];
    p.contact.max_contacts = 16;
    p.max_iter = 64;
    p.pairs_per_body = 32;
    p.use_graph = 1;
    return p;
}

}  // namespace

// Bodies in this order: single-sphere particles (radius diameter / 2, solid
// sphere inertia), hulls (inertia of equal point masses at the vertices, in
// double), mesh triangles (static; centre = the corners' mean, corners
// relative to it); masses at or above the static threshold are uploaded as 0.
void DEMEngine::transfer_bodies() {
    EntityHandler& h = entity_handler;
    const int np = h.getNumberOfParticles();
    std::vector<float> pos4, vel4, quat4, angvel4, inertia6, radius, verts4;
    std::vector<int> vstart, nv;
    auto body = [&](float x, float y, float z, float m, const float4& v, const float4& q, const double I[6], float r,
                    int vs, int n) {
        const float mm = m >= m_static_mass ? 0.0f : m;
        pos4.insert(pos4.end(), {x, y, z, mm});
        vel4.insert(vel4.end(), {v.x, v.y, v.z, 0.0f});
        quat4.insert(quat4.end(), {q.x, q.y, q.z, q.w});
        angvel4.insert(angvel4.end(), {0.0f, 0.0f, 0.0f, 0.0f});
        for (int k = 0; k < 6; ++k) inertia6.push_back((float)I[k]);
        radius.push_back(r);
        vstart.push_back(vs);
        nv.push_back(n);
    };
    const float4 ident = make_float4(1.0f, 0.0f, 0.0f, 0.0f), zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int b = 0; b < np; ++b) {
        if (h.entity_length[b] != 1)
            throw std::runtime_error("DEMEngine::transfer: multi-sphere particles cannot share a scene with meshes "
                                     "or hulls");
        const int k = h.entity_start[b];
        const float r = 0.5f * h.p_size[k], m = h.rb_mass[b];
        const double Is = 0.4 * (double)m * r * r;
        const double I[6] = {Is, Is, Is, 0.0, 0.0, 0.0};
        body(h.rb_pos[b].x, h.rb_pos[b].y, h.rb_pos[b].z, m, h.rb_vel[b], ident, I, r, 0, 0);
    }
    for (const DEMHull& hl : h.getHulls()) {
        const int vs = (int)(verts4.size() / 4), n = (int)hl.vertices.size();
        double I[6] = {0, 0, 0, 0, 0, 0};
        const double pm = (double)hl.mass / n;
        for (const float4& v : hl.vertices) {
            verts4.insert(verts4.end(), {v.x, v.y, v.z, 0.0f});
            const double x = v.x, y = v.y, z = v.z;
            I[0] += pm * (y * y + z * z);
            I[1] += pm * (x * x + z * z);
            I[2] += pm * (x * x + y * y);
            I[3] -= pm * x * y;
            I[4] -= pm * x * z;
            I[5] -= pm * y * z;
        }
        body(hl.position.x, hl.position.y, hl.position.z, hl.mass, hl.velocity, hl.quat, I, 0.0f, vs, n);
    }
    const double I0[6] = {0, 0, 0, 0, 0, 0};
    for (const DEMMesh& mesh : h.getMeshes()) {
        for (int t = 0; t < mesh.n_triangles; ++t) {
            float c[9];
            mesh.worldTriangle(t, c);
            const float cx = (float)(((double)c[0] + c[3] + c[6]) / 3.0), cy = (float)(((double)c[1] + c[4] + c[7]) / 3.0),
                        cz = (float)(((double)c[2] + c[5] + c[8]) / 3.0);
            const int vs = (int)(verts4.size() / 4);
            for (int k = 0; k < 3; ++k) verts4.insert(verts4.end(), {c[3 * k] - cx, c[3 * k + 1] - cy, c[3 * k + 2] - cz, 0.0f});
            body(cx, cy, cz, 0.0f, zero, ident, I0, 0.0f, vs, 3);
        }
    }
    const int n = (int)radius.size();
    if (!m_hctx) {
        const sd_hull_params p = hull_law(m_kn, m_zeta, m_mu, 1e-5f, m_g);  // dt of the first update replaces it
        check(sd_hull_create(&p, &m_hctx), "DEMEngine::transfer");
    }
    check(sd_hull_upload(m_hctx, n, pos4.data(), vel4.data(), quat4.data(), angvel4.data(), inertia6.data(),
                         radius.data(), vstart.data(), nv.data(), (int)(verts4.size() / 4),
                         verts4.empty() ? nullptr : verts4.data(), nullptr, 1),
          "DEMEngine::transfer");
}

void DEMEngine::ensure_context(float dt) {
    if (dt == m_dt) return;
    if (m_hull_mode) {
        const sd_hull_params p = hull_law(m_kn, m_zeta, m_mu, dt, m_g);
        check(sd_hull_set_params(m_hctx, &p), "DEMEngine::update");
        m_dt = dt;
        return;
    }
    if (m_sphere_mode) {
        const sd_contact_params cp = sphere_law(m_kn, m_zeta, m_mu, m_kt_ratio, dt, m_g, m_sgrid);
        check(sd_ctx_set_contact(m_sctx, &cp), "DEMEngine::update");
        m_dt = dt;
        return;
    }
    sd_clump_params p{};
    p.contact.kn = m_kn;
    p.contact.damping_ratio = m_zeta;
    p.contact.kt_ratio = m_kt_ratio;
    p.contact.mu = m_mu;
    p.contact.dt = dt;
    for (int a = 0; a < 3; ++a) p.contact.gravity[a] = m_g[a];
    p.contact.max_contacts = 16;
    check(sd_clump_set_params(m_ctx, &p), "DEMEngine::update");
    m_dt = dt;
}

void DEMEngine::update(float dt) {
    if (!m_uploaded) throw std::runtime_error("DEMEngine::update: call prep() and transfer() first");
    if (!(dt > 0.0f)) throw std::invalid_argument("DEMEngine::update: dt must be positive");
    ensure_context(dt);
    if (m_hull_mode) check(sd_hull_step(m_hctx, 1), "DEMEngine::update");
    else if (m_sphere_mode) check(sd_ctx_step(m_sctx, 1), "DEMEngine::update");
    else check(sd_clump_step(m_ctx, 1), "DEMEngine::update");
    ++m_steps;
    is_first_step = false;
}

void DEMEngine::sync() {
    if (!m_uploaded) throw std::runtime_error("DEMEngine::sync: nothing transferred");
    EntityHandler& h = entity_handler;
    const int nb = h.getNumberOfParticles(), ns = h.getNumberOfSpheres();
    if (m_hull_mode) {
        const int nh = h.getNumberOfHulls();
        int n = nb + nh;
        for (const DEMMesh& mesh : h.getMeshes()) n += mesh.n_triangles;
        std::vector<float> pos4(4 * (size_t)n), vel4(4 * (size_t)n), quat4(4 * (size_t)n), angm4(4 * (size_t)n);
        check(sd_hull_download(m_hctx, pos4.data(), vel4.data(), quat4.data(), angm4.data()), "DEMEngine::sync");
        auto f4 = [](const std::vector<float>& a, int i) {
            return make_float4(a[4 * (size_t)i], a[4 * (size_t)i + 1], a[4 * (size_t)i + 2], a[4 * (size_t)i + 3]);
        };
        for (int b = 0; b < nb; ++b) {
            const int k = h.entity_start[b];
            h.rb_pos[b] = f4(pos4, b);
            h.rb_vel[b] = f4(vel4, b);
            h.rb_quat[b] = f4(quat4, b);
            h.rb_angmom[b] = f4(angm4, b);
            h.rb_linmom[b] = h.rb_vel[b] * h.rb_mass[b];
            h.p_pos[k] = h.rb_pos[b];
        }
        for (int j = 0; j < nh; ++j) {
            DEMHull& hl = h.getHulls()[j];
            hl.position = f4(pos4, nb + j);
            hl.velocity = f4(vel4, nb + j);
            hl.quat = f4(quat4, nb + j);
            hl.angular_momentum = f4(angm4, nb + j);
        }
        return;
    }
    if (m_sphere_mode) {
        std::vector<float> x(nb), y(nb), z(nb), vx(nb), vy(nb), vz(nb), wx(nb), wy(nb), wz(nb);
        check(sd_ctx_download_spheres(m_sctx, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(),
                                      wx.data(), wy.data(), wz.data()),
              "DEMEngine::sync");
        for (int b = 0; b < nb; ++b) {
            const int k = h.entity_start[b];
            const float m = h.rb_mass[b], r = 0.5f * h.p_size[k];
            const float I = 0.4f * m * r * r;
            h.rb_pos[b] = make_float4(x[b], y[b], z[b], 0.0f);
            h.rb_vel[b] = make_float4(vx[b], vy[b], vz[b], 0.0f);
            h.rb_linmom[b] = h.rb_vel[b] * m;
            h.rb_angmom[b] = make_float4(I * wx[b], I * wy[b], I * wz[b], 0.0f);
            h.p_pos[k] = h.rb_pos[b];
            h.p_rel[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        return;
    }
    std::vector<float> pos4(4 * (size_t)nb), vel4(4 * (size_t)nb), quat4(4 * (size_t)nb), angm4(4 * (size_t)nb);
    std::vector<float> sph(4 * (size_t)ns);
    check(sd_clump_download(m_ctx, pos4.data(), vel4.data(), quat4.data(), angm4.data()), "DEMEngine::sync");
    check(sd_clump_spheres(m_ctx, sph.data()), "DEMEngine::sync");
    for (int b = 0; b < nb; ++b) {
        h.rb_pos[b] = make_float4(pos4[4 * b], pos4[4 * b + 1], pos4[4 * b + 2], 0.0f);
        h.rb_vel[b] = make_float4(vel4[4 * b], vel4[4 * b + 1], vel4[4 * b + 2], 0.0f);
        h.rb_quat[b] = make_float4(quat4[4 * b], quat4[4 * b + 1], quat4[4 * b + 2], quat4[4 * b + 3]);
        h.rb_angmom[b] = make_float4(angm4[4 * b], angm4[4 * b + 1], angm4[4 * b + 2], 0.0f);
        const float m = h.rb_mass[b];
        h.rb_linmom[b] = h.rb_vel[b] * m;
    }
    // sphere positions of the last pose evaluation (before the last integration)
    for (int k = 0; k < ns; ++k) {
        h.p_pos[k] = make_float4(sph[4 * (size_t)k], sph[4 * (size_t)k + 1], sph[4 * (size_t)k + 2], 0.0f);
        const float4 X = h.rb_pos[h.particle_to_rigid[k]];
        h.p_rel[k] = h.p_pos[k] - X;
    }
}

void DEMEngine::loadJSONSetup(const std::string& filepath) {
    std::ifstream f(filepath);
    if (!f) throw std::runtime_error("DEMEngine::loadJSONSetup: cannot open " + filepath);
    std::stringstream buf;
    buf << f.rdbuf();
    const json::Value data = json::parse(buf.str());
    const json::Value& scene = data["scene"];
    if (scene.has("domain")) m_domain = (Scalar)scene["domain"].number();
    if (scene.has("gravity")) setGravity(0.0f, 0.0f, -(float)scene["gravity"].number());
    const json::Value& ents = scene["entities"];
    if (ents.kind != json::Value::Array) throw std::runtime_error("DEMEngine::loadJSONSetup: entities must be a list");
    for (const json::Value& e : ents.arr) {
        const std::string& type = e["type"].string();
        if (type != "generic") throw std::runtime_error("DEMEngine::loadJSONSetup: unsupported entity type " + type);
        const json::Value &p = e["position"], &v = e["velocity"];
        const float4 pos = make_float4((float)p[0].number(), (float)p[1].number(), (float)p[2].number(), 0.0f);
        const float4 vel = make_float4((float)v[0].number(), (float)v[1].number(), (float)v[2].number(), 0.0f);
        const Scalar size = (Scalar)e["size"].number();
        const Scalar mass = e.has("mass") ? (Scalar)e["mass"].number() : 1.0f;
        const unsigned id = (unsigned)entity_handler.getNumberOfParticles() + 1u;
        add(DEMParticle(id, 10, size, size, mass, pos, vel));
    }
}

}  // namespace STARDUST
