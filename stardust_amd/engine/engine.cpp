// DEMEngine forwarding to libstardust_hip.so (reference: src/engine/engine.cpp:601-609,
// src/engine/cuda/simulation_kernel.cu:52-153).
#include "engine.hpp"

#include <cmath>
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
}

void DEMEngine::add(DEMParticle particle) {
    if (m_uploaded) throw std::runtime_error("DEMEngine::add: entities cannot be added after transfer()");
    entity_handler.addEntity(particle);
}

void DEMEngine::add(DEMMesh mesh) { entity_handler.addEntity(mesh); }

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

// Packs bodies for sd_clump_upload: centre = entity position, spheres relative
// to it with radius diameter / 2, inertia about the centre from the member
// spheres (own 2/5 m r^2 plus the parallel-axis term, in double).
void DEMEngine::transfer() {
    EntityHandler& h = entity_handler;
    if (!h.prepared) throw std::runtime_error("DEMEngine::transfer: call prep() first");
    const int nb = h.getNumberOfParticles(), ns = h.getNumberOfSpheres();
    if (nb == 0) throw std::runtime_error("DEMEngine::transfer: no entities");
    up_pos4.assign(4 * (size_t)nb, 0.0f);
    up_vel4.assign(4 * (size_t)nb, 0.0f);
    up_quat4.assign(4 * (size_t)nb, 0.0f);
    up_angvel4.assign(4 * (size_t)nb, 0.0f);
    up_inertia6.assign(6 * (size_t)nb, 0.0f);
    up_first.assign(nb, 0);
    up_count.assign(nb, 0);
    up_local4.assign(4 * (size_t)ns, 0.0f);
    for (int b = 0; b < nb; ++b) {
        const float4 X = h.rb_pos[b], V = h.rb_vel[b];
        const float m = h.rb_mass[b];
        const bool fixed = m >= m_static_mass;
        up_pos4[4 * b] = X.x;
        up_pos4[4 * b + 1] = X.y;
        up_pos4[4 * b + 2] = X.z;
        up_pos4[4 * b + 3] = fixed ? 0.0f : m;
        up_vel4[4 * b] = V.x;
        up_vel4[4 * b + 1] = V.y;
        up_vel4[4 * b + 2] = V.z;
        up_quat4[4 * b] = 1.0f;
        up_first[b] = h.entity_start[b];
        up_count[b] = h.entity_length[b];
        double I[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < h.entity_length[b]; ++j) {
            const int k = h.entity_start[b] + j;
            const float4 d = h.p_rel[k];
            const float r = 0.5f * h.p_size[k];
            up_local4[4 * (size_t)k] = d.x;
            up_local4[4 * (size_t)k + 1] = d.y;
            up_local4[4 * (size_t)k + 2] = d.z;
            up_local4[4 * (size_t)k + 3] = r;
            const double mj = h.p_mass[k], dx = d.x, dy = d.y, dz = d.z;
            const double own = 0.4 * mj * (double)r * (double)r;
            I[0] += own + mj * (dy * dy + dz * dz);
            I[1] += own + mj * (dx * dx + dz * dz);
            I[2] += own + mj * (dx * dx + dy * dy);
            I[3] += -mj * dx * dy;
            I[4] += -mj * dx * dz;
            I[5] += -mj * dy * dz;
        }
        for (int q = 0; q < 6; ++q) up_inertia6[6 * (size_t)b + q] = (float)I[q];
    }
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
    check(sd_clump_upload(m_ctx, nb, up_pos4.data(), up_vel4.data(), up_quat4.data(), up_angvel4.data(),
                          up_inertia6.data(), up_first.data(), up_count.data(), ns, up_local4.data()),
          "DEMEngine::transfer");
    m_uploaded = true;
    m_dt = 0.0f;
    m_steps = 0;
}

void DEMEngine::ensure_context(float dt) {
    if (dt == m_dt) return;
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
    check(sd_clump_step(m_ctx, 1), "DEMEngine::update");
    ++m_steps;
    is_first_step = false;
}

void DEMEngine::sync() {
    if (!m_uploaded) throw std::runtime_error("DEMEngine::sync: nothing transferred");
    EntityHandler& h = entity_handler;
    const int nb = h.getNumberOfParticles(), ns = h.getNumberOfSpheres();
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
