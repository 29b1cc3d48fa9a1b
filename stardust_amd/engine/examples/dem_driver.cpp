// Driver of the STARDUST engine API in the style of sandbox/testDEM.cpp:80-128
// (the reference's own usage pattern: build entities, add, prep, transfer,
// update(dt) in a loop). Prints JSON so that tests can check it.
//
//   sd_dem_driver pack                      the reference driver's entities
//                                           (testDEM.cpp:76, :95), packed arrays only
//   sd_dem_driver pack-json <scene.json>    a scene file, packed arrays only
//   sd_dem_driver run <steps> <dt> [scene.json]
//                                           steps the scene (default: two
//                                           approaching lattice bodies) on the GPU
//                                           and prints the packed input and the
//                                           final body state
//   sd_dem_driver run-spheres <steps> <dt> <n> <domain> <size> <lo> <hi> <seed>
//                                           testDEM.cpp:80-128's pattern: n single
//                                           spheres of diameter size at seeded
//                                           positions uniform in [lo, hi]^3 (masses
//                                           1..9, every 97th 100: static), engine
//                                           domain <domain>; steps on the GPU and
//                                           prints the packed input, the step path
//                                           with its grid, and the final state
//   sd_dem_driver edit-spheres <same arguments as run-spheres> <every>
//                                           as run-spheres, with runtime edits:
//                                           <steps> steps, removeEntityRuntime of
//                                           every <every>-th entity (highest index
//                                           first), <steps> steps, addEntityRuntime
//                                           of 3 spheres, <steps> steps
//   sd_dem_driver edit-clumps <steps> <dt>  the default "run" bodies plus a fourth,
//                                           with body 2 removed and a fifth body
//                                           added at runtime (steps between)
//   sd_dem_driver mesh-info <file.stl>      a DEMMesh at the origin: its vertices,
//                                           indices, normals and bounding spheres
//   sd_dem_driver mesh-drop <file.stl> <steps> <dt>
//                                           the mesh as static walls, spheres and a
//                                           cube hull dropped onto its top face
//                                           (rigid-body path); prints the inputs
//                                           and the final states
//   sd_dem_driver checkpoint <steps> <dt> <file> [vtk]
//                                           the default "run" bodies: <steps> steps,
//                                           saveState, <steps> steps (final A); a
//                                           second engine with the same entities,
//                                           loadState, <steps> steps (final B);
//                                           optionally writeVTK of the first
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"

using namespace STARDUST;

namespace {

void print_f4(const char* name, const float4* a, int n, bool last = false) {
    std::printf("\"%s\": [", name);
    for (int i = 0; i < n; ++i)
        std::printf("%s[%.9g, %.9g, %.9g, %.9g]", i ? ", " : "", a[i].x, a[i].y, a[i].z, a[i].w);
    std::printf("]%s\n", last ? "" : ",");
}
void print_f(const char* name, const float* a, int n) {
    std::printf("\"%s\": [", name);
    for (int i = 0; i < n; ++i) std::printf("%s%.9g", i ? ", " : "", a[i]);
    std::printf("],\n");
}
void print_i(const char* name, const int* a, int n) {
    std::printf("\"%s\": [", name);
    for (int i = 0; i < n; ++i) std::printf("%s%d", i ? ", " : "", a[i]);
    std::printf("],\n");
}

void print_pack(EntityHandler& h) {
    const int nb = h.getNumberOfParticles(), ns = h.getNumberOfSpheres();
    std::printf("\"bodies\": %d, \"spheres\": %d,\n", nb, ns);
    print_i("entity_start", h.h_entity_start_ptr, nb);
    print_i("entity_length", h.h_entity_length_ptr, nb);
    print_i("particle_to_rigid", h.h_particle_to_rigid_idx_ptr, ns);
    print_f("rigid_body_mass", h.h_rigid_body_mass_ptr, nb);
    print_f("particle_mass", h.h_particle_mass_ptr, ns);
    print_f("particle_size", h.h_particle_size_ptr, ns);
    std::printf("\"inertia\": [");
    for (int b = 0; b < nb; ++b) {
        std::printf("%s[", b ? ", " : "");
        for (int k = 0; k < 9; ++k) std::printf("%s%.9g", k ? ", " : "", h.h_rigid_body_inertia_tensor_ptr[b][k]);
        std::printf("]");
    }
    std::printf("],\n");
    print_f4("rigid_body_position", h.h_rigid_body_position_ptr, nb);
    print_f4("rigid_body_velocity", h.h_rigid_body_velocity_ptr, nb);
    print_f4("rigid_body_quaternion", h.h_rigid_body_quaternion_ptr, nb);
    print_f4("rigid_body_angular_velocity", h.h_rigid_body_angular_velocity_ptr, nb);
    print_f4("particle_relative_position", h.h_particle_relative_position_ptr, ns);
    print_f4("particle_init_relative_position", h.h_particle_init_relative_position_ptr, ns);
}

// Clumps from a text file: the body count, then per body a line
//   n px py pz vx vy vz qw qx qy qz wx wy wz
// and n lines "rx ry rz diameter mass" (body-frame sphere centres).
void read_clumps(const char* path, DEMEngine& engine) {
    std::FILE* f = std::fopen(path, "r");
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    int nb = 0;
    if (std::fscanf(f, "%d", &nb) != 1 || nb < 1) throw std::runtime_error("bad clump file");
    for (int b = 0; b < nb; ++b) {
        int n = 0;
        float v[13];
        if (std::fscanf(f, "%d", &n) != 1 || n < 1) throw std::runtime_error("bad clump file");
        for (float& x : v)
            if (std::fscanf(f, "%f", &x) != 1) throw std::runtime_error("bad clump file");
        std::vector<DEMSphere> sph(n);
        for (DEMSphere& q : sph) {
            float r[5];
            for (float& x : r)
                if (std::fscanf(f, "%f", &x) != 1) throw std::runtime_error("bad clump file");
            q.position = make_float4(r[0], r[1], r[2], 0.0f);
            q.size = r[3];
            q.mass = r[4];
        }
        engine.add(DEMParticle((unsigned)b, sph, make_float4(v[0], v[1], v[2], 0.0f), make_float4(v[3], v[4], v[5], 0.0f),
                               make_float4(v[6], v[7], v[8], v[9]), make_float4(v[10], v[11], v[12], 0.0f)));
    }
    std::fclose(f);
}

// two 2x2x2 lattice bodies approaching along x, slightly offset, and a static sphere
void two_bodies(DEMEngine& engine) {
    engine.add(DEMParticle(0, 2, 0.04f, 0.02f, 0.5f, make_float4(0.0f, 0.0f, 0.0f, 0.0f),
                           make_float4(1.0f, 0.0f, 0.0f, 0.0f)));
    engine.add(DEMParticle(1, 2, 0.04f, 0.02f, 0.5f, make_float4(0.032f, 0.003f, 0.002f, 0.0f),
                           make_float4(-1.0f, 0.0f, 0.0f, 0.0f)));
    engine.add(DEMParticle(2, 1, 0.02f, 0.02f, 50.0f, make_float4(0.016f, -0.02f, 0.0f, 0.0f),
                           make_float4(0.0f, 0.0f, 0.0f, 0.0f)));  // mass >= 10: static
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s pack | pack-json <scene> | run <steps> <dt> [scene]\n", argv[0]);
        return 2;
    }
    try {
        const std::string mode = argv[1];
        const engineConfig config{SAP, GPU};
        if (mode == "pack") {
            DEMEngine engine(config, 0.005f);
            const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            engine.add(DEMParticle(1, 2, 1.0f, 0.5f, 9.0f, make_float4(0.5f, 0.0f, 3.0f, 0.0f), zero));
            engine.add(DEMParticle(0, 1, 0.5f, 0.5f, 9.0f, make_float4(0.1f, 0.2f, 0.3f, 0.0f), zero));
            engine.prep();
            std::printf("{\n");
            print_pack(engine.getEntityHandler());
            print_f4("particle_position", engine.getEntityHandler().h_particle_position_ptr,
                     engine.getNumberOfSpheres(), true);
            std::printf("}\n");
            return 0;
        }
        if (mode == "pack-json" && argc >= 3) {
            DEMEngine engine(config, 0.0f);
            engine.loadJSONSetup(argv[2]);
            engine.prep();
            std::printf("{\n");
            print_pack(engine.getEntityHandler());
            print_f4("particle_position", engine.getEntityHandler().h_particle_position_ptr,
                     engine.getNumberOfSpheres(), true);
            std::printf("}\n");
            return 0;
        }
        if (mode == "run" && argc >= 4) {
            const int steps = std::atoi(argv[2]);
            const float dt = (float)std::atof(argv[3]);
            DEMEngine engine(config, 1.0f);
            if (argc >= 5) {
                engine.loadJSONSetup(argv[4]);
            } else {
                two_bodies(engine);
            }
            engine.prep();
            std::printf("{\n");
            print_pack(engine.getEntityHandler());
            engine.transfer();
            for (int i = 0; i < steps; ++i) engine.update(dt);
            engine.sync();
            EntityHandler& h = engine.getEntityHandler();
            const int nb = h.getNumberOfParticles();
            std::printf("\"steps\": %lld,\n", engine.stepsDone());
            print_f4("final_position", h.h_rigid_body_position_ptr, nb);
            print_f4("final_velocity", h.h_rigid_body_velocity_ptr, nb);
            print_f4("final_quaternion", h.h_rigid_body_quaternion_ptr, nb);
            print_f4("final_angular_momentum", h.h_rigid_body_angular_momentum_ptr, nb, true);
            std::printf("}\n");
            return 0;
        }
        if ((mode == "pack-clumps" && argc >= 3) || (mode == "run-clumps" && argc >= 5)) {
            // general clumps (DEMParticle from spheres, orientation and spin)
            const bool run = mode == "run-clumps";
            DEMEngine engine(config, 0.0f);
            read_clumps(argv[run ? 4 : 2], engine);
            engine.prep();
            std::printf("{\n");
            print_pack(engine.getEntityHandler());
            if (run) {
                const int steps = std::atoi(argv[2]);
                const float dt = (float)std::atof(argv[3]);
                engine.transfer();
                for (int i = 0; i < steps; ++i) engine.update(dt);
                engine.sync();
                EntityHandler& h = engine.getEntityHandler();
                const int nb = h.getNumberOfParticles();
                std::printf("\"steps\": %lld,\n", engine.stepsDone());
                print_f4("final_position", h.h_rigid_body_position_ptr, nb);
                print_f4("final_velocity", h.h_rigid_body_velocity_ptr, nb);
                print_f4("final_quaternion", h.h_rigid_body_quaternion_ptr, nb);
                print_f4("final_angular_momentum", h.h_rigid_body_angular_momentum_ptr, nb);
            }
            print_f4("particle_position", engine.getEntityHandler().h_particle_position_ptr,
                     engine.getNumberOfSpheres(), true);
            std::printf("}\n");
            return 0;
        }
        if (mode == "run-legacy" && argc >= 4) {
            // the two-body scene through the reference's other names
            const int steps = std::atoi(argv[2]);
            const float dt = (float)std::atof(argv[3]);
            DEMEngine engine(config, 1.0f);
            engine.addParticle(DEMParticle(0, 2, 0.04f, 0.02f, 0.5f, make_float4(0.0f, 0.0f, 0.0f, 0.0f),
                                           make_float4(1.0f, 0.0f, 0.0f, 0.0f)));
            engine.addParticle(DEMParticle(1, 2, 0.04f, 0.02f, 0.5f, make_float4(0.032f, 0.003f, 0.002f, 0.0f),
                                           make_float4(-1.0f, 0.0f, 0.0f, 0.0f)));
            engine.addParticle(DEMParticle(2, 1, 0.02f, 0.02f, 50.0f, make_float4(0.016f, -0.02f, 0.0f, 0.0f),
                                           make_float4(0.0f, 0.0f, 0.0f, 0.0f)));
            engine.prepArrays();
            engine.setUpGrid();
            engine.transferDataToDevice();
            for (int i = 0; i < steps; ++i) engine.step(dt);
            engine.sync();
            EntityHandler& h = engine.getEntityHandler();
            const int nb = h.getNumberOfParticles();
            engine.cleanBuffers();
            bool cleaned = false;
            try {
                engine.update(dt);
            } catch (const std::runtime_error&) {
                cleaned = true;  // no device state after cleanBuffers
            }
            std::printf("{\n\"steps\": %d, \"cleaned\": %s,\n", steps, cleaned ? "true" : "false");
            print_f4("final_position", h.h_rigid_body_position_ptr, nb);
            print_f4("final_velocity", h.h_rigid_body_velocity_ptr, nb);
            print_f4("final_quaternion", h.h_rigid_body_quaternion_ptr, nb);
            print_f4("final_angular_momentum", h.h_rigid_body_angular_momentum_ptr, nb, true);
            std::printf("}\n");
            return 0;
        }
        if (mode == "mesh-info" && argc >= 3) {
            const DEMMesh mesh(argv[2], make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(1.0f, 0.0f, 0.0f, 0.0f));
            std::printf("{\n\"n_triangles\": %d,\n", mesh.n_triangles);
            print_f4("vertices", mesh.vertices.data(), (int)mesh.vertices.size());
            print_f4("normals", mesh.normals.data(), (int)mesh.normals.size());
            print_i("indices", mesh.indicies.data(), (int)mesh.indicies.size());
            print_f("radii", mesh.radii.data(), (int)mesh.radii.size());
            print_f4("centres", mesh.positions.data(), (int)mesh.positions.size(), true);
            std::printf("}\n");
            return 0;
        }
        if (mode == "mesh-drop" && argc >= 5) {
            const int steps = std::atoi(argv[3]);
            const float dt = (float)std::atof(argv[4]);
            DEMEngine engine(config, 0.0f);
            // the box of the reference's test mesh: x in [-33, -28], y in [16, 21], z in [0, 1]
            engine.add(DEMMesh(argv[2], make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(1.0f, 0.0f, 0.0f, 0.0f)));
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j)
                    engine.add(DEMParticle((unsigned)(4 * i + j), 1, 0.5f, 0.5f, 1.0f,
                                           make_float4(-32.0f + 1.1f * i, 17.0f + 1.1f * j, 1.4f + 0.05f * (i + j), 0.0f),
                                           make_float4(0.1f * i, -0.1f * j, -1.0f, 0.0f)));
            std::vector<float4> cube;
            for (int k = 0; k < 8; ++k)
                cube.push_back(make_float4((k & 1) ? 0.3f : -0.3f, (k & 2) ? 0.3f : -0.3f, (k & 4) ? 0.3f : -0.3f, 0.0f));
            engine.add(DEMHull(100, cube, 2.0f, make_float4(-30.5f, 18.5f, 2.5f, 0.0f), make_float4(0.0f, 0.0f, -2.0f, 0.0f),
                               make_float4(0.9659258f, 0.2588190f, 0.0f, 0.0f)));
            engine.setContactLaw(2e4f, 0.3f, 0.5f, 2.0f / 7.0f);
            engine.prep();
            std::printf("{\n");
            print_pack(engine.getEntityHandler());
            engine.transfer();
            std::printf("\"hull_mode\": %s,\n", engine.hullMode() ? "true" : "false");
            for (int i = 0; i < steps; ++i) engine.update(dt);
            engine.sync();
            EntityHandler& h = engine.getEntityHandler();
            const int nb = h.getNumberOfParticles();
            std::printf("\"steps\": %lld,\n", engine.stepsDone());
            const DEMHull& hl = h.getHulls()[0];
            print_f4("hull_final", &hl.position, 1);
            print_f4("hull_final_velocity", &hl.velocity, 1);
            print_f4("hull_final_quaternion", &hl.quat, 1);
            print_f4("hull_final_angular_momentum", &hl.angular_momentum, 1);
            print_f4("final_position", h.h_rigid_body_position_ptr, nb);
            print_f4("final_velocity", h.h_rigid_body_velocity_ptr, nb);
            print_f4("final_quaternion", h.h_rigid_body_quaternion_ptr, nb);
            print_f4("final_angular_momentum", h.h_rigid_body_angular_momentum_ptr, nb, true);
            std::printf("}\n");
            return 0;
        }
        if (mode == "checkpoint" && argc >= 5) {
            const int steps = std::atoi(argv[2]);
            const float dt = (float)std::atof(argv[3]);
            auto build = [&](DEMEngine& e) {
                two_bodies(e);
                e.prep();
                e.transfer();
            };
            DEMEngine a(config, 1.0f), b(config, 1.0f);
            build(a);
            for (int i = 0; i < steps; ++i) a.update(dt);
            a.saveState(argv[4]);
            for (int i = 0; i < steps; ++i) a.update(dt);
            a.sync();
            build(b);
            b.loadState(argv[4]);
            const long long resumed_at = b.stepsDone();
            for (int i = 0; i < steps; ++i) b.update(dt);
            b.sync();
            if (argc >= 6) a.writeVTK(argv[5]);
            std::printf("{\n\"resumed_at\": %lld, \"steps_a\": %lld, \"steps_b\": %lld,\n", resumed_at,
                        a.stepsDone(), b.stepsDone());
            EntityHandler &ha = a.getEntityHandler(), &hb = b.getEntityHandler();
            const int nb = ha.getNumberOfParticles(), ns = ha.getNumberOfSpheres();
            print_f4("a_position", ha.h_rigid_body_position_ptr, nb);
            print_f4("a_quaternion", ha.h_rigid_body_quaternion_ptr, nb);
            print_f4("a_angular_momentum", ha.h_rigid_body_angular_momentum_ptr, nb);
            print_f4("a_velocity", ha.h_rigid_body_velocity_ptr, nb);
            print_f4("a_sphere_position", ha.h_particle_position_ptr, ns);
            print_f("a_sphere_size", ha.h_particle_size_ptr, ns);
            print_i("a_particle_to_rigid", ha.h_particle_to_rigid_idx_ptr, ns);
            print_f4("b_position", hb.h_rigid_body_position_ptr, nb);
            print_f4("b_quaternion", hb.h_rigid_body_quaternion_ptr, nb);
            print_f4("b_angular_momentum", hb.h_rigid_body_angular_momentum_ptr, nb);
            print_f4("b_velocity", hb.h_rigid_body_velocity_ptr, nb, true);
            std::printf("}\n");
            return 0;
        }
        if (mode == "edit-clumps" && argc >= 4) {
            const int steps = std::atoi(argv[2]);
            const float dt = (float)std::atof(argv[3]);
            DEMEngine engine(config, 1.0f);
            two_bodies(engine);
            engine.add(DEMParticle(3, 2, 0.04f, 0.02f, 0.5f, make_float4(0.0f, 0.05f, 0.0f, 0.0f),
                                   make_float4(0.0f, -0.5f, 0.0f, 0.0f)));
            engine.prep();
            std::printf("{\n");
            print_pack(engine.getEntityHandler());
            engine.transfer();
            for (int i = 0; i < steps; ++i) engine.update(dt);
            engine.getEntityHandler().removeEntityRuntime(0, 2);
            for (int i = 0; i < steps; ++i) engine.update(dt);
            const DEMParticle extra(4, 2, 0.04f, 0.02f, 0.5f, make_float4(0.03f, 0.045f, 0.002f, 0.0f),
                                    make_float4(0.0f, 0.0f, 0.0f, 0.0f));
            engine.getEntityHandler().addEntityRuntime(extra);
            for (int i = 0; i < steps; ++i) engine.update(dt);
            engine.sync();
            {
                DEMEngine one(config, 1.0f);
                one.add(extra);
                one.prep();
                std::printf("\"added\": {\n");
                print_pack(one.getEntityHandler());
                std::printf("\"end\": 0},\n");
            }
            EntityHandler& h = engine.getEntityHandler();
            const int nb = h.getNumberOfParticles();
            std::printf("\"steps\": %lld, \"bodies_after\": %d,\n", engine.stepsDone(), nb);
            print_f4("final_position", h.h_rigid_body_position_ptr, nb);
            print_f4("final_velocity", h.h_rigid_body_velocity_ptr, nb);
            print_f4("final_quaternion", h.h_rigid_body_quaternion_ptr, nb);
            print_f4("final_angular_momentum", h.h_rigid_body_angular_momentum_ptr, nb, true);
            std::printf("}\n");
            return 0;
        }
        if ((mode == "run-spheres" && argc >= 10) || (mode == "edit-spheres" && argc >= 11)) {
            const int steps = std::atoi(argv[2]);
            const float dt = (float)std::atof(argv[3]);
            const int n = std::atoi(argv[4]);
            const float domain = (float)std::atof(argv[5]), size = (float)std::atof(argv[6]);
            const float lo = (float)std::atof(argv[7]), hi = (float)std::atof(argv[8]);
            unsigned long long state = std::strtoull(argv[9], nullptr, 10);
            const int n0 = n;
            (void)n0;
            auto uniform = [&state]() {  // SplitMix64, 24-bit uniform in [0, 1)
                state += 0x9E3779B97F4A7C15ull;
                unsigned long long z = state;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                return (float)(z >> 40) * (1.0f / 16777216.0f);
            };
            DEMEngine engine(config, domain);
            for (int i = 0; i < n; ++i) {
                const float4 pos = make_float4(lo + (hi - lo) * uniform(), lo + (hi - lo) * uniform(),
                                               lo + (hi - lo) * uniform(), 0.0f);
                const float4 vel = make_float4(uniform() - 0.5f, uniform() - 0.5f, uniform() - 0.5f, 0.0f);
                const float mass = (i % 97 == 0) ? 100.0f : 1.0f + 8.0f * uniform();
                engine.add(DEMParticle((unsigned)i, 1, size, size, mass, pos, vel));
            }
            engine.prep();
            std::printf("{\n");
            print_pack(engine.getEntityHandler());
            engine.transfer();
            const DEMEngine::SphereGrid& g = engine.sphereGrid();
            std::printf("\"sphere_mode\": %s,\n", engine.sphereMode() ? "true" : "false");
            std::printf("\"grid\": {\"lo\": [%.9g, %.9g, %.9g], \"cell\": %.9g, \"n\": [%d, %d, %d], "
                        "\"walls\": %s, \"wall_lo\": [%.9g, %.9g, %.9g], \"wall_hi\": [%.9g, %.9g, %.9g]},\n",
                        g.lo[0], g.lo[1], g.lo[2], g.cell, g.n[0], g.n[1], g.n[2], g.walls ? "true" : "false",
                        g.wall_lo[0], g.wall_lo[1], g.wall_lo[2], g.wall_hi[0], g.wall_hi[1], g.wall_hi[2]);
            for (int i = 0; i < steps; ++i) engine.update(dt);
            if (mode == "edit-spheres") {
                const int every = std::atoi(argv[10]);
                std::printf("\"removed\": [");
                bool first = true;
                for (int i = n - 1; i >= 0; --i) {
                    if (i % every) continue;
                    engine.getEntityHandler().removeEntityRuntime(0, i);
                    std::printf("%s%d", first ? "" : ", ", i);
                    first = false;
                }
                std::printf("],\n");
                for (int i = 0; i < steps; ++i) engine.update(dt);
                std::printf("\"added\": [");
                for (int k = 0; k < 3; ++k) {
                    const float4 pos = make_float4(lo + (hi - lo) * uniform(), lo + (hi - lo) * uniform(),
                                                   lo + (hi - lo) * uniform(), 0.0f);
                    const float4 vel = make_float4(uniform() - 0.5f, uniform() - 0.5f, uniform() - 0.5f, 0.0f);
                    const float mass = 1.0f + 8.0f * uniform();
                    engine.getEntityHandler().addEntityRuntime(DEMParticle((unsigned)(n + k), 1, size, size, mass,
                                                                           pos, vel));
                    std::printf("%s[%.9g, %.9g, %.9g, %.9g, %.9g, %.9g, %.9g]", k ? ", " : "", pos.x, pos.y, pos.z,
                                vel.x, vel.y, vel.z, mass);
                }
                std::printf("],\n");
                for (int i = 0; i < steps; ++i) engine.update(dt);
            }
            engine.sync();
            EntityHandler& h = engine.getEntityHandler();
            const int nb = h.getNumberOfParticles();
            std::printf("\"steps\": %lld,\n", engine.stepsDone());
            print_f4("final_position", h.h_rigid_body_position_ptr, nb);
            print_f4("final_velocity", h.h_rigid_body_velocity_ptr, nb);
            print_f4("final_angular_momentum", h.h_rigid_body_angular_momentum_ptr, nb, true);
            std::printf("}\n");
            return 0;
        }
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
