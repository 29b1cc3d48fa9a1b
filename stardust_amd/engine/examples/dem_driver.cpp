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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

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
    print_f4("particle_relative_position", h.h_particle_relative_position_ptr, ns);
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
                // two 2x2x2 lattice bodies approaching along x, slightly offset
                engine.add(DEMParticle(0, 2, 0.04f, 0.02f, 0.5f, make_float4(0.0f, 0.0f, 0.0f, 0.0f),
                                       make_float4(1.0f, 0.0f, 0.0f, 0.0f)));
                engine.add(DEMParticle(1, 2, 0.04f, 0.02f, 0.5f, make_float4(0.032f, 0.003f, 0.002f, 0.0f),
                                       make_float4(-1.0f, 0.0f, 0.0f, 0.0f)));
                engine.add(DEMParticle(2, 1, 0.02f, 0.02f, 50.0f, make_float4(0.016f, -0.02f, 0.0f, 0.0f),
                                       make_float4(0.0f, 0.0f, 0.0f, 0.0f)));  // mass >= 10: static
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
        if (mode == "run-spheres" && argc >= 10) {
            const int steps = std::atoi(argv[2]);
            const float dt = (float)std::atof(argv[3]);
            const int n = std::atoi(argv[4]);
            const float domain = (float)std::atof(argv[5]), size = (float)std::atof(argv[6]);
            const float lo = (float)std::atof(argv[7]), hi = (float)std::atof(argv[8]);
            unsigned long long state = std::strtoull(argv[9], nullptr, 10);
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
