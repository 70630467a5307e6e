// Drives the C++ mirror (include/ipic3d_gpu/ipic3d.hpp) the way the reference's simulation
// tests drive theirs: build a Universe, place particles, set fields, simulateSteps.
//
//   ipic3d_mirror_test <input.bin> <output.bin>   run the case of input.bin on the GPU; also
//                                                 writes <output.bin>.counts.out and
//                                                 <output.bin>.log in the reference's formats
//   ipic3d_mirror_test --expect-no-device          check the error path without a GPU
//   ipic3d_mirror_test --driver-case <prefix>       the main driver's set-up of inputs/test.inp
//                                                 (src/ipic3d.cxx:41-99): createUniverseFrom-
//                                                 Distribution, 30 steps of the loop as shipped,
//                                                 <prefix>.out in the reference's format
//   ipic3d_mirror_test --params-case <n> <ppc>      createUniverseFromParams of a lattice
//                                                 set-up (qom -25, uth 0.045, u0 0.02,
//                                                 B0 = (0, 0, 1e-4), cells of 10 / n): prints
//                                                 the conserved quantities of cycle 0
//
// input.bin (little-endian): int64 n, int64 steps, int64 nparticles, double dt, then per
// particle 8 doubles (position, velocity, q, qom), then n^3 x 3 doubles E and n^3 x 3 Bc.
// output.bin: int64 particle count, n^3 int64 cell counts, double kinetic energy, double
// firstStep, double remainingSteps, int64 1 if the operator-policy guard threw.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <set>

#include "ipic3d_gpu/io.hpp"

using namespace ipic3d_gpu;

int main(int argc, char** argv) {
    if (argc == 2 && std::string(argv[1]) == "--expect-no-device") {
        UniverseProperties p;
        p.size = {4, 4, 4};
        try {
            Universe u(p, {{-1.0, -1.0}});
            std::printf("{\"created\": true}\n");
            return 1;
        } catch (const Error& e) {
            std::printf("{\"error\": %d}\n", e.code());
            return e.code() == IPIC_E_HIP ? 0 : 1;
        }
    }
    if (argc == 3 && std::string(argv[1]) == "--driver-case") {
        // ipic3d.cxx:47-85 with test.inp: 8^3 cells, L = 10, 8 x 8 x 8 particles per cell
        const double R = 6378137.0, c = 299792458.0;
        UniverseProperties p;
        p.size = {8, 8, 8};
        p.cellWidth = {10.0 / 8 * R, 10.0 / 8 * R, 10.0 / 8 * R};
        p.dt = 0.01;
        p.speedOfLight = c;
        p.planetRadius = R;
        p.objectCenter = {0.0, 0.0, 0.0};
        Vector3 size{}, low{}, hig{};
        for (int k = 0; k < 3; ++k) {
            size[k] = p.cellWidth[k] * (double)p.size[k];
            p.origin[k] = p.objectCenter[k] - size[k] / 2.0;
            low[k] = p.origin[k] + 0.125 * size[k];
            hig[k] = low[k] + 0.75 * size[k];
        }
        p.externalMagneticField = {0.0, 0.0, 3.07e-5};
        p.shippedDipoleMover = true;
        p.fieldSubsteps = 0;
        p.flags = IPIC_FLAG_REFLECT;
        const double K = 1e7 * kElementaryCharge, m = kProtonMass;
        const double v_mod = c / std::sqrt(1.0 + (m * c * c) / K);
        const std::uint64_t numParticles = 8 * 8 * 8 * 8 * 8 * 8;
        try {
            Universe u = createUniverseFromDistribution(
                p, numParticles, distribution::uniform_pos_normal_speed(low, hig, {0, 0, 0}, {v_mod, v_mod, v_mod}));
            const DurationMeasurement d = simulateSteps(30, u);
            outputNumberOfParticlesPerCell(u, std::string(argv[2]) + ".out");
            std::printf("{\"particles\": %lld, \"first\": %g, \"remaining\": %g}\n",
                        (long long)u.countParticlesInDomain(), d.firstStep, d.remainingSteps);
        } catch (const Error& e) {
            std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
            return 1;
        }
        return 0;
    }
    if (argc == 4 && std::string(argv[1]) == "--params-case") {
        const int n = std::atoi(argv[2]);
        const unsigned ppc = (unsigned)std::atoi(argv[3]);
        UniverseProperties p;
        p.size = {n, n, n};
        p.cellWidth = {10.0 / n, 10.0 / n, 10.0 / n};
        p.dt = 0.15;
        p.fieldSubsteps = 0;
        InitProperties init;
        init.particlesPerCell = {{ppc, ppc, ppc}};
        init.driftVelocity = {{0.02, 0.0, 0.0}};
        init.magneticField = {0.0, 0.0, 1e-4};
        try {
            Universe u = createUniverseFromParams(p, init, -25.0, {0.045, 0.045, 0.045});
            const std::pair<double, double> en = u.initFields(init);
            const ipic_conservation c = u.conservation();
            std::printf("{\"particles\": %lld, \"momentum\": %.17g, \"kinetic_energy\": %.17g, "
                        "\"E_energy\": %.17g, \"B_energy\": %.17g}\n",
                        (long long)u.countParticlesInDomain(), c.momentum, c.kinetic_energy, en.first, en.second);
        } catch (const Error& e) {
            std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
            return 1;
        }
        return 0;
    }
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s input.bin output.bin | --expect-no-device\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    int64_t n = 0, steps = 0, np = 0;
    double dt = 0;
    in.read((char*)&n, 8).read((char*)&steps, 8).read((char*)&np, 8).read((char*)&dt, 8);
    std::vector<Particle> ps((size_t)np);
    std::set<std::pair<double, double>> kinds;
    for (auto& p : ps) {
        double b[8];
        in.read((char*)b, sizeof(b));
        p.position = {b[0], b[1], b[2]};
        p.velocity = {b[3], b[4], b[5]};
        p.q = b[6];
        p.qom = b[7];
        kinds.insert({p.q, p.qom});
    }
    const size_t nc = (size_t)(n * n * n);
    std::vector<Vector3> E(nc), Bc(nc);
    in.read((char*)E[0].data(), nc * 24);
    in.read((char*)Bc[0].data(), nc * 24);
    if (!in) {
        std::fprintf(stderr, "short input\n");
        return 2;
    }
    try {
        UniverseProperties props;
        props.size = {n, n, n};
        props.dt = dt;
        Universe u(props, std::vector<std::pair<double, double>>(kinds.begin(), kinds.end()));
        u.setParticles(ps);
        u.setFields(E, Bc);
        int64_t guard = 0;
        try {
            simulateSteps<detail::gpu_particle_to_field_projector, detail::gpu_field_operator_solver>(1, u);
        } catch (const Error& e) {
            guard = e.code() == IPIC_E_STATE ? 1 : 0;
        }
        std::ofstream log(std::string(argv[2]) + ".log");
        writeOutputHeader(log);
        writeOutputData(0, (int)steps, u, log, std::string(argv[2]) + ".density", 1, 0);
        const DurationMeasurement d = simulateSteps(steps, u);
        writeOutputData((int)steps - 1, (int)steps, u, log, std::string(argv[2]) + ".density", 1, 0);
        outputNumberOfParticlesPerCell(u, std::string(argv[2]) + ".counts.out");
        const int64_t count = u.countParticlesInDomain();
        const std::vector<int64_t> cells = u.getParticleCountsPerCell();
        const double ke = u.conservation().kinetic_energy;
        std::ofstream out(argv[2], std::ios::binary);
        out.write((const char*)&count, 8);
        out.write((const char*)cells.data(), (std::streamsize)(cells.size() * 8));
        out.write((const char*)&ke, 8);
        out.write((const char*)&d.firstStep, 8);
        out.write((const char*)&d.remainingSteps, 8);
        out.write((const char*)&guard, 8);
        std::printf("{\"particles\": %lld, \"kinetic_energy\": %.17g, \"guard\": %lld}\n", (long long)count, ke,
                    (long long)guard);
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
