// io.hpp -- the reference's output formats for an ipic3d_gpu::Universe (C++ mirror).
//
//   outputNumberOfParticlesPerCell   cell.h:1360-1395
//   outputParticlePositions          cell.h:1397-1429
//   writeOutputHeader / writeOutputData (conserved quantities, density frames)  simulator.h:67-115
//   getFieldEnergy                   field.h:443-459
//   outputFieldGrids                 field.h:464-530
//
// Doubles print with the stream's defaults (6 significant digits), vectors and sizes as
// "[x,y,z]".  The grids here are N-periodic, so field dumps hold N^3 entries.
#ifndef IPIC3D_GPU_IO_HPP
#define IPIC3D_GPU_IO_HPP

#include <cmath>
#include <cstdio>
#include <fstream>
#include <ostream>
#include <string>
#include <vector>

#include "ipic3d.hpp"

namespace ipic3d_gpu {

inline std::ostream& operator<<(std::ostream& out, const Vector3& v) {
    return out << "[" << v[0] << "," << v[1] << "," << v[2] << "]";
}

inline void outputNumberOfParticlesPerCell(const Universe& u, const std::string& filename) {
    const auto& n = u.getProperties().size;
    const std::vector<std::int64_t> counts = u.getParticleCountsPerCell();
    std::ofstream out(filename);
    out << "[" << n[0] << "," << n[1] << "," << n[2] << "]\n";
    std::uint64_t total = 0;
    std::size_t i = 0;
    for (std::int64_t x = 0; x < n[0]; ++x)
        for (std::int64_t y = 0; y < n[1]; ++y)
            for (std::int64_t z = 0; z < n[2]; ++z) {
                out << x << "," << y << "," << z << ":" << counts[i] << "\n";
                total += (std::uint64_t)counts[i++];
            }
    out << "\nTotal: " << total << "\n";
}

inline void outputParticlePositions(const Universe& u, const std::string& filename) {
    std::ofstream out(filename);
    for (const Particle& p : u.getParticles())
        out << p.position[0] << " " << p.position[1] << " " << p.position[2] << "\n";
}

// getFieldEnergy (field.h:443-459): sum over nodes of 0.5 vol |F|^2 / (4 pi)
inline double getFieldEnergy(const std::vector<Vector3>& F, const UniverseProperties& props) {
    const double vol = 0.5 * props.cellWidth[0] * props.cellWidth[1] * props.cellWidth[2];
    const double fourPI = 16.0 * std::atan(1.0);
    double r = 0.0;
    for (const Vector3& v : F) r += vol * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / fourPI;
    return r;
}

inline void writeOutputHeader(std::ostream& out) { out << "Cycle \t Total Moment \t E energy \t B energy \t Total KE \n"; }

// writeOutputData (simulator.h:69-114): conserved quantities every fieldOutputCycle cycles
// and at the last one; a density frame file densityPrefix%06d at cycle 0 and every
// particleOutputCycle cycles
inline void writeOutputData(int cycle, int numSteps, const Universe& u, std::ostream& out, const std::string& densityPrefix,
                            int fieldOutputCycle, int particleOutputCycle) {
    const UniverseProperties& props = u.getProperties();
    if (fieldOutputCycle > 0 && ((cycle % fieldOutputCycle == 0) || (cycle + 1 == numSteps))) {
        std::vector<Vector3> E, Bn;
        u.getFields(&E, nullptr, &Bn, nullptr);
        const ipic_conservation c = u.conservation();
        out << cycle << "\t" << c.momentum << "\t" << getFieldEnergy(E, props) << "\t" << getFieldEnergy(Bn, props)
            << "\t" << c.kinetic_energy << "\n";
    }
    if (particleOutputCycle > 0 && ((cycle == 0) || ((cycle + 1) % particleOutputCycle == 0))) {
        const int t = (cycle + 1) / particleOutputCycle;
        char name[4096];
        std::snprintf(name, sizeof(name), "%s%06d", densityPrefix.c_str(), t);
        std::ofstream f(name);
        f << "t,x,y,z,density\n";
        const std::vector<std::int64_t> counts = u.getParticleCountsPerCell();
        std::size_t i = 0;
        for (std::int64_t x = 0; x < props.size[0]; ++x)
            for (std::int64_t y = 0; y < props.size[1]; ++y)
                for (std::int64_t z = 0; z < props.size[2]; ++z)
                    f << t << "," << x * props.cellWidth[0] << "," << y * props.cellWidth[1] << ","
                      << z * props.cellWidth[2] << "," << counts[i++] << "\n";
    }
}

// outputFieldGrids (field.h:464-530): node grid as i,j,k:E|B|Bext, then the centre grid
inline void outputFieldGrids(const Universe& u, const std::string& filename) {
    const auto& n = u.getProperties().size;
    std::vector<Vector3> E, Bc, Bn;
    u.getFields(&E, &Bc, &Bn, nullptr);
    std::ofstream out(filename);
    const Vector3 zero{0.0, 0.0, 0.0};
    for (int grid = 0; grid < 2; ++grid) {
        out << "[" << n[0] << "," << n[1] << "," << n[2] << "]\n";
        std::size_t i = 0;
        for (std::int64_t x = 0; x < n[0]; ++x)
            for (std::int64_t y = 0; y < n[1]; ++y)
                for (std::int64_t z = 0; z < n[2]; ++z, ++i) {
                    out << x << "," << y << "," << z << ":";
                    if (grid == 0) out << E[i] << "|" << Bn[i] << "|" << zero << "\n";
                    else out << Bc[i] << "\n";
                }
        out << "\n";
    }
}

}  // namespace ipic3d_gpu

#endif
