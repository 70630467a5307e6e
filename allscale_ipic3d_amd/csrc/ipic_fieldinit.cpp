// ipic_fieldinit.cpp -- the reference's initial field of the dipole use case, on the host
// (one-time set-up; the result is uploaded as the field state).
//
//   getDipoleFieldAt (field.h:43-78): E = -driftVelocity[0] x magneticField; outside the planet
//     (r^2 > a^2, r = location - objectCenter) Bext = fac (3 x z, 3 y z, 2 z^2 - x^2 - y^2) with
//     fac = -externalMagneticField.z a^3 / (r^2)^2.5, inside Bext = 0; B = magneticField + Bext.
//   initFields (field.h:81-118): that field at every node location origin + i * cellWidth
//     (getLocationForFields, universe_properties.h:132-137), i in [0, N]^3.
//   initBcFields (field.h:120-138): Bc of cell i = 1/8 of node B over the 8 nodes i + {0,1}^3
//     (interpN2C, field.h:252-265).
//   getFieldEnergy (field.h:443-459) over those (N+1)^3 nodes of E and of B + Bext, as
//     writeOutputData logs them (simulator.h:73-79): sum (w_x w_y w_z / 2) |F|^2 / (4 pi).
//
// On the library's N-periodic grid node i < N and centre i < N keep these values; the centres
// of the last layer average the node locations at N, as the reference's do.
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "ipic_gpu.h"

namespace {

thread_local std::string g_field_error;

int field_fail(int rc, const std::string& msg) {
    g_field_error = msg;
    return rc;
}

struct NodeField {
    double E[3], B[3], Bext[3];
};

NodeField dipole_at(const double loc[3], const ipic_field_init& f) {
    NodeField r{};
    const double* u = f.drift;
    const double* b = f.B0;
    r.E[0] = -(u[1] * b[2] - u[2] * b[1]);
    r.E[1] = -(u[2] * b[0] - u[0] * b[2]);
    r.E[2] = -(u[0] * b[1] - u[1] * b[0]);
    const double dx = loc[0] - f.object_center[0], dy = loc[1] - f.object_center[1],
                 dz = loc[2] - f.object_center[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double a = f.planet_radius;
    if (r2 > a * a) {
        const double fac = -f.external_Bz * std::pow(a, 3) / std::pow(r2, 2.5);
        r.Bext[0] = 3.0 * dx * dz * fac;
        r.Bext[1] = 3.0 * dy * dz * fac;
        r.Bext[2] = (2.0 * dz * dz - dx * dx - dy * dy) * fac;
    }
    for (int k = 0; k < 3; ++k) r.B[k] = b[k] + r.Bext[k];
    return r;
}

}  // namespace

extern "C" {

const char* ipic_field_init_last_error(void) { return g_field_error.c_str(); }

int ipic_field_init_host(const ipic_desc* desc, const ipic_field_init* init, double* E, double* Bn, double* Bc,
                         double energies[2]) {
    try {
        if (!desc || !init) return field_fail(IPIC_E_ARG, "null argument");
        int64_t n[3];
        for (int k = 0; k < 3; ++k) {
            n[k] = desc->n[k];
            if (n[k] <= 0) return field_fail(IPIC_E_ARG, "grid size must be positive");
            if (!(desc->cell_width[k] > 0)) return field_fail(IPIC_E_ARG, "cell width must be positive");
        }
        if ((n[0] + 1) * (n[1] + 1) * (n[2] + 1) > (int64_t(1) << 31))
            return field_fail(IPIC_E_RANGE, "grid too large");
        if (!(init->planet_radius >= 0)) return field_fail(IPIC_E_ARG, "planet radius must be >= 0");
        const int64_t m[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
        // node field on the (N+1)^3 node set of initFields
        std::vector<NodeField> nodes((size_t)(m[0] * m[1] * m[2]));
        const double four_pi = 16.0 * std::atan(1.0);
        const double hv = 0.5 * desc->cell_width[0] * desc->cell_width[1] * desc->cell_width[2];
        double eE = 0.0, eB = 0.0;
        for (int64_t x = 0; x < m[0]; ++x)
            for (int64_t y = 0; y < m[1]; ++y)
                for (int64_t z = 0; z < m[2]; ++z) {
                    const double loc[3] = {desc->origin[0] + (double)x * desc->cell_width[0],
                                           desc->origin[1] + (double)y * desc->cell_width[1],
                                           desc->origin[2] + (double)z * desc->cell_width[2]};
                    const NodeField f = dipole_at(loc, *init);
                    nodes[(size_t)((x * m[1] + y) * m[2] + z)] = f;
                    double se = 0.0, sb = 0.0;
                    for (int k = 0; k < 3; ++k) {
                        se += f.E[k] * f.E[k];
                        const double bb = f.B[k] + f.Bext[k];
                        sb += bb * bb;
                    }
                    eE += hv * se / four_pi;
                    eB += hv * sb / four_pi;
                }
        auto node = [&](int64_t x, int64_t y, int64_t z) -> const NodeField& {
            return nodes[(size_t)((x * m[1] + y) * m[2] + z)];
        };
        for (int64_t x = 0; x < n[0]; ++x)
            for (int64_t y = 0; y < n[1]; ++y)
                for (int64_t z = 0; z < n[2]; ++z) {
                    const size_t i = (size_t)((x * n[1] + y) * n[2] + z);
                    const NodeField& f = node(x, y, z);
                    for (int k = 0; k < 3; ++k) {
                        if (E) E[3 * i + k] = f.E[k];
                        if (Bn) Bn[3 * i + k] = f.B[k];
                    }
                    if (Bc) {
                        double acc[3] = {0.0, 0.0, 0.0};
                        for (int a = 0; a < 2; ++a)
                            for (int b = 0; b < 2; ++b)
                                for (int c = 0; c < 2; ++c) {
                                    const NodeField& g = node(x + a, y + b, z + c);
                                    for (int k = 0; k < 3; ++k) acc[k] += g.B[k];
                                }
                        for (int k = 0; k < 3; ++k) Bc[3 * i + k] = 0.125 * acc[k];
                    }
                }
        if (energies) {
            energies[0] = eE;
            energies[1] = eB;
        }
        return 0;
    } catch (const std::exception&) {
        return field_fail(IPIC_E_RANGE, "out of host memory");
    }
}

}  // extern "C"
