// ipic3d.hpp -- C++ host mirror of the reference's step API over the C ABI of libipic_gpu.so.
//
// Mirrors, with the same names and argument meaning:
//   Particle                    particle.h:9-19 (absolute position, velocity, q, qom)
//   UniverseProperties          universe_properties.h:52-142 (the members the step uses)
//   DurationMeasurement         simulator.h:37-40 (seconds, millisecond resolution)
//   simulateSteps / simulateStep  simulator.h:42-57, 127-242, with GPU policy tags in place of
//                               the per-cell projector / field solver / mover of :244-276
//   Universe                    universe.h:12-51 (state resident on the GPU)
//   distribution::uniform_pos_normal_speed, createUniverseFromDistribution, Universe::initCells
//                               cell.h:405-440, universe.h:85-103, cell.h:512-654 (the
//                               reference's own random streams)
//   InitProperties, createUniverseFromParams, Universe::initCells(InitProperties, ...),
//   Universe::initFields        init_properties.h:14-45, universe.h:53-75, cell.h:699-776,
//                               field.h:43-138 (the parameter-file set-up)
//
// The reference aborts on a failed assert_*; here every failed C call throws
// ipic3d_gpu::Error carrying the IPIC_E_* code and the library's message.  Header-only,
// C++17; link with -lipic_gpu.
#ifndef IPIC3D_GPU_IPIC3D_HPP
#define IPIC3D_GPU_IPIC3D_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../ipic_gpu.h"

namespace ipic3d_gpu {

using Vector3 = std::array<double, 3>;
using Coordinate3 = std::array<std::int64_t, 3>;

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

// particle.h:9-19
struct Particle {
    Vector3 position{};   // absolute
    Vector3 velocity{};
    double q = 0.0;
    double qom = 0.0;
};

// universe_properties.h:52-142: the members the step reads, plus the GPU's layout knobs
struct UniverseProperties {
    Coordinate3 size{1, 1, 1};
    Vector3 cellWidth{1.0, 1.0, 1.0};
    double dt = 1.0;
    double speedOfLight = 1.0;
    double planetRadius = 0.0;       // > 0: particles within it of objectCenter are removed
    Vector3 objectCenter{0.0, 0.0, 0.0};
    Vector3 origin{0.0, 0.0, 0.0};
    Vector3 externalMagneticField{0.0, 0.0, 0.0};
    bool shippedDipoleMover = false; // the mover as shipped (cell.h:919-945): analytic dipole of
                                     // externalMagneticField.z and planetRadius at the origin, E = 0
    int fieldSubsteps = 1;           // leapfrog substeps per particle step (J held fixed); 0: static
    Coordinate3 tile{0, 0, 0};       // cells per GPU tile, 0: automatic
    std::uint32_t flags = 0;         // IPIC_FLAG_*
};

// the constants initCells gives every particle (cell.h:514-515, 644-645)
constexpr double kElementaryCharge = 1.602176565e-19;
constexpr double kProtonMass = 1.672621777e-27;

namespace distribution {
// uniform_pos_normal_speed (cell.h:405-440): positions uniform in [minPos, maxPos), velocities
// center + stddev * the reference's ziggurat normal, per component
struct uniform_pos_normal_speed {
    Vector3 minPos{}, maxPos{}, center{}, stddev{};
    std::uint32_t seed = 0;
    uniform_pos_normal_speed(const Vector3& minPos_, const Vector3& maxPos_, const Vector3& center_,
                             const Vector3& stddev_, std::uint32_t seed_ = 0)
        : minPos(minPos_), maxPos(maxPos_), center(center_), stddev(stddev_), seed(seed_) {}
};
}  // namespace distribution

// init_properties.h:14-45: the properties that only initialisation reads, one entry per species
// in particlesPerCell / driftVelocity (the set-up uses the first)
struct InitProperties {
    std::uint64_t numSteps = 1;
    std::vector<std::array<unsigned, 3>> particlesPerCell;
    std::vector<Vector3> driftVelocity;
    Vector3 magneticField{0.0, 0.0, 0.0};
    double rhoInit = 1.0;
};

// simulator.h:37-40
struct DurationMeasurement {
    double firstStep;
    double remainingSteps;
};

namespace detail {
// policy tags of simulateSteps: the GPU versions of the reference's per-cell hooks
struct gpu_particle_to_field_projector {};  // projectToDensityField, cell.h:787-824
struct gpu_leapfrog_field_solver {};        // solveFieldLeapfrog, field.h:342-402
struct gpu_field_operator_solver {};        // the learned field step (with leapfrog fallback)
struct gpu_particle_mover {};               // moveParticles + exportParticles, cell.h:896-1159
}  // namespace detail

class Universe;

// simulateSteps (simulator.h:127-242): deposit, field solve, move and rebin on the GPU.  The
// policies select the GPU implementation at compile time; the leapfrog policy refuses a
// universe that carries a field operator and the operator policy needs one.
template <class ParticleToFieldProjector = detail::gpu_particle_to_field_projector,
          class FieldSolver = detail::gpu_leapfrog_field_solver,
          class ParticleMover = detail::gpu_particle_mover>
DurationMeasurement simulateSteps(std::uint64_t numSteps, Universe& universe);

inline void check(ipic_ctx* ctx, int rc) {
    if (rc != IPIC_OK) throw Error(rc, std::string("libipic_gpu: ") + ipic_gpu_last_error(ctx));
}

// universe.h:12-51 -- the cells, fields and properties, resident on one GPU.  Particles are
// grouped into species by their (q, qom), as the GPU keeps both per species.
class Universe {
  public:
    Universe(const UniverseProperties& props, const std::vector<std::pair<double, double>>& species)
        : props_(props), species_(species) {
        if (species.empty() || species.size() > IPIC_MAX_SPECIES)
            throw Error(IPIC_E_ARG, "between 1 and IPIC_MAX_SPECIES species (q, qom) are needed");
        ipic_desc d{};
        for (int k = 0; k < 3; ++k) {
            d.n[k] = (int32_t)props.size[k];
            d.cell_width[k] = props.cellWidth[k];
            d.origin[k] = props.origin[k];
            d.sink_center[k] = props.objectCenter[k];
            d.tile[k] = (int32_t)props.tile[k];
        }
        d.nspecies = (int32_t)species.size();
        d.dt = props.dt;
        d.c = props.speedOfLight;
        for (size_t s = 0; s < species.size(); ++s) d.species[s] = ipic_species_desc{species[s].first, species[s].second};
        d.field_substeps = props.fieldSubsteps;
        d.sink = props.planetRadius > 0.0 ? 1 : 0;
        d.sink_radius = props.planetRadius;
        d.flags = props.flags;
        if (props.shippedDipoleMover) {
            d.flags |= IPIC_FLAG_DIPOLE_MOVER;
            d.dipole_moment = -props.externalMagneticField[2] * props.planetRadius * props.planetRadius * props.planetRadius;
        }
        check(nullptr, ipic_gpu_create(&d, &ctx_));
    }
    ~Universe() {
        if (ctx_) ipic_gpu_destroy(ctx_);
    }
    Universe(const Universe&) = delete;
    Universe& operator=(const Universe&) = delete;
    Universe(Universe&& o) noexcept : props_(o.props_), species_(std::move(o.species_)), ctx_(o.ctx_) { o.ctx_ = nullptr; }

    const UniverseProperties& getProperties() const { return props_; }
    ipic_ctx* handle() const { return ctx_; }
    std::int64_t numCells() const { return props_.size[0] * props_.size[1] * props_.size[2]; }

    int speciesOf(const Particle& p) const {
        for (size_t s = 0; s < species_.size(); ++s)
            if (species_[s].first == p.q && species_[s].second == p.qom) return (int)s;
        return -1;
    }

    // replaces every particle; each is binned into the cell floor((x - origin) / width)
    void setParticles(const std::vector<Particle>& ps) {
        std::vector<std::vector<double>> pos(species_.size()), vel(species_.size());
        for (const Particle& p : ps) {
            const int s = speciesOf(p);
            if (s < 0) throw Error(IPIC_E_ARG, "particle (q, qom) matches no species of the universe");
            for (int k = 0; k < 3; ++k) {
                pos[s].push_back(p.position[k]);
                vel[s].push_back(p.velocity[k]);
            }
        }
        for (size_t s = 0; s < species_.size(); ++s)
            check(ctx_, ipic_gpu_upload_particles(ctx_, (int32_t)s, (int64_t)(pos[s].size() / 3), pos[s].data(),
                                                  vel[s].data()));
    }

    std::vector<Particle> getParticles() const {
        std::vector<Particle> out;
        for (size_t s = 0; s < species_.size(); ++s) {
            int64_t n = 0;
            check(ctx_, ipic_gpu_num_particles(ctx_, (int32_t)s, &n));
            std::vector<double> pos((size_t)n * 3), vel((size_t)n * 3);
            check(ctx_, ipic_gpu_download_particles(ctx_, (int32_t)s, pos.data(), vel.data(), nullptr, nullptr));
            for (int64_t i = 0; i < n; ++i) {
                Particle p;
                for (int k = 0; k < 3; ++k) {
                    p.position[k] = pos[3 * i + k];
                    p.velocity[k] = vel[3 * i + k];
                }
                p.q = species_[s].first;
                p.qom = species_[s].second;
                out.push_back(p);
            }
        }
        return out;
    }

    // initCells (cell.h:512-654) over the distribution with the reference's random streams, into
    // `species` (0: the first); returns the particles per cell (row-major)
    std::vector<std::int64_t> initCells(std::uint64_t numParticles, const distribution::uniform_pos_normal_speed& d,
                                        int species = 0) {
        ipic_ref_distribution r{};
        for (int k = 0; k < 3; ++k) {
            r.pos_min[k] = d.minPos[k];
            r.pos_max[k] = d.maxPos[k];
            r.vel_mean[k] = d.center[k];
            r.vel_stddev[k] = d.stddev[k];
        }
        r.num_particles = numParticles;
        r.seed = d.seed;
        std::vector<std::int64_t> counts((size_t)numCells());
        check(ctx_, ipic_gpu_init_reference(ctx_, (int32_t)species, &r, counts.data()));
        return counts;
    }

    // the parameter-file initCells (cell.h:699-776): init.particlesPerCell[0] particles on a
    // lattice in every cell, velocities driftVelocity[0] + thermalVelocity * the Box-Muller
    // Maxwellian of each cell's own std::minstd_rand stream, into `species`; the species' q should
    // be latticeCharge(...) of the same lattice (the charge every such particle carries)
    void initCells(const InitProperties& init, const Vector3& thermalVelocity, int species = 0) {
        if (init.particlesPerCell.empty() || init.driftVelocity.empty())
            throw Error(IPIC_E_ARG, "InitProperties needs particlesPerCell and driftVelocity of the first species");
        ipic_ref_lattice lat{};
        for (int k = 0; k < 3; ++k) {
            lat.ppc[k] = (int32_t)init.particlesPerCell[0][k];
            lat.vel_thermal[k] = thermalVelocity[k];
            lat.vel_drift[k] = init.driftVelocity[0][k];
        }
        check(ctx_, ipic_gpu_init_lattice(ctx_, (int32_t)species, &lat));
    }

    // initFields + initBcFields (field.h:81-138) of the dipole use case: E = -drift x B0, B = B0 plus
    // the dipole of externalMagneticField.z and planetRadius at objectCenter; returns the field
    // energies writeOutputData logs for them (E, B + Bext)
    std::pair<double, double> initFields(const InitProperties& init) {
        ipic_field_init f{};
        for (int k = 0; k < 3; ++k) {
            f.drift[k] = init.driftVelocity.empty() ? 0.0 : init.driftVelocity[0][k];
            f.B0[k] = init.magneticField[k];
            f.object_center[k] = props_.objectCenter[k];
        }
        f.planet_radius = props_.planetRadius;
        f.external_Bz = props_.externalMagneticField[2];
        double en[2] = {0.0, 0.0};
        check(ctx_, ipic_gpu_init_fields(ctx_, &f, en));
        return {en[0], en[1]};
    }

    // countParticlesInDomain, cell.h:87-100
    std::int64_t countParticlesInDomain() const {
        int64_t n = 0;
        check(ctx_, ipic_gpu_num_particles(ctx_, -1, &n));
        return n;
    }

    // outputNumberOfParticlesPerCell, cell.h:1360-1395 (row-major cells)
    std::vector<std::int64_t> getParticleCountsPerCell() const {
        std::vector<std::int64_t> c((size_t)numCells());
        check(ctx_, ipic_gpu_cell_counts(ctx_, -1, c.data()));
        return c;
    }

    // Field (nodes) and BcField (centres) state, row-major, one Vector3 per entry
    void setFields(const std::vector<Vector3>& E, const std::vector<Vector3>& Bc) {
        if ((std::int64_t)E.size() != numCells() || (std::int64_t)Bc.size() != numCells())
            throw Error(IPIC_E_ARG, "field arrays need one entry per cell");
        check(ctx_, ipic_gpu_upload_fields(ctx_, E[0].data(), Bc[0].data()));
    }

    void getFields(std::vector<Vector3>* E, std::vector<Vector3>* Bc, std::vector<Vector3>* Bn,
                   std::vector<Vector3>* J) const {
        const size_t n = (size_t)numCells();
        for (auto* v : {E, Bc, Bn, J})
            if (v) v->assign(n, Vector3{});
        check(ctx_, ipic_gpu_download_fields(ctx_, E ? (*E)[0].data() : nullptr, Bc ? (*Bc)[0].data() : nullptr,
                                             Bn ? (*Bn)[0].data() : nullptr, J ? (*J)[0].data() : nullptr));
    }

    // getParticlesKineticEnergy / Momentum (cell.h:1315-1338), getFieldEnergy (field.h:443-459)
    ipic_conservation conservation() const {
        ipic_conservation c{};
        check(ctx_, ipic_gpu_conservation(ctx_, &c));
        return c;
    }

    bool hasFieldOperator() const { return has_operator_; }

    void setFieldOperator(const ipic_field_operator* op) {
        check(ctx_, ipic_gpu_set_field_operator(ctx_, op));
        has_operator_ = op != nullptr;
    }

    ipic_step_stats lastStats() const { return last_; }

  private:
    template <class P, class F, class M>
    friend DurationMeasurement simulateSteps(std::uint64_t, Universe&);

    UniverseProperties props_;
    std::vector<std::pair<double, double>> species_;
    ipic_ctx* ctx_ = nullptr;
    bool has_operator_ = false;
    ipic_step_stats last_{};
};

// createUniverseFromDistribution (universe.h:85-103): the particles initCells draws, all with
// q = e and qom = e / m_p as there; the fields start at zero (Universe::setFields)
inline Universe createUniverseFromDistribution(const UniverseProperties& props, std::uint64_t numParticles,
                                               const distribution::uniform_pos_normal_speed& dist) {
    Universe u(props, {{kElementaryCharge, kElementaryCharge / kProtonMass}});
    u.initCells(numParticles, dist);
    return u;
}

// q_factor of cell.h:711-716: the charge of every particle of the parameter-file initCells
inline double latticeCharge(const UniverseProperties& props, const std::array<unsigned, 3>& particlesPerCell,
                            double qom, double rhoInit) {
    const double w[3] = {props.cellWidth[0], props.cellWidth[1], props.cellWidth[2]};
    const int32_t ppc[3] = {(int32_t)particlesPerCell[0], (int32_t)particlesPerCell[1], (int32_t)particlesPerCell[2]};
    return ipic_ref_lattice_charge(w, ppc, qom, rhoInit);
}

// createUniverseFromParams (universe.h:53-75) for what it reads of Parameters besides
// InitProperties: the first species' qom and thermal velocity (uth, vth, wth).  One species of
// (latticeCharge, qom); cells from the parameter-file initCells, fields from initFields.
inline Universe createUniverseFromParams(const UniverseProperties& props, const InitProperties& init, double qom,
                                         const Vector3& thermalVelocity) {
    if (init.particlesPerCell.empty()) throw Error(IPIC_E_ARG, "InitProperties needs particlesPerCell");
    Universe u(props, {{latticeCharge(props, init.particlesPerCell[0], qom, init.rhoInit), qom}});
    u.initCells(init, thermalVelocity);
    u.initFields(init);
    return u;
}

template <class ParticleToFieldProjector, class FieldSolver, class ParticleMover>
DurationMeasurement simulateSteps(std::uint64_t numSteps, Universe& universe) {
    static_assert(std::is_same<ParticleToFieldProjector, detail::gpu_particle_to_field_projector>::value,
                  "the GPU step deposits with its own projector");
    static_assert(std::is_same<ParticleMover, detail::gpu_particle_mover>::value,
                  "the GPU step moves with its own mover");
    constexpr bool op = std::is_same<FieldSolver, detail::gpu_field_operator_solver>::value;
    static_assert(op || std::is_same<FieldSolver, detail::gpu_leapfrog_field_solver>::value,
                  "field solver policy must be gpu_leapfrog_field_solver or gpu_field_operator_solver");
    if (op != universe.has_operator_)
        throw Error(IPIC_E_STATE, op ? "gpu_field_operator_solver needs Universe::setFieldOperator"
                                     : "the universe carries a field operator; use gpu_field_operator_solver");
    ipic_step_stats st{};
    check(universe.ctx_, ipic_gpu_step(universe.ctx_, (int64_t)numSteps, &st));
    universe.last_ = st;
    // millisecond resolution, as getTimeCount of simulator.h:118-123
    auto ms = [](double s) { return std::floor(s * 1000.0) / 1000.0; };
    return {ms(st.first_step_s), ms(st.remaining_steps_s)};
}

template <class ParticleToFieldProjector = detail::gpu_particle_to_field_projector,
          class FieldSolver = detail::gpu_leapfrog_field_solver,
          class ParticleMover = detail::gpu_particle_mover>
void simulateStep(Universe& universe) {
    simulateSteps<ParticleToFieldProjector, FieldSolver, ParticleMover>(1, universe);
}

}  // namespace ipic3d_gpu

#endif
