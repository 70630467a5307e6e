/*
 * ipic_gpu.h -- C ABI of the MI355X (gfx950) particle-in-cell step library libipic_gpu.so.
 *
 * This is the drop-in boundary for the reference's step API:
 *   template<P,F,M> DurationMeasurement simulateSteps(std::uint64_t, Universe&)
 *   (code/app/include/ipic3d/app/simulator.h:42-47, 127-242)
 * with the projector / field solver / mover policies of simulator.h:244-276 replaced by
 * the GPU implementation.  The C++ mirror in include/ipic3d_gpu/ipic3d.hpp binds it to a
 * Universe (universe.h:12-51); any other host language binds these plain-C entry points.
 *
 * Conventions: plain pointers and sizes, no C++ types; every call returns 0 on success and
 * a negative IPIC_E_* code on failure, with a message from ipic_gpu_last_error(ctx)
 * (ctx may be NULL for a failed create).  No exception crosses the ABI.
 *
 * Grid conventions: the physics grid is N-periodic (nx*ny*nz cells).  Node, centre and
 * cell arrays are row-major with index (x*ny + y)*nz + z and 3 doubles per entry.
 *   E  -- electric field on nodes       (reference Field::E at index pos+1, field.h:14-18)
 *   Bc -- magnetic field on cell centres (reference BcField::Bc at pos+1, field.h:20-23)
 *   Bn -- magnetic field on nodes = 1/8 sum of the 8 surrounding Bc (field.h:234-247)
 *   J  -- current density on nodes      (reference CurrentDensity::J at pos, field.h:25-27)
 */
#ifndef IPIC_GPU_H
#define IPIC_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPIC_MAX_SPECIES 8

#define IPIC_OK 0
#define IPIC_E_ARG -1      /* invalid argument */
#define IPIC_E_HIP -2      /* HIP runtime error (no device, allocation, launch) */
#define IPIC_E_STATE -3    /* call not valid in the current state */
#define IPIC_E_RANGE -4    /* a size exceeds a documented limit */

/* flags */
#define IPIC_FLAG_TRACK_IDS 1u    /* carry a uint32 id per particle (input order) */
#define IPIC_FLAG_PHASE_TIMING 4u /* record per-phase HIP events (synchronises after each step) */
#define IPIC_FLAG_REFLECT 8u      /* particles reflect at the global faces (exportParticles,
                                     cell.h:1103-1121) instead of wrapping; fields stay periodic */
#define IPIC_FLAG_DIPOLE_MOVER 32u /* the mover as shipped (cell.h:919-945): E = 0 and the analytic
                                     field of a z dipole at the coordinate origin, strength
                                     ipic_desc.dipole_moment; the field grids are not gathered */
#define IPIC_FLAG_FORWARD_SOLVER 64u /* field policy: the explicit forward solver (solveFieldForward,
                                     field.h:293-337) instead of leapfrog */
#define IPIC_FLAG_REPRODUCIBLE 128u /* after every pass that moves particles, each cell's arrivals
                                     (particles that changed cell) are put in the order of their
                                     contents rather than the order of atomic slot counters, so
                                     steps are bitwise reproducible from run to run; the reference's
                                     import (cell.h:1272-1306) is sequential, hence fixed, too */
#define IPIC_FLAG_UNIFORM_BACKGROUND 16u /* upload_fields keeps only the uniform part of the
                                     fields as static background; the rest becomes the evolving
                                     FP32 state (needed by the field operator) */

/* Particle species: every particle of a species shares Particle::q and Particle::qom
 * (particle.h:15-16). */
typedef struct ipic_species_desc {
    double q;
    double qom;
} ipic_species_desc;

/* UniverseProperties (universe_properties.h:52-142) plus the GPU layout parameters. */
typedef struct ipic_desc {
    int32_t n[3];                 /* cells per axis (UniverseProperties::size); each <= 2^21, product <= 2^30 */
    int32_t nspecies;             /* 1..IPIC_MAX_SPECIES */
    double cell_width[3];         /* UniverseProperties::cellWidth */
    double origin[3];             /* UniverseProperties::origin */
    double dt;                    /* UniverseProperties::dt */
    double c;                     /* UniverseProperties::speedOfLight */
    ipic_species_desc species[IPIC_MAX_SPECIES];
    int32_t field_substeps;       /* leapfrog substeps per particle step with J held fixed; 0: static
                                     field (solveFieldStatically, field.h:272-284) */
    int32_t sink;                 /* 1: remove particles with |x - sink_center|^2 <= sink_radius^2
                                     (cell.h:1123-1128, planetRadius / objectCenter) */
    double sink_center[3];
    double sink_radius;
    int32_t tile[3];              /* cells per tile along each axis; 0 = automatic (<= 8, divides n) */
    uint32_t flags;               /* IPIC_FLAG_* */
    double dipole_moment;         /* -externalMagneticField.z * planetRadius^3 (cell.h:921), read
                                     with IPIC_FLAG_DIPOLE_MOVER */
} ipic_desc;

/* DurationMeasurement (simulator.h:37-40) plus counters of the call. */
typedef struct ipic_step_stats {
    double first_step_s;          /* wall time of the first step of the call */
    double remaining_steps_s;     /* wall time of the other steps */
    int64_t steps;
    int64_t particles;            /* live particles after the call, all species */
    int64_t removed;              /* particles removed by the sink during the call */
    /* filled when IPIC_FLAG_PHASE_TIMING is set: summed GPU milliseconds per phase */
    double ms_field;              /* field phase: leapfrog substeps or the field operator with its
                                     checks, interpC2N, cross-tile J sum */
    double ms_push;               /* fused gather + Boris + rebin + deposit kernel */
    double ms_scan;               /* cell-count prefix sums */
    double ms_deposit;            /* initial deposit of a fresh particle state */
    /* field operator (ipic_gpu_set_field_operator) */
    int64_t operator_steps;       /* field phases whose result came from the operator */
    int64_t fallback_steps;       /* operator results replaced by leapfrog (failed check or audit) */
    int64_t audits;               /* operator steps compared against leapfrog */
    double max_audit_error;       /* largest audited |F_op - F_lf| / |F_lf| of the total field
                                     (background included), |F|^2 = sum |E|^2 + c^2 |Bc|^2 */
    double max_audit_error_increment; /* the same error relative to |F_lf - F_before| */
    double max_divb_residual;     /* largest max |div+ (Bc_after - Bc_before)| of an operator step */
    double max_energy_residual;   /* largest field-energy residual of an operator step */
    double max_audit_error_evolving; /* the audited error relative to the evolving field |F_lf|
                                     without the uniform background (dE, dBc): the audit's gate */
} ipic_step_stats;

/* FP64 conservation sums, reduced on the GPU (cell.h:1315-1355, field.h:443-459). */
typedef struct ipic_conservation {
    double kinetic_energy;        /* sum 0.5 (q/qom) |v|^2      (cell.h:1315-1324) */
    double momentum;              /* sum (q/qom) |v|            (cell.h:1329-1338) */
    double charge;                /* sum q */
    double field_energy_E;        /* 0.5 * sum |E|^2 * vol over nodes */
    double field_energy_B;        /* 0.5 * c^2 * sum |Bc|^2 * vol over centres */
    double divB_max;              /* max |div+ Bc| on the index lattice */
    double divB_l2;               /* l2 norm of div+ Bc */
    double EdotJ;                 /* sum E . J * vol over nodes (work of the field on J) */
    int64_t particles;            /* live particles (countParticlesInDomain, cell.h:87-100) */
} ipic_conservation;

typedef struct ipic_ctx ipic_ctx;

/* Universe(properties) (universe.h:40-51): allocates device state for the grid and species.
 * Fields start at zero; no particles. */
int ipic_gpu_create(const ipic_desc* desc, ipic_ctx** out);
int ipic_gpu_destroy(ipic_ctx* ctx);
const char* ipic_gpu_last_error(const ipic_ctx* ctx);

/* Replaces the particles of one species (Cell::particles of every cell, cell.h:31-38).
 * pos, vel: n x 3 doubles, absolute positions (Particle::position / velocity).  Positions
 * outside the box are wrapped periodically; each particle is binned into the cell
 * floor((x - origin) / width). */
int ipic_gpu_upload_particles(ipic_ctx* ctx, int32_t species, int64_t n,
                              const double* pos, const double* vel);

/* Synthetic initialisation on the GPU (the role of initCells, cell.h:700-776): for every
 * cell, ppc[cell] particles (ppc == NULL: uniform_ppc each) with offsets uniform in the
 * cell and velocities drift + vth * N(0,1) per component, drawn from the counter-based
 * generator keyed by (seed, species, cell, index). */
int ipic_gpu_init_particles(ipic_ctx* ctx, int32_t species, const int32_t* ppc, int32_t uniform_ppc,
                            const double vth[3], const double drift[3], uint64_t seed);

/* Field (nodes) and BcField (centres) state, n^3 x 3 doubles each; NULL leaves one as is.
 * The uploaded state becomes a static background held in FP64 (its uniform part as
 * constants, the rest as arrays with FP64 curl sources) plus a zero FP32 perturbation that the
 * leapfrog evolves, so a large static field costs no precision in the curls; with
 * IPIC_FLAG_UNIFORM_BACKGROUND only the uniform part is background. */
int ipic_gpu_upload_fields(ipic_ctx* ctx, const double* E, const double* Bc);

/* simulateSteps (simulator.h:127-242) with the intended full loop of simulator.h:189-223:
 * per step deposit J (cell.h:787-824), leapfrog E then Bc (field.h:356-372) and
 * interpC2N (field.h:234-247), then gather + Boris move (cell.h:896-964) and rebin
 * (cell.h:975-1310) with periodic boundaries.  stats may be NULL. */
int ipic_gpu_step(ipic_ctx* ctx, int64_t nsteps, ipic_step_stats* stats);

/* countParticlesInDomain (cell.h:87-100) per species (species < 0: all). */
int ipic_gpu_num_particles(ipic_ctx* ctx, int32_t species, int64_t* n);

/* Per-cell particle counts (outputNumberOfParticlesPerCell, cell.h:1360-1395), ncell int64,
 * species < 0: all species. */
int ipic_gpu_cell_counts(ipic_ctx* ctx, int32_t species, int64_t* counts);

/* Particles of one species sorted by cell (row-major cell index), each cell's particles in
 * the order held on the device.  pos, vel: n x 3 doubles (absolute); cell, ids nullable
 * (ids need IPIC_FLAG_TRACK_IDS). */
int ipic_gpu_download_particles(ipic_ctx* ctx, int32_t species, double* pos, double* vel,
                                int64_t* cell, uint32_t* ids);

/* Field state; any pointer may be NULL.  J is the current density used by the last
 * field update; with the static field (field_substeps = 0, no update uses J) the current
 * density of the current particle state. */
int ipic_gpu_download_fields(ipic_ctx* ctx, double* E, double* Bc, double* Bn, double* J);

/* FP64 conservation sums of the current state. */
int ipic_gpu_conservation(ipic_ctx* ctx, ipic_conservation* out);

/* The two halves of a step (ipic_gpu_step runs field then particles, nsteps times).
 * Field phase: the J of the last deposit summed over tiles, then `field_substeps` leapfrog
 * substeps (field.h:356-372) with J held fixed -- or, with use_operator = 1 and an operator
 * set, the operator with its checks -- then interpC2N for the gather.  Particle phase: gather
 * + Boris + rebin + deposit and the cell-count scans.  stats (nullable) accumulate. */
int ipic_gpu_field_step(ipic_ctx* ctx, int32_t use_operator, ipic_step_stats* stats);
int ipic_gpu_particle_step(ipic_ctx* ctx, ipic_step_stats* stats);

/* Learned field step (no reference counterpart; operator and parameter layout in
 * ipic_nop.h).  With an operator set, every field phase of ipic_gpu_step runs it in place of
 * the leapfrog substeps and falls back to leapfrog when max |div+ (Bc' - Bc)| > divb_tol or
 * when the field-energy residual |W' - W + dt (E.J + E'.J)/2 vol| / max(W, W') > energy_tol,
 * where W = 1/2 sum (|E0 + dE|^2 + c^2 |dBc|^2) vol is the energy of the field without its
 * uniform magnetic background.  Every audit_every-th operator step is also advanced by
 * leapfrog from the same state; the operator's result is kept when its error relative to the
 * evolving leapfrog field (|F_op - F_lf| / |F_lf - F0|, F0 the uniform background) is at most
 * audit_tol, else the leapfrog result is.  Needs a uniform
 * background field (upload_fields with spatially constant E, Bc), nx % 4 == 0, ny % 4 == 0,
 * nz % 16 == 0 and substeps == field_substeps.  op == NULL switches the operator off. */
typedef struct ipic_field_operator {
    const float* params;          /* host FP32 parameter vector (ipic_nop_param_count() floats) */
    float in_E, in_B, in_J;       /* input scales */
    float out_E, out_P;           /* output scales */
    int32_t substeps;             /* leapfrog substeps one application stands for */
    int32_t audit_every;          /* 0: no audits */
    double divb_tol;
    double energy_tol;
    double audit_tol;
} ipic_field_operator;
int ipic_gpu_set_field_operator(ipic_ctx* ctx, const ipic_field_operator* op);

/* Device view of the field state for a trainer (valid until the next call on ctx): float4
 * arrays of n^3 entries, (x, y, z, unused).  J is formed from the last deposit first. */
typedef struct ipic_field_state {
    float* E;                     /* E - E0 on nodes */
    float* Bc;                    /* Bc - B0 on centres */
    float* J;                     /* current density on nodes */
    double E0[3];                 /* uniform background */
    double B0[3];
    int32_t uniform_background;   /* 1: the background is E0, B0 only */
} ipic_field_state;
int ipic_gpu_field_state(ipic_ctx* ctx, ipic_field_state* out);

/* The reference's particle initialisation with its own random streams: initCells
 * (cell.h:512-654) over distribution::uniform_pos_normal_speed (cell.h:405-440), as the main
 * driver calls it through createUniverseFromDistribution (src/ipic3d.cxx:74-85,
 * universe.h:85-103).  Positions are uniform in [pos_min, pos_max) per component
 * (std::minstd_rand + std::uniform_real_distribution<double>, cell.h:138-175), velocities
 * vel_mean + vel_stddev * the reference's ziggurat normal (ziggurat_normal_distribution.h,
 * cell.h:223-257).  Phase 1 (cell.h:525-610) estimates the per-cell counts from 100 pseudo
 * particles per cell of the distribution's own stream; phase 2 (cell.h:613-650) draws each
 * cell's particles from streams seeded by the cell coordinate.  Host computation, bit-identical
 * to the reference's streams; the geometry is desc's n, cell_width and origin.  Every
 * particle gets the species the caller names (the reference sets q = e, qom = e / m_p). */
typedef struct ipic_ref_distribution {
    double pos_min[3], pos_max[3];     /* vector::uniform bounds of the positions */
    double vel_mean[3], vel_stddev[3]; /* vector::normal of the velocities */
    uint64_t num_particles;            /* total particles, initCells' numParticles */
    uint32_t seed;                     /* the distribution's seed (positions seed+1, velocities seed+2) */
    int32_t nthreads;                  /* host threads, 0: all */
} ipic_ref_distribution;

/* Phase 1: the number of particles of each cell (ncell int64, row-major).  Fails with
 * IPIC_E_ARG where the reference's initCells asserts (no remainder to spread, cell.h:557). */
int ipic_ref_cell_counts(const ipic_desc* desc, const ipic_ref_distribution* dist, int64_t* counts);
/* Phase 2: the particles of all cells, row-major cells, each cell's particles in draw order;
 * pos, vel: sum(counts) x 3 doubles (absolute positions). */
int ipic_ref_particles(const ipic_desc* desc, const ipic_ref_distribution* dist, const int64_t* counts,
                       double* pos, double* vel);
/* Message of the last failed ipic_ref_* call on this thread. */
const char* ipic_ref_last_error(void);
/* Both phases, then ipic_gpu_upload_particles of the result into `species`.  counts (nullable)
 * receives the phase-1 counts. */
int ipic_gpu_init_reference(ipic_ctx* ctx, int32_t species, const ipic_ref_distribution* dist, int64_t* counts);

/* The parameter-file form of initCells (cell.h:699-776), which createUniverseFromParams
 * (universe.h:53-75) calls: every cell holds ppc[0] x ppc[1] x ppc[2] particles at the centres
 * of a regular sub-lattice, position (l + 0.5) * cell_width / ppc + cell corner per axis; each
 * cell (x, y, z) draws from its own std::minstd_rand seeded with (unsigned)(10000 x + 100 y + z),
 * four draws u / max() per particle, and sets the Box-Muller Maxwellian
 *   v = vel_drift + vel_thermal * (r0 cos t0, r0 sin t0, r1 cos t1),
 *   r = sqrt(-2 ln(1 - 0.999999 u)), t = 2 pi u      (draw order r0, t0, r1, t1).
 * The members are the Parameters keys the reference reads: npcel{x,y,z}, {u,v,w}th, {u,v,w}0 of
 * the first species.  Host computation, bit-identical to the reference's streams. */
typedef struct ipic_ref_lattice {
    int32_t ppc[3];               /* particles per cell along each axis */
    int32_t nthreads;             /* host threads, 0: all */
    double vel_thermal[3];        /* uth, vth, wth */
    double vel_drift[3];          /* u0, v0, w0 */
} ipic_ref_lattice;

/* The particles of all cells: ncell * ppc[0] * ppc[1] * ppc[2] of them, row-major cells, each
 * cell's in lattice order (x slowest); pos, vel: that many x 3 doubles (absolute positions). */
int ipic_ref_lattice_particles(const ipic_desc* desc, const ipic_ref_lattice* lat, double* pos, double* vel);
/* The charge every particle of that initialisation carries (q_factor, cell.h:711-716):
 * sign(qom) * cell volume / (ppc product) * rho / (4 pi).  Set it as the species' q. */
double ipic_ref_lattice_charge(const double cell_width[3], const int32_t ppc[3], double qom, double rho);
/* ipic_ref_lattice_particles, then ipic_gpu_upload_particles of the result into `species`. */
int ipic_gpu_init_lattice(ipic_ctx* ctx, int32_t species, const ipic_ref_lattice* lat);

/* The initial field of the reference's dipole use case (UseCase::Dipole):
 *   getDipoleFieldAt (field.h:43-78): E = -drift x B0; B = B0 + Bext with
 *     Bext = -external_Bz a^3 (3 x z, 3 y z, 2 z^2 - x^2 - y^2) / r^5 for r > a (r = location -
 *     object_center, a = planet_radius), Bext = 0 inside the planet;
 *   initFields (field.h:81-118): that field at the node locations origin + i * cell_width;
 *   initBcFields (field.h:120-138): Bc of cell i = 1/8 of node B over nodes i + {0,1}^3
 *     (interpN2C, field.h:252-265).
 * The members are the InitProperties / UniverseProperties fields these functions read. */
typedef struct ipic_field_init {
    double drift[3];              /* InitProperties::driftVelocity[0] */
    double B0[3];                 /* InitProperties::magneticField */
    double object_center[3];      /* UniverseProperties::objectCenter */
    double planet_radius;         /* UniverseProperties::planetRadius */
    double external_Bz;           /* UniverseProperties::externalMagneticField.z */
} ipic_field_init;

/* Host computation on desc's grid: E and node B at the N^3 nodes, Bc at the N^3 centres
 * (n^3 x 3 doubles each, row-major, any may be NULL; the last centre layer averages the node
 * locations at N as the reference's does).  energies (nullable) receives getFieldEnergy
 * (field.h:443-459) of E and of B + Bext over the (N+1)^3 nodes of initFields, the values
 * writeOutputData logs (simulator.h:73-79). */
int ipic_field_init_host(const ipic_desc* desc, const ipic_field_init* init, double* E, double* Bn, double* Bc,
                         double energies[2]);
/* Message of the last failed ipic_field_init_host call on this thread. */
const char* ipic_field_init_last_error(void);
/* initFields + initBcFields as the context's field state: ipic_gpu_upload_fields of the
 * host computation's E and Bc.  energies as above (nullable). */
int ipic_gpu_init_fields(ipic_ctx* ctx, const ipic_field_init* init, double energies[2]);

/* Device properties used by the tests and benchmark. */
int ipic_gpu_device_info(char* name, int32_t name_len, int32_t* cu_count, int64_t* hbm_bytes);

#ifdef __cplusplus
}
#endif
#endif
