/*
 * ipic_oracle.c -- FP64 CPU restatement of the AllScale iPIC3D particle-in-cell step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See ipic_oracle.h for the
 * composition decisions D1-D7.  Reference citations are file:line into
 * code/app/include/ipic3d/app/ unless a path is given.
 */
#include "ipic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------ */
/* small vector helpers                                                                 */
/* ------------------------------------------------------------------------------------ */

static inline void cross3(const double a[3], const double b[3], double r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

static inline int64_t wrapi(int64_t i, int64_t n) {
    int64_t r = i % n;
    return r < 0 ? r + n : r;
}

static inline int64_t idx3(const int64_t n[3], int64_t x, int64_t y, int64_t z) {
    return (x * n[1] + y) * n[2] + z;
}

static inline int64_t pidx(const int64_t n[3], int64_t x, int64_t y, int64_t z) {
    return idx3(n, wrapi(x, n[0]), wrapi(y, n[1]), wrapi(z, n[2]));
}

/* ------------------------------------------------------------------------------------ */
/* particle kernels                                                                     */
/* ------------------------------------------------------------------------------------ */

/* particle.h:24-26  position += velocity * dt */
void or_update_position(or_particle* p, double dt) {
    for (int d = 0; d < 3; ++d) p->pos[d] += p->vel[d] * dt;
}

/* particle.h:31-49  Boris rotation:
 *   k = qom*dt/2, t = kB, s = 2t/(1+|t|^2), v- = v + kE, v' = v- + v- x t,
 *   v+ = v- + v' x s, v = v+ + kE */
void or_update_velocity(or_particle* p, const double E[3], const double B[3], double dt) {
    const double k = p->qom * 0.5 * dt;
    double t[3], s[3], vm[3], vp[3], cr[3];
    for (int d = 0; d < 3; ++d) t[d] = k * B[d];
    const double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    for (int d = 0; d < 3; ++d) s[d] = (2.0 * t[d]) / (1.0 + t2);
    for (int d = 0; d < 3; ++d) vm[d] = p->vel[d] + k * E[d];
    cross3(vm, t, cr);
    for (int d = 0; d < 3; ++d) vp[d] = vm[d] + cr[d];
    cross3(vp, s, cr);
    for (int d = 0; d < 3; ++d) p->vel[d] = vm[d] + cr[d] + k * E[d];
}

/* cell.h:867-885  sum_ijk corner*fac / vol, fac = (i?x:1-x)(j?y:1-y)(k?z:1-z) */
double or_trilinear(const double corners[8], const double pos[3], double vol) {
    double res = 0.0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                const double fac = (i == 0 ? (1 - pos[0]) : pos[0]) *
                                   (j == 0 ? (1 - pos[1]) : pos[1]) *
                                   (k == 0 ? (1 - pos[2]) : pos[2]);
                res += corners[(i * 2 + j) * 2 + k] * fac;
            }
    return res / vol;
}

/* ------------------------------------------------------------------------------------ */
/* reference-layout stencils (dense [dx][dy][dz][3] grids)                              */
/* ------------------------------------------------------------------------------------ */

static inline const double* at3(const int64_t dims[3], const double* a, int64_t x, int64_t y, int64_t z) {
    return a + 3 * idx3(dims, x, y, z);
}
static inline double* at3w(const int64_t dims[3], double* a, int64_t x, int64_t y, int64_t z) {
    return a + 3 * idx3(dims, x, y, z);
}

/* Half-sum difference of component `comp` along axis `ax` over a 2x2x2 stencil:
 * 0.25 * sum_ijk sign * S[ijk].comp / w[ax], sign = +1 where the index along ax equals
 * `plus_index`, -1 otherwise.  field.h:180-191 (plus_index 0) and field.h:216-227 (1). */
static double stencil_diff(const double S[8][3], int ax, int comp, int plus_index, const double w[3]) {
    double acc = 0.0;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                const int id[3] = {i, j, k};
                const double v = S[(i * 2 + j) * 2 + k][comp];
                acc += (id[ax] == plus_index) ? v : -v;
            }
    return 0.25 * acc / w[ax];
}

/* field.h:162-193: Bs[ijk] = Bc[pos - (i,j,k)] */
void or_ref_curlB(const int64_t dims[3], const double* Bc, const int64_t pos[3],
                  const double w[3], double curl[3]) {
    double S[8][3];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                memcpy(S[(i * 2 + j) * 2 + k], at3(dims, Bc, pos[0] - i, pos[1] - j, pos[2] - k), 3 * sizeof(double));
    curl[0] = stencil_diff(S, 1, 2, 0, w) - stencil_diff(S, 2, 1, 0, w);
    curl[1] = stencil_diff(S, 2, 0, 0, w) - stencil_diff(S, 0, 2, 0, w);
    curl[2] = stencil_diff(S, 0, 1, 0, w) - stencil_diff(S, 1, 0, 0, w);
}

/* field.h:198-229: Es[ijk] = E[pos + (i,j,k)] */
void or_ref_curlE(const int64_t dims[3], const double* E, const int64_t pos[3],
                  const double w[3], double curl[3]) {
    double S[8][3];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                memcpy(S[(i * 2 + j) * 2 + k], at3(dims, E, pos[0] + i, pos[1] + j, pos[2] + k), 3 * sizeof(double));
    curl[0] = stencil_diff(S, 1, 2, 1, w) - stencil_diff(S, 2, 1, 1, w);
    curl[1] = stencil_diff(S, 2, 0, 1, w) - stencil_diff(S, 0, 2, 1, w);
    curl[2] = stencil_diff(S, 0, 1, 1, w) - stencil_diff(S, 1, 0, 1, w);
}

/* field.h:234-247: node B = 1/8 sum Bc[pos - (i,j,k)] */
void or_ref_interpC2N(const int64_t bc_dims[3], const double* Bc, const int64_t node_dims[3],
                      double* B, const int64_t pos[3]) {
    double acc[3] = {0, 0, 0};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                const double* v = at3(bc_dims, Bc, pos[0] - i, pos[1] - j, pos[2] - k);
                for (int d = 0; d < 3; ++d) acc[d] += v[d];
            }
    double* o = at3w(node_dims, B, pos[0], pos[1], pos[2]);
    for (int d = 0; d < 3; ++d) o[d] = 0.125 * acc[d];
}

/* field.h:252-265: Bc = 1/8 sum B[pos + (i,j,k)] */
void or_ref_interpN2C(const int64_t node_dims[3], const double* B, const int64_t bc_dims[3],
                      double* Bc, const int64_t pos[3]) {
    double acc[3] = {0, 0, 0};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                const double* v = at3(node_dims, B, pos[0] + i, pos[1] + j, pos[2] + k);
                for (int d = 0; d < 3; ++d) acc[d] += v[d];
            }
    double* o = at3w(bc_dims, Bc, pos[0], pos[1], pos[2]);
    for (int d = 0; d < 3; ++d) o[d] = 0.125 * acc[d];
}

/* field.h:407-440: for every (a,b) in [1,end)^2 copy f[0]=f[end-1], f[end]=f[1] on the
 * three face pairs; end = size-1.  Corners and edges are not written. */
static void ref_faces(int64_t size, double* f) {
    const int64_t end = size - 1;
    const int64_t d[3] = {size, size, size};
    for (int64_t a = 1; a < end; ++a)
        for (int64_t b = 1; b < end; ++b) {
            memcpy(at3w(d, f, 0, a, b), at3(d, f, end - 1, a, b), 3 * sizeof(double));
            memcpy(at3w(d, f, end, a, b), at3(d, f, 1, a, b), 3 * sizeof(double));
            memcpy(at3w(d, f, a, 0, b), at3(d, f, a, end - 1, b), 3 * sizeof(double));
            memcpy(at3w(d, f, a, end, b), at3(d, f, a, 1, b), 3 * sizeof(double));
            memcpy(at3w(d, f, a, b, 0), at3(d, f, a, b, end - 1), 3 * sizeof(double));
            memcpy(at3w(d, f, a, b, end), at3(d, f, a, b, 1), 3 * sizeof(double));
        }
}

void or_ref_update_boundaries(int64_t field_size, double* E, double* B, int64_t bc_size, double* Bc) {
    if (E) ref_faces(field_size, E);
    if (B) ref_faces(field_size, B);
    if (Bc) ref_faces(bc_size, Bc);
}

/* field.h:356-372 (PSC branch), one node in place, reference indexing.
 * The B update guard `pos < bcfield.size() - 1` is read as every component strictly less. */
void or_ref_leapfrog_node(const int64_t fd[3], double* E, const int64_t bd[3], double* Bc,
                          const int64_t jd[3], const double* J, const int64_t p[3],
                          const double w[3], double dt, double c) {
    const double c2 = c * c;
    double* e = at3w(fd, E, p[0], p[1], p[2]);
    const double* b = at3(bd, Bc, p[0], p[1], p[2]);
    const double* bmx = at3(bd, Bc, p[0] - 1, p[1], p[2]);
    const double* bmy = at3(bd, Bc, p[0], p[1] - 1, p[2]);
    const double* bmz = at3(bd, Bc, p[0], p[1], p[2] - 1);
    const double* j = at3(jd, J, p[0] - 1, p[1] - 1, p[2] - 1);
    e[0] = e[0] + dt * (c2 * ((b[2] - bmy[2]) / w[1] - (b[1] - bmz[1]) / w[2]) - j[0]);
    e[1] = e[1] + dt * (c2 * ((b[0] - bmz[0]) / w[2] - (b[2] - bmx[2]) / w[0]) - j[1]);
    e[2] = e[2] + dt * (c2 * ((b[1] - bmx[1]) / w[0] - (b[0] - bmy[0]) / w[1]) - j[2]);
    if (p[0] < bd[0] - 1 && p[1] < bd[1] - 1 && p[2] < bd[2] - 1) {
        double* bw = at3w(bd, Bc, p[0], p[1], p[2]);
        const double* epx = at3(fd, E, p[0] + 1, p[1], p[2]);
        const double* epy = at3(fd, E, p[0], p[1] + 1, p[2]);
        const double* epz = at3(fd, E, p[0], p[1], p[2] + 1);
        bw[0] = bw[0] - dt * ((epy[2] - e[2]) / w[1] - (epz[1] - e[1]) / w[2]);
        bw[1] = bw[1] - dt * ((epz[0] - e[0]) / w[2] - (epx[2] - e[2]) / w[0]);
        bw[2] = bw[2] - dt * ((epx[1] - e[1]) / w[0] - (epy[0] - e[0]) / w[1]);
    }
}

/* field.h:43-78 */
void or_dipole_field(const double loc[3], const double drift[3], const double B0[3],
                     const double center[3], double a, double B1z,
                     double E[3], double Bext[3], double B[3]) {
    double cr[3];
    cross3(drift, B0, cr);
    for (int d = 0; d < 3; ++d) E[d] = -1.0 * cr[d];
    const double dx = loc[0] - center[0], dy = loc[1] - center[1], dz = loc[2] - center[2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 > a * a) {
        const double fac1 = -B1z * pow(a, 3) / pow(r2, 2.5);
        Bext[0] = 3.0 * dx * dz * fac1;
        Bext[1] = 3.0 * dy * dz * fac1;
        Bext[2] = (2.0 * dz * dz - dx * dx - dy * dy) * fac1;
    } else {
        Bext[0] = Bext[1] = Bext[2] = 0.0;
    }
    for (int d = 0; d < 3; ++d) B[d] = B0[d] + Bext[d];
}

/* ------------------------------------------------------------------------------------ */
/* universe                                                                             */
/* ------------------------------------------------------------------------------------ */

struct or_universe {
    or_props p;
    int64_t ncell;
    int64_t* count;
    int64_t* cap;
    or_particle** cell;
    double* E;   /* nodes   [ncell][3] */
    double* Bc;  /* centres [ncell][3] */
    double* Bn;  /* nodes   [ncell][3] (interpC2N of Bc) */
    double* J;   /* nodes   [ncell][3] */
};

or_universe* or_create(const or_props* props) {
    if (!props || props->n[0] <= 0 || props->n[1] <= 0 || props->n[2] <= 0) return NULL;
    or_universe* u = (or_universe*)calloc(1, sizeof(or_universe));
    u->p = *props;
    if (u->p.field_substeps < 0) u->p.field_substeps = 0;
    u->ncell = props->n[0] * props->n[1] * props->n[2];
    u->count = (int64_t*)calloc((size_t)u->ncell, sizeof(int64_t));
    u->cap = (int64_t*)calloc((size_t)u->ncell, sizeof(int64_t));
    u->cell = (or_particle**)calloc((size_t)u->ncell, sizeof(or_particle*));
    u->E = (double*)calloc((size_t)u->ncell * 3, sizeof(double));
    u->Bc = (double*)calloc((size_t)u->ncell * 3, sizeof(double));
    u->Bn = (double*)calloc((size_t)u->ncell * 3, sizeof(double));
    u->J = (double*)calloc((size_t)u->ncell * 3, sizeof(double));
    return u;
}

void or_destroy(or_universe* u) {
    if (!u) return;
    for (int64_t c = 0; c < u->ncell; ++c) free(u->cell[c]);
    free(u->cell); free(u->count); free(u->cap);
    free(u->E); free(u->Bc); free(u->Bn); free(u->J);
    free(u);
}

int64_t or_num_cells(const or_universe* u) { return u->ncell; }

static void cell_push(or_universe* u, int64_t c, const or_particle* p) {
    if (u->count[c] == u->cap[c]) {
        int64_t nc = u->cap[c] ? 2 * u->cap[c] : 8;
        u->cell[c] = (or_particle*)realloc(u->cell[c], (size_t)nc * sizeof(or_particle));
        u->cap[c] = nc;
    }
    u->cell[c][u->count[c]++] = *p;
}

/* getCellCoordinates (cell.h:64-67) uses truncation; here floor with periodic wrap so
 * that any position is assigned to exactly one cell of the box. */
int or_add_particles(or_universe* u, int64_t n, const or_particle* ps) {
    const or_props* P = &u->p;
    for (int64_t i = 0; i < n; ++i) {
        or_particle p = ps[i];
        int64_t ci[3];
        for (int d = 0; d < 3; ++d) {
            const double L = (double)P->n[d] * P->w[d];
            double x = p.pos[d];
            while (x < P->origin[d]) x += L;
            while (x >= P->origin[d] + L) x -= L;
            p.pos[d] = x;
            int64_t k = (int64_t)floor((x - P->origin[d]) / P->w[d]);
            if (k < 0) k = 0;
            if (k >= P->n[d]) k = P->n[d] - 1;
            ci[d] = k;
        }
        cell_push(u, idx3(P->n, ci[0], ci[1], ci[2]), &p);
    }
    return 0;
}

int64_t or_num_particles(const or_universe* u) {
    int64_t s = 0;
    for (int64_t c = 0; c < u->ncell; ++c) s += u->count[c];
    return s;
}

void or_get_particles(const or_universe* u, or_particle* out, int64_t* cell_of) {
    int64_t k = 0;
    for (int64_t c = 0; c < u->ncell; ++c) {
        memcpy(out + k, u->cell[c], (size_t)u->count[c] * sizeof(or_particle));
        if (cell_of)
            for (int64_t i = 0; i < u->count[c]; ++i) cell_of[k + i] = c;
        k += u->count[c];
    }
}

void or_cell_counts(const or_universe* u, int64_t* counts) {
    memcpy(counts, u->count, (size_t)u->ncell * sizeof(int64_t));
}

void or_set_fields(or_universe* u, const double* E, const double* Bc) {
    if (E) memcpy(u->E, E, (size_t)u->ncell * 3 * sizeof(double));
    if (Bc) memcpy(u->Bc, Bc, (size_t)u->ncell * 3 * sizeof(double));
}

void or_get_fields(const or_universe* u, double* E, double* Bc, double* Bn, double* J) {
    const size_t sz = (size_t)u->ncell * 3 * sizeof(double);
    if (E) memcpy(E, u->E, sz);
    if (Bc) memcpy(Bc, u->Bc, sz);
    if (Bn) memcpy(Bn, u->Bn, sz);
    if (J) memcpy(J, u->J, sz);
}

void or_set_J(or_universe* u, const double* J) {
    memcpy(u->J, J, (size_t)u->ncell * 3 * sizeof(double));
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* cell.h:787-824 restated on the N-periodic node grid.  Node `pos` gathers from the 8 cells
 * pos + {-1,0}^3 (wrapped, cell.h:799-805); a particle of cell `cur` weighs
 * (i==0 ? 1-r : r) per axis with r its fractional position in `cur` (cell.h:813-817);
 * J = sum(q v fac) / vol / 8 (cell.h:822-823). */
void or_deposit(or_universe* u, int nthreads) {
    const or_props* P = &u->p;
    const int64_t* n = P->n;
    const double vol = P->w[0] * P->w[1] * P->w[2];
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t node = 0; node < u->ncell; ++node) {
        const int64_t x = node / (n[1] * n[2]), y = (node / n[2]) % n[1], z = node % n[2];
        double Js[3] = {0, 0, 0};
        for (int i = -1; i < 1; ++i)
            for (int j = -1; j < 1; ++j)
                for (int k = -1; k < 1; ++k) {
                    const int64_t cx = wrapi(x + i, n[0]), cy = wrapi(y + j, n[1]), cz = wrapi(z + k, n[2]);
                    const int64_t c = idx3(n, cx, cy, cz);
                    const double org[3] = {P->origin[0] + (double)cx * P->w[0],
                                           P->origin[1] + (double)cy * P->w[1],
                                           P->origin[2] + (double)cz * P->w[2]};
                    const or_particle* ps = u->cell[c];
                    for (int64_t q = 0; q < u->count[c]; ++q) {
                        const or_particle* p = &ps[q];
                        const double r0 = (p->pos[0] - org[0]) / P->w[0];
                        const double r1 = (p->pos[1] - org[1]) / P->w[1];
                        const double r2 = (p->pos[2] - org[2]) / P->w[2];
                        const double fac = (i == 0 ? (1 - r0) : r0) * (j == 0 ? (1 - r1) : r1) *
                                           (k == 0 ? (1 - r2) : r2);
                        for (int d = 0; d < 3; ++d) Js[d] += p->q * p->vel[d] * fac;
                    }
                }
        for (int d = 0; d < 3; ++d) u->J[3 * node + d] = (Js[d] / vol) / 8.0;
    }
}

/* field.h:356-372 (solver 0) or field.h:293-337 (solver 1, the explicit forward solver) as
 * two passes over the N-periodic grid (D2, D3), repeated
 * field_substeps times with dt/field_substeps and J held fixed, then interpC2N
 * (field.h:234-247) to refresh node B for the gather. */
void or_field_solve(or_universe* u, int nthreads) {
    const or_props* P = &u->p;
    const int64_t* n = P->n;
    const double* w = P->w;
    const double c2 = P->c * P->c;
    const int ks = P->field_substeps;  /* 0: static field (field.h:272-284) */
    const double dt = ks > 0 ? P->dt / (double)ks : 0.0;
    /* leapfrog: c^2 curl- B in the E pass, 1 x curl+ E in the B pass; forward solver
     * (field.h:308, 322): c curl B and c curl E */
    const double eF = P->solver == 1 ? P->c : c2;
    const double bF = P->solver == 1 ? P->c : 1.0;
    set_threads(nthreads);
    for (int s = 0; s < ks; ++s) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < u->ncell; ++i) {
            const int64_t x = i / (n[1] * n[2]), y = (i / n[2]) % n[1], z = i % n[2];
            const double* b = u->Bc + 3 * i;
            const double* bmx = u->Bc + 3 * pidx(n, x - 1, y, z);
            const double* bmy = u->Bc + 3 * pidx(n, x, y - 1, z);
            const double* bmz = u->Bc + 3 * pidx(n, x, y, z - 1);
            const double* j = u->J + 3 * i;
            double* e = u->E + 3 * i;
            e[0] = e[0] + dt * (eF * ((b[2] - bmy[2]) / w[1] - (b[1] - bmz[1]) / w[2]) - j[0]);
            e[1] = e[1] + dt * (eF * ((b[0] - bmz[0]) / w[2] - (b[2] - bmx[2]) / w[0]) - j[1]);
            e[2] = e[2] + dt * (eF * ((b[1] - bmx[1]) / w[0] - (b[0] - bmy[0]) / w[1]) - j[2]);
        }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < u->ncell; ++i) {
            const int64_t x = i / (n[1] * n[2]), y = (i / n[2]) % n[1], z = i % n[2];
            const double* e = u->E + 3 * i;
            const double* epx = u->E + 3 * pidx(n, x + 1, y, z);
            const double* epy = u->E + 3 * pidx(n, x, y + 1, z);
            const double* epz = u->E + 3 * pidx(n, x, y, z + 1);
            double* b = u->Bc + 3 * i;
            b[0] = b[0] - dt * bF * ((epy[2] - e[2]) / w[1] - (epz[1] - e[1]) / w[2]);
            b[1] = b[1] - dt * bF * ((epz[0] - e[0]) / w[2] - (epx[2] - e[2]) / w[0]);
            b[2] = b[2] - dt * bF * ((epx[1] - e[1]) / w[0] - (epy[0] - e[0]) / w[1]);
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < u->ncell; ++i) {
        const int64_t x = i / (n[1] * n[2]), y = (i / n[2]) % n[1], z = i % n[2];
        double acc[3] = {0, 0, 0};
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                for (int c = 0; c < 2; ++c) {
                    const double* v = u->Bc + 3 * pidx(n, x - a, y - b, z - c);
                    for (int d = 0; d < 3; ++d) acc[d] += v[d];
                }
        for (int d = 0; d < 3; ++d) u->Bn[3 * i + d] = 0.125 * acc[d];
    }
}

/* cell.h:896-964.  mover 0: the field gather re-enabled (cell.h:905-917, 932-937): corners
 * are the 8 nodes of the particle's cell, interpolated with trilinearInterpolationF2P and the
 * cell volume.  mover 1: the mover as shipped, an analytic dipole B at the particle's absolute
 * position with E = 0 (cell.h:919-945).  Adaptive sub-cycling per cell.h:948-952 (|B|^2 as
 * written there), B held over the sub-cycles. */
void or_move(or_universe* u, int nthreads) {
    const or_props* P = &u->p;
    const int64_t* n = P->n;
    const double vol = P->w[0] * P->w[1] * P->w[2];
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < u->ncell; ++c) {
        if (u->count[c] == 0) continue;
        const int64_t x = c / (n[1] * n[2]), y = (c / n[2]) % n[1], z = c % n[2];
        double Es[3][8], Bs[3][8];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) {
                    const int64_t nd = pidx(n, x + i, y + j, z + k);
                    for (int d = 0; d < 3; ++d) {
                        Es[d][(i * 2 + j) * 2 + k] = u->E[3 * nd + d];
                        Bs[d][(i * 2 + j) * 2 + k] = u->Bn[3 * nd + d];
                    }
                }
        const double org[3] = {P->origin[0] + (double)x * P->w[0], P->origin[1] + (double)y * P->w[1],
                               P->origin[2] + (double)z * P->w[2]};
        for (int64_t q = 0; q < u->count[c]; ++q) {
            or_particle* p = &u->cell[c][q];
            double rel[3], E[3], B[3];
            if (P->mover == 1) {
                /* cell.h:919-945: dipole at the coordinate origin, E = 0 */
                const double* x = p->pos;
                const double fac1 = P->dipole_moment / pow(x[0] * x[0] + x[1] * x[1] + x[2] * x[2], 2.5);
                E[0] = E[1] = E[2] = 0.0;
                B[0] = 3.0 * x[0] * x[2] * fac1;
                B[1] = 3.0 * x[1] * x[2] * fac1;
                B[2] = (2.0 * pow(x[2], 2) - pow(x[0], 2) - pow(x[1], 2)) * fac1;
            } else {
                for (int d = 0; d < 3; ++d) rel[d] = (p->pos[d] - org[d]) / P->w[d];
                for (int d = 0; d < 3; ++d) {
                    E[d] = or_trilinear(Es[d], rel, vol);
                    B[d] = or_trilinear(Bs[d], rel, vol);
                }
            }
            const double B_mag = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
            double dt_sub = M_PI * P->c / (4.0 * fabs(p->qom) * B_mag);
            int sub_cycles = (int)(P->dt / dt_sub) + 1;
            if (sub_cycles > 100) sub_cycles = 100;
            dt_sub = P->dt / (double)sub_cycles;
            for (int s = 0; s < sub_cycles; ++s) {
                or_update_velocity(p, E, B, dt_sub);
                or_update_position(p, dt_sub);
            }
        }
    }
}

/* One-cell classification of cell.h:1134-1143 (relPos vs +-halfWidth, strict), repeated
 * for moves of more than one cell (D6).  Returns the cell displacement along axis d. */
static inline int64_t classify_axis(double x, double center, double w) {
    const double hw = w / 2.0;
    double rel = x - center;
    int64_t dcell = 0;
    while (rel > hw) { rel -= w; ++dcell; }
    while (rel < -hw) { rel += w; --dcell; }
    return dcell;
}

/* cell.h:975-1159 (export) and cell.h:1195-1310 (import) with periodic boundaries (D1):
 * a particle keeps its place if it stays; leavers go to the wrapped neighbour
 * (pos + d + size) % size with their position shifted by one box length when wrapping;
 * a particle inside the sink sphere is dropped (cell.h:1123-1128).  The new content of a
 * cell is its kept particles in order followed by the imports of the 26 directions in
 * the fixed order of cell.h:1272-1306, each in its source order; displacements larger
 * than one cell (not representable in the reference) follow, by source cell order. */
int64_t or_rebin(or_universe* u, int nthreads) {
    const or_props* P = &u->p;
    const int64_t* n = P->n;
    const int64_t nc = u->ncell;
    set_threads(nthreads);
    /* per source cell: leavers grouped by direction 0..26 (13 unused) plus 27 = far */
    int64_t* lstart = (int64_t*)calloc((size_t)nc * 29, sizeof(int64_t));
    or_particle** lbuf = (or_particle**)calloc((size_t)nc, sizeof(or_particle*));
    int64_t** lfar_dest = (int64_t**)calloc((size_t)nc, sizeof(int64_t*));
    int64_t removed = 0;

#pragma omp parallel for schedule(dynamic, 64) reduction(+ : removed)
    for (int64_t c = 0; c < nc; ++c) {
        const int64_t cnt = u->count[c];
        if (cnt == 0) continue;
        const int64_t x = c / (n[1] * n[2]), y = (c / n[2]) % n[1], z = c % n[2];
        const int64_t cc[3] = {x, y, z};
        int* code = (int*)malloc((size_t)cnt * sizeof(int));
        int64_t* fdest = (int64_t*)malloc((size_t)cnt * sizeof(int64_t));
        int64_t* st = lstart + 29 * c;
        int64_t kept = 0;
        for (int64_t q = 0; q < cnt; ++q) {
            or_particle* p = &u->cell[c][q];
            int64_t dv[3];
            int far = 0;
            if (P->bc == 1) {
                /* cell.h:1103-1121: at a global face, shift the position back by half a
                 * cell width and invert the whole velocity once */
                int invert = 0;
                for (int d = 0; d < 3; ++d) {
                    const double hw = P->w[d] / 2.0;
                    const double rel = p->pos[d] - (P->origin[d] + (double)cc[d] * P->w[d] + hw);
                    if (cc[d] == 0 && rel < -hw) { p->pos[d] = p->pos[d] + hw; invert = 1; }
                    else if (cc[d] == n[d] - 1 && rel > hw) { p->pos[d] = p->pos[d] - hw; invert = 1; }
                }
                if (invert)
                    for (int d = 0; d < 3; ++d) p->vel[d] *= -1;
            }
            for (int d = 0; d < 3; ++d) {
                const double center = P->origin[d] + (double)cc[d] * P->w[d] + P->w[d] / 2.0;
                dv[d] = classify_axis(p->pos[d], center, P->w[d]);
                if (dv[d] < -1 || dv[d] > 1) far = 1;
            }
            /* periodic wrap of the destination and of the position */
            int64_t dest[3];
            for (int d = 0; d < 3; ++d) {
                const int64_t raw = cc[d] + dv[d];
                dest[d] = wrapi(raw, n[d]);
                if (dest[d] != raw) p->pos[d] += (double)(dest[d] - raw) * P->w[d];
            }
            if (P->sink) {
                double r2 = 0;
                for (int d = 0; d < 3; ++d) {
                    const double df = p->pos[d] - P->sink_center[d];
                    r2 += df * df;
                }
                if (r2 <= P->sink_radius * P->sink_radius) { code[q] = -1; ++removed; continue; }
            }
            if (dv[0] == 0 && dv[1] == 0 && dv[2] == 0) {
                code[q] = 13;
            } else if (far) {
                code[q] = 27;
                fdest[q] = idx3(n, dest[0], dest[1], dest[2]);
            } else {
                code[q] = (int)((dv[0] + 1) * 9 + (dv[1] + 1) * 3 + (dv[2] + 1));
            }
        }
        /* counting sort of leavers by direction code, stable */
        int64_t hist[28] = {0};
        for (int64_t q = 0; q < cnt; ++q)
            if (code[q] >= 0 && code[q] != 13) hist[code[q]]++;
        st[0] = 0;
        for (int k = 0; k < 28; ++k) st[k + 1] = st[k] + hist[k];
        const int64_t nleave = st[28];
        lbuf[c] = nleave ? (or_particle*)malloc((size_t)nleave * sizeof(or_particle)) : NULL;
        lfar_dest[c] = hist[27] ? (int64_t*)malloc((size_t)hist[27] * sizeof(int64_t)) : NULL;
        int64_t fill[28];
        for (int k = 0; k < 28; ++k) fill[k] = st[k];
        for (int64_t q = 0; q < cnt; ++q) {
            const int k = code[q];
            if (k < 0) continue;
            if (k == 13) { u->cell[c][kept++] = u->cell[c][q]; continue; }
            if (k == 27) lfar_dest[c][fill[27] - st[27]] = fdest[q];
            lbuf[c][fill[k]++] = u->cell[c][q];
        }
        u->count[c] = kept;
        free(code);
        free(fdest);
    }

    /* import: destination-parallel, fixed direction order (cell.h:1220-1306) */
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t c = 0; c < nc; ++c) {
        const int64_t x = c / (n[1] * n[2]), y = (c / n[2]) % n[1], z = c % n[2];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int e = 0; e < 3; ++e) {
                    if (a == 1 && b == 1 && e == 1) continue;
                    /* buffer TransferDirection(a,b,e) of cell c holds the particles of the
                     * source c + (a-1, b-1, e-1) that moved by (1-a, 1-b, 1-e) */
                    const int64_t src = pidx(n, x + a - 1, y + b - 1, z + e - 1);
                    const int k = (2 - a) * 9 + (2 - b) * 3 + (2 - e);
                    const int64_t* st = lstart + 29 * src;
                    for (int64_t q = st[k]; q < st[k + 1]; ++q) cell_push(u, c, &lbuf[src][q]);
                }
    }
    /* far movers (more than one cell along some axis), sequential by source order */
    for (int64_t c = 0; c < nc; ++c) {
        const int64_t* st = lstart + 29 * c;
        for (int64_t q = st[27]; q < st[28]; ++q) cell_push(u, lfar_dest[c][q - st[27]], &lbuf[c][q]);
    }
    for (int64_t c = 0; c < nc; ++c) { free(lbuf[c]); free(lfar_dest[c]); }
    free(lbuf);
    free(lfar_dest);
    free(lstart);
    return removed;
}

/* simulator.h:181-223 intended full loop (STEP 1a/1b deposit, STEP 2 field solve,
 * STEP 3 move, STEP 4 import) */
void or_step(or_universe* u, int64_t nsteps, int nthreads) {
    for (int64_t s = 0; s < nsteps; ++s) {
        or_deposit(u, nthreads);
        or_field_solve(u, nthreads);
        or_move(u, nthreads);
        or_rebin(u, nthreads);
    }
}

/* ------------------------------------------------------------------------------------ */
/* reductions                                                                           */
/* ------------------------------------------------------------------------------------ */

/* cell.h:1315-1324: sum 0.5 (q/qom) |v|^2 */
double or_kinetic_energy(const or_universe* u) {
    double s = 0;
    for (int64_t c = 0; c < u->ncell; ++c)
        for (int64_t q = 0; q < u->count[c]; ++q) {
            const or_particle* p = &u->cell[c][q];
            s += 0.5 * (p->q / p->qom) *
                 (p->vel[0] * p->vel[0] + p->vel[1] * p->vel[1] + p->vel[2] * p->vel[2]);
        }
    return s;
}

/* cell.h:1329-1338: sum (q/qom) |v| */
double or_momentum(const or_universe* u) {
    double s = 0;
    for (int64_t c = 0; c < u->ncell; ++c)
        for (int64_t q = 0; q < u->count[c]; ++q) {
            const or_particle* p = &u->cell[c][q];
            s += (p->q / p->qom) * sqrt(p->vel[0] * p->vel[0] + p->vel[1] * p->vel[1] + p->vel[2] * p->vel[2]);
        }
    return s;
}

double or_total_charge(const or_universe* u) {
    double s = 0;
    for (int64_t c = 0; c < u->ncell; ++c)
        for (int64_t q = 0; q < u->count[c]; ++q) s += u->cell[c][q].q;
    return s;
}

void or_field_energy(const or_universe* u, double* wE, double* wB) {
    const double vol = u->p.w[0] * u->p.w[1] * u->p.w[2];
    const double c2 = u->p.c * u->p.c;
    double se = 0, sb = 0;
    for (int64_t i = 0; i < 3 * u->ncell; ++i) {
        se += u->E[i] * u->E[i];
        sb += u->Bc[i] * u->Bc[i];
    }
    if (wE) *wE = 0.5 * se * vol;
    if (wB) *wB = 0.5 * c2 * sb * vol;
}

double or_divB(const or_universe* u, double* l2) {
    const int64_t* n = u->p.n;
    const double* w = u->p.w;
    double mx = 0, s2 = 0;
    for (int64_t i = 0; i < u->ncell; ++i) {
        const int64_t x = i / (n[1] * n[2]), y = (i / n[2]) % n[1], z = i % n[2];
        const double dv = (u->Bc[3 * pidx(n, x + 1, y, z) + 0] - u->Bc[3 * i + 0]) / w[0] +
                          (u->Bc[3 * pidx(n, x, y + 1, z) + 1] - u->Bc[3 * i + 1]) / w[1] +
                          (u->Bc[3 * pidx(n, x, y, z + 1) + 2] - u->Bc[3 * i + 2]) / w[2];
        if (fabs(dv) > mx) mx = fabs(dv);
        s2 += dv * dv;
    }
    if (l2) *l2 = sqrt(s2);
    return mx;
}

/* field.h:443-459 with vol = 0.5*wx*wy*wz and 4 pi = 16 atan(1) */
double or_ref_field_energy(const or_universe* u, int which) {
    const double vol = 0.5 * u->p.w[0] * u->p.w[1] * u->p.w[2];
    const double fourPI = 16.0 * atan(1.0);
    const double* F = which ? u->Bn : u->E;
    double s = 0;
    for (int64_t i = 0; i < u->ncell; ++i)
        s += vol * (F[3 * i] * F[3 * i] + F[3 * i + 1] * F[3 * i + 1] + F[3 * i + 2] * F[3 * i + 2]) / fourPI;
    return s;
}

/* ------------------------------------------------------------------------------------ */
/* synthetic initialisation (shared definition with the GPU init kernel)                */
/* ------------------------------------------------------------------------------------ */

/* SplitMix64 finaliser (Steele, Lea, Flood 2014) */
uint64_t or_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline double u53(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

void or_draw(uint64_t seed, int32_t species, int64_t cell, int64_t index, double u[3], double nrm[3]) {
    uint64_t key = or_mix64(or_mix64(or_mix64(seed) ^ (uint64_t)(uint32_t)species) ^ (uint64_t)cell);
    key = or_mix64(key ^ (uint64_t)index);
    double r[7];
    for (int k = 0; k < 7; ++k) r[k] = u53(or_mix64(key + (uint64_t)(k + 1) * 0xD1B54A32D192ED03ULL));
    u[0] = r[0]; u[1] = r[1]; u[2] = r[2];
    /* Box-Muller on (1 - r) in (0,1] so the logarithm is finite */
    const double twopi = 6.283185307179586;
    const double m0 = sqrt(-2.0 * log(1.0 - r[3]));
    const double m1 = sqrt(-2.0 * log(1.0 - r[5]));
    nrm[0] = m0 * cos(twopi * r[4]);
    nrm[1] = m0 * sin(twopi * r[4]);
    nrm[2] = m1 * cos(twopi * r[6]);
}

int or_init_species(or_universe* u, int32_t species, uint64_t seed, const int64_t* counts,
                    double q, double qom, const double vth[3], const double drift[3]) {
    const or_props* P = &u->p;
    const int64_t* n = P->n;
    for (int64_t c = 0; c < u->ncell; ++c) {
        const int64_t cc[3] = {c / (n[1] * n[2]), (c / n[2]) % n[1], c % n[2]};
        for (int64_t i = 0; i < counts[c]; ++i) {
            double uu[3], nn[3];
            or_draw(seed, species, c, i, uu, nn);
            or_particle p;
            for (int d = 0; d < 3; ++d) {
                p.pos[d] = P->origin[d] + ((double)cc[d] + uu[d]) * P->w[d];
                p.vel[d] = drift[d] + vth[d] * nn[d];
            }
            p.q = q;
            p.qom = qom;
            cell_push(u, c, &p);
        }
    }
    return 0;
}
