// ipic_gpu.hip -- MI355X (gfx950) particle-in-cell step behind the C ABI of include/ipic_gpu.h.
//
// Device layout (one HIP stream per context):
//   particles  per species, structure of arrays of FP32: cell-relative offsets ox,oy,oz (cell
//              units) and velocities vx,vy,vz; 24 B per particle; sorted by cell in tile-major
//              cell order (cid = tile * TC + local cell), ping-pong buffers X[0], X[1].
//   fields     E (nodes), Bc (centres), Bn (nodes), J (nodes) as float4 rows, row-major.
//   J partials per tile and block slice: the tile's nodes plus IPIC_JHALO node layers on each
//              side, (T+1+2h)^3 x 3 FP32, summed across tiles and slices in FP64 by the E pass.
//
// The particle array lags its sort key by one step: X_n is sorted by the cell each particle
// occupied when the previous push began (its "segment" a) and stores offsets relative to a,
// so its current cell is b = a + dcell(offset).  One pass over X_n (k_particles<PUSH>)
// gathers E,B at b, pushes, writes the particle into X_{n+1} sorted by b (positions known
// from the counts gathered by the previous pass), deposits J at the new cell e, and counts
// particles per e for the next pass.  Every particle is read once and written once per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <exception>
#include <new>
#include <vector>

#include "ipic_gpu.h"
#include "ipic_nop.h"

namespace {

// tuning knobs of the particle kernel (compile-time)
#ifndef IPIC_PUSH_THREADS
#define IPIC_PUSH_THREADS 256    // threads per block of the particle kernel
#endif
#ifndef IPIC_KRUN
#define IPIC_KRUN 5              // max particles per run (one lane, held in registers)
#endif
#ifndef IPIC_JHALO
#define IPIC_JHALO 0             // node layers of the tile J accumulator beyond the tile (0 or 1);
                                 // deposits beyond it go to the global overflow accumulator
#endif
#ifndef IPIC_PUSH_WAVES_PER_EU
#define IPIC_PUSH_WAVES_PER_EU 3 // register budget of the particle kernel, waves per SIMD (0: the
                                 // compiler's choice); see push_waves()
#endif

constexpr int NT = 256;          // threads per block of the scan kernels
constexpr int WS = 64;           // wavefront size
constexpr int NTP = IPIC_PUSH_THREADS;
constexpr int NW = NTP / WS;     // waves per block of the particle kernel
constexpr int KRUN = IPIC_KRUN;
constexpr int WCH = WS * KRUN;   // max particles per wave chunk
constexpr int K0CAP = 32;        // leaver buffer entries per wave chunk
constexpr int QCAP = WCH + WS;   // arrival queue entries per wave (fewer than WS carried over + one chunk)
constexpr int TCMAX = 512;       // max cells per tile
constexpr int NJMAX = 11 * 11 * 11;  // max J nodes per tile ((T+1+2h)^3, T <= 8, h <= 1)
constexpr int JPMAX = 4;         // max (tile, local) pairs per axis coordinate

enum { MODE_PUSH = 0, MODE_DEPOSIT = 1, MODE_REBIN = 2 };

struct Geo {
    int n[3];
    int T[3];
    int NTl[3];
    int TC;
    int ntiles;
    int nsl;              // blocks per tile of the particle kernel (slices of the tile's cells)
    int ncell;
    int NJ;
    int nj[3];            // J accumulator nodes per axis: T + 1 + 2 jh
    int jh;               // its halo (node layers beyond the tile on each side)
    int nmax;
    int s0;               // n[1] * n[2], row stride of x
    float w[3];
    float inv_w[3];
    float origin[3];
    float dt;
    float inv_vol;
    float pi_c;
    int sink;
    int reflect;          // 1: reflect at the global faces (cell.h:1103-1121) instead of wrapping
    float dip_m;          // -B1z R^3 / L^3 of the mover as shipped (analytic dipole at the origin)
    float dip_o[3];       // origin / L and cell width / L: the dipole is evaluated in units of
    float dip_w[3];       // L = cell_width[0], as |x|^5 of a box in metres overflows FP32
    float sink_c[3];
    float sink_r2;
    float E0[3];          // uniform part of the background field (added in the gather)
    float B0[3];
    float inv_T[3];       // 1 / T (exact small-integer division by a float reciprocal)
    float inv_T12;        // 1 / (T[1] * T[2])
    float fix_lim;        // 2^(50 - fix_bits): J in fixed point with fix_bits fractional bits
    double fix_scale;     // 2^fix_bits
    double fix_inv;       // 2^-fix_bits
};

struct SpArgs {
    const float* in;      // input arrays, `cap` floats each, back to back: offsets ox oy oz
                          // (cell units), velocities vx vy vz
    float* out;           // output arrays, same layout
    const uint32_t* id_in;
    uint32_t* id_out;
    long long cap;
    const int* seg;       // segment starts of `in`, ncell+1
    const int* start;     // output starts by current cell, ncell+1
    const int* stay;      // particles of segment c that are in cell c
    int* arr;             // arrival counters per cell
    int* cnt_acc;         // next-step counts by new cell
    int* stay_acc;        // next-step stayers by output segment
    float qom;
    float qs;             // q / (8 vol)
    float subk;           // 4 |qom| dt / (pi c): sub-cycle ratio per unit |B|^2
};

struct PushArgs {
    Geo g;
    SpArgs sp[IPIC_MAX_SPECIES];
    int nsp;
    const float4* E;
    const float4* Bn;
    float* jpart;
    long long* jover;     // overflow deposits, 64-bit fixed point like the tile accumulator
    unsigned char* jmark; // jmark[i] = 1: node i holds overflow deposits
    unsigned long long* removed;
    int* err;             // set to nonzero when a consistency guard fires
    const int* tile_order;  // block -> tile (a permutation), or null for the XCD-aware default
};

// ---------------------------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------------------------

// One-cell rule of cell.h:1134-1143 in offset form (a particle stays while its offset lies in
// the closed interval [0,1]), repeated for moves of more than one cell.
__device__ __forceinline__ int dcell(float o) {
    if (o > 1.0f) return (int)fminf(ceilf(o - 1.0f), 1048576.0f);
    if (o < 0.0f) return (int)fmaxf(floorf(o), -1048576.0f);
    return 0;
}

__device__ __forceinline__ int wrapc(int i, int n) {
    if ((unsigned)i < (unsigned)n) return i;
    if (i < 0 && i >= -n) return i + n;
    if (i >= n && i < 2 * n) return i - n;
    const int r = i % n;
    return r < 0 ? r + n : r;
}

// floor(a / b) for 0 <= a < 2^21 and b >= 1 through a reciprocal inv_b within one ulp of 1/b:
// the quotient is offset by 0.5/b, more than the error (a + 0.5) * 2^-22 / b of the product.
__device__ __forceinline__ int fdivi(int a, float inv_b) { return (int)(((float)a + 0.5f) * inv_b); }

__device__ __forceinline__ int cid_of(const Geo& g, int x, int y, int z) {
    const int tx = fdivi(x, g.inv_T[0]), ty = fdivi(y, g.inv_T[1]), tz = fdivi(z, g.inv_T[2]);
    const int lx = x - tx * g.T[0], ly = y - ty * g.T[1], lz = z - tz * g.T[2];
    return ((tx * g.NTl[1] + ty) * g.NTl[2] + tz) * g.TC + (lx * g.T[1] + ly) * g.T[2] + lz;
}

__device__ __forceinline__ int nid_of(const Geo& g, int x, int y, int z) {
    return (x * g.n[1] + y) * g.n[2] + z;
}

// XCD-aware block -> tile map: consecutive tiles (neighbours in x-slabs) go to the same XCD
// (blocks are dealt round-robin over the 8 XCDs), so halo nodes shared by neighbouring tiles
// stay in one L2.
__device__ __forceinline__ int xcd_tile(int bid, int ntiles) {
    const int per = ntiles / 8, rem = ntiles % 8;
    const int xcd = bid % 8, slot = bid / 8;
    return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

// Inclusive prefix sum over the wave with DPP lane moves (no LDS round trips): shifts by 1, 2,
// 4, 8 within each row of 16 lanes (lanes shifted in from outside the row read 0), then lane
// 15's sum added to row 1 and row 3, then lane 31's sum added to rows 2 and 3.
__device__ __forceinline__ int wave_incl_scan(int x, int) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return x;
}

// Inclusive prefix maximum over the wave of non-negative values, with the same lane moves.
__device__ __forceinline__ int wave_incl_max(int x) {
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true));  // row_shr:1
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true));  // row_shr:2
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true));  // row_shr:4
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true));  // row_shr:8
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));  // row_bcast:15 into rows 1, 3
    x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));  // row_bcast:31 into rows 2, 3
    return x;
}

// Exclusive prefix sum over the block (NT threads); returns the prefix, sets total.
__device__ __forceinline__ int block_exscan(int v, int* s_w, int& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int x = wave_incl_scan(v, lane);
    if (lane == 63) s_w[wid] = x;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        const int sw = s_w[w];
        off += (w < wid) ? sw : 0;
        tot += sw;
    }
    total = tot;
    __syncthreads();
    return off + x - v;
}

// J is accumulated in LDS as 64-bit fixed point (fix_bits fractional bits): integer atomics
// are fast where float atomics on shared addresses serialise, and the sum is exact and
// independent of the order of the adds.  The conversion rounds v * 2^fix_bits to the nearest
// integer x in one FP64 fused multiply-add with 1.5 * 2^52, whose unit in the last place is 1:
// for |x| < 2^51 the bit pattern of the result is bits(1.5 * 2^52) + x, and as the constant's
// low word is zero, x is that pattern with the constant's high word subtracted.  The callers
// keep |v| <= fix_lim = 2^(50 - fix_bits) (|x| <= 2^50) or raise ERR_JSAT: every deposited
// value is a sum of at most run_in terms q v w with weights in [0, 1], checked against the
// run's largest |q v| component, and velocities are finite (a non-finite position removes the
// particle; uploads refuse non-finite values).
__device__ __forceinline__ long long fix_of(float v, double scale) {
    const double d = fma((double)v, scale, 6755399441055744.0);
    return (long long)((unsigned long long)__double_as_longlong(d) - 0x4338000000000000ull);
}

__device__ __forceinline__ void fix_add(long long* p, float v, double scale) {
    __hip_atomic_fetch_add(p, fix_of(v, scale), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// device error bits (the host turns them into an error of the next synchronising call)
enum { ERR_SLOT = 1, ERR_CHUNK = 2, ERR_JSAT = 4 };

// Trilinear weights (cell.h:876-880): w[(i*2+j)*2+k] = (i?x:1-x)(j?y:1-y)(k?z:1-z)
__device__ __forceinline__ void tri_weights(float x, float y, float z, float wt[8]) {
    const float x0 = 1.0f - x, y0 = 1.0f - y, z0 = 1.0f - z;
    const float a00 = x0 * y0, a01 = x0 * y, a10 = x * y0, a11 = x * y;
    wt[0] = a00 * z0; wt[1] = a00 * z;
    wt[2] = a01 * z0; wt[3] = a01 * z;
    wt[4] = a10 * z0; wt[5] = a10 * z;
    wt[6] = a11 * z0; wt[7] = a11 * z;
}

constexpr int JH = IPIC_JHALO;

// Tile geometry of the particle kernel, from Geo at run time or, for 8x8x8 tiles, as
// compile-time constants (cell splits and node indices fold into shifts and offsets).
template <bool FIX8>
struct TileGeo {
    int T0, T1, T2, TC, nj1, nj2, NJ;
    int nj0;
    float iT2, iT12;
    __device__ __forceinline__ explicit TileGeo(const Geo& g)
        : T0(g.T[0]), T1(g.T[1]), T2(g.T[2]), TC(g.TC), nj1(g.nj[1]), nj2(g.nj[2]), NJ(g.NJ), nj0(g.nj[0]),
          iT2(g.inv_T[2]), iT12(g.inv_T12) {}
    // tile-local cell index -> coordinates
    __device__ __forceinline__ void split(int c, int& lx, int& ly, int& lz) const {
        lx = fdivi(c, iT12);
        const int r = c - lx * T1 * T2;
        ly = fdivi(r, iT2);
        lz = r - ly * T2;
    }
    // wrapped global cell -> tile-major cell id
    __device__ __forceinline__ int cid(const Geo& g, int x, int y, int z) const { return cid_of(g, x, y, z); }
};

template <>
struct TileGeo<true> {
    static constexpr int T0 = 8, T1 = 8, T2 = 8, TC = 512;
    static constexpr int nj0 = 9 + 2 * JH, nj1 = nj0, nj2 = nj0, NJ = nj0 * nj0 * nj0;
    __device__ __forceinline__ explicit TileGeo(const Geo&) {}
    __device__ __forceinline__ void split(int c, int& lx, int& ly, int& lz) const {
        lx = c >> 6;
        ly = (c >> 3) & 7;
        lz = c & 7;
    }
    __device__ __forceinline__ int cid(const Geo& g, int x, int y, int z) const {
        return (((x >> 3) * g.NTl[1] + (y >> 3)) * g.NTl[2] + (z >> 3)) * 512 + ((x & 7) * 8 + (y & 7)) * 8 + (z & 7);
    }
};

// Deposit of one particle outside the run's own cell: corner (i,j,k) of cell e goes to the
// tile's LDS accumulator when inside its halo, else to the global overflow accumulator.
template <class TG>
__device__ __forceinline__ void deposit_direct(const PushArgs& A, const TG& G, long long* sJ, int x0, int y0,
                                               int z0, int elx, int ely, int elz, const float wt[8], float jx,
                                               float jy, float jz) {
    const Geo& g = A.g;
    // every corner value is at most max |j| (the weights lie in [0, 1])
    if (!(fmaxf(fabsf(jx), fmaxf(fabsf(jy), fabsf(jz))) <= g.fix_lim)) atomicOr(A.err, ERR_JSAT);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const float ww = wt[(i * 2 + j) * 2 + k];
                const int nx = elx + JH + i, ny = ely + JH + j, nz = elz + JH + k;
                if (nx >= 0 && nx < G.nj0 && ny >= 0 && ny < G.nj1 && nz >= 0 && nz < G.nj2) {
                    const int l = (nx * G.nj1 + ny) * G.nj2 + nz;
                    fix_add(&sJ[l], jx * ww, g.fix_scale);
                    fix_add(&sJ[G.NJ + l], jy * ww, g.fix_scale);
                    fix_add(&sJ[2 * G.NJ + l], jz * ww, g.fix_scale);
                } else {
                    // beyond the tile: the global accumulator, in the same fixed point, so J does
                    // not depend on the order of these adds either
                    const int gid = nid_of(g, wrapc(x0 + elx + i, g.n[0]), wrapc(y0 + ely + j, g.n[1]),
                                           wrapc(z0 + elz + k, g.n[2]));
                    long long* o = A.jover + 3 * (long long)gid;
                    // every writer stores the same value: a relaxed atomic byte store
                    __hip_atomic_store(&A.jmark[gid], (unsigned char)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_add(o + 0, fix_of(jx * ww, g.fix_scale), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_add(o + 1, fix_of(jy * ww, g.fix_scale), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_add(o + 2, fix_of(jz * ww, g.fix_scale), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
            }
}

template <class TG>
__device__ __forceinline__ void count_add(const Geo& g, const TG& G, int* s_arr, int* gl, int x0, int y0, int z0,
                                          int ex, int ey, int ez) {
    const int lx = ex - x0, ly = ey - y0, lz = ez - z0;
    if (lx >= 0 && lx < G.T0 && ly >= 0 && ly < G.T1 && lz >= 0 && lz < G.T2) {
        atomicAdd(&s_arr[(lx * G.T1 + ly) * G.T2 + lz], 1);
    } else {
        atomicAdd(&gl[G.cid(g, wrapc(ex, g.n[0]), wrapc(ey, g.n[1]), wrapc(ez, g.n[2]))], 1);
    }
}

// ---------------------------------------------------------------------------------------------
// particle kernel: PUSH (gather + Boris + rebin + deposit), DEPOSIT (J only), REBIN (sort only)
//
// One block per tile of cells (or per slice of a tile's cells on grids of few tiles).  The
// block's cells are split among its waves by particle count and every wave streams its cells on its own: it cuts them into runs of at most KRUN
// particles of one cell, one run per lane, and each lane loads its run straight into registers
// (all its loads in flight at once).  The fast path covers particles that start and end the
// step in the run's cell, in two passes over the run registers: the push (gather from the run
// cell's 8 corner nodes, held in registers) updates them in place, then the deposit accumulates
// per-run moments in registers and the updated particle goes directly to its slot of the output
// array.  The corner nodes and the moments are never live together (168 VGPRs: 3 waves per
// SIMD).  Every other particle (it sits in another cell, or leaves the
// cell) is appended to a wave queue by ballot compaction and handled afterwards with all lanes
// busy.  The tile's J accumulator (64-bit fixed point) and per-cell counters live in LDS.
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
    for (int d = WS / 2; d > 0; d >>= 1) x = max(x, __shfl_xor(x, d, WS));
    return x;
}

__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// xyz of a float4 node value (a 12-byte load: the fourth component is padding)
__device__ __forceinline__ float3 ld3(const float4* p) {
    struct F3 { float x, y, z; };
    const F3 v = *reinterpret_cast<const F3*>(p);
    return make_float3(v.x, v.y, v.z);
}

__device__ __forceinline__ int ballot_rank(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Trilinear interpolation written as the polynomial it expands to:
// f(x,y,z) = c0 + c1 x + c2 y + c3 z + c4 xy + c5 xz + c6 yz + c7 xyz for corner values
// f[(i*2+j)*2+k] with the weights (i?x:1-x)(j?y:1-y)(k?z:1-z) of cell.h:876-880.
// `add` is added to every corner value and the result is multiplied by `scale`.
__device__ __forceinline__ void tri_coef(const float f[8], float add, float scale, float c[8]) {
    c[0] = (f[0] + add) * scale;
    c[1] = (f[4] - f[0]) * scale;
    c[2] = (f[2] - f[0]) * scale;
    c[3] = (f[1] - f[0]) * scale;
    c[4] = ((f[6] - f[4]) - (f[2] - f[0])) * scale;
    c[5] = ((f[5] - f[4]) - (f[1] - f[0])) * scale;
    c[6] = ((f[3] - f[2]) - (f[1] - f[0])) * scale;
    c[7] = (((f[7] - f[6]) - (f[5] - f[4])) - ((f[3] - f[2]) - (f[1] - f[0]))) * scale;
}

__device__ __forceinline__ float tri_eval(const float c[8], float x, float y, float z) {
    const float a = fmaf(c[1], x, c[0]);
    const float b = fmaf(c[4], x, c[2]);
    const float d = fmaf(c[5], x, c[3]);
    const float e = fmaf(c[7], x, c[6]);
    return fmaf(z, fmaf(e, y, d), fmaf(b, y, a));
}

// Moments M = sum j * {1, x, y, z, xy, xz, yz, xyz} to the 8 corner sums
// N[(i*2+j)*2+k] = sum j * (i?x:1-x)(j?y:1-y)(k?z:1-z), one axis at a time.
__device__ __forceinline__ void mom_to_nodes(const float M[8], float N[8]) {
    const float a0_1 = M[0] - M[1], a1_1 = M[1];
    const float a0_y = M[2] - M[4], a1_y = M[4];
    const float a0_z = M[3] - M[5], a1_z = M[5];
    const float a0_yz = M[6] - M[7], a1_yz = M[7];
    const float b00_1 = a0_1 - a0_y, b01_1 = a0_y, b00_z = a0_z - a0_yz, b01_z = a0_yz;
    const float b10_1 = a1_1 - a1_y, b11_1 = a1_y, b10_z = a1_z - a1_yz, b11_z = a1_yz;
    N[0] = b00_1 - b00_z; N[1] = b00_z;
    N[2] = b01_1 - b01_z; N[3] = b01_z;
    N[4] = b10_1 - b10_z; N[5] = b10_z;
    N[6] = b11_1 - b11_z; N[7] = b11_z;
}

// Adaptive sub-cycling (cell.h:948-952: n = min(100, int(dt/dt_sub)+1) with
// dt_sub = pi c / (4 |qom| |B|^2)), Boris rotation (particle.h:31-49) and position update
// (particle.h:24-26) in cell units.
__device__ __forceinline__ void boris_move(const Geo& g, float qom, float subk, float Ex, float Ey, float Ez,
                                           float Bx, float By, float Bz, float& vx, float& vy, float& vz,
                                           float& px, float& py, float& pz) {
    const float ratio = (Bx * Bx + By * By + Bz * Bz) * subk;
    const int nsub = (ratio < 99.0f) ? (int)ratio + 1 : 100;
    const float dts = (nsub == 1) ? g.dt : g.dt / (float)nsub;
    const float kq = qom * 0.5f * dts;
    const float sx = dts * g.inv_w[0], sy = dts * g.inv_w[1], sz = dts * g.inv_w[2];
    const float tx = kq * Bx, ty = kq * By, tz = kq * Bz;
    const float f = 2.0f * __builtin_amdgcn_rcpf(1.0f + (tx * tx + ty * ty + tz * tz));
    const float ux = f * tx, uy = f * ty, uz = f * tz;
    const float kex = kq * Ex, key = kq * Ey, kez = kq * Ez;
    auto sub = [&]() {
        const float mx = vx + kex, my = vy + key, mz = vz + kez;
        const float qx = mx + (my * tz - mz * ty);
        const float qy = my + (mz * tx - mx * tz);
        const float qz = mz + (mx * ty - my * tx);
        vx = mx + (qy * uz - qz * uy) + kex;
        vy = my + (qz * ux - qx * uz) + key;
        vz = mz + (qx * uy - qy * ux) + kez;
        px += vx * sx;
        py += vy * sy;
        pz += vz * sz;
    };
    sub();  // the one sub-cycle of the common case outside the loop
    for (int sc = 1; sc < nsub; ++sc) sub();
}

// Reflection at the global faces (cell.h:1103-1121) for a particle of cell (cx,cy,cz) at offsets
// (px,py,pz): past the low face of cell 0 (offset < 0) it moves half a cell back (+0.5), past
// the high face of cell n-1 (offset > 1) half a cell back (-0.5); either inverts the whole
// velocity once.
__device__ __forceinline__ void reflect_faces(const Geo& g, int cx, int cy, int cz, float& px, float& py, float& pz,
                                              float& vx, float& vy, float& vz) {
    bool inv = false;
    if (cx == 0 && px < 0.0f) { px += 0.5f; inv = true; } else if (cx == g.n[0] - 1 && px > 1.0f) { px -= 0.5f; inv = true; }
    if (cy == 0 && py < 0.0f) { py += 0.5f; inv = true; } else if (cy == g.n[1] - 1 && py > 1.0f) { py -= 0.5f; inv = true; }
    if (cz == 0 && pz < 0.0f) { pz += 0.5f; inv = true; } else if (cz == g.n[2] - 1 && pz > 1.0f) { pz -= 0.5f; inv = true; }
    if (inv) {
        vx = -vx; vy = -vy; vz = -vz;
    }
}

// The mover as shipped (cell.h:919-945): E = 0 and the field of a z dipole at the coordinate
// origin, B = m (3xz, 3yz, 2z^2 - x^2 - y^2) / |x|^5 with m = -B1z R^3, at the absolute
// position of offsets (px,py,pz) in cell (cx,cy,cz).  Lengths are in units of L (x' = x / L,
// m' = m / L^3), which leaves B unchanged and keeps |x'|^5 in range.
__device__ __forceinline__ void dipole_field(const Geo& g, int cx, int cy, int cz, float px, float py, float pz,
                                             float& Bx, float& By, float& Bz) {
    const float x = g.dip_o[0] + ((float)cx + px) * g.dip_w[0];
    const float y = g.dip_o[1] + ((float)cy + py) * g.dip_w[1];
    const float z = g.dip_o[2] + ((float)cz + pz) * g.dip_w[2];
    const float r2 = x * x + y * y + z * z;
    const float fac = g.dip_m / (r2 * r2 * sqrtf(r2));
    Bx = 3.0f * x * z * fac;
    By = 3.0f * y * z * fac;
    Bz = (2.0f * z * z - x * x - y * y) * fac;
}

// E and B (gather arrays plus the uniform background, divided by the cell volume) at offsets
// (x,y,z) of the cell with wrapped coordinates (cx,cy,cz), read straight from the node arrays.
__device__ __forceinline__ void gather_direct(const PushArgs& A, int cx, int cy, int cz, float x, float y,
                                              float z, float& Ex, float& Ey, float& Ez, float& Bx, float& By,
                                              float& Bz) {
    const Geo& g = A.g;
    const int xs[2] = {cx * g.s0, wrap1(cx + 1, g.n[0]) * g.s0};
    const int ys[2] = {cy * g.n[2], wrap1(cy + 1, g.n[1]) * g.n[2]};
    const int zs[2] = {cz, wrap1(cz + 1, g.n[2])};
    float wt[8];
    tri_weights(x, y, z, wt);
    Ex = Ey = Ez = Bx = By = Bz = 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int nd = xs[i] + ys[j] + zs[k];
                const float4 e4 = A.E[nd], b4 = A.Bn[nd];
                const float ww = wt[(i * 2 + j) * 2 + k];
                Ex += e4.x * ww; Ey += e4.y * ww; Ez += e4.z * ww;
                Bx += b4.x * ww; By += b4.y * ww; Bz += b4.z * ww;
            }
    Ex = (Ex + g.E0[0]) * g.inv_vol; Ey = (Ey + g.E0[1]) * g.inv_vol; Ez = (Ez + g.E0[2]) * g.inv_vol;
    Bx = (Bx + g.B0[0]) * g.inv_vol; By = (By + g.B0[1]) * g.inv_vol; Bz = (Bz + g.B0[2]) * g.inv_vol;
}

// A particle whose position is not finite after the push is removed like a sunk one: a slot
// with a NaN offset marks a removed particle in the next pass.
__device__ __forceinline__ bool finite3(float x, float y, float z) {
    return __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
}

__device__ __forceinline__ bool alive(float x, float y, float z) { return x == x && y == y && z == z; }

// 0 <= x, y, z <= 1 on the bit patterns: a non-negative float orders like its pattern as an
// unsigned integer, and a negative one (sign bit set) or a NaN lies above the pattern of 1.0.
// -0.0 counts as outside; offsets are never -0.0 (uploads store +0.0 for it, and the push only
// adds to offsets, which cannot produce -0.0 from a +0.0 or non-zero start).
__device__ __forceinline__ bool in_cell(float x, float y, float z) {
    const unsigned m = max(max(__float_as_uint(x), __float_as_uint(y)), __float_as_uint(z));
    return m <= 0x3f800000u;
}

// Particle arrays are read and written through buffer resources: one resource spans the six
// arrays of a species (offset from the array start in `voff`, array d at byte d * cs), so a
// lane keeps one 32-bit offset for all six and an out-of-range access reads 0 / is dropped.
typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc make_rsrc(const void* p, long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)(unsigned)bytes, 0x00020000);
}

__device__ __forceinline__ float ldp(Rsrc r, unsigned voff, unsigned soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

__device__ __forceinline__ void stp(Rsrc r, float v, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}

// v[0..3] to four consecutive floats in one dwordx4 store
template <int N>
__device__ __forceinline__ void st4(Rsrc r, unsigned voff, unsigned soff, const float (&v)[N]) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 q = {__float_as_uint(v[0]), __float_as_uint(v[1 % N]), __float_as_uint(v[2 % N]), __float_as_uint(v[3 % N])};
    __builtin_amdgcn_raw_buffer_store_b128(q, r, voff, soff, 0);
}

__device__ __forceinline__ void store_particle(Rsrc r, unsigned cs, int oidx, float px, float py, float pz,
                                               float vx, float vy, float vz) {
    const unsigned vo = (unsigned)oidx * 4u;
    stp(r, px, vo, 0);
    stp(r, py, vo, cs);
    stp(r, pz, vo, 2 * cs);
    stp(r, vx, vo, 3 * cs);
    stp(r, vy, vo, 4 * cs);
    stp(r, vz, vo, 5 * cs);
}

// E and B at offsets (x,y,z) of a cell from the tile's node field in LDS: F4[n] = (Ex,Ey,Ez,Bx),
// F2[n] = (By,Bz) at node n, background included and divided by the cell volume; nb is the
// cell's (0,0,0) corner, sx and sy the node strides of x and y.
__device__ __forceinline__ void gather_lds(const float4* F4, const float2* F2, int nb, int sx, int sy, float x,
                                           float y, float z, float& Ex, float& Ey, float& Ez, float& Bx,
                                           float& By, float& Bz) {
    float wt[8];
    tri_weights(x, y, z, wt);
    Ex = Ey = Ez = Bx = By = Bz = 0.0f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int n = nb + i * sx + j * sy + k;
                const float4 a = F4[n];
                const float2 b = F2[n];
                const float w = wt[(i * 2 + j) * 2 + k];
                Ex = fmaf(a.x, w, Ex); Ey = fmaf(a.y, w, Ey); Ez = fmaf(a.z, w, Ez);
                Bx = fmaf(a.w, w, Bx); By = fmaf(b.x, w, By); Bz = fmaf(b.y, w, Bz);
            }
}

// The same interpolation from the 8 corner values held in registers (corner (i,j,k) at
// index (i*2+j)*2+k).
__device__ __forceinline__ void gather_regs(const float4 n4[8], const float2 n2[8], float x, float y, float z,
                                            float& Ex, float& Ey, float& Ez, float& Bx, float& By, float& Bz) {
    float wt[8];
    tri_weights(x, y, z, wt);
    Ex = Ey = Ez = Bx = By = Bz = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        Ex = fmaf(n4[q].x, wt[q], Ex); Ey = fmaf(n4[q].y, wt[q], Ey); Ez = fmaf(n4[q].z, wt[q], Ez);
        Bx = fmaf(n4[q].w, wt[q], Bx); By = fmaf(n2[q].x, wt[q], By); Bz = fmaf(n2[q].y, wt[q], Bz);
    }
}

// Dynamic LDS layout of k_particles (bytes, 16-aligned sections)
struct PLayout {
    int off_f4, off_f2, off_seg, off_cnt, off_stay, off_wlo, off_wave, wave_bytes, total;
};

__host__ __device__ inline PLayout particle_layout(int NJ, int TC, int NF) {
    PLayout L;
    int o = 3 * NJ * 8;             // J, int64
    L.off_f4 = o; o += NF * 16;     // node field (Ex,Ey,Ez,Bx)
    L.off_f2 = o; o += NF * 8;      // node field (By,Bz)
    L.off_seg = o; o += (TC + 1) * 4;
    L.off_cnt = o; o += TC * 4;
    L.off_stay = o; o += TC * 4;
    L.off_wlo = o; o += (NW + 1) * 4;
    o = (o + 15) & ~15;
    L.off_wave = o;
    // per wave: run table (WS+1 ints), run heads (WS ints), arrival queue (QCAP ints), leaver
    // buffer (K0CAP x 8 floats)
    L.wave_bytes = ((WS + 1) * 4 + WS * 4 + QCAP * 4 + 15) & ~15;
    L.wave_bytes += K0CAP * 8 * 4;
    o += NW * L.wave_bytes;
    L.total = o;
    return L;
}

__host__ __device__ inline int field_nodes(const Geo& g) { return (g.T[0] + 1) * (g.T[1] + 1) * (g.T[2] + 1); }

// option bits of the particle kernel
enum { OPT_FIX8 = 1, OPT_SINK = 2, OPT_REFLECT = 4, OPT_DIPOLE = 8 };

// Register budget (waves per SIMD) of a particle-kernel instance: every instance fits 168 VGPRs
// (3 waves per SIMD) with runs of 5.
constexpr int push_waves(bool, int) { return IPIC_PUSH_WAVES_PER_EU; }

#if IPIC_PUSH_WAVES_PER_EU > 1
#define IPIC_PUSH_ATTR __attribute__((amdgpu_waves_per_eu(push_waves(TRACK, OPT))))
#else
#define IPIC_PUSH_ATTR
#endif

template <int MODE, bool TRACK, int OPT>
__global__ __launch_bounds__(NTP) IPIC_PUSH_ATTR void k_particles(PushArgs A) {
    const Geo& g = A.g;
    using TG = TileGeo<(OPT & OPT_FIX8) != 0>;
    const TG G(g);
    constexpr bool SINK = (OPT & OPT_SINK) != 0, REFLECT = (OPT & OPT_REFLECT) != 0;
    constexpr bool DIPOLE = (OPT & OPT_DIPOLE) != 0;  // analytic dipole B, E = 0: no field staging
    extern __shared__ __align__(16) unsigned char smem[];
    const PLayout L =
        particle_layout(G.NJ, G.TC, (MODE == MODE_PUSH && !DIPOLE) ? (G.T0 + 1) * (G.T1 + 1) * (G.T2 + 1) : 0);
    long long* sJ = (long long*)smem;
    float4* sF4 = (float4*)(smem + L.off_f4);
    float2* sF2 = (float2*)(smem + L.off_f2);
    int* s_seg = (int*)(smem + L.off_seg);
    int* s_cnt = (int*)(smem + L.off_cnt);
    int* s_stay = (int*)(smem + L.off_stay);
    int* s_wlo = (int*)(smem + L.off_wlo);

    const int tid = threadIdx.x, lane = tid & (WS - 1), wid = tid / WS;
    unsigned char* wbase = smem + L.off_wave + wid * L.wave_bytes;
    int* R = (int*)wbase;                                            // WS+1 ints
    int* H = (int*)(wbase + (WS + 1) * 4);                           // WS ints
    int* Q = (int*)(wbase + (2 * WS + 1) * 4);                       // arrival queue
    float* K0 = (float*)(wbase + (((2 * WS + 1) * 4 + QCAP * 4 + 15) & ~15));  // leaver buffer

    const int t = A.tile_order ? A.tile_order[blockIdx.x] : xcd_tile(blockIdx.x, g.ntiles);
    const int sl = blockIdx.y;  // this block's slice of the tile's cells
    const int tx = t / (g.NTl[1] * g.NTl[2]), ty = (t / g.NTl[2]) % g.NTl[1], tz = t % g.NTl[2];
    const int x0 = tx * G.T0, y0 = ty * G.T1, z0 = tz * G.T2;
    const int cbase = t * G.TC;
    const float NaNf = __builtin_nanf("");
    const int fsy = G.T2 + 1, fsx = (G.T1 + 1) * fsy;  // node strides of the LDS field

    if (MODE != MODE_REBIN)
        for (int i = tid; i < 3 * G.NJ; i += NTP) sJ[i] = 0;
    if (MODE == MODE_PUSH && !DIPOLE) {
        // the tile's (T+1)^3 nodes of E and B, background added, divided by the cell volume
        const int nf = (G.T0 + 1) * fsx;
        for (int i = tid; i < nf; i += NTP) {
            const int ix = i / fsx, r = i - ix * fsx, iy = r / fsy, iz = r - iy * fsy;
            const int nd = wrap1(x0 + ix, g.n[0]) * g.s0 + wrap1(y0 + iy, g.n[1]) * g.n[2] + wrap1(z0 + iz, g.n[2]);
            const float3 e = ld3(A.E + nd), b = ld3(A.Bn + nd);
            const float iv = g.inv_vol;
            sF4[i] = make_float4((e.x + g.E0[0]) * iv, (e.y + g.E0[1]) * iv, (e.z + g.E0[2]) * iv, (b.x + g.B0[0]) * iv);
            sF2[i] = make_float2((b.y + g.B0[1]) * iv, (b.z + g.B0[2]) * iv);
        }
    }

    for (int s = 0; s < A.nsp; ++s) {
        const SpArgs& S = A.sp[s];
        const Rsrc rin = make_rsrc(S.in, S.cap * 24);
        const Rsrc rout = make_rsrc(S.out, S.cap * 24);
        const unsigned cs = (unsigned)S.cap * 4u;  // byte stride between the six arrays
        __syncthreads();
        for (int i = tid; i <= G.TC; i += NTP) s_seg[i] = S.seg[cbase + i];
        for (int i = tid; i < G.TC; i += NTP) {
            s_cnt[i] = 0;
            s_stay[i] = 0;
        }
        __syncthreads();
        // the block's cells [clo, chi): slice sl of g.nsl at equal particle counts; cells of
        // wave w: [s_wlo[w], s_wlo[w+1]), the slice split again at equal particle counts
        if (tid <= NW) {
            auto first_ge = [&](long long target) {  // first cell c with s_seg[c] >= target
                int lo = 0, hi = G.TC;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_seg[mid] >= target) hi = mid; else lo = mid + 1;
                }
                return lo;
            };
            const long long tot = (long long)s_seg[G.TC] - s_seg[0];
            const int clo = (sl == 0) ? 0 : first_ge(s_seg[0] + tot * sl / g.nsl);
            const int chi = (sl == g.nsl - 1) ? G.TC : first_ge(s_seg[0] + tot * (sl + 1) / g.nsl);
            const long long sub = (long long)s_seg[chi] - s_seg[clo];
            s_wlo[tid] = (tid == 0) ? clo
                       : (tid == NW ? chi : max(clo, min(chi, first_ge(s_seg[clo] + sub * tid / NW))));
        }
        __syncthreads();
        const int chi = s_wlo[wid + 1];
        int c_cur = s_wlo[wid];
        int pos_cur = s_seg[c_cur];
        int carry = 0;  // stayers already placed for cell c_cur (when it continues)
        const float qom = S.qom, qs = S.qs, subk = S.subk;

        // A particle sitting in another cell b (an arrival of the lagged sort, input index inq):
        // full processing by one lane.  Its input cell is the last tile cell whose segment
        // starts at or below inq.
        auto arrive = [&](int inq) {
            int kc = 0, khi = G.TC - 1;
            while (kc < khi) {
                const int mid = (kc + khi + 1) >> 1;
                if (s_seg[mid] <= inq) kc = mid; else khi = mid - 1;
            }
            int qlx, qly, qlz;
            G.split(kc, qlx, qly, qlz);
            const unsigned vq = (unsigned)inq * 4u;
            const float ox = ldp(rin, vq, 0), oy = ldp(rin, vq, cs), oz = ldp(rin, vq, 2 * cs);
            float vx = ldp(rin, vq, 3 * cs), vy = ldp(rin, vq, 4 * cs), vz = ldp(rin, vq, 5 * cs);
            const int dbx = dcell(ox), dby = dcell(oy), dbz = dcell(oz);
            float px = ox - (float)dbx, py = oy - (float)dby, pz = oz - (float)dbz;
            const int blx = qlx + dbx, bly = qly + dby, blz = qlz + dbz;  // tile-local cell b
            const int bx = x0 + blx, by = y0 + bly, bz = z0 + blz;        // unwrapped
            const int bxw = wrapc(bx, g.n[0]), byw = wrapc(by, g.n[1]), bzw = wrapc(bz, g.n[2]);
            bool sunk = false;
            int dex = 0, dey = 0, dez = 0;
            if (MODE == MODE_PUSH) {
                float Ex, Ey, Ez, Bx, By, Bz;
                if (DIPOLE) {
                    Ex = Ey = Ez = 0.0f;
                    dipole_field(g, bxw, byw, bzw, px, py, pz, Bx, By, Bz);
                } else if ((unsigned)blx < (unsigned)G.T0 && (unsigned)bly < (unsigned)G.T1 &&
                           (unsigned)blz < (unsigned)G.T2)
                    gather_lds(sF4, sF2, blx * fsx + bly * fsy + blz, fsx, fsy, px, py, pz, Ex, Ey, Ez, Bx, By, Bz);
                else
                    gather_direct(A, bxw, byw, bzw, px, py, pz, Ex, Ey, Ez, Bx, By, Bz);
                boris_move(g, qom, subk, Ex, Ey, Ez, Bx, By, Bz, vx, vy, vz, px, py, pz);
                if (REFLECT) reflect_faces(g, bxw, byw, bzw, px, py, pz, vx, vy, vz);
                sunk = !finite3(px, py, pz);
                if (!sunk) {
                    dex = dcell(px); dey = dcell(py); dez = dcell(pz);
                }
                if (SINK && !sunk) {
                    const float rx_ = g.origin[0] + ((float)wrapc(bx + dex, g.n[0]) + px - (float)dex) * g.w[0] - g.sink_c[0];
                    const float ry_ = g.origin[1] + ((float)wrapc(by + dey, g.n[1]) + py - (float)dey) * g.w[1] - g.sink_c[1];
                    const float rz_ = g.origin[2] + ((float)wrapc(bz + dez, g.n[2]) + pz - (float)dez) * g.w[2] - g.sink_c[2];
                    sunk = (rx_ * rx_ + ry_ * ry_ + rz_ * rz_) <= g.sink_r2;
                }
            }
            bool ok = true;
            if (MODE != MODE_DEPOSIT) {
                const int bc = G.cid(g, bxw, byw, bzw);
                const int oidx = S.start[bc] + S.stay[bc] + atomicAdd(&S.arr[bc], 1);
                if (oidx < 0 || oidx >= S.start[bc + 1] || (long long)oidx >= S.cap) {
                    atomicOr(A.err, ERR_SLOT);
                    ok = false;
                } else {
                    store_particle(rout, cs, oidx, sunk ? NaNf : px, py, pz, vx, vy, vz);
                    if (TRACK) S.id_out[oidx] = S.id_in[inq];
                }
            }
            if (MODE == MODE_PUSH && sunk) {
                atomicAdd(A.removed, 1ull);
            } else if (MODE != MODE_REBIN && ok) {
                float wt[8];
                tri_weights(px - (float)dex, py - (float)dey, pz - (float)dez, wt);
                deposit_direct(A, G, sJ, x0, y0, z0, blx + dex, bly + dey, blz + dez, wt, qs * vx, qs * vy,
                               qs * vz);
                if (MODE == MODE_PUSH) {
                    count_add(g, G, s_cnt, S.cnt_acc, x0, y0, z0, bx + dex, by + dey, bz + dez);
                    if ((dex | dey | dez) == 0) count_add(g, G, s_stay, S.stay_acc, x0, y0, z0, bx, by, bz);
                }
            }
        };
        // leavers: deposit and count at the new cell, one per lane; empties the buffer
        int kn = 0;
        auto drain_leavers = [&]() {
            const int kq_n = min(kn, K0CAP);
            for (int qi = lane; qi < kq_n; qi += WS) {
                const float* e = K0 + qi * 8;
                const float px = e[0], py = e[1], pz = e[2], vx = e[3], vy = e[4], vz = e[5];
                const int kc = __float_as_int(e[6]);  // the tile cell it left
                int qlx, qly, qlz;
                G.split(kc, qlx, qly, qlz);
                const int dex = dcell(px), dey = dcell(py), dez = dcell(pz);
                float wt[8];
                tri_weights(px - (float)dex, py - (float)dey, pz - (float)dez, wt);
                deposit_direct(A, G, sJ, x0, y0, z0, qlx + dex, qly + dey, qlz + dez, wt, qs * vx, qs * vy, qs * vz);
                count_add(g, G, s_cnt, S.cnt_acc, x0, y0, z0, x0 + qlx + dex, y0 + qly + dey, z0 + qlz + dez);
            }
            wave_sync();
            kn = 0;
        };
        // processes the queue's first lim entries, one per lane, and moves the rest to the front
        int qn = 0;
        auto drain = [&](int lim) {
            for (int qi = lane; qi < lim; qi += WS) arrive(Q[qi]);
            if (lim > 0) {
                const int rest = qn - lim;
                const int v = lane < rest ? Q[lim + lane] : 0;
                wave_sync();
                if (lane < rest) Q[lane] = v;
                wave_sync();
                qn = rest;
            }
        };

        while (c_cur < chi) {
            // runs of the next 64 cells (lane l looks at cell c_cur + l)
            const int cl = c_cur + lane;
            int remp = 0;
            if (cl < chi) remp = s_seg[cl + 1] - (lane == 0 ? pos_cur : s_seg[cl]);
            const int rl = (remp + KRUN - 1) / KRUN;
            const int incl = wave_incl_scan(rl, lane);
            R[lane + 1] = incl;
            if (lane == 0) R[0] = 0;
            // run heads: H[r] = the cell whose first run is run r (0 elsewhere); the cell of run r
            // is then the running maximum of H up to r
            H[lane] = 0;
            wave_sync();
            if (rl > 0 && incl - rl < WS) H[incl - rl] = lane;
            const int all_runs = __shfl(incl, WS - 1, WS);
            wave_sync();
            const int nruns = min(all_runs, WS);
            if (nruns == 0) {  // the next 64 cells are empty
                c_cur = min(c_cur + WS, chi);
                pos_cur = c_cur < chi ? s_seg[c_cur] : 0;
                carry = 0;
                continue;
            }
            const bool active = lane < nruns;
            int k = wave_incl_max(H[lane]), len = 0, pst = 0;  // the run's cell: c_cur + k
            if (active) {
                const int j = lane - R[k];
                const int nr = R[k + 1] - R[k];
                const int p0 = (k == 0) ? pos_cur : s_seg[c_cur + k];
                const int nc = s_seg[c_cur + k + 1] - p0;
                const int bl = nc < (1 << 21) ? fdivi(nc, __builtin_amdgcn_rcpf((float)nr)) : nc / nr;
                const int rem = nc - bl * nr;
                len = bl + (j < rem ? 1 : 0);
                pst = p0 + j * bl + min(j, rem);
            }
            const int plo = __shfl(pst, 0, WS);
            const int phi = __shfl(pst + len, nruns - 1, WS);
            if (plo < 0 || (long long)phi > S.cap || phi - plo > WCH || phi < plo) {  // uniform over the wave
                if (lane == 0) atomicOr(A.err, ERR_CHUNK);
                break;
            }
            // the run cell's output range [a_start, a_end), loaded with the run so that both loads
            // are in flight together
            const int c_run = c_cur + k;
            const int a_start = (MODE != MODE_DEPOSIT && active) ? S.start[cbase + c_run] : 0;
            const int a_end = (MODE != MODE_DEPOSIT && active) ? S.start[cbase + c_run + 1] : 0;
            // this lane's run into registers
            float rx[KRUN], ry[KRUN], rz[KRUN], rvx[KRUN], rvy[KRUN], rvz[KRUN];
            uint32_t rid[TRACK ? KRUN : 1];
            const unsigned vrun = (unsigned)pst * 4u;
            // (dword loads: wider loads need aligned register tuples that push the kernel past its
            // register budget)
#pragma unroll
            for (int jj = 0; jj < KRUN; ++jj) {
                if (jj < len) {
                    const unsigned vo = vrun + 4u * jj;
                    rx[jj] = ldp(rin, vo, 0);
                    ry[jj] = ldp(rin, vo, cs);
                    rz[jj] = ldp(rin, vo, 2 * cs);
                    rvx[jj] = ldp(rin, vo, 3 * cs);
                    rvy[jj] = ldp(rin, vo, 4 * cs);
                    rvz[jj] = ldp(rin, vo, 5 * cs);
                    if (TRACK) rid[jj] = S.id_in[pst + jj];
                } else {
                    rx[jj] = ry[jj] = rz[jj] = NaNf;
                    rvx[jj] = rvy[jj] = rvz[jj] = 0.0f;
                    if (TRACK) rid[jj] = 0;
                }
            }
            const int c = c_cur + k;
            int lx, ly, lz;
            G.split(c, lx, ly, lz);
            const int ax = x0 + lx, ay = y0 + ly, az = z0 + lz;
            const int nb = lx * fsx + ly * fsy + lz;  // the run cell's first node in the LDS field

            // pass A: the run's home particles, bit jj of hm (particles of segment a still in cell
            // a, the fast path; NaN: no); they are the stayers that take the cell's first slots
            unsigned hm = 0;
#pragma unroll
            for (int jj = 0; jj < KRUN; ++jj)
                if (in_cell(rx[jj], ry[jj], rz[jj])) hm |= 1u << jj;
            const int my_stay = (MODE != MODE_DEPOSIT) ? __popc(hm) : 0;
            const int sincl = wave_incl_scan(my_stay, lane);
            const int sex = sincl - my_stay;
            const int sfirst = __shfl(sex, R[k], WS);
            const int base = (k == 0 ? carry : 0) + sex - sfirst;

            int run_in = 0, rank = 0;
            // slots of the run cell's stayers lie in [0, a_lim): below its end and the capacity
            const int a_lim = min(a_end, (int)S.cap);  // S.cap < 2^31 (one buffer resource < 4 GiB)

            // pass B1, push of the home particles in place in the run registers (gather, Boris,
            // reflection, sink); a removed particle gets a NaN x offset.  The run cell's 8 corner
            // nodes are read from LDS once per run and held in registers for this pass only, so
            // they never share the register file with the deposit moments of pass B2.
            if (MODE == MODE_PUSH) {
                float4 rn4[8];
                float2 rn2[8];
                if (!DIPOLE) {
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
#pragma unroll
                            for (int kk = 0; kk < 2; ++kk) {
                                const int nn = nb + i * fsx + j * fsy + kk;
                                rn4[(i * 2 + j) * 2 + kk] = sF4[nn];
                                rn2[(i * 2 + j) * 2 + kk] = sF2[nn];
                            }
                }
#pragma unroll
                for (int jj = 0; jj < KRUN; ++jj) {
                    if (hm & (1u << jj)) {
                        float px = rx[jj], py = ry[jj], pz = rz[jj];
                        float vx = rvx[jj], vy = rvy[jj], vz = rvz[jj];
                        float Ex, Ey, Ez, Bx, By, Bz;
                        if (DIPOLE) {
                            Ex = Ey = Ez = 0.0f;
                            dipole_field(g, ax, ay, az, px, py, pz, Bx, By, Bz);
                        } else {
                            gather_regs(rn4, rn2, px, py, pz, Ex, Ey, Ez, Bx, By, Bz);
                        }
                        boris_move(g, qom, subk, Ex, Ey, Ez, Bx, By, Bz, vx, vy, vz, px, py, pz);
                        if (REFLECT) reflect_faces(g, ax, ay, az, px, py, pz, vx, vy, vz);
                        bool sunk = !finite3(px, py, pz);
                        if (SINK && !sunk) {
                            const int dex = dcell(px), dey = dcell(py), dez = dcell(pz);
                            const float sx_ = g.origin[0] + ((float)wrapc(ax + dex, g.n[0]) + px - (float)dex) * g.w[0] - g.sink_c[0];
                            const float sy_ = g.origin[1] + ((float)wrapc(ay + dey, g.n[1]) + py - (float)dey) * g.w[1] - g.sink_c[1];
                            const float sz_ = g.origin[2] + ((float)wrapc(az + dez, g.n[2]) + pz - (float)dez) * g.w[2] - g.sink_c[2];
                            sunk = (sx_ * sx_ + sy_ * sy_ + sz_ * sz_) <= g.sink_r2;
                        }
                        rx[jj] = sunk ? NaNf : px;
                        ry[jj] = py;
                        rz[jj] = pz;
                        rvx[jj] = vx;
                        rvy[jj] = vy;
                        rvz[jj] = vz;
                    }
                }
            }

            // pass B2: slots, removal counts and deposit moments of the home particles; queue
            // entries for the others (their registers are untouched by pass B1)
            float mom[3][8];
#pragma unroll
            for (int m = 0; m < 8; ++m) mom[0][m] = mom[1][m] = mom[2][m] = 0.0f;
            float jmax = 0.0f;  // largest |j| component of the run: bounds every corner sum / run_in
            // this run's arrivals: particles that sit in another cell (queued by input index)
            unsigned amask = 0;
#pragma unroll
            for (int jj = 0; jj < KRUN; ++jj)
                if (!((hm >> jj) & 1u) && jj < len && alive(rx[jj], ry[jj], rz[jj])) amask |= 1u << jj;
            // A run with a single slot that is not home (an arrival, or a particle removed in the
            // previous step) closes the gap: the home particles move down one register, and the
            // run is stored like a run of home particles only.  The arrival itself is reloaded by
            // its input index when the queue is drained.  (Not in the instances that carry particle
            // ids, where the extra registers would spill; they store such runs particle by
            // particle.)
            if (MODE != MODE_DEPOSIT && !TRACK) {
                const unsigned gap = ((1u << len) - 1u) & ~hm;
                const bool one_gap = gap != 0u && (gap & (gap - 1u)) == 0u;
                if (__ballot(one_gap)) {
                    const int zg = one_gap ? __builtin_ctz(gap) : KRUN;
#pragma unroll
                    for (int i = 0; i + 1 < KRUN; ++i) {
                        const bool sh = i >= zg;
                        rx[i] = sh ? rx[i + 1] : rx[i];
                        ry[i] = sh ? ry[i + 1] : ry[i];
                        rz[i] = sh ? rz[i + 1] : rz[i];
                        rvx[i] = sh ? rvx[i + 1] : rvx[i];
                        rvy[i] = sh ? rvy[i + 1] : rvy[i];
                        rvz[i] = sh ? rvz[i + 1] : rvz[i];
                    }
                    if (one_gap) {
                        --len;
                        hm = (1u << len) - 1u;
                    }
                }
            }
            // a run of home particles only (the common case) takes consecutive slots: its six
            // arrays go out in a few stores covering whole lines across the wave
            const bool whole = (MODE != MODE_DEPOSIT) && len > 0 && hm == (1u << len) - 1u;
            if (whole) {
                const int o0 = a_start + base;
                if (o0 < 0 || o0 + len > a_lim) {
                    atomicOr(A.err, ERR_SLOT);
                } else {
                    const unsigned vo = (unsigned)o0 * 4u;
                    // the first four of each array in one dwordx4 store, the rest one by one
                    if (KRUN >= 4 && len >= 4) {
                        st4(rout, vo, 0, rx);
                        st4(rout, vo, cs, ry);
                        st4(rout, vo, 2 * cs, rz);
                        st4(rout, vo, 3 * cs, rvx);
                        st4(rout, vo, 4 * cs, rvy);
                        st4(rout, vo, 5 * cs, rvz);
                    }
                    const int jlo = (KRUN >= 4 && len >= 4) ? 4 : 0;
#pragma unroll
                    for (int jj = 0; jj < KRUN; ++jj)
                        if (jj >= jlo && jj < len)
                            store_particle(rout, cs, o0 + jj, rx[jj], ry[jj], rz[jj], rvx[jj], rvy[jj], rvz[jj]);
                    if (TRACK)
#pragma unroll
                        for (int jj = 0; jj < KRUN; ++jj)
                            if (jj < len) S.id_out[o0 + jj] = rid[jj];
                }
            }
            unsigned lmask = 0;  // this run's leavers (register positions after a closed gap)
#pragma unroll
            for (int jj = 0; jj < KRUN; ++jj) {
                const float px = rx[jj], py = ry[jj], pz = rz[jj];
                const float vx = rvx[jj], vy = rvy[jj], vz = rvz[jj];
                const bool home = (hm >> jj) & 1u;
                if (MODE != MODE_DEPOSIT && !whole && home) {
                    const int oidx = a_start + base + rank;
                    ++rank;
                    if ((unsigned)oidx >= (unsigned)a_lim) {
                        atomicOr(A.err, ERR_SLOT);
                    } else {
                        store_particle(rout, cs, oidx, px, py, pz, vx, vy, vz);
                        if (TRACK) S.id_out[oidx] = rid[jj];
                    }
                }
                // a home particle's x offset is NaN only when pass B1 removed it
                const bool sunk = (MODE == MODE_PUSH) && home && __builtin_isnan(px);
                if (sunk) atomicAdd(A.removed, 1ull);
                if (MODE != MODE_REBIN) {
                    const bool inc = home && (MODE == MODE_DEPOSIT || in_cell(px, py, pz));
                    if (home && !sunk && !inc) lmask |= 1u << jj;  // leaves cell a: queued below
                    // branch-free moments: a particle that deposits elsewhere (or not at all)
                    // adds zeros
                    const float dx = inc ? px : 0.0f, dy = inc ? py : 0.0f, dz = inc ? pz : 0.0f;
                    const float jx = inc ? qs * vx : 0.0f, jy = inc ? qs * vy : 0.0f, jz = inc ? qs * vz : 0.0f;
                    jmax = fmaxf(jmax, fmaxf(fabsf(jx), fmaxf(fabsf(jy), fabsf(jz))));
                    const float xy = dx * dy, xz = dx * dz, yz = dy * dz, xyz = xy * dz;
                    const float mono[8] = {1.0f, dx, dy, dz, xy, xz, yz, xyz};
#pragma unroll
                    for (int m = 0; m < 8; ++m) {
                        mom[0][m] = fmaf(jx, mono[m], mom[0][m]);
                        mom[1][m] = fmaf(jy, mono[m], mom[1][m]);
                        mom[2][m] = fmaf(jz, mono[m], mom[2][m]);
                    }
                    run_in += inc ? 1 : 0;
                }
            }
            // ballot compaction of both queues, for the chunks that have entries
            if (__ballot((amask | lmask) != 0)) {
#pragma unroll
                for (int jj = 0; jj < KRUN; ++jj) {
                    const bool arr = (amask >> jj) & 1u, leaver = (lmask >> jj) & 1u;
                    const unsigned long long bq = __ballot(arr);
                    if (arr) Q[qn + ballot_rank(bq)] = pst + jj;
                    qn += __popcll(bq);
                    const unsigned long long bk = __ballot(leaver);
                    if (leaver) {
                        const float px = rx[jj], py = ry[jj], pz = rz[jj];
                        const float vx = rvx[jj], vy = rvy[jj], vz = rvz[jj];
                        const int slot = kn + ballot_rank(bk);
                        if (slot < K0CAP) {
                            float* e = K0 + slot * 8;
                            e[0] = px; e[1] = py; e[2] = pz; e[3] = vx; e[4] = vy; e[5] = vz;
                            e[6] = __int_as_float(c);
                        } else {
                            // buffer full: handle in place
                            const int dex = dcell(px), dey = dcell(py), dez = dcell(pz);
                            float wt[8];
                            tri_weights(px - (float)dex, py - (float)dey, pz - (float)dez, wt);
                            deposit_direct(A, G, sJ, x0, y0, z0, lx + dex, ly + dey, lz + dez, wt, qs * vx,
                                           qs * vy, qs * vz);
                            count_add(g, G, s_cnt, S.cnt_acc, x0, y0, z0, ax + dex, ay + dey, az + dez);
                        }
                    }
                    kn += __popcll(bk);
                }
            }
            if (MODE != MODE_REBIN && run_in > 0) {
                if (!(jmax * (float)run_in <= g.fix_lim)) atomicOr(A.err, ERR_JSAT);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    float N[8];
                    mom_to_nodes(mom[d], N);
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
#pragma unroll
                            for (int kk = 0; kk < 2; ++kk) {
                                const int l = ((lx + JH + i) * G.nj1 + (ly + JH + j)) * G.nj2 + (lz + JH + kk);
                                fix_add(&sJ[d * G.NJ + l], N[(i * 2 + j) * 2 + kk], g.fix_scale);
                            }
                }
            }
            if (MODE == MODE_PUSH && run_in > 0) {
                atomicAdd(&s_cnt[c], run_in);
                atomicAdd(&s_stay[c], run_in);
            }
            wave_sync();

            // leavers wait in their buffer until it is half full, arrivals until a whole wave of
            // them is there
            if (kn >= K0CAP / 2) drain_leavers();
            drain(qn & ~(WS - 1));
            // advance the wave's cursor past the runs of this chunk
            const int klast = __shfl(k, nruns - 1, WS);
            if (all_runs <= WS) {
                c_cur = min(c_cur + WS, chi);
                pos_cur = c_cur < chi ? s_seg[c_cur] : 0;
                carry = 0;
            } else if (R[klast + 1] == WS) {  // the last cell ended with run 63
                c_cur = c_cur + klast + 1;
                pos_cur = c_cur < chi ? s_seg[c_cur] : 0;
                carry = 0;
            } else {  // cell c_cur + klast continues in the next chunk
                const int sl = __shfl(sincl, nruns - 1, WS);
                const int s0 = __shfl(sex, R[klast], WS);
                carry = (klast == 0 ? carry : 0) + sl - s0;
                c_cur = c_cur + klast;
                pos_cur = phi;
            }
            wave_sync();
        }
        if (kn > 0) drain_leavers();  // the leavers and arrivals still queued
        drain(qn);
        __syncthreads();
        if (MODE == MODE_PUSH) {
            for (int i = tid; i < G.TC; i += NTP) {
                if (s_cnt[i]) atomicAdd(&S.cnt_acc[cbase + i], s_cnt[i]);
                if (s_stay[i]) atomicAdd(&S.stay_acc[cbase + i], s_stay[i]);
            }
        }
    }
    __syncthreads();
    if (MODE != MODE_REBIN) {
        float* dst = A.jpart + ((long long)t * g.nsl + sl) * 3 * G.NJ;
        for (int i = tid; i < 3 * G.NJ; i += NTP) dst[i] = (float)((double)sJ[i] * g.fix_inv);
    }
}

// ---------------------------------------------------------------------------------------------
// field kernels (N-periodic, row-major float4)
// ---------------------------------------------------------------------------------------------

// The field is held as a static background (the state at upload) plus an FP32 perturbation
// that the leapfrog evolves: E = E0 + Ebg + dE, Bc = B0 + Bcbg + dBc.  The uniform parts E0,
// B0 are constants, the non-uniform parts are optional arrays, and the curls of the
// background enter as static sources SE = c^2 curl- Bc_bg and SB = curl+ E_bg computed in
// FP64 on the host.  A large uniform B0 therefore costs no precision in the curls.
struct FieldArgs {
    Geo g;
    float4* E;            // dE on nodes
    float4* Bc;           // dBc on centres
    float4* Bn;           // gather B minus B0: interpC2N(dBc) + Bnbg
    float4* J;
    float4* Eg;           // gather E minus E0 (dE + Ebg) when Ebg exists, else null
    const float4* Ebg;    // non-uniform background E on nodes (nullable)
    const float4* Bcbg;   // non-uniform background Bc on centres (nullable)
    const float4* Bnbg;   // interpC2N(Bcbg) on nodes (nullable)
    const float4* SE;     // static E source (nullable)
    const float4* SB;     // static B source (nullable)
    const float* jpart;
    long long* jover;
    unsigned char* jmark;
    const int* jpairs;  // [3][nmax][1+JPMAX]: count, then (tile << 8 | local)
    int nsl_read;       // slices of each tile's partials the sum reads: 1 once k_slice_sum ran
    float dt;           // field substep
    float c2;           // c^2 (field energies)
    float eF;           // factor of curl- B in the E pass: c^2 (leapfrog) or c (forward solver)
    float bF;           // factor of curl+ E in the B pass: 1 (leapfrog) or c (forward solver)
};

__device__ __forceinline__ void xyz_of(const Geo& g, int i, int& x, int& y, int& z) {
    x = i / (g.n[1] * g.n[2]);
    y = (i / g.n[2]) % g.n[1];
    z = i % g.n[2];
}

// The cross-tile moment sum: J of node (x,y,z) summed in FP64 over every tile partial covering
// it (every slice of the tile) plus the overflow accumulator (cleared here), rounded once to FP32.
// The (tile, local node) pairs of axis d whose J partial covers coordinate x, in increasing
// local index (the order of the jpairs table): without a halo, (x / T, x % T) and, on a tile's
// low face, (the previous tile, T); with a halo, from the table.
__device__ __forceinline__ int axis_pairs(const FieldArgs& F, int d, int x, int* p) {
    const Geo& g = F.g;
    if (g.jh == 0) {
        const int t = fdivi(x, g.inv_T[d]), l = x - t * g.T[d];
        p[0] = (t << 8) | l;
        if (l != 0) return 1;
        p[1] = ((t == 0 ? g.NTl[d] - 1 : t - 1) << 8) | g.T[d];
        return 2;
    }
    const int* e = F.jpairs + (d * g.nmax + x) * (1 + JPMAX);
    for (int q = 0; q < e[0]; ++q) p[q] = e[1 + q];
    return e[0];
}

__device__ __forceinline__ float4 sum_partials(const FieldArgs& F, int i, int x, int y, int z) {
    const Geo& g = F.g;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    int px[JPMAX], py[JPMAX], pz[JPMAX];
    const int nx = axis_pairs(F, 0, x, px), ny = axis_pairs(F, 1, y, py), nz = axis_pairs(F, 2, z, pz);
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b)
            for (int cc = 0; cc < nz; ++cc) {
                const int ta = px[a] >> 8, la = px[a] & 255;
                const int tb = py[b] >> 8, lb = py[b] & 255;
                const int tc = pz[cc] >> 8, lc = pz[cc] & 255;
                const int tile = (ta * g.NTl[1] + tb) * g.NTl[2] + tc;
                const int l = (la * g.nj[1] + lb) * g.nj[2] + lc;
                const float* base = F.jpart + (long long)tile * g.nsl * 3 * g.NJ;
                for (int s = 0; s < F.nsl_read; ++s, base += 3 * g.NJ) {
                    sx += (double)base[l];
                    sy += (double)base[g.NJ + l];
                    sz += (double)base[2 * g.NJ + l];
                }
            }
    // the overflow accumulator, read and cleared only at the nodes that hold deposits
    if (F.jmark[i]) {
        F.jmark[i] = 0;
        long long* o = F.jover + 3 * (long long)i;
        sx += (double)o[0] * g.fix_inv;
        sy += (double)o[1] * g.fix_inv;
        sz += (double)o[2] * g.fix_inv;
        o[0] = 0;
        o[1] = 0;
        o[2] = 0;
    }
    return make_float4((float)sx, (float)sy, (float)sz, 0.0f);
}

// The block slices of a tile (grids of few tiles split each tile among nsl blocks) summed in
// FP64 into the tile's first slice, one thread per partial entry: the node sums then read one
// slice per tile instead of nsl.
__global__ __launch_bounds__(256) void k_slice_sum(Geo g, float* jpart) {
    const long long per = 3LL * g.NJ;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)g.ntiles * per) return;
    const long long t = i / per, e = i - t * per;
    float* base = jpart + t * g.nsl * per + e;
    double acc = 0.0;
    for (int s = 0; s < g.nsl; ++s) acc += (double)base[s * per];
    base[0] = (float)acc;
}

// J from the tile partials, without a field update (ahead of the field operator).
__global__ __launch_bounds__(256) void k_jsum(FieldArgs F) {
    const Geo& g = F.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.ncell) return;
    int x, y, z;
    xyz_of(g, i, x, y, z);
    F.J[i] = sum_partials(F, i, x, y, z);
}

// BF16 of a finite float, rounded to nearest even (the network input's rounding).
__device__ __forceinline__ uint32_t bf16_rne(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b + 0x7FFFu + ((b >> 16) & 1u)) >> 16;
}

// The field operator's inputs in one pass: J from the tile partials when they are newer
// (stored, as k_jsum), the state before the phase (Eb, Bb: the fallback and the checks start
// from it), and the network input u, 16 BF16 channels per node (E iE, Bc iB, J iJ, zeros), in
// the layout of ipic_nop_pack.
__global__ __launch_bounds__(256) void k_op_prep(FieldArgs F, int from_partials, float4* Eb, float4* Bb, float iE,
                                                 float iB, float iJ, uint16_t* u) {
    const Geo& g = F.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.ncell) return;
    float4 j;
    if (from_partials) {
        int x, y, z;
        xyz_of(g, i, x, y, z);
        j = sum_partials(F, i, x, y, z);
        F.J[i] = j;
    } else {
        j = F.J[i];
    }
    const float4 e = F.E[i], b = F.Bc[i];
    Eb[i] = e;
    Bb[i] = b;
    uint4 lo, hi;
    lo.x = bf16_rne(e.x * iE) | (bf16_rne(e.y * iE) << 16);
    lo.y = bf16_rne(e.z * iE) | (bf16_rne(b.x * iB) << 16);
    lo.z = bf16_rne(b.y * iB) | (bf16_rne(b.z * iB) << 16);
    lo.w = bf16_rne(j.x * iJ) | (bf16_rne(j.y * iJ) << 16);
    hi.x = bf16_rne(j.z * iJ);
    hi.y = hi.z = hi.w = 0u;
    uint4* d = (uint4*)(u + (long long)i * 16);
    d[0] = lo;
    d[1] = hi;
}

// Drops overflow deposits that were never summed (a deposit pass replaces the partials).
__global__ void k_discard_overflow(unsigned char* jmark, long long* jover, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !jmark[i]) return;
    jmark[i] = 0;
    long long* o = jover + 3 * (long long)i;
    o[0] = 0;
    o[1] = 0;
    o[2] = 0;
}

// E pass of field.h:359-363: E += dt (c^2 curl- Bc - J).  With from_partials J is first
// formed from the tile partials (sum_partials) and stored.
__global__ __launch_bounds__(256) void k_field_E(FieldArgs F, int from_partials) {
    const Geo& g = F.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.ncell) return;
    int x, y, z;
    xyz_of(g, i, x, y, z);
    float Jx, Jy, Jz;
    if (from_partials) {
        const float4 j4 = sum_partials(F, i, x, y, z);
        Jx = j4.x; Jy = j4.y; Jz = j4.z;
        F.J[i] = j4;
    } else {
        const float4 j4 = F.J[i];
        Jx = j4.x; Jy = j4.y; Jz = j4.z;
    }
    const float4 b = F.Bc[i];
    const float4 bmx = F.Bc[nid_of(g, wrap1(x - 1, g.n[0]), y, z)];
    const float4 bmy = F.Bc[nid_of(g, x, wrap1(y - 1, g.n[1]), z)];
    const float4 bmz = F.Bc[nid_of(g, x, y, wrap1(z - 1, g.n[2]))];
    float4 e = F.E[i];
    const float dt = F.dt, c2 = F.eF;
    e.x = e.x + dt * (c2 * ((b.z - bmy.z) * g.inv_w[1] - (b.y - bmz.y) * g.inv_w[2]) - Jx);
    e.y = e.y + dt * (c2 * ((b.x - bmz.x) * g.inv_w[2] - (b.z - bmx.z) * g.inv_w[0]) - Jy);
    e.z = e.z + dt * (c2 * ((b.y - bmx.y) * g.inv_w[0] - (b.x - bmy.x) * g.inv_w[1]) - Jz);
    if (F.SE) {
        const float4 se = F.SE[i];
        e.x += dt * se.x; e.y += dt * se.y; e.z += dt * se.z;
    }
    F.E[i] = e;
    if (F.Eg) {
        const float4 eb = F.Ebg[i];
        F.Eg[i] = make_float4(e.x + eb.x, e.y + eb.y, e.z + eb.z, 0.0f);
    }
}

// B pass of field.h:366-371: Bc -= dt curl+ E.
__global__ __launch_bounds__(256) void k_field_B(FieldArgs F) {
    const Geo& g = F.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.ncell) return;
    int x, y, z;
    xyz_of(g, i, x, y, z);
    const float4 e = F.E[i];
    const float4 epx = F.E[nid_of(g, wrap1(x + 1, g.n[0]), y, z)];
    const float4 epy = F.E[nid_of(g, x, wrap1(y + 1, g.n[1]), z)];
    const float4 epz = F.E[nid_of(g, x, y, wrap1(z + 1, g.n[2]))];
    float4 b = F.Bc[i];
    const float dt = F.dt * F.bF;
    b.x = b.x - dt * ((epy.z - e.z) * g.inv_w[1] - (epz.y - e.y) * g.inv_w[2]);
    b.y = b.y - dt * ((epz.x - e.x) * g.inv_w[2] - (epx.z - e.z) * g.inv_w[0]);
    b.z = b.z - dt * ((epx.y - e.y) * g.inv_w[0] - (epy.x - e.x) * g.inv_w[1]);
    if (F.SB) {  // the source already carries the B-pass factor
        const float4 sb = F.SB[i];
        b.x -= F.dt * sb.x; b.y -= F.dt * sb.y; b.z -= F.dt * sb.z;
    }
    F.Bc[i] = b;
}

// interpC2N (field.h:234-247): node B = 1/8 sum of the 8 surrounding centres.
__global__ __launch_bounds__(256) void k_c2n(FieldArgs F) {
    const Geo& g = F.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.ncell) return;
    int x, y, z;
    xyz_of(g, i, x, y, z);
    float sx = 0, sy = 0, sz = 0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const float4 v = F.Bc[nid_of(g, wrap1(x - a, g.n[0]), wrap1(y - b, g.n[1]), wrap1(z - c, g.n[2]))];
                sx += v.x; sy += v.y; sz += v.z;
            }
    float4 r = make_float4(0.125f * sx, 0.125f * sy, 0.125f * sz, 0.0f);
    if (F.Bnbg) {
        const float4 bb = F.Bnbg[i];
        r.x += bb.x; r.y += bb.y; r.z += bb.z;
    }
    F.Bn[i] = r;
}

// ---------------------------------------------------------------------------------------------
// cell-count prefix sums (tile-major cell order)
// ---------------------------------------------------------------------------------------------

struct ScanSp {
    int* cnt_acc;
    int* stay_acc;
    int* seg;
    int* start;
    int* stay;
    int* arr;
    int* tile_total;
    int* tile_base;
};
struct ScanArgs {
    Geo g;
    ScanSp sp[IPIC_MAX_SPECIES];
};

__global__ __launch_bounds__(NT) void k_scan_tiles(ScanArgs A) {
    const Geo& g = A.g;
    const ScanSp& S = A.sp[blockIdx.y];
    __shared__ int s_w[NT / 64];
    const int t = blockIdx.x;
    int v = 0;
    for (int i = threadIdx.x; i < g.TC; i += NT) v += S.cnt_acc[t * g.TC + i];
    int tot;
    block_exscan(v, s_w, tot);
    if (threadIdx.x == 0) S.tile_total[t] = tot;
}

// Launch order of the particle kernel's tiles for the layout the scans just produced.  A
// block's time grows with its tile's particles.  With an uneven load (the heaviest tile more
// than 25 % above the mean) the heaviest tiles go first (longest-processing-time order over 256
// load classes), so that the blocks that start last are light.  Otherwise the tiles go in a
// scrambled order, tile (b * mul) mod ntiles for block b with mul coprime to ntiles: neighbouring
// tiles then run at different times, which on the uniform 128^3 x 32 workload takes 2 % off the
// push against launching neighbours together (the XCD-aware order).  Any permutation gives the
// same results: the order only schedules.
__device__ void tile_order_block(const ScanArgs& A, int nsp, int* order, int mul) {
    const Geo& g = A.g;
    __shared__ unsigned long long s_sum[16];
    __shared__ int s_max[16], s_w4[4];
    __shared__ int s_hist[256];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    auto load = [&](int t) {
        int v = 0;
        for (int sp = 0; sp < nsp; ++sp) v += A.sp[sp].tile_total[t];
        return v;
    };
    unsigned long long sum = 0;
    int mx = 0;
    for (int t = tid; t < g.ntiles; t += 1024) {
        const int v = load(t);
        sum += (unsigned long long)v;
        mx = max(mx, v);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        sum += __shfl_xor(sum, d, 64);
        mx = max(mx, __shfl_xor(mx, d, 64));
    }
    if (lane == 0) {
        s_sum[wid] = sum;
        s_max[wid] = mx;
    }
    if (tid < 256) s_hist[tid] = 0;
    __syncthreads();
    sum = 0;
    mx = 0;
    for (int w = 0; w < 16; ++w) {
        sum += s_sum[w];
        mx = max(mx, s_max[w]);
    }
    if ((double)mx * g.ntiles <= 1.25 * (double)sum || mx == 0) {
        for (int b = tid; b < g.ntiles; b += 1024) order[b] = (int)(((long long)b * mul) % g.ntiles);
        return;
    }
    auto cls = [&](int v) { return 255 - (int)(((long long)v * 255) / mx); };  // heaviest: class 0
    for (int t = tid; t < g.ntiles; t += 1024) atomicAdd(&s_hist[cls(load(t))], 1);
    __syncthreads();
    // exclusive prefix sum of the 256 class counts, by the first four waves
    int c = 0, incl = 0;
    if (tid < 256) {
        c = s_hist[tid];
        incl = wave_incl_scan(c, lane);
        if (lane == 63) s_w4[wid] = incl;
    }
    __syncthreads();
    if (tid < 256) {
        int off = 0;
        for (int w = 0; w < wid; ++w) off += s_w4[w];
        s_hist[tid] = off + incl - c;
    }
    __syncthreads();
    for (int t = tid; t < g.ntiles; t += 1024) order[atomicAdd(&s_hist[cls(load(t))], 1)] = t;
}

// Tile bases from the tile totals, one block per species; block nsp (the last) orders the
// tiles for the next push meanwhile.
__global__ __launch_bounds__(1024) void k_scan_top(ScanArgs A, int nsp, int* order, int mul) {
    if ((int)blockIdx.x == nsp) {
        tile_order_block(A, nsp, order, mul);
        return;
    }
    const Geo& g = A.g;
    const ScanSp& S = A.sp[blockIdx.x];
    __shared__ int s_part[1024];
    const int per = (g.ntiles + 1023) / 1024;
    const int b0 = threadIdx.x * per;
    int v = 0;
    for (int k = 0; k < per; ++k)
        if (b0 + k < g.ntiles) v += S.tile_total[b0 + k];
    s_part[threadIdx.x] = v;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partial sums
    for (int d = 1; d < 1024; d <<= 1) {
        const int y = threadIdx.x >= d ? s_part[threadIdx.x - d] : 0;
        __syncthreads();
        s_part[threadIdx.x] += y;
        __syncthreads();
    }
    int run = s_part[threadIdx.x] - v;
    for (int k = 0; k < per; ++k)
        if (b0 + k < g.ntiles) {
            S.tile_base[b0 + k] = run;
            run += S.tile_total[b0 + k];
        }
    if (threadIdx.x == 1023) S.tile_base[g.ntiles] = s_part[1023];
}

// New starts = tile base + in-tile exclusive prefix of the counts, written over the old
// segments; the host then swaps `seg` and `start` (the previous starts become the segments of
// the array just written, without a copy) and `stay` and `stay_acc` (the stayer counts become
// `stay`; the old `stay`, cleared here, the next accumulator).  The cell counters and arrival
// counters are cleared for the next pass.
__global__ __launch_bounds__(NT) void k_scan_cells(ScanArgs A) {
    const Geo& g = A.g;
    const ScanSp& S = A.sp[blockIdx.y];
    __shared__ int s_w[NT / 64];
    const int t = blockIdx.x;
    const int c0 = 2 * threadIdx.x, c1 = c0 + 1;
    const int id0 = t * g.TC + c0, id1 = t * g.TC + c1;
    int v0 = 0, v1 = 0;
    if (c0 < g.TC) v0 = S.cnt_acc[id0];
    if (c1 < g.TC) v1 = S.cnt_acc[id1];
    int tot;
    const int ex = block_exscan(v0 + v1, s_w, tot) + S.tile_base[t];
    if (c0 < g.TC) {
        S.seg[id0] = ex;
        S.stay[id0] = 0;
        S.cnt_acc[id0] = 0;
        S.arr[id0] = 0;
    }
    if (c1 < g.TC) {
        S.seg[id1] = ex + v0;
        S.stay[id1] = 0;
        S.cnt_acc[id1] = 0;
        S.arr[id1] = 0;
    }
    if (t == g.ntiles - 1 && threadIdx.x == 0) S.seg[g.ncell] = S.tile_base[g.ntiles];
}


// After a REBIN pass the array is sorted by current cell: segments = starts, all stay.
__global__ __launch_bounds__(256) void k_canon_fix(ScanArgs A) {
    const Geo& g = A.g;
    const ScanSp& S = A.sp[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > g.ncell) return;
    S.seg[i] = S.start[i];
    if (i < g.ncell) {
        S.stay[i] = S.start[i + 1] - S.start[i];
        S.arr[i] = 0;
    }
}

// Reproducible order of each cell's arrivals (IPIC_FLAG_REPRODUCIBLE).  A pass that moves
// particles writes the particles it takes into another cell (the arrivals of the lagged sort)
// to that cell's slots [start + stay, start + stay + arr) in the order of atomic counters, which
// follows timing.  These kernels reorder every such range by the particles' own contents (the
// bit patterns of the three offsets and three velocities, then the id when ids are carried),
// which do not depend on timing.  Every cell's order is then a function of the state, and so is
// the next pass, whose runs (and through them J's FP32 moment sums) follow that order.
struct CanonSp {
    float* out;          // the array the pass wrote: six arrays of cap floats
    float* spare;        // the array the pass read, dead now: holds a copy of each range
    uint32_t* id_out;    // particle ids (null when not carried)
    uint32_t* id_spare;
    long long cap;
    const int* start;
    const int* stay;
    const int* arr;
    int* work;           // [0]: number of cells left to k_canon_large, then their ids
};
struct CanonArgs {
    int ncell;
    int* err;
    CanonSp sp[IPIC_MAX_SPECIES];
};

constexpr int CANON_SMALL = 4;  // a cell with at most this many arrivals is ordered by one thread

// Particle in slot a comes before the one in slot b (both read from the copy).  Slots decide
// only between particles identical in every word, where either order gives the same array.
__device__ __forceinline__ bool canon_before(const CanonSp& S, long long a, long long b) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const uint32_t x = __float_as_uint(S.spare[k * S.cap + a]), y = __float_as_uint(S.spare[k * S.cap + b]);
        if (x != y) return x < y;
    }
    if (S.id_spare) {
        const uint32_t x = S.id_spare[a], y = S.id_spare[b];
        if (x != y) return x < y;
    }
    return a < b;
}

__device__ __forceinline__ void canon_copy(const CanonSp& S, long long i) {
#pragma unroll
    for (int k = 0; k < 6; ++k) S.spare[k * S.cap + i] = S.out[k * S.cap + i];
    if (S.id_out) S.id_spare[i] = S.id_out[i];
}

// The particle of copy slot i goes to its rank among the range [lo, lo + n).
__device__ __forceinline__ void canon_place(const CanonSp& S, long long lo, int n, long long i) {
    int r = 0;
    for (int j = 0; j < n; ++j) r += canon_before(S, lo + j, i) ? 1 : 0;  // j == i counts 0
#pragma unroll
    for (int k = 0; k < 6; ++k) S.out[k * S.cap + lo + r] = S.spare[k * S.cap + i];
    if (S.id_out) S.id_out[lo + r] = S.id_spare[i];
}

__global__ void k_canon_reset(CanonArgs A, int nsp) {
    if ((int)threadIdx.x < nsp) A.sp[threadIdx.x].work[0] = 0;
}

// One thread per cell and species: cells with 2..CANON_SMALL arrivals are ordered here, larger
// ones are listed for k_canon_large.
__global__ __launch_bounds__(256) void k_canon_small(CanonArgs A) {
    const CanonSp& S = A.sp[blockIdx.y];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A.ncell) return;
    const int n = S.arr[c];
    if (n < 2) return;
    const long long lo = (long long)S.start[c] + S.stay[c];
    if (lo < 0 || lo + n > S.start[c + 1] || lo + n > S.cap) {
        atomicOr(A.err, ERR_SLOT);
        return;
    }
    if (n > CANON_SMALL) {
        S.work[1 + atomicAdd(S.work, 1)] = c;  // at most one entry per cell: ncell + 1 ints
        return;
    }
    // the range in registers (every index below is a constant after unrolling), ranked there and
    // written back in rank order: all loads precede the first store
    uint32_t v[CANON_SMALL][7];
#pragma unroll
    for (int j = 0; j < CANON_SMALL; ++j)
        if (j < n) {
#pragma unroll
            for (int k = 0; k < 6; ++k) v[j][k] = __float_as_uint(S.out[k * S.cap + lo + j]);
            v[j][6] = S.id_out ? S.id_out[lo + j] : 0u;
        }
#pragma unroll
    for (int i = 0; i < CANON_SMALL; ++i)
        if (i < n) {
            int r = 0;
#pragma unroll
            for (int j = 0; j < CANON_SMALL; ++j)
                if (j < n && j != i) {
                    bool before = j < i;  // identical in every word: the slots decide
#pragma unroll
                    for (int k = 6; k >= 0; --k)
                        if (v[j][k] != v[i][k]) before = v[j][k] < v[i][k];
                    r += before ? 1 : 0;
                }
#pragma unroll
            for (int k = 0; k < 6; ++k) S.out[k * S.cap + lo + r] = __uint_as_float(v[i][k]);
            if (S.id_out) S.id_out[lo + r] = v[i][6];
        }
}

constexpr int CANON_LDS = 2048;  // arrivals whose leading key words k_canon_large keeps in LDS

// One block per listed cell (grid-stride over the list); ranges of different cells are disjoint.
// Up to CANON_LDS arrivals: the range is copied, the first two key words (x and y offset bits,
// one 64-bit word whose unsigned order is theirs) are sorted with their slots in LDS by a
// bitonic network, every group of equal words is then put in full-key order by one thread
// (insertion sort; such groups are rare and the network leaves them in an order that follows the
// input), and the particles are written back in sorted order.  Larger ranges rank each particle
// against all others in global memory.
__global__ __launch_bounds__(256) void k_canon_large(CanonArgs A) {
    __shared__ unsigned long long s_key[CANON_LDS];
    __shared__ int s_idx[CANON_LDS];
    const CanonSp& S = A.sp[blockIdx.y];
    const int nwork = S.work[0];
    const int tid = threadIdx.x;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int c = S.work[1 + w];
        const int n = S.arr[c];
        const long long lo = (long long)S.start[c] + S.stay[c];  // checked when listed
        if (n > CANON_LDS) {
            for (int i = tid; i < n; i += blockDim.x) canon_copy(S, lo + i);
            __syncthreads();
            for (int i = tid; i < n; i += blockDim.x) canon_place(S, lo, n, lo + i);
            __syncthreads();
            continue;
        }
        int np = 1;
        while (np < n) np <<= 1;
        for (int i = tid; i < np; i += blockDim.x) {
            if (i < n) {
                canon_copy(S, lo + i);
                s_key[i] = ((unsigned long long)__float_as_uint(S.out[lo + i]) << 32) |
                           __float_as_uint(S.out[S.cap + lo + i]);
            } else {
                s_key[i] = ~0ull;  // padding: after every real entry (the slot breaks a tie)
            }
            s_idx[i] = i;
        }
        __syncthreads();
        // ascending in (leading words, slot), all pairs distinct
        for (int k = 2; k <= np; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = tid; i < np; i += blockDim.x) {
                    const int p = i ^ j;
                    if (p > i) {
                        const unsigned long long a = s_key[i], b = s_key[p];
                        const bool asc = (i & k) == 0;
                        const bool gt = a > b || (a == b && s_idx[i] > s_idx[p]);
                        if (gt == asc) {
                            s_key[i] = b;
                            s_key[p] = a;
                            const int t = s_idx[i];
                            s_idx[i] = s_idx[p];
                            s_idx[p] = t;
                        }
                    }
                }
                __syncthreads();
            }
        // groups of equal leading words: full-key insertion sort by the group's first thread
        for (int i = tid; i < n; i += blockDim.x) {
            if ((i == 0 || s_key[i - 1] != s_key[i]) && i + 1 < n && s_key[i + 1] == s_key[i]) {
                int e = i + 1;
                while (e < n && s_key[e] == s_key[i]) ++e;
                for (int a = i + 1; a < e; ++a) {
                    const int v = s_idx[a];
                    int b = a - 1;
                    while (b >= i && canon_before(S, lo + v, lo + s_idx[b])) {
                        s_idx[b + 1] = s_idx[b];
                        --b;
                    }
                    s_idx[b + 1] = v;
                }
            }
        }
        __syncthreads();
        for (int i = tid; i < n; i += blockDim.x) {
            const long long src = lo + s_idx[i];
#pragma unroll
            for (int k = 0; k < 6; ++k) S.out[k * S.cap + lo + i] = S.spare[k * S.cap + src];
            if (S.id_out) S.id_out[lo + i] = S.id_spare[src];
        }
        __syncthreads();  // s_key and s_idx are rewritten for the next cell
    }
}

// ---------------------------------------------------------------------------------------------
// synthetic initialisation and FP64 reductions
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_init(Geo g, const int* start, float* X, uint32_t* ids, long long cap,
                                              long long total, uint64_t seed, int species, double vx_, double vy_,
                                              double vz_, double dx_, double dy_, double dz_) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    int lo = 0, hi = g.ncell - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= p) lo = mid; else hi = mid - 1;
    }
    const int cid = lo;
    const long long idx = p - start[cid];
    const int t = cid / g.TC, l = cid % g.TC;
    const int tx = t / (g.NTl[1] * g.NTl[2]), ty = (t / g.NTl[2]) % g.NTl[1], tz = t % g.NTl[2];
    const int x = tx * g.T[0] + l / (g.T[1] * g.T[2]);
    const int y = ty * g.T[1] + (l / g.T[2]) % g.T[1];
    const int z = tz * g.T[2] + l % g.T[2];
    const long long row = ((long long)x * g.n[1] + y) * g.n[2] + z;
    uint64_t key = mix64(mix64(mix64(seed) ^ (uint64_t)(uint32_t)species) ^ (uint64_t)row);
    key = mix64(key ^ (uint64_t)idx);
    double r[7];
    for (int k = 0; k < 7; ++k)
        r[k] = (double)(mix64(key + (uint64_t)(k + 1) * 0xD1B54A32D192ED03ULL) >> 11) * (1.0 / 9007199254740992.0);
    const double twopi = 6.283185307179586;
    const double m0 = sqrt(-2.0 * log(1.0 - r[3]));
    const double m1 = sqrt(-2.0 * log(1.0 - r[5]));
    const double n0 = m0 * cos(twopi * r[4]), n1 = m0 * sin(twopi * r[4]), n2 = m1 * cos(twopi * r[6]);
    X[p] = (float)r[0];
    X[cap + p] = (float)r[1];
    X[2 * cap + p] = (float)r[2];
    X[3 * cap + p] = (float)(dx_ + vx_ * n0);
    X[4 * cap + p] = (float)(dy_ + vy_ * n1);
    X[5 * cap + p] = (float)(dz_ + vz_ * n2);
    if (ids) ids[p] = (uint32_t)p;
}

__device__ __forceinline__ double block_sum_d(double v, double* s) {
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += s[w];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_max_d(double v, double* s) {
    for (int d = 32; d > 0; d >>= 1) v = fmax(v, __shfl_down(v, d, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, s[w]);
    __syncthreads();
    return r;
}

// FP64 per-block partials of sum 0.5 m |v|^2, sum m |v|, count (cell.h:1315-1338).
__global__ __launch_bounds__(256) void k_reduce_particles(const float* X, long long cap, long long nslots,
                                                          double mass, double* out) {
    __shared__ double s[4];
    double ke = 0, mo = 0, cnt = 0;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < nslots;
         p += (long long)gridDim.x * blockDim.x) {
        const float ox = X[p];
        if (!(ox == ox)) continue;
        const double vx = X[3 * cap + p], vy = X[4 * cap + p], vz = X[5 * cap + p];
        const double v2 = vx * vx + vy * vy + vz * vz;
        ke += 0.5 * mass * v2;
        mo += mass * sqrt(v2);
        cnt += 1.0;
    }
    ke = block_sum_d(ke, s);
    mo = block_sum_d(mo, s);
    cnt = block_sum_d(cnt, s);
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x + 0] = ke;
        out[3 * blockIdx.x + 1] = mo;
        out[3 * blockIdx.x + 2] = cnt;
    }
}

__device__ __forceinline__ void tot_field(const float4* d, const float4* bg, const float c0[3], int i, double r[3]) {
    const float4 v = d[i];
    r[0] = (double)c0[0] + v.x;
    r[1] = (double)c0[1] + v.y;
    r[2] = (double)c0[2] + v.z;
    if (bg) {
        const float4 u = bg[i];
        r[0] += u.x; r[1] += u.y; r[2] += u.z;
    }
}

// FP64 per-block partials of sum|E|^2, sum|Bc|^2, sum E.J, sum (div+ Bc)^2, max |div+ Bc|
// of the total field (background + perturbation).
__global__ __launch_bounds__(256) void k_reduce_fields(FieldArgs F, double* out) {
    const Geo& g = F.g;
    __shared__ double s[4];
    double se = 0, sb = 0, sej = 0, sd2 = 0, md = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.ncell; i += gridDim.x * blockDim.x) {
        int x, y, z;
        xyz_of(g, i, x, y, z);
        const float4 j = F.J[i];
        double e[3], b[3];
        tot_field(F.E, F.Ebg, g.E0, i, e);
        tot_field(F.Bc, F.Bcbg, g.B0, i, b);
        se += e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
        sb += b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        sej += e[0] * j.x + e[1] * j.y + e[2] * j.z;
        double bpx[3], bpy[3], bpz[3];
        tot_field(F.Bc, F.Bcbg, g.B0, nid_of(g, wrap1(x + 1, g.n[0]), y, z), bpx);
        tot_field(F.Bc, F.Bcbg, g.B0, nid_of(g, x, wrap1(y + 1, g.n[1]), z), bpy);
        tot_field(F.Bc, F.Bcbg, g.B0, nid_of(g, x, y, wrap1(z + 1, g.n[2])), bpz);
        const double dv = (bpx[0] - b[0]) * g.inv_w[0] + (bpy[1] - b[1]) * g.inv_w[1] +
                          (bpz[2] - b[2]) * g.inv_w[2];
        sd2 += dv * dv;
        md = fmax(md, fabs(dv));
    }
    se = block_sum_d(se, s);
    sb = block_sum_d(sb, s);
    sej = block_sum_d(sej, s);
    sd2 = block_sum_d(sd2, s);
    md = block_max_d(md, s);
    if (threadIdx.x == 0) {
        double* o = out + 5 * blockIdx.x;
        o[0] = se; o[1] = sb; o[2] = sej; o[3] = sd2; o[4] = md;
    }
}

// Checks of a field-operator step, FP64 per-block partials: sum over nodes/centres of
// |E0 + dE|^2 + c^2 |dBc|^2 before and after, of (E0 + dE) . J before and after, and
// max |div+ (dBc' - dBc)| (the divergence the update adds).
__global__ __launch_bounds__(256) void k_op_check(FieldArgs F, const float4* Eold, const float4* Bold, double* out) {
    const Geo& g = F.g;
    __shared__ double s[4];
    double wo = 0, wn = 0, ejo = 0, ejn = 0, md = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.ncell; i += gridDim.x * blockDim.x) {
        int x, y, z;
        xyz_of(g, i, x, y, z);
        const float4 eo = Eold[i], en = F.E[i], bo = Bold[i], bn = F.Bc[i], j = F.J[i];
        const double ox = (double)g.E0[0] + eo.x, oy = (double)g.E0[1] + eo.y, oz = (double)g.E0[2] + eo.z;
        const double nx = (double)g.E0[0] + en.x, ny = (double)g.E0[1] + en.y, nz = (double)g.E0[2] + en.z;
        wo += ox * ox + oy * oy + oz * oz + F.c2 * ((double)bo.x * bo.x + (double)bo.y * bo.y + (double)bo.z * bo.z);
        wn += nx * nx + ny * ny + nz * nz + F.c2 * ((double)bn.x * bn.x + (double)bn.y * bn.y + (double)bn.z * bn.z);
        ejo += ox * j.x + oy * j.y + oz * j.z;
        ejn += nx * j.x + ny * j.y + nz * j.z;
        const int ipx = nid_of(g, wrap1(x + 1, g.n[0]), y, z), ipy = nid_of(g, x, wrap1(y + 1, g.n[1]), z),
                  ipz = nid_of(g, x, y, wrap1(z + 1, g.n[2]));
        const double d0x = (double)bn.x - bo.x, d0y = (double)bn.y - bo.y, d0z = (double)bn.z - bo.z;
        const double dpx = (double)F.Bc[ipx].x - Bold[ipx].x;
        const double dpy = (double)F.Bc[ipy].y - Bold[ipy].y;
        const double dpz = (double)F.Bc[ipz].z - Bold[ipz].z;
        const double dv = (dpx - d0x) * g.inv_w[0] + (dpy - d0y) * g.inv_w[1] + (dpz - d0z) * g.inv_w[2];
        md = fmax(md, fabs(dv));
    }
    wo = block_sum_d(wo, s);
    wn = block_sum_d(wn, s);
    ejo = block_sum_d(ejo, s);
    ejn = block_sum_d(ejn, s);
    md = block_max_d(md, s);
    if (threadIdx.x == 0) {
        double* o = out + 5 * blockIdx.x;
        o[0] = wo; o[1] = wn; o[2] = ejo; o[3] = ejn; o[4] = md;
    }
}

// Audit sums: |Fa - Fb|^2, |Fb - Fc|^2, |Fb + F0|^2 and |Fb|^2 with |F|^2 = |E|^2 + c^2 |Bc|^2 (a:
// operator result, b: leapfrog result, c: state before the field phase, all without the uniform
// background F0).
__global__ __launch_bounds__(256) void k_audit(Geo g, float c2, const float4* Ea, const float4* Ba,
                                               const float4* Eb, const float4* Bb, const float4* Ec,
                                               const float4* Bc, double* out) {
    __shared__ double s[4];
    const int ncell = g.ncell;
    double num = 0, den = 0, tot = 0, evo = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ncell; i += gridDim.x * blockDim.x) {
        const float4 ea = Ea[i], ba = Ba[i], eb = Eb[i], bb = Bb[i], ec = Ec[i], bc = Bc[i];
        const double a0 = (double)ea.x - eb.x, a1 = (double)ea.y - eb.y, a2 = (double)ea.z - eb.z;
        const double a3 = (double)ba.x - bb.x, a4 = (double)ba.y - bb.y, a5 = (double)ba.z - bb.z;
        const double b0 = (double)eb.x - ec.x, b1 = (double)eb.y - ec.y, b2 = (double)eb.z - ec.z;
        const double b3 = (double)bb.x - bc.x, b4 = (double)bb.y - bc.y, b5 = (double)bb.z - bc.z;
        num += a0 * a0 + a1 * a1 + a2 * a2 + c2 * (a3 * a3 + a4 * a4 + a5 * a5);
        den += b0 * b0 + b1 * b1 + b2 * b2 + c2 * (b3 * b3 + b4 * b4 + b5 * b5);
        const double t0 = (double)g.E0[0] + eb.x, t1 = (double)g.E0[1] + eb.y, t2 = (double)g.E0[2] + eb.z;
        const double t3 = (double)g.B0[0] + bb.x, t4 = (double)g.B0[1] + bb.y, t5 = (double)g.B0[2] + bb.z;
        tot += t0 * t0 + t1 * t1 + t2 * t2 + c2 * (t3 * t3 + t4 * t4 + t5 * t5);
        evo += (double)eb.x * eb.x + (double)eb.y * eb.y + (double)eb.z * eb.z +
               c2 * ((double)bb.x * bb.x + (double)bb.y * bb.y + (double)bb.z * bb.z);
    }
    num = block_sum_d(num, s);
    den = block_sum_d(den, s);
    tot = block_sum_d(tot, s);
    evo = block_sum_d(evo, s);
    if (threadIdx.x == 0) {
        out[4 * blockIdx.x] = num;
        out[4 * blockIdx.x + 1] = den;
        out[4 * blockIdx.x + 2] = tot;
        out[4 * blockIdx.x + 3] = evo;
    }
}

__global__ void k_fill_f4(float4* p, long long n, float4 v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// =============================================================================================
// host side
// =============================================================================================

namespace {

thread_local std::string g_last_error;

struct Species {
    double q = 0, qom = 0;
    float* X[2] = {nullptr, nullptr};
    uint32_t* ids[2] = {nullptr, nullptr};
    long long cap = 0;
    int cur = 0;
    int* seg = nullptr;
    int* start = nullptr;
    int* stay = nullptr;
    int* cnt_acc = nullptr;
    int* stay_acc = nullptr;
    int* arr = nullptr;
    int* tile_total = nullptr;
    int* tile_base = nullptr;
    int* work = nullptr;  // cells whose arrivals k_canon_large orders (IPIC_FLAG_REPRODUCIBLE)
};

}  // namespace

struct ipic_ctx {
    ipic_desc d{};
    Geo g{};
    hipStream_t stream = nullptr;
    float4* E = nullptr;
    float4* Bc = nullptr;
    float4* Bn = nullptr;
    float4* J = nullptr;
    float4* Eg = nullptr;     // optional device arrays of the background split
    float4* Ebg = nullptr;
    float4* Bcbg = nullptr;
    float4* Bnbg = nullptr;
    float4* SE = nullptr;
    float4* SB = nullptr;
    std::vector<double> bgE, bgBc;  // background (state at the last field upload), FP64
    float* jpart = nullptr;
    long long* jover = nullptr;
    unsigned char* jmark = nullptr;  // nodes that hold overflow deposits (one byte each)
    int* jpairs = nullptr;
    unsigned long long* removed = nullptr;
    int* derr = nullptr;
    int* tile_order = nullptr;     // launch order of the particle kernel's tiles
    int order_mul = 1;             // multiplier of the scrambled tile order, coprime to ntiles
    bool order_valid = false;      // tile_order describes the current layout's loads
    double* red = nullptr;
    int red_blocks = 1024;
    Species sp[IPIC_MAX_SPECIES];
    bool j_valid = false;
    bool j_summed = true;          // J holds the sum of the current tile partials (and overflow)
    bool canonical = true;
    bool track = false;
    std::vector<void*> allocs;
    std::string err;
    hipEvent_t ev[8] = {};
    hipEvent_t ev_wait = nullptr;  // polled by wait_stream
    double* hred = nullptr;        // pinned host copy of the operator checks' block sums
    // field operator
    bool op_on = false;
    ipic_field_operator op{};      // settings (params pointer not kept)
    ipic_nop_scales op_scales{};
    long long op_count = 0;        // operator applications since it was set
    float* op_params = nullptr;
    uint16_t* op_wf = nullptr;
    unsigned char* op_ws = nullptr;
    uint16_t* op_u = nullptr;
    float* op_out = nullptr;
    float4* op_psi = nullptr;
    float4* Eb = nullptr;          // state before the field phase
    float4* Bb = nullptr;
    float4* Eo = nullptr;          // operator result during an audit
    float4* Bo = nullptr;
};

namespace {

int fail(ipic_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    g_last_error = msg;
    return code;
}

#define HIPCK(ctx, expr)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), IPIC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(ipic_ctx* ctx, T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
    HIPCK(ctx, hipMalloc(&q, bytes));
    HIPCK(ctx, hipMemsetAsync(q, 0, bytes, ctx->stream));
    ctx->allocs.push_back(q);
    *p = (T*)q;
    return 0;
}

void dfree(ipic_ctx* ctx, void* p) {
    if (!p) return;
    auto it = std::find(ctx->allocs.begin(), ctx->allocs.end(), p);
    if (it != ctx->allocs.end()) ctx->allocs.erase(it);
    (void)hipFree(p);
}

int largest_divisor_le(int n, int cap) {
    for (int t = std::min(n, cap); t >= 1; --t)
        if (n % t == 0) return t;
    return 1;
}

// per axis, the (tile, local node) pairs whose J partial region [tile*T - 1, tile*T + T + 2)
// contains coordinate x (periodic)
int build_jpairs(const Geo& g, std::vector<int>& tab) {
    tab.assign((size_t)3 * g.nmax * (1 + JPMAX), 0);
    for (int d = 0; d < 3; ++d)
        for (int x = 0; x < g.n[d]; ++x) {
            int* e = &tab[((size_t)d * g.nmax + x) * (1 + JPMAX)];
            for (int l = 0; l < g.nj[d]; ++l) {
                int s = (x + g.jh - l) % g.n[d];
                if (s < 0) s += g.n[d];
                if (s % g.T[d] == 0) {
                    if (e[0] >= JPMAX) return -1;
                    e[1 + e[0]] = ((s / g.T[d]) << 8) | l;
                    e[0]++;
                }
            }
        }
    return 0;
}

PushArgs push_args(ipic_ctx* ctx) {
    PushArgs A{};
    A.g = ctx->g;
    A.nsp = ctx->d.nspecies;
    const double vol = ctx->d.cell_width[0] * ctx->d.cell_width[1] * ctx->d.cell_width[2];
    for (int s = 0; s < A.nsp; ++s) {
        Species& S = ctx->sp[s];
        SpArgs& a = A.sp[s];
        a.in = S.X[S.cur];
        a.out = S.X[1 - S.cur];
        a.id_in = S.ids[S.cur];
        a.id_out = S.ids[1 - S.cur];
        a.cap = S.cap;
        a.seg = S.seg;
        a.start = S.start;
        a.stay = S.stay;
        a.arr = S.arr;
        a.cnt_acc = S.cnt_acc;
        a.stay_acc = S.stay_acc;
        a.qom = (float)S.qom;
        a.qs = (float)(S.q / (8.0 * vol));
        a.subk = (float)(4.0 * std::fabs(S.qom) * ctx->d.dt / (M_PI * ctx->d.c));
    }
    A.E = ctx->Eg ? ctx->Eg : ctx->E;
    A.Bn = ctx->Bn;
    A.jpart = ctx->jpart;
    A.jover = ctx->jover;
    A.jmark = ctx->jmark;
    A.removed = ctx->removed;
    A.err = ctx->derr;
    A.tile_order = ctx->order_valid ? ctx->tile_order : nullptr;
    return A;
}

FieldArgs field_args(ipic_ctx* ctx) {
    FieldArgs F{};
    F.g = ctx->g;
    F.E = ctx->E;
    F.Bc = ctx->Bc;
    F.Bn = ctx->Bn;
    F.J = ctx->J;
    F.Eg = ctx->Eg;
    F.Ebg = ctx->Ebg;
    F.Bcbg = ctx->Bcbg;
    F.Bnbg = ctx->Bnbg;
    F.SE = ctx->SE;
    F.SB = ctx->SB;
    F.jpart = ctx->jpart;
    F.jover = ctx->jover;
    F.jmark = ctx->jmark;
    F.jpairs = ctx->jpairs;
    F.nsl_read = ctx->g.nsl;
    F.dt = ctx->d.field_substeps > 0 ? (float)(ctx->d.dt / ctx->d.field_substeps) : 0.0f;
    F.c2 = (float)(ctx->d.c * ctx->d.c);
    const bool fwd = (ctx->d.flags & IPIC_FLAG_FORWARD_SOLVER) != 0;
    F.eF = fwd ? (float)ctx->d.c : F.c2;
    F.bF = fwd ? (float)ctx->d.c : 1.0f;
    return F;
}

ScanArgs scan_args(ipic_ctx* ctx) {
    ScanArgs A{};
    A.g = ctx->g;
    for (int s = 0; s < ctx->d.nspecies; ++s) {
        Species& S = ctx->sp[s];
        A.sp[s] = ScanSp{S.cnt_acc, S.stay_acc, S.seg, S.start, S.stay, S.arr, S.tile_total, S.tile_base};
    }
    return A;
}

template <int MODE, bool TRACK, int OPT>
void launch_k(dim3 grid, dim3 block, size_t lds, hipStream_t st, const PushArgs& A) {
    hipLaunchKernelGGL((k_particles<MODE, TRACK, OPT>), grid, block, lds, st, A);
}

// options of the push (tile shape, sink, reflection, dipole mover); the deposit and rebin
// passes depend on the tile shape only
template <int MODE, bool TRACK, int OPT>
void launch_opt_t(int opt, dim3 grid, dim3 block, size_t lds, hipStream_t st, const PushArgs& A) {
    if constexpr (OPT == 15) {
        launch_k<MODE, TRACK, 15>(grid, block, lds, st, A);
    } else {
        if (opt == OPT) launch_k<MODE, TRACK, OPT>(grid, block, lds, st, A);
        else launch_opt_t<MODE, TRACK, OPT + 1>(opt, grid, block, lds, st, A);
    }
}

template <int MODE, bool TRACK>
void launch_opt(int opt, dim3 grid, dim3 block, size_t lds, hipStream_t st, const PushArgs& A) {
    if constexpr (MODE != MODE_PUSH) {
        if (opt & OPT_FIX8) launch_k<MODE, TRACK, OPT_FIX8>(grid, block, lds, st, A);
        else launch_k<MODE, TRACK, 0>(grid, block, lds, st, A);
    } else {
        launch_opt_t<MODE, TRACK, 0>(opt, grid, block, lds, st, A);
    }
}

int launch_particles(ipic_ctx* ctx, int mode) {
    if (mode != MODE_REBIN && !ctx->j_summed) {
        hipLaunchKernelGGL(k_discard_overflow, dim3((unsigned)((ctx->g.ncell + 255) / 256)), dim3(256), 0,
                           ctx->stream, ctx->jmark, ctx->jover, ctx->g.ncell);
    }
    PushArgs A = push_args(ctx);
    const Geo& g = ctx->g;
    dim3 grid(g.ntiles, g.nsl), block(NTP);
    const bool dip = (ctx->d.flags & IPIC_FLAG_DIPOLE_MOVER) != 0;
    const size_t lds =
        (size_t)particle_layout(g.NJ, g.TC, (mode == MODE_PUSH && !dip) ? field_nodes(g) : 0).total;
    const int opt = ((g.T[0] == 8 && g.T[1] == 8 && g.T[2] == 8) ? OPT_FIX8 : 0) | (g.sink ? OPT_SINK : 0) |
                    (g.reflect ? OPT_REFLECT : 0) | (dip ? OPT_DIPOLE : 0);
    hipStream_t st = ctx->stream;
    if (ctx->track) {
        if (mode == MODE_PUSH) launch_opt<MODE_PUSH, true>(opt, grid, block, lds, st, A);
        else if (mode == MODE_DEPOSIT) launch_opt<MODE_DEPOSIT, true>(opt, grid, block, lds, st, A);
        else launch_opt<MODE_REBIN, true>(opt, grid, block, lds, st, A);
    } else {
        if (mode == MODE_PUSH) launch_opt<MODE_PUSH, false>(opt, grid, block, lds, st, A);
        else if (mode == MODE_DEPOSIT) launch_opt<MODE_DEPOSIT, false>(opt, grid, block, lds, st, A);
        else launch_opt<MODE_REBIN, false>(opt, grid, block, lds, st, A);
    }
    HIPCK(ctx, hipGetLastError());
    if (mode != MODE_REBIN) ctx->j_summed = false;
    if (mode != MODE_DEPOSIT)
        for (int s = 0; s < ctx->d.nspecies; ++s) ctx->sp[s].cur = 1 - ctx->sp[s].cur;
    return 0;
}

int launch_scan(ipic_ctx* ctx) {
    ScanArgs A = scan_args(ctx);
    const int ns = ctx->d.nspecies;
    hipLaunchKernelGGL(k_scan_tiles, dim3(ctx->g.ntiles, ns), dim3(NT), 0, ctx->stream, A);
    hipLaunchKernelGGL(k_scan_top, dim3(ns + 1), dim3(1024), 0, ctx->stream, A, ns, ctx->tile_order, ctx->order_mul);
    hipLaunchKernelGGL(k_scan_cells, dim3(ctx->g.ntiles, ns), dim3(NT), 0, ctx->stream, A);
    HIPCK(ctx, hipGetLastError());
    for (int s = 0; s < ns; ++s) {  // the roles k_scan_cells gave the arrays
        std::swap(ctx->sp[s].seg, ctx->sp[s].start);
        std::swap(ctx->sp[s].stay, ctx->sp[s].stay_acc);
    }
    ctx->order_valid = true;
    return 0;
}

// After a pass that moved particles (the buffers already swapped): each cell's arrivals in the
// order of their contents, when the context asks for reproducible steps.
int launch_canon(ipic_ctx* ctx) {
    if (!(ctx->d.flags & IPIC_FLAG_REPRODUCIBLE)) return 0;
    CanonArgs A{};
    const int ns = ctx->d.nspecies;
    A.ncell = ctx->g.ncell;
    A.err = ctx->derr;
    for (int s = 0; s < ns; ++s) {
        Species& S = ctx->sp[s];
        A.sp[s] = CanonSp{S.X[S.cur], S.X[1 - S.cur], S.ids[S.cur], S.ids[1 - S.cur], S.cap,
                          S.start, S.stay, S.arr, S.work};
    }
    hipLaunchKernelGGL(k_canon_reset, dim3(1), dim3(64), 0, ctx->stream, A, ns);
    hipLaunchKernelGGL(k_canon_small, dim3((unsigned)((ctx->g.ncell + 255) / 256), ns), dim3(256), 0, ctx->stream, A);
    hipLaunchKernelGGL(k_canon_large, dim3(2048, ns), dim3(256), 0, ctx->stream, A);
    HIPCK(ctx, hipGetLastError());
    return 0;
}

int launch_jsum(ipic_ctx* ctx);

// Reduces the block slices of the partials (a no-op with one slice per tile); F reads one slice
// afterwards.
int slice_sum(ipic_ctx* ctx, FieldArgs& F) {
    const Geo& g = ctx->g;
    if (g.nsl > 1) {
        const long long n = (long long)g.ntiles * 3 * g.NJ;
        hipLaunchKernelGGL(k_slice_sum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, g, ctx->jpart);
        HIPCK(ctx, hipGetLastError());
    }
    F.nsl_read = 1;
    return 0;
}

int launch_leapfrog(ipic_ctx* ctx) {
    // the static field uses no J: it is formed on demand (download, conservation, field state)
    // from the current deposit, and the next deposit pass clears the overflow accumulator
    if (ctx->d.field_substeps == 0) return 0;
    FieldArgs F = field_args(ctx);
    const int nb = (ctx->g.ncell + 255) / 256;
    for (int s = 0; s < ctx->d.field_substeps; ++s) {
        // the first E pass forms J from the partials unless that was done already
        const int from_partials = (s == 0 && !ctx->j_summed) ? 1 : 0;
        if (from_partials)
            if (int rc = slice_sum(ctx, F)) return rc;
        hipLaunchKernelGGL(k_field_E, dim3(nb), dim3(256), 0, ctx->stream, F, from_partials);
        hipLaunchKernelGGL(k_field_B, dim3(nb), dim3(256), 0, ctx->stream, F);
    }
    HIPCK(ctx, hipGetLastError());
    ctx->j_summed = true;
    return 0;
}

int launch_c2n(ipic_ctx* ctx) {
    FieldArgs F = field_args(ctx);
    const int nb = (ctx->g.ncell + 255) / 256;
    hipLaunchKernelGGL(k_c2n, dim3(nb), dim3(256), 0, ctx->stream, F);
    HIPCK(ctx, hipGetLastError());
    return 0;
}

// J formed from the tile partials and the overflow accumulator (which it consumes), once per
// deposit
int launch_jsum(ipic_ctx* ctx) {
    if (ctx->j_summed) return 0;
    FieldArgs F = field_args(ctx);
    if (int rc = slice_sum(ctx, F)) return rc;
    const int nb = (ctx->g.ncell + 255) / 256;
    hipLaunchKernelGGL(k_jsum, dim3(nb), dim3(256), 0, ctx->stream, F);
    HIPCK(ctx, hipGetLastError());
    ctx->j_summed = true;
    return 0;
}

int copy_f4(ipic_ctx* ctx, float4* dst, const float4* src) {
    HIPCK(ctx, hipMemcpyAsync(dst, src, (size_t)ctx->g.ncell * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
    return 0;
}

// Waits for the work queued on the context's stream by polling an event.  The operator phase
// waits in the middle of a field phase for its check sums; a blocking wait can leave the host
// thread asleep well past the end of the work, which then delays everything queued after it.
int wait_stream(ipic_ctx* ctx) {
    HIPCK(ctx, hipEventRecord(ctx->ev_wait, ctx->stream));
    hipError_t e;
    while ((e = hipEventQuery(ctx->ev_wait)) == hipErrorNotReady) {
    }
    HIPCK(ctx, e);
    return 0;
}

int nop_fail(ipic_ctx* ctx, int rc) { return fail(ctx, rc, std::string("field operator: ") + ipic_nop_last_error()); }

// One operator application with its checks, fallback and audit (the field phase without C2N).
int operator_phase(ipic_ctx* ctx, ipic_step_stats* st) {
    const Geo& g = ctx->g;
    if (ctx->SE || ctx->SB || ctx->Ebg || ctx->Bcbg)
        return fail(ctx, IPIC_E_STATE, "the field operator needs a uniform background field");
    {  // J (when the partials are newer), the state before the phase and the network input
        FieldArgs F0 = field_args(ctx);
        const int from = ctx->j_summed ? 0 : 1;
        if (from)
            if (int rc = slice_sum(ctx, F0)) return rc;
        const ipic_nop_scales& sc = ctx->op_scales;
        hipLaunchKernelGGL(k_op_prep, dim3((unsigned)((g.ncell + 255) / 256)), dim3(256), 0, ctx->stream, F0, from,
                           ctx->Eb, ctx->Bb, 1.0f / sc.in_E, 1.0f / sc.in_B, 1.0f / sc.in_J, ctx->op_u);
        HIPCK(ctx, hipGetLastError());
        ctx->j_summed = true;
    }
    const int32_t n3[3] = {g.n[0], g.n[1], g.n[2]};
    const float iw[3] = {g.inv_w[0], g.inv_w[1], g.inv_w[2]};
    if (int rc = ipic_nop_forward(n3, ctx->op_u, ctx->op_wf, ctx->op_params, ctx->op_ws, 0, ctx->op_out, ctx->stream))
        return nop_fail(ctx, rc);
    if (int rc = ipic_nop_apply(n3, ctx->op_out, &ctx->op_scales, iw, (float*)ctx->E, (float*)ctx->Bc,
                                (float*)ctx->op_psi, ctx->stream))
        return nop_fail(ctx, rc);
    // checks
    FieldArgs F = field_args(ctx);
    hipLaunchKernelGGL(k_op_check, dim3(ctx->red_blocks), dim3(256), 0, ctx->stream, F, ctx->Eb, ctx->Bb, ctx->red);
    HIPCK(ctx, hipGetLastError());
    const double* h = ctx->hred;
    HIPCK(ctx, hipMemcpyAsync(ctx->hred, ctx->red, (size_t)ctx->red_blocks * 5 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
    if (int rc = wait_stream(ctx)) return rc;
    double wo = 0, wn = 0, ejo = 0, ejn = 0, md = 0;
    for (int b = 0; b < ctx->red_blocks; ++b) {
        wo += h[5 * b];
        wn += h[5 * b + 1];
        ejo += h[5 * b + 2];
        ejn += h[5 * b + 3];
        md = std::max(md, h[5 * b + 4]);
    }
    const double vol = ctx->d.cell_width[0] * ctx->d.cell_width[1] * ctx->d.cell_width[2];
    const double W0 = 0.5 * wo * vol, W1 = 0.5 * wn * vol;
    const double eres = std::fabs(W1 - W0 + ctx->d.dt * 0.5 * (ejo + ejn) * vol) / std::max(std::max(W0, W1), 1e-300);
    const bool ok = md <= ctx->op.divb_tol && eres <= ctx->op.energy_tol;  // false for NaN
    if (st) {
        st->max_divb_residual = std::max(st->max_divb_residual, std::isfinite(md) ? md : HUGE_VAL);
        st->max_energy_residual = std::max(st->max_energy_residual, std::isfinite(eres) ? eres : HUGE_VAL);
    }
    ++ctx->op_count;
    if (!ok) {
        if (int rc = copy_f4(ctx, ctx->E, ctx->Eb)) return rc;
        if (int rc = copy_f4(ctx, ctx->Bc, ctx->Bb)) return rc;
        if (int rc = launch_leapfrog(ctx)) return rc;
        if (st) st->fallback_steps++;
        return 0;
    }
    if (ctx->op.audit_every > 0 && ctx->op_count % ctx->op.audit_every == 0) {
        if (int rc = copy_f4(ctx, ctx->Eo, ctx->E)) return rc;
        if (int rc = copy_f4(ctx, ctx->Bo, ctx->Bc)) return rc;
        if (int rc = copy_f4(ctx, ctx->E, ctx->Eb)) return rc;
        if (int rc = copy_f4(ctx, ctx->Bc, ctx->Bb)) return rc;
        if (int rc = launch_leapfrog(ctx)) return rc;
        hipLaunchKernelGGL(k_audit, dim3(ctx->red_blocks), dim3(256), 0, ctx->stream, g, F.c2, ctx->Eo, ctx->Bo,
                           ctx->E, ctx->Bc, ctx->Eb, ctx->Bb, ctx->red);
        HIPCK(ctx, hipGetLastError());
        HIPCK(ctx, hipMemcpyAsync(ctx->hred, ctx->red, (size_t)ctx->red_blocks * 4 * sizeof(double),
                                  hipMemcpyDeviceToHost, ctx->stream));
        if (int rc = wait_stream(ctx)) return rc;
        double num = 0, den = 0, tot = 0, evo = 0;
        for (int b = 0; b < ctx->red_blocks; ++b) {
            num += h[4 * b];
            den += h[4 * b + 1];
            tot += h[4 * b + 2];
            evo += h[4 * b + 3];
        }
        auto ratio = [](double a, double b) { return b > 0 ? std::sqrt(a / b) : (a > 0 ? HUGE_VAL : 0.0); };
        const double err = ratio(num, tot), err_inc = ratio(num, den), err_evo = ratio(num, evo);
        auto fin = [](double v) { return std::isfinite(v) ? v : HUGE_VAL; };
        if (st) {
            st->audits++;
            st->max_audit_error = std::max(st->max_audit_error, fin(err));
            st->max_audit_error_increment = std::max(st->max_audit_error_increment, fin(err_inc));
            st->max_audit_error_evolving = std::max(st->max_audit_error_evolving, fin(err_evo));
        }
        // the gate is the error of the evolving field: the uniform background, which the
        // operator cannot change, does not dilute it
        if (err_evo <= ctx->op.audit_tol) {
            if (int rc = copy_f4(ctx, ctx->E, ctx->Eo)) return rc;
            if (int rc = copy_f4(ctx, ctx->Bc, ctx->Bo)) return rc;
        } else {
            if (st) st->fallback_steps++;
            return 0;
        }
    }
    if (st) st->operator_steps++;
    return 0;
}

double ev_ms(ipic_ctx* ctx, int a, int b) {
    float ms = 0.0f;
    if (hipEventSynchronize(ctx->ev[b]) != hipSuccess) return 0.0;
    if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) != hipSuccess) return 0.0;
    return ms;
}

// Field phase of a step: initial deposit when needed, leapfrog or operator, interpC2N.
int field_phase(ipic_ctx* ctx, bool use_op, ipic_step_stats* st) {
    const bool timing = (ctx->d.flags & IPIC_FLAG_PHASE_TIMING) != 0;
    if (!ctx->j_valid) {
        if (timing) HIPCK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
        if (int rc = launch_particles(ctx, MODE_DEPOSIT)) return rc;
        if (timing) {
            HIPCK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
            if (st) st->ms_deposit += ev_ms(ctx, 0, 1);
        }
        ctx->j_valid = true;
    }
    if (timing) HIPCK(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
    if (use_op && ctx->op_on) {
        if (int rc = operator_phase(ctx, st)) return rc;
    } else {
        if (int rc = launch_leapfrog(ctx)) return rc;
    }
    if (int rc = launch_c2n(ctx)) return rc;
    if (timing) {
        HIPCK(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
        if (st) st->ms_field += ev_ms(ctx, 2, 3);
    }
    return 0;
}

// Particle phase: fused push kernel, then the cell-count scans.
int particle_phase(ipic_ctx* ctx, ipic_step_stats* st) {
    const bool timing = (ctx->d.flags & IPIC_FLAG_PHASE_TIMING) != 0;
    if (timing) HIPCK(ctx, hipEventRecord(ctx->ev[4], ctx->stream));
    if (int rc = launch_particles(ctx, MODE_PUSH)) return rc;
    if (timing) HIPCK(ctx, hipEventRecord(ctx->ev[5], ctx->stream));
    if (int rc = launch_canon(ctx)) return rc;
    if (int rc = launch_scan(ctx)) return rc;
    if (timing) {
        HIPCK(ctx, hipEventRecord(ctx->ev[6], ctx->stream));
        if (st) {
            st->ms_push += ev_ms(ctx, 4, 5);
            st->ms_scan += ev_ms(ctx, 5, 6);
        }
    }
    ctx->canonical = false;
    ctx->j_valid = true;  // the push deposited the J of the new state
    return 0;
}

int check_device_flag(ipic_ctx* ctx);

int canonicalize(ipic_ctx* ctx) {
    if (ctx->canonical) return 0;
    int rc = launch_particles(ctx, MODE_REBIN);
    if (!rc) rc = launch_canon(ctx);
    if (rc) return rc;
    ScanArgs A = scan_args(ctx);
    hipLaunchKernelGGL(k_canon_fix, dim3((ctx->g.ncell + 1 + 255) / 256, ctx->d.nspecies), dim3(256), 0,
                       ctx->stream, A);
    HIPCK(ctx, hipGetLastError());
    ctx->canonical = true;
    return check_device_flag(ctx);
}

// live particles of species s (the layout the next pass writes)
int read_total(ipic_ctx* ctx, int s, long long* total) {
    int v = 0;
    HIPCK(ctx, hipMemcpyAsync(&v, ctx->sp[s].start + ctx->g.ncell, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(ctx, hipStreamSynchronize(ctx->stream));
    *total = v;
    return 0;
}

// slots of the current particle array (its segments), including slots of particles the
// sink removed during the pass that wrote it
int read_slots(ipic_ctx* ctx, int s, long long* slots) {
    int v = 0;
    HIPCK(ctx, hipMemcpyAsync(&v, ctx->sp[s].seg + ctx->g.ncell, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(ctx, hipStreamSynchronize(ctx->stream));
    *slots = v;
    return 0;
}

// cell coordinates of a tile-major cell id
inline void cid_xyz(const Geo& g, int cid, int& x, int& y, int& z) {
    const int t = cid / g.TC, l = cid % g.TC;
    const int tx = t / (g.NTl[1] * g.NTl[2]), ty = (t / g.NTl[2]) % g.NTl[1], tz = t % g.NTl[2];
    x = tx * g.T[0] + l / (g.T[1] * g.T[2]);
    y = ty * g.T[1] + (l / g.T[2]) % g.T[1];
    z = tz * g.T[2] + l % g.T[2];
}

inline int host_cid(const Geo& g, int x, int y, int z) {
    const int tx = x / g.T[0], ty = y / g.T[1], tz = z / g.T[2];
    const int lx = x - tx * g.T[0], ly = y - ty * g.T[1], lz = z - tz * g.T[2];
    return ((tx * g.NTl[1] + ty) * g.NTl[2] + tz) * g.TC + (lx * g.T[1] + ly) * g.T[2] + lz;
}

int alloc_species_buffers(ipic_ctx* ctx, int s, long long cap) {
    Species& S = ctx->sp[s];
    // the six arrays of a species are addressed through one buffer resource (< 4 GiB); checked
    // before anything is freed, so a refused size leaves the species as it was
    if (std::max<long long>(cap, 1) * 24 >= (1LL << 32))
        return fail(ctx, IPIC_E_RANGE, "a species holds at most 178956970 particles");
    if (cap > S.cap) {
        for (int b = 0; b < 2; ++b) {
            dfree(ctx, S.X[b]);
            dfree(ctx, S.ids[b]);
            S.X[b] = nullptr;
            S.ids[b] = nullptr;
        }
        S.cap = 0;  // no buffers until both are allocated
        const long long c = std::max<long long>(cap, 1);
        for (int b = 0; b < 2; ++b) {
            if (int rc = dalloc(ctx, &S.X[b], (size_t)6 * c)) return rc;
            if (ctx->track)
                if (int rc = dalloc(ctx, &S.ids[b], (size_t)c)) return rc;
        }
        S.cap = c;
    }
    S.cur = 0;
    return 0;
}

// set the per-cell bookkeeping of a freshly sorted (canonical) array from host starts
int set_layout(ipic_ctx* ctx, int s, const std::vector<int>& start) {
    Species& S = ctx->sp[s];
    ctx->order_valid = false;  // the tile loads of the new layout are not known on the device
    const int nc = ctx->g.ncell;
    std::vector<int> stay(nc);
    for (int c = 0; c < nc; ++c) stay[c] = start[c + 1] - start[c];
    HIPCK(ctx, hipMemcpyAsync(S.seg, start.data(), (nc + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(ctx, hipMemcpyAsync(S.start, start.data(), (nc + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(ctx, hipMemcpyAsync(S.stay, stay.data(), nc * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    HIPCK(ctx, hipMemsetAsync(S.cnt_acc, 0, nc * sizeof(int), ctx->stream));
    HIPCK(ctx, hipMemsetAsync(S.stay_acc, 0, nc * sizeof(int), ctx->stream));
    HIPCK(ctx, hipMemsetAsync(S.arr, 0, nc * sizeof(int), ctx->stream));
    HIPCK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// total E (nodes) and Bc (centres) in FP64: background + device perturbation
int total_fields(ipic_ctx* ctx, double* E, double* Bc) {
    const int nc = ctx->g.ncell;
    std::vector<float4> h(nc);
    HIPCK(ctx, hipStreamSynchronize(ctx->stream));
    for (int f = 0; f < 2; ++f) {
        double* dst = f == 0 ? E : Bc;
        if (!dst) continue;
        const std::vector<double>& bg = f == 0 ? ctx->bgE : ctx->bgBc;
        HIPCK(ctx, hipMemcpy(h.data(), f == 0 ? ctx->E : ctx->Bc, nc * sizeof(float4), hipMemcpyDeviceToHost));
        for (int i = 0; i < nc; ++i) {
            const double b0 = bg.empty() ? 0.0 : bg[3 * i], b1 = bg.empty() ? 0.0 : bg[3 * i + 1],
                         b2 = bg.empty() ? 0.0 : bg[3 * i + 2];
            dst[3 * i] = b0 + h[i].x;
            dst[3 * i + 1] = b1 + h[i].y;
            dst[3 * i + 2] = b2 + h[i].z;
        }
    }
    return 0;
}

// reads and clears the device consistency flag set by the particle kernel guards
int check_device_flag(ipic_ctx* ctx) {
    int v = 0;
    HIPCK(ctx, hipMemcpyAsync(&v, ctx->derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCK(ctx, hipStreamSynchronize(ctx->stream));
    if (v) {
        const int z = 0;
        HIPCK(ctx, hipMemcpy(ctx->derr, &z, sizeof(int), hipMemcpyHostToDevice));
        if (v == ERR_JSAT)
            return fail(ctx, IPIC_E_STATE,
                        "current density out of the fixed-point range (a particle's q v exceeds 2^20 q c): "
                        "J is clamped and the state is invalid");
        return fail(ctx, IPIC_E_STATE,
                    std::string("particle layout guard fired (code ") + std::to_string(v) +
                        "): per-cell counts inconsistent; state is invalid");
    }
    return 0;
}

int check_ctx(ipic_ctx* ctx) {
    if (!ctx) return fail(nullptr, IPIC_E_ARG, "null context");
    return 0;
}

// Runs the body of a C-ABI entry point: no C++ exception (a failed host allocation, for
// instance) crosses the ABI; it becomes an error code with a message instead.
template <class F>
int guarded(ipic_ctx* ctx, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(ctx, IPIC_E_RANGE, "out of host memory");
    } catch (const std::exception& e) {
        return fail(ctx, IPIC_E_HIP, std::string("internal error: ") + e.what());
    } catch (...) {
        return fail(ctx, IPIC_E_HIP, "internal error");
    }
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================

extern "C" {

const char* ipic_gpu_last_error(const ipic_ctx* ctx) {
    if (ctx && !ctx->err.empty()) return ctx->err.c_str();
    return g_last_error.c_str();
}

int ipic_gpu_device_info(char* name, int32_t name_len, int32_t* cu_count, int64_t* hbm_bytes) {
    int dev = 0;
    hipDeviceProp_t p;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&p, dev);
    if (e != hipSuccess) return fail(nullptr, IPIC_E_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    if (name && name_len > 0) {
        std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
    return 0;
}

int ipic_gpu_create(const ipic_desc* desc, ipic_ctx** out) {
    return guarded(nullptr, [&]() -> int {
        if (!desc || !out) return fail(nullptr, IPIC_E_ARG, "null argument");
        *out = nullptr;
        const ipic_desc& d = *desc;
        for (int k = 0; k < 3; ++k) {
            if (d.n[k] <= 0) return fail(nullptr, IPIC_E_ARG, "grid size must be positive");
            // cell coordinates go through fdivi (exact below 2^21)
            if (d.n[k] > (1 << 21)) return fail(nullptr, IPIC_E_RANGE, "at most 2097152 cells per axis");
            if (!(d.cell_width[k] > 0)) return fail(nullptr, IPIC_E_ARG, "cell width must be positive");
        }
        if (!(d.dt > 0) || !(d.c > 0)) return fail(nullptr, IPIC_E_ARG, "dt and c must be positive");
        if (d.nspecies < 1 || d.nspecies > IPIC_MAX_SPECIES) return fail(nullptr, IPIC_E_ARG, "nspecies out of range");
        const long long ncell = (long long)d.n[0] * d.n[1] * d.n[2];
        if (ncell > (1LL << 30)) return fail(nullptr, IPIC_E_RANGE, "grid too large");

        ipic_ctx* ctx = new ipic_ctx();
        ctx->d = d;
        if (ctx->d.field_substeps < 0) {
            delete ctx;
            return fail(nullptr, IPIC_E_ARG, "field_substeps must be >= 0 (0: static field)");
        }
        ctx->track = (d.flags & IPIC_FLAG_TRACK_IDS) != 0;
        Geo& g = ctx->g;
        for (int k = 0; k < 3; ++k) {
            g.n[k] = d.n[k];
            int T = d.tile[k] > 0 ? d.tile[k] : largest_divisor_le(d.n[k], 8);
            if (T > 8 || d.n[k] % T != 0) {
                delete ctx;
                return fail(nullptr, IPIC_E_ARG, "tile size must divide the grid and be at most 8");
            }
            g.T[k] = T;
            g.NTl[k] = d.n[k] / T;
            g.jh = IPIC_JHALO;
            g.nj[k] = T + 1 + 2 * g.jh;
            g.w[k] = (float)d.cell_width[k];
            g.inv_w[k] = (float)(1.0 / d.cell_width[k]);
            g.origin[k] = (float)d.origin[k];
            g.sink_c[k] = (float)d.sink_center[k];
        }
        g.TC = g.T[0] * g.T[1] * g.T[2];
        g.ntiles = g.NTl[0] * g.NTl[1] * g.NTl[2];
        {
            // One block per tile leaves most CUs idle on a grid of few tiles (8^3 cells: one tile).
            // The particle kernel then splits each tile's cells among several blocks, up to two
            // blocks per CU and at least as many cells per block as it has waves.
            int cus = 256, dev = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
                cus = 256;
            const int want = (2 * cus + g.ntiles - 1) / g.ntiles;
            g.nsl = std::max(1, std::min(want, g.TC / NW));
        }
        g.ncell = (int)ncell;
        g.NJ = g.nj[0] * g.nj[1] * g.nj[2];
        g.nmax = std::max(g.n[0], std::max(g.n[1], g.n[2]));
        g.s0 = g.n[1] * g.n[2];
        g.dt = (float)d.dt;
        g.inv_vol = (float)(1.0 / (d.cell_width[0] * d.cell_width[1] * d.cell_width[2]));
        g.pi_c = (float)(M_PI * d.c);
        for (int k = 0; k < 3; ++k) g.inv_T[k] = 1.0f / (float)g.T[k];
        g.inv_T12 = 1.0f / (float)(g.T[1] * g.T[2]);
        {
            // fixed-point scale of the LDS J accumulator: one particle at speed c contributes about
            // 2^30 units, leaving 2^33 of headroom for the sums of a tile node
            const double vol = d.cell_width[0] * d.cell_width[1] * d.cell_width[2];
            double qmax = 0.0;
            for (int s = 0; s < d.nspecies; ++s) qmax = std::max(qmax, std::fabs(d.species[s].q) / (8.0 * vol));
            int bits = 40;
            if (qmax * d.c > 0.0 && std::isfinite(qmax * d.c)) bits = 30 - (int)std::ceil(std::log2(qmax * d.c));
            bits = std::max(-60, std::min(120, bits));
            g.fix_lim = (float)std::ldexp(1.0, 50 - bits);
            g.fix_scale = std::ldexp(1.0, bits);
            g.fix_inv = std::ldexp(1.0, -bits);
        }
        g.sink = d.sink ? 1 : 0;
        g.reflect = (d.flags & IPIC_FLAG_REFLECT) ? 1 : 0;
        {
            const double L = d.cell_width[0];
            g.dip_m = (float)(d.dipole_moment / (L * L * L));
            for (int k = 0; k < 3; ++k) {
                g.dip_o[k] = (float)(d.origin[k] / L);
                g.dip_w[k] = (float)(d.cell_width[k] / L);
            }
        }
        g.sink_r2 = (float)(d.sink_radius * d.sink_radius);

        std::vector<int> jp;
        if (build_jpairs(g, jp) != 0) {
            delete ctx;
            return fail(nullptr, IPIC_E_ARG, "tile/grid combination not supported");
        }
        {
            // multiplier of the scrambled tile order: near 0.618 ntiles, so that consecutive
            // blocks take tiles far apart, and coprime to ntiles, so that b -> b * mul mod
            // ntiles is a permutation
            auto gcd = [](int a, int b) {
                while (b) {
                    const int r = a % b;
                    a = b;
                    b = r;
                }
                return a;
            };
            int m = std::max(1, (int)(0.6180339887 * g.ntiles));
            while (gcd(m, g.ntiles) != 1) ++m;
            ctx->order_mul = m;
        }
        hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete ctx;
            return fail(nullptr, IPIC_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        int rc = 0;
        const size_t nc = (size_t)g.ncell;
        rc = rc ? rc : dalloc(ctx, &ctx->E, nc);
        rc = rc ? rc : dalloc(ctx, &ctx->Bc, nc);
        rc = rc ? rc : dalloc(ctx, &ctx->Bn, nc);
        rc = rc ? rc : dalloc(ctx, &ctx->J, nc);
        rc = rc ? rc : dalloc(ctx, &ctx->jpart, (size_t)g.ntiles * g.nsl * 3 * g.NJ);
        rc = rc ? rc : dalloc(ctx, &ctx->jover, nc * 3);
        rc = rc ? rc : dalloc(ctx, &ctx->jmark, nc);
        rc = rc ? rc : dalloc(ctx, &ctx->jpairs, jp.size());
        rc = rc ? rc : dalloc(ctx, &ctx->removed, 1);
        rc = rc ? rc : dalloc(ctx, &ctx->derr, 1);
        rc = rc ? rc : dalloc(ctx, &ctx->tile_order, (size_t)g.ntiles);
        rc = rc ? rc : dalloc(ctx, &ctx->red, (size_t)ctx->red_blocks * 5);
        for (int s = 0; s < d.nspecies && !rc; ++s) {
            Species& S = ctx->sp[s];
            S.q = d.species[s].q;
            S.qom = d.species[s].qom;
            rc = rc ? rc : dalloc(ctx, &S.seg, nc + 1);
            rc = rc ? rc : dalloc(ctx, &S.start, nc + 1);
            rc = rc ? rc : dalloc(ctx, &S.stay, nc);
            rc = rc ? rc : dalloc(ctx, &S.cnt_acc, nc);
            rc = rc ? rc : dalloc(ctx, &S.stay_acc, nc);
            rc = rc ? rc : dalloc(ctx, &S.arr, nc);
            rc = rc ? rc : dalloc(ctx, &S.tile_total, (size_t)g.ntiles);
            rc = rc ? rc : dalloc(ctx, &S.tile_base, (size_t)g.ntiles + 1);
            if (d.flags & IPIC_FLAG_REPRODUCIBLE) rc = rc ? rc : dalloc(ctx, &S.work, nc + 1);
        }
        if (!rc) {
            e = hipMemcpyAsync(ctx->jpairs, jp.data(), jp.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) rc = fail(ctx, IPIC_E_HIP, std::string("upload: ") + hipGetErrorString(e));
        }
        if (!rc)
            for (int k = 0; k < 8; ++k)
                if (hipEventCreate(&ctx->ev[k]) != hipSuccess) rc = fail(ctx, IPIC_E_HIP, "hipEventCreate");
        if (!rc && hipEventCreateWithFlags(&ctx->ev_wait, hipEventDisableTiming) != hipSuccess)
            rc = fail(ctx, IPIC_E_HIP, "hipEventCreateWithFlags");
        if (!rc && hipHostMalloc((void**)&ctx->hred, (size_t)ctx->red_blocks * 5 * sizeof(double)) != hipSuccess)
            rc = fail(ctx, IPIC_E_HIP, "hipHostMalloc");
        if (rc) {
            std::string msg = ctx->err;
            ipic_gpu_destroy(ctx);
            g_last_error = msg;
            return rc;
        }
        *out = ctx;
        return 0;
    });
}

int ipic_gpu_destroy(ipic_ctx* ctx) {
    return guarded(ctx, [&]() -> int {
        if (!ctx) return 0;
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        for (void* p : ctx->allocs) (void)hipFree(p);
        for (int k = 0; k < 8; ++k)
            if (ctx->ev[k]) (void)hipEventDestroy(ctx->ev[k]);
        if (ctx->ev_wait) (void)hipEventDestroy(ctx->ev_wait);
        if (ctx->hred) (void)hipHostFree(ctx->hred);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return 0;
    });
}

int ipic_gpu_upload_particles(ipic_ctx* ctx, int32_t species, int64_t n, const double* pos, const double* vel) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (species < 0 || species >= ctx->d.nspecies) return fail(ctx, IPIC_E_ARG, "species out of range");
        if (n < 0 || (n > 0 && (!pos || !vel))) return fail(ctx, IPIC_E_ARG, "invalid particle arrays");
        // the limit of alloc_species_buffers (one buffer resource spans the six arrays, < 4 GiB),
        // checked before any host buffer is built
        if (std::max<int64_t>(n, 1) * 24 >= (1LL << 32))
            return fail(ctx, IPIC_E_RANGE, "a species holds at most 178956970 particles");
        if (int rc = canonicalize(ctx)) return rc;
        const Geo& g = ctx->g;
        const int nc = g.ncell;
        std::vector<int> cell((size_t)n);
        std::vector<float> off((size_t)n * 3);
        std::vector<int> count(nc + 1, 0);
        for (int64_t i = 0; i < n; ++i) {
            int ci[3];
            for (int k = 0; k < 3; ++k) {
                const double L = (double)g.n[k] * ctx->d.cell_width[k];
                double x = pos[3 * i + k] - ctx->d.origin[k];
                if (!(std::isfinite(x))) return fail(ctx, IPIC_E_ARG, "non-finite particle position");
                x = std::fmod(x, L);
                if (x < 0) x += L;
                double qd = x / ctx->d.cell_width[k];
                long long cc = (long long)std::floor(qd);
                if (cc >= g.n[k]) cc = g.n[k] - 1;
                if (cc < 0) cc = 0;
                ci[k] = (int)cc;
                float o = (float)(qd - (double)cc);
                if (!(o > 0.0f)) o = 0.0f;  // also -0.0 -> +0.0: the kernel's cell test reads bit patterns
                if (o > 1.0f) o = 1.0f;
                off[3 * i + k] = o;
            }
            for (int k = 0; k < 3; ++k)
            if (!std::isfinite(vel[3 * i + k])) return fail(ctx, IPIC_E_ARG, "non-finite particle velocity");
        cell[i] = host_cid(g, ci[0], ci[1], ci[2]);
            count[cell[i] + 1]++;
        }
        std::vector<int> start(nc + 1, 0);
        for (int c = 0; c < nc; ++c) start[c + 1] = start[c] + count[c + 1];
        std::vector<int> fillp(start.begin(), start.end() - 1);
        const long long cap = std::max<long long>(n, 1);
        std::vector<float> X((size_t)6 * cap, 0.0f);
        std::vector<uint32_t> ids(ctx->track ? cap : 0);
        for (int64_t i = 0; i < n; ++i) {
            const int p = fillp[cell[i]]++;
            X[p] = off[3 * i];
            X[cap + p] = off[3 * i + 1];
            X[2 * cap + p] = off[3 * i + 2];
            X[3 * cap + p] = (float)vel[3 * i];
            X[4 * cap + p] = (float)vel[3 * i + 1];
            X[5 * cap + p] = (float)vel[3 * i + 2];
            if (ctx->track) ids[p] = (uint32_t)i;
        }
        if (int rc = alloc_species_buffers(ctx, species, cap)) return rc;
        Species& S = ctx->sp[species];
        if (S.cap != cap) {
            // keep the array stride of the allocation
            std::vector<float> Y((size_t)6 * S.cap, 0.0f);
            for (int k = 0; k < 6; ++k) std::memcpy(&Y[(size_t)k * S.cap], &X[(size_t)k * cap], (size_t)n * sizeof(float));
            X.swap(Y);
        }
        HIPCK(ctx, hipMemcpyAsync(S.X[S.cur], X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        if (ctx->track && n > 0)
            HIPCK(ctx, hipMemcpyAsync(S.ids[S.cur], ids.data(), (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        if (int rc = set_layout(ctx, species, start)) return rc;
        ctx->j_valid = false;
        return 0;
    });
}

int ipic_gpu_init_reference(ipic_ctx* ctx, int32_t species, const ipic_ref_distribution* dist, int64_t* counts) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (species < 0 || species >= ctx->d.nspecies) return fail(ctx, IPIC_E_ARG, "species out of range");
        const int64_t nc = (int64_t)ctx->d.n[0] * ctx->d.n[1] * ctx->d.n[2];
        std::vector<int64_t> cnt((size_t)nc);
        if (int rc = ipic_ref_cell_counts(&ctx->d, dist, cnt.data())) return fail(ctx, rc, ipic_ref_last_error());
        int64_t total = 0;
        for (int64_t c = 0; c < nc; ++c) total += cnt[(size_t)c];
        if (total >= (1LL << 31)) return fail(ctx, IPIC_E_RANGE, "at most 2^31-1 particles per species");
        std::vector<double> pos, vel;
        try {
            pos.resize((size_t)total * 3);
            vel.resize((size_t)total * 3);
        } catch (const std::exception&) {
            return fail(ctx, IPIC_E_RANGE, "host memory for the reference initialisation");
        }
        if (int rc = ipic_ref_particles(&ctx->d, dist, cnt.data(), pos.data(), vel.data()))
            return fail(ctx, rc, ipic_ref_last_error());
        if (counts) std::memcpy(counts, cnt.data(), (size_t)nc * sizeof(int64_t));
        return ipic_gpu_upload_particles(ctx, species, total, pos.data(), vel.data());
    });
}

int ipic_gpu_init_lattice(ipic_ctx* ctx, int32_t species, const ipic_ref_lattice* lat) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (species < 0 || species >= ctx->d.nspecies) return fail(ctx, IPIC_E_ARG, "species out of range");
        if (!lat) return fail(ctx, IPIC_E_ARG, "null lattice description");
        const int64_t nc = (int64_t)ctx->d.n[0] * ctx->d.n[1] * ctx->d.n[2];
        int64_t per = 1;
        for (int k = 0; k < 3; ++k) {
            if (lat->ppc[k] <= 0) return fail(ctx, IPIC_E_ARG, "particles per cell must be positive on every axis");
            per *= lat->ppc[k];
        }
        if (per >= (1LL << 31) / nc) return fail(ctx, IPIC_E_RANGE, "at most 2^31-1 particles per species");
        const int64_t total = nc * per;
        std::vector<double> pos((size_t)total * 3), vel((size_t)total * 3);
        if (int rc = ipic_ref_lattice_particles(&ctx->d, lat, pos.data(), vel.data()))
            return fail(ctx, rc, ipic_ref_last_error());
        return ipic_gpu_upload_particles(ctx, species, total, pos.data(), vel.data());
    });
}

int ipic_gpu_init_particles(ipic_ctx* ctx, int32_t species, const int32_t* ppc, int32_t uniform_ppc,
                            const double vth[3], const double drift[3], uint64_t seed) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (species < 0 || species >= ctx->d.nspecies) return fail(ctx, IPIC_E_ARG, "species out of range");
        if (!vth || !drift) return fail(ctx, IPIC_E_ARG, "null vth/drift");
        if (!ppc && uniform_ppc < 0) return fail(ctx, IPIC_E_ARG, "negative particles per cell");
        if (int rc = canonicalize(ctx)) return rc;
        const Geo& g = ctx->g;
        const int nc = g.ncell;
        std::vector<int> start(nc + 1, 0);
        long long total = 0;
        for (int cid = 0; cid < nc; ++cid) {
            int x, y, z;
            cid_xyz(g, cid, x, y, z);
            const int row = (x * g.n[1] + y) * g.n[2] + z;
            const long long k = ppc ? ppc[row] : uniform_ppc;
            if (k < 0) return fail(ctx, IPIC_E_ARG, "negative particles per cell");
            total += k;
            if (total >= (1LL << 31)) return fail(ctx, IPIC_E_RANGE, "at most 2^31-1 particles per species");
            start[cid + 1] = (int)total;
        }
        if (int rc = alloc_species_buffers(ctx, species, std::max<long long>(total, 1))) return rc;
        Species& S = ctx->sp[species];
        if (int rc = set_layout(ctx, species, start)) return rc;
        if (total > 0) {
            const unsigned nb = (unsigned)((total + 255) / 256);
            hipLaunchKernelGGL(k_init, dim3(nb), dim3(256), 0, ctx->stream, g, S.start, S.X[S.cur],
                               ctx->track ? S.ids[S.cur] : nullptr, S.cap, total, seed, (int)species, vth[0], vth[1], vth[2],
                               drift[0], drift[1], drift[2]);
            HIPCK(ctx, hipGetLastError());
        }
        HIPCK(ctx, hipStreamSynchronize(ctx->stream));
        ctx->j_valid = false;
        return 0;
    });
}

int ipic_gpu_upload_fields(ipic_ctx* ctx, const double* E, const double* Bc) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        const Geo& g = ctx->g;
        const int nc = g.ncell;
        const size_t n3 = (size_t)nc * 3;
        // total fields: the given ones, or the current state for a NULL argument
        std::vector<double> tE(n3), tB(n3);
        if (!E || !Bc)
            if (int rc = total_fields(ctx, E ? nullptr : tE.data(), Bc ? nullptr : tB.data())) return rc;
        if (E) std::memcpy(tE.data(), E, n3 * sizeof(double));
        if (Bc) std::memcpy(tB.data(), Bc, n3 * sizeof(double));
        for (size_t i = 0; i < n3; ++i)
            if (!std::isfinite(tE[i]) || !std::isfinite(tB[i])) return fail(ctx, IPIC_E_ARG, "non-finite field value");
        ctx->bgE.swap(tE);
        ctx->bgBc.swap(tB);
        // uniform parts: mean over the grid
        double e0[3] = {0, 0, 0}, b0[3] = {0, 0, 0};
        for (int i = 0; i < nc; ++i)
            for (int k = 0; k < 3; ++k) {
                e0[k] += ctx->bgE[3 * i + k];
                b0[k] += ctx->bgBc[3 * i + k];
            }
        for (int k = 0; k < 3; ++k) {
            e0[k] /= nc;
            b0[k] /= nc;
            ctx->g.E0[k] = (float)e0[k];
            ctx->g.B0[k] = (float)b0[k];
        }
        // non-uniform remainders relative to the FP32-rounded uniform parts
        std::vector<double> rE(n3), rB(n3);
        bool nuE = false, nuB = false;
        for (int i = 0; i < nc; ++i)
            for (int k = 0; k < 3; ++k) {
                rE[3 * i + k] = ctx->bgE[3 * i + k] - (double)ctx->g.E0[k];
                rB[3 * i + k] = ctx->bgBc[3 * i + k] - (double)ctx->g.B0[k];
                nuE |= rE[3 * i + k] != 0.0;
                nuB |= rB[3 * i + k] != 0.0;
            }
        if (ctx->d.flags & IPIC_FLAG_UNIFORM_BACKGROUND) {
            // only E0, B0 stay static; the remainders start the evolving FP32 state
            for (int i = 0; i < nc; ++i)
                for (int k = 0; k < 3; ++k) {
                    ctx->bgE[3 * i + k] = (double)ctx->g.E0[k];
                    ctx->bgBc[3 * i + k] = (double)ctx->g.B0[k];
                }
            std::vector<float4> he(nc), hb(nc);
            for (int i = 0; i < nc; ++i) {
                he[i] = make_float4((float)rE[3 * i], (float)rE[3 * i + 1], (float)rE[3 * i + 2], 0.0f);
                hb[i] = make_float4((float)rB[3 * i], (float)rB[3 * i + 1], (float)rB[3 * i + 2], 0.0f);
            }
            for (float4** a : {&ctx->Ebg, &ctx->Eg, &ctx->Bcbg, &ctx->Bnbg, &ctx->SE, &ctx->SB}) {
                dfree(ctx, *a);
                *a = nullptr;
            }
            HIPCK(ctx, hipMemcpyAsync(ctx->E, he.data(), nc * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
            HIPCK(ctx, hipMemcpyAsync(ctx->Bc, hb.data(), nc * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
            if (int rc = launch_c2n(ctx)) return rc;
            HIPCK(ctx, hipStreamSynchronize(ctx->stream));
            return 0;
        }
        // static sources of the background: SE = c^2 curl- Bc_bg (field.h:359-363),
        // SB = curl+ E_bg (field.h:367-371), FP64
        std::vector<double> sE(n3), sB(n3), rBn(n3);
        const int* n = g.n;
        const double* w = ctx->d.cell_width;
        // curl factors of the field policy: leapfrog c^2 and 1, forward solver c and c
        const bool fwd = (ctx->d.flags & IPIC_FLAG_FORWARD_SOLVER) != 0;
        const double c2 = fwd ? ctx->d.c : ctx->d.c * ctx->d.c;
        const double bF = fwd ? ctx->d.c : 1.0;
        auto id = [&](int x, int y, int z) {
            x = (x % n[0] + n[0]) % n[0];
            y = (y % n[1] + n[1]) % n[1];
            z = (z % n[2] + n[2]) % n[2];
            return (size_t)((x * n[1] + y) * n[2] + z);
        };
        bool hasSE = false, hasSB = false;
        const std::vector<double>& bB = ctx->bgBc;
        const std::vector<double>& bE = ctx->bgE;
        for (int x = 0; x < n[0]; ++x)
            for (int y = 0; y < n[1]; ++y)
                for (int z = 0; z < n[2]; ++z) {
                    const size_t i = id(x, y, z), mx = id(x - 1, y, z), my = id(x, y - 1, z), mz = id(x, y, z - 1);
                    const size_t px = id(x + 1, y, z), py = id(x, y + 1, z), pz = id(x, y, z + 1);
                    sE[3 * i + 0] = c2 * ((bB[3 * i + 2] - bB[3 * my + 2]) / w[1] - (bB[3 * i + 1] - bB[3 * mz + 1]) / w[2]);
                    sE[3 * i + 1] = c2 * ((bB[3 * i + 0] - bB[3 * mz + 0]) / w[2] - (bB[3 * i + 2] - bB[3 * mx + 2]) / w[0]);
                    sE[3 * i + 2] = c2 * ((bB[3 * i + 1] - bB[3 * mx + 1]) / w[0] - (bB[3 * i + 0] - bB[3 * my + 0]) / w[1]);
                    sB[3 * i + 0] = bF * ((bE[3 * py + 2] - bE[3 * i + 2]) / w[1] - (bE[3 * pz + 1] - bE[3 * i + 1]) / w[2]);
                    sB[3 * i + 1] = bF * ((bE[3 * pz + 0] - bE[3 * i + 0]) / w[2] - (bE[3 * px + 2] - bE[3 * i + 2]) / w[0]);
                    sB[3 * i + 2] = bF * ((bE[3 * px + 1] - bE[3 * i + 1]) / w[0] - (bE[3 * py + 0] - bE[3 * i + 0]) / w[1]);
                    for (int k = 0; k < 3; ++k) {
                        hasSE |= sE[3 * i + k] != 0.0;
                        hasSB |= sB[3 * i + k] != 0.0;
                    }
                    for (int k = 0; k < 3; ++k) {
                        double acc = 0;
                        for (int a = 0; a < 2; ++a)
                            for (int b = 0; b < 2; ++b)
                                for (int c = 0; c < 2; ++c) acc += rB[3 * id(x - a, y - b, z - c) + k];
                        rBn[3 * i + k] = 0.125 * acc;
                    }
                }
        auto put = [&](float4** dst, bool need, const std::vector<double>& src) -> int {
            if (!need) {
                dfree(ctx, *dst);
                *dst = nullptr;
                return 0;
            }
            if (!*dst)
                if (int rc = dalloc(ctx, dst, (size_t)nc)) return rc;
            std::vector<float4> h(nc);
            for (int i = 0; i < nc; ++i)
                h[i] = make_float4((float)src[3 * i], (float)src[3 * i + 1], (float)src[3 * i + 2], 0.0f);
            HIPCK(ctx, hipMemcpyAsync(*dst, h.data(), nc * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
            HIPCK(ctx, hipStreamSynchronize(ctx->stream));
            return 0;
        };
        if (int rc = put(&ctx->Ebg, nuE, rE)) return rc;
        if (int rc = put(&ctx->Eg, nuE, rE)) return rc;  // gather E = dE + Ebg with dE = 0
        if (int rc = put(&ctx->Bcbg, nuB, rB)) return rc;
        if (int rc = put(&ctx->Bnbg, nuB, rBn)) return rc;
        if (int rc = put(&ctx->SE, hasSE, sE)) return rc;
        if (int rc = put(&ctx->SB, hasSB, sB)) return rc;
        // perturbations restart from zero; gather B = interpC2N(Bcbg)
        HIPCK(ctx, hipMemsetAsync(ctx->E, 0, nc * sizeof(float4), ctx->stream));
        HIPCK(ctx, hipMemsetAsync(ctx->Bc, 0, nc * sizeof(float4), ctx->stream));
        HIPCK(ctx, hipMemsetAsync(ctx->Bn, 0, nc * sizeof(float4), ctx->stream));
        if (nuB) HIPCK(ctx, hipMemcpyAsync(ctx->Bn, ctx->Bnbg, nc * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
        HIPCK(ctx, hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int ipic_gpu_init_fields(ipic_ctx* ctx, const ipic_field_init* init, double energies[2]) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (!init) return fail(ctx, IPIC_E_ARG, "null field initialisation");
        const size_t n3 = (size_t)ctx->g.ncell * 3;
        std::vector<double> E(n3), Bc(n3);
        if (int rc = ipic_field_init_host(&ctx->d, init, E.data(), nullptr, Bc.data(), energies))
            return fail(ctx, rc, ipic_field_init_last_error());
        return ipic_gpu_upload_fields(ctx, E.data(), Bc.data());
    });
}

int ipic_gpu_step(ipic_ctx* ctx, int64_t nsteps, ipic_step_stats* stats) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (nsteps < 0) return fail(ctx, IPIC_E_ARG, "negative step count");
        ipic_step_stats st{};
        HIPCK(ctx, hipMemsetAsync(ctx->removed, 0, sizeof(unsigned long long), ctx->stream));
        HIPCK(ctx, hipStreamSynchronize(ctx->stream));
        auto t0 = std::chrono::steady_clock::now();
        auto t1 = t0;
        for (int64_t s = 0; s < nsteps; ++s) {
            if (int rc = field_phase(ctx, true, &st)) return rc;
            if (int rc = particle_phase(ctx, &st)) return rc;
            if (s == 0) {
                HIPCK(ctx, hipStreamSynchronize(ctx->stream));
                t1 = std::chrono::steady_clock::now();
            }
        }
        HIPCK(ctx, hipStreamSynchronize(ctx->stream));
        auto t2 = std::chrono::steady_clock::now();
        if (int rc = check_device_flag(ctx)) return rc;
        unsigned long long removed = 0;
        HIPCK(ctx, hipMemcpy(&removed, ctx->removed, sizeof(removed), hipMemcpyDeviceToHost));
        st.first_step_s = std::chrono::duration<double>(t1 - t0).count();
        st.remaining_steps_s = std::chrono::duration<double>(t2 - t1).count();
        st.steps = nsteps;
        st.removed = (int64_t)removed;
        long long tot = 0;
        for (int s = 0; s < ctx->d.nspecies; ++s) {
            long long t = 0;
            if (int rc = read_total(ctx, s, &t)) return rc;
            tot += t;
        }
        st.particles = tot;
        if (stats) *stats = st;
        return 0;
    });
}

int ipic_gpu_field_step(ipic_ctx* ctx, int32_t use_operator, ipic_step_stats* stats) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (use_operator && !ctx->op_on) return fail(ctx, IPIC_E_STATE, "no field operator set");
        ipic_step_stats st{};
        if (int rc = field_phase(ctx, use_operator != 0, &st)) return rc;
        HIPCK(ctx, hipStreamSynchronize(ctx->stream));
        if (stats) {
            stats->ms_field += st.ms_field;
            stats->ms_deposit += st.ms_deposit;
            stats->operator_steps += st.operator_steps;
            stats->fallback_steps += st.fallback_steps;
            stats->audits += st.audits;
            stats->max_audit_error = std::max(stats->max_audit_error, st.max_audit_error);
            stats->max_audit_error_increment = std::max(stats->max_audit_error_increment, st.max_audit_error_increment);
            stats->max_audit_error_evolving = std::max(stats->max_audit_error_evolving, st.max_audit_error_evolving);
            stats->max_divb_residual = std::max(stats->max_divb_residual, st.max_divb_residual);
            stats->max_energy_residual = std::max(stats->max_energy_residual, st.max_energy_residual);
        }
        return 0;
    });
}

int ipic_gpu_particle_step(ipic_ctx* ctx, ipic_step_stats* stats) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        ipic_step_stats st{};
        HIPCK(ctx, hipMemsetAsync(ctx->removed, 0, sizeof(unsigned long long), ctx->stream));
        if (int rc = particle_phase(ctx, &st)) return rc;
        HIPCK(ctx, hipStreamSynchronize(ctx->stream));
        if (int rc = check_device_flag(ctx)) return rc;
        if (stats) {
            unsigned long long removed = 0;
            HIPCK(ctx, hipMemcpy(&removed, ctx->removed, sizeof(removed), hipMemcpyDeviceToHost));
            stats->removed += (int64_t)removed;
            stats->ms_push += st.ms_push;
            stats->ms_scan += st.ms_scan;
            stats->steps += 1;
        }
        return 0;
    });
}

int ipic_gpu_set_field_operator(ipic_ctx* ctx, const ipic_field_operator* op) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (!op) {
            ctx->op_on = false;
            return 0;
        }
        if (!op->params) return fail(ctx, IPIC_E_ARG, "null operator parameters");
        if (ctx->d.flags & IPIC_FLAG_FORWARD_SOLVER)
            return fail(ctx, IPIC_E_STATE, "the field operator stands in for leapfrog substeps");
        if (op->substeps < 1 || op->substeps != ctx->d.field_substeps)
            return fail(ctx, IPIC_E_ARG, "operator substeps differ from the context's field_substeps");
        if (!(op->in_E > 0) || !(op->in_B > 0) || !(op->in_J > 0) || !std::isfinite(op->out_E) || !std::isfinite(op->out_P))
            return fail(ctx, IPIC_E_ARG, "operator scales must be positive and finite");
        if (op->audit_every < 0) return fail(ctx, IPIC_E_ARG, "negative audit interval");
        const Geo& g = ctx->g;
        const int32_t n3[3] = {g.n[0], g.n[1], g.n[2]};
        const long long wsb = ipic_nop_workspace_bytes(n3, 0);
        if (wsb < 0) return nop_fail(ctx, IPIC_E_ARG);
        const long long np = ipic_nop_param_count();
        const size_t nc = (size_t)g.ncell;
        if (!ctx->op_params) {
            if (int rc = dalloc(ctx, &ctx->op_params, (size_t)np)) return rc;
            if (int rc = dalloc(ctx, &ctx->op_wf, (size_t)np)) return rc;
            if (int rc = dalloc(ctx, &ctx->op_ws, (size_t)wsb)) return rc;
            if (int rc = dalloc(ctx, &ctx->op_u, nc * 16)) return rc;
            if (int rc = dalloc(ctx, &ctx->op_out, nc * 16)) return rc;
            if (int rc = dalloc(ctx, &ctx->op_psi, nc)) return rc;
            if (int rc = dalloc(ctx, &ctx->Eb, nc)) return rc;
            if (int rc = dalloc(ctx, &ctx->Bb, nc)) return rc;
            if (int rc = dalloc(ctx, &ctx->Eo, nc)) return rc;
            if (int rc = dalloc(ctx, &ctx->Bo, nc)) return rc;
        }
        for (long long i = 0; i < np; ++i)
            if (!std::isfinite(op->params[i])) return fail(ctx, IPIC_E_ARG, "non-finite operator parameter");
        HIPCK(ctx, hipMemcpyAsync(ctx->op_params, op->params, (size_t)np * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        if (int rc = ipic_nop_prepare(ctx->op_params, ctx->op_wf, nullptr, ctx->stream)) return nop_fail(ctx, rc);
        HIPCK(ctx, hipStreamSynchronize(ctx->stream));
        ctx->op = *op;
        ctx->op.params = nullptr;
        ctx->op_scales = ipic_nop_scales{op->in_E, op->in_B, op->in_J, op->out_E, op->out_P};
        ctx->op_count = 0;
        ctx->op_on = true;
        return 0;
    });
}

int ipic_gpu_field_state(ipic_ctx* ctx, ipic_field_state* out) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (!out) return fail(ctx, IPIC_E_ARG, "null output");
        if (!ctx->j_valid) {
            if (int rc = launch_particles(ctx, MODE_DEPOSIT)) return rc;
            ctx->j_valid = true;
        }
        if (int rc = launch_jsum(ctx)) return rc;
        HIPCK(ctx, hipStreamSynchronize(ctx->stream));
        out->E = (float*)ctx->E;
        out->Bc = (float*)ctx->Bc;
        out->J = (float*)ctx->J;
        for (int k = 0; k < 3; ++k) {
            out->E0[k] = ctx->g.E0[k];
            out->B0[k] = ctx->g.B0[k];
        }
        out->uniform_background = (ctx->Ebg || ctx->Bcbg || ctx->SE || ctx->SB) ? 0 : 1;
        return 0;
    });
}

int ipic_gpu_num_particles(ipic_ctx* ctx, int32_t species, int64_t* n) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (!n) return fail(ctx, IPIC_E_ARG, "null output");
        if (species >= ctx->d.nspecies) return fail(ctx, IPIC_E_ARG, "species out of range");
        if (int rc = canonicalize(ctx)) return rc;
        long long tot = 0;
        for (int s = 0; s < ctx->d.nspecies; ++s) {
            if (species >= 0 && s != species) continue;
            long long t = 0;
            if (int rc = read_total(ctx, s, &t)) return rc;
            tot += t;
        }
        *n = tot;
        return 0;
    });
}

int ipic_gpu_cell_counts(ipic_ctx* ctx, int32_t species, int64_t* counts) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (!counts) return fail(ctx, IPIC_E_ARG, "null output");
        if (species >= ctx->d.nspecies) return fail(ctx, IPIC_E_ARG, "species out of range");
        if (int rc = canonicalize(ctx)) return rc;
        const Geo& g = ctx->g;
        const int nc = g.ncell;
        std::vector<int> st(nc + 1);
        std::memset(counts, 0, (size_t)nc * sizeof(int64_t));
        for (int s = 0; s < ctx->d.nspecies; ++s) {
            if (species >= 0 && s != species) continue;
            HIPCK(ctx, hipMemcpy(st.data(), ctx->sp[s].start, (nc + 1) * sizeof(int), hipMemcpyDeviceToHost));
            for (int cid = 0; cid < nc; ++cid) {
                int x, y, z;
                cid_xyz(g, cid, x, y, z);
                counts[(x * g.n[1] + y) * g.n[2] + z] += st[cid + 1] - st[cid];
            }
        }
        return 0;
    });
}

int ipic_gpu_download_particles(ipic_ctx* ctx, int32_t species, double* pos, double* vel, int64_t* cell,
                                uint32_t* ids) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (species < 0 || species >= ctx->d.nspecies) return fail(ctx, IPIC_E_ARG, "species out of range");
        if (ids && !ctx->track) return fail(ctx, IPIC_E_STATE, "ids need IPIC_FLAG_TRACK_IDS");
        if (int rc = canonicalize(ctx)) return rc;
        const Geo& g = ctx->g;
        const int nc = g.ncell;
        Species& S = ctx->sp[species];
        std::vector<int> st(nc + 1);
        HIPCK(ctx, hipMemcpy(st.data(), S.start, (nc + 1) * sizeof(int), hipMemcpyDeviceToHost));
        const long long n = st[nc];
        std::vector<float> X((size_t)6 * std::max<long long>(n, 1));
        for (int k = 0; k < 6 && n > 0; ++k)
            HIPCK(ctx, hipMemcpy(&X[(size_t)k * n], S.X[S.cur] + (size_t)k * S.cap, (size_t)n * sizeof(float),
                                 hipMemcpyDeviceToHost));
        std::vector<uint32_t> hid(ids ? std::max<long long>(n, 1) : 0);
        if (ids && n > 0) HIPCK(ctx, hipMemcpy(hid.data(), S.ids[S.cur], (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        long long o = 0;
        for (int x = 0; x < g.n[0]; ++x)
            for (int y = 0; y < g.n[1]; ++y)
                for (int z = 0; z < g.n[2]; ++z) {
                    const int cid = host_cid(g, x, y, z);
                    const int cc[3] = {x, y, z};
                    for (int p = st[cid]; p < st[cid + 1]; ++p, ++o) {
                        for (int k = 0; k < 3; ++k) {
                            if (pos) pos[3 * o + k] = ctx->d.origin[k] + ((double)cc[k] + (double)X[(size_t)k * n + p]) * ctx->d.cell_width[k];
                            if (vel) vel[3 * o + k] = (double)X[(size_t)(3 + k) * n + p];
                        }
                        if (cell) cell[o] = ((long long)x * g.n[1] + y) * g.n[2] + z;
                        if (ids) ids[o] = hid[p];
                    }
                }
        return 0;
    });
}

// With the static field no field update uses J: it is the current density of the current
// particle state, formed here from the deposit when it is read.
int static_field_j(ipic_ctx* ctx) {
    if (ctx->d.field_substeps != 0) return 0;
    if (!ctx->j_valid) {
        if (int rc = launch_particles(ctx, MODE_DEPOSIT)) return rc;
        ctx->j_valid = true;
    }
    return launch_jsum(ctx);
}

int ipic_gpu_download_fields(ipic_ctx* ctx, double* E, double* Bc, double* Bn, double* J) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (J)
            if (int rc = static_field_j(ctx)) return rc;
        const Geo& g = ctx->g;
        const int nc = g.ncell;
        std::vector<double> tB;
        if (Bn && !Bc) tB.resize((size_t)nc * 3);
        double* bc = Bc ? Bc : (Bn ? tB.data() : nullptr);
        if (int rc = total_fields(ctx, E, bc)) return rc;
        if (Bn) {
            // interpC2N (field.h:234-247) of the total Bc, FP64
            const int* n = g.n;
            for (int x = 0; x < n[0]; ++x)
                for (int y = 0; y < n[1]; ++y)
                    for (int z = 0; z < n[2]; ++z) {
                        const size_t i = (size_t)((x * n[1] + y) * n[2] + z);
                        for (int k = 0; k < 3; ++k) {
                            double acc = 0;
                            for (int a = 0; a < 2; ++a)
                                for (int b = 0; b < 2; ++b)
                                    for (int c = 0; c < 2; ++c) {
                                        const int xx = (x - a + n[0]) % n[0], yy = (y - b + n[1]) % n[1],
                                                  zz = (z - c + n[2]) % n[2];
                                        acc += bc[3 * ((xx * n[1] + yy) * n[2] + zz) + k];
                                    }
                            Bn[3 * i + k] = 0.125 * acc;
                        }
                    }
        }
        if (J) {
            std::vector<float4> h(nc);
            HIPCK(ctx, hipMemcpy(h.data(), ctx->J, nc * sizeof(float4), hipMemcpyDeviceToHost));
            for (int i = 0; i < nc; ++i) {
                J[3 * i] = h[i].x;
                J[3 * i + 1] = h[i].y;
                J[3 * i + 2] = h[i].z;
            }
        }
        return 0;
    });
}

int ipic_gpu_conservation(ipic_ctx* ctx, ipic_conservation* out) {
    return guarded(ctx, [&]() -> int {
        if (int rc = check_ctx(ctx)) return rc;
        if (!out) return fail(ctx, IPIC_E_ARG, "null output");
        ipic_conservation r{};
        std::vector<double> h((size_t)ctx->red_blocks * 5);
        for (int s = 0; s < ctx->d.nspecies; ++s) {
            Species& S = ctx->sp[s];
            long long nslots = 0;
            if (int rc = read_slots(ctx, s, &nslots)) return rc;
            if (nslots == 0) continue;
            const double mass = S.q / S.qom;
            hipLaunchKernelGGL(k_reduce_particles, dim3(ctx->red_blocks), dim3(256), 0, ctx->stream, S.X[S.cur], S.cap,
                               nslots, mass, ctx->red);
            HIPCK(ctx, hipGetLastError());
            HIPCK(ctx, hipMemcpyAsync(h.data(), ctx->red, (size_t)ctx->red_blocks * 3 * sizeof(double),
                                      hipMemcpyDeviceToHost, ctx->stream));
            HIPCK(ctx, hipStreamSynchronize(ctx->stream));
            double ke = 0, mo = 0, cnt = 0;
            for (int b = 0; b < ctx->red_blocks; ++b) {
                ke += h[3 * b];
                mo += h[3 * b + 1];
                cnt += h[3 * b + 2];
            }
            r.kinetic_energy += ke;
            r.momentum += mo;
            r.charge += cnt * S.q;
            r.particles += (int64_t)cnt;
        }
        if (int rc = static_field_j(ctx)) return rc;
        FieldArgs F = field_args(ctx);
        hipLaunchKernelGGL(k_reduce_fields, dim3(ctx->red_blocks), dim3(256), 0, ctx->stream, F, ctx->red);
        HIPCK(ctx, hipGetLastError());
        HIPCK(ctx, hipMemcpyAsync(h.data(), ctx->red, (size_t)ctx->red_blocks * 5 * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
        HIPCK(ctx, hipStreamSynchronize(ctx->stream));
        double se = 0, sb = 0, sej = 0, sd2 = 0, md = 0;
        for (int b = 0; b < ctx->red_blocks; ++b) {
            se += h[5 * b];
            sb += h[5 * b + 1];
            sej += h[5 * b + 2];
            sd2 += h[5 * b + 3];
            md = std::max(md, h[5 * b + 4]);
        }
        const double vol = ctx->d.cell_width[0] * ctx->d.cell_width[1] * ctx->d.cell_width[2];
        r.field_energy_E = 0.5 * se * vol;
        r.field_energy_B = 0.5 * ctx->d.c * ctx->d.c * sb * vol;
        r.EdotJ = sej * vol;
        r.divB_max = md;
        r.divB_l2 = std::sqrt(sd2);
        *out = r;
        return 0;
    });
}

}  // extern "C"
