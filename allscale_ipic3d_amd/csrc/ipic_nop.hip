// ipic_nop.hip -- learned field step for gfx950: 3x3x3 periodic convolutions as implicit GEMMs
// on BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32 accumulation, their data and weight
// gradients, the field head (increments of E and of a potential whose forward curl updates Bc),
// the training loss and Adam.  C ABI in include/ipic_nop.h.
//
// Activations are channel-last on the n^3 index lattice: X[v * C + c], v = (x * ny + y) * nz + z.
//
// Forward convolution (k_conv): one persistent block per CU stages the layer's weights in LDS
// once and walks 4x4x16 bricks of output voxels, computing every output channel of a brick.  The
// brick's input halo (6x6x18 voxels x CIN channels) is staged in LDS with an 8-byte pad per
// voxel (conflict-free A-fragment reads; the weight rows carry the same pad), and the next
// brick's halo is loaded into registers while the current brick computes.  Each wave owns one
// x-plane of the brick, i.e. four 16-voxel z-rows, and for every tap and 16-channel slice issues
// one MFMA per (row, 16-output-channel tile): the A operand is 16 voxels x 16 input channels,
// the B operand 16 input x 16 output channels of the tap's weights, both from LDS.  The epilogue adds the bias, optionally multiplies by silu'(z_saved) and adds a
// residual, and writes FP32 and/or BF16 (optionally silu-activated) outputs.
//
// Data gradient: the same kernel with the weights mirrored in the taps and transposed in the
// channels.  Weight gradient (k_wgrad): per block and brick, dW[tap][co][ci] += sum over the
// brick's voxels of dY[v][co] X[v + tap][ci] as MFMAs with K = voxels; the operands are
// staged transposed (channel-major) in LDS, block partials go to global FP32 by atomics.

#include <hip/hip_runtime.h>

#include <atomic>

#include <cmath>
#include <cstdint>
#include <string>

#include "ipic_gpu.h"
#include "ipic_nop.h"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 8 consecutive BF16 values from LDS at an 8-byte aligned address, as an MFMA operand
__device__ __forceinline__ v8bf lds8(const unsigned char* p) {
    const v4s lo = *(const v4s*)p, hi = *(const v4s*)(p + 8);
    return __builtin_bit_cast(v8bf, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
}

constexpr int BX = 4, BY = 4, BZ = 16;               // output brick of a block
constexpr int BV = BX * BY * BZ;                     // 256 voxels
constexpr int HX = BX + 2, HY = BY + 2, HZ = BZ + 2;  // input halo brick
constexpr int HV = HX * HY * HZ;                     // 648 voxels
constexpr int NL = IPIC_NOP_LAYERS;

__host__ __device__ constexpr int lcin(int l) { return l == 0 ? 16 : 32; }
__host__ __device__ constexpr int lcout(int l) { return l == NL - 1 ? 16 : 32; }
__host__ __device__ constexpr long long lsize(int l) { return 27LL * lcin(l) * lcout(l) + lcout(l); }
__host__ __device__ constexpr long long loff(int l) { return l == 0 ? 0 : loff(l - 1) + lsize(l - 1); }
constexpr long long NPARAM = loff(NL);

thread_local std::string g_err;
thread_local const char* g_err_lit = nullptr;  // message of a failure that allocates nothing

int fail(int code, const std::string& m) {
    g_err_lit = nullptr;
    g_err = m;
    return code;
}

int fail_lit(int code, const char* m) {
    g_err_lit = m;
    return code;
}

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// silu and its derivative with the hardware reciprocal (1 ulp): the results feed BF16 values
// and an FP32 residual stream whose error budget is set by the BF16 rounding of the weights
__device__ __forceinline__ float silu(float x) { return x * __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }

__device__ __forceinline__ float dsilu(float x) {
    const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-x));
    return s * (1.0f + x * (1.0f - s));
}

__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

struct ConvArgs {
    int n[3];
    const uint16_t* x;    // [V][CIN] BF16
    const uint16_t* w;    // [27][COUT][CIN] BF16
    const float* bias;    // [COUT] (nullable)
    const float* zs;      // [V][COUT]: multiply by silu'(zs) (nullable)
    const float* res;     // [V][COUT]: add after the multiply (nullable; may equal outf)
    float* outf;          // [V][COUT] FP32 (nullable)
    uint16_t* outb;       // [V][COUT] BF16 (nullable)
    int act;              // outb = silu(v) (1) or v (0)
};

template <int CIN, int COUT>
__global__ __launch_bounds__(512) void k_conv(ConvArgs a, int nbricks) {
    constexpr int SV = CIN * 2 + 8;  // LDS bytes per halo voxel and per weight row
    constexpr int PV = CIN / 4;      // 8-byte pieces per voxel
    constexpr int NT = COUT / 16;
    constexpr int NP = HV * PV;      // halo pieces of a brick
    constexpr int NH = (NP + 255) / 256;
    __shared__ __align__(16) unsigned char sw[27 * COUT * SV];  // the layer's weights, all bricks
    __shared__ __align__(16) unsigned char sx[2][HV * SV];      // the two current bricks' halos
    const int n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    const int nbz = n2 / BZ, nby = n1 / BY;
    // the block's two halves (4 waves each) work on bricks 2 p and 2 p + 1 of pair p
    const int half = threadIdx.x >> 8, ht = threadIdx.x & 255;

    for (int i = threadIdx.x; i < 27 * COUT * PV; i += 512) {
        const int row = i / PV, p = i - row * PV;
        *(uint2*)(sw + row * SV + p * 8) = *(const uint2*)(a.w + (long long)row * CIN + p * 4);
    }

    // the next brick's halo travels in registers while the current brick computes
    uint2 pre[NH];
    auto fetch = [&](int bb) {
        const int bz = bb % nbz, by = (bb / nbz) % nby, bx = bb / (nbz * nby);
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            const int i = ht + j * 256;
            if (i < NP) {
                const int v = i / PV, p = i - v * PV;
                const int hx = v / (HY * HZ), r = v - hx * (HY * HZ), hy = r / HZ, hz = r - hy * HZ;
                const int gx = wrap1(bx * BX + hx - 1, n0), gy = wrap1(by * BY + hy - 1, n1);
                const int gz = wrap1(bz * BZ + hz - 1, n2);
                const long long gv = ((long long)gx * n1 + gy) * n2 + gz;
                pre[j] = *(const uint2*)(a.x + gv * CIN + p * 4);
            }
        }
    };

    const int lane = threadIdx.x & 63, w = (threadIdx.x >> 6) & 3;
    const int m = lane & 15, kq = lane >> 4;
    unsigned char* hx_ = sx[half];
    if (2 * (int)blockIdx.x + half < nbricks) fetch(2 * blockIdx.x + half);
    for (int pr = blockIdx.x; 2 * pr < nbricks; pr += gridDim.x) {
        const int bb = 2 * pr + half;
        const bool mine = bb < nbricks;
        __syncthreads();  // the previous bricks' halo reads are done (and the weights are in)
        if (mine) {
#pragma unroll
            for (int j = 0; j < NH; ++j) {
                const int i = ht + j * 256;
                if (i < NP) {
                    const int v = i / PV, p = i - v * PV;
                    *(uint2*)(hx_ + v * SV + p * 8) = pre[j];
                }
            }
        }
        __syncthreads();
        const int nb = 2 * (pr + (int)gridDim.x) + half;
        if (nb < nbricks) fetch(nb);
        if (!mine) continue;
        const int bz = bb % nbz, by = (bb / nbz) % nby, bx = bb / (nbz * nby);
        const int x0 = bx * BX, y0 = by * BY, z0 = bz * BZ;

        v4f acc[BY][NT];
#pragma unroll
        for (int y = 0; y < BY; ++y)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[y][t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};

        for (int dx = 0; dx < 3; ++dx)
            for (int dy = 0; dy < 3; ++dy) {
#pragma unroll
                for (int dz = 0; dz < 3; ++dz) {
                    const int tap = (dx * 3 + dy) * 3 + dz;
                    if constexpr (CIN % 32 == 0 && COUT >= 32) {
                        // 32 input channels per MFMA (v_mfma_f32_16x16x32_bf16): lane (m, kq)
                        // holds channels 8 kq .. 8 kq + 7 of its row / column
#pragma unroll
                        for (int s2 = 0; s2 < CIN / 32; ++s2) {
                            v8bf bf[NT];
#pragma unroll
                            for (int t = 0; t < NT; ++t)
                                bf[t] = lds8(sw + (tap * COUT + t * 16 + m) * SV + (s2 * 32 + kq * 8) * 2);
#pragma unroll
                            for (int y = 0; y < BY; ++y) {
                                const int hv = ((w + dx) * HY + (y + dy)) * HZ + (m + dz);
                                const v8bf af = lds8(hx_ + hv * SV + (s2 * 32 + kq * 8) * 2);
#pragma unroll
                                for (int t = 0; t < NT; ++t)
                                    acc[y][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf[t], acc[y][t], 0, 0, 0);
                            }
                        }
                    } else {
#pragma unroll
                        for (int s2 = 0; s2 < CIN / 16; ++s2) {
                            v4s bf[NT];
#pragma unroll
                            for (int t = 0; t < NT; ++t)
                                bf[t] = *(const v4s*)(sw + (tap * COUT + t * 16 + m) * SV + (s2 * 16 + kq * 4) * 2);
#pragma unroll
                            for (int y = 0; y < BY; ++y) {
                                const int hv = ((w + dx) * HY + (y + dy)) * HZ + (m + dz);
                                const v4s af = *(const v4s*)(hx_ + hv * SV + (s2 * 16 + kq * 4) * 2);
#pragma unroll
                                for (int t = 0; t < NT; ++t)
                                    acc[y][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af, bf[t], acc[y][t], 0, 0, 0);
                            }
                        }
                    }
                }
            }

        // D[row = z (4 kq + j)][col = output channel (t 16 + m)]
        if constexpr (NT == 2) {
            // lanes m and m ^ 1 swap one value (DPP quad permute 1,0,3,2): the even lane then
            // holds channels m, m + 1 and the odd lane 15 + m, 16 + m of the same voxel, and
            // each stores its pair at once, so a store instruction covers each voxel's whole
            // 32-channel row (128 B in FP32, 64 B in BF16) instead of two halves
            const bool odd = (m & 1) != 0;
            const long long vb0 = (((long long)(x0 + w) * n1 + y0) * n2 + z0 + kq * 4) * COUT;  // channel 0
            const float b0 = a.bias ? a.bias[m] : 0.0f, b1 = a.bias ? a.bias[16 + m] : 0.0f;
#pragma unroll
            for (int y = 0; y < BY; ++y)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const long long vb = vb0 + ((long long)y * n2 + j) * COUT;
                    float v0 = acc[y][0][j] + b0, v1 = acc[y][1][j] + b1;
                    if (a.zs) {
                        v0 *= dsilu(a.zs[vb + m]);
                        v1 *= dsilu(a.zs[vb + 16 + m]);
                    }
                    if (a.res) {
                        v0 += a.res[vb + m];
                        v1 += a.res[vb + 16 + m];
                    }
                    const float give = odd ? v0 : v1;
                    const float got =
                        __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(give), 0xB1, 0xF, 0xF, false));
                    const float plo = odd ? got : v0, phi = odd ? v1 : got;
                    const long long o = vb + (odd ? 15 + m : m);  // even: 8-byte aligned
                    if (a.outf) *(float2*)(a.outf + o) = make_float2(plo, phi);
                    if (a.outb)
                        *(uint32_t*)(a.outb + o) = (uint32_t)f2bf(a.act ? silu(plo) : plo) |
                                                   ((uint32_t)f2bf(a.act ? silu(phi) : phi) << 16);
                }
            continue;
        }
        const long long ob = (((long long)(x0 + w) * n1 + y0) * n2 + z0 + kq * 4) * COUT + m;  // (t, y, j) = 0
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const float bias = a.bias ? a.bias[t * 16 + m] : 0.0f;
#pragma unroll
            for (int y = 0; y < BY; ++y)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const long long o = ob + ((long long)y * n2 + j) * COUT + t * 16;
                    float v = acc[y][t][j] + bias;
                    if (a.zs) v *= dsilu(a.zs[o]);
                    if (a.res) v += a.res[o];
                    if (a.outf) a.outf[o] = v;
                    if (a.outb) a.outb[o] = f2bf(a.act ? silu(v) : v);
                }
        }
    }
}

struct WgradArgs {
    int n[3];
    const uint16_t* x;    // [V][CIN] BF16, input of the convolution
    const uint16_t* dy;   // [V][COUT] BF16, gradient of its output
    float* dw;            // [27][COUT][CIN] FP32, accumulated
    float* db;            // [COUT] FP32, accumulated
};

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
    constexpr int HVP = HV + 2;   // channel row of the transposed halo (odd dword stride)
    constexpr int YP = BV + 4;    // channel row of the transposed output gradient
    constexpr int NS = CIN / 16, NT = COUT / 16;
    constexpr int TPW = 7;        // taps per wave: w, w + 4, ..., < 27
    __shared__ __align__(16) uint16_t xt[CIN * HVP];
    __shared__ __align__(16) uint16_t yt[COUT * YP];
    const int n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    const int nbz = n2 / BZ, nby = n1 / BY;
    const int nbricks = (n0 / BX) * nby * nbz;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;

    v4f acc[TPW][NS][NT];
#pragma unroll
    for (int i = 0; i < TPW; ++i)
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[i][s][t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    constexpr int PARTS = 256 / COUT, PLEN = BV / PARTS;
    const int bco = threadIdx.x % COUT, bpart = threadIdx.x / COUT;
    float dbacc = 0.0f;

    for (int b = blockIdx.x; b < nbricks; b += gridDim.x) {
        const int bz = b % nbz, by = (b / nbz) % nby, bx = b / (nbz * nby);
        const int x0 = bx * BX, y0 = by * BY, z0 = bz * BZ;
        __syncthreads();
        for (int i = threadIdx.x; i < HV * (CIN / 4); i += 256) {
            const int v = i / (CIN / 4), p = i - v * (CIN / 4);
            const int hx = v / (HY * HZ), r = v - hx * (HY * HZ), hy = r / HZ, hz = r - hy * HZ;
            const int gx = wrap1(x0 + hx - 1, n0), gy = wrap1(y0 + hy - 1, n1), gz = wrap1(z0 + hz - 1, n2);
            const long long gv = ((long long)gx * n1 + gy) * n2 + gz;
            const uint2 q = *(const uint2*)(a.x + gv * CIN + p * 4);
            xt[(p * 4 + 0) * HVP + v] = (uint16_t)(q.x & 0xFFFFu);
            xt[(p * 4 + 1) * HVP + v] = (uint16_t)(q.x >> 16);
            xt[(p * 4 + 2) * HVP + v] = (uint16_t)(q.y & 0xFFFFu);
            xt[(p * 4 + 3) * HVP + v] = (uint16_t)(q.y >> 16);
        }
        for (int i = threadIdx.x; i < BV * (COUT / 4); i += 256) {
            const int v = i / (COUT / 4), p = i - v * (COUT / 4);
            const int lx = v / (BY * BZ), r = v - lx * (BY * BZ), ly = r / BZ, lz = r - ly * BZ;
            const long long gv = ((long long)(x0 + lx) * n1 + (y0 + ly)) * n2 + (z0 + lz);
            const uint2 q = *(const uint2*)(a.dy + gv * COUT + p * 4);
            yt[(p * 4 + 0) * YP + v] = (uint16_t)(q.x & 0xFFFFu);
            yt[(p * 4 + 1) * YP + v] = (uint16_t)(q.x >> 16);
            yt[(p * 4 + 2) * YP + v] = (uint16_t)(q.y & 0xFFFFu);
            yt[(p * 4 + 3) * YP + v] = (uint16_t)(q.y >> 16);
        }
        __syncthreads();
        for (int i = 0; i < PLEN; ++i)
            dbacc += __uint_as_float((uint32_t)yt[bco * YP + bpart * PLEN + i] << 16);

        for (int kk = 0; kk < BX * BY; ++kk) {  // one 16-voxel z-row of the brick per K step
            const int rx = kk / BY, ry = kk - rx * BY;
            v4s af[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) af[t] = *(const v4s*)(yt + (t * 16 + m) * YP + kk * 16 + kq * 4);
#pragma unroll
            for (int i = 0; i < TPW; ++i) {
                const int tap = w + 4 * i;
                if (tap < 27) {
                    const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;
                    const int hv = ((rx + dx) * HY + (ry + dy)) * HZ + kq * 4 + dz;
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        const uint16_t* p = xt + (s * 16 + m) * HVP + hv;
                        const v4s bfr = v4s{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
#pragma unroll
                        for (int t = 0; t < NT; ++t)
                            acc[i][s][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[t], bfr, acc[i][s][t], 0, 0, 0);
                    }
                }
            }
        }
    }
    // D[row = output channel (t 16 + 4 kq + j)][col = input channel (s 16 + m)]
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
        const int tap = w + 4 * i;
        if (tap < 27) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int co = t * 16 + kq * 4 + j, ci = s * 16 + m;
                        atomicAdd(&a.dw[(tap * COUT + co) * CIN + ci], acc[i][s][t][j]);
                    }
        }
    }
    atomicAdd(&a.db[bco], dbacc);
}

// BF16 weight copies: forward layout (as the parameters) and the data-gradient layout
// wb[26 - tap][ci][co] = W[tap][co][ci] of each layer.
__global__ void k_prepare(const float* p, uint16_t* wf, uint16_t* wb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    int l = 0;
    while (l + 1 < NL && i >= loff(l + 1)) ++l;
    const long long r = i - loff(l);
    const int ci_n = lcin(l), co_n = lcout(l);
    const uint16_t v = f2bf(p[i]);
    wf[i] = v;
    if (wb) {
        if (r < 27LL * ci_n * co_n) {
            const int tap = (int)(r / (ci_n * co_n)), q = (int)(r % (ci_n * co_n));
            const int co = q / ci_n, ci = q % ci_n;
            wb[loff(l) + ((26 - tap) * ci_n + ci) * co_n + co] = v;
        } else {
            wb[i] = 0;
        }
    }
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void k_pack(int n0, int n1, int n2, const float4* E, const float4* B, const float4* J, float iE,
                       float iB, float iJ, uint16_t* u) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n0 * n1 * n2) return;
    const float4 e = E[i], b = B[i], j = J[i];
    uint4 lo, hi;
    lo.x = f2bf(e.x * iE) | ((uint32_t)f2bf(e.y * iE) << 16);
    lo.y = f2bf(e.z * iE) | ((uint32_t)f2bf(b.x * iB) << 16);
    lo.z = f2bf(b.y * iB) | ((uint32_t)f2bf(b.z * iB) << 16);
    lo.w = f2bf(j.x * iJ) | ((uint32_t)f2bf(j.y * iJ) << 16);
    hi.x = f2bf(j.z * iJ);
    hi.y = hi.z = hi.w = 0u;
    uint4* d = (uint4*)(u + i * 16);
    d[0] = lo;
    d[1] = hi;
}

struct HeadArgs {
    int n[3];
    const float* out;     // [V][16]
    float soE, soP;
    float iw[3];
    float4* E;            // apply: dE (in place)
    float4* B;            // apply: dBc (in place)
    float4* psi;          // [V] scratch
    const float4* tE;     // loss: targets
    const float4* tB;
    float4* g;            // [V] scratch: dL/d(dB_pred)
    float kE, kB;         // 1 / (varE V), 1 / (varB V)
    double* loss;
    uint16_t* dout;       // [V][16]
};

__device__ __forceinline__ void xyz(const HeadArgs& h, long long i, int& x, int& y, int& z) {
    z = (int)(i % h.n[2]);
    const long long r = i / h.n[2];
    y = (int)(r % h.n[1]);
    x = (int)(r / h.n[1]);
}

__device__ __forceinline__ long long lin(const HeadArgs& h, int x, int y, int z) {
    return ((long long)x * h.n[1] + y) * h.n[2] + z;
}

__device__ __forceinline__ void block_add(double v, double* dst) {
    __shared__ double s[4];
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(dst, s[0] + s[1] + s[2] + s[3]);
}

// mode 0: apply dE, store psi; mode 1: loss on dE, dout[0:3], store psi
__global__ __launch_bounds__(256) void k_head_a(HeadArgs h, int mode) {
    const long long V = (long long)h.n[0] * h.n[1] * h.n[2];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < V) {
        const float* o = h.out + i * 16;
        const float ex = h.soE * o[0], ey = h.soE * o[1], ez = h.soE * o[2];
        h.psi[i] = make_float4(h.soP * o[3], h.soP * o[4], h.soP * o[5], 0.0f);
        if (mode == 0) {
            float4 e = h.E[i];
            e.x += ex; e.y += ey; e.z += ez;
            h.E[i] = e;
        } else {
            const float4 t = h.tE[i];
            const float rx = ex - t.x, ry = ey - t.y, rz = ez - t.z;
            l = (double)(h.kE * (rx * rx + ry * ry + rz * rz));
            uint16_t* d = h.dout + i * 16;
            d[0] = f2bf(2.0f * h.kE * rx * h.soE);
            d[1] = f2bf(2.0f * h.kE * ry * h.soE);
            d[2] = f2bf(2.0f * h.kE * rz * h.soE);
#pragma unroll
            for (int c = 6; c < 16; ++c) d[c] = 0;
        }
    }
    if (mode == 1) block_add(l, h.loss);
}

// dB_pred = -curl+ psi (the leapfrog B update with psi in place of dt E); mode 0 applies it,
// mode 1 adds its loss and stores g = dL/d(dB_pred)
__global__ __launch_bounds__(256) void k_head_b(HeadArgs h, int mode) {
    const long long V = (long long)h.n[0] * h.n[1] * h.n[2];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < V) {
        int x, y, z;
        xyz(h, i, x, y, z);
        const float4 p = h.psi[i];
        const float4 px = h.psi[lin(h, wrap1(x + 1, h.n[0]), y, z)];
        const float4 py = h.psi[lin(h, x, wrap1(y + 1, h.n[1]), z)];
        const float4 pz = h.psi[lin(h, x, y, wrap1(z + 1, h.n[2]))];
        const float bx = -((py.z - p.z) * h.iw[1] - (pz.y - p.y) * h.iw[2]);
        const float by = -((pz.x - p.x) * h.iw[2] - (px.z - p.z) * h.iw[0]);
        const float bz = -((px.y - p.y) * h.iw[0] - (py.x - p.x) * h.iw[1]);
        if (mode == 0) {
            float4 b = h.B[i];
            b.x += bx; b.y += by; b.z += bz;
            h.B[i] = b;
        } else {
            const float4 t = h.tB[i];
            const float rx = bx - t.x, ry = by - t.y, rz = bz - t.z;
            l = (double)(h.kB * (rx * rx + ry * ry + rz * rz));
            h.g[i] = make_float4(2.0f * h.kB * rx, 2.0f * h.kB * ry, 2.0f * h.kB * rz, 0.0f);
        }
    }
    if (mode == 1) block_add(l, h.loss);
}

// dL/dpsi = -(curl+)^T g = -curl- g (backward differences), scaled into dout[3:6]
__global__ __launch_bounds__(256) void k_head_c(HeadArgs h) {
    const long long V = (long long)h.n[0] * h.n[1] * h.n[2];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    int x, y, z;
    xyz(h, i, x, y, z);
    const float4 g = h.g[i];
    const float4 gx = h.g[lin(h, wrap1(x - 1, h.n[0]), y, z)];
    const float4 gy = h.g[lin(h, x, wrap1(y - 1, h.n[1]), z)];
    const float4 gz = h.g[lin(h, x, y, wrap1(z - 1, h.n[2]))];
    const float cx = (g.z - gy.z) * h.iw[1] - (g.y - gz.y) * h.iw[2];
    const float cy = (g.x - gz.x) * h.iw[2] - (g.z - gx.z) * h.iw[0];
    const float cz = (g.y - gx.y) * h.iw[0] - (g.x - gy.x) * h.iw[1];
    uint16_t* d = h.dout + i * 16;
    d[3] = f2bf(-cx * h.soP);
    d[4] = f2bf(-cy * h.soP);
    d[5] = f2bf(-cz * h.soP);
}

__global__ void k_sub4(float4* d, const float4* a, const float4* b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 x = a[i], y = b[i];
    d[i] = make_float4(x.x - y.x, x.y - y.y, x.z - y.z, 0.0f);
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------

struct Ws {
    float* x[5];          // residual stream x_0..x_4 (train); x[0] only for inference
    uint16_t* a[5];       // silu(x_b), BF16 (train); a[0] only for inference
    float* h[4];          // conv_b1 outputs (train)
    uint16_t* g[4];       // silu(h_b), BF16 (train); g[0] only for inference
    float* dx[2];         // backward: FP32 gradient of the residual stream (ping-pong)
    uint16_t* dxb;        // its BF16 copy
    uint16_t* dhb;        // BF16 gradient of conv_b1's output
};

long long voxels(const int32_t n[3]) { return (long long)n[0] * n[1] * n[2]; }

Ws carve(void* base, const int32_t n[3], int train, long long* bytes) {
    const long long V = voxels(n);
    const long long F = V * 32 * 4, H = V * 32 * 2;
    unsigned char* p = (unsigned char*)base;
    long long o = 0;
    Ws w{};
    auto takef = [&](float** d) { *d = (float*)(p ? p + o : nullptr); o += (F + 255) & ~255LL; };
    auto takeb = [&](uint16_t** d) { *d = (uint16_t*)(p ? p + o : nullptr); o += (H + 255) & ~255LL; };
    const int nb = train ? 5 : 1, nh = train ? 4 : 1;
    for (int k = 0; k < nb; ++k) takef(&w.x[k]);
    for (int k = 0; k < nb; ++k) takeb(&w.a[k]);
    if (train)
        for (int k = 0; k < 4; ++k) takef(&w.h[k]);
    for (int k = 0; k < nh; ++k) takeb(&w.g[k]);
    if (train) {
        takef(&w.dx[0]);
        takef(&w.dx[1]);
        takeb(&w.dxb);
        takeb(&w.dhb);
    }
    if (bytes) *bytes = o;
    return w;
}

int check_grid(const int32_t n[3]) {
    if (!n || n[0] <= 0 || n[1] <= 0 || n[2] <= 0 || n[0] % BX || n[1] % BY || n[2] % BZ)
        return fail(IPIC_E_ARG, "grid must be positive with nx % 4 == 0, ny % 4 == 0, nz % 16 == 0");
    if (voxels(n) * 32 >= (1LL << 31)) return fail(IPIC_E_RANGE, "grid too large for the operator");
    return 0;
}

#define NOPCK(expr)                                                                                 \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return fail(IPIC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// compute units of the current device (cached per device)
int cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cached[dev]) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        cached[dev] = c;
    }
    return cached[dev];
}

template <int CIN, int COUT>
int conv(const int32_t n[3], const uint16_t* x, const uint16_t* w, const float* bias, const float* zs,
         const float* res, float* outf, uint16_t* outb, int act, hipStream_t st) {
    ConvArgs a{{n[0], n[1], n[2]}, x, w, bias, zs, res, outf, outb, act};
    const int bricks = (n[0] / BX) * (n[1] / BY) * (n[2] / BZ);
    // one persistent block per CU: the weights are staged in LDS once per block, and its two
    // halves work on two bricks at a time
    const int pairs = (bricks + 1) / 2;
    const int blocks = pairs < cu_count() ? pairs : cu_count();
    hipLaunchKernelGGL((k_conv<CIN, COUT>), dim3(blocks), dim3(512), 0, st, a, bricks);
    NOPCK(hipGetLastError());
    return 0;
}

template <int CIN, int COUT>
int wgrad(const int32_t n[3], const uint16_t* x, const uint16_t* dy, float* dw, float* db, hipStream_t st) {
    WgradArgs a{{n[0], n[1], n[2]}, x, dy, dw, db};
    const int bricks = (n[0] / BX) * (n[1] / BY) * (n[2] / BZ);
    const int blocks = bricks < 256 ? bricks : 256;
    hipLaunchKernelGGL((k_wgrad<CIN, COUT>), dim3(blocks), dim3(256), 0, st, a);
    NOPCK(hipGetLastError());
    return 0;
}

HeadArgs head_args(const int32_t n[3], const float* out, const ipic_nop_scales* s, const float inv_w[3]) {
    HeadArgs h{};
    h.n[0] = n[0]; h.n[1] = n[1]; h.n[2] = n[2];
    h.out = out;
    h.soE = s->out_E;
    h.soP = s->out_P;
    for (int k = 0; k < 3; ++k) h.iw[k] = inv_w ? inv_w[k] : 1.0f;
    return h;
}

// Scratch of ipic_nop_apply_noalloc: the head's potential psi for grids of at most
// IPIC_NOP_NOALLOC_MAX_VOXELS voxels (32 MiB of float4), one copy per GPU the library runs on,
// held until the process ends.  Host code uses the per-device address from
// hipGetSymbolAddress, never the array's name (a host-side stand-in).
static __device__ __align__(256) float4 scratch[IPIC_NOP_NOALLOC_MAX_VOXELS];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::atomic<float4*> g_scratch_ptr[NOALLOC_MAX_DEVICES];

// The current device's copy of `scratch`, looked up on the first call on that device.
int noalloc_scratch(float4** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail_lit(IPIC_E_HIP, "hipGetDevice failed");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail_lit(IPIC_E_RANGE, "device index above 63");
    float4* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        void* q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q)
            return fail_lit(IPIC_E_HIP, "hipGetSymbolAddress of the scratch array failed");
        p = (float4*)q;
        g_scratch_ptr[dev].store(p, std::memory_order_release);
    }
    *out = p;
    return 0;
}

}  // namespace

extern "C" {

const char* ipic_nop_last_error(void) { return g_err_lit ? g_err_lit : g_err.c_str(); }

int64_t ipic_nop_param_count(void) { return NPARAM; }

int ipic_nop_layer(int32_t layer, int64_t* w_offset, int64_t* b_offset, int32_t* cin_p, int32_t* cout_p) {
    if (layer < 0 || layer >= NL) return fail(IPIC_E_ARG, "layer out of range");
    if (w_offset) *w_offset = loff(layer);
    if (b_offset) *b_offset = loff(layer) + 27LL * lcin(layer) * lcout(layer);
    if (cin_p) *cin_p = lcin(layer);
    if (cout_p) *cout_p = lcout(layer);
    return 0;
}

int64_t ipic_nop_workspace_bytes(const int32_t n[3], int32_t train) {
    if (check_grid(n)) return -1;
    long long b = 0;
    carve(nullptr, n, train ? 1 : 0, &b);
    return b;
}

int ipic_nop_prepare(const float* params, uint16_t* wf, uint16_t* wb, void* stream) {
    if (!params || !wf) return fail(IPIC_E_ARG, "null pointer");
    hipLaunchKernelGGL(k_prepare, dim3((unsigned)((NPARAM + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params,
                       wf, wb);
    NOPCK(hipGetLastError());
    return 0;
}

int ipic_nop_pack(const int32_t n[3], const float* dE4, const float* dBc4, const float* J4,
                  const ipic_nop_scales* s, uint16_t* u, void* stream) {
    if (int rc = check_grid(n)) return rc;
    if (!dE4 || !dBc4 || !J4 || !s || !u) return fail(IPIC_E_ARG, "null pointer");
    const long long V = voxels(n);
    hipLaunchKernelGGL(k_pack, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n[0], n[1],
                       n[2], (const float4*)dE4, (const float4*)dBc4, (const float4*)J4, 1.0f / s->in_E,
                       1.0f / s->in_B, 1.0f / s->in_J, u);
    NOPCK(hipGetLastError());
    return 0;
}

int ipic_nop_forward(const int32_t n[3], const uint16_t* u, const uint16_t* wf, const float* params, void* ws,
                     int32_t train, float* out, void* stream) {
    if (int rc = check_grid(n)) return rc;
    if (!u || !wf || !params || !ws || !out) return fail(IPIC_E_ARG, "null pointer");
    hipStream_t st = (hipStream_t)stream;
    const Ws w = carve(ws, n, train ? 1 : 0, nullptr);
    auto W = [&](int l) { return wf + loff(l); };
    auto Bi = [&](int l) { return params + loff(l) + 27LL * lcin(l) * lcout(l); };
    if (int rc = conv<16, 32>(n, u, W(0), Bi(0), nullptr, nullptr, w.x[0], w.a[0], 1, st)) return rc;
    for (int b = 0; b < 4; ++b) {
        const int xi = train ? b : 0, xo = train ? b + 1 : 0;
        uint16_t* g = train ? w.g[b] : w.g[0];
        if (int rc = conv<32, 32>(n, w.a[xi], W(1 + 2 * b), Bi(1 + 2 * b), nullptr, nullptr, train ? w.h[b] : nullptr,
                                  g, 1, st))
            return rc;
        if (int rc = conv<32, 32>(n, g, W(2 + 2 * b), Bi(2 + 2 * b), nullptr, w.x[xi], w.x[xo], w.a[xo], 1, st))
            return rc;
    }
    return conv<32, 16>(n, w.a[train ? 4 : 0], W(9), Bi(9), nullptr, nullptr, out, nullptr, 0, st);
}

int ipic_nop_apply(const int32_t n[3], const float* out, const ipic_nop_scales* s, const float inv_w[3],
                   float* dE4, float* dBc4, float* psi4, void* stream) {
    if (int rc = check_grid(n)) return rc;
    if (!out || !s || !dE4 || !dBc4 || !psi4) return fail(IPIC_E_ARG, "null pointer");
    HeadArgs h = head_args(n, out, s, inv_w);
    h.E = (float4*)dE4;
    h.B = (float4*)dBc4;
    h.psi = (float4*)psi4;
    const long long V = voxels(n);
    const dim3 g((unsigned)((V + 255) / 256));
    hipLaunchKernelGGL(k_head_a, g, dim3(256), 0, (hipStream_t)stream, h, 0);
    hipLaunchKernelGGL(k_head_b, g, dim3(256), 0, (hipStream_t)stream, h, 0);
    NOPCK(hipGetLastError());
    return 0;
}

int ipic_nop_apply_noalloc(const int32_t n[3], const float* out, const ipic_nop_scales* s, const float inv_w[3],
                           float* dE4, float* dBc4, void* stream) {
    if (!n || n[0] <= 0 || n[1] <= 0 || n[2] <= 0 || n[0] % BX || n[1] % BY || n[2] % BZ)
        return fail_lit(IPIC_E_ARG, "grid must be positive with nx % 4 == 0, ny % 4 == 0, nz % 16 == 0");
    const long long V = voxels(n);
    if (V > IPIC_NOP_NOALLOC_MAX_VOXELS)
        return fail_lit(IPIC_E_RANGE, "grid above IPIC_NOP_NOALLOC_MAX_VOXELS (2097152 voxels) for the "
                                      "no-allocation apply");
    if (!out || !s || !dE4 || !dBc4) return fail_lit(IPIC_E_ARG, "null pointer");
    float4* psi = nullptr;
    if (int rc = noalloc_scratch(&psi)) return rc;
    HeadArgs h = head_args(n, out, s, inv_w);
    h.E = (float4*)dE4;
    h.B = (float4*)dBc4;
    h.psi = psi;  // every entry below V is written by k_head_a before k_head_b reads it
    const dim3 g((unsigned)((V + 255) / 256));
    hipLaunchKernelGGL(k_head_a, g, dim3(256), 0, (hipStream_t)stream, h, 0);
    hipLaunchKernelGGL(k_head_b, g, dim3(256), 0, (hipStream_t)stream, h, 0);
    if (hipGetLastError() != hipSuccess) return fail_lit(IPIC_E_HIP, "kernel launch failed");
    return 0;
}

int ipic_nop_loss(const int32_t n[3], const float* out, const ipic_nop_scales* s, const float inv_w[3],
                  const float* tE4, const float* tB4, float varE, float varB, double* loss, uint16_t* dout,
                  float* g4, float* psi4, void* stream) {
    if (int rc = check_grid(n)) return rc;
    if (!out || !s || !tE4 || !tB4 || !loss || !dout || !g4 || !psi4) return fail(IPIC_E_ARG, "null pointer");
    if (!(varE > 0.0f) || !(varB > 0.0f)) return fail(IPIC_E_ARG, "variances must be positive");
    HeadArgs h = head_args(n, out, s, inv_w);
    const long long V = voxels(n);
    h.psi = (float4*)psi4;
    h.tE = (const float4*)tE4;
    h.tB = (const float4*)tB4;
    h.g = (float4*)g4;
    h.kE = (float)(1.0 / ((double)varE * (double)V));
    h.kB = (float)(1.0 / ((double)varB * (double)V));
    h.loss = loss;
    h.dout = dout;
    const dim3 g((unsigned)((V + 255) / 256));
    hipLaunchKernelGGL(k_head_a, g, dim3(256), 0, (hipStream_t)stream, h, 1);
    hipLaunchKernelGGL(k_head_b, g, dim3(256), 0, (hipStream_t)stream, h, 1);
    hipLaunchKernelGGL(k_head_c, g, dim3(256), 0, (hipStream_t)stream, h);
    NOPCK(hipGetLastError());
    return 0;
}

int ipic_nop_backward(const int32_t n[3], const uint16_t* u, const uint16_t* wb, const float* params, void* ws,
                      const uint16_t* dout, float* grad, void* stream) {
    if (int rc = check_grid(n)) return rc;
    if (!u || !wb || !params || !ws || !dout || !grad) return fail(IPIC_E_ARG, "null pointer");
    hipStream_t st = (hipStream_t)stream;
    const Ws w = carve(ws, n, 1, nullptr);
    auto Wb = [&](int l) { return wb + loff(l); };
    auto Gw = [&](int l) { return grad + loff(l); };
    auto Gb = [&](int l) { return grad + loff(l) + 27LL * lcin(l) * lcout(l); };
    // projection: out = conv9(a4)
    if (int rc = wgrad<32, 16>(n, w.a[4], dout, Gw(9), Gb(9), st)) return rc;
    if (int rc = conv<16, 32>(n, dout, Wb(9), nullptr, w.x[4], nullptr, w.dx[0], w.dxb, 0, st)) return rc;
    int cur = 0;
    for (int b = 3; b >= 0; --b) {
        // x_{b+1} = x_b + conv_b2(g_b), g_b = silu(h_b), h_b = conv_b1(a_b), a_b = silu(x_b)
        if (int rc = wgrad<32, 32>(n, w.g[b], w.dxb, Gw(2 + 2 * b), Gb(2 + 2 * b), st)) return rc;
        if (int rc = conv<32, 32>(n, w.dxb, Wb(2 + 2 * b), nullptr, w.h[b], nullptr, nullptr, w.dhb, 0, st)) return rc;
        if (int rc = wgrad<32, 32>(n, w.a[b], w.dhb, Gw(1 + 2 * b), Gb(1 + 2 * b), st)) return rc;
        if (int rc = conv<32, 32>(n, w.dhb, Wb(1 + 2 * b), nullptr, w.x[b], w.dx[cur], w.dx[1 - cur], w.dxb, 0, st))
            return rc;
        cur = 1 - cur;
    }
    return wgrad<16, 32>(n, u, w.dxb, Gw(0), Gb(0), st);
}

int ipic_nop_adam(float* params, const float* grad, float* m, float* v, int64_t step, float lr, float beta1,
                  float beta2, float eps, uint16_t* wf, uint16_t* wb, void* stream) {
    if (!params || !grad || !m || !v) return fail(IPIC_E_ARG, "null pointer");
    if (step < 1) return fail(IPIC_E_ARG, "Adam step counts from 1");
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, (double)step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, (double)step)));
    const dim3 g((unsigned)((NPARAM + 255) / 256));
    hipLaunchKernelGGL(k_adam, g, dim3(256), 0, (hipStream_t)stream, params, grad, m, v, lr, beta1, beta2, eps, c1,
                       c2);
    NOPCK(hipGetLastError());
    if (wf) return ipic_nop_prepare(params, wf, wb, stream);
    return 0;
}

int ipic_dev_alloc(int64_t bytes, void** out) {
    if (!out || bytes < 0) return fail(IPIC_E_ARG, "bad allocation request");
    *out = nullptr;
    NOPCK(hipMalloc(out, bytes > 0 ? (size_t)bytes : 256));
    return 0;
}

int ipic_dev_free(void* p) {
    if (p) NOPCK(hipFree(p));
    return 0;
}

int ipic_dev_copy(void* dst, const void* src, int64_t bytes) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return fail(IPIC_E_ARG, "bad copy");
    if (bytes) NOPCK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDefault));
    return 0;
}

int ipic_dev_zero(void* p, int64_t bytes, void* stream) {
    if (bytes < 0 || (bytes > 0 && !p)) return fail(IPIC_E_ARG, "bad fill");
    if (bytes) NOPCK(hipMemsetAsync(p, 0, (size_t)bytes, (hipStream_t)stream));
    return 0;
}

int ipic_dev_sync(void* stream) {
    NOPCK(hipStreamSynchronize((hipStream_t)stream));
    return 0;
}

int ipic_dev_sub4(float* dst, const float* a, const float* b, int64_t count, void* stream) {
    if (count < 0 || (count > 0 && (!dst || !a || !b))) return fail(IPIC_E_ARG, "bad difference");
    if (count == 0) return 0;
    hipLaunchKernelGGL(k_sub4, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (float4*)dst, (const float4*)a, (const float4*)b, (long long)count);
    NOPCK(hipGetLastError());
    return 0;
}

}  // extern "C"
