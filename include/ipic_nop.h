/*
 * ipic_nop.h -- C ABI of the learned field step in libipic_gpu.so: a 3D-convolutional neural
 * operator that stands in for k leapfrog substeps of the field update.
 *
 * The reference has no counterpart: its field policies are the leapfrog / forward solvers of
 * field.h:272-402 behind the solver hook of simulator.h:252-268.  The operator maps the field
 * state of a step to the state k substeps later, with J held fixed as in the leapfrog
 * substeps of ipic_gpu_step:
 *
 *   u   = (dE / sE, dBc / sB, J / sJ)           9 channels on the index lattice (node or
 *                                               centre i share index i), zero-padded to 16
 *   x0  = conv_lift(u)                          9 -> 32
 *   x_{b+1} = x_b + conv_b2(silu(conv_b1(silu(x_b))))      b = 0..3, 32 -> 32
 *   o   = conv_proj(silu(x_4))                  32 -> 6 (padded to 16)
 *   E'  = dE  + so_E   * o[0:3]
 *   Bc' = dBc - curl+ (so_P * o[3:6])           (divergence-free update by construction)
 *
 * Every conv is 3x3x3 with periodic wrap and a bias; weights and activations are BF16, the
 * MFMA accumulation and the residual stream FP32.
 *
 * All array arguments are DEVICE pointers; `stream` is a hipStream_t (NULL: the null stream).
 * Grids need nx % 4 == 0, ny % 4 == 0, nz % 16 == 0.  Calls return 0 or a negative
 * IPIC_E_* code (ipic_gpu.h), with a message from ipic_nop_last_error().
 *
 * Parameter vector (FP32, ipic_nop_param_count() floats): per layer l = 0..9 the weights
 * W[27][cout_p][cin_p] (tap = (dx+1)*9 + (dy+1)*3 + (dz+1), output channel, input channel;
 * padded channels are zero) followed by the bias[cout_p].  cin_p = 16 for layer 0, else 32;
 * cout_p = 16 for layer 9, else 32.  Layers 1+2b, 2+2b are conv_b1, conv_b2.
 */
#ifndef IPIC_NOP_H
#define IPIC_NOP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPIC_NOP_LAYERS 10
#define IPIC_NOP_WIDTH 32
#define IPIC_NOP_IN 9
#define IPIC_NOP_OUT 6

/* Scales of the operator's inputs and outputs (see the formula above). */
typedef struct ipic_nop_scales {
    float in_E, in_B, in_J;       /* sE, sB, sJ */
    float out_E, out_P;           /* so_E, so_P */
} ipic_nop_scales;

const char* ipic_nop_last_error(void);
int64_t ipic_nop_param_count(void);
/* offsets (in floats) of layer l's weights and bias, and its padded channel counts */
int ipic_nop_layer(int32_t layer, int64_t* w_offset, int64_t* b_offset, int32_t* cin_p, int32_t* cout_p);
/* device workspace of forward (train = 0) or forward + backward (train = 1) on an n^3 grid */
int64_t ipic_nop_workspace_bytes(const int32_t n[3], int32_t train);

/* BF16 weights from the FP32 parameters: wf in the forward layout, wb (nullable) with taps
 * mirrored and channels swapped for the data-gradient convolutions.  Each holds
 * ipic_nop_param_count() uint16 (bias slots unused). */
int ipic_nop_prepare(const float* params, uint16_t* wf, uint16_t* wb, void* stream);

/* u (n^3 x 16 BF16) from float4 arrays dE (nodes), dBc (centres), J (nodes). */
int ipic_nop_pack(const int32_t n[3], const float* dE4, const float* dBc4, const float* J4,
                  const ipic_nop_scales* s, uint16_t* u, void* stream);

/* Forward pass; out = o (n^3 x 16 FP32, channels 6..15 zero).  With train = 1 the
 * activations needed by ipic_nop_backward stay in ws. */
int ipic_nop_forward(const int32_t n[3], const uint16_t* u, const uint16_t* wf, const float* params,
                     void* ws, int32_t train, float* out, void* stream);

/* Applies o to the field state in place: dE4 += so_E o[0:3]; dBc4 -= curl+(so_P o[3:6]).
 * inv_w: 1 / cell width per axis.  psi4: n^3 float4 scratch. */
int ipic_nop_apply(const int32_t n[3], const float* out, const ipic_nop_scales* s, const float inv_w[3],
                   float* dE4, float* dBc4, float* psi4, void* stream);

/* ipic_nop_apply without the psi4 workspace, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops); it has no counterpart in the reference.
 * The potential lives in a file-scope device array of the library, IPIC_NOP_NOALLOC_MAX_VOXELS
 * float4 (32 MiB) on each GPU the library runs on, held until the process ends.  Allocates
 * nothing: no hipMalloc, hipFree or allocating library call, and its error messages are static
 * strings.  The first call on a device looks the array up (hipGetSymbolAddress): make it
 * before any capture.  Every call on one GPU shares that array, so calls that overlap on one
 * GPU (concurrent streams) race; each call writes every entry it reads.  Grids above the
 * maximum are refused with IPIC_E_RANGE. */
#define IPIC_NOP_NOALLOC_MAX_VOXELS 2097152
int ipic_nop_apply_noalloc(const int32_t n[3], const float* out, const ipic_nop_scales* s, const float inv_w[3],
                           float* dE4, float* dBc4, void* stream);

/* Training loss L = mean over i of |dE_pred - tE|^2 / varE + |dB_pred - tB|^2 / varB, where
 * (dE_pred, dB_pred) are the increments ipic_nop_apply would make and tE, tB (n^3 float4)
 * the targets.  Adds L to *loss (one FP64 on the device) and writes dout = dL/do
 * (n^3 x 16 BF16).  g4, psi4: n^3 float4 scratch each. */
int ipic_nop_loss(const int32_t n[3], const float* out, const ipic_nop_scales* s, const float inv_w[3],
                  const float* tE4, const float* tB4, float varE, float varB, double* loss,
                  uint16_t* dout, float* g4, float* psi4, void* stream);

/* Backward pass after ipic_nop_forward(train = 1) on the same ws: adds dL/dparams to grad. */
int ipic_nop_backward(const int32_t n[3], const uint16_t* u, const uint16_t* wb, const float* params,
                      void* ws, const uint16_t* dout, float* grad, void* stream);

/* One Adam step over the parameter vector (bias-corrected, step >= 1), then refreshes wf and
 * wb (nullable).  grad is left as is. */
int ipic_nop_adam(float* params, const float* grad, float* m, float* v, int64_t step, float lr, float beta1,
                  float beta2, float eps, uint16_t* wf, uint16_t* wb, void* stream);

/* Device memory plumbing for hosts without a GPU array library (HIP runtime calls). */
int ipic_dev_alloc(int64_t bytes, void** out);
int ipic_dev_free(void* p);
int ipic_dev_copy(void* dst, const void* src, int64_t bytes);   /* any direction, synchronous */
int ipic_dev_zero(void* p, int64_t bytes, void* stream);
int ipic_dev_sync(void* stream);
/* dst[i] = a[i] - b[i] for count float4 entries (field increments as training targets) */
int ipic_dev_sub4(float* dst, const float* a, const float* b, int64_t count, void* stream);

#ifdef __cplusplus
}
#endif
#endif
