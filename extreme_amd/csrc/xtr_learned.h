// Learned right preconditioner: device buffers and launchers (learned_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace xtr {

struct NetGeom {
    int nx, ny, nz;
    int64_t cells;   // nx ny nz
    int linear;      // 1: the hidden layer has no relu (the network is a fixed linear map)
};

// Per-network device buffers; sizes in elements.
struct NetBuffers {
    unsigned short *x0;    // [cells][8]  BF16 input channels
    unsigned short *h;     // [cells][16] BF16 relu(layer 1)
    float *pre;            // [cells][16] layer-1 output before relu
    float *y;              // [cells][16] network output (first 6 used)
    float *tgt;            // [cells][8]  training target
    float *dy;             // [cells][8]  loss gradient
    unsigned short *dyb;   // [cells][8]  BF16 copy
    float *dh;             // [cells][16] gradient at the hidden layer
    float *dp;             // [cells][16] after relu'
    unsigned short *dpb;   // [cells][16] BF16 copy (unused by the VALU weight gradient, kept for symmetry)
    float *part;           // [chunks][27*16*8] weight-gradient partials
    unsigned short *W1;    // [224][16]   BF16 layer-1 weights (K = 27*8 rows, zero rows to 224: 14 MFMA chunks)
    unsigned short *W2;    // [27*16][16] BF16 layer-2 weights, columns 8..15 zero
    unsigned short *WT2;   // [224][16]   BF16 layer-2 weights transposed per tap (backward), zero rows to 224
    const float *feat;     // [cells][2]  input channels 6, 7 (log10 sigma, log10 f), or null (zeros)
};

constexpr int kNetW1 = 27 * 8 * 16;    // layer-1 weights [t*8 + ci][co], ci < 8, co < 16
constexpr int kNetKpad8 = 224;         // rows of the BF16 operands with K = 27*8 (whole 16-row chunks)
constexpr int kNetW2 = 27 * 16 * 8;    // layer-2 weights [t*16 + ci][co], ci < 16, co < 8

void net_pack(const NetGeom &g, const double2 *v, const double2 *z, const double *nn, const NetBuffers &b,
              hipStream_t s);
void net_forward(const NetGeom &g, const NetBuffers &b, hipStream_t s);
void net_unpack(const NetGeom &g, const double2 *v, const double *nn, const NetBuffers &b, double2 *z, hipStream_t s);
int net_loss_blocks(const NetGeom &g);
void net_loss(const NetGeom &g, const NetBuffers &b, double *partial, hipStream_t s);
int net_wgrad_chunks(const NetGeom &g);
void net_backward(const NetGeom &g, const NetBuffers &b, float *g1, float *g2, float scale, int accumulate,
                  hipStream_t s);
void net_adam(float *w, float *m, float *v2, const float *grad, int n, float gscale, float lr, int step,
              hipStream_t s);
void net_weights(const float *w1, const float *w2, const NetBuffers &b, hipStream_t s);

// ---- spectral branch: per-lateral-mode channel mixing on the zero-padded lateral spectrum ----------
// Modes kx, ky in -K..K of the 2Nx x 2Ny padded plane, mode m = (kx + K)(2K + 1) + (ky + K); channels
// q = c Nz + k (component c, layer k), Q = 3 Nz. U[m][q] = s0 sum_{i<Nx, j<Ny} u(i, j, q)
// e^{-2 pi i (kx i / 2Nx + ky j / 2Ny)} with s0 = 1 / sqrt(Nx Ny), Y[m] = (W0[m] + phi W1[m]) bf16(U[m]),
// y(i, j, q) += s0 sum_m Y[m][q] e^{+2 pi i (...)}. W0, W1: complex Q x Q per mode (row = output).
struct SpecGeom {
    int nx, ny, nz, K;
    __host__ __device__ int nm() const { return (2 * K + 1) * (2 * K + 1); }
    __host__ __device__ int q() const { return 3 * nz; }
    __host__ __device__ int64_t weights() const { return 2LL * nm() * q() * q(); }   // complex entries of W0, W1
};
struct SpecBuffers {
    const float2 *twx, *twy;   // e^{-2 pi i t / 2Nx}, t < 2Nx; e^{-2 pi i t / 2Ny}, t < 2Ny (FP32)
    float2 *T;                 // [Nx][2K+1][Q] partial transform
    float2 *U;                 // [NM][Q] transform of the input, rounded to BF16 values (the MFMA operand)
    float2 *Y;                 // [NM][Q] mixed modes
    float2 *gY;                // [NM][Q] transform of the loss gradient
    unsigned short *WB;        // [2][NM][Q][Q][2] BF16 copies of W0, W1 (re, im)
};
// the network's input channels 0..5 (BF16 [cells][8]) -> U; mixing -> Y; y[cells][16] channels 0..5 += inverse
void spec_forward(const SpecGeom &g, const NetBuffers &b, const SpecBuffers &sb, float phi, hipStream_t s);
// dY[m] = transform of dy (FP32 [cells][8]); grad (+)= scale dY conj(U)^T into [2][NM][Q][Q] complex
// (W1's part times phi)
void spec_backward(const SpecGeom &g, const NetBuffers &b, const SpecBuffers &sb, float phi, float *grad, float scale,
                   int accumulate, hipStream_t s);
// BF16 copies of the FP32 master weights w [2][NM][Q][Q][2]
void spec_weights(const SpecGeom &g, const float *w, unsigned short *WB, hipStream_t s);

}  // namespace xtr
