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

}  // namespace xtr
