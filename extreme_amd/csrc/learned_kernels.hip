// Learned right preconditioner for the flexible FGMRES (no counterpart in the reference): a
// two-layer 3-D convolution network on the residual field, z = s (v/s + N(v/s)), s = rms(v),
//   N = conv3(H -> 8) o relu o conv3(8 -> H),   3 x 3 x 3 taps, zero padding, no biases,
// over the cells (i, j, k) of the anomaly grid with channels (Re, Im) x (x, y, z) of the vector
// plus two feature channels of the cell (log10 sigma, log10 f; zero when the caller sets none).
// Activations and weights are BF16 for the MFMAs, accumulation FP32.
//
// The convolutions run as implicit GEMMs on v_mfma_f32_16x16x16_bf16: one wave computes 16 cells x
// 16 output channels, K = 27 taps x C_in input channels in chunks of 16 (four consecutive channels
// of one tap per lane). The backward pass into the hidden layer is the same kernel with the tap
// offsets negated and the weight matrix transposed per tap; the weight gradients are FP32 sums over
// the cells, two-stage (per-chunk partials, then a fixed-order sum) so they are deterministic.
#include "xtr_common.h"
#include "xtr_learned.h"

#include <algorithm>
#include <cmath>

namespace xtr {
namespace lp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float f)   // round to nearest even
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline void cell_ijk(const NetGeom &g, int64_t cell, int &i, int &j, int &k)
{
    k = (int)(cell % g.nz);
    const int64_t t = cell / g.nz;
    j = (int)(t % g.ny);
    i = (int)(t / g.ny);
}

// neighbour of (i, j, k) through tap t (offsets t / 9 - 1, (t / 3) % 3 - 1, t % 3 - 1, negated when
// FLIP), or -1 outside the grid
template <bool FLIP>
__device__ inline int64_t neighbour(const NetGeom &g, int i, int j, int k, int t)
{
    int dx = t / 9 - 1, dy = (t / 3) % 3 - 1, dz = t % 3 - 1;
    if (FLIP) {
        dx = -dx;
        dy = -dy;
        dz = -dz;
    }
    const int ni = i + dx, nj = j + dy, nk = k + dz;
    if (ni < 0 || ni >= g.nx || nj < 0 || nj >= g.ny || nk < 0 || nk >= g.nz) return -1;
    return ((int64_t)ni * g.ny + nj) * g.nz + nk;
}

// out[cell][co] (16 FP32 columns) = sum over taps t and input channels ci of
// W[t CIN + ci][co] in[neighbour(cell, t)][ci]; in: BF16 [cells][CIN], W: BF16 [KPAD][16].
// One wave per 16 cells; four waves per workgroup. ACT 1 also writes relu(out) as BF16 [cells][16],
// ACT 2 writes out itself as BF16 (the linear network), ACT 0 nothing.
template <int CIN, bool FLIP, int ACT>
__global__ __launch_bounds__(256) void conv_kernel(NetGeom g, const unsigned short *__restrict__ in,
                                                   const unsigned short *__restrict__ W, float *__restrict__ out,
                                                   unsigned short *__restrict__ act)
{
    static_assert(CIN % 4 == 0, "four channels of one tap per lane");
    constexpr int K = 27 * CIN, NCH = (K + 15) / 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t c0 = ((int64_t)blockIdx.x * 4 + wave) * 16;
    if (c0 >= g.cells) return;                           // the whole wave: c0 is wave-uniform
    const int row = lane & 15, kg = lane >> 4;
    const int64_t cell = c0 + row;
    const bool live = cell < g.cells;
    int i = 0, j = 0, k = 0;
    if (live) cell_ijk(g, cell, i, j, k);
    f32x4 acc = {0, 0, 0, 0};
    for (int ch = 0; ch < NCH; ++ch) {
        const int k0 = ch * 16 + 4 * kg;                 // this lane's four K indices k0 .. k0 + 3
        bf16x4 a = {0, 0, 0, 0};
        if (live && k0 < K) {
            const int64_t nb = neighbour<FLIP>(g, i, j, k, k0 / CIN);
            if (nb >= 0) a = *reinterpret_cast<const bf16x4 *>(in + nb * CIN + (k0 % CIN));
        }
        bf16x4 b;
#pragma unroll
        for (int r = 0; r < 4; ++r) b[r] = (short)W[(int64_t)(k0 + r) * 16 + row];   // column = lane % 16
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {                        // D[4 kg + r][lane % 16]
        const int64_t oc = c0 + 4 * kg + r;
        if (oc < g.cells) {
            out[oc * 16 + row] = acc[r];
            if (ACT == 1) act[oc * 16 + row] = f2bf(acc[r] > 0 ? acc[r] : 0.0f);
            if (ACT == 2) act[oc * 16 + row] = f2bf(acc[r]);
        }
    }
}

// Input channels of v / s and the training target (z - v) / s, s = sqrt(nn / n); 0 when nn = 0.
__global__ void pack_kernel(NetGeom g, const double2 *__restrict__ v, const double2 *__restrict__ z,
                            const double *nn, const float *__restrict__ feat, unsigned short *__restrict__ x0,
                            float *__restrict__ tgt)
{
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= g.cells) return;
    const double n = (double)g.cells * 3;
    const double rs = *nn > 0 ? 1.0 / sqrt(*nn / n) : 0.0;
    const int64_t col = cell / g.nz, k = cell % g.nz;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int64_t e = (col * 3 + c) * g.nz + k;
        const double2 a = v[e];
        x0[cell * 8 + 2 * c] = f2bf((float)(a.x * rs));
        x0[cell * 8 + 2 * c + 1] = f2bf((float)(a.y * rs));
        if (tgt) {
            const double2 b = z[e];
            tgt[cell * 8 + 2 * c] = (float)((b.x - a.x) * rs);
            tgt[cell * 8 + 2 * c + 1] = (float)((b.y - a.y) * rs);
        }
    }
    // channels 6 and 7: the cell's features (log10 sigma, log10 f), not scaled by s; 0 without them
    x0[cell * 8 + 6] = feat ? f2bf(feat[2 * cell]) : 0;
    x0[cell * 8 + 7] = feat ? f2bf(feat[2 * cell + 1]) : 0;
    if (tgt) {
        tgt[cell * 8 + 6] = 0;
        tgt[cell * 8 + 7] = 0;
    }
}

// z = v + s y (y: the network's output channels, [cells][16], first six used)
__global__ void unpack_kernel(NetGeom g, const double2 *__restrict__ v, const float *__restrict__ y,
                              const double *nn, double2 *__restrict__ z)
{
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= g.cells) return;
    const double s = sqrt(*nn / ((double)g.cells * 3));
    const int64_t col = cell / g.nz, k = cell % g.nz;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int64_t e = (col * 3 + c) * g.nz + k;
        const double2 a = v[e];
        z[e] = make_double2(a.x + s * (double)y[cell * 16 + 2 * c], a.y + s * (double)y[cell * 16 + 2 * c + 1]);
    }
}

// Squared error of one sample over the six used channels: dY = 2 (y - t) (FP32 [cells][8] and a
// BF16 copy) and the per-block partial sums of (y - t)^2 and t^2. The loss is the relative error
// sum (y - t)^2 / sum t^2; its 1 / sum t^2 is applied to the weight gradients afterwards.
__global__ __launch_bounds__(256) void loss_kernel(NetGeom g, const float *__restrict__ y, const float *__restrict__ tgt,
                                                   float *__restrict__ dy, unsigned short *__restrict__ dyb,
                                                   double *partial)
{
    __shared__ double sh[2][4];
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double se = 0, st = 0;
    if (cell < g.cells) {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float d = c < 6 ? y[cell * 16 + c] - tgt[cell * 8 + c] : 0.0f;
            se += (double)d * d;
            st += (double)tgt[cell * 8 + c] * tgt[cell * 8 + c];
            const float gval = 2.0f * d;
            dy[cell * 8 + c] = gval;
            dyb[cell * 8 + c] = f2bf(gval);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        se += __shfl_down(se, o, 64);
        st += __shfl_down(st, o, 64);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[0][wv] = se;
        sh[1][wv] = st;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            a += sh[0][w];
            b += sh[1][w];
        }
        partial[2 * blockIdx.x] = a;
        partial[2 * blockIdx.x + 1] = b;
    }
}

// dH (FP32 [cells][16]) masked by relu'(pre) (h > 0), or passed through for the linear network,
// as FP32 and BF16
__global__ void relu_back_kernel(int64_t n, const float *__restrict__ dh, const unsigned short *__restrict__ h,
                                 float *__restrict__ dp, unsigned short *__restrict__ dpb, int linear)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float v = (linear || bf2f(h[e]) > 0) ? dh[e] : 0.0f;
    dp[e] = v;
    dpb[e] = f2bf(v);
}

// Weight-gradient partials of one chunk of cells: part[chunk][t CIN + ci][co] =
// sum over the chunk's cells of in[neighbour(cell, t)][ci] dOut[cell][co], with in BF16
// [cells][CIN] and dOut FP32 [cells][DSTRIDE], co < COUT. Thread e owns entries e, e + 256, ...
template <int CIN, int COUT, int DSTRIDE>
__global__ __launch_bounds__(256) void wgrad_kernel(NetGeom g, int64_t chunk, const unsigned short *__restrict__ in,
                                                    const float *__restrict__ dout, float *__restrict__ part)
{
    constexpr int NE = 27 * CIN, PER = (NE + 255) / 256, SUB = 64;
    __shared__ float sd[SUB][COUT];
    const int64_t cbeg = (int64_t)blockIdx.x * chunk;
    const int64_t cend = cbeg + chunk < g.cells ? cbeg + chunk : g.cells;
    float acc[PER][COUT];
#pragma unroll
    for (int p = 0; p < PER; ++p)
#pragma unroll
        for (int c = 0; c < COUT; ++c) acc[p][c] = 0;
    for (int64_t s0 = cbeg; s0 < cend; s0 += SUB) {
        __syncthreads();
        for (int e = threadIdx.x; e < SUB * COUT; e += blockDim.x) {
            const int64_t cell = s0 + e / COUT;
            sd[e / COUT][e % COUT] = cell < cend ? dout[cell * DSTRIDE + e % COUT] : 0.0f;
        }
        __syncthreads();
        const int nsub = cend - s0 < SUB ? (int)(cend - s0) : SUB;
        for (int q = 0; q < nsub; ++q) {
            int i, j, k;
            cell_ijk(g, s0 + q, i, j, k);
#pragma unroll
            for (int p = 0; p < PER; ++p) {
                const int en = threadIdx.x + 256 * p;
                if (en < NE) {
                    const int64_t nb = neighbour<false>(g, i, j, k, en / CIN);
                    if (nb >= 0) {
                        const float x = bf2f(in[nb * CIN + en % CIN]);
#pragma unroll
                        for (int c = 0; c < COUT; ++c) acc[p][c] += x * sd[q][c];
                    }
                }
            }
        }
    }
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int en = threadIdx.x + 256 * p;
        if (en < NE)
#pragma unroll
            for (int c = 0; c < COUT; ++c) part[((int64_t)blockIdx.x * NE + en) * COUT + c] = acc[p][c];
    }
}

// grad[e] (+)= scale x (sum over chunks of part[chunk][e], fixed order)
__global__ void wgrad_reduce_kernel(int nchunk, int ne, const float *__restrict__ part, float *__restrict__ grad,
                                    float scale, int accumulate)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    float s = 0;
    for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * ne + e];
    grad[e] = accumulate ? grad[e] + scale * s : scale * s;
}

// Adam on the FP32 master weights (bias-corrected), gradient scaled by gscale
__global__ void adam_kernel(int n, float *__restrict__ w, float *__restrict__ m, float *__restrict__ v2,
                            const float *__restrict__ grad, float gscale, float lr, float b1, float b2, float eps,
                            float c1, float c2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gv = grad[e] * gscale;
    m[e] = b1 * m[e] + (1 - b1) * gv;
    v2[e] = b2 * v2[e] + (1 - b2) * gv * gv;
    w[e] -= lr * (m[e] / c1) / (sqrtf(v2[e] / c2) + eps);
}

// BF16 MFMA operands from the master weights: W1 [27*8][16] <- w1 [27*8][16]; W2 [27*16][16] <- w2
// [27*16][8] (columns 8..15 zero); WT2 [27*8][16]: WT2[t 8 + co][ci] = w2[t 16 + ci][co].
__global__ void weights_kernel(const float *__restrict__ w1, const float *__restrict__ w2, unsigned short *W1,
                               unsigned short *W2, unsigned short *WT2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < 27 * 8 * 16) W1[e] = f2bf(w1[e]);
    if (e < 27 * 16 * 16) {
        const int r = e / 16, c = e % 16;
        W2[e] = c < 8 ? f2bf(w2[r * 8 + c]) : (unsigned short)0;
    }
    if (e < 27 * 8 * 16) {
        const int r = e / 16, ci = e % 16, t = r / 8, co = r % 8;
        WT2[e] = f2bf(w2[(t * 16 + ci) * 8 + co]);
    }
}

}  // namespace lp

using namespace lp;

static unsigned grid_of(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void net_forward(const NetGeom &g, const NetBuffers &b, hipStream_t s)
{
    const unsigned tiles = grid_of(grid_of(g.cells, 16), 4);
    if (g.linear)
        hipLaunchKernelGGL((conv_kernel<8, false, 2>), dim3(tiles), dim3(256), 0, s, g, b.x0, b.W1, b.pre, b.h);
    else
        hipLaunchKernelGGL((conv_kernel<8, false, 1>), dim3(tiles), dim3(256), 0, s, g, b.x0, b.W1, b.pre, b.h);
    XTR_LAUNCH_CHECK("conv_kernel (layer 1)");
    hipLaunchKernelGGL((conv_kernel<16, false, 0>), dim3(tiles), dim3(256), 0, s, g, b.h, b.W2, b.y,
                       (unsigned short *)nullptr);
    XTR_LAUNCH_CHECK("conv_kernel (layer 2)");
}

void net_pack(const NetGeom &g, const double2 *v, const double2 *z, const double *nn, const NetBuffers &b, hipStream_t s)
{
    hipLaunchKernelGGL(pack_kernel, dim3(grid_of(g.cells, 256)), dim3(256), 0, s, g, v, z, nn, b.feat, b.x0,
                       z ? b.tgt : nullptr);
    XTR_LAUNCH_CHECK("pack_kernel");
}

void net_unpack(const NetGeom &g, const double2 *v, const double *nn, const NetBuffers &b, double2 *z, hipStream_t s)
{
    hipLaunchKernelGGL(unpack_kernel, dim3(grid_of(g.cells, 256)), dim3(256), 0, s, g, v, b.y, nn, z);
    XTR_LAUNCH_CHECK("unpack_kernel");
}

int net_loss_blocks(const NetGeom &g) { return (int)grid_of(g.cells, 256); }

void net_loss(const NetGeom &g, const NetBuffers &b, double *partial, hipStream_t s)
{
    hipLaunchKernelGGL(loss_kernel, dim3(grid_of(g.cells, 256)), dim3(256), 0, s, g, b.y, b.tgt, b.dy, b.dyb,
                       partial);
    XTR_LAUNCH_CHECK("loss_kernel");
}

int net_wgrad_chunks(const NetGeom &g) { return (int)std::min<int64_t>(1024, grid_of(g.cells, 256)); }

void net_backward(const NetGeom &g, const NetBuffers &b, float *g1, float *g2, float scale, int accumulate,
                  hipStream_t s)
{
    const int nchunk = net_wgrad_chunks(g);
    const int64_t chunk = (g.cells + nchunk - 1) / nchunk;
    // layer 2: dW2[t 16 + ci][co] = sum h[nb][ci] dY[cell][co]
    hipLaunchKernelGGL((wgrad_kernel<16, 8, 8>), dim3(nchunk), dim3(256), 0, s, g, chunk, b.h, b.dy, b.part);
    XTR_LAUNCH_CHECK("wgrad_kernel (layer 2)");
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(grid_of(27 * 16 * 8, 256)), dim3(256), 0, s, nchunk, 27 * 16 * 8,
                       b.part, g2, scale, accumulate);
    XTR_LAUNCH_CHECK("wgrad_reduce_kernel (layer 2)");
    // into the hidden layer: dH[cell][ci] = sum W2[t 16 + ci][co] dY[cell - off t][co], then relu'
    const unsigned tiles = grid_of(grid_of(g.cells, 16), 4);
    hipLaunchKernelGGL((conv_kernel<8, true, 0>), dim3(tiles), dim3(256), 0, s, g, b.dyb, b.WT2, b.dh,
                       (unsigned short *)nullptr);
    XTR_LAUNCH_CHECK("conv_kernel (backward)");
    hipLaunchKernelGGL(relu_back_kernel, dim3(grid_of(g.cells * 16, 256)), dim3(256), 0, s, g.cells * 16, b.dh, b.h,
                       b.dp, b.dpb, g.linear);
    XTR_LAUNCH_CHECK("relu_back_kernel");
    // layer 1: dW1[t 8 + ci][co] = sum x0[nb][ci] dP[cell][co]
    hipLaunchKernelGGL((wgrad_kernel<8, 16, 16>), dim3(nchunk), dim3(256), 0, s, g, chunk, b.x0, b.dp, b.part);
    XTR_LAUNCH_CHECK("wgrad_kernel (layer 1)");
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(grid_of(27 * 8 * 16, 256)), dim3(256), 0, s, nchunk, 27 * 8 * 16,
                       b.part, g1, scale, accumulate);
    XTR_LAUNCH_CHECK("wgrad_reduce_kernel (layer 1)");
}

void net_adam(float *w, float *m, float *v2, const float *grad, int n, float gscale, float lr, int step, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1 - std::pow(b1, (float)step), c2 = 1 - std::pow(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, n, w, m, v2, grad, gscale, lr, b1, b2,
                       1e-8f, c1, c2);
    XTR_LAUNCH_CHECK("adam_kernel");
}

void net_weights(const float *w1, const float *w2, const NetBuffers &b, hipStream_t s)
{
    hipLaunchKernelGGL(weights_kernel, dim3(grid_of(27 * 16 * 16, 256)), dim3(256), 0, s, w1, w2, b.W1, b.W2, b.WT2);
    XTR_LAUNCH_CHECK("weights_kernel");
}

}  // namespace xtr
