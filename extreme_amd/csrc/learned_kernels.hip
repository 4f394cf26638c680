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

// ---- spectral branch ---------------------------------------------------------------------------------
// Per-lateral-mode channel mixing on the zero-padded 2Nx x 2Ny spectrum, modes |kx|, |ky| <= K. The
// truncated transforms are separable sums with FP32 twiddle tables (t = k n mod 2N); the mixing of
// the 3 Nz complex channels of each mode is a real GEMM on v_mfma_f32_16x16x16_bf16 in the block
// form [Wr -Wi; Wi Wr] [Re U; Im U], A = BF16(W0 + phi W1) per element, B = BF16(U).
constexpr int kSpecMaxK = 12;

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cmulf_conj(float2 a, float2 b)   // a conj(b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ int twidx(int kk, int n, int p)   // (kk n) mod p, kk may be negative
{
    int t = (kk * n) % p;
    return t < 0 ? t + p : t;
}
__device__ __forceinline__ float chan(const unsigned short *src, int64_t e) { return bf2f(src[e]); }
__device__ __forceinline__ float chan(const float *src, int64_t e) { return src[e]; }

// T[i][b][q] = sum_j u(i, j, q) e^{-2 pi i ky j / 2Ny}, ky = b - K; u from channels 2c, 2c + 1 of
// src[cells][8] (q = c Nz + k). One thread per (i, q).
template <typename SRC>
__global__ __launch_bounds__(256) void spec_dft_y_kernel(SpecGeom g, const SRC *__restrict__ src,
                                                         const float2 *__restrict__ twy, float2 *__restrict__ T)
{
    const int Q = g.q(), nb = 2 * g.K + 1, py = 2 * g.ny;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)g.nx * Q) return;
    const int i = (int)(e / Q), q = (int)(e % Q), c = q / g.nz, k = q % g.nz;
    float2 acc[2 * kSpecMaxK + 1];
#pragma unroll
    for (int b = 0; b < 2 * kSpecMaxK + 1; ++b) acc[b] = make_float2(0, 0);
    for (int j = 0; j < g.ny; ++j) {
        const int64_t cell = ((int64_t)i * g.ny + j) * g.nz + k;
        const float2 u = make_float2(chan(src, cell * 8 + 2 * c), chan(src, cell * 8 + 2 * c + 1));
#pragma unroll
        for (int b = 0; b < 2 * kSpecMaxK + 1; ++b)
            if (b < nb) {
                const float2 w = twy[twidx(b - g.K, j, py)];
                acc[b].x = fmaf(u.x, w.x, fmaf(-u.y, w.y, acc[b].x));
                acc[b].y = fmaf(u.x, w.y, fmaf(u.y, w.x, acc[b].y));
            }
    }
#pragma unroll
    for (int b = 0; b < 2 * kSpecMaxK + 1; ++b)
        if (b < nb) T[((int64_t)i * nb + b) * Q + q] = acc[b];
}

// U[a nb + b][q] = s0 sum_i T[i][b][q] e^{-2 pi i kx i / 2Nx}, kx = a - K, rounded to BF16 values when
// round (the MFMA operand). One thread per (b, q).
__global__ __launch_bounds__(256) void spec_dft_x_kernel(SpecGeom g, const float2 *__restrict__ T,
                                                         const float2 *__restrict__ twx, float s0,
                                                         float2 *__restrict__ U, int round)
{
    const int Q = g.q(), nb = 2 * g.K + 1, px = 2 * g.nx;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * Q) return;
    const int b = e / Q, q = e % Q;
    float2 acc[2 * kSpecMaxK + 1];
#pragma unroll
    for (int a = 0; a < 2 * kSpecMaxK + 1; ++a) acc[a] = make_float2(0, 0);
    for (int i = 0; i < g.nx; ++i) {
        const float2 t = T[((int64_t)i * nb + b) * Q + q];
#pragma unroll
        for (int a = 0; a < 2 * kSpecMaxK + 1; ++a)
            if (a < nb) {
                const float2 w = twx[twidx(a - g.K, i, px)];
                acc[a].x = fmaf(t.x, w.x, fmaf(-t.y, w.y, acc[a].x));
                acc[a].y = fmaf(t.x, w.y, fmaf(t.y, w.x, acc[a].y));
            }
    }
#pragma unroll
    for (int a = 0; a < 2 * kSpecMaxK + 1; ++a)
        if (a < nb) {
            float2 v = make_float2(s0 * acc[a].x, s0 * acc[a].y);
            if (round) v = make_float2(bf2f(f2bf(v.x)), bf2f(f2bf(v.y)));
            U[((int64_t)a * nb + b) * Q + q] = v;
        }
}

// Y[m] = A U[m] with A the 2Q x 2Q real block form of W0[m] + phi W1[m] (BF16 per element): one wave
// per (mode, 16-row tile), K = 2Q in chunks of 16, one live B column (lane % 16 == 0).
__global__ __launch_bounds__(256) void spec_mix_kernel(SpecGeom g, const unsigned short *__restrict__ WB, float phi,
                                                       const float2 *__restrict__ U, float2 *__restrict__ Y)
{
    const int Q = g.q(), Q2 = 2 * Q, tiles = Q2 / 16, nm = g.nm();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t wg = (int64_t)blockIdx.x * 4 + wave;
    if (wg >= (int64_t)nm * tiles) return;             // wave-uniform
    const int m = (int)(wg / tiles), tile = (int)(wg % tiles);
    const int row = lane & 15, kg = lane >> 4;
    const int R = tile * 16 + row, qr = R % Q;
    const bool top = R < Q;
    const unsigned short *w0 = WB + ((int64_t)m * Q + qr) * Q * 2;
    const unsigned short *w1 = WB + (((int64_t)nm + m) * Q + qr) * Q * 2;
    const float2 *u = U + (int64_t)m * Q;
    f32x4 acc = {0, 0, 0, 0};
    for (int ch = 0; ch < tiles; ++ch) {
        const int k0 = ch * 16 + 4 * kg;
        bf16x4 a, bv;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int kk = k0 + r, qk = kk % Q;
            const bool left = kk < Q;
            const float wr = bf2f(w0[2 * qk]) + phi * bf2f(w1[2 * qk]);
            const float wi = bf2f(w0[2 * qk + 1]) + phi * bf2f(w1[2 * qk + 1]);
            const float val = top ? (left ? wr : -wi) : (left ? wi : wr);
            a[r] = (short)f2bf(val);
            const float ub = row == 0 ? (left ? u[qk].x : u[qk].y) : 0.0f;
            bv[r] = (short)f2bf(ub);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc, 0, 0, 0);
    }
    if (row == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {                      // D[4 kg + r][0]
            const int RR = tile * 16 + 4 * kg + r;
            float *dst = reinterpret_cast<float *>(Y + (int64_t)m * Q + (RR % Q)) + (RR < Q ? 0 : 1);
            *dst = acc[r];
        }
    }
}

// T[i][b][q] = sum_a Y[a nb + b][q] e^{+2 pi i kx i / 2Nx}. One thread per (i, b, q).
__global__ __launch_bounds__(256) void spec_idft_x_kernel(SpecGeom g, const float2 *__restrict__ Y,
                                                          const float2 *__restrict__ twx, float2 *__restrict__ T)
{
    const int Q = g.q(), nb = 2 * g.K + 1, px = 2 * g.nx;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)g.nx * nb * Q) return;
    const int q = (int)(e % Q), b = (int)((e / Q) % nb), i = (int)(e / ((int64_t)Q * nb));
    float2 acc = make_float2(0, 0);
    for (int a = 0; a < nb; ++a) {
        const float2 y = Y[((int64_t)a * nb + b) * Q + q];
        const float2 w = twx[twidx(a - g.K, i, px)];
        acc = make_float2(fmaf(y.x, w.x, fmaf(y.y, w.y, acc.x)), fmaf(y.y, w.x, fmaf(-y.x, w.y, acc.y)));   // y conj(w)
    }
    T[e] = acc;
}

// y[cell][2c], y[cell][2c + 1] += Re, Im of s0 sum_b T[i][b][q] e^{+2 pi i ky j / 2Ny}. One thread per
// (i, j, q).
__global__ __launch_bounds__(256) void spec_idft_y_kernel(SpecGeom g, const float2 *__restrict__ T,
                                                          const float2 *__restrict__ twy, float s0, float *y)
{
    const int Q = g.q(), nb = 2 * g.K + 1, py = 2 * g.ny;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)g.nx * g.ny * Q) return;
    const int q = (int)(e % Q);
    const int64_t ij = e / Q;
    const int j = (int)(ij % g.ny), i = (int)(ij / g.ny), c = q / g.nz, k = q % g.nz;
    float2 acc = make_float2(0, 0);
    for (int b = 0; b < nb; ++b) {
        const float2 t = T[((int64_t)i * nb + b) * Q + q];
        const float2 w = twy[twidx(b - g.K, j, py)];
        acc = make_float2(fmaf(t.x, w.x, fmaf(t.y, w.y, acc.x)), fmaf(t.y, w.x, fmaf(-t.x, w.y, acc.y)));
    }
    const int64_t cell = ij * g.nz + k;
    y[cell * 16 + 2 * c] += s0 * acc.x;
    y[cell * 16 + 2 * c + 1] += s0 * acc.y;
}

// grad[w][m][a][b] (+)= scale f_w gY[m][a] conj(U[m][b]), f_0 = 1, f_1 = phi; complex entries (re, im).
__global__ __launch_bounds__(256) void spec_wgrad_kernel(SpecGeom g, const float2 *__restrict__ gY,
                                                         const float2 *__restrict__ U, float phi, float scale,
                                                         float *__restrict__ grad, int accumulate)
{
    const int Q = g.q();
    const int64_t per = (int64_t)g.nm() * Q * Q;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per) return;
    const int b = (int)(e % Q), a = (int)((e / Q) % Q);
    const int64_t m = e / ((int64_t)Q * Q);
    const float2 v = cmulf_conj(gY[m * Q + a], U[m * Q + b]);
    float2 *g0 = reinterpret_cast<float2 *>(grad) + e, *g1 = reinterpret_cast<float2 *>(grad) + per + e;
    const float2 d0 = make_float2(scale * v.x, scale * v.y), d1 = make_float2(scale * phi * v.x, scale * phi * v.y);
    *g0 = accumulate ? make_float2(g0->x + d0.x, g0->y + d0.y) : d0;
    *g1 = accumulate ? make_float2(g1->x + d1.x, g1->y + d1.y) : d1;
}

__global__ void spec_weights_kernel(int64_t n, const float *__restrict__ w, unsigned short *__restrict__ WB)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) WB[e] = f2bf(w[e]);
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

// ---- spectral branch launchers ------------------------------------------------------------------------

template <typename SRC>
static void spec_transform(const SpecGeom &g, const SRC *src, const SpecBuffers &sb, float2 *out, int round,
                           hipStream_t s)
{
    if (g.K < 1 || g.K > kSpecMaxK) throw hip_error("spectral modes K must be 1..12");
    const int Q = g.q(), nb = 2 * g.K + 1;
    const float s0 = (float)(1.0 / std::sqrt((double)g.nx * (double)g.ny));
    hipLaunchKernelGGL(spec_dft_y_kernel<SRC>, dim3(grid_of((int64_t)g.nx * Q, 256)), dim3(256), 0, s, g, src, sb.twy,
                       sb.T);
    XTR_LAUNCH_CHECK("spec_dft_y_kernel");
    hipLaunchKernelGGL(spec_dft_x_kernel, dim3(grid_of((int64_t)nb * Q, 256)), dim3(256), 0, s, g, sb.T, sb.twx, s0,
                       out, round);
    XTR_LAUNCH_CHECK("spec_dft_x_kernel");
}

void spec_forward(const SpecGeom &g, const NetBuffers &b, const SpecBuffers &sb, float phi, hipStream_t s)
{
    if (g.q() % 8 != 0) throw hip_error("the spectral branch needs Nz divisible by 8");
    const int Q = g.q(), nb = 2 * g.K + 1;
    spec_transform(g, b.x0, sb, sb.U, 1, s);
    const int64_t waves = (int64_t)g.nm() * (2 * Q / 16);
    hipLaunchKernelGGL(spec_mix_kernel, dim3(grid_of(waves, 4)), dim3(256), 0, s, g, sb.WB, phi, sb.U, sb.Y);
    XTR_LAUNCH_CHECK("spec_mix_kernel");
    hipLaunchKernelGGL(spec_idft_x_kernel, dim3(grid_of((int64_t)g.nx * nb * Q, 256)), dim3(256), 0, s, g, sb.Y, sb.twx,
                       sb.T);
    XTR_LAUNCH_CHECK("spec_idft_x_kernel");
    const float s0 = (float)(1.0 / std::sqrt((double)g.nx * (double)g.ny));
    hipLaunchKernelGGL(spec_idft_y_kernel, dim3(grid_of((int64_t)g.nx * g.ny * Q, 256)), dim3(256), 0, s, g, sb.T,
                       sb.twy, s0, b.y);
    XTR_LAUNCH_CHECK("spec_idft_y_kernel");
}

void spec_backward(const SpecGeom &g, const NetBuffers &b, const SpecBuffers &sb, float phi, float *grad, float scale,
                   int accumulate, hipStream_t s)
{
    spec_transform(g, b.dy, sb, sb.gY, 0, s);
    const int64_t per = (int64_t)g.nm() * g.q() * g.q();
    hipLaunchKernelGGL(spec_wgrad_kernel, dim3(grid_of(per, 256)), dim3(256), 0, s, g, sb.gY, sb.U, phi, scale, grad,
                       accumulate);
    XTR_LAUNCH_CHECK("spec_wgrad_kernel");
}

void spec_weights(const SpecGeom &g, const float *w, unsigned short *WB, hipStream_t s)
{
    const int64_t n = 2 * g.weights();
    hipLaunchKernelGGL(spec_weights_kernel, dim3(grid_of(n, 256)), dim3(256), 0, s, n, w, WB);
    XTR_LAUNCH_CHECK("spec_weights_kernel");
}

}  // namespace xtr
