// Krylov-basis kernels for the FGMRES driver: block dot V^H w, the classical Gram-Schmidt update
// with the new norm, norms, scaling and the solution update.
//
// Vectors are complex128 (the reference's arithmetic) or complex64 (the correction solves of the
// mixed-precision refinement): every kernel is a template on the stored element type and widens each
// element to FP64 on load, so products, dot sums and norms accumulate in FP64 either way; complex64
// results are rounded once, on store.
//
// Reductions are two-stage and deterministic: kRedBlocks workgroups write per-block
// partials (grid-stride loops, wavefront shuffles, then LDS), and one workgroup per output
// sums the partials in a fixed tree order, so repeated solves give identical iterates.
#include "xtr_common.h"
#include "xtr_kernels.h"

namespace xtr {

constexpr int kThreads = 256;
constexpr int kDotChunk = 8;   // basis columns per pass of block_dots

__device__ __forceinline__ double2 ld(const double2 &v) { return v; }
__device__ __forceinline__ double2 ld(const float2 &v) { return to_d(v); }
__device__ __forceinline__ void st(double2 &d, double2 v) { d = v; }
__device__ __forceinline__ void st(float2 &d, double2 v) { d = to_f(v); }

__device__ inline double warp_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// sum over the workgroup of v; valid in thread 0
__device__ inline double block_sum(double v, double *sh)
{
    v = warp_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
    __syncthreads();
    return r;
}

// partial[(k0 + k) * kRedBlocks + blk] = sum_i conj(V[k][i]) w[i] over this block's i
template <typename VT>
__global__ __launch_bounds__(kThreads) void block_dots_kernel(const VT *__restrict__ V, int cnt,
                                                              const VT *__restrict__ w, int64_t n,
                                                              double2 *partial)
{
    __shared__ double sh[kThreads / 64];
    const int k0 = blockIdx.y * kDotChunk;
    const int kc = min(kDotChunk, cnt - k0);
    double2 acc[kDotChunk];
#pragma unroll
    for (int k = 0; k < kDotChunk; ++k) acc[k] = make_double2(0, 0);
    // two elements per iteration, i and i + stride, all loads issued before the products (the
    // accumulation order per thread is the plain grid-stride order)
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    for (; i + stride < n; i += 2 * stride) {
        const VT w0 = w[i], w1 = w[i + stride];
        VT v0[kDotChunk], v1[kDotChunk];
#pragma unroll
        for (int k = 0; k < kDotChunk; ++k)
            if (k < kc) {
                v0[k] = V[(int64_t)(k0 + k) * n + i];
                v1[k] = V[(int64_t)(k0 + k) * n + i + stride];
            }
#pragma unroll
        for (int k = 0; k < kDotChunk; ++k)
            if (k < kc) {
                acc[k] = cfma_conj<double>(acc[k], ld(v0[k]), ld(w0));
                acc[k] = cfma_conj<double>(acc[k], ld(v1[k]), ld(w1));
            }
    }
    if (i < n) {
        const double2 wv = ld(w[i]);
#pragma unroll
        for (int k = 0; k < kDotChunk; ++k)
            if (k < kc) acc[k] = cfma_conj<double>(acc[k], ld(V[(int64_t)(k0 + k) * n + i]), wv);
    }
#pragma unroll
    for (int k = 0; k < kDotChunk; ++k) {
        if (k >= kc) break;
        const double re = block_sum(acc[k].x, sh);
        const double im = block_sum(acc[k].y, sh);
        if (threadIdx.x == 0) partial[(int64_t)(k0 + k) * kRedBlocks + blockIdx.x] = make_double2(re, im);
    }
}

template <typename VT>
static void block_dots(const VT *V, int cnt, const VT *w, int64_t n, double2 *partial, hipStream_t s)
{
    dim3 grid(kRedBlocks, (unsigned)ceil_div(cnt, kDotChunk));
    hipLaunchKernelGGL(block_dots_kernel<VT>, grid, dim3(kThreads), 0, s, V, cnt, w, n, partial);
    XTR_LAUNCH_CHECK("block_dots_kernel");
}
void launch_block_dots(const double2 *V, int cnt, const double2 *w, int64_t n, double2 *partial, hipStream_t s)
{
    block_dots(V, cnt, w, n, partial, s);
}
void launch_block_dots(const float2 *V, int cnt, const float2 *w, int64_t n, double2 *partial, hipStream_t s)
{
    block_dots(V, cnt, w, n, partial, s);
}

// out[k] = sum_b partial[k][b], one workgroup per k, fixed order; also into host_out[k] (pinned,
// device-mapped host memory) when given, which the host reads once the stream has passed this kernel
__global__ __launch_bounds__(kThreads) void reduce_c_kernel(const double2 *partial, double2 *out,
                                                            double2 *host_out)
{
    __shared__ double sh[kThreads / 64];
    const int k = blockIdx.x;
    double re = 0, im = 0;
    for (int b = threadIdx.x; b < kRedBlocks; b += kThreads) {
        re += partial[(int64_t)k * kRedBlocks + b].x;
        im += partial[(int64_t)k * kRedBlocks + b].y;
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) {
        out[k] = make_double2(re, im);
        if (host_out) host_out[k] = make_double2(re, im);
    }
}

void launch_reduce_c(const double2 *partial, int cnt, double2 *out, hipStream_t s, double2 *host_out)
{
    hipLaunchKernelGGL(reduce_c_kernel, dim3(cnt), dim3(kThreads), 0, s, partial, out, host_out);
    XTR_LAUNCH_CHECK("reduce_c_kernel");
}

// w[i] -= sum_k V[k][i] h[k]; partial[blk] = sum |w_new|^2 (of the stored, rounded w)
template <typename VT>
__global__ __launch_bounds__(kThreads) void cgs_update_kernel(const VT *__restrict__ V, int cnt,
                                                              const double2 *__restrict__ h, VT *w,
                                                              int64_t n, double *partial)
{
    __shared__ double sh[kThreads / 64];
    __shared__ double2 hs[32];
    if (threadIdx.x < cnt) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    double nn = 0;
    // two elements per iteration, i and i + stride, so that twice the loads are in flight; each
    // element's sum over k and the norm's sum over elements keep the plain grid-stride order
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x;
    for (; i + stride < n; i += 2 * stride) {
        double2 a0 = make_double2(0, 0), a1 = a0;
        int k = 0;
        for (; k + kDotChunk <= cnt; k += kDotChunk) {          // 2 x 8 loads issued together
            VT v0[kDotChunk], v1[kDotChunk];
#pragma unroll
            for (int j = 0; j < kDotChunk; ++j) {
                v0[j] = V[(int64_t)(k + j) * n + i];
                v1[j] = V[(int64_t)(k + j) * n + i + stride];
            }
#pragma unroll
            for (int j = 0; j < kDotChunk; ++j) {
                a0 = cfma<double>(a0, ld(v0[j]), hs[k + j]);
                a1 = cfma<double>(a1, ld(v1[j]), hs[k + j]);
            }
        }
        for (; k < cnt; ++k) {
            const VT v0 = V[(int64_t)k * n + i], v1 = V[(int64_t)k * n + i + stride];
            a0 = cfma<double>(a0, ld(v0), hs[k]);
            a1 = cfma<double>(a1, ld(v1), hs[k]);
        }
        VT t0, t1;
        st(t0, csub<double>(ld(w[i]), a0));
        st(t1, csub<double>(ld(w[i + stride]), a1));
        w[i] = t0;
        w[i + stride] = t1;
        const double2 n0 = ld(t0), n1 = ld(t1);
        nn = fma(n0.x, n0.x, nn);
        nn = fma(n0.y, n0.y, nn);
        nn = fma(n1.x, n1.x, nn);
        nn = fma(n1.y, n1.y, nn);
    }
    if (i < n) {
        double2 a = make_double2(0, 0);
        for (int k = 0; k < cnt; ++k) a = cfma<double>(a, ld(V[(int64_t)k * n + i]), hs[k]);
        VT t;
        st(t, csub<double>(ld(w[i]), a));
        w[i] = t;
        const double2 nw = ld(t);
        nn = fma(nw.x, nw.x, nn);
        nn = fma(nw.y, nw.y, nn);
    }
    nn = block_sum(nn, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = nn;
}

template <typename VT>
static void cgs_update(const VT *V, int cnt, const double2 *h, VT *w, int64_t n, double *partial, hipStream_t s)
{
    if (cnt > 32) throw hip_error("cgs_update: more than 32 basis vectors");
    hipLaunchKernelGGL(cgs_update_kernel<VT>, dim3(kRedBlocks), dim3(kThreads), 0, s, V, cnt, h, w, n, partial);
    XTR_LAUNCH_CHECK("cgs_update_kernel");
}
void launch_cgs_update(const double2 *V, int cnt, const double2 *h, double2 *w, int64_t n, double *partial,
                       hipStream_t s)
{
    cgs_update(V, cnt, h, w, n, partial, s);
}
void launch_cgs_update(const float2 *V, int cnt, const double2 *h, float2 *w, int64_t n, double *partial,
                       hipStream_t s)
{
    cgs_update(V, cnt, h, w, n, partial, s);
}

template <typename VT>
__global__ __launch_bounds__(kThreads) void norm2_kernel(const VT *__restrict__ w, int64_t n, double *partial)
{
    __shared__ double sh[kThreads / 64];
    double nn = 0;
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const double2 v = ld(w[i]);
        nn = fma(v.x, v.x, nn);
        nn = fma(v.y, v.y, nn);
    }
    nn = block_sum(nn, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = nn;
}

void launch_norm2(const double2 *w, int64_t n, double *partial, hipStream_t s)
{
    hipLaunchKernelGGL(norm2_kernel<double2>, dim3(kRedBlocks), dim3(kThreads), 0, s, w, n, partial);
    XTR_LAUNCH_CHECK("norm2_kernel");
}
void launch_norm2(const float2 *w, int64_t n, double *partial, hipStream_t s)
{
    hipLaunchKernelGGL(norm2_kernel<float2>, dim3(kRedBlocks), dim3(kThreads), 0, s, w, n, partial);
    XTR_LAUNCH_CHECK("norm2_kernel");
}

__global__ __launch_bounds__(kThreads) void reduce_r_kernel(const double *partial, double *out, double *host_out)
{
    __shared__ double sh[kThreads / 64];
    double v = 0;
    for (int b = threadIdx.x; b < kRedBlocks; b += kThreads) v += partial[b];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) {
        *out = v;
        if (host_out) *host_out = v;
    }
}

void launch_reduce_r(const double *partial, double *out, hipStream_t s, double *host_out)
{
    hipLaunchKernelGGL(reduce_r_kernel, dim3(1), dim3(kThreads), 0, s, partial, out, host_out);
    XTR_LAUNCH_CHECK("reduce_r_kernel");
}

template <typename VT>
__global__ void div_kernel(const VT *src, double d, VT *dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = ld(src[i]);
        st(dst[i], make_double2(v.x / d, v.y / d));
    }
}

void launch_div(const double2 *src, double d, double2 *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(div_kernel<double2>, dim3(2048), dim3(256), 0, s, src, d, dst, n);
    XTR_LAUNCH_CHECK("div_kernel");
}
void launch_div(const float2 *src, double d, float2 *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(div_kernel<float2>, dim3(2048), dim3(256), 0, s, src, d, dst, n);
    XTR_LAUNCH_CHECK("div_kernel");
}

template <typename VS, typename VD>
__global__ void mul_kernel(const VS *src, double a, VD *dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = ld(src[i]);
        st(dst[i], make_double2(a * v.x, a * v.y));
    }
}

void launch_mul(const double2 *src, double a, double2 *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL((mul_kernel<double2, double2>), dim3(2048), dim3(256), 0, s, src, a, dst, n);
    XTR_LAUNCH_CHECK("mul_kernel");
}
void launch_mul(const float2 *src, double a, float2 *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL((mul_kernel<float2, float2>), dim3(2048), dim3(256), 0, s, src, a, dst, n);
    XTR_LAUNCH_CHECK("mul_kernel");
}
void launch_mul(const double2 *src, double a, float2 *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL((mul_kernel<double2, float2>), dim3(2048), dim3(256), 0, s, src, a, dst, n);
    XTR_LAUNCH_CHECK("mul_kernel");
}

template <typename VT>
__global__ void sub_kernel(const VT *a, const VT *b, VT *dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        st(dst[i], csub<double>(ld(a[i]), ld(b[i])));
}

void launch_sub(const double2 *a, const double2 *b, double2 *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(sub_kernel<double2>, dim3(2048), dim3(256), 0, s, a, b, dst, n);
    XTR_LAUNCH_CHECK("sub_kernel");
}
void launch_sub(const float2 *a, const float2 *b, float2 *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(sub_kernel<float2>, dim3(2048), dim3(256), 0, s, a, b, dst, n);
    XTR_LAUNCH_CHECK("sub_kernel");
}

// x[i] += sum_k B[k][i] y[k]  (ConstructCurrentSolution: r0 = Z y; x = x + r0)
template <typename VB, typename VX>
__global__ __launch_bounds__(kThreads) void update_solution_kernel(const VB *__restrict__ B, int cnt, Coef32 y,
                                                                   VX *x, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        double2 a = make_double2(0, 0);
        for (int k = 0; k < cnt; ++k) a = cfma<double>(a, ld(B[(int64_t)k * n + i]), y.c[k]);
        st(x[i], cadd<double>(ld(x[i]), a));
    }
}

void launch_update_solution(const double2 *B, int cnt, const Coef32 &y, double2 *x, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL((update_solution_kernel<double2, double2>), dim3(2048), dim3(kThreads), 0, s, B, cnt, y, x, n);
    XTR_LAUNCH_CHECK("update_solution_kernel");
}
void launch_update_solution(const float2 *B, int cnt, const Coef32 &y, float2 *x, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL((update_solution_kernel<float2, float2>), dim3(2048), dim3(kThreads), 0, s, B, cnt, y, x, n);
    XTR_LAUNCH_CHECK("update_solution_kernel");
}
void launch_update_solution(const float2 *B, int cnt, const Coef32 &y, double2 *x, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL((update_solution_kernel<float2, double2>), dim3(2048), dim3(kThreads), 0, s, B, cnt, y, x, n);
    XTR_LAUNCH_CHECK("update_solution_kernel");
}

}  // namespace xtr
