// Observation-level kernels (FP64): the anomalous current jQ from the solved chi, and the
// anomaly-to-observation convolution of ToOCalculator.CalculateAnomalyField
// (Extreme.Cartesian/Forward/ToObs/ToOCalculator.cs:21-125) on the device.
#include "xtr_common.h"
#include "xtr_kernels.h"

namespace xtr {

namespace {

// jQ = E_s (zeta - zeta0) + j_s with E_s = (2 sqrt(Re zeta0) chi - j_s) / (zeta + conj zeta0)
// (ForwardSolver.CalculateEScatteredFrom / CalculateJqFrom, ForwardSolver.cs:319-436), written as
// jQ = R (2 f chi - j_s) + j_s with the A-type R = (zeta - zeta0)/(zeta + conj zeta0), f = sqrt(Re zeta0).
__global__ void jq_kernel(const double2 *chi, const double2 *js, const double2 *R, const double *f, int nz,
                          int64_t n, double2 *jq)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % nz);
        const int64_t cell = e / (3LL * nz);
        const double2 c = chi[e], s = js[e], r = R[cell * nz + k];
        const double2 t = make_double2(2 * f[k] * c.x - s.x, 2 * f[k] * c.y - s.y);
        jq[e] = cadd<double>(cmul<double>(r, t), s);
    }
}

// [Nx][Ny][C] -> zero-padded [2Nx][2Ny][C] (ToOCalculator.PrepareForForwardFft :143-157)
__global__ void pad_kernel(const double2 *src, double2 *dst, int nx, int ny, int C)
{
    const int64_t total = 4LL * nx * ny * C;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = e / C;
        const int q = (int)(e % C);
        const int i = (int)(row / (2 * ny)), j = (int)(row % (2 * ny));
        dst[e] = (i < nx && j < ny) ? src[((int64_t)i * ny + j) * C + q] : make_double2(0, 0);
    }
}

// One 64-lane wave per wavenumber, lane = transmitter layer: the nine (electric) or eight
// (magnetic) Zgemv of ApplyGreenTensorAlongZForE/H (:47-125), reduced over the layers.
// G: [L][ncomp][nz] (electric xx yy zz xy xz yz zx zy; magnetic xx xy yx xz yz zx zy); u: [L][3nz].
__global__ __launch_bounds__(256) void atoo_contract_kernel(const double2 *__restrict__ G, const double2 *__restrict__ u,
                                                            int64_t L, int nz, int electric, double2 *out)
{
    const int lane = threadIdx.x & 63;
    const int64_t l = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (l >= L) return;
    const int ncomp = electric ? 8 : 7;
    double2 ox = make_double2(0, 0), oy = ox, oz = ox;
    for (int k = lane; k < nz; k += 64) {
        const double2 *g = G + l * ncomp * nz + k;
        const double2 sx = u[l * 3 * nz + k], sy = u[l * 3 * nz + nz + k], sz = u[l * 3 * nz + 2 * nz + k];
        if (electric) {
            const double2 xx = g[0], yy = g[nz], zz = g[2 * nz], xy = g[3 * nz], xz = g[4 * nz], yz = g[5 * nz],
                          zx = g[6 * nz], zy = g[7 * nz];
            ox = cfma<double>(cfma<double>(cfma<double>(ox, xx, sx), xy, sy), xz, sz);
            oy = cfma<double>(cfma<double>(cfma<double>(oy, xy, sx), yy, sy), yz, sz);
            oz = cfma<double>(cfma<double>(cfma<double>(oz, zx, sx), zy, sy), zz, sz);
        } else {
            const double2 xx = g[0], xy = g[nz], yx = g[2 * nz], xz = g[3 * nz], yz = g[4 * nz], zx = g[5 * nz],
                          zy = g[6 * nz];
            ox = cfms<double>(cfms<double>(cfms<double>(ox, xx, sx), xy, sy), xz, sz);   // -(xx sx + xy sy + xz sz)
            oy = cfma<double>(cfma<double>(cfma<double>(oy, yx, sx), xx, sy), yz, sz);
            oz = cfma<double>(cfms<double>(oz, zx, sx), zy, sy);                       // -zx sx + zy sy
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        ox.x += __shfl_xor(ox.x, off, 64);
        ox.y += __shfl_xor(ox.y, off, 64);
        oy.x += __shfl_xor(oy.x, off, 64);
        oy.y += __shfl_xor(oy.y, off, 64);
        oz.x += __shfl_xor(oz.x, off, 64);
        oz.y += __shfl_xor(oz.y, off, 64);
    }
    if (lane == 0) {
        out[l * 3 + 0] = ox;
        out[l * 3 + 1] = oy;
        out[l * 3 + 2] = oz;
    }
}

// crop of the inverse transform with the 1/(4 Nx Ny) scale (ToOCalculator.cs:38-42, 159-172)
__global__ void crop_scale_kernel(const double2 *src, double2 *dst, int nx, int ny, double scale)
{
    const int64_t total = 3LL * nx * ny;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t cell = e / 3;
        const int c = (int)(e % 3);
        const int i = (int)(cell / ny), j = (int)(cell % ny);
        dst[e] = cscale<double>(src[((int64_t)i * 2 * ny + j) * 3 + c], scale);
    }
}

}  // namespace

void launch_jq(const double2 *chi, const double2 *js, const double2 *R, const double *f, const Geom &g, double2 *jq,
               hipStream_t s)
{
    hipLaunchKernelGGL(jq_kernel, dim3(2048), dim3(256), 0, s, chi, js, R, f, g.nz, g.n(), jq);
    XTR_LAUNCH_CHECK("jq_kernel");
}

void launch_level_field(const double2 *G, int electric, const double2 *jq, const Geom &g, double2 *spec,
                        double2 *small, const double2 *tw_x, const double2 *tw_y, double2 *field, hipStream_t s)
{
    const int C = 3 * g.nz;
    hipLaunchKernelGGL(pad_kernel, dim3(2048), dim3(256), 0, s, jq, spec, g.nx, g.ny, C);
    XTR_LAUNCH_CHECK("pad_kernel");
    launch_fft2_plain(spec, 2 * g.nx, 2 * g.ny, C, tw_x, tw_y, -1, s);
    const int64_t L = g.L();
    hipLaunchKernelGGL(atoo_contract_kernel, dim3((unsigned)ceil_div(L, 4)), dim3(256), 0, s, G, spec, L, g.nz,
                       electric, small);
    XTR_LAUNCH_CHECK("atoo_contract_kernel");
    launch_fft2_plain(small, 2 * g.nx, 2 * g.ny, 3, tw_x, tw_y, +1, s);
    hipLaunchKernelGGL(crop_scale_kernel, dim3(512), dim3(256), 0, s, small, field, g.nx, g.ny,
                       1.0 / ((double)g.nx * (double)g.ny * 4));
    XTR_LAUNCH_CHECK("crop_scale_kernel");
}

}  // namespace xtr
