// Operator-apply kernels for gfx950: LDS-staged pencil FFTs with the R/f scaling fused into
// the first forward pass and the b scaling plus the last step fused into the last inverse
// pass, and the per-wavenumber 3x3-block Green-tensor contraction.
//
// Spectrum layout (the reference's FFT buffer, Extreme.Cartesian/Fft/FftBuffer.cs:78-96,
// ConvolutionOperator.cs:250-272): buf[(i*2Ny + j)*3Nz + c*Nz + k], one buffer per
// right-hand side. 2-D transforms are unnormalised, sign -1 forward, +1 backward
// (Fft/FftW/FftWTransform.cs:20-27); the 1/(4 Nx Ny) lives in the backward factors b.
//
// A pencil kernel owns TC adjacent columns and all P points of each: it stages the
// TC x P tile in LDS (column index fastest, so global loads of a row of TC columns are
// contiguous), runs a Stockham radix-4/2/3/5 transform with every stage reading its
// butterfly inputs into registers before a barrier and writing after it, and writes
// the tile back. The zero padding of the forward passes and the cropping of the inverse
// passes are applied at load/store time, so padded rows are never read or written twice.
//
// The contraction kernels copy the Green tensor (read once per apply: 1.1 GB at 128 x 128 x 32,
// 17 GB at 256 x 256 x 64) HBM -> LDS with the non-temporal hint (nt) on global_load_lds_dwordx4,
// so the stream does not allocate in L2 and the Infinity Cache over the lines the kernel writes
// (the folded contraction at 128 x 128 x 32: 0.337 -> 0.300 ms).
#include <algorithm>
#include <map>
#include <mutex>
#include <utility>
#include <cstdlib>
#include <cstring>

#include "xtr_common.h"
#include "xtr_kernels.h"

namespace xtr {

constexpr int kFftThreads = 256;

// ---- compile-time radix plan: 4s first, then 2, then 3, then 5 ---------------------------------
constexpr int radix_of(int len)
{
    return (len % 4 == 0) ? 4 : (len % 2 == 0) ? 2 : (len % 3 == 0) ? 3 : 5;
}

template <typename T, int R, int DIR>
__host__ __device__ inline void dft_small(cplx<T> *a)
{
    if constexpr (R == 2) {
        cplx<T> t = a[0];
        a[0] = cadd<T>(t, a[1]);
        a[1] = csub<T>(t, a[1]);
    } else if constexpr (R == 4) {
        cplx<T> t0 = cadd<T>(a[0], a[2]), t1 = csub<T>(a[0], a[2]);
        cplx<T> t2 = cadd<T>(a[1], a[3]), d = csub<T>(a[1], a[3]);
        // t3 = d * (DIR i)
        cplx<T> t3 = DIR < 0 ? mk<T>(d.y, -d.x) : mk<T>(-d.y, d.x);
        a[0] = cadd<T>(t0, t2);
        a[1] = cadd<T>(t1, t3);
        a[2] = csub<T>(t0, t2);
        a[3] = csub<T>(t1, t3);
    } else if constexpr (R == 5) {
        // w = exp(DIR 2 pi i / 5): with s1 = a1 + a4, d1 = a1 - a4, s2 = a2 + a3, d2 = a2 - a3,
        // y1,4 = a0 + c1 s1 + c2 s2 +- DIR i (n1 d1 + n2 d2), y2,3 = a0 + c2 s1 + c1 s2 +- DIR i (n2 d1 - n1 d2),
        // c1 = cos(2 pi/5) = (sqrt5 - 1)/4, c2 = cos(4 pi/5) = -(sqrt5 + 1)/4, n1 = sin(2 pi/5), n2 = sin(4 pi/5)
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T n1 = (T)0.95105651629515357212, n2 = (T)0.58778525229247312917;
        const cplx<T> s1 = cadd<T>(a[1], a[4]), d1 = csub<T>(a[1], a[4]);
        const cplx<T> s2 = cadd<T>(a[2], a[3]), d2 = csub<T>(a[2], a[3]);
        const cplx<T> m1 = mk<T>(a[0].x + c1 * s1.x + c2 * s2.x, a[0].y + c1 * s1.y + c2 * s2.y);
        const cplx<T> m2 = mk<T>(a[0].x + c2 * s1.x + c1 * s2.x, a[0].y + c2 * s1.y + c1 * s2.y);
        const cplx<T> e1 = mk<T>(n1 * d1.x + n2 * d2.x, n1 * d1.y + n2 * d2.y);
        const cplx<T> e2 = mk<T>(n2 * d1.x - n1 * d2.x, n2 * d1.y - n1 * d2.y);
        // DIR i e
        const cplx<T> j1 = DIR < 0 ? mk<T>(e1.y, -e1.x) : mk<T>(-e1.y, e1.x);
        const cplx<T> j2 = DIR < 0 ? mk<T>(e2.y, -e2.x) : mk<T>(-e2.y, e2.x);
        a[0] = mk<T>(a[0].x + s1.x + s2.x, a[0].y + s1.y + s2.y);
        a[1] = cadd<T>(m1, j1);
        a[4] = csub<T>(m1, j1);
        a[2] = cadd<T>(m2, j2);
        a[3] = csub<T>(m2, j2);
    } else {  // R == 3, w = exp(DIR 2 pi i / 3)
        const T h = (T)0.86602540378443864676;  // sqrt(3)/2
        cplx<T> s = cadd<T>(a[1], a[2]), d = csub<T>(a[1], a[2]);
        cplx<T> m = mk<T>(a[0].x - (T)0.5 * s.x, a[0].y - (T)0.5 * s.y);
        // (DIR h) i d
        cplx<T> id = DIR < 0 ? mk<T>(h * d.y, -h * d.x) : mk<T>(-h * d.y, h * d.x);
        a[0] = cadd<T>(a[0], s);
        a[1] = cadd<T>(m, id);
        a[2] = csub<T>(m, id);
    }
}

// exp(DIR 2 pi i j / 16) applied to x, j in 0..15 known at compile time after unrolling. The
// cosines of multiples of pi/8 are 1, cos(pi/8), sqrt(1/2), sin(pi/8), 0 and their mirrors.
template <typename T, int DIR>
__host__ __device__ inline cplx<T> rot16(cplx<T> x, int j)
{
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
    const double cs[16] = {1, c1, h, s1, 0, -s1, -h, -c1, -1, -c1, -h, -s1, 0, s1, h, c1};
    j &= 15;
    if (j == 0) return x;
    if (j == 4) return DIR < 0 ? mk<T>(x.y, -x.x) : mk<T>(-x.y, x.x);       // DIR i
    if (j == 8) return mk<T>(-x.x, -x.y);
    if (j == 12) return DIR < 0 ? mk<T>(-x.y, x.x) : mk<T>(x.y, -x.x);      // -DIR i
    const T c = (T)cs[j], sn = (T)(DIR < 0 ? -cs[(j + 12) & 15] : cs[(j + 12) & 15]);   // sin = cos(. - pi/2)
    return mk<T>(x.x * c - x.y * sn, x.x * sn + x.y * c);
}

// In-register DFT of length R = 8 or 16 (sign DIR): n = (R/4) n1 + n2, k = k1 + 4 k2, a length-4
// DFT over n1 for each n2, the twiddles exp(DIR 2 pi i n2 k1 / R), then a length-R/4 DFT over n2.
template <typename T, int R, int DIR>
__host__ __device__ inline void dft_pow2(cplx<T> *a)
{
    static_assert(R == 8 || R == 16, "dft_pow2 covers 8 and 16");
    constexpr int R2 = R / 4;
    cplx<T> b[R];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) {
        cplx<T> t[4] = {a[n2], a[R2 + n2], a[2 * R2 + n2], a[3 * R2 + n2]};
        dft_small<T, 4, DIR>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) b[n2 * 4 + k1] = rot16<T, DIR>(t[k1], n2 * k1 * (16 / R));
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cplx<T> t[R2];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) t[n2] = b[n2 * 4 + k1];
        dft_small<T, R2, DIR>(t);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) a[k1 + 4 * k2] = t[k2];
    }
}

// One Stockham stage over the LDS tile s[p*TC + c] (NT threads) and the recursion to the next stage.
template <typename T, int P, int TC, int DIR, int LEN, int ST, int NT>
__device__ inline void fft_stage(cplx<T> *s, const cplx<T> *tw, int tid)
{
    constexpr int R = radix_of(LEN);
    constexpr int M = LEN / R;
    constexpr int NB = (P / R) * TC;
    constexpr int NPT = (NB + NT - 1) / NT;
    cplx<T> y[NPT][R];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const int b = tid + k * NT;
        if (NB % NT == 0 || b < NB) {
            const int cl = b % TC, bb = b / TC;
            const int p = bb / ST, q = bb % ST;
            cplx<T> a[R];
#pragma unroll
            for (int t = 0; t < R; ++t) a[t] = s[(q + ST * (p + t * M)) * TC + cl];
            dft_small<T, R, DIR>(a);
#pragma unroll
            for (int u = 1; u < R; ++u) a[u] = cmul<T>(a[u], tw[p * u * (P / LEN)]);
#pragma unroll
            for (int u = 0; u < R; ++u) y[k][u] = a[u];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const int b = tid + k * NT;
        if (NB % NT == 0 || b < NB) {
            const int cl = b % TC, bb = b / TC;
            const int p = bb / ST, q = bb % ST;
#pragma unroll
            for (int u = 0; u < R; ++u) s[(q + ST * (R * p + u)) * TC + cl] = y[k][u];
        }
    }
    __syncthreads();
    if constexpr (M > 1) fft_stage<T, P, TC, DIR, M, ST * R, NT>(s, tw, tid);
}

// ---- load/store policies -------------------------------------------------------------------

// Generic strided pencils of one array per right-hand side. Column c lives at
// (c / cpr) * row_stride + (c % cpr); point p at + p * pstride.
template <typename T>
struct IoPlain {
    cplx<T> *buf;
    int64_t rhs_stride;
    int64_t ncols, cpr, row_stride, pstride;
    int n_in, n_out;
    __device__ int64_t addr(int rhs, int64_t c, int p) const
    {
        return rhs * rhs_stride + (c / cpr) * row_stride + (c % cpr) + (int64_t)p * pstride;
    }
    __device__ cplx<T> load(int rhs, int64_t c, int p) const
    {
        if (c >= ncols || p >= n_in) return mk<T>(0, 0);
        return buf[addr(rhs, c, p)];
    }
    __device__ void store(int rhs, int64_t c, int p, cplx<T> v) const
    {
        if (c < ncols && p < n_out) buf[addr(rhs, c, p)] = v;
    }
};

__device__ __forceinline__ double2 widen(double2 v) { return v; }
__device__ __forceinline__ double2 widen(float2 v) { return to_d(v); }

// First forward pass along y, fused with ApplyOperatorR and PrepareForForwardFft
// (ConvolutionOperator.cs:233-272): reads x (complex128, or complex64 when V = float) and R, writes
// f * R * x into rows i < Nx of the padded buffer; columns c = i*3Nz + col; points j < Ny are data,
// j >= Ny are the padding. R * x is formed in FP64.
template <typename T, typename V>
struct IoFwdY {
    RhsPtrs v;                // x, and R per right-hand side
    const double *f;
    cplx<T> *buf;
    int64_t M;
    int nx, ny, nz;
    __device__ cplx<T> load(int rhs, int64_t c, int p) const
    {
        const int64_t C = 3LL * nz;
        if (c >= nx * C || p >= ny) return mk<T>(0, 0);
        const int64_t i = c / C, col = c % C;
        const int k = (int)(col % nz);
        const int64_t ij = i * ny + p;
        const cplx<V> *x = reinterpret_cast<const cplx<V> *>(v.x[rhs]);
        double2 rx = cmul<double>(v.R[rhs][ij * nz + k], widen(x[ij * C + col]));
        return from_d<T>(cscale<double>(rx, f[k]));
    }
    __device__ void store(int rhs, int64_t c, int p, cplx<T> val) const
    {
        const int64_t C = 3LL * nz;
        if (c >= nx * C) return;
        const int64_t i = c / C, col = c % C;
        buf[rhs * M + (i * 2 * ny + p) * C + col] = val;
    }
};

// Last inverse pass along y, fused with ExtractData and PerformLastStep
// (ConvolutionOperator.cs:187-231): keeps j < Ny of rows i < Nx, scales by b, and writes
// y = x - b*conv - R*x (A) or y = b*conv + R*x (Chi0), formed in FP64 and stored as complex128 (or
// complex64 when V = float).
template <typename T, typename V>
struct IoInvY {
    RhsPtrs v;                // x, y, and R per right-hand side
    const double *b;
    cplx<T> *buf;
    int64_t M;
    int nx, ny, nz, type;
    __device__ cplx<T> load(int rhs, int64_t c, int p) const
    {
        const int64_t C = 3LL * nz;
        if (c >= nx * C) return mk<T>(0, 0);
        const int64_t i = c / C, col = c % C;
        return buf[rhs * M + (i * 2 * ny + p) * C + col];
    }
    __device__ void store(int rhs, int64_t c, int p, cplx<T> val) const
    {
        const int64_t C = 3LL * nz;
        if (c >= nx * C || p >= ny) return;
        const int64_t i = c / C, col = c % C;
        const int k = (int)(col % nz);
        const int64_t ij = i * ny + p, t = ij * C + col;
        double2 conv;
        if constexpr (sizeof(T) == 8) conv = cscale<double>(val, b[k]);
        else conv = cscale<double>(to_d(val), b[k]);
        const double2 xv = widen(reinterpret_cast<const cplx<V> *>(v.x[rhs])[t]);
        const double2 rx = cmul<double>(v.R[rhs][ij * nz + k], xv);
        double2 out;
        if (type == 0) out = csub<double>(csub<double>(xv, conv), rx);
        else out = cadd<double>(conv, rx);
        reinterpret_cast<cplx<V> *>(v.y[rhs])[t] = from_d<V>(out);
    }
};

// Column tile of workgroup b among nb. Workgroups go to the 8 XCDs round-robin (b mod 8), and
// each XCD has its own L2. Giving XCD x a contiguous run of tiles puts neighbouring tiles, which
// share cache lines of narrow rows (and the R values of the three components of one cell), on one
// L2; in launch order they would sit on eight different ones. A bijection on [0, nb).
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb)
{
    const int64_t per = nb / 8, rem = nb % 8, x = b % 8, i = b / 8;
    return x * per + (x < rem ? x : rem) + i;
}

template <typename T, int P, int TC, int DIR, class IO, int NT>
__global__ __launch_bounds__(NT) void pencil_fft_kernel(IO io, const cplx<T> *__restrict__ twg)
{
    __shared__ cplx<T> s[P * TC];
    __shared__ cplx<T> tw[P];
    const int tid = threadIdx.x;
    for (int k = tid; k < P; k += NT) {
        cplx<T> w = twg[k];
        tw[k] = DIR < 0 ? w : cconj<T>(w);
    }
    const int64_t c0 = xcd_tile(blockIdx.x, gridDim.x) * TC;
    const int rhs = blockIdx.y;
    constexpr int NE = (P * TC + NT - 1) / NT;
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int e = tid + k * NT;
        if ((P * TC) % NT == 0 || e < P * TC) {
            const int p = e / TC, cl = e % TC;
            s[p * TC + cl] = io.load(rhs, c0 + cl, p);
        }
    }
    __syncthreads();
    fft_stage<T, P, TC, DIR, P, 1, NT>(s, tw, tid);
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int e = tid + k * NT;
        if ((P * TC) % NT == 0 || e < P * TC) {
            const int p = e / TC, cl = e % TC;
            io.store(rhs, c0 + cl, p, s[p * TC + cl]);
        }
    }
}

// ---- register-resident pencils for lengths 64, 128, 256, 512 -------------------------------
// The same Stockham stages with radix 16 or 8, so a length takes two or three stages: the first
// reads its butterfly inputs straight from global memory (through the load policy), the last
// writes its outputs straight to global memory, and only the exchanges between stages go through
// the LDS tile (one exchange for 64, 128 and 256; two for 512 = 8 x 8 x 8).
constexpr int reg_radix(int len)
{
    return (len == 512 || len == 64 || len % 16 != 0) ? 8 : 16;
}
constexpr bool reg_length(int p) { return p == 64 || p == 128 || p == 256 || p == 512; }
template <int P, int TC>
constexpr int reg_threads()
{
    const int n = TC * (P / reg_radix(P));
    return n < 64 ? 64 : n > 512 ? 512 : n;
}

template <typename T, int P, int TC, int DIR, int LEN, int ST, int NT, bool FIRST, class IO>
__device__ inline void reg_stage(const IO &io, int rhs, int64_t c0, cplx<T> *s, const cplx<T> *tw, int tid)
{
    constexpr int R = reg_radix(LEN);
    constexpr int M = LEN / R;
    constexpr int NB = (P / R) * TC;
    constexpr int NPT = (NB + NT - 1) / NT;
    constexpr bool LAST = M == 1;
    static_assert(LEN % R == 0 && !(FIRST && LAST), "a register pencil has at least two stages");
    cplx<T> y[NPT][R];
#pragma unroll
    for (int k = 0; k < NPT; ++k) {
        const int b = tid + k * NT;
        if (NB % NT == 0 || b < NB) {
            const int cl = b % TC, bb = b / TC;
            const int p = bb / ST, q = bb % ST;
            cplx<T> a[R];
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const int pt = q + ST * (p + t * M);
                if constexpr (FIRST) a[t] = io.load(rhs, c0 + cl, pt);
                else a[t] = s[pt * TC + cl];
            }
            dft_pow2<T, R, DIR>(a);
            if constexpr (LAST) {      // p == 0: no twiddles; point q + ST u of column cl
#pragma unroll
                for (int u = 0; u < R; ++u) io.store(rhs, c0 + cl, q + ST * u, a[u]);
            } else {
#pragma unroll
                for (int u = 1; u < R; ++u) a[u] = cmul<T>(a[u], tw[p * u * (P / LEN)]);
#pragma unroll
                for (int u = 0; u < R; ++u) y[k][u] = a[u];
            }
        }
    }
    if constexpr (!LAST) {
        if constexpr (!FIRST) __syncthreads();     // every read of this stage's inputs is done
#pragma unroll
        for (int k = 0; k < NPT; ++k) {
            const int b = tid + k * NT;
            if (NB % NT == 0 || b < NB) {
                const int cl = b % TC, bb = b / TC;
                const int p = bb / ST, q = bb % ST;
#pragma unroll
                for (int u = 0; u < R; ++u) s[(q + ST * (R * p + u)) * TC + cl] = y[k][u];
            }
        }
        __syncthreads();
        reg_stage<T, P, TC, DIR, M, ST * R, NT, false>(io, rhs, c0, s, tw, tid);
    }
}

template <typename T, int P, int TC, int DIR, class IO, int NT>
__global__ __launch_bounds__(NT) void pencil_fft_reg_kernel(IO io, const cplx<T> *__restrict__ twg)
{
    static_assert(reg_length(P), "register pencils cover 64, 128, 256 and 512");
    __shared__ cplx<T> s[P * TC];
    __shared__ cplx<T> tw[P];
    const int tid = threadIdx.x;
    for (int k = tid; k < P; k += NT) {
        cplx<T> w = twg[k];
        tw[k] = DIR < 0 ? w : cconj<T>(w);
    }
    __syncthreads();
    reg_stage<T, P, TC, DIR, P, 1, NT, true>(io, blockIdx.y, xcd_tile(blockIdx.x, gridDim.x) * TC, s, tw, tid);
}

// Column tiles per pencil length. The default tile has 16 columns while P x 16 stays within 32 KB,
// else 8. For lengths from 64 there are two more: the narrow tile (half the columns, more workgroups
// per CU) and the wide one (twice the columns, longer contiguous runs per row, within 128 KB); from
// 128, the narrow tile also runs with 512 threads (twice the waves per CU for the same LDS). Which
// one a pass uses is measured on the device (the operator's tile tuning); the choice does not change
// the arithmetic, only how the columns are grouped and how many threads share a tile.
template <typename T, int P>
constexpr int tile_cols()
{
    return (P * 16 * (int)sizeof(cplx<T>) <= 32768) ? 16 : 8;
}
template <typename T, int P>
constexpr int tile_cols_of(int tile)
{
    constexpr int d = tile_cols<T, P>();
    if (P < 64) return d;
    if (tile == FFT_TILE_WIDE) return P * 2 * d * (int)sizeof(cplx<T>) <= 131072 ? 2 * d : d;
    if (tile == FFT_TILE_NARROW || tile == FFT_TILE_NARROW_T512) return d / 2 >= 4 ? d / 2 : d;
    return d;
}
template <int P>
constexpr int tile_threads_of(int tile)
{
    return tile == FFT_TILE_NARROW_T512 && P >= 128 ? 512 : kFftThreads;
}
// register pencils: the default column count, or the alternative (twice that within 64 KB of LDS,
// else half)
template <typename T, int P>
constexpr int reg_cols_of(int tile)
{
    constexpr int d = tile_cols<T, P>();
    if (tile != FFT_TILE_REG_ALT) return d;
    return P * 2 * d * (int)sizeof(cplx<T>) <= 65536 ? 2 * d : d / 2;
}

template <typename T, int DIR, class IO>
void launch_pencil(int P, const IO &io, int64_t ncols, int nrhs, const void *tw, hipStream_t s, int tile)
{
    auto go = [&](auto pc) {
        constexpr int PP = decltype(pc)::value;
        constexpr int TD = tile_cols_of<T, PP>(FFT_TILE_DEFAULT), TW = tile_cols_of<T, PP>(FFT_TILE_WIDE),
                      TN = tile_cols_of<T, PP>(FFT_TILE_NARROW);
        constexpr int NW = tile_threads_of<PP>(FFT_TILE_NARROW_T512);
        auto run = [&](auto tc, auto nt) {
            constexpr int TC = decltype(tc)::value, NT = decltype(nt)::value;
            dim3 grid((unsigned)ceil_div(ncols, TC), nrhs);
            hipLaunchKernelGGL((pencil_fft_kernel<T, PP, TC, DIR, IO, NT>), grid, dim3(NT), 0, s, io,
                               (const cplx<T> *)tw);
        };
        using T256 = std::integral_constant<int, kFftThreads>;
        if constexpr (reg_length(PP)) {
            if (tile == FFT_TILE_REG || tile == FFT_TILE_REG_ALT) {
                auto reg = [&](auto tc) {
                    constexpr int TC = decltype(tc)::value, NT = reg_threads<PP, TC>();
                    dim3 grid((unsigned)ceil_div(ncols, TC), nrhs);
                    hipLaunchKernelGGL((pencil_fft_reg_kernel<T, PP, TC, DIR, IO, NT>), grid, dim3(NT), 0, s, io,
                                       (const cplx<T> *)tw);
                };
                if (tile == FFT_TILE_REG) reg(std::integral_constant<int, reg_cols_of<T, PP>(FFT_TILE_REG)>{});
                else reg(std::integral_constant<int, reg_cols_of<T, PP>(FFT_TILE_REG_ALT)>{});
                return;
            }
        }
        if (tile == FFT_TILE_WIDE && TW != TD) run(std::integral_constant<int, TW>{}, T256{});
        else if (tile == FFT_TILE_NARROW && TN != TD) run(std::integral_constant<int, TN>{}, T256{});
        else if (tile == FFT_TILE_NARROW_T512 && TN != TD && NW != kFftThreads)
            run(std::integral_constant<int, TN>{}, std::integral_constant<int, NW>{});
        else run(std::integral_constant<int, TD>{}, T256{});
    };
    switch (P) {
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 6: go(std::integral_constant<int, 6>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 10: go(std::integral_constant<int, 10>{}); break;
    case 12: go(std::integral_constant<int, 12>{}); break;
    case 16: go(std::integral_constant<int, 16>{}); break;
    case 20: go(std::integral_constant<int, 20>{}); break;
    case 24: go(std::integral_constant<int, 24>{}); break;
    case 30: go(std::integral_constant<int, 30>{}); break;
    case 32: go(std::integral_constant<int, 32>{}); break;
    case 40: go(std::integral_constant<int, 40>{}); break;
    case 48: go(std::integral_constant<int, 48>{}); break;
    case 50: go(std::integral_constant<int, 50>{}); break;
    case 60: go(std::integral_constant<int, 60>{}); break;
    case 64: go(std::integral_constant<int, 64>{}); break;
    case 80: go(std::integral_constant<int, 80>{}); break;
    case 96: go(std::integral_constant<int, 96>{}); break;
    case 100: go(std::integral_constant<int, 100>{}); break;
    case 120: go(std::integral_constant<int, 120>{}); break;
    case 128: go(std::integral_constant<int, 128>{}); break;
    case 150: go(std::integral_constant<int, 150>{}); break;
    case 160: go(std::integral_constant<int, 160>{}); break;
    case 192: go(std::integral_constant<int, 192>{}); break;
    case 200: go(std::integral_constant<int, 200>{}); break;
    case 240: go(std::integral_constant<int, 240>{}); break;
    case 256: go(std::integral_constant<int, 256>{}); break;
    case 300: go(std::integral_constant<int, 300>{}); break;
    case 320: go(std::integral_constant<int, 320>{}); break;
    case 384: go(std::integral_constant<int, 384>{}); break;
    case 400: go(std::integral_constant<int, 400>{}); break;
    case 480: go(std::integral_constant<int, 480>{}); break;
    case 512: go(std::integral_constant<int, 512>{}); break;
    default: throw hip_error("unsupported FFT length " + std::to_string(P));
    }
    XTR_LAUNCH_CHECK("pencil_fft_kernel");
}

// even lengths 2^a m, m in {1, 3, 5, 15, 25, 75}, up to 512 (Nx, Ny <= 256)
int fft_tile_options(int p, int precision, int *tiles)
{
    int n = 0;
    tiles[n++] = FFT_TILE_DEFAULT;
    if (p < 64) return n;
    const int e = precision ? 8 : 16;
    const int d = (p * 16 * e <= 32768) ? 16 : 8;
    tiles[n++] = FFT_TILE_NARROW;                    // d / 2 >= 4 always
    if (p * 2 * d * e <= 131072) tiles[n++] = FFT_TILE_WIDE;
    if (p >= 128) tiles[n++] = FFT_TILE_NARROW_T512;
    if (p == 64 || p == 128 || p == 256 || p == 512) {
        tiles[n++] = FFT_TILE_REG;
        tiles[n++] = FFT_TILE_REG_ALT;
    }
    return n;
}

bool fft_length_supported(int p)
{
    switch (p) {
    case 4:
    case 6:
    case 8:
    case 10:
    case 12:
    case 16:
    case 20:
    case 24:
    case 30:
    case 32:
    case 40:
    case 48:
    case 50:
    case 60:
    case 64:
    case 80:
    case 96:
    case 100:
    case 120:
    case 128:
    case 150:
    case 160:
    case 192:
    case 200:
    case 240:
    case 256:
    case 300:
    case 320:
    case 384:
    case 400:
    case 480:
    case 512:
        return true;
    default:
        return false;
    }
}

// ---- stage launchers -------------------------------------------------------------------------

// right-hand sides without a model R of their own use the handle's R
static RhsPtrs with_r(const ApplyArgs &a)
{
    RhsPtrs v = a.v;
    for (int i = 0; i < a.nrhs; ++i)
        if (!v.R[i]) v.R[i] = a.R;
    return v;
}

template <typename T, typename V>
static void fwd_y(const ApplyArgs &a, hipStream_t s)
{
    IoFwdY<T, V> io{with_r(a), a.f, (cplx<T> *)a.buf, a.g.M(), a.g.nx, a.g.ny, a.g.nz};
    launch_pencil<T, -1>(2 * a.g.ny, io, (int64_t)a.g.nx * a.g.C(), a.nrhs, a.tw_y, s, a.fft_tile[0]);
}

template <typename T>
static void fwd_x(const ApplyArgs &a, hipStream_t s)
{
    const int64_t ncols = 2LL * a.g.ny * a.g.C();
    IoPlain<T> io{(cplx<T> *)a.buf, a.g.M(), ncols, ncols, 0, ncols, a.g.nx, 2 * a.g.nx};
    launch_pencil<T, -1>(2 * a.g.nx, io, ncols, a.nrhs, a.tw_x, s, a.fft_tile[1]);
}

template <typename T>
static void inv_x(const ApplyArgs &a, hipStream_t s)
{
    const int64_t ncols = 2LL * a.g.ny * a.g.C();
    IoPlain<T> io{(cplx<T> *)a.buf, a.g.M(), ncols, ncols, 0, ncols, 2 * a.g.nx, a.g.nx};
    launch_pencil<T, +1>(2 * a.g.nx, io, ncols, a.nrhs, a.tw_x, s, a.fft_tile[2]);
}

template <typename T, typename V>
static void inv_y(const ApplyArgs &a, hipStream_t s)
{
    IoInvY<T, V> io{with_r(a), a.b, (cplx<T> *)a.buf, a.g.M(), a.g.nx, a.g.ny, a.g.nz, a.type};
    launch_pencil<T, +1>(2 * a.g.ny, io, (int64_t)a.g.nx * a.g.C(), a.nrhs, a.tw_y, s, a.fft_tile[3]);
}

// complex64 vectors go with the FP32 apply only
static void check_vec32(const ApplyArgs &a)
{
    if (a.vec32 && !a.precision) throw hip_error("complex64 vectors need the FP32 apply");
}

void launch_fwd_y(const ApplyArgs &a, hipStream_t s)
{
    check_vec32(a);
    if (a.vec32) fwd_y<float, float>(a, s);
    else if (a.precision) fwd_y<float, double>(a, s);
    else fwd_y<double, double>(a, s);
}
void launch_fwd_x(const ApplyArgs &a, hipStream_t s) { a.precision ? fwd_x<float>(a, s) : fwd_x<double>(a, s); }
void launch_inv_x(const ApplyArgs &a, hipStream_t s) { a.precision ? inv_x<float>(a, s) : inv_x<double>(a, s); }
void launch_inv_y(const ApplyArgs &a, hipStream_t s)
{
    check_vec32(a);
    if (a.vec32) inv_y<float, float>(a, s);
    else if (a.precision) inv_y<float, double>(a, s);
    else inv_y<double, double>(a, s);
}

void launch_fft2_plain(double2 *data, int n1, int n2, int64_t ncols, const double2 *tw1, const double2 *tw2,
                       int sign, hipStream_t s)
{
    // along n2: columns (i1, c), pencil stride ncols
    IoPlain<double> y{data, 0, (int64_t)n1 * ncols, ncols, (int64_t)n2 * ncols, ncols, n2, n2};
    // along n1: columns (i2, c) flattened, pencil stride n2*ncols
    IoPlain<double> x{data, 0, (int64_t)n2 * ncols, (int64_t)n2 * ncols, 0, (int64_t)n2 * ncols, n1, n1};
    // the register-resident pencils where the length has them (at 512 points 0.5 ms against
    // 0.63-0.95 ms for the LDS kernel's tiles, per 1.6 GB pass)
    const int t2 = reg_length(n2) ? FFT_TILE_REG : FFT_TILE_DEFAULT, t1 = reg_length(n1) ? FFT_TILE_REG : FFT_TILE_DEFAULT;
    if (sign < 0) {
        launch_pencil<double, -1>(n2, y, (int64_t)n1 * ncols, 1, tw2, s, t2);
        launch_pencil<double, -1>(n1, x, (int64_t)n2 * ncols, 1, tw1, s, t1);
    } else {
        launch_pencil<double, +1>(n2, y, (int64_t)n1 * ncols, 1, tw2, s, t2);
        launch_pencil<double, +1>(n1, x, (int64_t)n2 * ncols, 1, tw1, s, t1);
    }
}

// ---- contraction -----------------------------------------------------------------------------
// ApplyGreenTensorAlongZ (ConvolutionOperator.cs:274-327): for every wavenumber l
//   y_x = XX u_x + XY u_y + XZ u_z
//   y_y = XY u_x + YY u_y + YZ u_z
//   y_z = -XZ^T u_x - YZ^T u_y + ZZ u_z
// XX, XY, YY, ZZ complex-symmetric (stored packed), XZ, YZ full row-major [rc][tr].
// A workgroup owns W wavenumbers and all right-hand sides; thread = (w, rhs, component, row).
// Each of the six blocks is staged in LDS as a padded full nz x (nz+1) matrix (the packed
// blocks are expanded on the way in), so both row and column reads are conflict-light.
// The spectrum is updated in place: every workgroup reads its u into LDS before writing y.
template <typename T>
__global__ __launch_bounds__(1024) void contract_kernel(const cplx<T> *__restrict__ G, cplx<T> *buf, int64_t L,
                                                         int nz, int nrhs, int W, int64_t M,
                                                         const uint16_t *__restrict__ packed_rc, int fold_nx,
                                                         int fold_ny)
{
    // fold_nx > 0: G is folded (block b(l), mirrored components negated on u and y)
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *blk = reinterpret_cast<cplx<T> *>(smem_raw);   // [W][nz][nz+1]
    const int ld = nz + 1;
    cplx<T> *u = blk + (int64_t)W * nz * ld;                  // [W][nrhs][3nz]
    const int C = 3 * nz;
    const int S = nz * (nz + 1) / 2;
    const int64_t gblk = 4LL * S + 2LL * nz * nz;
    const int tid = threadIdx.x, nth = blockDim.x;
    const int64_t l0 = (int64_t)blockIdx.x * W;
    const int per_w = nrhs * C;
    const int swz = asym_swizzle_mask(nz);

    for (int e = tid; e < W * per_w; e += nth) {
        const int w = e / per_w, rest = e % per_w;
        const int rhs = rest / C, q = rest % C;
        const int64_t l = l0 + w;
        cplx<T> v = (l < L) ? buf[rhs * M + l * C + q] : mk<T>(0, 0);
        if (fold_nx > 0 && l < L && q < 2 * nz) {
            const FoldPos p = fold_pos(l, fold_nx, fold_ny);
            if (q < nz ? p.mx : p.my) v = mk<T>(-v.x, -v.y);
        }
        u[e] = v;
    }

    const bool active = tid < W * per_w;
    const int w = active ? tid / per_w : 0;
    const int rest = active ? tid % per_w : 0;
    const int rhs = rest / C, q = rest % C;
    const int a = q / nz, r = q % nz;
    const cplx<T> *uw = u + (w * nrhs + rhs) * C;
    const cplx<T> *bw = blk + (int64_t)w * nz * ld;
    cplx<T> acc = mk<T>(0, 0);

    // block order in memory: xx, xy, yy, zz (packed S), xz, yz (nz*nz)
    for (int bi = 0; bi < 6; ++bi) {
        const bool sym = bi < 4;
        const int64_t off = sym ? (int64_t)bi * S : 4LL * S + (int64_t)(bi - 4) * nz * nz;
        const int sz = sym ? S : nz * nz;
        __syncthreads();
        for (int e = tid; e < W * sz; e += nth) {
            const int ww = e / sz, o = e % sz;
            const int64_t l = l0 + ww;
            const int64_t gb = fold_nx > 0 ? fold_pos(l, fold_nx, fold_ny).b : l;
            const cplx<T> g = (l < L) ? G[gb * gblk + off + o] : mk<T>(0, 0);
            cplx<T> *bb = blk + (int64_t)ww * nz * ld;
            if (sym) {
                const int rc = packed_rc[2 * o], tr = packed_rc[2 * o + 1];
                bb[rc * ld + tr] = g;
                bb[tr * ld + rc] = g;
            } else {
                const int rc = o / nz;
                bb[rc * ld + ((o % nz) ^ (rc & swz))] = g;
            }
        }
        __syncthreads();
        if (!active) continue;
        // row product acc += B[r][:] . u_col  (col component cc) ; or transposed
        int cc = -1;       // component of u for the row product
        int ct = -1;       // component of u for the transposed product (subtracted)
        switch (bi) {
        case 0: cc = (a == 0) ? 0 : -1; break;                           // XX
        case 1: cc = (a == 0) ? 1 : (a == 1) ? 0 : -1; break;            // XY
        case 2: cc = (a == 1) ? 1 : -1; break;                           // YY
        case 3: cc = (a == 2) ? 2 : -1; break;                           // ZZ
        case 4: cc = (a == 0) ? 2 : -1; ct = (a == 2) ? 0 : -1; break;   // XZ
        case 5: cc = (a == 1) ? 2 : -1; ct = (a == 2) ? 1 : -1; break;   // YZ
        }
        if (cc >= 0) {
            const cplx<T> *uc = uw + cc * nz;
            const cplx<T> *row = bw + r * ld;
            cplx<T> s0 = mk<T>(0, 0);
            for (int c = 0; c < nz; ++c) s0 = cfma<T>(s0, row[c], uc[c]);
            acc = cadd<T>(acc, s0);
        }
        if (ct >= 0) {
            const cplx<T> *uc = uw + ct * nz;
            cplx<T> s0 = mk<T>(0, 0);
            for (int m = 0; m < nz; ++m) s0 = cfma<T>(s0, bw[m * ld + r], uc[m]);
            acc = csub<T>(acc, s0);
        }
    }
    if (active) {
        const int64_t l = l0 + w;
        if (l < L) {
            if (fold_nx > 0 && q < 2 * nz) {
                const FoldPos p = fold_pos(l, fold_nx, fold_ny);
                if (q < nz ? p.mx : p.my) acc = mk<T>(-acc.x, -acc.y);
            }
            buf[rhs * M + l * C + q] = acc;
        }
    }
}

// ---- streaming contraction (nz = 32) ---------------------------------------------------------------
// Persistent: one 256-thread workgroup per CU walks wavenumbers l = blockIdx.x, += gridDim.x.
// Each wavenumber's whole Green block (4 packed + 2 swizzled asymmetric blocks, 4160 elements) and
// its spectrum columns are copied straight from HBM into LDS by global_load_lds_dwordx4 (64 lanes x
// 16 B = 1 KB per wave-load), into a ring of NBUF buffers: the copies of the next NBUF-1 wavenumbers
// are in flight while one is contracted, so the HBM stream does not wait for the arithmetic
// (complex128: 68 KB per wavenumber, 2 buffers; complex64: 34 KB, 4 buffers). The copies are issued
// from inline asm so that the only waits on them are the explicit vmcnt waits below: vector-memory
// operations of one wave complete in order, so with k wave-loads issued after those of wavenumber l,
// vmcnt(k) means l has landed.
// Contraction: wave w owns columns [8w, 8w+8); lane (s, r) takes row r and columns 8w + 4s + j
// (j < 4) of all six blocks, for all three components and both right-hand sides, with its u slice
// in registers (36 FMAs per 8 LDS reads). Partial rows are summed across s by a lane shuffle and
// across the four waves through LDS.
namespace stream {
constexpr int NZ = 32, C = 3 * NZ, S = NZ * (NZ + 1) / 2, A = NZ * NZ, GB = 4 * S + 2 * A;
constexpr int MASK = 15;                    // asym_swizzle_mask(32)
template <typename T> struct Cfg {
    static constexpr int E = (int)sizeof(cplx<T>);            // bytes per element
    static constexpr int EPL = 16 / E;                        // elements per lane per wave-load
    static constexpr int EPC = 64 * EPL;                      // elements per wave-load
    static constexpr int RAW = (GB + 4 * C + EPC - 1) / EPC;  // wave-loads for Green + 4 rhs columns
    static constexpr int NCHUNK = (RAW + 3) / 4 * 4;          // padded to the 4 waves
    static constexpr int PER_WAVE = NCHUNK / 4;
    static constexpr int BUF = NCHUNK * EPC;                  // elements per LDS buffer
    static constexpr int NBUF = E == 16 ? 2 : 4;
};
static_assert(GB % 2 == 0 && C % 2 == 0, "element pairs never straddle the Green/spectrum boundary");
static_assert(Cfg<double>::PER_WAVE * (Cfg<double>::NBUF - 1) <= 63, "vmcnt range");
static_assert(Cfg<float>::PER_WAVE * (Cfg<float>::NBUF - 1) <= 63, "vmcnt range");
static_assert(Cfg<double>::NBUF * Cfg<double>::BUF * 16 + 4 * 2 * C * 16 <= 160 * 1024, "LDS");
static_assert(Cfg<float>::NBUF * Cfg<float>::BUF * 8 + 4 * 2 * C * 8 <= 160 * 1024, "LDS");
constexpr int kMaxStreamRhs = 4;
}  // namespace stream

template <int N>
__device__ __forceinline__ void wait_vm()
{
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// LDS writes of this wave (partial sums) must be complete before other waves pass the barrier and
// read them: s_barrier alone does not wait for them.
#define XTR_BARRIER() asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory")

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
// This wave's PER_WAVE wave-loads of wavenumber l (Green block and nrhs spectrum columns) into the
// LDS buffer at byte address dst.
template <typename T>
__device__ __forceinline__ void stream_issue(const cplx<T> *__restrict__ G, const cplx<T> *buf, int64_t M, int64_t l,
                                             int nrhs, unsigned dst, int wave, int lane)
{
    using namespace stream;
    using K = Cfg<T>;
#pragma unroll
    for (int t = 0; t < K::PER_WAVE; ++t) {
        const int k = wave + 4 * t;
        const int e = k * K::EPC + lane * K::EPL;                  // first element this lane copies
        const cplx<T> *src;
        if (e < GB) {
            src = G + l * GB + e;
        } else {
            int q = e - GB;                                         // spectrum column of this wavenumber
            if (q > nrhs * C - K::EPL) q = nrhs * C - K::EPL;       // padding lanes re-read valid data
            const int rhs = q / C;
            src = buf + rhs * M + l * C + (q - rhs * C);
        }
        const unsigned m0 = __builtin_amdgcn_readfirstlane(dst + (unsigned)k * 1024u);
        asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off nt" ::"s"(m0), "v"(src) : "memory", "m0");
    }
}
#pragma clang diagnostic pop

// Right-hand sides h0 .. h0 + NRHS - 1 of the wavenumber in blk; partial rows into part [4][NRHS][C].
template <typename T, int NRHS>
__device__ __forceinline__ void stream_compute(const cplx<T> *__restrict__ blk, cplx<T> *__restrict__ part, int wave,
                                               int lane, int h0)
{
    using namespace stream;
    const cplx<T> *u = blk + GB + h0 * C;                        // [NRHS][C] of this pass
    const int s = lane >> 5, r = lane & 31, r15 = r & MASK;
    const int cb = wave * 8 + s * 4;
    const int trir = r * (r + 1) / 2;
    cplx<T> ux[NRHS][4], uy[NRHS][4], uz[NRHS][4];
#pragma unroll
    for (int h = 0; h < NRHS; ++h)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            ux[h][j] = u[h * C + cb + j];
            uy[h][j] = u[h * C + NZ + cb + j];
            uz[h][j] = u[h * C + 2 * NZ + cb + j];
        }
    cplx<T> yx[NRHS], yy[NRHS], yz[NRHS];
#pragma unroll
    for (int h = 0; h < NRHS; ++h) {
        yx[h] = mk<T>(0, 0);
        yy[h] = yx[h];
        yz[h] = yx[h];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int c = cb + j;
        const int pidx = c >= r ? r + c * (c + 1) / 2 : c + trir;    // packed (min, max)
        const cplx<T> gxx = blk[pidx], gxy = blk[S + pidx], gyy = blk[2 * S + pidx], gzz = blk[3 * S + pidx];
        const int ro = r * NZ + (c ^ r15), co = c * NZ + (r ^ (c & MASK));   // (r, c) and (c, r), swizzled
        const cplx<T> gxz = blk[4 * S + ro], gyz = blk[4 * S + A + ro];
        const cplx<T> gzx = blk[4 * S + co], gzy = blk[4 * S + A + co];
#pragma unroll
        for (int h = 0; h < NRHS; ++h) {
            yx[h] = cfma<T>(yx[h], gxx, ux[h][j]);
            yx[h] = cfma<T>(yx[h], gxy, uy[h][j]);
            yx[h] = cfma<T>(yx[h], gxz, uz[h][j]);
            yy[h] = cfma<T>(yy[h], gxy, ux[h][j]);
            yy[h] = cfma<T>(yy[h], gyy, uy[h][j]);
            yy[h] = cfma<T>(yy[h], gyz, uz[h][j]);
            yz[h] = cfms<T>(yz[h], gzx, ux[h][j]);             // -XZ^T u_x
            yz[h] = cfms<T>(yz[h], gzy, uy[h][j]);             // -YZ^T u_y
            yz[h] = cfma<T>(yz[h], gzz, uz[h][j]);
        }
    }
#pragma unroll
    for (int h = 0; h < NRHS; ++h) {
        yx[h].x += __shfl_xor(yx[h].x, 32, 64);
        yx[h].y += __shfl_xor(yx[h].y, 32, 64);
        yy[h].x += __shfl_xor(yy[h].x, 32, 64);
        yy[h].y += __shfl_xor(yy[h].y, 32, 64);
        yz[h].x += __shfl_xor(yz[h].x, 32, 64);
        yz[h].y += __shfl_xor(yz[h].y, 32, 64);
    }
    // both halves now hold the wave's row sums; half s stores right-hand side s
    if (NRHS == 1 && s) return;
    const int h = NRHS == 1 ? 0 : s;
    cplx<T> *pw = part + (wave * NRHS + h) * C;
    // component-wise selects (a select of whole vector values would go through scratch)
    const T vxx = h ? yx[NRHS - 1].x : yx[0].x, vxy = h ? yx[NRHS - 1].y : yx[0].y;
    const T vyx = h ? yy[NRHS - 1].x : yy[0].x, vyy = h ? yy[NRHS - 1].y : yy[0].y;
    const T vzx = h ? yz[NRHS - 1].x : yz[0].x, vzy = h ? yz[NRHS - 1].y : yz[0].y;
    pw[r] = mk<T>(vxx, vxy);
    pw[NZ + r] = mk<T>(vyx, vyy);
    pw[2 * NZ + r] = mk<T>(vzx, vzy);
}

// waits until at most `after` groups of PER_WAVE wave-loads are outstanding
template <int PW>
__device__ __forceinline__ void wait_groups(int after)
{
    switch (after) {
    case 0: wait_vm<0>(); break;
    case 1: wait_vm<PW>(); break;
    case 2: wait_vm<2 * PW>(); break;
    default: wait_vm<3 * PW>(); break;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void contract_stream_kernel(const cplx<T> *__restrict__ G, cplx<T> *buf, int64_t L,
                                                              int64_t M, int nrhs)
{
    using namespace stream;
    using K = Cfg<T>;
    static_assert(K::NBUF <= 4, "wait_groups covers up to three groups in flight after the current one");
    __shared__ cplx<T> ring[K::NBUF * K::BUF];
    __shared__ cplx<T> part[4 * 2 * C];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, t = threadIdx.x;
    const unsigned base = (unsigned)(size_t)(__attribute__((address_space(3))) cplx<T> *)ring;
    const int64_t g = gridDim.x;                                 // the launch keeps gridDim.x <= L
    const int64_t l0 = blockIdx.x;
    const int64_t cnt = (L - 1 - l0) / g + 1;                    // wavenumbers of this workgroup
#pragma unroll
    for (int p = 0; p < K::NBUF - 1; ++p)
        if (p < cnt)
            stream_issue<T>(G, buf, M, l0 + p * g, nrhs, base + (unsigned)(p * K::BUF * K::E), wave, lane);
    for (int64_t i = 0; i < cnt; ++i) {
        const int64_t l = l0 + i * g;
        const int64_t ahead = i + K::NBUF - 1;                   // the wavenumber index to prefetch now
        if (ahead < cnt)
            stream_issue<T>(G, buf, M, l0 + ahead * g, nrhs, base + (unsigned)((ahead % K::NBUF) * K::BUF * K::E),
                            wave, lane);
        // groups issued after wavenumber i's: i+1 .. min(cnt-1, i+NBUF-1)
        const int64_t after = (cnt - 1 - i) < (K::NBUF - 1) ? (cnt - 1 - i) : (K::NBUF - 1);
        wait_groups<K::PER_WAVE>((int)after);
        const cplx<T> *blk = ring + (i % K::NBUF) * K::BUF;
        // right-hand sides in passes of two over the same LDS-resident Green block
        for (int h0 = 0; h0 < nrhs; h0 += 2) {
            const int nh = nrhs - h0 >= 2 ? 2 : 1;
            XTR_BARRIER();                                       // block in LDS / previous pass read
            if (nh == 2) stream_compute<T, 2>(blk, part, wave, lane, h0);
            else stream_compute<T, 1>(blk, part, wave, lane, h0);
            XTR_BARRIER();                                       // partials complete
            if (t < nh * C) {
                cplx<T> acc = part[t];
#pragma unroll
                for (int w = 1; w < 4; ++w) acc = cadd<T>(acc, part[w * nh * C + t]);
                const int h = h0 + t / C;
                buf[h * M + l * C + (t % C)] = acc;
            }
        }
    }
    wait_vm<0>();                                                // no copy may land after exit
}

// ---- streaming contraction, nz = 64 --------------------------------------------------------------------
// A wavenumber's Green block (264 KB in FP64, 132 KB in FP32) is streamed as four units through a
// ring of 72-wave-load slots: U0 = xx|xy (+ the right-hand sides' spectrum columns), U1 = yy|zz,
// U2 = xz, U3 = yz. The units of all the workgroup's wavenumbers form one stream: step s contracts
// unit s while units s+1 .. s+NSLOT-1 are in flight (FP64: 2 slots of 72 KB; FP32: 4 slots of 36 KB).
// Wave w owns columns [16w, 16w+16); lane (hc, r) owns rows r and r + 32 and columns 16w + 8hc + j
// (j < 8); accumulators persist across a wavenumber's four units. The spectrum columns are copied
// from U0's slot into a small LDS area that outlives the slot; partial rows are reduced in U3's slot.
namespace stream64 {
constexpr int NZ = 64, C = 3 * NZ, S = NZ * (NZ + 1) / 2, A = NZ * NZ, GB = 4 * S + 2 * A;
constexpr int MASK = 15;                  // asym_swizzle_mask(64)
constexpr int SLOT = 72 * 64;             // elements per slot: 72 wave-loads of 1 KB (FP64), 36 (FP32)
constexpr int MAXRHS = 2;
__host__ __device__ constexpr int ubegin(int u) { return u == 0 ? 0 : u == 1 ? 2 * S : u == 2 ? 4 * S : 4 * S + A; }
__host__ __device__ constexpr int ulen(int u) { return u < 2 ? 2 * S : A; }
__host__ __device__ constexpr int uimage(int u) { return u == 0 ? 2 * S + MAXRHS * C : ulen(u); }
template <typename T>
struct Cfg {
    static constexpr int E = (int)sizeof(cplx<T>);   // bytes per element
    static constexpr int EPL = 16 / E;               // elements per lane and wave-load
    static constexpr int EPC = 64 * EPL;             // elements per wave-load
    static constexpr int NSLOT = E == 16 ? 2 : 4;
    __host__ __device__ static constexpr int upw(int u) { return ((uimage(u) + EPC - 1) / EPC + 3) / 4; }
    // wave-loads a wave has issued after unit u's when step u waits (`last`: no next wavenumber)
    __host__ __device__ static constexpr int ahead(int u, bool last)
    {
        int n = 0;
        for (int k = 1; k < NSLOT; ++k) {
            if (last && u + k > 3) break;
            n += upw((u + k) % 4);
        }
        return n;
    }
};
static_assert(4 * Cfg<double>::upw(0) * Cfg<double>::EPC <= SLOT && 4 * Cfg<float>::upw(0) * Cfg<float>::EPC <= SLOT,
              "slot size");
static_assert(2 * SLOT * 16 + MAXRHS * C * 16 <= 160 * 1024 && 4 * SLOT * 8 + MAXRHS * C * 8 <= 160 * 1024, "LDS");
static_assert(Cfg<double>::ahead(0, false) < 64 && Cfg<float>::ahead(0, false) < 64, "vmcnt range");
}  // namespace stream64

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
template <typename T, int U>
__device__ __forceinline__ void s64_issue(const cplx<T> *__restrict__ G, const cplx<T> *buf, int64_t M, int64_t l,
                                          int nrhs, unsigned dst, int wave, int lane)
{
    using namespace stream64;
    using K = Cfg<T>;
#pragma unroll
    for (int t = 0; t < K::upw(U); ++t) {
        const int k = wave + 4 * t;
        const int e = k * K::EPC + lane * K::EPL;        // first element of this unit's LDS image this lane copies
        const cplx<T> *src;
        if (e < ulen(U)) {
            src = G + l * GB + ubegin(U) + e;
        } else if (U == 0) {
            int q = e - ulen(0);                         // spectrum column q of this wavenumber
            if (q > nrhs * C - K::EPL) q = nrhs * C - K::EPL;   // padding lanes re-read valid data
            const int rhs = q / C;
            src = buf + rhs * M + l * C + (q - rhs * C);
        } else {
            src = G + l * GB + ubegin(U) + ulen(U) - K::EPL;     // padding lanes re-read valid data
        }
        const unsigned m0 = __builtin_amdgcn_readfirstlane(dst + (unsigned)k * 1024u);
        asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off nt" ::"s"(m0), "v"(src) : "memory", "m0");
    }
}
#pragma clang diagnostic pop

template <typename T, int U, int NRHS>
__device__ __forceinline__ void s64_compute(const cplx<T> *__restrict__ blk, const cplx<T> *__restrict__ ua, int wave,
                                            int lane, cplx<T> (&yx)[2][NRHS], cplx<T> (&yy)[2][NRHS],
                                            cplx<T> (&yz)[2][NRHS])
{
    using namespace stream64;
    const int hc = lane >> 5, r0 = lane & 31;
    const int cb = wave * 16 + hc * 8;
#pragma unroll 2
    for (int j = 0; j < 8; ++j) {
        const int c = cb + j;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int r = r0 + 32 * q;
            cplx<T> g0, g1;
            if (U < 2) {
                const int p = c >= r ? r + c * (c + 1) / 2 : c + r * (r + 1) / 2;   // packed (min, max)
                g0 = blk[p];
                g1 = blk[S + p];
            } else {
                g0 = blk[r * NZ + (c ^ (r & MASK))];       // B(r, c)
                g1 = blk[c * NZ + (r ^ (c & MASK))];       // B(c, r)
            }
#pragma unroll
            for (int h = 0; h < NRHS; ++h) {
                const cplx<T> *u = ua + h * C;
                if (U == 0) {          // xx, xy
                    const cplx<T> ux = u[c], uy = u[NZ + c];
                    yx[q][h] = cfma<T>(cfma<T>(yx[q][h], g0, ux), g1, uy);
                    yy[q][h] = cfma<T>(yy[q][h], g1, ux);
                } else if (U == 1) {   // yy, zz
                    yy[q][h] = cfma<T>(yy[q][h], g0, u[NZ + c]);
                    yz[q][h] = cfma<T>(yz[q][h], g1, u[2 * NZ + c]);
                } else if (U == 2) {   // xz: row product, then -xz^T
                    yx[q][h] = cfma<T>(yx[q][h], g0, u[2 * NZ + c]);
                    yz[q][h] = cfms<T>(yz[q][h], g1, u[c]);
                } else {               // yz: row product, then -yz^T
                    yy[q][h] = cfma<T>(yy[q][h], g0, u[2 * NZ + c]);
                    yz[q][h] = cfms<T>(yz[q][h], g1, u[NZ + c]);
                }
            }
        }
    }
}

// One step of the unit stream: contract unit U of wavenumber l (slot U % NSLOT) after issuing unit
// U + NSLOT - 1 into the slot the previous step freed.
template <typename T, int NRHS, int U>
__device__ __forceinline__ void s64_step(const cplx<T> *__restrict__ G, cplx<T> *buf, int64_t M, int64_t l,
                                         int64_t lnext, bool has_next, cplx<T> *ring, cplx<T> *ua, unsigned base,
                                         int wave, int lane, int t, cplx<T> (&yx)[2][NRHS], cplx<T> (&yy)[2][NRHS],
                                         cplx<T> (&yz)[2][NRHS])
{
    using namespace stream64;
    using K = Cfg<T>;
    constexpr int V = U + K::NSLOT - 1;                 // unit issued now
    constexpr unsigned vdst = (unsigned)((V % K::NSLOT) * SLOT * K::E);
    if (V <= 3) s64_issue<T, V % 4>(G, buf, M, l, NRHS, base + vdst, wave, lane);
    else if (has_next) s64_issue<T, V % 4>(G, buf, M, lnext, NRHS, base + vdst, wave, lane);
    if (has_next) wait_vm<K::ahead(U, false)>();
    else wait_vm<K::ahead(U, true)>();
    XTR_BARRIER();                                      // unit U in LDS for every wave
    cplx<T> *blk = ring + (U % K::NSLOT) * SLOT;
    if (U == 0) {
        for (int o = t; o < NRHS * C; o += 256) ua[o] = blk[ulen(0) + o];   // spectrum columns -> u area
        XTR_BARRIER();
    }
    s64_compute<T, U, NRHS>(blk, ua, wave, lane, yx, yy, yz);
    if (U == 3) {
        // rows r0 and r0 + 32 summed over the lane halves, then over the waves in this slot
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int h = 0; h < NRHS; ++h) {
                yx[q][h].x += __shfl_xor(yx[q][h].x, 32, 64);
                yx[q][h].y += __shfl_xor(yx[q][h].y, 32, 64);
                yy[q][h].x += __shfl_xor(yy[q][h].x, 32, 64);
                yy[q][h].y += __shfl_xor(yy[q][h].y, 32, 64);
                yz[q][h].x += __shfl_xor(yz[q][h].x, 32, 64);
                yz[q][h].y += __shfl_xor(yz[q][h].y, 32, 64);
            }
        XTR_BARRIER();                                  // slot read by every wave
        cplx<T> *part = blk;                            // [4][NRHS][C]
        const int hc = lane >> 5, r = (lane & 31) + 32 * hc;   // half hc stores row r0 + 32 hc
#pragma unroll
        for (int h = 0; h < NRHS; ++h) {
            cplx<T> *pw = part + (wave * NRHS + h) * C;
            // component-wise selects (a select of whole vector values would go through scratch)
            pw[r] = mk<T>(hc ? yx[1][h].x : yx[0][h].x, hc ? yx[1][h].y : yx[0][h].y);
            pw[NZ + r] = mk<T>(hc ? yy[1][h].x : yy[0][h].x, hc ? yy[1][h].y : yy[0][h].y);
            pw[2 * NZ + r] = mk<T>(hc ? yz[1][h].x : yz[0][h].x, hc ? yz[1][h].y : yz[0][h].y);
        }
        XTR_BARRIER();
        for (int o = t; o < NRHS * C; o += 256) {
            const int h = o / C, qq = o % C;
            cplx<T> acc = part[h * C + qq];
#pragma unroll
            for (int w = 1; w < 4; ++w) acc = cadd<T>(acc, part[(w * NRHS + h) * C + qq]);
            buf[h * M + l * C + qq] = acc;
        }
    }
    XTR_BARRIER();                                      // slot free for the next step's copy
}

template <typename T, int NRHS>
__global__ __launch_bounds__(256) void contract_stream64_kernel(const cplx<T> *__restrict__ G, cplx<T> *buf,
                                                                int64_t L, int64_t M)
{
    using namespace stream64;
    using K = Cfg<T>;
    __shared__ cplx<T> ring[K::NSLOT * SLOT];
    __shared__ cplx<T> ua[MAXRHS * C];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, t = threadIdx.x;
    const unsigned base = (unsigned)(size_t)(__attribute__((address_space(3))) cplx<T> *)ring;
    const int64_t g = gridDim.x;                                 // the launch keeps gridDim.x <= L
    const int64_t l0 = blockIdx.x;
    const int64_t cnt = (L - 1 - l0) / g + 1;
    // units 0 .. NSLOT-2 of the first wavenumber (NSLOT - 1 <= 3)
    s64_issue<T, 0>(G, buf, M, l0, NRHS, base, wave, lane);
    if (K::NSLOT > 2) {
        s64_issue<T, 1>(G, buf, M, l0, NRHS, base + (unsigned)(SLOT * K::E), wave, lane);
        s64_issue<T, 2>(G, buf, M, l0, NRHS, base + (unsigned)(2 * SLOT * K::E), wave, lane);
    }
    for (int64_t i = 0; i < cnt; ++i) {
        const int64_t l = l0 + i * g, ln = l + g;
        const bool nx = i + 1 < cnt;
        cplx<T> yx[2][NRHS], yy[2][NRHS], yz[2][NRHS];
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int h = 0; h < NRHS; ++h) {
                yx[q][h] = mk<T>(0, 0);
                yy[q][h] = yx[q][h];
                yz[q][h] = yx[q][h];
            }
        s64_step<T, NRHS, 0>(G, buf, M, l, ln, nx, ring, ua, base, wave, lane, t, yx, yy, yz);
        s64_step<T, NRHS, 1>(G, buf, M, l, ln, nx, ring, ua, base, wave, lane, t, yx, yy, yz);
        s64_step<T, NRHS, 2>(G, buf, M, l, ln, nx, ring, ua, base, wave, lane, t, yx, yy, yz);
        s64_step<T, NRHS, 3>(G, buf, M, l, ln, nx, ring, ua, base, wave, lane, t, yx, yy, yz);
    }
    wait_vm<0>();                                                // no copy may land after exit
}

static bool contract_stream64(const ApplyArgs &a, hipStream_t s);

// XTR_CONTRACT_KERNEL=generic selects contract_kernel for every shape (to compare the two kernels).
static bool contract_force_generic()
{
    const char *e = std::getenv("XTR_CONTRACT_KERNEL");
    return e && std::strcmp(e, "generic") == 0;
}

static int device_cus()
{
    static int cus[64] = {0};
    int dev = 0;
    XTR_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw hip_error("device index out of range");
    if (!cus[dev]) XTR_CHECK(hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
    return cus[dev];
}

// nz = 64, one or two right-hand sides, FP64 or FP32
template <typename T>
static void launch_stream64(const ApplyArgs &a, unsigned grid, hipStream_t s)
{
    const int64_t L = a.g.L();
    if (a.nrhs == 2)
        hipLaunchKernelGGL((contract_stream64_kernel<T, 2>), dim3(grid), dim3(256), 0, s, (const cplx<T> *)a.G,
                           (cplx<T> *)a.buf, L, a.g.M());
    else
        hipLaunchKernelGGL((contract_stream64_kernel<T, 1>), dim3(grid), dim3(256), 0, s, (const cplx<T> *)a.G,
                           (cplx<T> *)a.buf, L, a.g.M());
}

static bool contract_stream64(const ApplyArgs &a, hipStream_t s)
{
    if (a.g.nz != stream64::NZ || (a.nrhs != 1 && a.nrhs != 2) || contract_force_generic()) return false;
    const unsigned grid = (unsigned)std::min<int64_t>(a.g.L(), device_cus());
    if (a.precision == 0) launch_stream64<double>(a, grid, s);
    else launch_stream64<float>(a, grid, s);
    XTR_LAUNCH_CHECK("contract_stream64_kernel");
    return true;
}

// The streaming kernel covers nz = 32 with one to four right-hand sides; other shapes use contract_kernel.
template <typename T>
static bool contract_stream(const ApplyArgs &a, hipStream_t s)
{
    if (a.g.nz != stream::NZ || a.nrhs < 1 || a.nrhs > stream::kMaxStreamRhs || contract_force_generic()) return false;
    const int64_t L = a.g.L();
    const unsigned grid = (unsigned)std::min<int64_t>(L, device_cus());
    hipLaunchKernelGGL((contract_stream_kernel<T>), dim3(grid), dim3(256), 0, s, (const cplx<T> *)a.G,
                       (cplx<T> *)a.buf, L, a.g.M(), a.nrhs);
    XTR_LAUNCH_CHECK("contract_stream_kernel");
    return true;
}

// ---- folded contraction on MFMA, nz = 32 -------------------------------------------------------------
// Folded block b serves up to 4 wavenumbers (its mirror images), so one LDS-resident block meets up to
// 4 images x 2 right-hand sides = 8 complex columns U (96 x 8), and Y = G U is a small GEMM for the
// 16x16x4 MFMA: B = [Re U | Im U] (96 x 16 real), A = Re G and Im G (96 x 96 each), D1 = Re G . B,
// D2 = Im G . B, then Re Y = D1[:, :8] - D2[:, 8:] and Im Y = D1[:, 8:] + D2[:, :8] (a lane and its
// xor-8 partner). Mirror signs: u_x (u_y) of an image with kx (ky) mirrored is negated on the way
// into B, and y_x (y_y) on the way out. Persistent: one 256-thread workgroup per CU walks blocks; each
// block and its spectrum columns are copied HBM -> LDS by global_load_lds_dwordx4 into a ring of NBUF
// slots (FP64 2 x 77 KB, FP32 4 x 39 KB) exactly as in the streaming kernel. The six 16-row tiles
// times (Re, Im) make 12 accumulators, three per wave; tiles 1 and 4 have their two parts on two
// waves and meet through the block's slot once the MFMAs have read it. Two workgroups share a CU,
// so one workgroup's copy runs under the other's MFMAs (FP64: one slot per workgroup; FP32: two,
// the next block copied during the current one's MFMA steps).
namespace fold32 {
constexpr int NZ = 32, C = 3 * NZ, S = NZ * (NZ + 1) / 2, A = NZ * NZ, GB = 4 * S + 2 * A;
constexpr int MASK = 15;                  // asym_swizzle_mask(32)
constexpr int NCOL = 8;                   // complex columns per block (images x right-hand sides)
constexpr int KSTEPS = C / 4;             // 24
template <typename T>
struct Cfg {
    static constexpr int E = (int)sizeof(cplx<T>);
    static constexpr int EPL = 16 / E;
    static constexpr int EPC = 64 * EPL;
    static constexpr int IMG = GB + NCOL * C;
    static constexpr int NLOAD = (IMG + EPC - 1) / EPC;     // wave-loads per block
    static constexpr int SLOT = NLOAD * EPC;
    // two workgroups per CU: FP64 one slot each (one computes while the other copies), FP32 two
    static constexpr int NBUF = E == 16 ? 1 : 2;
    static constexpr int PW_HI = (NLOAD + 3) / 4;           // wave-loads of the waves with the most
};
static_assert(2 * 1 * Cfg<double>::SLOT * 16 <= 160 * 1024, "LDS (FP64, two workgroups per CU)");
static_assert(2 * 2 * Cfg<float>::SLOT * 8 <= 160 * 1024, "LDS (FP32, two workgroups per CU)");
static_assert(2 * 64 * 4 * 8 <= Cfg<double>::SLOT * 16 && 2 * 64 * 4 * 4 <= Cfg<float>::SLOT * 8, "exchange fits a slot");
static_assert((Cfg<double>::SLOT * 16) % 32 == 0 && (Cfg<float>::SLOT * 8) % 32 == 0, "slots stay 32-byte aligned");
static_assert(Cfg<double>::PW_HI < 64 && Cfg<float>::PW_HI < 64, "vmcnt range");
static_assert(GB % 2 == 0 && C % 2 == 0, "two-element lanes stay inside one column");
static_assert(2 * KSTEPS >= Cfg<double>::PW_HI && 2 * KSTEPS >= Cfg<float>::PW_HI, "copy spread over the steps");
}  // namespace fold32

typedef double mfma_d4 __attribute__((ext_vector_type(4)));
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
template <typename T> struct MfmaAcc;
template <> struct MfmaAcc<double> { typedef mfma_d4 type; };
template <> struct MfmaAcc<float> { typedef mfma_f4 type; };
__device__ __forceinline__ mfma_d4 mfma16(double a, double b, mfma_d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mfma_f4 mfma16(float a, float b, mfma_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// Row of D held in register r of lane l (16x16x4 results; column = l % 16): the FP64 instruction
// spreads consecutive rows over the lane quarters, the FP32 one over the registers.
template <typename T>
__device__ __forceinline__ int mfma_row(int lane, int r)
{
    return sizeof(T) == 8 ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r;
}

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
// Copy of the next block, issued one wave-load at a time between the MFMA steps of the current one
// (all at once, a wave stalls on the full memory queues before it can start its MFMAs).
template <typename T>
struct Fold32Copy {
    const cplx<T> *G;
    const cplx<T> *buf;
    int64_t M, b;
    FoldImages im;
    int nrhs;
    unsigned dst;
    bool active;
};

template <typename T>
__device__ __forceinline__ void fold32_issue_one(const Fold32Copy<T> &cp, int t, int wave, int lane)
{
    using namespace fold32;
    using K = Cfg<T>;
    const int k = __builtin_amdgcn_readfirstlane(wave + 4 * t);   // uniform: the branches below are scalar
    if (!cp.active || k >= K::NLOAD) return;
    const int e = k * K::EPC + lane * K::EPL;
    const cplx<T> *src;
    if ((k + 1) * K::EPC <= GB || (k * K::EPC < GB && e < GB)) {
        // Green data (the second test: the one FP32 wave-load that straddles GB, its lower lanes)
        src = cp.G + cp.b * GB + e;
    } else {
        // spectrum column q at LDS position GB + q C, element kk at (kk + 2q) mod C (the rotation
        // spreads the columns over the LDS banks for the B reads)
        const int u = e - GB;
        int q = u / C, j = u - q * C;
        const int ncol = cp.im.n * cp.nrhs;
        if (q >= ncol) q = j = 0;                            // unused column / padding: valid data
        int kk = j - 2 * q;
        if (kk < 0) kk += C;
        const int img = q / cp.nrhs, h = q - img * cp.nrhs;
        src = cp.buf + h * cp.M + cp.im.l(img) * C + kk;
    }
    const unsigned m0 = __builtin_amdgcn_readfirstlane(cp.dst + (unsigned)k * 1024u);
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off nt" ::"s"(m0), "v"(src) : "memory", "m0");
}

template <typename T>
__device__ __forceinline__ void fold32_issue(const Fold32Copy<T> &cp, int wave, int lane)
{
#pragma unroll
    for (int t = 0; t < fold32::Cfg<T>::PW_HI; ++t) fold32_issue_one<T>(cp, t, wave, lane);
}
#pragma clang diagnostic pop

// A element (row 16 TILE + i, column 4 STEP + kl) of the 96 x 96 block matrix
// [xx xy xz; xy yy yz; -xz^T -yz^T zz]: its index in the folded block (complex elements), and whether
// the stored value enters negated (the -xz^T, -yz^T part).
template <int TILE, int STEP>
__device__ __forceinline__ int fold32_a_idx(int i, int kl)
{
    using namespace fold32;
    constexpr int o = TILE / 2, ci = STEP / 8;
    const int r = 16 * (TILE % 2) + i, c = 4 * (STEP % 8) + kl;
    if constexpr (o == ci || (o < 2 && ci < 2)) {
        constexpr int base = o != ci ? S : o == 0 ? 0 : o == 1 ? 2 * S : 3 * S;   // xy | xx, yy, zz
        return base + (c >= r ? r + c * (c + 1) / 2 : c + r * (r + 1) / 2);
    } else if constexpr (ci == 2) {
        constexpr int base = o == 0 ? 4 * S : 4 * S + A;                       // xz, yz (r, c)
        return base + r * NZ + (c ^ (r & MASK));
    } else {
        constexpr int base = ci == 0 ? 4 * S : 4 * S + A;                      // -xz^T, -yz^T
        return base + c * NZ + (r ^ (c & MASK));
    }
}
template <int TILE, int STEP>
constexpr bool fold32_a_neg = TILE / 2 == 2 && STEP / 8 < 2;

// D += (NEG ? -a : a) b. The FP64 instruction negates its A operand itself (the neg modifier);
// the FP32 one has no such modifier, the sign goes on the value.
template <bool NEG>
__device__ __forceinline__ mfma_d4 mfma16s(double a, double b, mfma_d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, NEG ? 1 : 0);
}
template <bool NEG>
__device__ __forceinline__ mfma_f4 mfma16s(float a, float b, mfma_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(NEG ? -a : a, b, c, 0, 0, 0);
}

// Per-lane byte offsets, within a slot, of the A values a wave reads at each k-step: the whole
// complex element of tile TF and part PH of tile TH. They depend only on the lane, so a wave
// computes them once and keeps them in registers for all its blocks.
template <typename T, int TF, int TH, int PH, int... Ss>
__device__ __forceinline__ void fold32_offsets(std::integer_sequence<int, Ss...>, int i, int kl,
                                               unsigned (&of)[fold32::KSTEPS], unsigned (&oh)[fold32::KSTEPS])
{
    constexpr unsigned E = sizeof(cplx<T>);
    ((of[Ss] = E * (unsigned)fold32_a_idx<TF, Ss>(i, kl)), ...);
    ((oh[Ss] = E * (unsigned)fold32_a_idx<TH, Ss>(i, kl) + (unsigned)(PH * sizeof(T))), ...);
}

template <typename T>
__device__ __forceinline__ cplx<T> lds_cplx(const cplx<T> *blk, unsigned off)
{
    return *reinterpret_cast<const cplx<T> *>(reinterpret_cast<const char *>(blk) + off);
}
template <typename T>
__device__ __forceinline__ T lds_part(const cplx<T> *blk, unsigned off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(blk) + off);
}

// One k-step: the A values of the next step are read before this step's MFMAs, so the LDS latency
// is covered by three MFMAs; two wave-loads of the next block's copy go out over the first steps.
template <typename T, int TF, int TH, int PH, int STEP>
__device__ __forceinline__ void fold32_step(const cplx<T> *__restrict__ blk, const T (&bf)[fold32::KSTEPS],
                                            const unsigned (&of)[fold32::KSTEPS],
                                            const unsigned (&oh)[fold32::KSTEPS], typename MfmaAcc<T>::type &f1,
                                            typename MfmaAcc<T>::type &f2, typename MfmaAcc<T>::type &hh,
                                            cplx<T> &gf, T &gh, const Fold32Copy<T> &cp, int wave, int lane)
{
    cplx<T> nf = gf;
    T nh = gh;
    if constexpr (STEP + 1 < fold32::KSTEPS) {
        nf = lds_cplx<T>(blk, of[STEP + 1]);
        nh = lds_part<T>(blk, oh[STEP + 1]);
    }
    if constexpr (2 * STEP < fold32::Cfg<T>::PW_HI) fold32_issue_one<T>(cp, 2 * STEP, wave, lane);
    if constexpr (2 * STEP + 1 < fold32::Cfg<T>::PW_HI) fold32_issue_one<T>(cp, 2 * STEP + 1, wave, lane);
    f1 = mfma16s<fold32_a_neg<TF, STEP>>(gf.x, bf[STEP], f1);
    f2 = mfma16s<fold32_a_neg<TF, STEP>>(gf.y, bf[STEP], f2);
    hh = mfma16s<fold32_a_neg<TH, STEP>>(gh, bf[STEP], hh);
    gf = nf;
    gh = nh;
}

// Wave's three accumulators over one block: tile TF (Re G and Im G parts) and part PH of tile TH.
template <typename T, int TF, int TH, int PH, int... Ss>
__device__ __forceinline__ void fold32_wave(std::integer_sequence<int, Ss...>, const cplx<T> *__restrict__ blk,
                                            const T (&bf)[fold32::KSTEPS], const unsigned (&of)[fold32::KSTEPS],
                                            const unsigned (&oh)[fold32::KSTEPS], typename MfmaAcc<T>::type &f1,
                                            typename MfmaAcc<T>::type &f2, typename MfmaAcc<T>::type &hh,
                                            const Fold32Copy<T> &cp, int wave, int lane)
{
    cplx<T> gf = lds_cplx<T>(blk, of[0]);
    T gh = lds_part<T>(blk, oh[0]);
    (fold32_step<T, TF, TH, PH, Ss>(blk, bf, of, oh, f1, f2, hh, gf, gh, cp, wave, lane), ...);
}

// Y of tile TILE from D1 (Re G . B) and D2 (Im G . B), stored with the image signs.
template <typename T, int TILE>
__device__ __forceinline__ void fold32_store(typename MfmaAcc<T>::type d1, typename MfmaAcc<T>::type d2,
                                             cplx<T> *buf, int64_t M, const FoldImages &im, int nrhs, int lane)
{
    using namespace fold32;
    constexpr int o = TILE / 2;
    const int n = lane & 15, q = n & 7, part = n >> 3;
    const int img = (q / nrhs) & 3, h = q % nrhs;
    const bool live = q < im.n * nrhs;
    const bool flip = (o == 0 && im.mx(img)) || (o == 1 && im.my(img));
    T *dst = reinterpret_cast<T *>(buf + h * M + im.l(img) * C + o * NZ) + part;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const T other = __shfl_xor(d2[r], 8, 64);
        T v = part ? d1[r] + other : d1[r] - other;
        if (flip) v = -v;
        const int rr = 16 * (TILE % 2) + mfma_row<T>(lane, r);
        if (live) dst[2 * rr] = v;
    }
}

// The block loop of wave `wave` (tiles TF, TH): every wave runs the same loop and barriers.
template <typename T, int TF, int TH, int PH>
__device__ __forceinline__ void fold32_run(const cplx<T> *__restrict__ G, cplx<T> *buf, int64_t Lq, int64_t M,
                                           int nrhs, int nx, int ny, cplx<T> *ring, unsigned base, int wave,
                                           int lane)
{
    using namespace fold32;
    using K = Cfg<T>;
    using acc_t = typename MfmaAcc<T>::type;
    const int64_t g = gridDim.x;                                 // the launch keeps gridDim.x <= Lq
    const int64_t b0 = blockIdx.x;
    const int64_t cnt = (Lq - 1 - b0) / g + 1;
    using Seq = std::make_integer_sequence<int, KSTEPS>;
    unsigned of[KSTEPS], oh[KSTEPS];
    fold32_offsets<T, TF, TH, PH>(Seq{}, lane & 15, lane >> 4, of, oh);
    Fold32Copy<T> cp{G, buf, M, 0, FoldImages{}, nrhs, 0u, false};
    if (K::NBUF == 2) {                                          // block 0 now, block it + 1 during block it
        cp.b = b0;
        cp.im = fold_images(b0, nx, ny);
        cp.dst = base;
        cp.active = true;
        fold32_issue<T>(cp, wave, lane);
    }
    for (int64_t it = 0; it < cnt; ++it) {
        const int64_t b = b0 + it * g;
        if (K::NBUF == 1) {                                      // the whole block now
            cp.b = b;
            cp.im = fold_images(b, nx, ny);
            cp.dst = base;
            cp.active = true;
            fold32_issue<T>(cp, wave, lane);
            cp.active = false;
        } else {
            cp.active = it + 1 < cnt;
            if (cp.active) {
                cp.b = b + g;
                cp.im = fold_images(cp.b, nx, ny);
                cp.dst = base + (unsigned)(((it + 1) % K::NBUF) * K::SLOT * K::E);
            }
        }
        wait_vm<0>();                                            // block it (the only copy in flight)
        XTR_BARRIER();                                           // block it in LDS for every wave
        cplx<T> *blk = ring + (it % K::NBUF) * K::SLOT;
        const FoldImages im = fold_images(b, nx, ny);
        // B fragments: lane (kl = lane / 16, n = lane % 16) holds B[4s + kl][n] for each k-step s
        T bf[KSTEPS];
        {
            const int n = lane & 15, q = n & 7, part = n >> 3, kl = lane >> 4;
            const int img = (q / nrhs) & 3;
            const T *ucol = reinterpret_cast<const T *>(blk + GB + q * C) + part;
            const bool fx = im.mx(img), fy = im.my(img);
#pragma unroll
            for (int st = 0; st < KSTEPS; ++st) {
                int j = 4 * st + kl + 2 * q;                     // rotated position of element 4 st + kl
                if (j >= C) j -= C;
                const T v = ucol[2 * j];
                const bool flip = (st < 8 && fx) || (st >= 8 && st < 16 && fy);
                bf[st] = flip ? -v : v;
            }
        }
        acc_t f1 = {0, 0, 0, 0}, f2 = f1, hh = f1;
        fold32_wave<T, TF, TH, PH>(Seq{}, blk, bf, of, oh, f1, f2, hh, cp, wave, lane);
        // tiles 1 and 4: the Im G part (waves 1, 3) joins the Re G part (waves 0, 2), through the slot
        acc_t (*exch)[64] = reinterpret_cast<acc_t (*)[64]>(blk);
        XTR_BARRIER();                                           // every wave has read the slot
        if (wave & 1) exch[wave >> 1][lane] = hh;
        XTR_BARRIER();
        acc_t h2 = f1;
        if (!(wave & 1)) h2 = exch[wave >> 1][lane];
        fold32_store<T, TF>(f1, f2, buf, M, im, nrhs, lane);
        if constexpr (PH == 0) fold32_store<T, TH>(hh, h2, buf, M, im, nrhs, lane);
        XTR_BARRIER();                                           // slot free (exchange read)
    }
    wait_vm<0>();                                                // no copy may land after exit
}

template <typename T>
__global__ __launch_bounds__(256, 2) void contract_fold32_kernel(const cplx<T> *__restrict__ G, cplx<T> *buf,
                                                                 int64_t Lq, int64_t M, int nrhs, int nx, int ny)
{
    using K = fold32::Cfg<T>;
    __shared__ __attribute__((aligned(32))) cplx<T> ring[K::NBUF * K::SLOT];   // 32: exchange vectors
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned base = (unsigned)(size_t)(__attribute__((address_space(3))) cplx<T> *)ring;
    switch (wave) {
    case 0: fold32_run<T, 0, 1, 0>(G, buf, Lq, M, nrhs, nx, ny, ring, base, wave, lane); break;
    case 1: fold32_run<T, 2, 1, 1>(G, buf, Lq, M, nrhs, nx, ny, ring, base, wave, lane); break;
    case 2: fold32_run<T, 3, 4, 0>(G, buf, Lq, M, nrhs, nx, ny, ring, base, wave, lane); break;
    default: fold32_run<T, 5, 4, 1>(G, buf, Lq, M, nrhs, nx, ny, ring, base, wave, lane); break;
    }
}

// ---- folded contraction on MFMA, nz = 64 -------------------------------------------------------------
// A folded block (264 KB in FP64, 132 KB in FP32) is streamed through LDS in units, through a ring of
// four slots of 33 wave-loads (33 KB) with three units in flight. FP64: nine units (the spectrum
// columns of the block's images, xx, xy, yy, zz, and the xz and yz blocks each as a diagonal and an
// off-diagonal pair of 32 x 32 quadrants); FP32: five (the spectrum, xx + xy, yy + zz, xz, yz: each
// a contiguous run of the block). Wave w = 4 g + t owns rows 16t..16t+15 of x, y and z and the k-half
// g (k in 32g..32g+31 of every input component); waves t and 4 + t add their partial sums at the end
// of the block. The spectrum unit is read once per block into registers: lane (kl, n) keeps B[k][n]
// of its 24 k-steps, B = [Re U | Im U] with the image's mirror signs applied. Two accumulators per
// tile, D1 = Re G . B and D2 = Im G . B, give Re Y = D1[:, :8] - D2[:, 8:] and Im Y = D1[:, 8:] +
// D2[:, :8] through one lane exchange at the end; xy's A values serve both of its products. The
// -xz^T, -yz^T blocks enter through the FP64 MFMA's negate modifier (FP32: negated values). In FP64 a
// quadrant pair holds the quadrants (row half h, column half h) (diagonal unit) or (h, 1 - h); the
// direct and the transposed product of wave (t, g) both need the quadrant (t / 2, g) or (g, t / 2),
// so each quadrant unit is contracted by the four waves with g == t / 2 (diagonal) or g != t / 2, one
// per SIMD. In FP32 every wave works on every unit.
namespace fold64 {
constexpr int NZ = 64, C = 3 * NZ, S = NZ * (NZ + 1) / 2, A = NZ * NZ, GB = 4 * S + 2 * A;
constexpr int MASK = 15;                  // asym_swizzle_mask(64)
constexpr int NCOL = 8, UCOLS = NCOL * C;
constexpr int NWAVE = 8, NTHR = 64 * NWAVE;
constexpr int KS = 8;                     // k-steps of one component's k-half (32 values, 4 per step)
constexpr int NSLOT = 4;
template <typename T>
struct Cfg {
    static constexpr bool WIDE = sizeof(T) == 4;            // FP32: whole xz, yz blocks in one slot
    static constexpr int NUNIT = WIDE ? 5 : 9;
    static constexpr int E = (int)sizeof(cplx<T>);
    static constexpr int EPL = 16 / E;
    static constexpr int EPC = 64 * EPL;
    __host__ __device__ static constexpr int uelems(int u)
    {
        return u == 0 ? UCOLS : WIDE ? (u <= 2 ? 2 * S : A) : (u <= 4 ? S : A / 2);
    }
    __host__ __device__ static constexpr int nw(int u) { return (uelems(u) + EPC - 1) / EPC; }
    static constexpr int SLOTW = 33;                        // the largest unit, in wave-loads
    static constexpr int SLOT = SLOTW * EPC;
    static constexpr int EXCH = 4 * 3 * 64;                 // exchange: [t][tile][lane] of 4 T
    // wave-loads of unit u issued by wave w
    __host__ __device__ static constexpr int cnt(int u, int w) { return (nw(u) - w + NWAVE - 1) / NWAVE; }
    // wave-loads a wave has issued after unit u's when step u waits (units u+1 .. u+NSLOT-2: unit
    // u+NSLOT-1 goes out after that step's barrier)
    __host__ __device__ static constexpr int ahead(int u, bool has_next, int w)
    {
        int n = 0;
        for (int k = 1; k < NSLOT - 1; ++k)
            if (u + k < NUNIT || has_next) n += cnt((u + k) % NUNIT, w);
        return n;
    }
    __host__ __device__ static constexpr bool fits()
    {
        for (int u = 0; u < NUNIT; ++u)
            if (nw(u) > SLOTW) return false;
        return true;
    }
};
static_assert(Cfg<double>::fits() && Cfg<float>::fits(), "every unit fits a slot");
static_assert(NSLOT * Cfg<double>::SLOT * 16 + Cfg<double>::EXCH * 4 * 8 <= 160 * 1024, "LDS (FP64)");
static_assert(NSLOT * Cfg<float>::SLOT * 8 + Cfg<float>::EXCH * 4 * 4 <= 160 * 1024, "LDS (FP32)");
static_assert(Cfg<double>::ahead(0, true, 0) < 64 && Cfg<float>::ahead(0, true, 0) < 64, "vmcnt range");
static_assert(NSLOT - 1 < Cfg<double>::NUNIT && NSLOT - 1 < Cfg<float>::NUNIT, "look-ahead within one block");
static_assert(S % 2 == 0 && C % 2 == 0, "two-element lanes stay inside one run");
static_assert(Cfg<double>::nw(0) * Cfg<double>::EPC == UCOLS && Cfg<float>::nw(0) * Cfg<float>::EPC == UCOLS,
              "the spectrum unit has no padding lanes");
}  // namespace fold64

// Step U's wait: wave w lets through the wave-loads it issued after unit U's (counts differ by wave;
// `wave` is uniform, so the dispatch is scalar).
template <typename T, int U, bool HN, int W = 0>
__device__ __forceinline__ void fold64_wait_wave(int wave)
{
    if constexpr (W < fold64::NWAVE) {
        if (wave == W) wait_vm<fold64::Cfg<T>::ahead(U, HN, W)>();
        else fold64_wait_wave<T, U, HN, W + 1>(wave);
    }
}

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
// Copy of unit UT of block b into the slot at LDS address dst. In the spectrum unit, column q's
// element kk sits at position (kk + 2q) mod C, which spreads the columns over the LDS banks for the
// B reads.
template <typename T, int UT>
__device__ __forceinline__ void fold64_issue(const cplx<T> *__restrict__ G, const cplx<T> *buf, int64_t M, int64_t b,
                                             const FoldImages &im, int nrhs, unsigned dst, int wave, int lane)
{
    using namespace fold64;
    using K = Cfg<T>;
    constexpr int NWU = K::nw(UT);
#pragma unroll
    for (int t = 0; t < (NWU + NWAVE - 1) / NWAVE; ++t) {
        const int k = __builtin_amdgcn_readfirstlane(wave + NWAVE * t);
        if (k >= NWU) break;
        const int e = k * K::EPC + lane * K::EPL;
        const cplx<T> *src;
        if constexpr (UT == 0) {
            int q = e / C, j = e - q * C;
            if (q >= im.n * nrhs) q = j = 0;                 // unused column: valid data
            int kk = j - 2 * q;                              // even, so a lane's two FP32 elements
            if (kk < 0) kk += C;                             // stay contiguous
            const int img = q / nrhs, h = q - img * nrhs;
            src = buf + h * M + im.l(img) * C + kk;
        } else if constexpr (K::WIDE) {
            // contiguous runs: xx + xy, yy + zz (2S elements), xz, yz (A)
            constexpr int base = UT == 1 ? 0 : UT == 2 ? 2 * S : UT == 3 ? 4 * S : 4 * S + A;
            constexpr int len = K::uelems(UT);
            src = G + b * GB + base + (e < len ? e : len - K::EPL);   // padding lanes re-read
        } else if constexpr (UT <= 4) {
            src = G + b * GB + (UT - 1) * S + (e < S ? e : S - K::EPL);   // padding lanes re-read
        } else {
            // quadrant pair: rows 0-31 with column half H0, rows 32-63 with half 1 - H0 (diag: H0 = 0,
            // rows 32-63 with half 1; off: the other halves); 32 elements per quadrant row
            constexpr int base = UT <= 6 ? 4 * S : 4 * S + A;
            constexpr bool diag = UT == 5 || UT == 7;
            const int qd = e >> 10, w = e & 1023;
            const int rc = qd * 32 + (w >> 5), half = diag ? qd : 1 - qd;
            src = G + b * GB + base + rc * NZ + 32 * half + (w & 31);
        }
        const unsigned m0 = __builtin_amdgcn_readfirstlane(dst + (unsigned)k * 1024u);
        asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off nt" ::"s"(m0), "v"(src) : "memory", "m0");
    }
}
#pragma clang diagnostic pop

// Per-lane byte offsets of the A values of k-step s (k = 32 g + 4 s + kl, local to the component)
// for row r = 16 t + i, within a slot: the packed symmetric layout (po), and the direct (row r,
// column k: qo) and the transposed (row k, column r: to) element of an xz or yz block, as a
// quadrant pair (FP64) or the whole swizzled block (FP32). They depend only on the lane and the
// wave, so each wave computes them once.
template <typename T>
__device__ __forceinline__ void fold64_offsets(int t, int g, int lane, unsigned (&po)[fold64::KS],
                                               unsigned (&qo)[fold64::KS], unsigned (&to)[fold64::KS])
{
    using namespace fold64;
    constexpr unsigned E = sizeof(cplx<T>);
    const int r = 16 * t + (lane & 15), kl = lane >> 4;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int c = 32 * g + 4 * s + kl, cq = 4 * s + kl;  // cq = c & 31
        po[s] = E * (unsigned)(c >= r ? r + c * (c + 1) / 2 : c + r * (r + 1) / 2);
        if constexpr (Cfg<T>::WIDE) {
            qo[s] = E * (unsigned)(r * NZ + (c ^ (r & MASK)));
            to[s] = E * (unsigned)(c * NZ + (r ^ (c & MASK)));
        } else {
            qo[s] = E * (unsigned)((r >> 5) * 1024 + (r & 31) * 32 + (cq ^ (r & MASK)));
            to[s] = E * (unsigned)(g * 1024 + cq * 32 + ((r & 31) ^ (cq & MASK)));
        }
    }
}

// D1 += (NEG ? -Re a : Re a) b, D2 += (NEG ? -Im a : Im a) b over the 8 k-steps of one A tile at byte
// offset OFF + off[s] of the slot.
template <typename T, bool NEG, unsigned OFF = 0>
__device__ __forceinline__ void fold64_job(const cplx<T> *__restrict__ blk, const unsigned (&off)[fold64::KS],
                                           const T (&b)[fold64::KS], typename MfmaAcc<T>::type &d1,
                                           typename MfmaAcc<T>::type &d2)
{
    cplx<T> a[fold64::KS];
#pragma unroll
    for (int s = 0; s < fold64::KS; ++s) a[s] = lds_cplx<T>(blk, OFF + off[s]);
#pragma unroll
    for (int s = 0; s < fold64::KS; ++s) {
        d1 = mfma16s<NEG>(a[s].x, b[s], d1);
        d2 = mfma16s<NEG>(a[s].y, b[s], d2);
    }
}

// xy . u_y -> y_x and xy . u_x -> y_y from the same A values (xy at byte offset OFF of the slot).
template <typename T, unsigned OFF>
__device__ __forceinline__ void fold64_xy(const cplx<T> *__restrict__ blk, const unsigned (&po)[fold64::KS],
                                          const T (&bf)[3][fold64::KS], typename MfmaAcc<T>::type (&d1)[3],
                                          typename MfmaAcc<T>::type (&d2)[3])
{
    cplx<T> a[fold64::KS];
#pragma unroll
    for (int s = 0; s < fold64::KS; ++s) a[s] = lds_cplx<T>(blk, OFF + po[s]);
#pragma unroll
    for (int s = 0; s < fold64::KS; ++s) {
        d1[0] = mfma16s<false>(a[s].x, bf[1][s], d1[0]);
        d2[0] = mfma16s<false>(a[s].y, bf[1][s], d2[0]);
        d1[1] = mfma16s<false>(a[s].x, bf[0][s], d1[1]);
        d2[1] = mfma16s<false>(a[s].y, bf[0][s], d2[1]);
    }
}

// The MFMAs of unit UT for wave (t, g).
template <typename T, int UT>
__device__ __forceinline__ void fold64_compute(const cplx<T> *__restrict__ blk, const T (&bf)[3][fold64::KS],
                                               const unsigned (&po)[fold64::KS], const unsigned (&qo)[fold64::KS],
                                               const unsigned (&to)[fold64::KS], int t, int g,
                                               typename MfmaAcc<T>::type (&d1)[3], typename MfmaAcc<T>::type (&d2)[3])
{
    using namespace fold64;
    constexpr unsigned SB = (unsigned)S * sizeof(cplx<T>);     // bytes of one packed block
    if constexpr (Cfg<T>::WIDE) {
        if constexpr (UT == 1) {
            fold64_job<T, false>(blk, po, bf[0], d1[0], d2[0]);      // xx . u_x -> y_x
            fold64_xy<T, SB>(blk, po, bf, d1, d2);
        } else if constexpr (UT == 2) {
            fold64_job<T, false>(blk, po, bf[1], d1[1], d2[1]);      // yy . u_y -> y_y
            fold64_job<T, false, SB>(blk, po, bf[2], d1[2], d2[2]);  // zz . u_z -> y_z
        } else {
            constexpr int o = UT == 3 ? 0 : 1;                       // xz (x) or yz (y)
            fold64_job<T, false>(blk, qo, bf[2], d1[o], d2[o]);      // xz . u_z -> y_x
            fold64_job<T, true>(blk, to, bf[o], d1[2], d2[2]);       // -xz^T . u_x -> y_z
        }
    } else if constexpr (UT == 1) {
        fold64_job<T, false>(blk, po, bf[0], d1[0], d2[0]);
    } else if constexpr (UT == 2) {
        fold64_xy<T, 0>(blk, po, bf, d1, d2);
    } else if constexpr (UT == 3) {
        fold64_job<T, false>(blk, po, bf[1], d1[1], d2[1]);
    } else if constexpr (UT == 4) {
        fold64_job<T, false>(blk, po, bf[2], d1[2], d2[2]);
    } else {
        constexpr int o = UT <= 6 ? 0 : 1;
        constexpr bool diag = UT == 5 || UT == 7;
        if ((g == (t >> 1)) == diag) {
            fold64_job<T, false>(blk, qo, bf[2], d1[o], d2[o]);
            fold64_job<T, true>(blk, to, bf[o], d1[2], d2[2]);
        }
    }
}

// One step of the unit stream: wait for unit U of block b (slot s % NSLOT); after the barrier every
// wave is done with the previous step's slot, so unit U + NSLOT - 1 is copied into it; then contract
// (unit 0: read this wave's B values).
template <typename T, int U>
__device__ __forceinline__ void fold64_step(const cplx<T> *__restrict__ G, cplx<T> *buf, int64_t M, int64_t b,
                                            int64_t bn, bool has_next, const FoldImages &im, const FoldImages &imn,
                                            int nrhs, int64_t s, cplx<T> *ring, unsigned base, int wave, int lane,
                                            T (&bf)[3][fold64::KS], const unsigned (&po)[fold64::KS],
                                            const unsigned (&qo)[fold64::KS], const unsigned (&to)[fold64::KS],
                                            typename MfmaAcc<T>::type (&d1)[3], typename MfmaAcc<T>::type (&d2)[3])
{
    using namespace fold64;
    using K = Cfg<T>;
    if (has_next) fold64_wait_wave<T, U, true>(wave);
    else fold64_wait_wave<T, U, false>(wave);
    XTR_BARRIER();                                         // unit U in LDS; previous slot's reads done
    constexpr int V = U + NSLOT - 1;                       // unit copied now, into the previous step's slot
    const unsigned vdst = base + (unsigned)(((s + NSLOT - 1) % NSLOT) * K::SLOT * K::E);
    if constexpr (V < K::NUNIT) fold64_issue<T, V>(G, buf, M, b, im, nrhs, vdst, wave, lane);
    else if (has_next) fold64_issue<T, V - K::NUNIT>(G, buf, M, bn, imn, nrhs, vdst, wave, lane);
    const cplx<T> *blk = ring + (s % NSLOT) * K::SLOT;
    const int t = wave & 3, g = wave >> 2;
    if constexpr (U == 0) {
        // lane (kl, n): B[k][n] for k = 64 ci + 32 g + 4 s + kl, column q = n % 8, Re (n < 8) or Im
        const int n = lane & 15, q = n & 7, part = n >> 3, kl = lane >> 4;
        const int img = (q / nrhs) & 3;
        const bool fx = im.mx(img) != 0, fy = im.my(img) != 0;
        const T *ucol = reinterpret_cast<const T *>(blk + q * C) + part;
#pragma unroll
        for (int ci = 0; ci < 3; ++ci)
#pragma unroll
            for (int st = 0; st < KS; ++st) {
                int j = 64 * ci + 32 * g + 4 * st + kl + 2 * q;
                if (j >= C) j -= C;
                const T v = ucol[2 * j];
                const bool flip = (ci == 0 && fx) || (ci == 1 && fy);
                bf[ci][st] = flip ? -v : v;
            }
    } else {
        fold64_compute<T, U>(blk, bf, po, qo, to, t, g, d1, d2);
    }
}

// Tile of rows 16t.. from D1 and D2 (the Re/Im exchange between lanes n and n ^ 8).
template <typename T>
__device__ __forceinline__ typename MfmaAcc<T>::type fold64_combine(typename MfmaAcc<T>::type d1,
                                                                   typename MfmaAcc<T>::type d2, int lane)
{
    const int part = (lane >> 3) & 1;
    typename MfmaAcc<T>::type y;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const T other = __shfl_xor(d2[r], 8, 64);
        y[r] = part ? d1[r] + other : d1[r] - other;
    }
    return y;
}

template <typename T, int O>
__device__ __forceinline__ void fold64_store(typename MfmaAcc<T>::type y, cplx<T> *buf, int64_t M,
                                             const FoldImages &im, int nrhs, int t, int lane)
{
    using namespace fold64;
    const int n = lane & 15, q = n & 7, part = n >> 3;
    const int img = (q / nrhs) & 3, h = q % nrhs;
    const bool live = q < im.n * nrhs;
    const bool flip = (O == 0 && im.mx(img)) || (O == 1 && im.my(img));
    T *dst = reinterpret_cast<T *>(buf + h * M + im.l(img) * C + O * NZ) + part;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int rr = 16 * t + mfma_row<T>(lane, r);
        const T v = flip ? -y[r] : y[r];
        if (live) dst[2 * rr] = v;
    }
}

// All the block's steps, 0 .. NUNIT - 1.
template <typename T, int... Us>
__device__ __forceinline__ void fold64_block(std::integer_sequence<int, Us...>, const cplx<T> *__restrict__ G,
                                             cplx<T> *buf, int64_t M, int64_t b, int64_t bn, bool hn,
                                             const FoldImages &im, const FoldImages &imn, int nrhs, int64_t s,
                                             cplx<T> *ring, unsigned base, int wave, int lane,
                                             T (&bf)[3][fold64::KS], const unsigned (&po)[fold64::KS],
                                             const unsigned (&qo)[fold64::KS], const unsigned (&to)[fold64::KS],
                                             typename MfmaAcc<T>::type (&d1)[3], typename MfmaAcc<T>::type (&d2)[3])
{
    (fold64_step<T, Us>(G, buf, M, b, bn, hn, im, imn, nrhs, s + Us, ring, base, wave, lane, bf, po, qo, to, d1, d2),
     ...);
}

template <typename T>
__global__ __launch_bounds__(fold64::NTHR) void contract_fold64_kernel(const cplx<T> *__restrict__ G, cplx<T> *buf, int64_t Lq,
                                                              int64_t M, int nrhs, int nx, int ny)
{
    using namespace fold64;
    using K = Cfg<T>;
    using acc_t = typename MfmaAcc<T>::type;
    __shared__ cplx<T> ring[NSLOT * K::SLOT];
    __shared__ acc_t exch[K::EXCH];                              // [t][tile][lane]: waves 4..7 -> 0..3
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int t = wave & 3, g = wave >> 2;
    const unsigned base = (unsigned)(size_t)(__attribute__((address_space(3))) cplx<T> *)ring;
    const int64_t gs = gridDim.x;                                // the launch keeps gridDim.x <= Lq
    const int64_t b0 = blockIdx.x;
    const int64_t cnt = (Lq - 1 - b0) / gs + 1;
    unsigned po[KS], qo[KS], to[KS];
    fold64_offsets<T>(t, g, lane, po, qo, to);
    {
        // units 0 .. NSLOT-2 of the first block into slots 0 .. NSLOT-2
        const FoldImages im0 = fold_images(b0, nx, ny);
        fold64_issue<T, 0>(G, buf, M, b0, im0, nrhs, base, wave, lane);
        fold64_issue<T, 1>(G, buf, M, b0, im0, nrhs, base + 1u * K::SLOT * K::E, wave, lane);
        fold64_issue<T, 2>(G, buf, M, b0, im0, nrhs, base + 2u * K::SLOT * K::E, wave, lane);
    }
    static_assert(NSLOT == 4, "the prologue issues units 0 .. 2");
    T bf[3][KS];
    for (int64_t i = 0; i < cnt; ++i) {
        const int64_t b = b0 + i * gs, bn = b + gs;
        const bool hn = i + 1 < cnt;
        const FoldImages im = fold_images(b, nx, ny);
        const FoldImages imn = fold_images(hn ? bn : b, nx, ny);
        acc_t d1[3], d2[3];
#pragma unroll
        for (int o = 0; o < 3; ++o) d1[o] = d2[o] = acc_t{0, 0, 0, 0};
        fold64_block<T>(std::make_integer_sequence<int, K::NUNIT>{}, G, buf, M, b, bn, hn, im, imn, nrhs,
                        i * K::NUNIT, ring, base, wave, lane, bf, po, qo, to, d1, d2);
        acc_t y[3];
#pragma unroll
        for (int o = 0; o < 3; ++o) y[o] = fold64_combine<T>(d1[o], d2[o], lane);
        // waves 4..7 hand their partial sums to waves 0..3; the exchange area was last read before
        // this block's first barrier
        if (g == 1)
#pragma unroll
            for (int o = 0; o < 3; ++o) exch[(t * 3 + o) * 64 + lane] = y[o];
        XTR_BARRIER();
        if (g == 0) {
#pragma unroll
            for (int o = 0; o < 3; ++o) y[o] += exch[(t * 3 + o) * 64 + lane];
            fold64_store<T, 0>(y[0], buf, M, im, nrhs, t, lane);
            fold64_store<T, 1>(y[1], buf, M, im, nrhs, t, lane);
            fold64_store<T, 2>(y[2], buf, M, im, nrhs, t, lane);
        }
    }
    wait_vm<0>();                                                // no copy may land after exit
}

// Folded tensor, nz = 64: right-hand sides in launches of up to two.
static bool contract_fold64(const ApplyArgs &a, hipStream_t s)
{
    if (!a.folded || a.g.nz != fold64::NZ || contract_force_generic()) return false;
    const int64_t Lq = a.g.Lq();
    const unsigned grid = (unsigned)std::min<int64_t>(Lq, device_cus());
    for (int h0 = 0; h0 < a.nrhs; h0 += 2) {
        const int nh = std::min(2, a.nrhs - h0);
        if (a.precision == 0)
            hipLaunchKernelGGL((contract_fold64_kernel<double>), dim3(grid), dim3(fold64::NTHR), 0, s,
                               (const double2 *)a.G, (double2 *)a.buf + h0 * a.g.M(), Lq, a.g.M(), nh, a.g.nx, a.g.ny);
        else
            hipLaunchKernelGGL((contract_fold64_kernel<float>), dim3(grid), dim3(fold64::NTHR), 0, s, (const float2 *)a.G,
                               (float2 *)a.buf + h0 * a.g.M(), Lq, a.g.M(), nh, a.g.nx, a.g.ny);
        XTR_LAUNCH_CHECK("contract_fold64_kernel");
    }
    return true;
}

// Folded tensor, nz = 32: right-hand sides in launches of up to two.
static bool contract_fold32(const ApplyArgs &a, hipStream_t s)
{
    if (!a.folded || a.g.nz != fold32::NZ || contract_force_generic()) return false;
    const int64_t Lq = a.g.Lq();
    const unsigned grid = (unsigned)std::min<int64_t>(Lq, 2 * (int64_t)device_cus());   // two per CU
    for (int h0 = 0; h0 < a.nrhs; h0 += 2) {
        const int nh = std::min(2, a.nrhs - h0);
        if (a.precision == 0)
            hipLaunchKernelGGL((contract_fold32_kernel<double>), dim3(grid), dim3(256), 0, s, (const double2 *)a.G,
                               (double2 *)a.buf + h0 * a.g.M(), Lq, a.g.M(), nh, a.g.nx, a.g.ny);
        else
            hipLaunchKernelGGL((contract_fold32_kernel<float>), dim3(grid), dim3(256), 0, s, (const float2 *)a.G,
                               (float2 *)a.buf + h0 * a.g.M(), Lq, a.g.M(), nh, a.g.nx, a.g.ny);
        XTR_LAUNCH_CHECK("contract_fold32_kernel");
    }
    return true;
}

// ---- folded contraction for small nz ---------------------------------------------------------------
// ApplyGreenTensorAlongZ (ConvolutionOperator.cs:274-327) over a folded tensor for the depths that
// have no kernel of their own (nz <= 31; from nz = 4 to 24 the operator handle times it against
// contract_foldm_kernel once per shape and takes the faster). A workgroup takes folded blocks b = blockIdx.x,
// += gridDim.x; each block is read from HBM once, expanded into six full nz x ld matrices in LDS
// (xx, xy, yy, zz, xz, yz; ld = nz rounded up to odd, so that rows and columns are both read without
// bank conflicts), and applied to the spectrum columns of all its images (<= 4) and right-hand sides
// (<= 8) at once, with the image's mirror signs on the x and y components of u and y. Thread
// (group cg, row q = (a, i)) forms row q of y for columns cg, cg + ngroups, ...: J accumulators.
// Each thread holds E elements of the block in registers, and the next block's are loaded while this
// one is contracted.
// The mirror signs are formed as integer flags with bitwise operators and applied as a factor of
// +-1: with hipcc (ROCm 7.2, gfx950) the short-circuit form
//   if ((a == 0 && mx) || (a == 1 && my)) v = -v;
// in this divergent code compiled to an empty conditional block for the second term, so lanes where
// only a == 1 && my held stored v unnegated (the same source run on CPU threads is exact).
namespace foldn {
// f(integral_constant<int, t>) for t = 0 .. N-1: array indices known at compile time, so per-thread
// arrays stay in registers
template <typename F, int... Ts>
__device__ __forceinline__ void static_for(std::integer_sequence<int, Ts...>, F &&f)
{
    (f(std::integral_constant<int, Ts>{}), ...);
}
constexpr int NT = 256;
constexpr int MAX_NZ = 31;
static_assert(4 * (MAX_NZ * (MAX_NZ + 1) / 2) + 2 * MAX_NZ * MAX_NZ <= 16 * NT, "a block is at most 16 elements per thread");
__host__ __device__ inline int ld_of(int nz) { return nz | 1; }
// row stride of the spectrum columns in LDS: 3nz rounded up to odd, so that the columns of a block's
// images (3nz complex apart; a multiple of 256 bytes at nz = 16) fall on different LDS banks
__host__ __device__ inline int ldu_of(int nz) { return (3 * nz) | 1; }

// A thread's share of one folded block's loads: Green elements e = tid + t NT (t < E) and spectrum
// elements tid + t NT (t < U; a block with more spectrum values loads the rest directly). What does
// not depend on the block is computed once per kernel: each Green element's LDS destinations (two for
// an off-diagonal element of a symmetric block), each spectrum element's column (image, right-hand
// side) and k, and the divisions behind them. The values of the next block are loaded while the
// current one is contracted.
template <typename T, int E>
struct Loads {
    static constexpr int U = 4;
    int gd1[E], gd2[E];                  // LDS destinations, -1: none
    T gre[E], gim[E];
    int uimg[U], uh[U], uk[U], ucomp[U]; // uimg -1: outside the largest block
    int us[U];                           // LDS position: column * ldu + k
    T ure[U], uim[U];

    __device__ void init(int tid, int nz, int nrhs, const uint16_t *__restrict__ packed_rc)
    {
        const int ld = ld_of(nz), msz = nz * ld, S = nz * (nz + 1) / 2, A = nz * nz, gblk = 4 * S + 2 * A;
        const int C = 3 * nz, swz = asym_swizzle_mask(nz);
        static_for(std::make_integer_sequence<int, E>{}, [&](auto t) {
            const int e = tid + t * NT;
            gd1[t] = gd2[t] = -1;
            if (e < 4 * S) {
                const int bi = e / S, o = e - bi * S;
                const int rc = packed_rc[2 * o], tr = packed_rc[2 * o + 1];
                gd1[t] = bi * msz + rc * ld + tr;
                if (rc != tr) gd2[t] = bi * msz + tr * ld + rc;
            } else if (e < gblk) {
                const int o = e - 4 * S, bi = 4 + o / A, pos = o % A;
                const int rc = pos / nz, tr = (pos % nz) ^ (rc & swz);
                gd1[t] = bi * msz + rc * ld + tr;
            }
        });
        static_for(std::make_integer_sequence<int, U>{}, [&](auto t) {
            const int e = tid + t * NT;
            uimg[t] = -1;
            uh[t] = uk[t] = ucomp[t] = us[t] = 0;
            if (e < 4 * nrhs * C) {
                const int j = e / C;
                uk[t] = e - j * C;
                us[t] = j * ldu_of(nz) + uk[t];
                uimg[t] = j / nrhs;
                uh[t] = j - uimg[t] * nrhs;
                ucomp[t] = uk[t] < nz ? 0 : uk[t] < 2 * nz ? 1 : 2;
            }
        });
    }
    // block b's values into registers; spectrum values with the image's mirror signs
    __device__ void load(const cplx<T> *__restrict__ G, int64_t b, int gblk, const cplx<T> *buf, int64_t M, int C,
                         const FoldImages &im)
    {
        const cplx<T> *gb = G + b * (int64_t)gblk;
        static_for(std::make_integer_sequence<int, E>{}, [&](auto t) {
            if (gd1[t] >= 0) {
                const cplx<T> g = gb[threadIdx.x + t * NT];
                gre[t] = g.x;
                gim[t] = g.y;
            }
        });
        static_for(std::make_integer_sequence<int, U>{}, [&](auto t) {
            ure[t] = uim[t] = T(0);
            if (uimg[t] >= 0 && uimg[t] < im.n) {
                const cplx<T> v = buf[uh[t] * M + im.l(uimg[t]) * C + uk[t]];
                const int flip = ((ucomp[t] == 0) & im.mx(uimg[t])) | ((ucomp[t] == 1) & im.my(uimg[t]));
                const T sg = flip ? T(-1) : T(1);
                ure[t] = sg * v.x;
                uim[t] = sg * v.y;
            }
        });
    }
    // the registers into LDS (gm: six nz x ld matrices, u: [images * nrhs][ldu])
    __device__ void store(cplx<T> *gm, cplx<T> *u, int ncols) const
    {
        static_for(std::make_integer_sequence<int, E>{}, [&](auto t) {
            const cplx<T> g = mk<T>(gre[t], gim[t]);
            if (gd1[t] >= 0) gm[gd1[t]] = g;
            if (gd2[t] >= 0) gm[gd2[t]] = g;
        });
        static_for(std::make_integer_sequence<int, U>{}, [&](auto t) {
            if (uimg[t] >= 0 && threadIdx.x + t * NT < ncols) u[us[t]] = mk<T>(ure[t], uim[t]);
        });
    }
};

// Spectrum values of block im beyond the U NT that Loads carries (more than 1024 values: several
// right-hand sides at larger nz), loaded directly.
template <typename T>
__device__ void load_rest(cplx<T> *u, const cplx<T> *buf, int64_t M, int nz, int nrhs, const FoldImages &im, int from)
{
    const int C = 3 * nz, ncols = im.n * nrhs * C;
    for (int e = from + (int)threadIdx.x; e < ncols; e += NT) {
        const int j = e / C, k = e - j * C;
        const int img = j / nrhs, h = j - img * nrhs;
        const cplx<T> v = buf[h * M + im.l(img) * C + k];
        const int flip = ((k < nz) & im.mx(img)) | ((k >= nz) & (k < 2 * nz) & im.my(img));
        const T sg = flip ? T(-1) : T(1);
        u[j * ldu_of(nz) + k] = mk<T>(sg * v.x, sg * v.y);
    }
}
}  // namespace foldn

template <typename T, int J, int E>
__global__ __launch_bounds__(foldn::NT) void contract_foldn_kernel(const cplx<T> *__restrict__ G, cplx<T> *buf,
                                                                  int64_t Lq, int64_t M, int nz, int nrhs, int nx,
                                                                  int ny, const uint16_t *__restrict__ packed_rc)
{
    extern __shared__ unsigned char smem_raw[];
    const int ld = foldn::ld_of(nz), C = 3 * nz, S = nz * (nz + 1) / 2, A = nz * nz;
    const int gblk = 4 * S + 2 * A;                        // the launch keeps gblk <= E * NT
    const int msz = nz * ld;
    cplx<T> *gm = reinterpret_cast<cplx<T> *>(smem_raw);   // [6][nz][ld]
    cplx<T> *u = gm + 6 * msz;                             // [images * nrhs][ldu]
    const int ldu = foldn::ldu_of(nz);
    const int tid = threadIdx.x;
    const int ngroups = foldn::NT / C;                     // the launch keeps C <= NT
    const int q = tid % C, cg = tid / C;
    const int a = q / nz, i = q % nz;
    // row q of the 3x3 block operator, one segment per component of u:
    //   y_x = XX u_x + XY u_y + XZ u_z; y_y = XY u_x + YY u_y + YZ u_z; y_z = -XZ^T u_x - YZ^T u_y + ZZ u_z
    int mat[3];
    bool trn[3];
    mat[0] = a == 0 ? 0 : a == 1 ? 1 : 4;
    mat[1] = a == 0 ? 1 : a == 1 ? 2 : 5;
    mat[2] = a == 0 ? 4 : a == 1 ? 5 : 3;
    trn[0] = a == 2;
    trn[1] = a == 2;
    trn[2] = false;
    using Lo = foldn::Loads<T, E>;
    Lo lo;
    lo.init(tid, nz, nrhs, packed_rc);
    if ((int64_t)blockIdx.x < Lq) lo.load(G, blockIdx.x, gblk, buf, M, C, fold_images(blockIdx.x, nx, ny));
    for (int64_t b = blockIdx.x; b < Lq; b += gridDim.x) {
        const FoldImages im = fold_images(b, nx, ny);
        const int ncol = im.n * nrhs;
        __syncthreads();                                   // the previous block's LDS reads are done
        lo.store(gm, u, ncol * C);
        if (ncol * C > Lo::U * foldn::NT) foldn::load_rest<T>(u, buf, M, nz, nrhs, im, Lo::U * foldn::NT);
        __syncthreads();
        if (b + gridDim.x < Lq) lo.load(G, b + gridDim.x, gblk, buf, M, C, fold_images(b + gridDim.x, nx, ny));
        if (cg >= ngroups) continue;
        cplx<T> acc[J];
#pragma unroll
        for (int jj = 0; jj < J; ++jj) acc[jj] = mk<T>(0, 0);
#pragma unroll
        for (int sgm = 0; sgm < 3; ++sgm) {
            const cplx<T> *m = gm + mat[sgm] * msz;
            const cplx<T> *row = trn[sgm] ? m + i : m + i * ld;
            const int step = trn[sgm] ? ld : 1;
            const cplx<T> *us = u + sgm * nz;
            for (int c = 0; c < nz; ++c) {
                cplx<T> g = row[c * step];
                if (trn[sgm]) g = mk<T>(-g.x, -g.y);
#pragma unroll
                for (int jj = 0; jj < J; ++jj) {
                    const int j = cg + jj * ngroups;
                    if (j < ncol) acc[jj] = cfma<T>(acc[jj], g, us[j * ldu + c]);
                }
            }
        }
#pragma unroll
        for (int jj = 0; jj < J; ++jj) {
            const int j = cg + jj * ngroups;
            if (j < ncol) {
                const int img = j / nrhs, h = j - img * nrhs;
                const int flip = ((a == 0) & im.mx(img)) | ((a == 1) & im.my(img));
                const T sg = flip ? T(-1) : T(1);
                buf[h * M + im.l(img) * C + q] = mk<T>(sg * acc[jj].x, sg * acc[jj].y);
            }
        }
    }
}

// ---- folded contraction for small nz on MFMA ---------------------------------------------------
// The same work as contract_foldn_kernel (same LDS layout, same loads) with the products on
// 16x16x4 MFMAs: per (row tile of 16 rows of the 3nz x 3nz block operator, column tile of 8 complex
// columns) a wave forms D1 = Re A . B and D2 = Im A . B with B = [Re U | Im U] (the 8 columns' spectrum
// values, held in registers for all k-steps), and Y = D1 - D2' / D1' + D2 by a lane exchange. Each
// A operand is read from the six expanded matrices through a per-lane address: row (a, i) of the
// operator and k = (segment, c) pick matrix, orientation and sign (-XZ^T, -YZ^T in the z rows);
// (segment, c) advance by 4 per k-step. The operands come from registers, so LDS carries about one
// Green value per lane and MFMA instead of one spectrum value per complex MAC.
namespace foldm {
constexpr int NT = foldn::NT;               // four waves; foldn::Loads assumes this count
constexpr int MIN_NZ = 4;                   // one k-step of 4 advances c across at most one segment
constexpr int MAX_NZ = 24;
constexpr int MAX_KS = (3 * MAX_NZ + 3) / 4;
static_assert(4 * (MAX_NZ * (MAX_NZ + 1) / 2) + 2 * MAX_NZ * MAX_NZ <= 12 * NT, "a block is at most 12 elements per thread");
}  // namespace foldm

template <typename T, int E>
__global__ __launch_bounds__(foldm::NT) void contract_foldm_kernel(const cplx<T> *__restrict__ G, cplx<T> *buf,
                                                                  int64_t Lq, int64_t M, int nz, int nrhs, int nx,
                                                                  int ny, const uint16_t *__restrict__ packed_rc)
{
    using acc_t = typename MfmaAcc<T>::type;
    extern __shared__ unsigned char smem_raw[];
    const int ld = foldn::ld_of(nz), C = 3 * nz, S = nz * (nz + 1) / 2, A = nz * nz;
    const int gblk = 4 * S + 2 * A;                        // the launch keeps gblk <= E * NT
    const int msz = nz * ld;
    const int KS = (C + 3) / 4, RT = (C + 15) / 16;        // k-steps, row tiles
    cplx<T> *gm = reinterpret_cast<cplx<T> *>(smem_raw);   // [6][nz][ld]: xx, xy, yy, zz, xz, yz
    cplx<T> *u = gm + 6 * msz;                             // [images * nrhs][ldu]
    const int ldu = foldn::ldu_of(nz);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n = lane & 15, kk = lane >> 4, part = n >> 3;
    using Lo = foldn::Loads<T, E>;
    Lo lo;
    lo.init(tid, nz, nrhs, packed_rc);
    if ((int64_t)blockIdx.x < Lq) lo.load(G, blockIdx.x, gblk, buf, M, C, fold_images(blockIdx.x, nx, ny));
    for (int64_t b = blockIdx.x; b < Lq; b += gridDim.x) {
        const FoldImages im = fold_images(b, nx, ny);
        const int ncol = im.n * nrhs;
        __syncthreads();                                   // the previous block's LDS reads are done
        lo.store(gm, u, ncol * C);
        if (ncol * C > Lo::U * foldn::NT) foldn::load_rest<T>(u, buf, M, nz, nrhs, im, Lo::U * foldn::NT);
        __syncthreads();
        if (b + gridDim.x < Lq) lo.load(G, b + gridDim.x, gblk, buf, M, C, fold_images(b + gridDim.x, nx, ny));
        const int nct = (ncol + 7) / 8;
        for (int it = wave; it < RT * nct; it += 4) {      // (row tile, column tile) items of this wave
            const int rt = it % RT, ct = it / RT;
            const int col = 8 * ct + (n & 7);
            const bool cval = col < ncol;
            // B[k][n] = Re (n < 8) or Im (n >= 8) of column col's value at k = 4 ks + kk
            T bfr[foldm::MAX_KS];
#pragma unroll
            for (int ks = 0; ks < foldm::MAX_KS; ++ks) {
                const int k = 4 * ks + kk;
                T v = T(0);
                if (ks < KS && cval && k < C) {
                    const cplx<T> w = u[col * ldu + k];
                    v = part ? w.y : w.x;
                }
                bfr[ks] = v;
            }
            // lane's A row: operator row 16 rt + n = (a, i); per segment the matrix base, the step along
            // c and the sign (row-major XX XY XZ / XY YY YZ / ZZ; the z rows read -XZ, -YZ by columns)
            const int row = 16 * rt + n;
            const bool rval = row < C;
            const int a = rval ? row / nz : 0, i = rval ? row - a * nz : 0;
            const int m0 = a == 0 ? 0 : a == 1 ? 1 : 4;
            const int m1 = a == 0 ? 1 : a == 1 ? 2 : 5;
            const int m2 = a == 0 ? 4 : a == 1 ? 5 : 3;
            const int zr = a == 2;                          // the z rows: segments 0, 1 transposed and negated
            const int base0 = m0 * msz + (zr ? i : i * ld), step0 = zr ? ld : 1;
            const int base1 = m1 * msz + (zr ? i : i * ld), step1 = zr ? ld : 1;
            const int base2 = m2 * msz + i * ld;
            const T s01 = zr ? T(-1) : T(1);
            const T rv = rval ? T(1) : T(0);
            int seg = 0, c = kk;                            // k = 4 ks + kk = seg nz + c (nz >= 4)
            if (c >= nz) { c -= nz; seg = 1; }
            acc_t d1 = {0, 0, 0, 0}, d2 = {0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < foldm::MAX_KS; ++ks) {
                if (ks < KS) {
                    const int base = seg == 0 ? base0 : seg == 1 ? base1 : base2;
                    const int step = seg == 0 ? step0 : seg == 1 ? step1 : 1;
                    const T f = seg < 2 ? s01 * rv : (seg == 2 ? rv : T(0));
                    const cplx<T> g = gm[seg < 3 ? base + c * step : 0];
                    d1 = mfma16(f * g.x, bfr[ks], d1);
                    d2 = mfma16(f * g.y, bfr[ks], d2);
                    c += 4;
                    if (c >= nz) { c -= nz; ++seg; }
                    if (c >= nz) { c -= nz; ++seg; }
                }
            }
            const acc_t y = fold64_combine<T>(d1, d2, lane);
            if (cval) {
                const int img = col / nrhs, h = col - img * nrhs;
                const int fx = im.mx(img), fy = im.my(img);
                T *dst = reinterpret_cast<T *>(buf + h * M + im.l(img) * C) + part;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int rr = 16 * rt + mfma_row<T>(lane, r);
                    if (rr < C) {
                        const int ar = rr / nz;
                        const int flip = ((ar == 0) & fx) | ((ar == 1) & fy);
                        dst[2 * rr] = flip ? -y[r] : y[r];
                    }
                }
            }
        }
    }
}

// Workgroups of `kernel` (NT threads, lds bytes of dynamic LDS) resident per CU, from the compiled
// kernel's registers and LDS; cached per kernel and LDS size. A persistent grid larger than
// CUs x this would run in rounds, each workgroup paying its prologue and first loads again.
static int resident_per_cu(const void *kernel, int nt, size_t lds)
{
    static std::mutex mu;
    static std::map<std::pair<const void *, size_t>, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({kernel, lds});
    if (it != cache.end()) return it->second;
    int nb = 0;
    XTR_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, nt, lds));
    nb = std::max(1, nb);
    cache[{kernel, lds}] = nb;
    return nb;
}

static int device_max_lds()
{
    static int lds[64] = {0};
    int dev = 0;
    XTR_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw hip_error("device index out of range");
    if (!lds[dev]) XTR_CHECK(hipDeviceGetAttribute(&lds[dev], hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return lds[dev];
}

template <typename T, int J, int E>
static void launch_foldn(const ApplyArgs &a, unsigned grid, size_t lds, hipStream_t s)
{
    const void *k = (const void *)contract_foldn_kernel<T, J, E>;
    grid = (unsigned)std::min<int64_t>(grid, (int64_t)resident_per_cu(k, foldn::NT, lds) * device_cus());
    hipLaunchKernelGGL((contract_foldn_kernel<T, J, E>), dim3(grid), dim3(foldn::NT), lds, s, (const cplx<T> *)a.G,
                       (cplx<T> *)a.buf, a.g.Lq(), a.g.M(), a.g.nz, a.nrhs, a.g.nx, a.g.ny, a.packed_rc);
}

template <typename T, int J>
static void launch_foldn_e(const ApplyArgs &a, unsigned grid, size_t lds, hipStream_t s)
{
    const int64_t per_thread = ceil_div(a.g.gblk(), foldn::NT);   // <= 16 for nz <= 31
    if (per_thread <= 4) launch_foldn<T, J, 4>(a, grid, lds, s);
    else if (per_thread <= 8) launch_foldn<T, J, 8>(a, grid, lds, s);
    else if (per_thread <= 12) launch_foldn<T, J, 12>(a, grid, lds, s);
    else launch_foldn<T, J, 16>(a, grid, lds, s);
}

// Folded tensor, nz <= 31, 1-8 right-hand sides in one launch (one Green-tensor read).
template <typename T>
static bool contract_foldn(const ApplyArgs &a, hipStream_t s)
{
    const int nz = a.g.nz;
    if (!a.folded || nz < 1 || nz > foldn::MAX_NZ || a.nrhs < 1 || a.nrhs > kMaxRhs || contract_force_generic())
        return false;
    const int C = 3 * nz;
    const size_t lds = sizeof(cplx<T>) * ((size_t)6 * nz * foldn::ld_of(nz) + (size_t)4 * a.nrhs * foldn::ldu_of(nz));
    if (lds > (size_t)device_max_lds()) return false;
    const int ngroups = foldn::NT / C;
    const int need = (4 * a.nrhs + ngroups - 1) / ngroups;    // columns per thread
    const int per_cu = std::max(1, std::min(8, (int)(device_max_lds() / lds)));
    const unsigned grid = (unsigned)std::min<int64_t>(a.g.Lq(), (int64_t)per_cu * device_cus());
    if (need <= 2) launch_foldn_e<T, 2>(a, grid, lds, s);
    else if (need <= 4) launch_foldn_e<T, 4>(a, grid, lds, s);
    else if (need <= 8) launch_foldn_e<T, 8>(a, grid, lds, s);
    else launch_foldn_e<T, 16>(a, grid, lds, s);
    XTR_LAUNCH_CHECK("contract_foldn_kernel");
    return true;
}

// XTR_CONTRACT_KERNEL=vector keeps small-nz folded tensors on contract_foldn_kernel, =mfma puts every
// depth contract_foldm_kernel supports on it (to compare the two).
static bool contract_kernel_is(const char *name)
{
    const char *e = std::getenv("XTR_CONTRACT_KERNEL");
    return e && std::strcmp(e, name) == 0;
}

template <typename T, int E>
static void launch_foldm(const ApplyArgs &a, unsigned grid, size_t lds, hipStream_t s)
{
    const void *k = (const void *)contract_foldm_kernel<T, E>;
    grid = (unsigned)std::min<int64_t>(grid, (int64_t)resident_per_cu(k, foldm::NT, lds) * device_cus());
    hipLaunchKernelGGL((contract_foldm_kernel<T, E>), dim3(grid), dim3(foldm::NT), lds, s, (const cplx<T> *)a.G,
                       (cplx<T> *)a.buf, a.g.Lq(), a.g.M(), a.g.nz, a.nrhs, a.g.nx, a.g.ny, a.packed_rc);
}

// Folded tensor, 1-8 right-hand sides in one launch, on MFMA: by the fixed rule from nz = 17 to 24.
// Measured against contract_foldn_kernel (two right-hand sides, FP64): 163 against 186 us per launch
// at nz = 24, but 63 against 48 us at nz = 16 and 38 against 30 us at nz = 8 and 12. The operator
// handle measures the two once per process and shape and sets a.contract_variant.
template <typename T>
static bool contract_foldm(const ApplyArgs &a, hipStream_t s)
{
    const int nz = a.g.nz;
    const int lowest = contract_kernel_is("mfma") || a.contract_variant == CONTRACT_MFMA ? foldm::MIN_NZ : 17;
    if (!a.folded || nz < lowest || nz > foldm::MAX_NZ || a.nrhs < 1 || a.nrhs > kMaxRhs ||
        contract_force_generic() || contract_kernel_is("vector") || a.contract_variant == CONTRACT_VECTOR)
        return false;
    const size_t lds = sizeof(cplx<T>) * ((size_t)6 * nz * foldn::ld_of(nz) + (size_t)4 * a.nrhs * foldn::ldu_of(nz));
    if (lds > (size_t)device_max_lds()) return false;
    const int per_cu = std::max(1, std::min(8, (int)(device_max_lds() / lds)));
    const unsigned grid = (unsigned)std::min<int64_t>(a.g.Lq(), (int64_t)per_cu * device_cus());
    const int64_t per_thread = ceil_div(a.g.gblk(), foldm::NT);
    if (per_thread <= 4) launch_foldm<T, 4>(a, grid, lds, s);
    else if (per_thread <= 8) launch_foldm<T, 8>(a, grid, lds, s);
    else launch_foldm<T, 12>(a, grid, lds, s);
    XTR_LAUNCH_CHECK("contract_foldm_kernel");
    return true;
}

bool contract_variants_apply(const ApplyArgs &a)
{
    return a.folded && a.g.nz >= foldm::MIN_NZ && a.g.nz <= foldm::MAX_NZ && a.nrhs >= 1 && a.nrhs <= kMaxRhs &&
           !std::getenv("XTR_CONTRACT_KERNEL");
}

template <typename T>
static void contract(const ApplyArgs &a, hipStream_t s)
{
    if (!a.folded && contract_stream<T>(a, s)) return;
    if (!a.folded && contract_stream64(a, s)) return;
    if (contract_fold32(a, s)) return;
    if (contract_fold64(a, s)) return;
    if (contract_foldm<T>(a, s)) return;
    if (contract_foldn<T>(a, s)) return;
    const int nz = a.g.nz, C = 3 * nz;
    const int per_w = a.nrhs * C;
    int W = per_w >= 256 ? 1 : 256 / per_w;
    if (W < 1) W = 1;
    int threads = (int)ceil_div((int64_t)W * per_w, 64) * 64;
    if (threads > 1024) throw hip_error("contraction: 3*nz*nrhs exceeds 1024 threads");
    const size_t lds = sizeof(cplx<T>) * ((size_t)W * nz * (nz + 1) + (size_t)W * per_w);
    const int64_t L = a.g.L();
    dim3 grid((unsigned)ceil_div(L, W));
    hipLaunchKernelGGL(contract_kernel<T>, grid, dim3(threads), lds, s, (const cplx<T> *)a.G, (cplx<T> *)a.buf, L,
                       nz, a.nrhs, W, a.g.M(), a.packed_rc, a.folded ? a.g.nx : 0, a.g.ny);
    XTR_LAUNCH_CHECK("contract_kernel");
}

void launch_contract(const ApplyArgs &a, hipStream_t s) { a.precision ? contract<float>(a, s) : contract<double>(a, s); }

// ---- Green tensor layout conversion -----------------------------------------------------------
__global__ void green_interleave_kernel(const double2 *xx, const double2 *xy, const double2 *xz, const double2 *yy,
                                        const double2 *yz, const double2 *zz, double2 *dev, int64_t L, int nz,
                                        int dir)
{
    const int S = nz * (nz + 1) / 2, A = nz * nz;
    const int64_t gblk = 4LL * S + 2LL * A;
    const int64_t total = L * gblk;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t l = e / gblk;
        const int64_t o = e % gblk;
        const double2 *src;
        int64_t si;
        if (o < 4LL * S) {
            const int b = (int)(o / S);
            const double2 *comps[4] = {xx, xy, yy, zz};
            src = comps[b];
            si = l * S + (o % S);
        } else {
            const int b = (int)((o - 4LL * S) / A);
            const int pos = (int)((o - 4LL * S) % A);          // device (swizzled) position
            const int rc = pos / nz, tr = (pos % nz) ^ (rc & asym_swizzle_mask(nz));
            src = b == 0 ? xz : yz;
            si = l * A + (int64_t)rc * nz + tr;
        }
        if (dir == 0) dev[e] = src[si];
        else const_cast<double2 *>(src)[si] = dev[e];
    }
}

void launch_green_interleave(const double2 *comp[6], double2 *dev, const Geom &g, hipStream_t s)
{
    hipLaunchKernelGGL(green_interleave_kernel, dim3(2048), dim3(256), 0, s, comp[0], comp[1], comp[2], comp[3],
                       comp[4], comp[5], dev, g.L(), g.nz, 0);
    XTR_LAUNCH_CHECK("green_interleave_kernel");
}

void launch_green_deinterleave(const double2 *dev, double2 *comp[6], const Geom &g, hipStream_t s)
{
    hipLaunchKernelGGL(green_interleave_kernel, dim3(2048), dim3(256), 0, s, comp[0], comp[1], comp[2], comp[3],
                       comp[4], comp[5], const_cast<double2 *>(dev), g.L(), g.nz, 1);
    XTR_LAUNCH_CHECK("green_deinterleave_kernel");
}

// ---- folded Green tensor -------------------------------------------------------------------------
// Sign of element o of a device block (xx, xy, yy, zz, xz, yz) in the image (mx, my) of its block.
__device__ __forceinline__ double fold_sign(int64_t o, int64_t S, int64_t A, int mx, int my)
{
    if (o < S) return 1.0;
    if (o < 2 * S) return (mx ^ my) ? -1.0 : 1.0;
    if (o < 4 * S) return 1.0;
    if (o < 4 * S + A) return mx ? -1.0 : 1.0;
    return my ? -1.0 : 1.0;
}

// dir 0: folded[b] = full[source wavenumber of b]; dir 1: full[l] = sign * folded[b(l)]
__global__ void green_fold_kernel(double2 *full, double2 *folded, int nx, int ny, int nz, int dir)
{
    const int64_t S = (int64_t)nz * (nz + 1) / 2, A = (int64_t)nz * nz, gblk = 4 * S + 2 * A;
    const int64_t total = (dir == 0 ? (int64_t)(nx + 1) * (ny + 1) : 4LL * nx * ny) * gblk;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = e / gblk, o = e % gblk;
        if (dir == 0) {
            const int64_t qx = k / (ny + 1), qy = k % (ny + 1);
            folded[e] = full[(qx * 2 * ny + qy) * gblk + o];
        } else {
            const FoldPos p = fold_pos(k, nx, ny);
            const double sg = fold_sign(o, S, A, p.mx, p.my);
            const double2 v = folded[p.b * gblk + o];
            full[e] = make_double2(sg * v.x, sg * v.y);
        }
    }
}

void launch_green_fold(const double2 *full, double2 *folded, const Geom &g, hipStream_t s)
{
    hipLaunchKernelGGL(green_fold_kernel, dim3(2048), dim3(256), 0, s, const_cast<double2 *>(full), folded, g.nx, g.ny,
                       g.nz, 0);
    XTR_LAUNCH_CHECK("green_fold_kernel");
}

void launch_green_unfold(const double2 *folded, double2 *full, const Geom &g, hipStream_t s)
{
    hipLaunchKernelGGL(green_fold_kernel, dim3(2048), dim3(256), 0, s, full, const_cast<double2 *>(folded), g.nx, g.ny,
                       g.nz, 1);
    XTR_LAUNCH_CHECK("green_unfold_kernel");
}

// Non-negative doubles order like their bit patterns, so an unsigned 64-bit atomic max is a max.
__global__ __launch_bounds__(256) void green_mirror_check_kernel(const double2 *full, int nx, int ny, int nz,
                                                                 unsigned long long *stat)
{
    const int64_t S = (int64_t)nz * (nz + 1) / 2, A = (int64_t)nz * nz, gblk = 4 * S + 2 * A;
    const int64_t total = 4LL * nx * ny * gblk;
    double m[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) m[c] = 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t l = e / gblk, o = e % gblk;
        const int c = o < 4 * S ? (int)(o / S) : 4 + (int)((o - 4 * S) / A);
        const double2 v = full[e];
        const double mag = fmax(fabs(v.x), fabs(v.y));
        const FoldPos p = fold_pos(l, nx, ny);
        double d = 0.0;
        if (p.mx || p.my) {
            const int64_t qx = p.b / (ny + 1), qy = p.b % (ny + 1);
            const double sg = fold_sign(o, S, A, p.mx, p.my);
            const double2 w = full[(qx * 2 * ny + qy) * gblk + o];
            d = fmax(fabs(v.x - sg * w.x), fabs(v.y - sg * w.y));
        }
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (k == c) {
                m[2 * k] = fmax(m[2 * k], d);
                m[2 * k + 1] = fmax(m[2 * k + 1], mag);
            }
    }
    __shared__ double red[12][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 12; ++c) {
        double v = m[c];
        for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
        if (lane == 0) red[c][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const double v = fmax(fmax(red[threadIdx.x][0], red[threadIdx.x][1]), fmax(red[threadIdx.x][2], red[threadIdx.x][3]));
        atomicMax(stat + threadIdx.x, (unsigned long long)__double_as_longlong(v));
    }
}

void launch_green_mirror_check(const double2 *full, const Geom &g, double *stat, hipStream_t s)
{
    XTR_CHECK(hipMemsetAsync(stat, 0, 12 * sizeof(double), s));
    hipLaunchKernelGGL(green_mirror_check_kernel, dim3(1024), dim3(256), 0, s, full, g.nx, g.ny, g.nz,
                       reinterpret_cast<unsigned long long *>(stat));
    XTR_LAUNCH_CHECK("green_mirror_check_kernel");
}

__global__ void to_f32_kernel(const double2 *src, float2 *dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = to_f(src[i]);
}

void launch_to_f32(const double2 *src, float2 *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(to_f32_kernel, dim3(4096), dim3(256), 0, s, src, dst, n);
    XTR_LAUNCH_CHECK("to_f32_kernel");
}

}  // namespace xtr
