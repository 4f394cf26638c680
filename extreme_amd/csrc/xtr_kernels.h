// Host-side launchers of the gfx950 kernels. Every launcher enqueues on the given stream and
// checks the launch; none synchronises.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace xtr {

constexpr int kMaxRhs = 8;

// Geometry of one operator (Nx, Ny, Nz) and derived sizes.
struct Geom {
    int nx, ny, nz;
    int64_t n() const { return 3LL * nx * ny * nz; }          // unknowns
    int64_t L() const { return 4LL * nx * ny; }               // padded lateral plane
    int64_t C() const { return 3LL * nz; }                    // spectrum columns per wavenumber
    int64_t M() const { return L() * C(); }                   // spectrum elements
    int64_t S() const { return (int64_t)nz * (nz + 1) / 2; }  // packed symmetric block
    int64_t gblk() const { return 4 * S() + 2LL * nz * nz; }  // Green elements per wavenumber
    int64_t Lq() const { return (int64_t)(nx + 1) * (ny + 1); } // blocks of a folded Green tensor
};

// Folded Green tensor. The mirrored lateral tensor is even or odd in x and in y per component
// (xx, yy, zz even/even; xy odd/odd; xz odd/even; yz even/odd), so its 2-D spectrum satisfies
// G(2Nx-kx, ky) = D_x G(kx, ky) D_x and G(kx, 2Ny-ky) = D_y G(kx, ky) D_y with D_x = diag(-I, I, I),
// D_y = diag(I, -I, I) on the (x, y, z) components. A folded tensor stores only kx <= Nx, ky <= Ny,
// block b = qx (Ny+1) + qy; wavenumber l = kx 2Ny + ky uses block b with u_x, y_x negated when
// kx > Nx (mx) and u_y, y_y negated when ky > Ny (my).
struct FoldPos {
    int64_t b;
    int mx, my;
};
__host__ __device__ inline FoldPos fold_pos(int64_t l, int nx, int ny)
{
    const int kx = (int)(l / (2 * ny)), ky = (int)(l % (2 * ny));
    FoldPos p;
    p.mx = kx > nx;
    p.my = ky > ny;
    const int qx = p.mx ? 2 * nx - kx : kx, qy = p.my ? 2 * ny - ky : ky;
    p.b = (int64_t)qx * (ny + 1) + qy;
    return p;
}
// Wavenumbers that use folded block b (its images): n = 1, 2 or 4. Image i < n has kx mirrored
// (mx) and/or ky mirrored (my); image 0 is the unmirrored wavenumber. Computed from i rather than
// stored in arrays: a lane-dependent index into a private array would put it in scratch.
struct FoldImages {
    int n;
    int qx, qy, nx, ny;
    bool two_x, two_y;
    __host__ __device__ int mx(int i) const { return two_x ? (i & 1) : 0; }
    __host__ __device__ int my(int i) const { return two_y ? (two_x ? (i >> 1) & 1 : i & 1) : 0; }
    __host__ __device__ int64_t l(int i) const
    {
        const int kx = mx(i) ? 2 * nx - qx : qx, ky = my(i) ? 2 * ny - qy : qy;
        return (int64_t)kx * 2 * ny + ky;
    }
};
__host__ __device__ inline FoldImages fold_images(int64_t b, int nx, int ny)
{
    FoldImages im;
    im.qx = (int)(b / (ny + 1));
    im.qy = (int)(b % (ny + 1));
    im.nx = nx;
    im.ny = ny;
    im.two_x = im.qx > 0 && im.qx < nx;
    im.two_y = im.qy > 0 && im.qy < ny;
    im.n = (im.two_x ? 2 : 1) * (im.two_y ? 2 : 1);
    return im;
}

// Device position of the asymmetric-block element (rc, tr) (xz and yz, nz*nz each): rows are
// XOR-swizzled, tr ^ (rc & mask), so that a block copied linearly into LDS can be read along rows
// and along columns without bank conflicts. mask = 15 when 16 | nz, 7 when 8 | nz, else 0
// (plain row-major). The reference layout at the set/get boundary is plain row-major [rc][tr].
__host__ __device__ inline int asym_swizzle_mask(int nz) { return nz % 16 == 0 ? 15 : (nz % 8 == 0 ? 7 : 0); }
__host__ __device__ inline int asym_pos(int rc, int tr, int nz, int mask) { return rc * nz + (tr ^ (rc & mask)); }

// Per-call vector pointers (device; complex128, or complex64 when the apply's vec32 is set) for up
// to kMaxRhs right-hand sides.
// R[i] is the R function of right-hand side i's model ([nx*ny*nz]); models that share the grid,
// background and frequency share the Green tensor, so one apply can serve several of them.
struct RhsPtrs {
    const double2 *x[kMaxRhs];
    double2 *y[kMaxRhs];
    const double2 *R[kMaxRhs];
};

// ---- operator apply (op_kernels.hip) -------------------------------------------------------
// Twiddle tables exp(-2 pi i k / P), k < P, in the apply precision.
// precision: 0 = FP64 (double2 buffers / Green), 1 = FP32 (float2).
struct ApplyArgs {
    Geom g;
    int precision;
    int nrhs;
    int type;                 // 0 = A, 1 = Chi0
    const double2 *R;         // [nx*ny*nz] FP64, used for right-hand sides whose v.R[i] is null
    const double *f;          // [nz] forward factors (real)
    const double *b;          // [nz] backward factors (real)
    const void *G;            // [nblk][gblk] in apply precision
    int folded;               // G holds the folded tensor (nblk = Lq) rather than all L wavenumbers
    const uint16_t *packed_rc;// [S][2] (rc, tr) of each packed element
    void *buf;                // nrhs * M spectrum in apply precision
    const void *tw_x;         // twiddles for P = 2 nx
    const void *tw_y;         // twiddles for P = 2 ny
    int fft_tile[4];          // column tile of fwd_y, fwd_x, inv_x, inv_y (FFT_TILE_*)
    int vec32;                // v.x and v.y point to complex64 vectors (float2; precision FP32 only)
    int contract_variant;     // folded nz <= 31: CONTRACT_* (CONTRACT_RULE: MFMA from nz = 17)
    RhsPtrs v;
};

enum { CONTRACT_RULE = 0, CONTRACT_VECTOR = 1, CONTRACT_MFMA = 2 };
// true when the apply can take either small-nz contraction (folded, 4 <= nz <= 24) and no
// XTR_CONTRACT_KERNEL override fixes the kernel
bool contract_variants_apply(const ApplyArgs &a);

enum {
    FFT_TILE_DEFAULT = 0,
    FFT_TILE_NARROW = 1,
    FFT_TILE_WIDE = 2,
    FFT_TILE_NARROW_T512 = 3,
    FFT_TILE_REG = 4,        // register-resident stages (lengths 64..512), the default column count
    FFT_TILE_REG_ALT = 5,    // the same with twice the columns (half where that exceeds 64 KB of LDS)
};
// the column tiles a pencil length offers in this precision (FFT_TILE_*), into tiles[8]; returns how many
int fft_tile_options(int p, int precision, int *tiles);

// Stage launchers (used separately for per-stage timing).
void launch_fwd_y(const ApplyArgs &a, hipStream_t s);
void launch_fwd_x(const ApplyArgs &a, hipStream_t s);
void launch_contract(const ApplyArgs &a, hipStream_t s);
void launch_inv_x(const ApplyArgs &a, hipStream_t s);
void launch_inv_y(const ApplyArgs &a, hipStream_t s);

// In-place forward 2-D FFT of a [2nx][2ny][ncols] array (Green tensor), FP64.
void launch_fft2_plain(double2 *data, int n1, int n2, int64_t ncols, const double2 *tw1, const double2 *tw2,
                       int sign, hipStream_t s);
bool fft_length_supported(int p);

// Reorders the Green tensor between the reference layout (six component arrays) and the
// device layout [L][xx S | xy S | yy S | zz S | xz nz^2 | yz nz^2].
void launch_green_interleave(const double2 *comp[6], double2 *dev, const Geom &g, hipStream_t s);
void launch_green_deinterleave(const double2 *dev, double2 *comp[6], const Geom &g, hipStream_t s);
// full [L][gblk] -> folded [Lq][gblk] (blocks kx <= Nx, ky <= Ny), and back with the mirror signs
void launch_green_fold(const double2 *full, double2 *folded, const Geom &g, hipStream_t s);
void launch_green_unfold(const double2 *folded, double2 *full, const Geom &g, hipStream_t s);
// stat[2c] = max |G_c(l) - s_c G_c(mirror source of l)|, stat[2c+1] = max |G_c| over a full
// tensor, c in device block order (xx, xy, yy, zz, xz, yz); stat (12 doubles) cleared here
void launch_green_mirror_check(const double2 *full, const Geom &g, double *stat, hipStream_t s);
void launch_to_f32(const double2 *src, float2 *dst, int64_t n, hipStream_t s);
void launch_probe_stream(const void *src, void *dst, int64_t blocks, int read_kb, int write_kb, hipStream_t s);

// Observation level (obs_kernels.hip): jQ from chi (operator prepared as A: R and f of that type),
// and the anomalous field at a level from jQ with the AtoO tensor G ([L][ncomp][nz]); spec holds
// L * 3nz and small L * 3 complex128 of scratch; field is [Nx][Ny][3].
void launch_jq(const double2 *chi, const double2 *js, const double2 *R, const double *f, const Geom &g, double2 *jq,
               hipStream_t s);
void launch_level_field(const double2 *G, int electric, const double2 *jq, const Geom &g, double2 *spec,
                        double2 *small, const double2 *tw_x, const double2 *tw_y, double2 *field, hipStream_t s);

// ---- Krylov (krylov_kernels.hip) -----------------------------------------------------------
// Vectors are complex128 (double2) or complex64 (float2, the mixed-precision correction solves);
// every kernel accumulates in FP64.
constexpr int kRedBlocks = 512;   // partial-sum blocks of the reductions
// partial[k][blk] of conj(V_k) . w, k < cnt (V columns contiguous, stride n)
void launch_block_dots(const double2 *V, int cnt, const double2 *w, int64_t n, double2 *partial, hipStream_t s);
void launch_block_dots(const float2 *V, int cnt, const float2 *w, int64_t n, double2 *partial, hipStream_t s);
// out[k] = sum_blk partial[k][blk]
void launch_reduce_c(const double2 *partial, int cnt, double2 *out, hipStream_t s, double2 *host_out = nullptr);
// w -= sum_k V_k h_k (h on device), partial norms of the new w
void launch_cgs_update(const double2 *V, int cnt, const double2 *h, double2 *w, int64_t n, double *partial,
                       hipStream_t s);
void launch_cgs_update(const float2 *V, int cnt, const double2 *h, float2 *w, int64_t n, double *partial,
                       hipStream_t s);
// partial sums of |w|^2 ; out = sum of partials
void launch_norm2(const double2 *w, int64_t n, double *partial, hipStream_t s);
void launch_norm2(const float2 *w, int64_t n, double *partial, hipStream_t s);
void launch_reduce_r(const double *partial, double *out, hipStream_t s, double *host_out = nullptr);
// dst = src / d  (v_{j+1} = w / |w|, Gmres.cs:267-268)
void launch_div(const double2 *src, double d, double2 *dst, int64_t n, hipStream_t s);
void launch_div(const float2 *src, double d, float2 *dst, int64_t n, hipStream_t s);
// dst = a * src  (V_0 = (1/beta) w, Gmres.cs:117; and the complex64 copy of a scaled residual)
void launch_mul(const double2 *src, double a, double2 *dst, int64_t n, hipStream_t s);
void launch_mul(const float2 *src, double a, float2 *dst, int64_t n, hipStream_t s);
void launch_mul(const double2 *src, double a, float2 *dst, int64_t n, hipStream_t s);
// dst = a - b
void launch_sub(const double2 *a, const double2 *b, double2 *dst, int64_t n, hipStream_t s);
void launch_sub(const float2 *a, const float2 *b, float2 *dst, int64_t n, hipStream_t s);
// x += sum_k B_k y_k  (ConstructCurrentSolution, y passed by value)
struct Coef32 { double2 c[32]; };
void launch_update_solution(const double2 *B, int cnt, const Coef32 &y, double2 *x, int64_t n, hipStream_t s);
void launch_update_solution(const float2 *B, int cnt, const Coef32 &y, float2 *x, int64_t n, hipStream_t s);
void launch_update_solution(const float2 *B, int cnt, const Coef32 &y, double2 *x, int64_t n, hipStream_t s);

}  // namespace xtr
