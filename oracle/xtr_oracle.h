/*
 * xtr_oracle.h -- FP64 CPU restatement of extrEMe's integral-equation operator,
 * its Green-tensor assembly and the FGMRES solve around it.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * path: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it. The product never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the extrEMe checkout). All arithmetic is complex128, as in the reference.
 */
#ifndef XTR_ORACLE_H
#define XTR_ORACLE_H

#include <complex.h>
#include <stdint.h>

typedef double _Complex xo_c;

/* Background 1-D section: layer 0 is the upper (air) half-space, the last
 * layer the lower half-space; half-spaces carry thickness 0
 * (Extreme.Core/Model/Section1D.cs, EndUser/ExtremeMt/Examples/COMMEMI3D2/commemi.xmodel:4-10). */
typedef struct {
    int nlayers;
    double zero_level;        /* depth of the top of layer 0 */
    const double *thickness;  /* [nlayers] */
    const double *sigma;      /* [nlayers] S/m */
} xo_section;

/* Anomalous domain: Nx*Ny*Nz cells, layer k spans [depth[k], depth[k]+thick[k]). */
typedef struct {
    int nx, ny, nz;
    double dx, dy;
    const double *depth;      /* [nz] */
    const double *thick;      /* [nz] */
} xo_grid;

/* ---- model (P1) -------------------------------------------------------- */
double xo_mu0(void);
double xo_eps0(void);
/* zeta = sigma - i*omega*eps0  (Model/Omega/OmegaModelUtils.cs:17-21,41-42) */
xo_c xo_zeta(double sigma, double omega);
/* ModelUtils.FindCorrespondingBackgroundLayerIndex (Model/ModelUtils.cs:35-52); -1 if none */
int xo_background_layer(const xo_section *s, double start, double end);
/* ScalarMathUtils.DetermineCorrespondingBackgroundLayer, half-open (Green/Scalar/Impl/ScalarMathUtils.cs:15-31) */
int xo_layer_at_depth(const xo_section *s, double depth);

/* ---- plane wave (P2): PlaneWaveCalculator.CalculateFieldE/H ------------- */
xo_c xo_plane_wave_e(const xo_section *s, double omega, double depth, double thick);
xo_c xo_plane_wave_h(const xo_section *s, double omega, double depth);

/* ---- FFT (A4/A6/G13) ---------------------------------------------------- */
/* In-place unnormalised 2-D DFT over (n1, n2) of each of nc interleaved
 * columns of buf[n1][n2][nc]; sign -1 forward, +1 backward
 * (Fft/FftW/FftWTransform.cs:20-27, FftWrap/FftWrap/Direction.cs:10-14). */
void xo_fft2_batch(xo_c *buf, int n1, int n2, int64_t nc, int sign);
/* 1-D in-place DFT of length n (n = 2^a 3^b), stride 1. */
void xo_fft1(xo_c *x, int n, int sign);

/* ---- Green tensor (G1-G13) --------------------------------------------- */
typedef struct {
    int nitems;           /* NumberOfHankels */
    int *nrho;            /* [nitems] */
    double **rho;         /* [nitems][nrho] */
    int *nlam;            /* [nitems] = 431 + nrho */
    double **lam;         /* [nitems][nlam] */
    int nrho_merged;
    double *rho_merged;   /* sorted unique */
    int *merge_item;      /* [nrho_merged] first-wins source item */
    int *merge_index;     /* [nrho_merged] index inside that item */
    int zero_rho;         /* rho_merged[0] = 0, merge_item[0] = -1 (ScalarPlan.CalculateZeroRho) */
} xo_plan;

/* ScalarPlansCreater.CreateForAnomalyToAnomaly + ScalarPlan.CalculateLambdasForLog10
 * + InnerResultsCombiner.CreateFromSample (Green/Scalar/ScalarPlansCreater.cs:102-176,
 * Green/Scalar/ScalarPlan.cs:76-97, Green/Scalar/Impl/InnerResultsCombiner.cs:19-33). */
int xo_plan_create(const xo_grid *g, int number_of_hankels, xo_plan *plan);
/* ScalarPlansCreater.CreateFor (:102-110) for an explicit rho range; zero_rho adds rho = 0 in front
 * of the merged radii (ScalarPlan.GetSortedRho :56-65). */
int xo_plan_create_range(int number_of_hankels, double rho_min, double rho_max, int zero_rho, xo_plan *plan);
/* rho range of one observation level shifted by (sx, sy) metres (ScalarPlansCreater.cs:51-66,
 * 180-266): returns 1 when the minimum is zero (then *rho_min = 0.001 min(dx, dy) and the plan
 * takes the rho = 0 point), else 0. */
int xo_atoo_level_rho_range(const xo_grid *g, double sx, double sy, double *rho_min, double *rho_max);
void xo_plan_free(xo_plan *plan);

/* Green scalars of the AtoA tensor on the merged rho grid.
 * asym = 1: pairs in Combine order (rc-major, nz*nz), component I3 only -> out[0][pair][rho].
 * asym = 0: pairs in CombineSymm order (tr-major, rc<=tr, nz(nz+1)/2), components
 *           I1, I2, I5 -> out[0..2][pair][rho].
 * (GreenScalarCalculator.Calculate, PlanCalculatorFast.CalculateForE,
 *  VolumeToVolumeCalculatorUFast, ScalarMathUtils.NormalizeAndPerformHankelFast) */
int xo_green_scalars(const xo_section *s, const xo_grid *g, double omega,
                     const double *hank0, const double *hank1, const xo_plan *plan,
                     int asym, xo_c *out);

/* Filter alone (test hooks): xo_hankel_lambdas fills lam[0 .. 431 + nrho - 1] for the output
 * grid rho[i] = rho0 * step^i, step = xo_log10_rho_step() (ScalarPlan.cs:76-97,
 * HankelCoefficients.cs:67), and returns the count; xo_hankel_filter gives
 * out[i] = sum_j hank[430 - j] fk[i + j] (HankelCoefficients.ConvoluteWithFast :103-118), so that
 * out[i] / rho[i] approximates the integral of fk(lambda) J(lambda rho[i]) over lambda. */
double xo_log10_rho_step(void);
int xo_hankel_lambdas(double rho0, int nrho, double *lam);
void xo_hankel_filter(const double *hank, const xo_c *fk, int nrho, xo_c *out);

/* Whole AtoA tensor after mirroring and forward FFT, reference layout:
 * each component [2Nx][2Ny][P] with P = nz*nz (xz, yz) or nz(nz+1)/2 (xx, xy, yy, zz).
 * fft = 0 stops before the FFT (the mirrored spatial tensor). */
int xo_green_atoa(const xo_section *s, const xo_grid *g, double omega, int number_of_hankels,
                  const double *hank0, const double *hank1, int fft,
                  xo_c *xx, xo_c *xy, xo_c *xz, xo_c *yy, xo_c *yz, xo_c *zz);

/* AtoO level tensors (P7): anomaly layers to the observation level at depth zl whose sites sit at
 * the cell centres shifted by (sx, sy) metres, with the level's own scalar plan
 * (AtoOCalculator.cs:124-160). electric = 1: components xx yy zz xy - xz yz zx zy; electric = 0
 * (magnetic): xx - - xy yx xz yz zx zy (AtoOLevelGreenTensorCalculator.cs:21-75). out[9] in the
 * order xx yy zz xy yx xz yz zx zy, each [2Nx][2Ny][nz] (NULL for an absent component); fft = 0
 * stops before the FFT. */
int xo_green_atoo_level(const xo_section *s, const xo_grid *g, double omega, int number_of_hankels,
                        const double *hank0, const double *hank1, double zl, double sx, double sy, int electric,
                        int fft, xo_c **out);
/* the level's scalars I1..I5 (electric) or I1..I4 (magnetic) of every anomaly layer on the merged
 * rho grid of the plan (rho_min, rho_max, zero_rho): out[comp][layer][rho], rho_out[rho]; returns the
 * number of radii (or minus the number needed when it exceeds nrho_cap, -1 on error). */
int xo_atoo_level_scalars(const xo_section *s, const xo_grid *g, double omega, int number_of_hankels,
                          const double *hank0, const double *hank1, double zl, double rho_min, double rho_max,
                          int zero_rho, int electric, xo_c *out, int nrho_cap, double *rho_out);
/* the same with the scalar plan's rho range given (rho_min, rho_max, zero_rho) */
int xo_green_atoo_level_plan(const xo_section *s, const xo_grid *g, double omega, int number_of_hankels,
                             const double *hank0, const double *hank1, double zl, double sx, double sy,
                             double rho_min, double rho_max, int zero_rho, int electric, int fft, xo_c **out);

/* Anomalous field at the level from jQ [Nx][Ny][3][Nz] (ToOCalculator.CalculateAnomalyField,
 * ToOCalculator.cs:21-45): field [Nx][Ny][3]. */
void xo_atoo_field(int nx, int ny, int nz, xo_c **G, int electric, const xo_c *jq, xo_c *field);

/* ---- operator (A1-A9) --------------------------------------------------- */
typedef struct {
    int nx, ny, nz;
    const xo_c *xx, *xy, *xz, *yy, *yz, *zz;  /* reference layout (above) */
    xo_c *r;        /* [nx][ny][nz] */
    xo_c *f;        /* [nz] forward factors */
    xo_c *b;        /* [nz] backward factors */
    int type;       /* 0 = A, 1 = Chi0 */
    xo_c *buf1, *buf2;  /* [2nx][2ny][3nz] */
    xo_c *rx;       /* [n] */
} xo_op;

int xo_op_init(xo_op *op, int nx, int ny, int nz,
               const xo_c *xx, const xo_c *xy, const xo_c *xz,
               const xo_c *yy, const xo_c *yz, const xo_c *zz);
/* ConvolutionOperator.PrepareOperator (Forward/AtoA/ConvolutionOperator.cs:64-143) */
void xo_op_prepare(xo_op *op, int type, const xo_c *zeta /*[nx][ny][nz]*/,
                   const xo_c *zeta0 /*[nz]*/, const double *dz /*[nz]*/);
/* ConvolutionOperator.Apply (Forward/AtoA/ConvolutionOperator.cs:45-62,162-327) */
void xo_op_apply(xo_op *op, const xo_c *x, xo_c *y);
void xo_op_free(xo_op *op);
/* The contraction alone (ApplyGreenTensorAlongZ, ConvolutionOperator.cs:274-327)
 * over `length` wavenumbers: dst[l] = G(l) * src[l]. */
void xo_contract(int nz, int64_t length, const xo_c *xx, const xo_c *xy, const xo_c *xz,
                 const xo_c *yy, const xo_c *yz, const xo_c *zz, const xo_c *src, xo_c *dst);

/* ---- FGMRES (K1-K13) ---------------------------------------------------- */
typedef void (*xo_matvec_fn)(void *ctx, const xo_c *x, xo_c *y);
typedef void (*xo_iter_fn)(void *ctx, int outer, int iteration, double bea);

typedef struct {
    int64_t n;
    int m_outer, m_inner, max_repeat;
    double tol;
    int backward_check;  /* 0 Arnoldi only, 1 warning, 2 error */
} xo_gmres_params;

typedef struct {
    int converged;
    double bea;          /* Arnoldi backward error of the last cycle */
    double be;           /* true backward error, -1 if not computed */
    int iterations;      /* outer iterations (FGMRES counter) */
    int inner_iterations;
    int64_t matvecs;
    int64_t dots;
} xo_result;

/* FlexibleGmresWithGmresPreconditioner.Solve(rhs, x0, result) with x0 aliased to
 * result as AnomalyCurrentFgmresSolver.RunMainPart does (:73-85). x holds x0 on entry. */
int xo_fgmres_solve(const xo_gmres_params *p, xo_matvec_fn mv, void *mv_ctx,
                    xo_iter_fn it, void *it_ctx, const xo_c *b, xo_c *x, xo_result *res);
/* ClassicGmresSolver with the inner parameters only (one cycle, x0 = rhs). */
int xo_gmres_inner(const xo_gmres_params *p, xo_matvec_fn mv, void *mv_ctx,
                   const xo_c *v, xo_c *z, xo_result *res);

/* Reference-BLAS complex Givens generation (zrotg, green_tensor.c:21-32). */
void xo_zrotg(xo_c a, xo_c b, double *c, xo_c *s);

/* ---- forward pipeline pieces (P3-P6) ------------------------------------ */
/* E0 per anomaly layer (Mt3DForwardSolver.CalculateNormalFieldFromSource :251-275) */
void xo_normal_field(const xo_section *s, const xo_grid *g, double omega, xo_c *e0 /*[nz]*/);
/* j_s = (zeta - zeta0) E0 for polarization pol (0 = X, 1 = Y) (ForwardSolver.cs:186-237) */
void xo_jscattered(const xo_grid *g, int pol, const xo_c *zeta, const xo_c *zeta0,
                   const xo_c *e0, xo_c *js);
/* E_s = (2 sqrt(Re zeta0) chi - j_s)/(zeta + conj zeta0); jQ = E_s (zeta - zeta0) + j_s
 * (ForwardSolver.cs:319-449) */
void xo_escattered_jq(const xo_grid *g, const xo_c *zeta, const xo_c *zeta0,
                      const xo_c *chi, const xo_c *js, xo_c *es, xo_c *jq);

void xo_set_threads(int n);

#endif
