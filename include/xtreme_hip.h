/*
 * xtreme_hip.h -- C ABI of libxtreme_hip.so, the MI355X (gfx950) engine for extrEMe's
 * integral-equation operator apply and the FGMRES solve around it.
 *
 * Plain C: pointers, sizes, int status codes (0 = OK). No C++ types cross this boundary and
 * no exception unwinds through it; on failure the message is kept per thread and returned by
 * xtr_last_error().
 *
 * Complex numbers are interleaved {double re, im} (the reference's complex16,
 * Native/Native.Matrix/algebra.h:29, = System.Numerics.Complex).
 *
 * Vector layout is AnomalyCurrent's: index ((i*Ny + j)*3 + c)*Nz + k, i.e. [Nx][Ny][3][Nz]
 * (Extreme.Cartesian/Core/LayerAccessor.cs:46-47,82-86).
 *
 * Thread model: one host thread drives one handle at a time; host-pointer calls return after
 * the handle's stream has finished the work. xtr_fgmres_solve_batch runs several right-hand
 * sides on their own host threads and streams internally.
 */
#ifndef XTREME_HIP_H
#define XTREME_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } xtr_complex;

typedef struct xtr_op xtr_op;
typedef struct xtr_net xtr_net;   /* learned right preconditioner (see the end of this header) */

/* Arithmetic of the operator apply. */
enum {
    XTR_FP64 = 0,   /* complex128 Green tensor, FFT and contraction (the reference's arithmetic) */
    XTR_FP32 = 1,   /* complex64 Green tensor, FFT and contraction; Krylov vectors stay FP64 */
    XTR_MIXED = 2   /* both tensors kept: FP64 applies by default, complex64 applies inside
                       xtr_refined_solve_batch_device's correction solves */
};

/* Operator type (Extreme.Cartesian/Forward/AtoA/OperatorType.cs). */
enum { XTR_OP_A = 0, XTR_OP_CHI0 = 1 };

/* Model and discretisation: the fields of OmegaModel that the operator and the AtoA Green
 * tensor read (Extreme.Cartesian/Model/Omega/OmegaModel.cs, Extreme.Core/Model/Section1D.cs,
 * LateralDimensions.cs, AnomalyLayer.cs). Arrays are copied by xtr_op_create. */
typedef struct {
    int nx, ny, nz;
    double dx, dy;
    const double *depth;          /* [nz] top of each anomaly layer, m */
    const double *thick;          /* [nz] thickness of each anomaly layer, m */
    int nlayers;                  /* background layers incl. both half-spaces */
    double zero_level;            /* depth of the top of layer 0 */
    const double *layer_thickness;/* [nlayers], 0 for the half-spaces */
    const double *layer_sigma;    /* [nlayers], S/m */
    double omega;                 /* 2 pi f */
    int number_of_hankels;        /* ForwardSettings.NumberOfHankels (> 0) */
    const double *hank0;          /* [ntaps] J0 filter weights (HankelCoefficients.LoadN40) */
    const double *hank1;          /* [ntaps] J1 filter weights */
    int ntaps;                    /* 431 */
    int precision;                /* XTR_FP64 or XTR_FP32 */
    int max_rhs;                  /* right-hand sides one apply may carry (>= 1) */
} xtr_op_desc;

/* GmresParams as AnomalyCurrentFgmresSolver fills it (Extreme.Cartesian/Forward/
 * AnomalyCurrentFgmresSolver.cs:40-50, Extreme.Fgmres/GmresParams.cs:4-23). */
typedef struct {
    int outer_buffer_length;      /* FGMRES restart m (OuterBufferLength) */
    int inner_buffer_length;      /* inner GMRES m (InnerBufferLength) */
    int max_repeat_number;        /* MaxRepeatsNumber */
    double tolerance;             /* Residual */
    int backward_error_checking;  /* 0 Arnoldi only, 1 warning, 2 error (BackwardErrorChecking) */
} xtr_gmres_params;

/* ResultInfo (Extreme.Fgmres/ResultInfo.cs:4-36) plus operation counts. */
typedef struct {
    int converged;
    double arnoldi_backward_error;   /* bea */
    double backward_error;           /* be = |b - A x| / |b|, -1 when not computed */
    int number_of_iterations;        /* outer FGMRES iterations */
    int64_t matvecs;                 /* operator applies */
    int64_t dot_products;
    double seconds;                  /* wall time of the solve */
} xtr_result_info;

/* IterationComplete (Extreme.Fgmres/Gmres.cs:64-66, AnomalyCurrentFgmresSolver.cs:132-139):
 * rhs = index of the right-hand side in a batch, iteration = outer counter, bea. */
typedef void (*xtr_iteration_fn)(void *user, int rhs, int iteration, double bea);

/* Replaces: ConvolutionOperator ctor + FftBuffersPool.PrepareBuffersForModel
 * (Extreme.Cartesian/Forward/AtoA/ConvolutionOperator.cs:33-43, Fft/FftBufferPool.cs:64-96),
 * the GIEM2G-style handle pattern of Forward/AtoA/GIEM2G/Giem2gGreenTensor.cs:272-298. */
int xtr_op_create(const xtr_op_desc *desc, xtr_op **op);

/* Replaces: AtoAGreenTensorCalculatorComponent.CalculateGreenTensor
 * (Extreme.Cartesian/Forward/AtoA/AtoAGreenTensorCalculatorComponent.cs:69-182):
 * the FP64 AtoA Green tensor assembled, mirrored and FFT'd on the device. */
int xtr_op_build_green(xtr_op *op);

/* Replaces: handing a precomputed GreenTensor to ConvolutionOperator.PrepareOperator
 * (ConvolutionOperator.cs:64-68; TensorCache reuse, Forward/TensorCache.cs:8-35).
 * Host arrays in the reference layout: each component [2Nx*2Ny][P], P = Nz*Nz for xz, yz and
 * Nz(Nz+1)/2 (packed upper, element (rc,tr) at rc + tr(tr+1)/2) for xx, xy, yy, zz
 * (Green/Tensor/GreenTensor.cs:77-99, Model/TranceiverUtils.cs:10-16,55-74). */
int xtr_op_set_green(xtr_op *op, const xtr_complex *xx, const xtr_complex *xy, const xtr_complex *xz,
                     const xtr_complex *yy, const xtr_complex *yz, const xtr_complex *zz);

/* Copies the device Green tensor back in the reference layout above (test and export hook;
 * the reference's GreenTensor component pointers, GreenTensor.cs:128-139). */
int xtr_op_get_green(xtr_op *op, xtr_complex *xx, xtr_complex *xy, xtr_complex *xz,
                     xtr_complex *yy, xtr_complex *yz, xtr_complex *zz);

/* Replaces: ConvolutionOperator.PrepareOperator (ConvolutionOperator.cs:64-143):
 * R, forward and backward factors for type XTR_OP_A or XTR_OP_CHI0.
 * zeta: host [Nx][Ny][Nz] complex conductivity of the anomaly (OmegaAnomaly.Zeta). */
int xtr_op_prepare(xtr_op *op, int type, const xtr_complex *zeta);

/* Replaces: ConvolutionOperator.PrepareOperator (ConvolutionOperator.cs:64-143) for nmodels
 * conductivity models that share the grid, the background and the frequency, hence the Green
 * tensor and the forward/backward factors: zeta[m] is model m's [Nx][Ny][Nz] (the reference
 * prepares one model per operator; Mt3DForwardSolver.cs:78-112 reuses its Green tensor through
 * TensorCache). xtr_op_prepare is the nmodels = 1 case. */
int xtr_op_prepare_models(xtr_op *op, int type, int nmodels, const xtr_complex *const *zeta);

/* One apply of nrhs device vectors x[r] -> y[r], right-hand side r on model model[r]: the Green
 * tensor is streamed once for all of them. No y[r] may overlap any x[q]. */
int xtr_op_apply_models_device(xtr_op *op, int nrhs, const int *model, const void *const *x, void *const *y);

/* Replaces: ConvolutionOperator.Apply (ConvolutionOperator.cs:45-62,162-327).
 * x, y: host arrays of nrhs consecutive vectors of n = 3*Nx*Ny*Nz complex values. */
int xtr_op_apply(xtr_op *op, const xtr_complex *x, xtr_complex *y, int nrhs);

/* Same on device-resident vectors (HIP device pointers), ordered on the handle's stream
 * and complete on return. y must not overlap x (an error otherwise). */
int xtr_op_apply_device(xtr_op *op, const void *x, void *y, int nrhs);

/* The apply in complex64 on complex64 vectors (interleaved {float re, im}, nrhs consecutive device
 * vectors of n values): the FP32 Green tensor, FFTs and contraction, with R x and the last step
 * formed in FP64 and rounded once on store. Needs a handle created with XTR_FP32 or XTR_MIXED. No
 * counterpart in the reference (complex128 throughout); the operator of the mixed-precision
 * correction solves, for callers that run their own. */
int xtr_op_apply_c64_device(xtr_op *op, const void *x, void *y, int nrhs);

/* Replaces: AnomalyCurrentFgmresSolver.Solve -> FlexibleGmresWithGmresPreconditioner.Solve
 * (Extreme.Cartesian/Forward/AnomalyCurrentFgmresSolver.cs:59-85,
 *  Extreme.Fgmres/FlexibleGmresWithGmresPreconditioner.cs:29-63): FGMRES with an inner GMRES
 * right preconditioner, iterated classical Gram-Schmidt, true residual at restarts.
 * The operator must be prepared as XTR_OP_A. x holds the initial guess on entry (the
 * reference passes x0 = chi0, ForwardSolver.cs:300-307) and the solution on return.
 * Host pointers. */
int xtr_fgmres_solve(xtr_op *op, const xtr_gmres_params *params, const xtr_complex *b, xtr_complex *x,
                     xtr_result_info *info, xtr_iteration_fn on_iteration, void *user);

/* Same with device pointers. */
int xtr_fgmres_solve_device(xtr_op *op, const xtr_gmres_params *params, const void *b, void *x,
                            xtr_result_info *info, xtr_iteration_fn on_iteration, void *user);

/* nrhs independent solves with the same operator (the two polarisations of
 * Mt3DForwardSolver.SolvePrivate, Extreme.Cartesian.Magnetotellurics/Solver/Mt3DForwardSolver.cs:95-129),
 * run side by side so that every operator apply streams the Green tensor once for all of them.
 * Each right-hand side follows exactly the iteration of a separate xtr_fgmres_solve.
 * b[r], x[r]: device pointers. info: [nrhs]. */
int xtr_fgmres_solve_batch_device(xtr_op *op, const xtr_gmres_params *params, int nrhs,
                                  const void *const *b, void *const *x, xtr_result_info *info,
                                  xtr_iteration_fn on_iteration, void *user);

/* xtr_fgmres_solve_batch_device for right-hand sides of different models sharing the Green
 * tensor (prepared by xtr_op_prepare_models as XTR_OP_A): right-hand side r solves model model[r]'s
 * system; every apply serves all running right-hand sides (the batch scheduler's grouping of
 * (model, polarisation) pairs, Mt3DForwardSolver.cs:78-129 per model). */
int xtr_fgmres_solve_models_device(xtr_op *op, const xtr_gmres_params *params, int nrhs, const int *model,
                                   const void *const *b, void *const *x, xtr_result_info *info,
                                   xtr_iteration_fn on_iteration, void *user);

/* Mixed-precision solve: FP64 iterative refinement around complex64 FGMRES correction solves.
 * No counterpart in the reference (its arithmetic is complex128 throughout); it solves the same
 * system A x = b as xtr_fgmres_solve_batch_device and stops on the FP64 true residual. */
typedef struct {
    double tolerance;             /* stop when |b - A x| / |b| <= tolerance, A applied in FP64 */
    double inner_floor;           /* smallest FGMRES tolerance asked of a correction solve */
    int max_refinements;          /* correction solves per right-hand side (>= 1) */
    xtr_gmres_params inner;       /* FGMRES(m) / GMRES(m_in) settings of the correction solves;
                                     inner.tolerance is ignored (set per refinement) */
    int vector_precision;         /* Krylov vectors of the correction solves: XTR_FP64 (complex128)
                                     or XTR_FP32 (complex64, the apply reading and writing complex64;
                                     dots, norms, Hessenberg and rotations stay FP64) */
} xtr_refine_params;

typedef struct {
    int converged;
    double residual;              /* final FP64 true relative residual */
    int refinements;              /* correction solves run */
    int64_t matvecs64;            /* FP64 applies (true residuals) */
    int64_t matvecs32;            /* complex64 applies (inside the correction solves) */
    int inner_iterations;         /* outer FGMRES iterations summed over the correction solves */
    double seconds;
} xtr_refine_info;

/* Solves A x[r] = b[r] for nrhs right-hand sides side by side, x holding the initial guess. The
 * operator must be created with XTR_MIXED and prepared as XTR_OP_A. Each refinement computes
 * r = b - A x with the FP64 tensor (one batched apply), solves A d = r by FGMRES with the complex64
 * tensor to max(inner_floor, tolerance |b| / (2 |r|)), and adds d to x.
 * Replaces, for callers that accept it: AnomalyCurrentFgmresSolver.Solve
 * (Extreme.Cartesian/Forward/AnomalyCurrentFgmresSolver.cs:59-85) for both polarisations. */
int xtr_refined_solve_batch_device(xtr_op *op, const xtr_refine_params *params, int nrhs,
                                   const void *const *b, void *const *x, xtr_refine_info *info);

/* xtr_refined_solve_batch_device for right-hand sides of different models sharing the Green tensor
 * (prepared by xtr_op_prepare_models as XTR_OP_A): right-hand side r solves model model[r]'s system,
 * its FP64 residuals and complex64 corrections using that model's R. No counterpart in the
 * reference (complex128 throughout, one model at a time, Mt3DForwardSolver.cs:78-129). */
int xtr_refined_solve_models_device(xtr_op *op, const xtr_refine_params *params, int nrhs, const int *model,
                                    const void *const *b, void *const *x, xtr_refine_info *info);

/* Device time of the operator applies since xtr_op_set_profiling(op, 1), summed by stage, in
 * milliseconds: out[0] forward pass along y (with R and f fused), out[1] forward pass along x,
 * out[2] contraction, out[3] inverse pass along x, out[4] inverse pass along y (with b and the
 * last step fused); *applies = number of applies summed. Recorded with HIP events on the stream
 * the kernels run on (the reference's OperatorA* profiler events, ProfilerEvent.cs:40-49). */
int xtr_op_set_profiling(xtr_op *op, int enable);
int xtr_op_stage_times(xtr_op *op, double *out5, int64_t *applies);
/* The same summed over the applies of exactly nrhs right-hand sides (nrhs = 0: all applies, as
 * xtr_op_stage_times). No counterpart in the reference, whose profiler does not split by right-hand
 * sides; a launch's bytes depend on them. */
int xtr_op_stage_times_rhs(xtr_op *op, int nrhs, double *out5, int64_t *applies);

/* No counterpart in the reference (its FFTs are FFTW plans): the column tile each of the four FFT
 * passes (y forward, x forward, x inverse, y inverse) used in the latest apply, 0 = default,
 * 1 = narrow (half the columns), 2 = wide (twice the columns), 3 = narrow with 512 threads per
 * workgroup, 4 = register-resident radix-16/8 stages (pencil lengths 64, 128, 256, 512), 5 = the
 * same with the other column count, -1 before the first apply. The tile is measured once per
 * process for each precision and grid, on the first apply of that shape (a few milliseconds at
 * 256 x 256 x 64); the environment variable XTR_FFT_TILES (four letters d, n, w, m, r, s) fixes it
 * instead. Tiles 0-3 change only how columns are grouped into workgroups; tiles 4 and 5 also
 * change the radix plan, so results move by rounding only. */
int xtr_op_fft_tiles(xtr_op *op, int *tiles4);

/* ---- observation level (AtoO) ----------------------------------------------------------------- */

/* Replaces: AtoOCalculator.CalculateGreenTensorAtoOLevelElectric / ...Magnetic
 * (Extreme.Cartesian/Forward/ToObs/AtoOCalculator.cs:124-160): the FP64 level tensor of the
 * ObservationLevel at depth level_z whose sites are the cell centres shifted by (x_shift, y_shift)
 * metres (ObservationLevel.ShiftAlongX/Y), with the level's own scalar plan, assembled and FFT'd on
 * the device and kept in the handle. electric = 1: E (xx yy zz xy xz yz zx zy); 0: H (xx xy yx xz
 * yz zx zy). */
int xtr_op_build_green_atoo(xtr_op *op, double level_z, double x_shift, double y_shift, int electric);

/* Copies a level tensor back: comps[9] in the order xx yy zz xy yx xz yz zx zy, each [2Nx*2Ny][Nz]
 * (the reference's AlongVertical layout, AtoOLevelGreenTensorCalculator.cs:52-57); NULL entries
 * and components the tensor lacks are skipped (test and export hook). */
int xtr_op_get_green_atoo(xtr_op *op, int electric, xtr_complex *const *comps);

/* Replaces: ForwardSolver.CalculateEScatteredFrom + CalculateJqFrom (ForwardSolver.cs:319-436):
 * jQ = E_s (zeta - zeta0) + j_s from chi and j_s (device vectors of n values). The operator must be
 * prepared as XTR_OP_A. */
int xtr_op_jq_device(xtr_op *op, const void *chi, const void *js, void *jq);
/* The same for model `model` of an operator prepared by xtr_op_prepare_models (its R); the
 * reference computes jQ per model (ForwardSolver.cs:364-389), one model per operator.
 * xtr_op_jq_device refuses an operator that holds more than one model. */
int xtr_op_jq_model_device(xtr_op *op, int model, const void *chi, const void *js, void *jq);

/* Replaces: AtoOCalculator.CalculateAnomalyFieldE/H(level, jQ, field) (AtoOCalculator.cs:63-99)
 * through ToOCalculator.CalculateAnomalyField (ToOCalculator.cs:21-45): the anomalous E (electric
 * = 1) or H at the level, field [Nx][Ny][3] on the device, from the tensor built above. */
int xtr_op_level_field_device(xtr_op *op, int electric, const void *jq, void *field);

/* xtr_op_level_field_device that allocates nothing, for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid allocation). No counterpart in
 * the reference. Its kernels take their scratch (the zero-padded spectrum of jQ and the level
 * spectrum, 4 Nx Ny * 3 (Nz + 1) complex128) from a 128 MiB array the library holds on each device
 * until the process ends; grids needing more are refused with an error (4 Nx Ny * 3 (Nz + 1) <=
 * 8388608: 128 x 128 x 32 fits, 256 x 256 x 64 does not). Every call on a device shares that array:
 * calls that overlap on one device race, so a caller must not run two at once on one GPU. */
int xtr_op_level_field_noalloc_device(xtr_op *op, int electric, const void *jq, void *field);

/* Sizes: n = 3 Nx Ny Nz, the Green tensor's device bytes, the spectrum buffer bytes. */
int xtr_op_sizes(xtr_op *op, int64_t *n, int64_t *green_bytes, int64_t *buffer_bytes);

/* Device storage of the AtoA Green tensor. No counterpart in the reference, which keeps all
 * 4 Nx Ny wavenumber blocks (GreenTensor.cs:77-99). The mirrored lateral tensor is even or odd in
 * x and y per component, so block (2Nx-kx, ky) equals block (kx, ky) with xy and xz negated and
 * block (kx, 2Ny-ky) equals it with xy and yz negated. A folded tensor keeps only the
 * (Nx+1)(Ny+1) blocks kx <= Nx, ky <= Ny, and the contraction applies the signs. xtr_op_build_green
 * always folds; xtr_op_set_green folds when every mirrored block is within 1e-12 (relative to the
 * component's largest element) of its source block, and otherwise keeps all blocks. The
 * environment variable XTR_GREEN_LAYOUT=full disables folding. folded = 1 when folded;
 * blocks = the number of blocks stored. */
int xtr_op_green_layout(xtr_op *op, int *folded, int64_t *blocks);

/* The mirror symmetry that folding relies on, measured on a handle holding every block (built or set
 * with XTR_GREEN_LAYOUT=full, or set from a tensor without the symmetry). No counterpart in the
 * reference, whose populator writes the mirrored blocks (GreenTensorFftPopulator.cs:88-148).
 * out12[2c] = max over wavenumbers l with kx > Nx or ky > Ny of |G_c(l) - s_c G_c(mirror source)|
 * (real and imaginary parts separately), out12[2c + 1] = max |G_c|, for c in device block order
 * xx, xy, yy, zz, xz, yz, with the signs s_c of the component's parities. */
int xtr_op_green_mirror_check(xtr_op *op, double *out12);

/* Replaces: ConvolutionOperator.Dispose / GreenTensor.Dispose (ConvolutionOperator.cs:338-344). */
void xtr_op_destroy(xtr_op *op);

/* Message of the last failed call on this thread; returns its length. */
int xtr_last_error(char *buf, int len);

/* Measurement aid with no counterpart in the reference: the copy path of the streaming contraction
 * kernels without their arithmetic. `blocks` blocks of read_kb KB (a positive multiple of 32) are
 * read from src (device memory of blocks * read_kb KB) into LDS with non-temporal LDS-DMA loads, one
 * 512-thread workgroup per CU, a ring of four 32 KB units; after each block write_kb KB are stored to
 * dst (blocks * write_kb KB). Launched on `stream` (a hipStream_t; null: the default stream) and not
 * waited for. Timing it gives the HBM rate the part sustains for the contraction's read/write mix. */
int xtr_probe_stream(const void *src, void *dst, int64_t blocks, int read_kb, int write_kb, void *stream);

/* ---- learned right preconditioner (no counterpart in the reference) --------------------------------
 * A two-layer 3-D convolution network on the residual field, z = s (v/s + N(v/s)) with s the RMS of
 * v, N = conv(16 -> 6) o relu o conv(8 -> 16) over 3 x 3 x 3 cell neighbourhoods of the anomaly grid,
 * input channels (Re, Im) x (x, y, z) and the cell's features (log10 sigma, log10 f; see
 * xtr_net_set_features), output channels (Re, Im) x (x, y, z); BF16 MFMA convolutions with FP32
 * accumulation, FP32 master weights
 * and Adam. It stands in for the inner GMRES of FlexibleGmresWithGmresPreconditioner
 * (Extreme.Fgmres/FlexibleGmresWithGmresPreconditioner.cs:29-63) through the flexible right
 * preconditioner hook (FlexibleGmresSolver.cs:24-29). For callers that train and use such a
 * network; the reference's inner GMRES stays the default. */

/* A network for an nx x ny x nz grid: layer 1 seeded (He-scaled normal values), layer 2 zero, so
 * that it starts as the identity preconditioner. */
int xtr_net_create(int nx, int ny, int nz, uint64_t seed, xtr_net **net);
void xtr_net_destroy(xtr_net *net);
/* linear = 1 drops the relu of the hidden layer (forward and gradient): the preconditioner is then
 * a fixed linear map z = v + N v, as the flexible FGMRES's Krylov relation prefers; 0 (the
 * default) keeps it. Weights are kept; set it before training. */
int xtr_net_set_linear(xtr_net *net, int linear);
/* Input channels 6 and 7 of every cell: feat[2*cell] and feat[2*cell + 1] (host, [nx][ny][nz][2];
 * log10 of the cell's conductivity and log10 of the frequency), used by every later apply and
 * training step until set again; null returns to zeros (the default). Set the features of the
 * model a sample comes from before training on it, and of the model being solved before a solve. */
int xtr_net_set_features(xtr_net *net, const float *feat);
/* FP32 master weights: w1[(t*8 + ci)*16 + co] (27 taps t, input channels ci < 8: six of the vector,
 * two features, 16 hidden channels co), w2[(t*16 + ci)*8 + co] (16 hidden, 8 outputs of which 6 are used).
 * Tap t is the offset (t/9 - 1, (t/3)%3 - 1, t%3 - 1) in (i, j, k); channel 2c + p is part p (Re, Im)
 * of component c (x, y, z). */
int xtr_net_set_weights(xtr_net *net, const float *w1, const float *w2);
int xtr_net_get_weights(xtr_net *net, float *w1, float *w2);
/* Spectral branch: a per-lateral-mode channel mixing on the zero-padded 2 nx x 2 ny spectrum added to
 * the network's output, for the non-local reach a 3 x 3 x 3 convolution lacks. For the modes kx, ky in
 * -K..K and the 3 nz complex channels q = c nz + k of the scaled residual u = v / s:
 *   U[m][q] = s0 sum_{i<nx, j<ny} u(i, j, q) e^{-2 pi i (kx i / 2nx + ky j / 2ny)}, s0 = 1 / sqrt(nx ny),
 *   Y[m] = (W0[m] + phi W1[m]) U[m],  N(u)(i, j, q) += s0 sum_m Y[m][q] e^{+2 pi i (kx i / 2nx + ky j / 2ny)},
 * with phi the frequency feature (log10 f, channel 7 of xtr_net_set_features; 0 without features),
 * mode m = (kx + K)(2K + 1) + (ky + K), W0, W1 complex 3nz x 3nz per mode (row = output channel). The
 * mixing runs on BF16 MFMA with FP32 accumulation (BF16 weights and BF16-rounded U), the transforms in
 * FP32. modes = K in 1..12 (0 removes the branch) needs nz divisible by 8 and K < nx, ny; the
 * weights start at zero. Weight layout: [2 (W0, W1)][NM][3nz][3nz][2 (re, im)] FP32, NM = (2K + 1)^2;
 * *floats = 2 * 2 NM (3nz)^2. The gradient is that of the last xtr_net_gradients_device or
 * xtr_net_train_device call (before its Adam step), for W1 already multiplied by phi. */
int xtr_net_set_spectral(xtr_net *net, int modes);
int xtr_net_spectral_size(xtr_net *net, int *modes, int64_t *floats);
int xtr_net_set_spectral_weights(xtr_net *net, const float *w);
int xtr_net_get_spectral_weights(xtr_net *net, float *w);
int xtr_net_spectral_gradient(xtr_net *net, float *g);
/* z = the preconditioner applied to v (device vectors of n = 3 nx ny nz complex values). */
int xtr_net_apply_device(xtr_net *net, const void *v, void *z);
/* Gradient of the mean over the samples of |N(v/s) - (z - v)/s|^2 / |(z - v)/s|^2 (relative
 * squared error of the correction) with respect to w1 and w2, into host arrays, and the loss. */
int xtr_net_gradients_device(xtr_net *net, int nsamples, const void *const *v, const void *const *z, float *g1,
                             float *g2, double *loss);
/* One Adam step (beta 0.9, 0.999, eps 1e-8, rate lr) on that gradient; returns the loss before it. */
int xtr_net_train_device(xtr_net *net, int nsamples, const void *const *v, const void *const *z, float lr,
                         double *loss);
/* Flexible FGMRES(OuterBufferLength) with the network as its right preconditioner in place of the
 * inner GMRES; convergence as xtr_fgmres_solve_device (Arnoldi estimate, then the FP64 true residual). */
int xtr_fgmres_solve_learned_device(xtr_op *op, const xtr_gmres_params *params, xtr_net *net, const void *b, void *x,
                                    xtr_result_info *info);
/* xtr_fgmres_solve_device that also copies the first `capacity` pairs (v_j, z_j) of the flexible
 * preconditioner (v_j the Krylov vector, z_j the inner GMRES's result) into v_out[k], z_out[k]
 * (device vectors of n values); *count = pairs written. The training data of the network. */
int xtr_fgmres_solve_harvest_device(xtr_op *op, const xtr_gmres_params *params, const void *b, void *x,
                                    xtr_result_info *info, int capacity, void *const *v_out, void *const *z_out,
                                    int *count);

#ifdef __cplusplus
}
#endif

#endif
