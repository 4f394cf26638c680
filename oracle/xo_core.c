/*
 * xo_core.c -- oracle: model constants, plane wave, CPU FFT, the IE operator
 * and FGMRES.  TEST INFRASTRUCTURE ONLY (see xtr_oracle.h).
 */
#include "xtr_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define XO_PI 3.14159265358979323846

void xo_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ======================================================================
 * Model (P1)
 * ====================================================================== */

/* Mu0 = 4*pi*1e-7 (Model/Omega/OmegaModelUtils.cs:12) */
double xo_mu0(void) { return 4.0 * XO_PI * 1.0E-07; }

/* Epsilon0 = 1/(Mu0 c^2), c = 299792458 (OmegaModelUtils.cs:13) */
double xo_eps0(void) { return 1.0 / (xo_mu0() * 299792458.0 * 299792458.0); }

/* zeta = sigma - i omega eps0 (OmegaModelUtils.cs:41-42) */
xo_c xo_zeta(double sigma, double omega) { return sigma - I * (omega * xo_eps0()); }

/* Depths and thicknesses are System.Decimal in the reference (Extreme.Core/Model/Layer1D.cs,
 * Model/ModelUtils.cs:28-52), so its layer lookups add and compare exactly. The same comparisons on
 * doubles can misjudge a layer boundary (2.2 + 99.9 > 102.1 in binary), so they run here on whole
 * micrometres: exact for every depth written with at most six decimals. */
static long long xo_um(double v) { return llround(v * 1e6); }

/* Model/ModelUtils.cs:35-52: first layer whose closed interval [ss, es] holds [start, end] */
int xo_background_layer(const xo_section *s, double start, double end)
{
    long long ss = xo_um(s->zero_level), a = xo_um(start), b = xo_um(end);
    for (int i = 0; i < s->nlayers; i++) {
        long long es = ss + xo_um(s->thickness[i]);
        if (a >= ss && b <= es) return i;
        ss = es;
    }
    return -1;
}

/* Green/Scalar/Impl/ScalarMathUtils.cs:15-31: half-open [d1, d2) */
int xo_layer_at_depth(const xo_section *s, double depth)
{
    long long prev = xo_um(s->zero_level), z = xo_um(depth);
    for (int k = 0; k < s->nlayers; k++) {
        long long d1 = prev, d2 = prev + xo_um(s->thickness[k]);
        if (z >= d1 && z < d2) return k;
        prev = d2;
    }
    return -1;
}

/* OmegaModel.GetLayerDepth (Model/Omega/OmegaModel.cs:27-38) */
static double layer_depth(const xo_section *s, int index)
{
    double d = s->zero_level;
    for (int i = 0; i < index; i++) d += s->thickness[i];
    return d;
}

/* ======================================================================
 * Plane wave (P2) at lambda = 0 with the non-fast per-lambda recursions
 * (Green/Scalar/Impl/PlanCalculator.cs:195-314,
 *  Extreme.Cartesian.Magnetotellurics/PlaneWaveCalculator.cs:11-83)
 * ====================================================================== */

typedef struct {
    int nl;
    xo_c *eta, *ex, *p, *q, *a; /* a is nl x nl */
} pw_aux;

static void pw_prepare(const xo_section *s, double omega, int rc, int tr, pw_aux *c)
{
    int nl = s->nlayers;
    double mu0 = xo_mu0();
    c->nl = nl;
    c->eta = calloc(nl, sizeof(xo_c));
    c->ex = calloc(nl, sizeof(xo_c));
    c->p = calloc(nl, sizeof(xo_c));
    c->q = calloc(nl, sizeof(xo_c));
    c->a = calloc((size_t)nl * nl, sizeof(xo_c));
    for (int k = 0; k < nl; k++) {
        xo_c zeta = xo_zeta(s->sigma[k], omega);
        c->eta[k] = csqrt(0.0 - (I * omega) * zeta * mu0);   /* PlanCalculator.cs:227-239 */
        c->ex[k] = cexp(-2.0 * c->eta[k] * s->thickness[k]);  /* :241-252 */
    }
    /* CalculateQ1 (:288-314), beta = 1 */
    int n = nl - 1;
    c->q[n] = 0;
    xo_c part = c->eta[n] / c->eta[n - 1];
    c->q[n - 1] = (1.0 - part) / (1.0 + part);
    for (int k = n - 1; k > 1; k--) {
        xo_c qe = c->q[k] * c->ex[k];
        xo_c be = c->eta[k] / c->eta[k - 1];
        part = be * ((qe - 1.0) / (qe + 1.0));
        c->q[k - 1] = (1.0 + part) / (1.0 - part);
    }
    /* CalculateP1 (:254-286), alpha = 1 */
    c->p[1] = (c->eta[1] - c->eta[0]) / (c->eta[1] + c->eta[0]);
    for (int k = 1; k < nl - 1; k++) {
        xo_c pe = c->p[k] * c->ex[k];
        xo_c ae = c->eta[k] / c->eta[k + 1];
        xo_c pt = ae * ((pe - 1.0) / (pe + 1.0));
        c->p[k + 1] = (1.0 + pt) / (1.0 - pt);
    }
    /* CalculateA (:195-225) */
    int smax = rc > tr ? rc : tr, rmin = rc < tr ? rc : tr;
    for (int nn = smax; nn >= rmin; nn--) {
        xo_c prev = 1.0 / (c->eta[nn] * (1.0 - c->p[nn] * c->q[nn] * c->ex[nn]));
        c->a[nn * nl + nn] = prev;
        for (int k = nn - 1; k >= rmin; k--) {
            double dk1 = layer_depth(s, k + 1);
            xo_c qq = ((1.0 + c->p[k + 1]) / (1.0 + c->p[k] * c->ex[k])) *
                      cexp((c->eta[k + 1] - c->eta[k]) * dk1);
            xo_c next = qq * prev;
            c->a[k * nl + nn] = next;
            c->a[nn * nl + k] = next;
            prev = next;
        }
    }
}

static void pw_free(pw_aux *c)
{
    free(c->eta); free(c->ex); free(c->p); free(c->q); free(c->a);
}

/* PlaneWaveCalculator.CalculateFieldE (:11-40); transmitter at depth 0 */
xo_c xo_plane_wave_e(const xo_section *s, double omega, double depth, double thick)
{
    double mu0 = xo_mu0();
    int tr = xo_layer_at_depth(s, 0.0);
    int rc = xo_layer_at_depth(s, depth);
    if (rc < 0 || tr < 0) return NAN;   /* the reference throws: not within the 1-D section */
    pw_aux c;
    pw_prepare(s, omega, rc, tr, &c);
    int nl = c.nl;
    xo_c kkr = I * omega * mu0 * xo_zeta(s->sigma[rc], omega);
    xo_c ikr = I * csqrt(kkr);
    xo_c kk1 = I * omega * mu0 * xo_zeta(s->sigma[tr], omega);
    xo_c ik1 = I * csqrt(kk1);
    double dr1 = layer_depth(s, rc + 1);
    double d2 = layer_depth(s, tr + 1);
    double zr = depth;
    xo_c a = c.a[rc * nl + tr] / c.a[tr * nl + tr];
    xo_c upper;
    if (thick == 0) {
        upper = cexp(ikr * zr) + c.q[rc] * cexp(ikr * (2 * dr1 - zr));
    } else {
        double z1 = zr - thick / 2, z2 = zr + thick / 2;
        xo_c u1 = cexp(ikr * z1) - c.q[rc] * cexp(ikr * (2 * dr1 - z1));
        xo_c u2 = cexp(ikr * z2) - c.q[rc] * cexp(ikr * (2 * dr1 - z2));
        upper = (u2 - u1) / (z2 - z1) / ikr;
    }
    xo_c lower = 1.0 + c.q[tr] * cexp(ik1 * (2 * d2));
    pw_free(&c);
    return a * (upper / lower);
}

/* PlaneWaveCalculator.CalculateFieldH (:42-64) */
xo_c xo_plane_wave_h(const xo_section *s, double omega, double depth)
{
    double mu0 = xo_mu0();
    int tr = xo_layer_at_depth(s, 0.0);
    int rc = xo_layer_at_depth(s, depth);
    if (rc < 0 || tr < 0) return NAN;   /* the reference throws: not within the 1-D section */
    pw_aux c;
    pw_prepare(s, omega, rc, tr, &c);
    int nl = c.nl;
    xo_c kkr = I * omega * mu0 * xo_zeta(s->sigma[rc], omega);
    xo_c ikr = I * csqrt(kkr);
    xo_c kk1 = I * omega * mu0 * xo_zeta(s->sigma[tr], omega);
    xo_c ik1 = I * csqrt(kk1);
    double dr1 = layer_depth(s, rc + 1);
    double d2 = layer_depth(s, tr + 1);
    xo_c a = c.a[rc * nl + tr] / c.a[tr * nl + tr];
    xo_c upper = cexp(ikr * depth) - c.q[rc] * cexp(ikr * (2 * dr1 - depth));
    xo_c lower = 1.0 + c.q[tr] * cexp(ik1 * (2 * d2));
    pw_free(&c);
    return (csqrt(kkr) / (omega * mu0)) * a * (upper / lower);
}

/* ======================================================================
 * FFT: mixed-radix (4, 2, 3, 5) Stockham, decimation in frequency, natural
 * order output.  Textbook DFT X[k] = sum_j x[j] exp(sign 2 pi i jk/n).
 * ====================================================================== */

static int fft_factor(int n, int *radix)
{
    int m = 0;
    while (n % 4 == 0) { radix[m++] = 4; n /= 4; }
    while (n % 2 == 0) { radix[m++] = 2; n /= 2; }
    while (n % 3 == 0) { radix[m++] = 3; n /= 3; }
    while (n % 5 == 0) { radix[m++] = 5; n /= 5; }
    return n == 1 ? m : -1;
}

static void fft1_work(xo_c *x, int n, int sign, xo_c *y, const xo_c *tw)
{
    /* tw[k] = exp(sign 2 pi i k / n) */
    int radix[40];
    int nr = fft_factor(n, radix);
    if (nr < 0) { fprintf(stderr, "xo_fft1: unsupported length %d\n", n); abort(); }
    xo_c *src = x, *dst = y;
    int len = n, s = 1;
    const xo_c w3 = cexp(sign * 2.0 * XO_PI * I / 3.0);
    xo_c w5[5]; /* exp(sign 2 pi i k / 5) */
    for (int k = 0; k < 5; k++) w5[k] = cexp(sign * 2.0 * XO_PI * I * k / 5.0);
    for (int st = 0; st < nr; st++) {
        int r = radix[st];
        int m = len / r;
        int tstep = n / len; /* twiddle stride: exp(sign 2 pi i p u / len) = tw[p u tstep] */
        for (int p = 0; p < m; p++) {
            for (int q = 0; q < s; q++) {
                if (r == 2) {
                    xo_c a = src[q + s * (p + 0)];
                    xo_c b = src[q + s * (p + m)];
                    dst[q + s * (2 * p + 0)] = a + b;
                    dst[q + s * (2 * p + 1)] = (a - b) * tw[(int64_t)p * tstep % n];
                } else if (r == 4) {
                    xo_c a0 = src[q + s * (p + 0 * m)];
                    xo_c a1 = src[q + s * (p + 1 * m)];
                    xo_c a2 = src[q + s * (p + 2 * m)];
                    xo_c a3 = src[q + s * (p + 3 * m)];
                    xo_c jj = sign * I; /* exp(sign i pi/2) */
                    xo_c t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = (a1 - a3) * jj;
                    xo_c y0 = t0 + t2, y1 = t1 + t3, y2 = t0 - t2, y3 = t1 - t3;
                    dst[q + s * (4 * p + 0)] = y0;
                    dst[q + s * (4 * p + 1)] = y1 * tw[(int64_t)p * 1 * tstep % n];
                    dst[q + s * (4 * p + 2)] = y2 * tw[(int64_t)p * 2 * tstep % n];
                    dst[q + s * (4 * p + 3)] = y3 * tw[(int64_t)p * 3 * tstep % n];
                } else if (r == 5) { /* y_u = sum_t a_t w5^(t u), textbook */
                    xo_c a[5];
                    for (int t = 0; t < 5; t++) a[t] = src[q + s * (p + t * m)];
                    for (int u = 0; u < 5; u++) {
                        xo_c yv = 0;
                        for (int t = 0; t < 5; t++) yv += a[t] * w5[(t * u) % 5];
                        dst[q + s * (5 * p + u)] = u == 0 ? yv : yv * tw[(int64_t)p * u * tstep % n];
                    }
                } else { /* r == 3 */
                    xo_c a0 = src[q + s * (p + 0 * m)];
                    xo_c a1 = src[q + s * (p + 1 * m)];
                    xo_c a2 = src[q + s * (p + 2 * m)];
                    xo_c y0 = a0 + a1 + a2;
                    xo_c y1 = a0 + a1 * w3 + a2 * w3 * w3;
                    xo_c y2 = a0 + a1 * w3 * w3 + a2 * w3;
                    dst[q + s * (3 * p + 0)] = y0;
                    dst[q + s * (3 * p + 1)] = y1 * tw[(int64_t)p * 1 * tstep % n];
                    dst[q + s * (3 * p + 2)] = y2 * tw[(int64_t)p * 2 * tstep % n];
                }
            }
        }
        len = m;
        s *= r;
        xo_c *t = src; src = dst; dst = t;
    }
    if (src != x) memcpy(x, src, sizeof(xo_c) * n);
}

static xo_c *make_twiddles(int n, int sign)
{
    xo_c *tw = malloc(sizeof(xo_c) * n);
    for (int k = 0; k < n; k++) {
        double a = 2.0 * XO_PI * (double)k / (double)n;
        tw[k] = cos(a) + I * (sign * sin(a));
    }
    return tw;
}

void xo_fft1(xo_c *x, int n, int sign)
{
    xo_c *y = malloc(sizeof(xo_c) * n);
    xo_c *tw = make_twiddles(n, sign);
    fft1_work(x, n, sign, y, tw);
    free(tw);
    free(y);
}

#define XO_FFT_CB 8

void xo_fft2_batch(xo_c *buf, int n1, int n2, int64_t nc, int sign)
{
    xo_c *tw1 = make_twiddles(n1, sign);
    xo_c *tw2 = make_twiddles(n2, sign);
    /* along n2 for each (i1, column) */
#pragma omp parallel
    {
        int nmax = n1 > n2 ? n1 : n2;
        xo_c *tmp = malloc(sizeof(xo_c) * nmax * XO_FFT_CB);
        xo_c *wk = malloc(sizeof(xo_c) * nmax);
#pragma omp for schedule(static)
        for (int64_t t = 0; t < (int64_t)n1 * ((nc + XO_FFT_CB - 1) / XO_FFT_CB); t++) {
            int64_t i1 = t / ((nc + XO_FFT_CB - 1) / XO_FFT_CB);
            int64_t c0 = (t % ((nc + XO_FFT_CB - 1) / XO_FFT_CB)) * XO_FFT_CB;
            int cb = (int)((nc - c0) < XO_FFT_CB ? (nc - c0) : XO_FFT_CB);
            xo_c *base = buf + i1 * n2 * nc + c0;
            for (int j = 0; j < n2; j++)
                for (int c = 0; c < cb; c++) tmp[c * n2 + j] = base[(int64_t)j * nc + c];
            for (int c = 0; c < cb; c++) fft1_work(tmp + c * n2, n2, sign, wk, tw2);
            for (int j = 0; j < n2; j++)
                for (int c = 0; c < cb; c++) base[(int64_t)j * nc + c] = tmp[c * n2 + j];
        }
        /* along n1 for each (i2, column) */
#pragma omp for schedule(static)
        for (int64_t t = 0; t < (int64_t)n2 * ((nc + XO_FFT_CB - 1) / XO_FFT_CB); t++) {
            int64_t i2 = t / ((nc + XO_FFT_CB - 1) / XO_FFT_CB);
            int64_t c0 = (t % ((nc + XO_FFT_CB - 1) / XO_FFT_CB)) * XO_FFT_CB;
            int cb = (int)((nc - c0) < XO_FFT_CB ? (nc - c0) : XO_FFT_CB);
            xo_c *base = buf + i2 * nc + c0;
            int64_t st = (int64_t)n2 * nc;
            for (int i = 0; i < n1; i++)
                for (int c = 0; c < cb; c++) tmp[c * n1 + i] = base[i * st + c];
            for (int c = 0; c < cb; c++) fft1_work(tmp + c * n1, n1, sign, wk, tw1);
            for (int i = 0; i < n1; i++)
                for (int c = 0; c < cb; c++) base[i * st + c] = tmp[c * n1 + i];
        }
        free(tmp);
        free(wk);
    }
    free(tw1);
    free(tw2);
}

/* ======================================================================
 * Operator (A1-A9)
 * ====================================================================== */

/* zspmv 'U' (full_zgemv.c:28-36): y = alpha A x + beta y, A symmetric packed
 * upper column-major, A(i,j) = ap[i + j(j+1)/2] for i <= j. */
static void spmv_u(int nz, xo_c alpha, xo_c beta, const xo_c *ap, const xo_c *x, xo_c *y)
{
    for (int i = 0; i < nz; i++) {
        xo_c s = 0;
        for (int j = 0; j < nz; j++) {
            int lo = i < j ? i : j, hi = i < j ? j : i;
            s += ap[lo + hi * (hi + 1) / 2] * x[j];
        }
        y[i] = (beta == 0 ? 0 : beta * y[i]) + alpha * s;
    }
}

/* cblas_zgemv RowMajor NoTrans (full_zgemv.c:16-20): y = alpha G x + beta y */
static void gemv_n(int nz, xo_c alpha, xo_c beta, const xo_c *g, const xo_c *x, xo_c *y)
{
    for (int r = 0; r < nz; r++) {
        xo_c s = 0;
        for (int c = 0; c < nz; c++) s += g[r * nz + c] * x[c];
        y[r] = (beta == 0 ? 0 : beta * y[r]) + alpha * s;
    }
}

/* cblas_zgemv RowMajor Trans (full_zgemv.c:10-14): y = alpha G^T x + beta y */
static void gemv_t(int nz, xo_c alpha, xo_c beta, const xo_c *g, const xo_c *x, xo_c *y)
{
    for (int c = 0; c < nz; c++) {
        xo_c s = 0;
        for (int r = 0; r < nz; r++) s += g[r * nz + c] * x[r];
        y[c] = (beta == 0 ? 0 : beta * y[c]) + alpha * s;
    }
}

/* ApplyGreenTensorAlongZ (ConvolutionOperator.cs:274-327) */
void xo_contract(int nz, int64_t length, const xo_c *xx, const xo_c *xy, const xo_c *xz,
                 const xo_c *yy, const xo_c *yz, const xo_c *zz, const xo_c *src, xo_c *dst)
{
    int64_t asymNz = (int64_t)nz * nz;
    int64_t symmNz = nz + (int64_t)nz * (nz - 1) / 2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < length; i++) {
        int64_t ds = i * 3 * nz, ss = i * symmNz, as = i * asymNz;
        xo_c *dX = dst + ds, *dY = dst + ds + nz, *dZ = dst + ds + 2 * nz;
        const xo_c *sX = src + ds, *sY = src + ds + nz, *sZ = src + ds + 2 * nz;
        spmv_u(nz, 1, 0, xx + ss, sX, dX);
        spmv_u(nz, 1, 1, xy + ss, sY, dX);
        gemv_n(nz, 1, 1, xz + as, sZ, dX);
        spmv_u(nz, 1, 0, xy + ss, sX, dY);
        spmv_u(nz, 1, 1, yy + ss, sY, dY);
        gemv_n(nz, 1, 1, yz + as, sZ, dY);
        gemv_t(nz, -1, 0, xz + as, sX, dZ);
        gemv_t(nz, -1, 1, yz + as, sY, dZ);
        spmv_u(nz, 1, 1, zz + ss, sZ, dZ);
    }
}

int xo_op_init(xo_op *op, int nx, int ny, int nz,
               const xo_c *xx, const xo_c *xy, const xo_c *xz,
               const xo_c *yy, const xo_c *yz, const xo_c *zz)
{
    memset(op, 0, sizeof(*op));
    op->nx = nx; op->ny = ny; op->nz = nz;
    op->xx = xx; op->xy = xy; op->xz = xz; op->yy = yy; op->yz = yz; op->zz = zz;
    int64_t m = (int64_t)4 * nx * ny * 3 * nz;
    int64_t n = (int64_t)3 * nx * ny * nz;
    op->r = calloc((size_t)nx * ny * nz, sizeof(xo_c));
    op->f = calloc(nz, sizeof(xo_c));
    op->b = calloc(nz, sizeof(xo_c));
    op->buf1 = calloc(m, sizeof(xo_c));
    op->buf2 = calloc(m, sizeof(xo_c));
    op->rx = calloc(n, sizeof(xo_c));
    return (op->r && op->buf1 && op->buf2 && op->rx) ? 0 : -1;
}

void xo_op_free(xo_op *op)
{
    free(op->r); free(op->f); free(op->b); free(op->buf1); free(op->buf2); free(op->rx);
    memset(op, 0, sizeof(*op));
}

/* PrepareOperator (ConvolutionOperator.cs:64-143) */
void xo_op_prepare(xo_op *op, int type, const xo_c *zeta, const xo_c *zeta0, const double *dz)
{
    int nx = op->nx, ny = op->ny, nz = op->nz;
    op->type = type;
    for (int k = 0; k < nz; k++) {
        op->f[k] = sqrt(creal(zeta0[k]));                           /* :90-97 */
        double nf = 1.0 / ((double)nx * (double)ny * 4);              /* :101 */
        xo_c sc = 2 * sqrt(creal(zeta0[k]));
        sc /= dz[k];
        sc *= nf;
        op->b[k] = sc;                                               /* :99-111 */
    }
    int64_t idx = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                xo_c z = zeta[idx];
                xo_c cz0 = conj(zeta0[k]);
                if (type == 1) op->r[idx] = csqrt(creal(zeta0[k])) / (z + cz0);   /* :138-139 */
                else op->r[idx] = (z - zeta0[k]) / (z + cz0);                    /* :140-141 */
                idx++;
            }
}

void xo_op_apply(xo_op *op, const xo_c *x, xo_c *y)
{
    int nx = op->nx, ny = op->ny, nz = op->nz;
    int64_t n = (int64_t)3 * nx * ny * nz;
    int64_t m = (int64_t)4 * nx * ny * 3 * nz;
    /* ApplyOperatorR (:233-248) */
#pragma omp parallel for schedule(static)
    for (int64_t ij = 0; ij < (int64_t)nx * ny; ij++)
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < nz; k++)
                op->rx[ij * 3 * nz + c * nz + k] = op->r[ij * nz + k] * x[ij * 3 * nz + c * nz + k];
    /* PrepareForForwardFft (:250-272) */
    memset(op->buf1, 0, sizeof(xo_c) * m);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            int64_t s = ((int64_t)i * ny + j) * 3 * nz;
            int64_t d = ((int64_t)i * ny * 2 + j) * 3 * nz;
            for (int c = 0; c < 3; c++)
                for (int k = 0; k < nz; k++)
                    op->buf1[d + c * nz + k] = op->f[k] * op->rx[s + c * nz + k];
        }
    /* PerformForwardFft (:206-207) */
    xo_fft2_batch(op->buf1, 2 * nx, 2 * ny, 3 * nz, -1);
    /* ApplyGreenTensorAlongZ (:274-327) */
    xo_contract(nz, (int64_t)4 * nx * ny, op->xx, op->xy, op->xz, op->yy, op->yz, op->zz,
                op->buf1, op->buf2);
    /* PerformBackwardFft (:203-204) */
    xo_fft2_batch(op->buf2, 2 * nx, 2 * ny, 3 * nz, +1);
    /* ExtractData (:209-231) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            int64_t d = ((int64_t)i * ny + j) * 3 * nz;
            int64_t s = ((int64_t)i * ny * 2 + j) * 3 * nz;
            for (int c = 0; c < 3; c++)
                for (int k = 0; k < nz; k++)
                    y[d + c * nz + k] = op->b[k] * op->buf2[s + c * nz + k];
        }
    /* PerformLastStep (:187-201) */
    if (op->type == 0) {
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < n; t++) {
            xo_c v = x[t] - y[t];
            y[t] = v - op->rx[t];
        }
    } else {
#pragma omp parallel for schedule(static)
        for (int64_t t = 0; t < n; t++) y[t] = y[t] + op->rx[t];
    }
}

/* ======================================================================
 * FGMRES (K1-K13): Extreme.Fgmres/Gmres.cs, FlexibleGmresSolver.cs,
 * ClassicGmresSolver.cs, FlexibleGmresWithGmresPreconditioner.cs
 * ====================================================================== */

/* Reference BLAS zrotg (the routine ntv_math's Zrotg wraps, green_tensor.c:21-32) */
void xo_zrotg(xo_c ca, xo_c cb, double *c, xo_c *s)
{
    double aca = cabs(ca);
    if (aca == 0.0) {
        *c = 0.0;
        *s = 1.0;
        return;
    }
    double scale = aca + cabs(cb);
    double t1 = cabs(ca / scale), t2 = cabs(cb / scale);
    double norm = scale * sqrt(t1 * t1 + t2 * t2);
    xo_c alpha = ca / aca;
    *c = aca / norm;
    *s = alpha * conj(cb) / norm;
}

/* zrot on one pair (green_tensor.c:21-28): x' = c x + s y ; y' = c y - conj(s) x */
static void zrot1(xo_c *x, xo_c *y, double c, xo_c s)
{
    xo_c t = c * (*x) + s * (*y);
    *y = c * (*y) - conj(s) * (*x);
    *x = t;
}

static xo_c zdotc(int64_t n, const xo_c *x, const xo_c *y)
{
    double re = 0, im = 0;
#pragma omp parallel for reduction(+ : re, im) schedule(static)
    for (int64_t i = 0; i < n; i++) {
        xo_c v = conj(x[i]) * y[i];
        re += creal(v);
        im += cimag(v);
    }
    return re + I * im;
}

typedef struct gm {
    int64_t n;
    int m;
    int flexible;
    int max_repeat;
    double tol;
    int backward_check;
    xo_c *V, *Z, *w, *r0, *xcur, *ycur, *H, *dots;
    double *gc;
    xo_c *gs;
    int iter_counter;
    int64_t *matvecs, *ndots;
    xo_matvec_fn mv;
    void *mv_ctx;
    xo_iter_fn it;
    void *it_ctx;
    struct gm *inner;  /* flexible: the right preconditioner */
} gm;

#define HIDX(s, i, j) ((int64_t)(j) * ((s)->m + 1) + (i))

static int gm_alloc(gm *s, int64_t n, int m, int flexible)
{
    memset(s, 0, sizeof(*s));
    s->n = n; s->m = m; s->flexible = flexible;
    s->V = calloc((size_t)n * m, sizeof(xo_c));  /* ComputeTheTrue: m columns (Gmres.cs:51-52) */
    s->Z = flexible ? calloc((size_t)n * m, sizeof(xo_c)) : NULL;
    s->w = calloc(n, sizeof(xo_c));
    s->r0 = calloc(n, sizeof(xo_c));
    s->xcur = calloc(n, sizeof(xo_c));
    s->ycur = calloc(m, sizeof(xo_c));
    s->H = calloc((size_t)(m + 1) * (m + 1), sizeof(xo_c));
    s->dots = calloc(m, sizeof(xo_c));
    s->gc = calloc(m, sizeof(double));
    s->gs = calloc(m, sizeof(xo_c));
    return (s->V && s->w && s->r0 && s->xcur && (!flexible || s->Z)) ? 0 : -1;
}

static void gm_free(gm *s)
{
    free(s->V); free(s->Z); free(s->w); free(s->r0); free(s->xcur); free(s->ycur);
    free(s->H); free(s->dots); free(s->gc); free(s->gs);
}

static void gm_matvec(gm *s, const xo_c *x, xo_c *y)
{
    s->mv(s->mv_ctx, x, y);
    (*s->matvecs)++;
}

static double gm_norm(gm *s, const xo_c *v)
{
    (*s->ndots)++;
    return sqrt(creal(zdotc(s->n, v, v)));
}

static int gm_solve(gm *s, const xo_c *rhs, const xo_c *x0, xo_c *result,
                    double *out_bea, double *out_be, int *out_conv);

/* ApplyPreconditionerRight: flexible -> inner GMRES on Z[:,jH]; classic -> copy */
static void gm_precond(gm *s, const xo_c *src, xo_c *dst, int jH)
{
    if (s->flexible) {
        xo_c *col = s->Z + (int64_t)jH * s->n;
        double bea, be;
        int conv;
        gm_solve(s->inner, src, src, col, &bea, &be, &conv);   /* FlexibleGmresWithGmresPreconditioner.cs:57 */
        memcpy(dst, col, sizeof(xo_c) * s->n);                 /* FlexibleGmresSolver.cs:24-29 */
    } else {
        memcpy(dst, src, sizeof(xo_c) * s->n);                  /* ClassicGmresSolver.cs:24-30 */
    }
}

/* ConstructCurrentSolution (Gmres.cs:187-196) */
static void gm_construct(gm *s, int jH, xo_c *result)
{
    int m = s->m;
    int64_t n = s->n;
    for (int i = 0; i < jH; i++) s->ycur[i] = s->H[HIDX(s, i, m)];
    /* SimplifiedZtrsv: upper, no-trans, non-unit, lda = m+1 (green_tensor.c:34-37) */
    for (int i = jH - 1; i >= 0; i--) {
        xo_c t = s->ycur[i];
        for (int k = i + 1; k < jH; k++) t -= s->H[HIDX(s, i, k)] * s->ycur[k];
        s->ycur[i] = t / s->H[HIDX(s, i, i)];
    }
    /* CalculateSolutionProjectionToKrylovSubspace: r0 = Z y (flexible) / V y (classic) */
    const xo_c *B = s->flexible ? s->Z : s->V;
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < n; t++) {
        xo_c a = 0;
        for (int k = 0; k < jH; k++) a += B[(int64_t)k * n + t] * s->ycur[k];
        s->r0[t] = a;
    }
    for (int64_t t = 0; t < n; t++) s->xcur[t] = result[t] + s->r0[t];
    memcpy(result, s->xcur, sizeof(xo_c) * n);
}

/* CheckBackwardError (Gmres.cs:200-230) */
static double gm_check(gm *s, double bn, const xo_c *rhs)
{
    memcpy(s->r0, s->xcur, sizeof(xo_c) * s->n);
    gm_matvec(s, s->r0, s->w);
    for (int64_t t = 0; t < s->n; t++) s->w[t] = rhs[t] - s->w[t];
    double tn = gm_norm(s, s->w);
    double be = tn / bn;
    if (be >= s->tol && s->backward_check == 1)
        fprintf(stderr, "oracle fgmres: be = %g, tolerance = %g\n", be, s->tol);
    return be;
}

static double gm_final(gm *s, int jH, double bn, const xo_c *rhs, xo_c *result)
{
    gm_construct(s, jH, result);
    if (s->backward_check != 0) return gm_check(s, bn, rhs);
    return -1;
}

/* ClassicalGramSchmidtIteration (Gmres.cs:302-313) */
static double gm_cgs(gm *s, int jH)
{
    int64_t n = s->n;
    int cnt = jH + 1;
    for (int k = 0; k < cnt; k++) s->dots[k] = zdotc(n, s->V + (int64_t)k * n, s->w);
    *s->ndots += cnt;
    for (int k = 0; k < cnt; k++) s->H[HIDX(s, k, jH)] += s->dots[k];
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < n; t++) {
        xo_c a = 0;
        for (int k = 0; k < cnt; k++) a += s->V[(int64_t)k * n + t] * s->dots[k];
        s->w[t] = s->w[t] - a;
    }
    double d = 0;
    for (int k = 0; k < cnt; k++) d += creal(s->dots[k]) * creal(s->dots[k]) + cimag(s->dots[k]) * cimag(s->dots[k]);
    return d; /* Dznrm2^2 */
}

/* KrylovSubspaceConstruction (Gmres.cs:233-295) */
static double gm_krylov(gm *s, double bn, int *jHp)
{
    int m = s->m;
    int64_t n = s->n;
    double bea = -1;
    int jH = *jHp;
    do {
        gm_precond(s, s->V + (int64_t)jH * n, s->w, jH);
        gm_matvec(s, s->w, s->r0);
        memcpy(s->w, s->r0, sizeof(xo_c) * n);
        for (int j = 0; j <= jH; j++) s->H[HIDX(s, j, jH)] = 0;
        int north = 0;
        double dloo, dnormw;
        do {
            north++;
            dloo = gm_cgs(s, jH);
            dnormw = gm_norm(s, s->w);
        } while ((2.0 * dnormw <= sqrt(dloo)) && (north < 3));
        s->H[HIDX(s, jH + 1, jH)] = dnormw;
        if (jH < m - 1) {
            xo_c *vn = s->V + (int64_t)(jH + 1) * n;
            for (int64_t t = 0; t < n; t++) vn[t] = s->w[t] / dnormw;
        }
        for (int j = 0; j < jH; j++)
            zrot1(&s->H[HIDX(s, j, jH)], &s->H[HIDX(s, j + 1, jH)], s->gc[j], s->gs[j]);
        xo_zrotg(s->H[HIDX(s, jH, jH)], s->H[HIDX(s, jH + 1, jH)], &s->gc[jH], &s->gs[jH]);
        zrot1(&s->H[HIDX(s, jH, m)], &s->H[HIDX(s, jH + 1, m)], s->gc[jH], s->gs[jH]);
        zrot1(&s->H[HIDX(s, jH, jH)], &s->H[HIDX(s, jH + 1, jH)], s->gc[jH], s->gs[jH]);
        s->H[HIDX(s, jH + 1, jH)] = 0;
        double dnormres = cabs(s->H[HIDX(s, jH + 1, m)]);
        bea = dnormres / bn;
        s->iter_counter++;
        if (s->it) s->it(s->it_ctx, s->flexible, s->iter_counter, bea);
        jH++;
    } while (bea > s->tol && jH < m);
    *jHp = jH;
    return bea;
}

/* PrepareRestart, ComputeTheTrue branch (Gmres.cs:145-162) */
static double gm_restart(gm *s, const xo_c *rhs, xo_c *x, int jH)
{
    gm_construct(s, jH, x);
    memcpy(s->w, x, sizeof(xo_c) * s->n);
    gm_matvec(s, s->w, s->r0);
    for (int64_t t = 0; t < s->n; t++) s->r0[t] = rhs[t] - s->r0[t];
    memcpy(s->w, s->r0, sizeof(xo_c) * s->n);
    return gm_norm(s, s->w);
}

/* GmresSolver.Solve (Gmres.cs:77-143), InitialGuess = UserSupplied */
static int gm_solve(gm *s, const xo_c *rhs, const xo_c *x0, xo_c *result,
                    double *out_bea, double *out_be, int *out_conv)
{
    int m = s->m;
    int64_t n = s->n;
    s->iter_counter = 0;
    double bn = gm_norm(s, rhs);
    if (bn == 0) { *out_bea = 0; *out_be = 0; *out_conv = 1; return 0; }
    gm_matvec(s, x0, s->r0);
    for (int64_t t = 0; t < n; t++) s->r0[t] = rhs[t] - s->r0[t];
    memcpy(s->w, s->r0, sizeof(xo_c) * n);
    double beta = gm_norm(s, s->w);
    if (beta == 0) {   /* CheckForExactSolution: ResultInfo.Converged(0, 0, 1) (Gmres.cs:104-107) */
        if (result != x0) memcpy(result, x0, sizeof(xo_c) * n);
        *out_bea = 0; *out_be = 0; *out_conv = 1;
        s->iter_counter = 1;
        return 0;
    }
    double bea = -1, be = -1;
    int iter = 0;
    if (result != x0) memcpy(result, x0, sizeof(xo_c) * n);
    for (;;) {
        for (int64_t t = 0; t < n; t++) s->V[t] = (1 / beta) * s->w[t];
        s->H[HIDX(s, 0, m)] = beta;
        for (int i = 1; i <= m; i++) s->H[HIDX(s, i, m)] = 0;
        int jH = 0;
        bea = gm_krylov(s, bn, &jH);
        if (bea < s->tol) {
            be = gm_final(s, jH, bn, rhs, result);
            *out_bea = bea; *out_be = be; *out_conv = 1;
            return 0;
        }
        iter++;
        if (iter >= s->max_repeat) break;
        beta = gm_restart(s, rhs, result, jH);
    }
    be = gm_final(s, m, bn, rhs, result);
    *out_bea = bea; *out_be = be; *out_conv = 0;
    return 0;
}

int xo_fgmres_solve(const xo_gmres_params *p, xo_matvec_fn mv, void *mv_ctx,
                    xo_iter_fn it, void *it_ctx, const xo_c *b, xo_c *x, xo_result *res)
{
    gm outer, inner;
    int64_t matvecs = 0, ndots = 0;
    if (gm_alloc(&outer, p->n, p->m_outer, 1) || gm_alloc(&inner, p->n, p->m_inner, 0)) return -1;
    /* FlexibleGmresWithGmresPreconditioner ctor (:29-58) */
    outer.tol = p->tol; outer.max_repeat = p->max_repeat; outer.backward_check = p->backward_check;
    inner.tol = p->tol; inner.max_repeat = 1; inner.backward_check = 0;
    outer.mv = inner.mv = mv;
    outer.mv_ctx = inner.mv_ctx = mv_ctx;
    outer.it = it; outer.it_ctx = it_ctx;
    inner.it = NULL;
    outer.matvecs = inner.matvecs = &matvecs;
    outer.ndots = inner.ndots = &ndots;
    outer.inner = &inner;
    double bea, be;
    int conv;
    gm_solve(&outer, b, x, x, &bea, &be, &conv);
    res->converged = conv;
    res->bea = bea;
    res->be = be;
    res->iterations = outer.iter_counter;
    res->inner_iterations = inner.iter_counter;
    res->matvecs = matvecs;
    res->dots = ndots;
    gm_free(&outer);
    gm_free(&inner);
    return 0;
}

int xo_gmres_inner(const xo_gmres_params *p, xo_matvec_fn mv, void *mv_ctx,
                   const xo_c *v, xo_c *z, xo_result *res)
{
    gm inner;
    int64_t matvecs = 0, ndots = 0;
    if (gm_alloc(&inner, p->n, p->m_inner, 0)) return -1;
    inner.tol = p->tol; inner.max_repeat = 1; inner.backward_check = 0;
    inner.mv = mv; inner.mv_ctx = mv_ctx;
    inner.matvecs = &matvecs; inner.ndots = &ndots;
    double bea, be;
    int conv;
    gm_solve(&inner, v, v, z, &bea, &be, &conv);
    res->converged = conv; res->bea = bea; res->be = be;
    res->iterations = inner.iter_counter; res->inner_iterations = 0;
    res->matvecs = matvecs; res->dots = ndots;
    gm_free(&inner);
    return 0;
}

/* ======================================================================
 * Forward pipeline pieces (P3-P6)
 * ====================================================================== */

/* Mt3DForwardSolver.CalculateEFieldForPlaneWaveOnAnomalies (:271-275) */
void xo_normal_field(const xo_section *s, const xo_grid *g, double omega, xo_c *e0)
{
    for (int k = 0; k < g->nz; k++) {
        double wd = g->depth[k] + g->thick[k] / 2;
        e0[k] = xo_plane_wave_e(s, omega, wd, g->thick[k]);
    }
}

/* ForwardSolver.CalculateScatteredCurrentFromBackgroundField (:225-237) */
void xo_jscattered(const xo_grid *g, int pol, const xo_c *zeta, const xo_c *zeta0,
                   const xo_c *e0, xo_c *js)
{
    int nz = g->nz;
    for (int64_t ij = 0; ij < (int64_t)g->nx * g->ny; ij++)
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < nz; k++) {
                xo_c f = (c == pol) ? e0[k] : 0;
                js[ij * 3 * nz + c * nz + k] = (zeta[ij * nz + k] - zeta0[k]) * f;
            }
}

/* CalculateScatteredFieldFromChi (:391-406) and CalculateJqFromScatteredField (:424-436) */
void xo_escattered_jq(const xo_grid *g, const xo_c *zeta, const xo_c *zeta0,
                      const xo_c *chi, const xo_c *js, xo_c *es, xo_c *jq)
{
    int nz = g->nz;
    for (int64_t ij = 0; ij < (int64_t)g->nx * g->ny; ij++)
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < nz; k++) {
                int64_t t = ij * 3 * nz + c * nz + k;
                xo_c z = zeta[ij * nz + k];
                xo_c sq = csqrt(creal(zeta0[k]));
                es[t] = (2 * sq * chi[t] - js[t]) / (z + conj(zeta0[k]));
                jq[t] = es[t] * (z - zeta0[k]) + js[t];
            }
}
