/*
 * xo_green.c -- oracle: anomaly-to-anomaly (AtoA) electric Green tensor.
 * TEST INFRASTRUCTURE ONLY (see xtr_oracle.h).
 *
 * Restates, in order:
 *   scalar plans (ScalarPlansCreater.cs, ScalarPlan.cs, HankelCoefficients.cs:67-76),
 *   Green scalars (GreenScalarCalculator.cs:54-131, PlanCalculatorFast.cs:80-318,
 *                  VolumeToVolumeCalculatorUFast.cs:34-337, green_scalar.c:30-386,
 *                  ScalarMathUtils.cs:99-263, InnerResultsCombiner.cs:19-62),
 *   segments (ScalarSegments.cs:86-128),
 *   lateral cell integration (GreenTensorCalculator.cs:166-552, IntegralFactors.cs:17-30,
 *                  QSign.cs:64-67, AtoAGreenTensorCalculator.cs:229-321),
 *   mirroring (GreenTensorFftPopulator.cs:37-148) and the forward FFT
 *                 (AtoAGreenTensorCalculatorComponent.cs:197-279).
 */
#include "xtr_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define XO_PI 3.14159265358979323846
#define NTAP 431
#define HANK_L2 130 /* HankelCoefficients._l2; n1 = M1 - l2 = -130 */

/* ---------------------------------------------------------------------- */
/* plans                                                                   */
/* ---------------------------------------------------------------------- */

/* HankelCoefficients.GetLog10RhoStep (:67): (float)Math.Pow(10, 1/_ndec), _ndec float 40 */
static double log10_rho_step(void)
{
    float inv = 1.0f / 40.0f;
    return (double)(float)pow(10.0, (double)inv);
}

typedef struct { double *v; int n, cap; } dvec;

static void dpush(dvec *a, double x)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 64;
        a->v = realloc(a->v, sizeof(double) * a->cap);
    }
    a->v[a->n++] = x;
}

/* CreateRhoFromMin (ScalarPlansCreater.cs:142-157) */
static dvec rho_from_min(double rmin, double rmax, double step)
{
    dvec r = {0};
    double next = rmin;
    while (next <= rmax) { dpush(&r, next); next *= step; }
    dpush(&r, next);
    return r;
}

/* CreateRhoFromMax (ScalarPlansCreater.cs:159-176) */
static dvec rho_from_max(double rmin, double rmax, double step)
{
    dvec r = {0};
    double next = rmax;
    while (next >= rmin) { dpush(&r, next); next /= step; }
    dpush(&r, next);
    for (int i = 0; i < r.n / 2; i++) { double t = r.v[i]; r.v[i] = r.v[r.n - 1 - i]; r.v[r.n - 1 - i] = t; }
    return r;
}

/* ScalarPlan.CalculateLambdasForLog10 (ScalarPlan.cs:76-97) */
static double *lambdas_for(const double *rho, int nrho, int *nlam_out)
{
    double step = log10_rho_step();
    int nlam = NTAP + nrho;
    double *k = malloc(sizeof(double) * nlam);
    k[HANK_L2] = 1 / rho[0];
    for (int i = HANK_L2 + 1; i <= nlam - 1; i++) k[i] = k[i - 1] / step;
    for (int i = HANK_L2 - 1; i >= 0; i--) k[i] = k[i + 1] * step;
    *nlam_out = nlam;
    return k;
}

static int cmp_rho(const void *a, const void *b)
{
    const double *x = a, *y = b;
    if (x[0] < y[0]) return -1;
    if (x[0] > y[0]) return 1;
    /* first occurrence wins: order by (item, index) */
    if (x[1] < y[1]) return -1;
    if (x[1] > y[1]) return 1;
    return (x[2] < y[2]) ? -1 : (x[2] > y[2]);
}

int xo_plan_create(const xo_grid *g, int number_of_hankels, xo_plan *plan)
{
    /* FindRhoMin / FindRhoMax for the anomaly (ScalarPlansCreater.cs:360-380) */
    double h1 = g->dx / 2, h2 = g->dy / 2;
    double rmin = h1 < h2 ? h1 : h2;
    double mx = g->nx * g->dx - h1, my = g->ny * g->dy - h2;
    double rmax = sqrt(mx * mx + my * my);
    return xo_plan_create_range(number_of_hankels, rmin, rmax, 0, plan);
}

/* FindRhoMin (ScalarPlansCreater.cs:245-266) along one axis of a level shifted by `shift` against
 * the anomaly's cell centres 0, h, ..., (n - 1) h. The reference works in System.Decimal; these
 * doubles give the same values whenever the cell size and the shift are whole multiples of a power
 * of two fraction of a metre (fmod and the sums are then exact). */
static double level_rho_min_axis(int n, double h, double shift)
{
    double a0 = -h / 2, a1 = n * h - h / 2;                  /* anomaly edges */
    double y0 = shift - h / 2, y1 = y0 + h * n;              /* level edges */
    if ((y0 >= a0 && y0 <= a1) || (y1 >= a0 && y1 <= a1)) {
        double rem = fmod(fabs(a0 - y0), h);
        if (rem == 0) return h / 2;
        return fabs(h / 2 - rem);
    }
    double c0 = 0, c1 = h * (n - 1);                         /* first and last cell centres */
    double r = fabs(y0 - c0);
    if (fabs(y0 - c1) < r) r = fabs(y0 - c1);
    if (fabs(y1 - c0) < r) r = fabs(y1 - c0);
    if (fabs(y1 - c1) < r) r = fabs(y1 - c1);
    return r;
}

/* FindRhoMaxByFourPoints (:220-236): farthest corner of the level's footprint from (x0, y0) */
static double level_rho_max_four(const xo_grid *g, double sx, double sy, double x0, double y0)
{
    double x1 = sx - g->dx / 2, y1 = sy - g->dy / 2;
    double x2 = x1 + g->dx * g->nx, y2 = y1 + g->dy * g->ny;
    double c[4][2] = {{x1, y1}, {x2, y1}, {x2, y2}, {x1, y2}};
    double m = 0;
    for (int k = 0; k < 4; k++) {
        double r = sqrt((c[k][0] - x0) * (c[k][0] - x0) + (c[k][1] - y0) * (c[k][1] - y0));
        if (r > m) m = r;
    }
    return m;
}

int xo_atoo_level_rho_range(const xo_grid *g, double sx, double sy, double *rho_min, double *rho_max)
{
    /* FindRhoMax (:207-218): the four anomaly corner cells' centres */
    double ex = g->dx * (g->nx - 1), ey = g->dy * (g->ny - 1);
    double pts[4][2] = {{0, 0}, {0, ey}, {ex, ey}, {ex, 0}};
    double m = 0;
    for (int k = 0; k < 4; k++) {
        double r = level_rho_max_four(g, sx, sy, pts[k][0], pts[k][1]);
        if (r > m) m = r;
    }
    *rho_max = m;
    /* FindRhoMin (:238-243) and CreateForAnomalyToObservationLevels (:51-66) */
    double rx = level_rho_min_axis(g->nx, g->dx, sx), ry = level_rho_min_axis(g->ny, g->dy, sy);
    double r = rx < ry ? rx : ry;
    if (r <= 0) {
        *rho_min = (g->dx < g->dy ? g->dx : g->dy) * 0.001;
        return 1;
    }
    *rho_min = r;
    return 0;
}

int xo_plan_create_range(int number_of_hankels, double rmin, double rmax, int zero_rho, xo_plan *plan)
{
    memset(plan, 0, sizeof(*plan));
    int nh = abs(number_of_hankels);
    if (nh < 1 || !(rmin > 0)) return -1;
    double step = log10_rho_step();
    /* AddLog10Plans (:123-140) */
    plan->nitems = nh;
    plan->nrho = calloc(nh, sizeof(int));
    plan->rho = calloc(nh, sizeof(double *));
    plan->nlam = calloc(nh, sizeof(int));
    plan->lam = calloc(nh, sizeof(double *));
    dvec r1 = rho_from_min(rmin, rmax, step);
    double rr2 = r1.v[r1.n - 1], rr1 = r1.v[r1.n - 2];
    double st = (rr2 - rr1) / nh;
    plan->rho[0] = r1.v;
    plan->nrho[0] = r1.n;
    for (int i = 1; i < nh; i++) {
        dvec rn = rho_from_max(rmin, rr2 - i * st, step);
        plan->rho[i] = rn.v;
        plan->nrho[i] = rn.n;
    }
    for (int i = 0; i < nh; i++) plan->lam[i] = lambdas_for(plan->rho[i], plan->nrho[i], &plan->nlam[i]);
    /* InnerResultsCombiner.CreateFromSample (:19-33): sorted unique rho, first wins */
    int tot = 0;
    for (int i = 0; i < nh; i++) tot += plan->nrho[i];
    double *all = malloc(sizeof(double) * 3 * tot);
    int c = 0;
    for (int i = 0; i < nh; i++)
        for (int j = 0; j < plan->nrho[i]; j++) {
            all[3 * c] = plan->rho[i][j]; all[3 * c + 1] = i; all[3 * c + 2] = j; c++;
        }
    qsort(all, tot, 3 * sizeof(double), cmp_rho);
    plan->rho_merged = malloc(sizeof(double) * (tot + 1));
    plan->merge_item = malloc(sizeof(int) * (tot + 1));
    plan->merge_index = malloc(sizeof(int) * (tot + 1));
    int u = 0;
    plan->zero_rho = zero_rho != 0;
    if (plan->zero_rho) {     /* GreenScalarCalculator.CalculateZeroRho (:147-157): rho = 0 first */
        plan->rho_merged[0] = 0;
        plan->merge_item[0] = -1;
        plan->merge_index[0] = 0;
        u = 1;
    }
    for (int t = 0; t < tot; t++) {
        if (u > plan->zero_rho && all[3 * t] == plan->rho_merged[u - 1]) continue;
        plan->rho_merged[u] = all[3 * t];
        plan->merge_item[u] = (int)all[3 * t + 1];
        plan->merge_index[u] = (int)all[3 * t + 2];
        u++;
    }
    plan->nrho_merged = u;
    free(all);
    /* ScalarSegments takes its radii from ScalarPlan.GetSortedRho (all values, duplicates kept);
     * the two grids coincide only without duplicates, which is what the reference assumes. */
    if (u != tot + plan->zero_rho) {
        fprintf(stderr, "xo_plan_create: duplicate rho values (%d of %d unique)\n", u, tot);
        return -2;
    }
    return 0;
}

void xo_plan_free(xo_plan *plan)
{
    for (int i = 0; i < plan->nitems; i++) { free(plan->rho[i]); free(plan->lam[i]); }
    free(plan->rho); free(plan->lam); free(plan->nrho); free(plan->nlam);
    free(plan->rho_merged); free(plan->merge_item); free(plan->merge_index);
    memset(plan, 0, sizeof(*plan));
}

/* ---------------------------------------------------------------------- */
/* scalars                                                                 */
/* ---------------------------------------------------------------------- */

static double layer_depth(const xo_section *s, int index)
{
    double d = s->zero_level;
    for (int i = 0; i < index; i++) d += s->thickness[i];
    return d;
}

/* Per-item, per-lambda layer quantities (PlanCalculatorFast.cs:203-318) */
typedef struct {
    int nl, nlam;
    xo_c *eta, *ex, *p1, *q1, *ps, *qs; /* [layer][lam] */
} item_aux;

static void item_aux_build(const xo_section *s, double omega, const double *lam, int nlam, item_aux *a)
{
    int nl = s->nlayers;
    double mu0 = xo_mu0();
    a->nl = nl; a->nlam = nlam;
    size_t sz = (size_t)nl * nlam;
    a->eta = malloc(sizeof(xo_c) * sz); a->ex = malloc(sizeof(xo_c) * sz);
    a->p1 = calloc(sz, sizeof(xo_c)); a->q1 = calloc(sz, sizeof(xo_c));
    a->ps = calloc(sz, sizeof(xo_c)); a->qs = calloc(sz, sizeof(xo_c));
    xo_c *zeta = malloc(sizeof(xo_c) * nl);
    for (int k = 0; k < nl; k++) zeta[k] = xo_zeta(s->sigma[k], omega);
    for (int k = 0; k < nl; k++) {
        xo_c value = (I * omega) * zeta[k] * mu0;          /* iOmega * zeta * Mu0 (:222-229) */
        for (int m = 0; m < nlam; m++) {
            a->eta[k * nlam + m] = csqrt(lam[m] * lam[m] - value);          /* CalcEta */
            a->ex[k * nlam + m] = cexp(a->eta[k * nlam + m] * (-2 * s->thickness[k])); /* CalcExp */
        }
    }
#define E(k, m) a->eta[(k) * nlam + (m)]
#define X(k, m) a->ex[(k) * nlam + (m)]
    for (int m = 0; m < nlam; m++) {
        /* CalculateP1 (:236-249) + CalculatePBelow1 (:266-280), alpha1 = 1 */
        a->p1[1 * nlam + m] = (E(1, m) - E(0, m)) / (E(1, m) + E(0, m));
        /* CalculatePSigma (:251-264), alphaSigma[k] = zeta[k+1]/zeta[k] (AlphaBeta.cs:24-42) */
        a->ps[1 * nlam + m] = (zeta[0] * E(1, m) - zeta[1] * E(0, m)) / (zeta[0] * E(1, m) + zeta[1] * E(0, m));
        for (int k = 1; k < nl - 1; k++) {
            xo_c pe = a->p1[k * nlam + m] * X(k, m);
            xo_c part = 1.0 * (E(k, m) / E(k + 1, m)) * ((pe - 1) / (pe + 1));
            a->p1[(k + 1) * nlam + m] = (1 + part) / (1 - part);
            pe = a->ps[k * nlam + m] * X(k, m);
            part = (zeta[k + 1] / zeta[k]) * (E(k, m) / E(k + 1, m)) * ((pe - 1) / (pe + 1));
            a->ps[(k + 1) * nlam + m] = (1 + part) / (1 - part);
        }
        /* CalculateQ (:292-318), beta1 = 1, betaSigma[k] = zeta[k-1]/zeta[k] */
        int n = nl - 1;
        a->q1[n * nlam + m] = 0;
        a->qs[n * nlam + m] = 0;
        xo_c part = 1.0 * (E(n, m) / E(n - 1, m));
        a->q1[(n - 1) * nlam + m] = (1 - part) / (1 + part);
        part = (zeta[n - 1] / zeta[n]) * (E(n, m) / E(n - 1, m));
        a->qs[(n - 1) * nlam + m] = (1 - part) / (1 + part);
        for (int k = n - 1; k > 1; k--) {
            xo_c qe = a->q1[k * nlam + m] * X(k, m);
            part = 1.0 * (E(k, m) / E(k - 1, m)) * ((qe - 1) / (qe + 1));
            a->q1[(k - 1) * nlam + m] = (1 + part) / (1 - part);
            qe = a->qs[k * nlam + m] * X(k, m);
            part = (zeta[k - 1] / zeta[k]) * (E(k, m) / E(k - 1, m)) * ((qe - 1) / (qe + 1));
            a->qs[(k - 1) * nlam + m] = (1 + part) / (1 - part);
        }
    }
#undef E
#undef X
    free(zeta);
}

static void item_aux_free(item_aux *a)
{
    free(a->eta); free(a->ex); free(a->p1); free(a->q1); free(a->ps); free(a->qs);
}

/* A[rc][tr] at lambda m for the P/Q set given (CalculateA :160-200, CalculateASigma :138-158) */
static xo_c ann_at(const xo_section *s, double omega, const item_aux *a, const xo_c *P, const xo_c *Q,
                   int rc, int tr, int m, int sigma)
{
    int nlam = a->nlam;
    int smax = rc > tr ? rc : tr, rmin = rc < tr ? rc : tr;
    int n = smax;
    xo_c prev = 1.0 / (a->eta[n * nlam + m] * (1 - P[n * nlam + m] * Q[n * nlam + m] * a->ex[n * nlam + m]));
    for (int k = n - 1; k >= rmin; k--) {
        double dk1 = layer_depth(s, k + 1);
        xo_c q = ((1 + P[(k + 1) * nlam + m]) / (1 + P[k * nlam + m] * a->ex[k * nlam + m])) *
                 cexp((a->eta[(k + 1) * nlam + m] - a->eta[k * nlam + m]) * dk1);
        prev = q * prev;
    }
    /* prev = A[rmin][smax] = A[smax][rmin] before the Sigma rescaling */
    if (sigma && rc > tr) {
        xo_c zrc = xo_zeta(s->sigma[rc], omega), ztr = xo_zeta(s->sigma[tr], omega);
        /* a[n,k] = (zeta_n / zeta_k) a[k,n] with n = rc > k = tr */
        return (zrc / ztr) * prev;
    }
    return prev;
}

typedef struct {
    int r, s, t, b;
    double dt, db1, r1, r2, s1, s2;
} envelop; /* NativeEnvelop (green_scalar.c:13-28) */

static xo_c e4(xo_c a, xo_c b, xo_c c, xo_c d) { return cexp(a) - cexp(b) - cexp(c) + cexp(d); }

/* U1..U5 for one transceiver pair at one lambda (VolumeToVolumeCalculatorUFast.cs:34-337).
 * which: bit0 U1, bit1 U2, bit2 U3, bit3 U4, bit4 U5. */
static void pair_u(const xo_section *sec, double omega, const item_aux *a, const envelop *ne,
                   double lambda, int m, int which, int kase, xo_c *u)
{
    int nlam = a->nlam;
    double mu0 = xo_mu0();
    xo_c er = a->eta[ne->r * nlam + m], es = a->eta[ne->s * nlam + m];
    xo_c et = a->eta[ne->t * nlam + m], eb = a->eta[ne->b * nlam + m];
    /* PreCalculate: T, R, Ctop, Cbot, Dtop, Dbot (green_scalar.c:86-386) */
    xo_c T = e4(-er * ne->r1 - es * ne->s1 + 2 * ne->dt * et, -er * ne->r1 - es * ne->s2 + 2 * ne->dt * et,
                -er * ne->r2 - es * ne->s1 + 2 * ne->dt * et, -er * ne->r2 - es * ne->s2 + 2 * ne->dt * et);
    xo_c R = e4(-2 * ne->db1 * eb + ne->s2 * es + ne->r2 * er, -2 * ne->db1 * eb + ne->s1 * es + ne->r2 * er,
                -2 * ne->db1 * eb + ne->s2 * es + ne->r1 * er, -2 * ne->db1 * eb + ne->s1 * es + ne->r1 * er);
    xo_c Ctop = 0, Cbot = 0, Dtop = 0, Dbot = 0;
    xo_c F = -2 * ne->db1 * eb + 2 * ne->dt * et;
    if (kase == 2) { /* rc above tr */
        Ctop = e4(-ne->s1 * es + ne->r2 * er, -ne->s2 * es + ne->r2 * er,
                  -ne->s1 * es + ne->r1 * er, -ne->s2 * es + ne->r1 * er);
        Dtop = e4(F + ne->s2 * es - ne->r1 * er, F + ne->s1 * es - ne->r1 * er,
                  F + ne->s2 * es - ne->r2 * er, F + ne->s1 * es - ne->r2 * er);
    } else if (kase == 1) { /* rc below tr */
        Cbot = e4(-ne->r1 * er + ne->s2 * es, -ne->r1 * er + ne->s1 * es,
                  -ne->r2 * er + ne->s2 * es, -ne->r2 * er + ne->s1 * es);
        Dbot = e4(F - ne->s1 * es + ne->r2 * er, F - ne->s2 * es + ne->r2 * er,
                  F - ne->s1 * es + ne->r1 * er, F - ne->s2 * es + ne->r1 * er);
    }
    for (int set = 0; set < 2; set++) {
        /* set 0: P1/Q1/A1 (U1); set 1: PSigma/QSigma/ASigma (U2..U5) */
        if (set == 0 && !(which & 1)) continue;
        if (set == 1 && !(which & 30)) continue;
        const xo_c *P = set ? a->ps : a->p1;
        const xo_c *Q = set ? a->qs : a->q1;
        xo_c Pt = P[ne->t * nlam + m], Qb = Q[ne->b * nlam + m];
        xo_c ann = ann_at(sec, omega, a, P, Q, ne->r, ne->s, m, set);
        /* psi for U1/U2/U5 (Psi1And5...) and U3/U4 (Psi3And4...) */
        xo_c psi1, psi3;
        if (kase == 0) {
            int n = ne->r;
            double tau = ne->r2 - ne->r1;
            double dn1 = layer_depth(sec, n + 1) - layer_depth(sec, n);
            xo_c en = a->eta[n * nlam + m];
            xo_c Pn = P[n * nlam + m], Qn = Q[n * nlam + m];
            /* CalcPsi1RcAndTrMatch (:313-332) */
            xo_c part1 = tau * en - (1 - cexp(-en * tau));
            xo_c part2 = Pn * Qn * ((tau * en + 1) * cexp(-2 * en * dn1) - cexp(-2 * en * dn1 + en * tau));
            psi1 = 2 * (part1 - part2);
            psi3 = 0; /* CalcPsi3And4RcAndTrMatch (:255-260) */
        } else if (kase == 1) {
            psi1 = Cbot + Pt * Qb * Dbot;   /* CalcPsi1And5RcBelowTr (:242-253) */
            psi3 = -Cbot + Pt * Qb * Dbot;  /* CalcPsi3And4RcBelowTr (:275-286) */
        } else {
            psi1 = Ctop + Pt * Qb * Dtop;   /* CalcPsi1And5RcAboveTr (:229-240) */
            psi3 = Ctop - Pt * Qb * Dtop;   /* CalcPsi3And4RcAboveTr (:262-273) */
        }
        xo_c fac = 1 / (er * es);
        if (set == 0) {
            /* CalculateU1And5Special (:144-159) */
            u[0] = ann * fac * (1 * Pt * T + Qb * R + 1 * psi1);
            continue;
        }
        if (which & 2) u[1] = ann * (1 * Pt * T + Qb * R - 1 * psi1);          /* U2Special (:183-193) */
        if (which & 4) u[2] = ann * (1 / es) * (-1 * Pt * T + Qb * R + 1 * psi3); /* U3Special (:195-210) */
        if (which & 8) u[3] = ann * (1 / er) * (-1 * Pt * T + Qb * R - 1 * psi3); /* U4Special (:212-227) */
        if (which & 16) {
            if (kase == 0) {
                /* CalculateU5SpecialForTrAndRcMatch (:161-181) + CalcPsi5RcAndTrMatch (:290-311) */
                int n = ne->r;
                double tau = ne->r2 - ne->r1;
                double ln = sec->thickness[n];
                xo_c kk = (I * omega) * xo_zeta(sec->sigma[n], omega) * mu0;
                xo_c en = a->eta[n * nlam + m];
                xo_c Pn = P[n * nlam + m], Qn = Q[n * nlam + m];
                xo_c part1 = -(1 - cexp(-en * tau));
                xo_c part2 = Pn * Qn * (cexp(-en * (2 * ln - tau)) - cexp(-2 * en * ln));
                xo_c part3 = (kk / (lambda * lambda)) * (1 - Pn * Qn * cexp(-2 * en * ln)) * tau * en;
                xo_c psi5 = 2 * (part1 + part2 + part3);
                u[4] = (ann * fac) * (Pt * T + Qb * R + psi5);
            } else {
                u[4] = ann * fac * (1 * Pt * T + Qb * R + 1 * psi1);
            }
        }
    }
}

/* HankelCoefficients.ConvoluteWithFast (:103-118): fr[i] = sum_j revHank[j] fk[i + j] */
static xo_c hankel_dot(const double *hank, const xo_c *fk, int i)
{
    double re = 0, im = 0;
    for (int j = 0; j < NTAP; j++) {
        double h = hank[NTAP - 1 - j];
        re += creal(fk[i + j]) * h;   /* CalcDotProduct (green_scalar.c:43-55) */
        im += cimag(fk[i + j]) * h;
    }
    return re + I * im;
}

/* Test hooks for the filter alone: the lambda grid of an output grid rho[i] = rho0 * step^i
 * (ScalarPlan.CalculateLambdasForLog10) and the filter sums on it (ConvoluteWithFast). */
double xo_log10_rho_step(void) { return log10_rho_step(); }

int xo_hankel_lambdas(double rho0, int nrho, double *lam)
{
    int nlam;
    double *k = lambdas_for(&rho0, nrho, &nlam);
    memcpy(lam, k, sizeof(double) * nlam);
    free(k);
    return nlam;
}

void xo_hankel_filter(const double *hank, const xo_c *fk, int nrho, xo_c *out)
{
    for (int i = 0; i < nrho; i++) out[i] = hankel_dot(hank, fk, i);
}

int xo_green_scalars(const xo_section *sec, const xo_grid *g, double omega,
                     const double *hank0, const double *hank1, const xo_plan *plan,
                     int asym, xo_c *out)
{
    int nz = g->nz;
    int npairs = asym ? nz * nz : nz * (nz + 1) / 2;
    int nrho = plan->nrho_merged;
    int ncomp = asym ? 1 : 3;
    double mu0 = xo_mu0();
    int *pr = malloc(sizeof(int) * npairs), *pt = malloc(sizeof(int) * npairs);
    if (asym) {
        for (int rc = 0; rc < nz; rc++)
            for (int tr = 0; tr < nz; tr++) { pr[rc * nz + tr] = rc; pt[rc * nz + tr] = tr; }
    } else {
        int c = 0;
        for (int tr = 0; tr < nz; tr++)
            for (int rc = 0; rc <= tr; rc++) { pr[c] = rc; pt[c] = tr; c++; }
    }
    item_aux *aux = malloc(sizeof(item_aux) * plan->nitems);
    for (int it = 0; it < plan->nitems; it++)
        item_aux_build(sec, omega, plan->lam[it], plan->nlam[it], &aux[it]);
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < npairs; p++) {
        int rc = pr[p], tr = pt[p];
        double r1 = g->depth[rc], r2 = g->depth[rc] + g->thick[rc];
        double s1 = g->depth[tr], s2 = g->depth[tr] + g->thick[tr];
        envelop ne;
        ne.r = xo_layer_at_depth(sec, r1 + (r2 - r1) / 2);  /* AuxContainerFast.CreateContainer */
        ne.s = xo_layer_at_depth(sec, s1 + (s2 - s1) / 2);
        if (ne.r < 0 || ne.s < 0) { err = 1; continue; }
        ne.r1 = r1; ne.r2 = r2; ne.s1 = s1; ne.s2 = s2;
        ne.t = ne.r < ne.s ? ne.r : ne.s;
        ne.b = ne.r > ne.s ? ne.r : ne.s;
        ne.dt = layer_depth(sec, ne.t);
        ne.db1 = layer_depth(sec, ne.b + 1);
        /* CalcUGeneral (:90-103) */
        double zw = r1 + (r2 - r1) / 2;
        int kase = (zw > s1 && zw < s2) ? 0 : (zw >= s2 ? 1 : 2);
        xo_c zrc2 = 2 * xo_zeta(sec->sigma[ne.r], omega);
        xo_c iwm = I * omega * mu0;
        xo_c **items = malloc(sizeof(xo_c *) * plan->nitems);
        for (int it = 0; it < plan->nitems; it++) {
            int nlam = plan->nlam[it], nr = plan->nrho[it];
            const double *lam = plan->lam[it];
            xo_c *U = malloc(sizeof(xo_c) * nlam * ncomp);  /* [comp][lam] */
            for (int m = 0; m < nlam; m++) {
                xo_c u[5];
                if (asym) {
                    pair_u(sec, omega, &aux[it], &ne, lam[m], m, 4, kase, u);
                    U[m] = u[2] * (lam[m] / zrc2);                         /* NormalizeBeforeHankelForE */
                } else {
                    pair_u(sec, omega, &aux[it], &ne, lam[m], m, 1 | 2 | 16, kase, u);
                    U[0 * nlam + m] = u[0] * (iwm / 2);
                    U[1 * nlam + m] = u[1] / zrc2;
                    U[2 * nlam + m] = u[4] * ((lam[m] * lam[m]) / zrc2);
                }
            }
            xo_c *Ir = malloc(sizeof(xo_c) * nr * ncomp);
            for (int i = 0; i < nr; i++) {
                double rho = plan->rho[it][i];
                if (asym) {
                    Ir[i] = hankel_dot(hank0, U, i) / (2 * XO_PI * rho);                 /* I3: J0, +1/(2 pi rho) */
                } else {
                    Ir[0 * nr + i] = -hankel_dot(hank1, U + 0 * nlam, i) / (2 * XO_PI * rho); /* I1: J1, -1/(2 pi rho) */
                    Ir[1 * nr + i] = hankel_dot(hank1, U + 1 * nlam, i) / (2 * XO_PI * rho);  /* I2: J1 */
                    Ir[2 * nr + i] = hankel_dot(hank1, U + 2 * nlam, i) / (2 * XO_PI * rho);  /* I5: J1 */
                }
            }
            free(U);
            items[it] = Ir;
        }
        for (int k = 0; k < nrho; k++) {
            int it = plan->merge_item[k], idx = plan->merge_index[k], nr = plan->nrho[it];
            for (int c = 0; c < ncomp; c++)
                out[((int64_t)c * npairs + p) * nrho + k] = items[it][c * nr + idx];
        }
        for (int it = 0; it < plan->nitems; it++) free(items[it]);
        free(items);
    }
    for (int it = 0; it < plan->nitems; it++) item_aux_free(&aux[it]);
    free(aux); free(pr); free(pt);
    return err ? -1 : 0;
}

/* ---------------------------------------------------------------------- */
/* lateral integration                                                     */
/* ---------------------------------------------------------------------- */

typedef struct { int index; double radius; } knot;

/* PrepareKnot (GreenTensorCalculator.cs:207-228) */
static knot make_knot(const double *radii, int nr, double x, double y)
{
    double r = sqrt(x * x + y * y);
    int lo = 0, hi = nr - 1, idx = -1;
    /* Array.BinarySearch: exact index or (insertion point - 1) */
    while (lo <= hi) {
        int mid = lo + (hi - lo) / 2;
        if (radii[mid] == r) { idx = mid; break; }
        if (radii[mid] < r) lo = mid + 1; else hi = mid - 1;
    }
    if (idx < 0) idx = lo - 1;
    if (idx == nr - 1) idx--;
    if (idx < 0) idx = 0;
    knot k = {idx, r};
    return k;
}

typedef struct { double d1a, d1b, d2a, d2b, d3a, d3b; } factors;

/* IntegralFactors.Prepare (IntegralFactors.cs:17-30) */
static factors prep_factors(double s, double u, double rs, double ru, double v)
{
    factors f;
    f.d1a = v == 0 ? (u * ru - s * rs) / 2 : (u * ru - s * rs + v * v * (log((u + ru) / (s + rs)))) / 2;
    f.d1b = u - s;
    f.d2a = v == 0 ? 0 : -(u - s) * v;
    f.d2b = v == 0 ? 0 : -(v * log((u + ru) / (s + rs)));
    f.d3a = -(u * u - s * s) / 2;
    f.d3b = -(ru - rs);
    return f;
}

/* One pair-chunk integration context */
typedef struct {
    int asym;
    int npairs, p0, pc;          /* chunk [p0, p0 + pc) */
    int nseg;
    const xo_c *segA, *segB;     /* [seg][comp][pair] */
    int ncomp;
    const double *radii;
    int nr;
    xo_c *q[4];                  /* asym: q[0]=xz q[1]=yz; symm: xx yy zz xy */
    xo_c *G[4];                  /* component tensors [2Nx][2Ny][P] */
    int ny2;
} ictx;

/* AddToQBuffAlong{X,Y}Electric (GreenTensorCalculator.cs:505-528) */
static void add_seg(ictx *c, const factors *f, int seg, int alongY)
{
    int np = c->npairs;
    for (int p = c->p0; p < c->p0 + c->pc; p++) {
        int lp = p - c->p0;
        if (c->asym) {
            xo_c a3 = c->segA[((int64_t)seg * 1 + 0) * np + p], b3 = c->segB[((int64_t)seg * 1 + 0) * np + p];
            if (alongY) { c->q[0][lp] += f->d1a * a3; c->q[0][lp] += f->d1b * b3; }  /* Qxz */
            else { c->q[1][lp] += f->d1a * a3; c->q[1][lp] += f->d1b * b3; }          /* Qyz */
        } else {
            xo_c a1 = c->segA[((int64_t)seg * 3 + 0) * np + p], b1 = c->segB[((int64_t)seg * 3 + 0) * np + p];
            xo_c a2 = c->segA[((int64_t)seg * 3 + 1) * np + p], b2 = c->segB[((int64_t)seg * 3 + 1) * np + p];
            xo_c a5 = c->segA[((int64_t)seg * 3 + 2) * np + p], b5 = c->segB[((int64_t)seg * 3 + 2) * np + p];
            if (!alongY) {
                c->q[0][lp] += f->d2a * a1; c->q[0][lp] += f->d2b * b1;   /* Qxx <- I1, D2 */
                c->q[1][lp] += f->d2a * a2; c->q[1][lp] += f->d2b * b2;   /* Qyy <- I2, D2 */
                c->q[2][lp] += f->d2a * a5; c->q[2][lp] += f->d2b * b5;   /* Qzz <- I5, D2 */
                c->q[3][lp] += f->d3a * a2; c->q[3][lp] += f->d3b * b2;   /* Qxy <- I2, D3 */
            } else {
                c->q[0][lp] += f->d2a * a2; c->q[0][lp] += f->d2b * b2;   /* Qxx <- I2, D2 */
                c->q[1][lp] += f->d2a * a1; c->q[1][lp] += f->d2b * b1;   /* Qyy <- I1, D2 */
                c->q[2][lp] += f->d2a * a5; c->q[2][lp] += f->d2b * b5;   /* Qzz <- I5, D2 */
                c->q[3][lp] += f->d3a * a1; c->q[3][lp] += f->d3b * b1;   /* Qxy <- I1, D3 */
            }
        }
    }
}

/* Integrate (GreenTensorCalculator.cs:367-411) */
static void integrate(ictx *c, knot left, knot right, double s, double u, double v, int alongY)
{
    int nq = c->asym ? 2 : 4;
    for (int t = 0; t < nq; t++) memset(c->q[t], 0, sizeof(xo_c) * c->pc);
    int i1 = left.index, i2 = right.index;
    if (i1 == i2) {
        factors f = prep_factors(s, u, left.radius, right.radius, v);
        add_seg(c, &f, i1, alongY);
    } else {
        double r1 = left.radius, r2 = c->radii[i1 + 1];
        double sj = s, uj = sqrt(r2 * r2 - v * v);
        factors f = prep_factors(sj, uj, r1, r2, v);
        add_seg(c, &f, i1, alongY);
        for (int i = i1 + 1; i < i2; i++) {
            r1 = r2; r2 = c->radii[i + 1];
            sj = uj; uj = sqrt(r2 * r2 - v * v);
            f = prep_factors(sj, uj, r1, r2, v);
            add_seg(c, &f, i, alongY);
        }
        sj = uj; uj = u; r1 = r2; r2 = right.radius;
        f = prep_factors(sj, uj, r1, r2, v);
        add_seg(c, &f, i2, alongY);
    }
}

/* QSigns (QSign.cs:64-67), per component in q-order */
static const int QS_ASYM[4][2] = {   /* xz, yz */
    {1, 1},   /* AlongX */
    {1, 1},   /* AlongY */
    {-1, 1},  /* AlongXSymm */
    {1, -1},  /* AlongYSymm */
};
static const int QS_SYMM[4][4] = {   /* xx, yy, zz, xy */
    {1, -1, -1, -1},
    {-1, 1, -1, -1},
    {1, -1, -1, 1},
    {-1, 1, -1, 1},
};

/* FillGreen (GreenTensorCalculator.cs:460-497) */
static void fill_green(ictx *c, int i, int j, int sign, int qsel)
{
    int nq = c->asym ? 2 : 4;
    int64_t shift = ((int64_t)i * c->ny2 + j) * c->npairs;
    for (int t = 0; t < nq; t++) {
        int qs = c->asym ? QS_ASYM[qsel][t] : QS_SYMM[qsel][t];
        double al = (double)(sign * qs);
        xo_c *dst = c->G[t] + shift + c->p0;
        for (int lp = 0; lp < c->pc; lp++) dst[lp] += al * c->q[t][lp];
    }
}

static void fill_along_x(ictx *c, int i, int j, int left, int right, int qsel)
{
    if (j == left) fill_green(c, i, left, -1, qsel);
    else if (j == right) fill_green(c, i, right - 1, 1, qsel);
    else { fill_green(c, i, j - 1, 1, qsel); fill_green(c, i, j, -1, qsel); }
}

static void fill_along_y(ictx *c, int i, int j, int left, int right, int qsel)
{
    if (i == left) fill_green(c, left, j, -1, qsel);
    else if (i == right) fill_green(c, right - 1, j, 1, qsel);
    else { fill_green(c, i - 1, j, 1, qsel); fill_green(c, i, j, -1, qsel); }
}

static void run_chunk(ictx *c, const xo_grid *g, const knot *K, const knot *ZX, const knot *ZY)
{
    int nx = g->nx, ny = g->ny;
    double dx = g->dx, dy = g->dy;
#define KN(i, j) K[(i) * (ny + 1) + (j)]
    /* RunAlongX (GreenTensorCalculator.cs:254-298) */
    for (int j = 0; j <= ny; j++) {
        double y = dy * (j - 0.5);
        for (int i = 0; i <= nx - 1; i++) {
            double x1 = dx * (i - 0.5), x2 = dx * (i + 0.5);
            if (x1 > 0 && x2 > 0) {
                integrate(c, KN(i, j), KN(i + 1, j), x1, x2, y, 0);
                fill_along_x(c, i, j, 0, ny, 0);
            } else if (x1 < 0 && x2 < 0) {
                integrate(c, KN(i + 1, j), KN(i, j), -x2, -x1, y, 0);
                fill_along_x(c, i, j, 0, ny, 2);
            } else {
                integrate(c, ZX[j], KN(i, j), 0, -x1, y, 0);
                fill_along_x(c, i, j, 0, ny, 2);
                integrate(c, ZX[j], KN(i + 1, j), 0, x2, y, 0);
                fill_along_x(c, i, j, 0, ny, 0);
            }
        }
    }
    /* RunAlongY (GreenTensorCalculator.cs:318-361) */
    for (int i = 0; i <= nx; i++) {
        double x = dx * (i - 0.5);
        for (int j = 0; j <= ny - 1; j++) {
            double y1 = dy * (j - 0.5), y2 = dy * (j + 0.5);
            if (y1 > 0 && y2 > 0) {
                integrate(c, KN(i, j), KN(i, j + 1), y1, y2, x, 1);
                fill_along_y(c, i, j, 0, nx, 1);
            } else if (y1 < 0 && y2 < 0) {
                integrate(c, KN(i, j + 1), KN(i, j), -y2, -y1, x, 1);
                fill_along_y(c, i, j, 0, nx, 3);
            } else {
                integrate(c, ZY[i], KN(i, j), 0, -y1, x, 1);
                fill_along_y(c, i, j, 0, nx, 3);
                integrate(c, ZY[i], KN(i, j + 1), 0, y2, x, 1);
                fill_along_y(c, i, j, 0, nx, 1);
            }
        }
    }
#undef KN
}

/* GreenTensorFftPopulator.PopulateForFft (:37-52, 88-148) */
static void mirror(xo_c *G, int nx, int ny, int64_t len, int xs, int ys)
{
    int64_t nx2 = 2 * nx, ny2 = 2 * ny;
    for (int i = 1; i < nx; i++)
        for (int j = 1; j < ny; j++) {
            const xo_c *src = G + len * (i * ny2 + j);
            xo_c *dx_ = G + len * ((nx2 - i) * ny2 + j);
            xo_c *dy_ = G + len * (i * ny2 + ny2 - j);
            xo_c *dxy = G + len * ((nx2 - i) * ny2 + ny2 - j);
            for (int64_t p = 0; p < len; p++) {
                dx_[p] += xs * src[p];
                dy_[p] += ys * src[p];
                dxy[p] += (xs * ys) * src[p];
            }
        }
    for (int i = 1; i < nx; i++) {
        const xo_c *src = G + len * (i * ny2);
        xo_c *dst = G + len * ((nx2 - i) * ny2);
        for (int64_t p = 0; p < len; p++) dst[p] += xs * src[p];
    }
    for (int j = 1; j < ny; j++) {
        const xo_c *src = G + len * j;
        xo_c *dst = G + len * (ny2 - j);
        for (int64_t p = 0; p < len; p++) dst[p] += ys * src[p];
    }
}

static int tensor_group(const xo_section *sec, const xo_grid *g, double omega, const double *hank0,
                        const double *hank1, const xo_plan *plan, int asym, int fft, xo_c **G)
{
    int nz = g->nz, nx = g->nx, ny = g->ny;
    int npairs = asym ? nz * nz : nz * (nz + 1) / 2;
    int ncomp = asym ? 1 : 3;
    int nrho = plan->nrho_merged;
    int nseg = nrho - 1;
    xo_c *Isc = malloc(sizeof(xo_c) * (size_t)ncomp * npairs * nrho);
    if (xo_green_scalars(sec, g, omega, hank0, hank1, plan, asym, Isc)) { free(Isc); return -1; }
    /* ScalarSegments.CalculateScalarSegment (ScalarSegments.cs:86-128) */
    xo_c *segA = malloc(sizeof(xo_c) * (size_t)nseg * ncomp * npairs);
    xo_c *segB = malloc(sizeof(xo_c) * (size_t)nseg * ncomp * npairs);
    const double *radii = plan->rho_merged;
    for (int k = 0; k < nseg; k++) {
        double s = radii[k], u = radii[k + 1], div = u - s;
        for (int c = 0; c < ncomp; c++)
            for (int p = 0; p < npairs; p++) {
                xo_c i0 = Isc[((int64_t)c * npairs + p) * nrho + k], i1 = Isc[((int64_t)c * npairs + p) * nrho + k + 1];
                xo_c a = (i1 - i0) / div;
                segA[((int64_t)k * ncomp + c) * npairs + p] = a;
                segB[((int64_t)k * ncomp + c) * npairs + p] = i0 - a * s;
            }
    }
    free(Isc);
    /* PrepareKnots (GreenTensorCalculator.cs:166-205) */
    knot *K = malloc(sizeof(knot) * (nx + 1) * (ny + 1));
    knot *ZX = malloc(sizeof(knot) * (ny + 1)), *ZY = malloc(sizeof(knot) * (nx + 1));
    for (int i = 0; i <= nx; i++)
        for (int j = 0; j <= ny; j++)
            K[i * (ny + 1) + j] = make_knot(radii, nrho, g->dx * (i - 0.5), g->dy * (j - 0.5));
    for (int j = 0; j <= ny; j++) ZX[j] = make_knot(radii, nrho, 0, g->dy * (j - 0.5));
    for (int i = 0; i <= nx; i++) ZY[i] = make_knot(radii, nrho, g->dx * (i - 0.5), 0);
    int nq = asym ? 2 : 4;
    int64_t compsize = (int64_t)4 * nx * ny * npairs;
    for (int t = 0; t < nq; t++) memset(G[t], 0, sizeof(xo_c) * compsize);
    const int CH = 32;
    int nchunks = (npairs + CH - 1) / CH;
#pragma omp parallel for schedule(dynamic, 1)
    for (int ch = 0; ch < nchunks; ch++) {
        ictx c;
        memset(&c, 0, sizeof(c));
        c.asym = asym; c.npairs = npairs; c.p0 = ch * CH;
        c.pc = (c.p0 + CH <= npairs) ? CH : npairs - c.p0;
        c.nseg = nseg; c.segA = segA; c.segB = segB; c.ncomp = ncomp;
        c.radii = radii; c.nr = nrho; c.ny2 = 2 * ny;
        for (int t = 0; t < nq; t++) { c.q[t] = calloc(CH, sizeof(xo_c)); c.G[t] = G[t]; }
        run_chunk(&c, g, K, ZX, ZY);
        for (int t = 0; t < nq; t++) free(c.q[t]);
    }
    free(K); free(ZX); free(ZY); free(segA); free(segB);
    /* mirror with parities; q-order asym {xz, yz}, symm {xx, yy, zz, xy} */
    if (asym) {
        mirror(G[0], nx, ny, npairs, -1, 1);   /* XZ */
        mirror(G[1], nx, ny, npairs, 1, -1);   /* YZ */
    } else {
        mirror(G[0], nx, ny, npairs, 1, 1);    /* XX */
        mirror(G[1], nx, ny, npairs, 1, 1);    /* YY */
        mirror(G[2], nx, ny, npairs, 1, 1);    /* ZZ */
        mirror(G[3], nx, ny, npairs, -1, -1);  /* XY */
    }
    if (fft)
        for (int t = 0; t < nq; t++) xo_fft2_batch(G[t], 2 * nx, 2 * ny, npairs, -1);
    return 0;
}

int xo_green_atoa(const xo_section *s, const xo_grid *g, double omega, int number_of_hankels,
                  const double *hank0, const double *hank1, int fft,
                  xo_c *xx, xo_c *xy, xo_c *xz, xo_c *yy, xo_c *yz, xo_c *zz)
{
    xo_plan plan;
    int rc = xo_plan_create(g, number_of_hankels, &plan);
    if (rc) return rc;
    xo_c *ga[2] = {xz, yz};
    xo_c *gs[4] = {xx, yy, zz, xy};
    rc = tensor_group(s, g, omega, hank0, hank1, &plan, 1, fft, ga);
    if (!rc) rc = tensor_group(s, g, omega, hank0, hank1, &plan, 0, fft, gs);
    xo_plan_free(&plan);
    return rc;
}

/* ====================================================================== */
/* AtoO: anomaly (volume transmitters) to an observation level (points)   */
/* ====================================================================== */

/* VolumeToPointCalculatorU (VolumeToPointCalculatorU.cs:34-337): U1..U5 of transmitter layer
 * [z1, z2] (background layer s) seen at the point zw (background layer r), at lambda index m,
 * with the P/Q set `sigma` (0: P1/Q1 with A1, 1: PSigma/QSigma with ASigma). */
typedef struct { xo_c ann, er, es, et, eb, pt, qb, pn, qn; double dt, db1; } pu_ctx;

static xo_c ctop(const xo_c *eta, int nlam, int m, double zw, double z1, double z2, int r, int s)
{
    xo_c er = eta[r * nlam + m], es = eta[s * nlam + m];
    return cexp(-es * z1 + er * zw) - cexp(-es * z2 + er * zw);
}
static xo_c cbot(const xo_c *eta, int nlam, int m, double zw, double z1, double z2, int r, int s)
{
    xo_c er = eta[r * nlam + m], es = eta[s * nlam + m];
    return cexp(-er * zw + es * z2) - cexp(-er * zw + es * z1);
}
static xo_c dtop(const xo_section *sec, const xo_c *eta, int nlam, int m, double zw, double z1, double z2, int r, int s)
{
    int t = r < s ? r : s, b = r > s ? r : s;
    double dt = layer_depth(sec, t), db1 = layer_depth(sec, b + 1);
    xo_c er = eta[r * nlam + m], es = eta[s * nlam + m], et = eta[t * nlam + m], eb = eta[b * nlam + m];
    xo_c base = -2 * (eb * db1 - et * dt);
    return cexp(base + es * z2 - er * zw) - cexp(base + es * z1 - er * zw);
}
static xo_c dbot(const xo_section *sec, const xo_c *eta, int nlam, int m, double zw, double z1, double z2, int r, int s)
{
    int t = r < s ? r : s, b = r > s ? r : s;
    double dt = layer_depth(sec, t), db1 = layer_depth(sec, b + 1);
    xo_c er = eta[r * nlam + m], es = eta[s * nlam + m], et = eta[t * nlam + m], eb = eta[b * nlam + m];
    xo_c base = -2 * (eb * db1 - et * dt);
    return cexp(base - es * z1 + er * zw) - cexp(base - es * z2 + er * zw);
}

/* which: bit0 U1, bit1 U2, bit2 U3, bit3 U4, bit4 U5 */
static void point_u(const xo_section *sec, double omega, const item_aux *a, int m, double lambda, int r, int s,
                    double zw, double z1, double z2, int sigma, int which, xo_c *u)
{
    int nlam = a->nlam;
    const xo_c *P = sigma ? a->ps : a->p1, *Q = sigma ? a->qs : a->q1, *eta = a->eta;
    int t = r < s ? r : s, b = r > s ? r : s;
    xo_c ann = ann_at(sec, omega, a, P, Q, r, s, m, sigma);
    xo_c er = eta[r * nlam + m], es = eta[s * nlam + m], et = eta[t * nlam + m], eb = eta[b * nlam + m];
    double dt = layer_depth(sec, t), db1 = layer_depth(sec, b + 1);
    /* CalculateT / CalculateR */
    xo_c T = cexp(-er * zw - es * z1 + 2 * et * dt) - cexp(-er * zw - es * z2 + 2 * et * dt);
    xo_c R = cexp(-2 * eb * db1 + es * z2 + er * zw) - cexp(-2 * eb * db1 + es * z1 + er * zw);
    xo_c pt = P[t * nlam + m], qb = Q[b * nlam + m];
    int kase = (zw > z1 && zw < z2) ? 0 : (zw >= z2 ? 1 : 2);   /* CalcUGeneral: match, below, above */
    xo_c psi1, psi3;
    if (kase == 0) {
        int n = r;
        xo_c pq = P[n * nlam + m] * Q[n * nlam + m];
        psi1 = ctop(eta, nlam, m, zw, zw, z2, n, n) + cbot(eta, nlam, m, zw, z1, zw, n, n) +
               pq * (dtop(sec, eta, nlam, m, zw, zw, z2, n, n) + dbot(sec, eta, nlam, m, zw, z1, zw, n, n));
        psi3 = ctop(eta, nlam, m, zw, zw, z2, n, n) - cbot(eta, nlam, m, zw, z1, zw, n, n) +
               pq * (-dtop(sec, eta, nlam, m, zw, zw, z2, n, n) + dtop(sec, eta, nlam, m, zw, z1, zw, n, n));
    } else if (kase == 1) {
        psi1 = cbot(eta, nlam, m, zw, z1, z2, r, s) + pt * qb * dbot(sec, eta, nlam, m, zw, z1, z2, r, s);
        psi3 = -cbot(eta, nlam, m, zw, z1, z2, r, s) + pt * qb * dbot(sec, eta, nlam, m, zw, z1, z2, r, s);
    } else {
        psi1 = ctop(eta, nlam, m, zw, z1, z2, r, s) + pt * qb * dtop(sec, eta, nlam, m, zw, z1, z2, r, s);
        psi3 = ctop(eta, nlam, m, zw, z1, z2, r, s) - pt * qb * dtop(sec, eta, nlam, m, zw, z1, z2, r, s);
    }
    if (which & 1) u[0] = (ann / es) * (pt * T + qb * R + psi1);                 /* CalculateU1And5Special */
    if (which & 2) u[1] = (ann * er) * (pt * T + qb * R - psi1);                 /* CalculateU2Special */
    if (which & 4) u[2] = (ann * (er / es)) * (-pt * T + qb * R + psi3);         /* CalculateU3Special */
    if (which & 8) u[3] = ann * (-pt * T + qb * R - psi3);                       /* CalculateU4Special */
    if (which & 16) {
        if (kase == 0) {                                                         /* CalculateU5SpecialForTrAndRcMatch */
            int n = r;
            double ln = sec->thickness[n];
            xo_c kk = I * omega * xo_zeta(sec->sigma[n], omega) * xo_mu0();
            xo_c en = eta[n * nlam + m], pq = P[n * nlam + m] * Q[n * nlam + m];
            xo_c psi5 = pq * (cexp(-en * (2 * ln - (z2 - zw))) + cexp(-en * (2 * ln - (zw - z1)))) -
                        (cexp(-en * (z2 - zw)) + cexp(-en * (zw - z1))) +
                        2 * kk * ((1 - pq * cexp(-2 * en * ln)) / (lambda * lambda));
            u[4] = (ann / es) * (pt * T + qb * R + psi5);
        } else {
            u[4] = (ann / es) * (pt * T + qb * R + psi1);
        }
    }
}

/* HankelCoefficients.ConvoluteWith (:86-99), the non-fast form used by PlanCalculator:
 * fr[i] = sum_{j ascending} fk[i + 430 - j] * hank[j] */
static xo_c hankel_dot_ascending(const double *hank, const xo_c *fk, int i)
{
    double re = 0, im = 0;
    for (int j = 0; j < NTAP; j++) {
        re += creal(fk[i + NTAP - 1 - j]) * hank[j];
        im += cimag(fk[i + NTAP - 1 - j]) * hank[j];
    }
    return re + I * im;
}

/* Scalars I1..I5 (electric) or I1..I4 (magnetic) of the nz transmitter layers seen at the level
 * depth zl, on the merged rho grid: out[c][tr][rho] (PlanCalculator.CalculateForE/H :58-141,
 * ScalarMathUtils.NormalizeAndPerformHankel :76-92, NormalizeBefore/AfterHankelForE/H). */
static int atoo_scalars(const xo_section *sec, const xo_grid *g, double omega, const double *hank0,
                        const double *hank1, const xo_plan *plan, double zl, int electric, xo_c *out)
{
    int nz = g->nz, nrho = plan->nrho_merged;
    int ncomp = electric ? 5 : 4;
    int r = xo_layer_at_depth(sec, zl);
    if (r < 0) return -1;
    double mu0 = xo_mu0();
    xo_c zrc2 = 2 * xo_zeta(sec->sigma[r], omega);
    xo_c iwm = I * omega * mu0;
    item_aux *aux = malloc(sizeof(item_aux) * plan->nitems);
    for (int it = 0; it < plan->nitems; it++) item_aux_build(sec, omega, plan->lam[it], plan->nlam[it], &aux[it]);
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int tr = 0; tr < nz; tr++) {
        double z1 = g->depth[tr], z2 = g->depth[tr] + g->thick[tr];
        int s = xo_layer_at_depth(sec, z1 + (z2 - z1) / 2);   /* transmitter working depth */
        if (s < 0) { err = 1; continue; }
        for (int it = 0; it < plan->nitems; it++) {
            int nlam = plan->nlam[it], nr = plan->nrho[it];
            const double *lam = plan->lam[it];
            xo_c *U = malloc(sizeof(xo_c) * (size_t)nlam * ncomp);   /* [comp][lam], comp = I index - 1 */
            for (int m = 0; m < nlam; m++) {
                xo_c u1[5], us[5];
                if (electric) {
                    point_u(sec, omega, &aux[it], m, lam[m], r, s, zl, z1, z2, 0, 1, u1);
                    point_u(sec, omega, &aux[it], m, lam[m], r, s, zl, z1, z2, 1, 2 | 4 | 8 | 16, us);
                    U[0 * nlam + m] = u1[0] * (iwm / 2);                        /* I1 <- U1 (A1 set) */
                    U[1 * nlam + m] = us[1] / zrc2;                             /* I2 <- U2 */
                    U[2 * nlam + m] = us[2] * (lam[m] / zrc2);                  /* I3 <- U3 */
                    U[3 * nlam + m] = us[3] * (-lam[m] / zrc2);                 /* I4 <- U4 */
                    U[4 * nlam + m] = us[4] * ((lam[m] * lam[m]) / zrc2);       /* I5 <- U5 */
                } else {
                    point_u(sec, omega, &aux[it], m, lam[m], r, s, zl, z1, z2, 0, 1 | 4, u1);
                    point_u(sec, omega, &aux[it], m, lam[m], r, s, zl, z1, z2, 1, 1 | 8, us);
                    U[0 * nlam + m] = u1[2];                                    /* I1 <- U31 */
                    U[1 * nlam + m] = us[3];                                    /* I2 <- U4Sigma */
                    U[2 * nlam + m] = us[0] * lam[m];                           /* I3 <- U1Sigma * lambda */
                    U[3 * nlam + m] = u1[0] * lam[m];                           /* I4 <- U11 * lambda */
                }
            }
            xo_c *Ir = malloc(sizeof(xo_c) * (size_t)nr * ncomp);
            for (int i = 0; i < nr; i++) {
                double rho = plan->rho[it][i];
                if (electric) {
                    double d = 2 * XO_PI * rho;
                    Ir[0 * nr + i] = -hankel_dot_ascending(hank1, U + 0 * nlam, i) / d;
                    Ir[1 * nr + i] = hankel_dot_ascending(hank1, U + 1 * nlam, i) / d;
                    Ir[2 * nr + i] = hankel_dot_ascending(hank0, U + 2 * nlam, i) / d;
                    Ir[3 * nr + i] = hankel_dot_ascending(hank0, U + 3 * nlam, i) / d;
                    Ir[4 * nr + i] = hankel_dot_ascending(hank1, U + 4 * nlam, i) / d;
                } else {
                    double d = 2 * (2 * XO_PI) * rho;
                    Ir[0 * nr + i] = -hankel_dot_ascending(hank1, U + 0 * nlam, i) / d;
                    Ir[1 * nr + i] = -hankel_dot_ascending(hank1, U + 1 * nlam, i) / d;
                    Ir[2 * nr + i] = -hankel_dot_ascending(hank0, U + 2 * nlam, i) / d;
                    Ir[3 * nr + i] = hankel_dot_ascending(hank0, U + 3 * nlam, i) / d;
                }
            }
            free(U);
            for (int k = 0; k < nrho; k++)
                if (plan->merge_item[k] == it)
                    for (int c = 0; c < ncomp; c++)
                        out[((int64_t)c * nz + tr) * nrho + k] = Ir[c * nr + plan->merge_index[k]];
            /* CalculateZeroRhoForE/H (ScalarMathUtils.cs:33-81) on the first item's results: I1, I2
             * (and I5) are 0 at rho = 0, I3 and I4 the value a + b rho^2 through its 2nd and 3rd
             * radii gives there: (f[1] a^2 - f[2]) / (a^2 - 1), a the rho step. */
            if (plan->zero_rho && it == 0) {
                double a2 = log10_rho_step() * log10_rho_step();
                for (int c = 0; c < ncomp; c++)
                    out[((int64_t)c * nz + tr) * nrho + 0] =
                        (c == 2 || c == 3) ? (Ir[c * nr + 1] * a2 - Ir[c * nr + 2]) / (a2 - 1) : 0;
            }
            free(Ir);
        }
    }
    for (int it = 0; it < plan->nitems; it++) item_aux_free(&aux[it]);
    free(aux);
    return err ? -1 : 0;
}

int xo_atoo_level_scalars(const xo_section *sec, const xo_grid *g, double omega, int number_of_hankels,
                          const double *hank0, const double *hank1, double zl, double rho_min, double rho_max,
                          int zero_rho, int electric, xo_c *out, int nrho_cap, double *rho_out)
{
    xo_plan plan;
    if (xo_plan_create_range(number_of_hankels, rho_min, rho_max, zero_rho, &plan)) return -1;
    int n = plan.nrho_merged;
    if (n > nrho_cap) { xo_plan_free(&plan); return -n; }
    memcpy(rho_out, plan.rho_merged, sizeof(double) * n);
    int rc = atoo_scalars(sec, g, omega, hank0, hank1, &plan, zl, electric, out);
    xo_plan_free(&plan);
    return rc ? -1 : n;
}

/* component order of the AtoO tensors */
enum { AO_XX, AO_YY, AO_ZZ, AO_XY, AO_YX, AO_XZ, AO_YZ, AO_ZX, AO_ZY, AO_N };

/* QSigns (QSign.cs:31-38) in AO order; [AlongX, AlongY, AlongXSymm, AlongYSymm] */
static const int AO_QS_E[4][AO_N] = {
    {1, -1, -1, -1, 0, 1, 1, 1, 1},
    {-1, 1, -1, -1, 0, 1, 1, 1, 1},
    {1, -1, -1, 1, 0, -1, 1, -1, 1},
    {-1, 1, -1, 1, 0, 1, -1, 1, -1},
};
static const int AO_QS_H[4][AO_N] = {
    {-1, -1, 0, -1, 1, 1, 1, 1, 1},
    {-1, -1, 0, 1, -1, 1, 1, 1, 1},
    {1, 1, 0, -1, 1, 1, -1, 1, -1},
    {1, 1, 0, 1, -1, 1, 1, -1, 1},
};

typedef struct {
    int electric, nz, nseg, ncomp;
    const xo_c *segA, *segB;      /* [seg][comp][tr] */
    const double *radii;
    xo_c *q[AO_N];                /* [nz] each; NULL when the component is not computed */
    xo_c *G[AO_N];                /* [2Nx][2Ny][nz] */
    int nx, ny;
} actx;

#define SEGV(arr, seg, comp, tr) (arr)[((int64_t)(seg) * c->ncomp + (comp)) * c->nz + (tr)]

/* AddToQBuffAlong{X,Y}{Electric,Magnetic} (GreenTensorCalculator.cs:505-560) */
static void ao_add_seg(actx *c, const factors *f, int seg, int alongY)
{
    for (int p = 0; p < c->nz; p++) {
#define ADD(comp, Ic, DA, DB) c->q[comp][p] += f->DA * SEGV(c->segA, seg, (Ic) - 1, p), \
                              c->q[comp][p] += f->DB * SEGV(c->segB, seg, (Ic) - 1, p)
        if (c->electric) {
            if (!alongY) {
                ADD(AO_XX, 1, d2a, d2b); ADD(AO_YY, 2, d2a, d2b); ADD(AO_ZZ, 5, d2a, d2b);
                ADD(AO_XY, 2, d3a, d3b); ADD(AO_YZ, 3, d1a, d1b); ADD(AO_ZY, 4, d1a, d1b);
            } else {
                ADD(AO_XX, 2, d2a, d2b); ADD(AO_YY, 1, d2a, d2b); ADD(AO_ZZ, 5, d2a, d2b);
                ADD(AO_XY, 1, d3a, d3b); ADD(AO_XZ, 3, d1a, d1b); ADD(AO_ZX, 4, d1a, d1b);
            }
        } else {
            if (!alongY) {
                ADD(AO_XX, 2, d3a, d3b); ADD(AO_YX, 1, d2a, d2b); ADD(AO_XY, 2, d2a, d2b);
                ADD(AO_XZ, 3, d1a, d1b); ADD(AO_ZX, 4, d1a, d1b);
            } else {
                ADD(AO_YX, 2, d2a, d2b); ADD(AO_XY, 1, d2a, d2b); ADD(AO_XX, 1, d3a, d3b);
                ADD(AO_YZ, 3, d1a, d1b); ADD(AO_ZY, 4, d1a, d1b);
            }
        }
#undef ADD
    }
}

static void ao_integrate(actx *c, knot left, knot right, double s, double u, double v, int alongY)
{
    for (int t = 0; t < AO_N; t++)
        if (c->q[t]) memset(c->q[t], 0, sizeof(xo_c) * c->nz);
    int i1 = left.index, i2 = right.index;
    if (i1 == i2) {
        factors f = prep_factors(s, u, left.radius, right.radius, v);
        ao_add_seg(c, &f, i1, alongY);
    } else {
        double r1 = left.radius, r2 = c->radii[i1 + 1];
        double sj = s, uj = sqrt(r2 * r2 - v * v);
        factors f = prep_factors(sj, uj, r1, r2, v);
        ao_add_seg(c, &f, i1, alongY);
        for (int i = i1 + 1; i < i2; i++) {
            r1 = r2; r2 = c->radii[i + 1];
            sj = uj; uj = sqrt(r2 * r2 - v * v);
            f = prep_factors(sj, uj, r1, r2, v);
            ao_add_seg(c, &f, i, alongY);
        }
        sj = uj; uj = u; r1 = r2; r2 = right.radius;
        f = prep_factors(sj, uj, r1, r2, v);
        ao_add_seg(c, &f, i2, alongY);
    }
}

/* FillGreen with AtoOLevelGreenTensorCalculator.TransformI/J (:106-117): negative offsets wrap */
static void ao_fill(actx *c, int i, int j, int sign, int qsel)
{
    int ii = i < 0 ? 2 * c->nx + i % (2 * c->nx) : i;
    int jj = j < 0 ? 2 * c->ny + j % (2 * c->ny) : j;
    int64_t shift = ((int64_t)ii * 2 * c->ny + jj) * c->nz;
    const int (*qs)[AO_N] = c->electric ? AO_QS_E : AO_QS_H;
    for (int t = 0; t < AO_N; t++) {
        if (!c->q[t]) continue;
        double al = (double)(sign * qs[qsel][t]);
        for (int p = 0; p < c->nz; p++) c->G[t][shift + p] += al * c->q[t][p];
    }
}

static void ao_fill_x(actx *c, int i, int j, int left, int right, int qsel)
{
    if (j == left) ao_fill(c, i, left, -1, qsel);
    else if (j == right) ao_fill(c, i, right - 1, 1, qsel);
    else { ao_fill(c, i, j - 1, 1, qsel); ao_fill(c, i, j, -1, qsel); }
}

static void ao_fill_y(actx *c, int i, int j, int left, int right, int qsel)
{
    if (i == left) ao_fill(c, left, j, -1, qsel);
    else if (i == right) ao_fill(c, right - 1, j, 1, qsel);
    else { ao_fill(c, i - 1, j, 1, qsel); ao_fill(c, i, j, -1, qsel); }
}

int xo_green_atoo_level(const xo_section *sec, const xo_grid *g, double omega, int number_of_hankels,
                        const double *hank0, const double *hank1, double zl, double sx, double sy, int electric,
                        int fft, xo_c **out)
{
    /* CreateForAnomalyToObservationLevels (ScalarPlansCreater.cs:51-66) for this level alone
     * (AtoOCalculator.cs:129, 151) */
    double rmin, rmax;
    int zero = xo_atoo_level_rho_range(g, sx, sy, &rmin, &rmax);
    return xo_green_atoo_level_plan(sec, g, omega, number_of_hankels, hank0, hank1, zl, sx, sy, rmin, rmax, zero,
                                    electric, fft, out);
}

int xo_green_atoo_level_plan(const xo_section *sec, const xo_grid *g, double omega, int number_of_hankels,
                             const double *hank0, const double *hank1, double zl, double sx, double sy,
                             double rho_min, double rho_max, int zero_rho, int electric, int fft, xo_c **out)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    xo_plan plan;
    if (xo_plan_create_range(number_of_hankels, rho_min, rho_max, zero_rho, &plan)) return -1;
    int nrho = plan.nrho_merged, nseg = nrho - 1;
    int ncomp = electric ? 5 : 4;
    xo_c *Isc = malloc(sizeof(xo_c) * (size_t)ncomp * nz * nrho);
    if (atoo_scalars(sec, g, omega, hank0, hank1, &plan, zl, electric, Isc)) {
        free(Isc); xo_plan_free(&plan); return -2;
    }
    xo_c *segA = malloc(sizeof(xo_c) * (size_t)nseg * ncomp * nz);
    xo_c *segB = malloc(sizeof(xo_c) * (size_t)nseg * ncomp * nz);
    const double *radii = plan.rho_merged;
    for (int k = 0; k < nseg; k++) {                      /* ScalarSegments.CalculateScalarSegment */
        double s = radii[k], u = radii[k + 1], div = u - s;
        for (int cc = 0; cc < ncomp; cc++)
            for (int p = 0; p < nz; p++) {
                xo_c i0 = Isc[((int64_t)cc * nz + p) * nrho + k], i1 = Isc[((int64_t)cc * nz + p) * nrho + k + 1];
                xo_c a = (i1 - i0) / div;
                segA[((int64_t)k * ncomp + cc) * nz + p] = a;
                segB[((int64_t)k * ncomp + cc) * nz + p] = i0 - a * s;
            }
    }
    free(Isc);
    /* which components: electric all but yx; magnetic all but yy, zz (AtoOLevelGreenTensorCalculator.cs:60-75) */
    int have[AO_N];
    for (int t = 0; t < AO_N; t++) have[t] = 1;
    if (electric) have[AO_YX] = 0; else { have[AO_YY] = 0; have[AO_ZZ] = 0; }
    int64_t csize = (int64_t)4 * nx * ny * nz;
    for (int t = 0; t < AO_N; t++)
        if (out[t]) memset(out[t], 0, sizeof(xo_c) * csize);
    /* knots over offsets [-Nx+1, Nx] x [-Ny+1, Ny] (PrepareKnotsAtoO, TensorPlan serial case:
     * DistributedUtils.cs:54-61); the edge coordinates are dx (i - 1/2) - xShift with
     * xShift = -ShiftAlongX (AtoOLevelGreenTensorCalculator.cs:119-125, TensorPlan.cs:44-54) */
    const double xs = -sx, ys = -sy;
    int lx = -nx + 1, rx = nx, ly = -ny + 1, ry = ny;
    int kx = rx - lx + 1, ky = ry - ly + 1;
    knot *K = malloc(sizeof(knot) * kx * ky), *ZX = malloc(sizeof(knot) * ky), *ZY = malloc(sizeof(knot) * kx);
    for (int i = lx; i <= rx; i++)
        for (int j = ly; j <= ry; j++)
            K[(i - lx) * ky + (j - ly)] = make_knot(radii, nrho, g->dx * (i - 0.5) - xs, g->dy * (j - 0.5) - ys);
    for (int j = ly; j <= ry; j++) ZX[j - ly] = make_knot(radii, nrho, 0, g->dy * (j - 0.5) - ys);
    for (int i = lx; i <= rx; i++) ZY[i - lx] = make_knot(radii, nrho, g->dx * (i - 0.5) - xs, 0);
#define KN(i, j) K[((i) - lx) * ky + ((j) - ly)]
    actx c;
    memset(&c, 0, sizeof(c));
    c.electric = electric; c.nz = nz; c.nseg = nseg; c.ncomp = ncomp; c.segA = segA; c.segB = segB;
    c.radii = radii; c.nx = nx; c.ny = ny;
    for (int t = 0; t < AO_N; t++) {
        c.q[t] = have[t] ? calloc(nz, sizeof(xo_c)) : NULL;
        c.G[t] = have[t] ? out[t] : NULL;
        if (have[t] && !out[t]) { c.q[t] = NULL; }
    }
    double dx = g->dx, dy = g->dy;
    /* RunAlongX (GreenTensorCalculator.cs:192-230) */
    for (int j = ly; j <= ry; j++) {
        double y = dy * (j - 0.5) - ys;
        for (int i = lx; i <= rx - 1; i++) {
            double x1 = dx * (i - 0.5) - xs, x2 = dx * (i + 0.5) - xs;
            if (x1 > 0 && x2 > 0) {
                ao_integrate(&c, KN(i, j), KN(i + 1, j), x1, x2, y, 0);
                ao_fill_x(&c, i, j, ly, ry, 0);
            } else if (x1 < 0 && x2 < 0) {
                ao_integrate(&c, KN(i + 1, j), KN(i, j), -x2, -x1, y, 0);
                ao_fill_x(&c, i, j, ly, ry, 2);
            } else {
                ao_integrate(&c, ZX[j - ly], KN(i, j), 0, -x1, y, 0);
                ao_fill_x(&c, i, j, ly, ry, 2);
                ao_integrate(&c, ZX[j - ly], KN(i + 1, j), 0, x2, y, 0);
                ao_fill_x(&c, i, j, ly, ry, 0);
            }
        }
    }
    /* RunAlongY (:246-281) */
    for (int i = lx; i <= rx; i++) {
        double x = dx * (i - 0.5) - xs;
        for (int j = ly; j <= ry - 1; j++) {
            double y1 = dy * (j - 0.5) - ys, y2 = dy * (j + 0.5) - ys;
            if (y1 > 0 && y2 > 0) {
                ao_integrate(&c, KN(i, j), KN(i, j + 1), y1, y2, x, 1);
                ao_fill_y(&c, i, j, lx, rx, 1);
            } else if (y1 < 0 && y2 < 0) {
                ao_integrate(&c, KN(i, j + 1), KN(i, j), -y2, -y1, x, 1);
                ao_fill_y(&c, i, j, lx, rx, 3);
            } else {
                ao_integrate(&c, ZY[i - lx], KN(i, j), 0, -y1, x, 1);
                ao_fill_y(&c, i, j, lx, rx, 3);
                ao_integrate(&c, ZY[i - lx], KN(i, j + 1), 0, y2, x, 1);
                ao_fill_y(&c, i, j, lx, rx, 1);
            }
        }
    }
#undef KN
    for (int t = 0; t < AO_N; t++) free(c.q[t]);
    free(K); free(ZX); free(ZY); free(segA); free(segB);
    xo_plan_free(&plan);
    /* AtoOCalculator.PerformGreenTensorFft (:138-147) */
    if (fft)
        for (int t = 0; t < AO_N; t++)
            if (have[t] && out[t]) xo_fft2_batch(out[t], 2 * nx, 2 * ny, nz, -1);
    return 0;
}

/* ToOCalculator.CalculateAnomalyField (ToOCalculator.cs:21-45) for one level: FFT of jQ, the
 * per-wavenumber nz -> 1 contraction (ApplyGreenTensorAlongZForE/H :47-125), inverse FFT,
 * 1/(4 Nx Ny), crop. G[t]: AO-ordered FFT'd level tensors [2Nx][2Ny][nz] (NULL if absent).
 * field: [Nx][Ny][3]. */
void xo_atoo_field(int nx, int ny, int nz, xo_c **G, int electric, const xo_c *jq, xo_c *field)
{
    int64_t L = (int64_t)4 * nx * ny;
    xo_c *b1 = calloc((size_t)L * 3 * nz, sizeof(xo_c));
    xo_c *b2 = calloc((size_t)L * 3, sizeof(xo_c));
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            memcpy(b1 + ((int64_t)i * 2 * ny + j) * 3 * nz, jq + ((int64_t)i * ny + j) * 3 * nz, sizeof(xo_c) * 3 * nz);
    xo_fft2_batch(b1, 2 * nx, 2 * ny, 3 * nz, -1);
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < L; l++) {
        const xo_c *sx = b1 + l * 3 * nz, *sy = sx + nz, *sz = sy + nz;
        xo_c dxv = 0, dyv = 0, dzv = 0;
#define G_(t, k) G[t][l * nz + (k)]
        for (int k = 0; k < nz; k++) {
            if (electric) {
                dxv += G_(AO_XX, k) * sx[k] + G_(AO_XY, k) * sy[k] + G_(AO_XZ, k) * sz[k];
                dyv += G_(AO_XY, k) * sx[k] + G_(AO_YY, k) * sy[k] + G_(AO_YZ, k) * sz[k];
                dzv += G_(AO_ZX, k) * sx[k] + G_(AO_ZY, k) * sy[k] + G_(AO_ZZ, k) * sz[k];
            } else {
                dxv += -(G_(AO_XX, k) * sx[k] + G_(AO_XY, k) * sy[k] + G_(AO_XZ, k) * sz[k]);
                dyv += G_(AO_YX, k) * sx[k] + G_(AO_XX, k) * sy[k] + G_(AO_YZ, k) * sz[k];
                dzv += -G_(AO_ZX, k) * sx[k] + G_(AO_ZY, k) * sy[k];
            }
        }
#undef G_
        b2[l * 3 + 0] = dxv; b2[l * 3 + 1] = dyv; b2[l * 3 + 2] = dzv;
    }
    xo_fft2_batch(b2, 2 * nx, 2 * ny, 3, +1);
    double sc = 1.0 / ((double)nx * (double)ny * 4);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++)
            for (int cc = 0; cc < 3; cc++)
                field[((int64_t)i * ny + j) * 3 + cc] = b2[((int64_t)i * 2 * ny + j) * 3 + cc] * sc;
    free(b1); free(b2);
}
