// Device assembly of the anomaly-to-anomaly (AtoA) electric Green tensor in FP64.
//
// Stages (reference files in brackets):
//   host  scalar plans: NumberOfHankels interleaved log-rho grids and their lambda grids
//         [Green/Scalar/ScalarPlansCreater.cs:102-176, ScalarPlan.cs:76-97, HankelCoefficients.cs:67-76]
//   K1    per (plan item, lambda): eta, exp(-2 eta h), the P/Q reflection recursions for the
//         "1" and "Sigma" coefficient sets  [PlanCalculatorFast.cs:203-318, AlphaBeta.cs:24-42]
//   K2    per (transceiver pair, item, lambda): the vertical integrals T, R, C, D, the transfer
//         coefficient A[rc][tr] and U1..U5, normalised for the Hankel transform
//         [VolumeToVolumeCalculatorUFast.cs:34-337, green_scalar.c:86-386, ScalarMathUtils.cs:243-263]
//   K3    431-tap Hankel filter onto the merged rho grid, normalised by 1/(2 pi rho)
//         [HankelCoefficients.cs:103-118, ScalarMathUtils.cs:189-241, InnerResultsCombiner.cs:19-62]
//   K4    piecewise-linear segments I(rho) = A rho + B  [ScalarSegments.cs:86-128]
//   K5    lateral integration of every cell of the positive quadrant by line integrals along
//         its four edges, written with the mirror parities straight into the device layout
//         [GreenTensorCalculator.cs:166-552, IntegralFactors.cs:17-30, QSign.cs:64-67,
//          GreenTensorFftPopulator.cs:37-148]
//   FFT   forward 2-D pencil FFT of every column  [AtoAGreenTensorCalculatorComponent.cs:197-279]
//
// Threads of K5 own one transceiver pair each and share the cell, so the edge geometry, the
// knot search and the segment walk are uniform across a wavefront while the per-pair segment
// coefficients are read coalesced.
#include "xtr_common.h"
#include "xtr_green.h"
#include "xtr_kernels.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

namespace xtr {

namespace {

constexpr int kTaps = 431;
constexpr int kL2 = 130;          // HankelCoefficients._l2: lambda index of 1/rho[0]
constexpr int kMaxLayers = 16;
constexpr double kPi = 3.14159265358979323846;
constexpr double kMu0 = 4.0 * kPi * 1.0E-07;

// ---- complex helpers (FP64) ---------------------------------------------------------------
__device__ inline double2 cexp_d(double2 z)
{
    double s, c;
    sincos(z.y, &s, &c);
    const double e = exp(z.x);
    return make_double2(e * c, e * s);
}

// principal square root
__device__ inline double2 csqrt_d(double2 z)
{
    const double a = z.x, b = z.y;
    if (a == 0 && b == 0) return make_double2(0, b);
    const double r = hypot(a, b);
    if (a >= 0) {
        const double t = sqrt(0.5 * (r + a));
        return make_double2(t, b / (2 * t));
    }
    const double t = sqrt(0.5 * (r - a));
    return make_double2(fabs(b) / (2 * t), copysign(t, b));
}

// a / b (Smith)
__device__ inline double2 cdiv_d(double2 a, double2 b)
{
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, den = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    const double r = b.x / b.y, den = b.x * r + b.y;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}

__device__ inline double2 c_(double re, double im = 0) { return make_double2(re, im); }
__device__ inline double2 operator+(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 operator-(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 operator-(double2 a) { return make_double2(-a.x, -a.y); }
__device__ inline double2 operator*(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 operator*(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ inline double2 operator/(double2 a, double2 b) { return cdiv_d(a, b); }

// ---- K1: per (item, lambda) layer recursions ----------------------------------------------
struct LayerModel {
    int nl;
    double2 zeta[kMaxLayers];
    double thick[kMaxLayers];
    double depth[kMaxLayers + 1];   // depth[k] = top of layer k
    double omega;
};

__global__ void layer_aux_kernel(LayerModel lm, int nitems, int nlam_max, const int *nlam, const double *lam,
                                 double2 *eta, double2 *ex, double2 *p1, double2 *q1, double2 *ps, double2 *qs)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    if (it >= nitems || m >= nlam[it]) return;
    const int nl = lm.nl;
    const double l = lam[(int64_t)it * nlam_max + m];
    const int64_t base = ((int64_t)it * nlam_max + m) * nl;
    double2 E[kMaxLayers], X[kMaxLayers];
    for (int k = 0; k < nl; ++k) {
        // value = i omega * zeta * Mu0; eta = sqrt(lambda^2 - value)  (CalcEta, green_scalar.c:30-34)
        const double2 value = (c_(0, lm.omega) * lm.zeta[k]) * c_(kMu0);
        E[k] = csqrt_d(c_(l * l) - value);
        X[k] = cexp_d((-2 * lm.thick[k]) * E[k]);    // CalcExp (green_scalar.c:36-41)
        eta[base + k] = E[k];
        ex[base + k] = X[k];
    }
    // P: p[1] from the top interface, then downwards (CalculatePBelow1, PlanCalculatorFast.cs:266-280)
    double2 P1[kMaxLayers], PS[kMaxLayers], Q1[kMaxLayers], QS[kMaxLayers];
    for (int k = 0; k < nl; ++k) P1[k] = PS[k] = Q1[k] = QS[k] = c_(0);
    P1[1] = (E[1] - E[0]) / (E[1] + E[0]);
    PS[1] = (lm.zeta[0] * E[1] - lm.zeta[1] * E[0]) / (lm.zeta[0] * E[1] + lm.zeta[1] * E[0]);
    for (int k = 1; k < nl - 1; ++k) {
        double2 pe = P1[k] * X[k];
        double2 part = (c_(1) * (E[k] / E[k + 1])) * ((pe - c_(1)) / (pe + c_(1)));
        P1[k + 1] = (c_(1) + part) / (c_(1) - part);
        pe = PS[k] * X[k];
        part = ((lm.zeta[k + 1] / lm.zeta[k]) * (E[k] / E[k + 1])) * ((pe - c_(1)) / (pe + c_(1)));
        PS[k + 1] = (c_(1) + part) / (c_(1) - part);
    }
    // Q: q[n] = 0 at the bottom, then upwards (CalculateQ, PlanCalculatorFast.cs:292-318)
    const int n = nl - 1;
    double2 part = c_(1) * (E[n] / E[n - 1]);
    Q1[n - 1] = (c_(1) - part) / (c_(1) + part);
    part = (lm.zeta[n - 1] / lm.zeta[n]) * (E[n] / E[n - 1]);
    QS[n - 1] = (c_(1) - part) / (c_(1) + part);
    for (int k = n - 1; k > 1; --k) {
        double2 qe = Q1[k] * X[k];
        part = (c_(1) * (E[k] / E[k - 1])) * ((qe - c_(1)) / (qe + c_(1)));
        Q1[k - 1] = (c_(1) + part) / (c_(1) - part);
        qe = QS[k] * X[k];
        part = ((lm.zeta[k - 1] / lm.zeta[k]) * (E[k] / E[k - 1])) * ((qe - c_(1)) / (qe + c_(1)));
        QS[k - 1] = (c_(1) + part) / (c_(1) - part);
    }
    for (int k = 0; k < nl; ++k) {
        p1[base + k] = P1[k];
        q1[base + k] = Q1[k];
        ps[base + k] = PS[k];
        qs[base + k] = QS[k];
    }
}

// ---- K2: U1..U5 per (pair, item, lambda) ----------------------------------------------------
struct PairInfo {
    int r, s, t, b, kase;     // background layers of rc/tr, min/max; 0 match, 1 rc below, 2 rc above
    double r1, r2, s1, s2, dt, db1;
};

__device__ inline double2 e4(double2 a, double2 b, double2 c, double2 d)
{
    return cexp_d(a) - cexp_d(b) - cexp_d(c) + cexp_d(d);
}

// A[rc][tr] (CalculateA :160-200; CalculateASigma :138-158 for sigma = 1)
__device__ inline double2 ann_at(const LayerModel &lm, const double2 *E, const double2 *X, const double2 *P,
                                 const double2 *Q, int rc, int tr, int sigma)
{
    const int smax = rc > tr ? rc : tr, rmin = rc < tr ? rc : tr;
    double2 prev = c_(1) / (E[smax] * (c_(1) - P[smax] * Q[smax] * X[smax]));
    for (int k = smax - 1; k >= rmin; --k) {
        const double2 q = ((c_(1) + P[k + 1]) / (c_(1) + P[k] * X[k])) * cexp_d(lm.depth[k + 1] * (E[k + 1] - E[k]));
        prev = q * prev;
    }
    if (sigma && rc > tr) return (lm.zeta[rc] / lm.zeta[tr]) * prev;
    return prev;
}

__global__ void pair_u_kernel(LayerModel lm, int npairs, const PairInfo *pairs, int asym, int nitems,
                              int nlam_max, const int *nlam, const double *lam, const double2 *eta,
                              const double2 *ex, const double2 *p1, const double2 *q1, const double2 *ps,
                              const double2 *qs, double2 *U)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    const int p = blockIdx.z;
    if (it >= nitems || m >= nlam[it] || p >= npairs) return;
    const PairInfo pi = pairs[p];
    const int nl = lm.nl;
    const int64_t base = ((int64_t)it * nlam_max + m) * nl;
    const double2 *E = eta + base, *X = ex + base;
    const double l = lam[(int64_t)it * nlam_max + m];
    const double2 er = E[pi.r], es = E[pi.s], et = E[pi.t], eb = E[pi.b];
    // PreCalculate (VolumeToVolumeCalculatorUFast.cs:34-46)
    const double2 T = e4(-pi.r1 * er - pi.s1 * es + 2 * pi.dt * et, -pi.r1 * er - pi.s2 * es + 2 * pi.dt * et,
                         -pi.r2 * er - pi.s1 * es + 2 * pi.dt * et, -pi.r2 * er - pi.s2 * es + 2 * pi.dt * et);
    const double2 Rv = e4(-2 * pi.db1 * eb + pi.s2 * es + pi.r2 * er, -2 * pi.db1 * eb + pi.s1 * es + pi.r2 * er,
                          -2 * pi.db1 * eb + pi.s2 * es + pi.r1 * er, -2 * pi.db1 * eb + pi.s1 * es + pi.r1 * er);
    double2 C = c_(0), D = c_(0);
    const double2 F = -2 * pi.db1 * eb + 2 * pi.dt * et;
    if (pi.kase == 2) {   // rc above tr: Ctop, Dtop
        C = e4(-pi.s1 * es + pi.r2 * er, -pi.s2 * es + pi.r2 * er, -pi.s1 * es + pi.r1 * er, -pi.s2 * es + pi.r1 * er);
        D = e4(F + pi.s2 * es - pi.r1 * er, F + pi.s1 * es - pi.r1 * er, F + pi.s2 * es - pi.r2 * er,
               F + pi.s1 * es - pi.r2 * er);
    } else if (pi.kase == 1) {   // rc below tr: Cbot, Dbot
        C = e4(-pi.r1 * er + pi.s2 * es, -pi.r1 * er + pi.s1 * es, -pi.r2 * er + pi.s2 * es, -pi.r2 * er + pi.s1 * es);
        D = e4(F - pi.s1 * es + pi.r2 * er, F - pi.s2 * es + pi.r2 * er, F - pi.s1 * es + pi.r1 * er,
               F - pi.s2 * es + pi.r1 * er);
    }
    const double2 zrc2 = 2.0 * lm.zeta[pi.r];
    const int ncomp = asym ? 1 : 3;
    double2 *out = U + (((int64_t)p * nitems + it) * ncomp) * nlam_max + m;
    const double2 fac = c_(1) / (er * es);
    for (int set = asym ? 1 : 0; set < 2; ++set) {
        const double2 *P = (set ? ps : p1) + base;
        const double2 *Q = (set ? qs : q1) + base;
        const double2 Pt = P[pi.t], Qb = Q[pi.b];
        const double2 ann = ann_at(lm, E, X, P, Q, pi.r, pi.s, set);
        double2 psi1, psi3;
        if (pi.kase == 0) {
            const int n = pi.r;
            const double tau = pi.r2 - pi.r1;
            const double dn1 = lm.depth[n + 1] - lm.depth[n];
            const double2 en = E[n];
            const double2 PQ = P[n] * Q[n];
            // CalcPsi1RcAndTrMatch (:313-332)
            const double2 part1 = tau * en - (c_(1) - cexp_d(-tau * en));
            const double2 part2 = PQ * ((tau * en + c_(1)) * cexp_d(-2 * dn1 * en) - cexp_d(-2 * dn1 * en + tau * en));
            psi1 = 2.0 * (part1 - part2);
            psi3 = c_(0);
        } else if (pi.kase == 1) {
            psi1 = C + Pt * Qb * D;
            psi3 = -C + Pt * Qb * D;
        } else {
            psi1 = C + Pt * Qb * D;
            psi3 = C - Pt * Qb * D;
        }
        if (set == 0) {
            // U1 (:144-159), normalised by i omega Mu0 / 2
            const double2 u1 = ann * fac * (Pt * T + Qb * Rv + psi1);
            out[0] = u1 * (c_(0, lm.omega * kMu0) * c_(0.5));
            continue;
        }
        if (asym) {
            // U3 (:195-210), normalised by lambda / (2 zeta_rc)
            const double2 u3 = ann * (c_(1) / es) * (-1.0 * (Pt * T) + Qb * Rv + psi3);
            out[0] = u3 * (c_(l) / zrc2);
            return;
        }
        // U2 (:183-193), normalised by 1 / (2 zeta_rc)
        const double2 u2 = ann * (Pt * T + Qb * Rv - psi1);
        out[(int64_t)1 * nlam_max] = u2 / zrc2;
        // U5 (:161-181, 290-311), normalised by lambda^2 / (2 zeta_rc)
        double2 u5;
        if (pi.kase == 0) {
            const int n = pi.r;
            const double tau = pi.r2 - pi.r1;
            const double ln = lm.thick[n];
            const double2 kk = (c_(0, lm.omega) * lm.zeta[n]) * c_(kMu0);
            const double2 en = E[n];
            const double2 PQ = P[n] * Q[n];
            const double2 part1 = -(c_(1) - cexp_d(-tau * en));
            const double2 part2 = PQ * (cexp_d(-(2 * ln - tau) * en) - cexp_d(-2 * ln * en));
            const double2 part3 = (kk / c_(l * l)) * (c_(1) - PQ * cexp_d(-2 * ln * en)) * (tau * en);
            const double2 psi5 = 2.0 * (part1 + part2 + part3);
            u5 = (ann * fac) * (Pt * T + Qb * Rv + psi5);
        } else {
            u5 = ann * fac * (Pt * T + Qb * Rv + psi1);
        }
        out[(int64_t)2 * nlam_max] = u5 * (c_(l * l) / zrc2);
    }
}

// ---- K3: Hankel filter onto the merged rho grid -------------------------------------------
// I[comp][pair][k] = sgn * (sum_j hank[430-j] U[pair][item][comp][idx+j]) / (2 pi rho_k)
__global__ void hankel_kernel(int npairs, int nitems, int ncomp, int nlam_max, const double2 *U, int nmerged,
                              const int *m_item, const int *m_idx, const double *radii, const double *hank0,
                              const double *hank1, int asym, double2 *I)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    const int c = blockIdx.z;
    if (k >= nmerged || p >= npairs || c >= ncomp) return;
    const int it = m_item[k], idx = m_idx[k];
    const double *hank = asym ? hank0 : hank1;      // I3 with J0; I1, I2, I5 with J1
    const double2 *f = U + (((int64_t)p * nitems + it) * ncomp + c) * nlam_max + idx;
    double re = 0, im = 0;
    for (int j = 0; j < kTaps; ++j) {
        const double h = hank[kTaps - 1 - j];
        const double2 v = f[j];
        re += v.x * h;
        im += v.y * h;
    }
    const double den = 2 * kPi * radii[k];
    const double sgn = (!asym && c == 0) ? -1.0 : 1.0;   // I1 = -I1/(2 pi rho)
    I[((int64_t)c * npairs + p) * nmerged + k] = make_double2(sgn * re / den, sgn * im / den);
}

// ---- K4: segments --------------------------------------------------------------------------
__global__ void segments_kernel(int npairs, int ncomp, int nmerged, const double *radii, const double2 *I,
                                double2 *segA, double2 *segB)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int k = blockIdx.z;
    if (p >= npairs) return;
    const double s = radii[k], u = radii[k + 1], div = u - s;
    const double2 i0 = I[((int64_t)c * npairs + p) * nmerged + k];
    const double2 i1 = I[((int64_t)c * npairs + p) * nmerged + k + 1];
    const double2 a = make_double2((i1.x - i0.x) / div, (i1.y - i0.y) / div);
    const int64_t o = ((int64_t)k * ncomp + c) * npairs + p;
    segA[o] = a;
    segB[o] = make_double2(i0.x - a.x * s, i0.y - a.y * s);
}

// ---- K5: lateral integration, mirroring, device layout ---------------------------------------
struct Knot { int index; double radius; };

__device__ inline Knot make_knot(const double *radii, int nr, double x, double y)
{
    const double r = sqrt(x * x + y * y);
    int lo = 0, hi = nr - 1, idx = -1;
    while (lo <= hi) {
        const int mid = lo + (hi - lo) / 2;
        const double v = radii[mid];
        if (v == r) { idx = mid; break; }
        if (v < r) lo = mid + 1; else hi = mid - 1;
    }
    if (idx < 0) idx = lo - 1;
    if (idx == nr - 1) idx--;
    if (idx < 0) idx = 0;
    return Knot{idx, r};
}

struct Factors { double d1a, d1b, d2a, d2b, d3a, d3b; };

__device__ inline Factors factors(double s, double u, double rs, double ru, double v)
{
    Factors f;
    const double lg = log((u + ru) / (s + rs));
    f.d1a = v == 0 ? (u * ru - s * rs) / 2 : (u * ru - s * rs + v * v * lg) / 2;
    f.d1b = u - s;
    f.d2a = v == 0 ? 0 : -(u - s) * v;
    f.d2b = v == 0 ? 0 : -(v * lg);
    f.d3a = -(u * u - s * s) / 2;
    f.d3b = -(ru - rs);
    return f;
}

struct LatArgs {
    int nx, ny, nz, npairs, asym, nr;
    double dx, dy;
    const double2 *segA, *segB;   // [seg][comp][pair]
    const double *radii;
    double2 *G;
    int64_t gblk;
    int64_t off[4];               // component offsets in the per-wavenumber block, q-order
    int xs[4], ys[4];             // mirror parities, q-order
};

// q-order: asym {xz, yz}; symm {xx, yy, zz, xy}
struct Q4 { double2 q[4]; };

__device__ inline void add_seg(const LatArgs &a, int p, const Factors &f, int seg, bool alongY, Q4 &Q)
{
    const int np = a.npairs;
    if (a.asym) {
        const int64_t o = (int64_t)seg * np + p;
        const double2 A3 = a.segA[o], B3 = a.segB[o];
        double2 &q = alongY ? Q.q[0] : Q.q[1];       // AlongY -> Qxz, AlongX -> Qyz (D1)
        q = q + f.d1a * A3;
        q = q + f.d1b * B3;
        return;
    }
    const int64_t o1 = ((int64_t)seg * 3 + 0) * np + p, o2 = o1 + np, o5 = o2 + np;
    const double2 A1 = a.segA[o1], B1 = a.segB[o1], A2 = a.segA[o2], B2 = a.segB[o2];
    const double2 A5 = a.segA[o5], B5 = a.segB[o5];
    if (!alongY) {   // AddToQBuffAlongXElectric (GreenTensorCalculator.cs:505-516)
        Q.q[0] = Q.q[0] + f.d2a * A1; Q.q[0] = Q.q[0] + f.d2b * B1;
        Q.q[1] = Q.q[1] + f.d2a * A2; Q.q[1] = Q.q[1] + f.d2b * B2;
        Q.q[2] = Q.q[2] + f.d2a * A5; Q.q[2] = Q.q[2] + f.d2b * B5;
        Q.q[3] = Q.q[3] + f.d3a * A2; Q.q[3] = Q.q[3] + f.d3b * B2;
    } else {         // AddToQBuffAlongYElectric (:517-528)
        Q.q[0] = Q.q[0] + f.d2a * A2; Q.q[0] = Q.q[0] + f.d2b * B2;
        Q.q[1] = Q.q[1] + f.d2a * A1; Q.q[1] = Q.q[1] + f.d2b * B1;
        Q.q[2] = Q.q[2] + f.d2a * A5; Q.q[2] = Q.q[2] + f.d2b * B5;
        Q.q[3] = Q.q[3] + f.d3a * A1; Q.q[3] = Q.q[3] + f.d3b * B1;
    }
}

// Integrate (GreenTensorCalculator.cs:367-411)
__device__ inline Q4 integrate(const LatArgs &a, int p, Knot left, Knot right, double s, double u, double v, bool alongY)
{
    Q4 Q;
    for (int t = 0; t < 4; ++t) Q.q[t] = c_(0);
    const int i1 = left.index, i2 = right.index;
    if (i1 == i2) {
        add_seg(a, p, factors(s, u, left.radius, right.radius, v), i1, alongY, Q);
        return Q;
    }
    double r1 = left.radius, r2 = a.radii[i1 + 1];
    double sj = s, uj = sqrt(r2 * r2 - v * v);
    add_seg(a, p, factors(sj, uj, r1, r2, v), i1, alongY, Q);
    for (int i = i1 + 1; i < i2; ++i) {
        r1 = r2;
        r2 = a.radii[i + 1];
        sj = uj;
        uj = sqrt(r2 * r2 - v * v);
        add_seg(a, p, factors(sj, uj, r1, r2, v), i, alongY, Q);
    }
    sj = uj; uj = u; r1 = r2; r2 = right.radius;
    add_seg(a, p, factors(sj, uj, r1, r2, v), i2, alongY, Q);
    return Q;
}

// QSigns (QSign.cs:64-67) in q-order; rows AlongX, AlongY, AlongXSymm, AlongYSymm
__constant__ int kQsAsym[4][2] = {{1, 1}, {1, 1}, {-1, 1}, {1, -1}};
__constant__ int kQsSymm[4][4] = {{1, -1, -1, -1}, {-1, 1, -1, -1}, {1, -1, -1, 1}, {-1, 1, -1, 1}};

__device__ inline void fill(const LatArgs &a, Q4 &G, const Q4 &Q, int sign, int qsel)
{
    const int nq = a.asym ? 2 : 4;
    for (int t = 0; t < nq; ++t) {
        const int qs = a.asym ? kQsAsym[qsel][t] : kQsSymm[qsel][t];
        G.q[t] = G.q[t] + (double)(sign * qs) * Q.q[t];
    }
}

__global__ __launch_bounds__(256) void lateral_kernel(LatArgs a)
{
    const int64_t cell = blockIdx.x;
    const int i = (int)(cell / a.ny), j = (int)(cell % a.ny);
    const int p = blockIdx.y * blockDim.x + threadIdx.x;
    if (p >= a.npairs) return;
    const double dx = a.dx, dy = a.dy;
    Q4 G;
    for (int t = 0; t < 4; ++t) G.q[t] = c_(0);
    // edges along x at y = dy(jj - 0.5), jj = j (sign -1) and j + 1 (sign +1)   [RunAlongX :254-298]
    for (int e = 0; e < 2; ++e) {
        const int jj = j + e;
        const int sign = e == 0 ? -1 : 1;
        const double y = dy * (jj - 0.5);
        const double x1 = dx * (i - 0.5), x2 = dx * (i + 0.5);
        if (x1 > 0 && x2 > 0) {
            const Knot kl = make_knot(a.radii, a.nr, x1, y), kr = make_knot(a.radii, a.nr, x2, y);
            fill(a, G, integrate(a, p, kl, kr, x1, x2, y, false), sign, 0);
        } else {
            const Knot kz = make_knot(a.radii, a.nr, 0, y);
            const Knot kl = make_knot(a.radii, a.nr, x1, y), kr = make_knot(a.radii, a.nr, x2, y);
            fill(a, G, integrate(a, p, kz, kl, 0, -x1, y, false), sign, 2);
            fill(a, G, integrate(a, p, kz, kr, 0, x2, y, false), sign, 0);
        }
    }
    // edges along y at x = dx(ii - 0.5), ii = i (sign -1) and i + 1 (sign +1)   [RunAlongY :318-361]
    for (int e = 0; e < 2; ++e) {
        const int ii = i + e;
        const int sign = e == 0 ? -1 : 1;
        const double x = dx * (ii - 0.5);
        const double y1 = dy * (j - 0.5), y2 = dy * (j + 0.5);
        if (y1 > 0 && y2 > 0) {
            const Knot kl = make_knot(a.radii, a.nr, x, y1), kr = make_knot(a.radii, a.nr, x, y2);
            fill(a, G, integrate(a, p, kl, kr, y1, y2, x, true), sign, 1);
        } else {
            const Knot kz = make_knot(a.radii, a.nr, x, 0);
            const Knot kl = make_knot(a.radii, a.nr, x, y1), kr = make_knot(a.radii, a.nr, x, y2);
            fill(a, G, integrate(a, p, kz, kl, 0, -y1, x, true), sign, 3);
            fill(a, G, integrate(a, p, kz, kr, 0, y2, x, true), sign, 1);
        }
    }
    // write G(i,j) and its mirror images (GreenTensorFftPopulator.cs:88-148)
    const int nq = a.asym ? 2 : 4;
    const int64_t nx2 = 2LL * a.nx, ny2 = 2LL * a.ny;
    // asym pairs are rc-major (p = rc*nz + tr); the device block stores them swizzled
    const int pd = a.asym ? asym_pos(p / a.nz, p % a.nz, a.nz, asym_swizzle_mask(a.nz)) : p;
    for (int t = 0; t < nq; ++t) {
        const int64_t o = a.off[t] + pd;
        const double2 g = G.q[t];
        a.G[((int64_t)i * ny2 + j) * a.gblk + o] = g;
        if (i >= 1) a.G[((nx2 - i) * ny2 + j) * a.gblk + o] = (double)a.xs[t] * g;
        if (j >= 1) a.G[((int64_t)i * ny2 + (ny2 - j)) * a.gblk + o] = (double)a.ys[t] * g;
        if (i >= 1 && j >= 1) a.G[((nx2 - i) * ny2 + (ny2 - j)) * a.gblk + o] = (double)(a.xs[t] * a.ys[t]) * g;
    }
}

// ---- host: plans ------------------------------------------------------------------------------
// HankelCoefficients.GetLog10RhoStep (:67): (float)Math.Pow(10, 1/_ndec) with _ndec a float 40
double log10_rho_step()
{
    const float inv = 1.0f / 40.0f;
    return (double)(float)std::pow(10.0, (double)inv);
}

struct HostPlan {
    std::vector<std::vector<double>> rho, lam;
    std::vector<double> merged;           // with a zero-rho plan, merged[0] = 0 and m_item[0] = -1
    std::vector<int> m_item, m_idx;
    bool zero_rho = false;
};

// ScalarPlansCreater.CreateFor (:102-110) over [rmin, rmax]; zero_rho puts rho = 0 in front of the
// merged radii (ScalarPlan.GetSortedRho :56-65, GreenScalarCalculator.CalculateZeroRho :147-157)
HostPlan make_plan(int number_of_hankels, double rmin, double rmax, bool zero_rho)
{
    HostPlan hp;
    hp.zero_rho = zero_rho;
    const int nh = std::abs(number_of_hankels);
    if (nh < 1 || !(rmin > 0)) throw hip_error("scalar plan needs NumberOfHankels >= 1 and rho_min > 0");
    const double step = log10_rho_step();
    std::vector<double> r1;                                                 // CreateRhoFromMin :142-157
    for (double nr = rmin; nr <= rmax; nr *= step) r1.push_back(nr);
    r1.push_back(r1.back() * step);
    const double rr2 = r1[r1.size() - 1], rr1 = r1[r1.size() - 2];
    const double st = (rr2 - rr1) / nh;
    hp.rho.push_back(r1);
    for (int i = 1; i < nh; ++i) {                                          // CreateRhoFromMax :159-176
        std::vector<double> rn;
        double nr = rr2 - i * st;
        while (nr >= rmin) {
            rn.push_back(nr);
            nr /= step;
        }
        rn.push_back(nr);
        std::reverse(rn.begin(), rn.end());
        hp.rho.push_back(rn);
    }
    for (auto &rho : hp.rho) {                                              // ScalarPlan.cs:76-97
        const int nlam = kTaps + (int)rho.size();
        std::vector<double> k(nlam);
        k[kL2] = 1 / rho[0];
        for (int i = kL2 + 1; i <= nlam - 1; ++i) k[i] = k[i - 1] / step;
        for (int i = kL2 - 1; i >= 0; --i) k[i] = k[i + 1] * step;
        hp.lam.push_back(k);
    }
    struct E { double r; int it, ix; };                                     // InnerResultsCombiner.cs:19-33
    std::vector<E> all;
    for (int it = 0; it < nh; ++it)
        for (int ix = 0; ix < (int)hp.rho[it].size(); ++ix) all.push_back({hp.rho[it][ix], it, ix});
    std::stable_sort(all.begin(), all.end(), [](const E &a, const E &b) { return a.r < b.r; });
    if (zero_rho) {
        hp.merged.push_back(0.0);
        hp.m_item.push_back(-1);
        hp.m_idx.push_back(0);
    }
    for (auto &e : all) {
        if ((int)hp.merged.size() > (int)zero_rho && hp.merged.back() == e.r) continue;
        hp.merged.push_back(e.r);
        hp.m_item.push_back(e.it);
        hp.m_idx.push_back(e.ix);
    }
    if (hp.merged.size() != all.size() + (zero_rho ? 1 : 0))
        throw hip_error("scalar plan has duplicate rho values; the reference's radii and scalars would misalign");
    return hp;
}

// the anomaly's own range, CreateForAnomalyToAnomaly (ScalarPlansCreater.cs:95-100, 360-380)
HostPlan make_plan(const GreenInput &in)
{
    const double h1 = in.dx / 2, h2 = in.dy / 2;
    const double mx = in.g.nx * in.dx - h1, my = in.g.ny * in.dy - h2;
    return make_plan(in.number_of_hankels, std::min(h1, h2), std::sqrt(mx * mx + my * my), false);
}

// FindRhoMin along one axis (ScalarPlansCreater.cs:245-266) for a level shifted by `shift` against
// the cell centres 0, h, ..., (n - 1) h. The reference's System.Decimal arithmetic and these doubles
// agree whenever h and the shift are whole multiples of a power-of-two fraction of a metre.
double level_rho_min_axis(int n, double h, double shift)
{
    const double a0 = -h / 2, a1 = n * h - h / 2;
    const double y0 = shift - h / 2, y1 = y0 + h * n;
    if ((y0 >= a0 && y0 <= a1) || (y1 >= a0 && y1 <= a1)) {
        const double rem = std::fmod(std::fabs(a0 - y0), h);
        return rem == 0 ? h / 2 : std::fabs(h / 2 - rem);
    }
    const double c1 = h * (n - 1);
    return std::min(std::min(std::fabs(y0), std::fabs(y0 - c1)), std::min(std::fabs(y1), std::fabs(y1 - c1)));
}

// CreateForAnomalyToObservationLevels (:51-66) for one level shifted by (sx, sy): FindRhoMax over
// the anomaly's corner cells and the level footprint's corners (:200-236), FindRhoMin (:238-243),
// and a zero-rho plan from 0.001 min(dx, dy) when the minimum is 0.
HostPlan make_level_plan(const GreenInput &in, double sx, double sy)
{
    const double ex = in.dx * (in.g.nx - 1), ey = in.dy * (in.g.ny - 1);
    const double x1 = sx - in.dx / 2, y1 = sy - in.dy / 2;
    const double x2 = x1 + in.dx * in.g.nx, y2 = y1 + in.dy * in.g.ny;
    const double ax[4] = {0, 0, ex, ex}, ay[4] = {0, ey, ey, 0};
    const double cx[4] = {x1, x2, x2, x1}, cy[4] = {y1, y1, y2, y2};
    double rmax = 0;
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c)
            rmax = std::max(rmax, std::sqrt((cx[c] - ax[a]) * (cx[c] - ax[a]) + (cy[c] - ay[a]) * (cy[c] - ay[a])));
    double rmin = std::min(level_rho_min_axis(in.g.nx, in.dx, sx), level_rho_min_axis(in.g.ny, in.dy, sy));
    const bool zero = rmin <= 0;
    if (zero) rmin = std::min(in.dx, in.dy) * 0.001;
    return make_plan(in.number_of_hankels, rmin, rmax, zero);
}

template <typename T>
struct Dev {
    T *p = nullptr;
    explicit Dev(size_t n) { XTR_CHECK(hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T))); }
    Dev(const std::vector<T> &h) : Dev(h.size())
    {
        XTR_CHECK(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    }
    ~Dev() { (void)hipFree(p); }
    Dev(const Dev &) = delete;
};

int layer_at_depth(const Section &s, double depth)   // ScalarMathUtils.cs:15-31
{
    const int k = layer_index_at_depth(s, depth);
    if (k < 0) throw hip_error("anomaly is not within 1D section");
    return k;
}

}  // namespace

void build_green_atoa(const GreenInput &in, double2 *G, const double2 *tw_x, const double2 *tw_y, hipStream_t s)
{
    const Geom &g = in.g;
    const int nz = g.nz;
    const int nl = (int)in.section.thickness.size();
    if (nl > kMaxLayers) throw hip_error("too many background layers for the device recursion");
    LayerModel lm{};
    lm.nl = nl;
    lm.omega = in.omega;
    const double eps0 = 1.0 / (kMu0 * 299792458.0 * 299792458.0);
    double d = in.section.zero_level;
    for (int k = 0; k < nl; ++k) {
        lm.zeta[k] = make_double2(in.section.sigma[k], -in.omega * eps0);
        lm.thick[k] = in.section.thickness[k];
        lm.depth[k] = d;
        d += in.section.thickness[k];
    }
    lm.depth[nl] = d;

    const HostPlan hp = make_plan(in);
    const int nitems = (int)hp.rho.size();
    int nlam_max = 0;
    std::vector<int> nlam(nitems);
    for (int it = 0; it < nitems; ++it) {
        nlam[it] = (int)hp.lam[it].size();
        nlam_max = std::max(nlam_max, nlam[it]);
    }
    std::vector<double> lam_flat((size_t)nitems * nlam_max, 0.0);
    for (int it = 0; it < nitems; ++it) std::copy(hp.lam[it].begin(), hp.lam[it].end(), lam_flat.begin() + (size_t)it * nlam_max);
    const int nmerged = (int)hp.merged.size();

    Dev<int> d_nlam(nlam);
    Dev<double> d_lam(lam_flat);
    Dev<double> d_radii(hp.merged);
    Dev<int> d_mitem(hp.m_item), d_midx(hp.m_idx);
    Dev<double> d_h0(in.hank0), d_h1(in.hank1);
    const size_t auxn = (size_t)nitems * nlam_max * nl;
    Dev<double2> eta(auxn), ex(auxn), p1(auxn), q1(auxn), ps(auxn), qs(auxn);
    {
        dim3 grid((unsigned)ceil_div(nlam_max, 128), nitems);
        hipLaunchKernelGGL(layer_aux_kernel, grid, dim3(128), 0, s, lm, nitems, nlam_max, d_nlam.p, d_lam.p, eta.p,
                           ex.p, p1.p, q1.p, ps.p, qs.p);
        XTR_LAUNCH_CHECK("layer_aux_kernel");
    }
    XTR_CHECK(hipMemsetAsync(G, 0, (size_t)g.L() * g.gblk() * sizeof(double2), s));

    for (int asym = 1; asym >= 0; --asym) {
        // transceiver pairs: Combine (rc-major) for asym, CombineSymm (tr-major, rc <= tr) for symm
        std::vector<PairInfo> pairs;
        auto add = [&](int rc, int tr) {
            PairInfo pi{};
            pi.r1 = in.depth[rc];
            pi.r2 = in.depth[rc] + in.thick[rc];
            pi.s1 = in.depth[tr];
            pi.s2 = in.depth[tr] + in.thick[tr];
            pi.r = layer_at_depth(in.section, pi.r1 + (pi.r2 - pi.r1) / 2);
            pi.s = layer_at_depth(in.section, pi.s1 + (pi.s2 - pi.s1) / 2);
            pi.t = std::min(pi.r, pi.s);
            pi.b = std::max(pi.r, pi.s);
            pi.dt = lm.depth[pi.t];
            pi.db1 = lm.depth[pi.b + 1];
            const double zw = pi.r1 + (pi.r2 - pi.r1) / 2;
            pi.kase = (zw > pi.s1 && zw < pi.s2) ? 0 : (zw >= pi.s2 ? 1 : 2);
            pairs.push_back(pi);
        };
        if (asym) {
            for (int rc = 0; rc < nz; ++rc)
                for (int tr = 0; tr < nz; ++tr) add(rc, tr);
        } else {
            for (int tr = 0; tr < nz; ++tr)
                for (int rc = 0; rc <= tr; ++rc) add(rc, tr);
        }
        const int npairs = (int)pairs.size();
        const int ncomp = asym ? 1 : 3;
        Dev<PairInfo> d_pairs(pairs);
        Dev<double2> U((size_t)npairs * nitems * ncomp * nlam_max);
        {
            dim3 grid((unsigned)ceil_div(nlam_max, 64), nitems, npairs);
            hipLaunchKernelGGL(pair_u_kernel, grid, dim3(64), 0, s, lm, npairs, d_pairs.p, asym, nitems, nlam_max,
                               d_nlam.p, d_lam.p, eta.p, ex.p, p1.p, q1.p, ps.p, qs.p, U.p);
            XTR_LAUNCH_CHECK("pair_u_kernel");
        }
        Dev<double2> I((size_t)ncomp * npairs * nmerged);
        {
            dim3 grid((unsigned)ceil_div(nmerged, 128), npairs, ncomp);
            hipLaunchKernelGGL(hankel_kernel, grid, dim3(128), 0, s, npairs, nitems, ncomp, nlam_max, U.p, nmerged,
                               d_mitem.p, d_midx.p, d_radii.p, d_h0.p, d_h1.p, asym, I.p);
            XTR_LAUNCH_CHECK("hankel_kernel");
        }
        const int nseg = nmerged - 1;
        Dev<double2> segA((size_t)nseg * ncomp * npairs), segB((size_t)nseg * ncomp * npairs);
        {
            dim3 grid((unsigned)ceil_div(npairs, 128), ncomp, nseg);
            hipLaunchKernelGGL(segments_kernel, grid, dim3(128), 0, s, npairs, ncomp, nmerged, d_radii.p, I.p, segA.p,
                               segB.p);
            XTR_LAUNCH_CHECK("segments_kernel");
        }
        LatArgs la{};
        la.nx = g.nx;
        la.ny = g.ny;
        la.nz = nz;
        la.npairs = npairs;
        la.asym = asym;
        la.nr = nmerged;
        la.dx = in.dx;
        la.dy = in.dy;
        la.segA = segA.p;
        la.segB = segB.p;
        la.radii = d_radii.p;
        la.G = G;
        la.gblk = g.gblk();
        const int64_t S = g.S(), A = (int64_t)nz * nz;
        if (asym) {   // q-order {xz, yz}; device block [xx|xy|yy|zz|xz|yz]; parities XZ (-,+), YZ (+,-)
            la.off[0] = 4 * S; la.off[1] = 4 * S + A;
            la.xs[0] = -1; la.ys[0] = 1;
            la.xs[1] = 1; la.ys[1] = -1;
        } else {      // q-order {xx, yy, zz, xy}; parities (+,+) x3, XY (-,-)
            la.off[0] = 0; la.off[1] = 2 * S; la.off[2] = 3 * S; la.off[3] = S;
            for (int t = 0; t < 3; ++t) { la.xs[t] = 1; la.ys[t] = 1; }
            la.xs[3] = -1; la.ys[3] = -1;
        }
        dim3 grid((unsigned)((int64_t)g.nx * g.ny), (unsigned)ceil_div(npairs, 256));
        hipLaunchKernelGGL(lateral_kernel, grid, dim3(std::min(256, ((npairs + 63) / 64) * 64)), 0, s, la);
        XTR_LAUNCH_CHECK("lateral_kernel");
        XTR_CHECK(hipStreamSynchronize(s));   // the temporaries above are freed on scope exit
    }
    launch_fft2_plain(G, 2 * g.nx, 2 * g.ny, g.gblk(), tw_x, tw_y, -1, s);
}


// =============================================================================================
// AtoO: anomaly layers (volume transmitters) to one observation level (point receivers)
// =============================================================================================
namespace {

// slots of the AtoO components (AtoOLevelGreenTensorCalculator.cs:21-75)
enum { AO_XX, AO_YY, AO_ZZ, AO_XY, AO_YX, AO_XZ, AO_YZ, AO_ZX, AO_ZY, AO_N };

struct PointInfo {            // one transmitter layer seen from the level
    int r, s, t, b, kase;     // background layers of receiver / transmitter, min / max; 0 match 1 below 2 above
    double zw, z1, z2, dt, db1;
};

// CalculateCTop/CBot/DTop/DBot (VolumeToPointCalculatorU.cs:166-208) for layers (r, s)
__device__ inline double2 ctop_(const double2 *E, int r, int s, double zw, double z1, double z2)
{
    return cexp_d(-z1 * E[s] + zw * E[r]) - cexp_d(-z2 * E[s] + zw * E[r]);
}
__device__ inline double2 cbot_(const double2 *E, int r, int s, double zw, double z1, double z2)
{
    return cexp_d(-zw * E[r] + z2 * E[s]) - cexp_d(-zw * E[r] + z1 * E[s]);
}
__device__ inline double2 dtop_(const LayerModel &lm, const double2 *E, int r, int s, double zw, double z1, double z2)
{
    const int t = r < s ? r : s, b = r > s ? r : s;
    const double2 base = -2.0 * (lm.depth[b + 1] * E[b] - lm.depth[t] * E[t]);
    return cexp_d(base + z2 * E[s] - zw * E[r]) - cexp_d(base + z1 * E[s] - zw * E[r]);
}
__device__ inline double2 dbot_(const LayerModel &lm, const double2 *E, int r, int s, double zw, double z1, double z2)
{
    const int t = r < s ? r : s, b = r > s ? r : s;
    const double2 base = -2.0 * (lm.depth[b + 1] * E[b] - lm.depth[t] * E[t]);
    return cexp_d(base - z1 * E[s] + zw * E[r]) - cexp_d(base - z2 * E[s] + zw * E[r]);
}

// U[tr][item][comp][lambda], normalised before the Hankel filter (PlanCalculator.CalculateForE/H
// :58-108, VolumeToPointCalculatorU.cs:34-337, ScalarMathUtils.NormalizeBeforeHankelForE/H).
// electric: comps I1..I5 <- U1 (A1 set) * i w mu0 / 2, U2 / 2z, U3 l / 2z, U4 (-l) / 2z, U5 l^2 / 2z
// magnetic: comps I1..I4 <- U31, U4Sigma, U1Sigma * l, U11 * l
__global__ void point_u_kernel(LayerModel lm, int nz, const PointInfo *pts, int electric, int nitems, int nlam_max,
                               const int *nlam, const double *lam, const double2 *eta, const double2 *ex,
                               const double2 *p1, const double2 *q1, const double2 *ps, const double2 *qs,
                               double2 *U)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    const int tr = blockIdx.z;
    if (it >= nitems || m >= nlam[it] || tr >= nz) return;
    const PointInfo pi = pts[tr];
    const int nl = lm.nl;
    const int64_t base = ((int64_t)it * nlam_max + m) * nl;
    const double2 *E = eta + base, *X = ex + base;
    const double l = lam[(int64_t)it * nlam_max + m];
    const int r = pi.r, s = pi.s, t = pi.t, b = pi.b;
    const double zw = pi.zw, z1 = pi.z1, z2 = pi.z2;
    const double2 er = E[r], es = E[s], et = E[t], eb = E[b];
    // CalculateT / CalculateR (:150-165)
    const double2 T = cexp_d(-zw * er - z1 * es + 2 * pi.dt * et) - cexp_d(-zw * er - z2 * es + 2 * pi.dt * et);
    const double2 R = cexp_d(-2 * pi.db1 * eb + z2 * es + zw * er) - cexp_d(-2 * pi.db1 * eb + z1 * es + zw * er);
    const double2 zrc2 = 2.0 * lm.zeta[r];
    const int ncomp = electric ? 5 : 4;
    double2 *out = U + (((int64_t)tr * nitems + it) * ncomp) * nlam_max + m;
    for (int set = 0; set < 2; ++set) {
        const double2 *P = (set ? ps : p1) + base;
        const double2 *Q = (set ? qs : q1) + base;
        const double2 pt = P[t], qb = Q[b];
        const double2 ann = ann_at(lm, E, X, P, Q, r, s, set);
        double2 psi1, psi3;
        if (pi.kase == 0) {          // receiver inside the transmitter layer (:210-240)
            const int n = r;
            const double2 pq = P[n] * Q[n];
            psi1 = ctop_(E, n, n, zw, zw, z2) + cbot_(E, n, n, zw, z1, zw) +
                   pq * (dtop_(lm, E, n, n, zw, zw, z2) + dbot_(lm, E, n, n, zw, z1, zw));
            psi3 = ctop_(E, n, n, zw, zw, z2) - cbot_(E, n, n, zw, z1, zw) +
                   pq * (dtop_(lm, E, n, n, zw, z1, zw) - dtop_(lm, E, n, n, zw, zw, z2));
        } else if (pi.kase == 1) {   // receiver below (:227-233, 249-256)
            psi1 = cbot_(E, r, s, zw, z1, z2) + pt * qb * dbot_(lm, E, r, s, zw, z1, z2);
            psi3 = -1.0 * cbot_(E, r, s, zw, z1, z2) + pt * qb * dbot_(lm, E, r, s, zw, z1, z2);
        } else {                     // receiver above (:219-226, 241-248)
            psi1 = ctop_(E, r, s, zw, z1, z2) + pt * qb * dtop_(lm, E, r, s, zw, z1, z2);
            psi3 = ctop_(E, r, s, zw, z1, z2) - pt * qb * dtop_(lm, E, r, s, zw, z1, z2);
        }
        const double2 u1 = (ann / es) * (pt * T + qb * R + psi1);
        const double2 u3 = (ann * (er / es)) * (-1.0 * (pt * T) + qb * R + psi3);
        const double2 u4 = ann * (-1.0 * (pt * T) + qb * R - psi3);
        if (electric) {
            if (set == 0) {
                out[0] = u1 * (c_(0, lm.omega * kMu0) * c_(0.5));
                continue;
            }
            const double2 u2 = (ann * er) * (pt * T + qb * R - psi1);
            double2 u5;
            if (pi.kase == 0) {      // CalcPsi5RcAndTrMatch (:258-272)
                const int n = r;
                const double ln = lm.thick[n];
                const double2 kk = (c_(0, lm.omega) * lm.zeta[n]) * c_(kMu0);
                const double2 en = E[n], pq = P[n] * Q[n];
                const double2 psi5 = pq * (cexp_d(-(2 * ln - (z2 - zw)) * en) + cexp_d(-(2 * ln - (zw - z1)) * en)) -
                                     (cexp_d(-(z2 - zw) * en) + cexp_d(-(zw - z1) * en)) +
                                     2.0 * (kk * ((c_(1) - pq * cexp_d(-2 * ln * en)) / c_(l * l)));
                u5 = (ann / es) * (pt * T + qb * R + psi5);
            } else {
                u5 = u1;
            }
            out[(int64_t)1 * nlam_max] = u2 / zrc2;
            out[(int64_t)2 * nlam_max] = u3 * (c_(l) / zrc2);
            out[(int64_t)3 * nlam_max] = u4 * (c_(-l) / zrc2);
            out[(int64_t)4 * nlam_max] = u5 * (c_(l * l) / zrc2);
        } else {
            if (set == 0) {
                out[0] = u3;                                 // I1 <- U31
                out[(int64_t)3 * nlam_max] = u1 * c_(l);     // I4 <- U11 * lambda
            } else {
                out[(int64_t)1 * nlam_max] = u4;             // I2 <- U4Sigma
                out[(int64_t)2 * nlam_max] = u1 * c_(l);     // I3 <- U1Sigma * lambda
            }
        }
    }
}

// Hankel filter in the non-fast order of HankelCoefficients.ConvoluteWith (:86-99) and
// NormalizeAfterHankelForE/H (ScalarMathUtils.cs:131-155): comp c uses J1 or J0 and the factor
// sign / (den rho). electric: J1 -, J1 +, J0 +, J0 +, J1 + with den 2 pi; magnetic: J1 -, J1 -,
// J0 -, J0 + with den 4 pi.
__device__ inline double2 atoo_hankel_at(int nitems, int ncomp, int nlam_max, const double2 *U, int tr, int it,
                                         int idx, int c, double rho, const double *hank0, const double *hank1,
                                         int electric)
{
    const bool j0 = c == 2 || c == 3;
    const double *hank = j0 ? hank0 : hank1;
    const double sgn = electric ? (c == 0 ? -1.0 : 1.0) : (c == 3 ? 1.0 : -1.0);
    const double den = (electric ? 2 * kPi : 4 * kPi) * rho;
    const double2 *f = U + (((int64_t)tr * nitems + it) * ncomp + c) * nlam_max + idx;
    double re = 0, im = 0;
    for (int j = 0; j < kTaps; ++j) {
        const double h = hank[j];
        const double2 v = f[kTaps - 1 - j];
        re += v.x * h;
        im += v.y * h;
    }
    return make_double2(sgn * re / den, sgn * im / den);
}

// A merged radius with item -1 is the zero-rho point (CalculateZeroRhoForE/H, ScalarMathUtils.cs:33-81):
// I1, I2 (and I5) are 0 there, I3 and I4 are (f[1] a^2 - f[2]) / (a^2 - 1) from the first item's 2nd and
// 3rd radii rho01, rho02, a the rho step.
__global__ void atoo_hankel_kernel(int nz, int nitems, int ncomp, int nlam_max, const double2 *U, int nmerged,
                                   const int *m_item, const int *m_idx, const double *radii, const double *hank0,
                                   const double *hank1, int electric, double rho01, double rho02, double alpha2,
                                   double2 *I)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int tr = blockIdx.y;
    const int c = blockIdx.z;
    if (k >= nmerged || tr >= nz || c >= ncomp) return;
    const int it = m_item[k], idx = m_idx[k];
    double2 v;
    if (it >= 0) {
        v = atoo_hankel_at(nitems, ncomp, nlam_max, U, tr, it, idx, c, radii[k], hank0, hank1, electric);
    } else if (c == 2 || c == 3) {
        const double2 f1 = atoo_hankel_at(nitems, ncomp, nlam_max, U, tr, 0, 1, c, rho01, hank0, hank1, electric);
        const double2 f2 = atoo_hankel_at(nitems, ncomp, nlam_max, U, tr, 0, 2, c, rho02, hank0, hank1, electric);
        v = make_double2((f1.x * alpha2 - f2.x) / (alpha2 - 1), (f1.y * alpha2 - f2.y) / (alpha2 - 1));
    } else {
        v = make_double2(0, 0);
    }
    I[((int64_t)c * nz + tr) * nmerged + k] = v;
}

struct AoArgs {
    int nx, ny, nz, nr, electric, ncomp, ntc;
    double dx, dy;
    double xs, ys;                // minus the level's shift (AtoOLevelGreenTensorCalculator.cs:119-125)
    const double2 *segA, *segB;   // [seg][comp][tr]
    const double *radii;
    double2 *G;                   // [2Nx][2Ny][ntc][nz]
    int slot[AO_N];               // compact component index of each slot, -1 if not computed
};

struct Q9 { double2 q[AO_N]; };

// AddToQBuffAlong{X,Y}{Electric,Magnetic} (GreenTensorCalculator.cs:505-560); comps I1..I5 at 0..4
__device__ inline void ao_add(const AoArgs &a, int tr, const Factors &f, int seg, bool alongY, Q9 &Q)
{
    auto A = [&](int ic) { return a.segA[((int64_t)seg * a.ncomp + (ic - 1)) * a.nz + tr]; };
    auto B = [&](int ic) { return a.segB[((int64_t)seg * a.ncomp + (ic - 1)) * a.nz + tr]; };
    auto add = [&](int slot, int ic, double da, double db) {
        Q.q[slot] = Q.q[slot] + da * A(ic);
        Q.q[slot] = Q.q[slot] + db * B(ic);
    };
    if (a.electric) {
        if (!alongY) {
            add(AO_XX, 1, f.d2a, f.d2b); add(AO_YY, 2, f.d2a, f.d2b); add(AO_ZZ, 5, f.d2a, f.d2b);
            add(AO_XY, 2, f.d3a, f.d3b); add(AO_YZ, 3, f.d1a, f.d1b); add(AO_ZY, 4, f.d1a, f.d1b);
        } else {
            add(AO_XX, 2, f.d2a, f.d2b); add(AO_YY, 1, f.d2a, f.d2b); add(AO_ZZ, 5, f.d2a, f.d2b);
            add(AO_XY, 1, f.d3a, f.d3b); add(AO_XZ, 3, f.d1a, f.d1b); add(AO_ZX, 4, f.d1a, f.d1b);
        }
    } else {
        if (!alongY) {
            add(AO_XX, 2, f.d3a, f.d3b); add(AO_YX, 1, f.d2a, f.d2b); add(AO_XY, 2, f.d2a, f.d2b);
            add(AO_XZ, 3, f.d1a, f.d1b); add(AO_ZX, 4, f.d1a, f.d1b);
        } else {
            add(AO_YX, 2, f.d2a, f.d2b); add(AO_XY, 1, f.d2a, f.d2b); add(AO_XX, 1, f.d3a, f.d3b);
            add(AO_YZ, 3, f.d1a, f.d1b); add(AO_ZY, 4, f.d1a, f.d1b);
        }
    }
}

__device__ inline Q9 ao_integrate(const AoArgs &a, int tr, Knot left, Knot right, double s, double u, double v,
                                  bool alongY)
{
    Q9 Q;
    for (int t = 0; t < AO_N; ++t) Q.q[t] = c_(0);
    const int i1 = left.index, i2 = right.index;
    if (i1 == i2) {
        ao_add(a, tr, factors(s, u, left.radius, right.radius, v), i1, alongY, Q);
        return Q;
    }
    double r1 = left.radius, r2 = a.radii[i1 + 1];
    double sj = s, uj = sqrt(r2 * r2 - v * v);
    ao_add(a, tr, factors(sj, uj, r1, r2, v), i1, alongY, Q);
    for (int i = i1 + 1; i < i2; ++i) {
        r1 = r2;
        r2 = a.radii[i + 1];
        sj = uj;
        uj = sqrt(r2 * r2 - v * v);
        ao_add(a, tr, factors(sj, uj, r1, r2, v), i, alongY, Q);
    }
    sj = uj; uj = u; r1 = r2; r2 = right.radius;
    ao_add(a, tr, factors(sj, uj, r1, r2, v), i2, alongY, Q);
    return Q;
}

// QSigns (QSign.cs:31-38) by slot; rows AlongX, AlongY, AlongXSymm, AlongYSymm
__constant__ int kQsE[4][AO_N] = {{1, -1, -1, -1, 0, 1, 1, 1, 1},
                                  {-1, 1, -1, -1, 0, 1, 1, 1, 1},
                                  {1, -1, -1, 1, 0, -1, 1, -1, 1},
                                  {-1, 1, -1, 1, 0, 1, -1, 1, -1}};
__constant__ int kQsH[4][AO_N] = {{-1, -1, 0, -1, 1, 1, 1, 1, 1},
                                  {-1, -1, 0, 1, -1, 1, 1, 1, 1},
                                  {1, 1, 0, -1, 1, 1, -1, 1, -1},
                                  {1, 1, 0, 1, -1, 1, 1, -1, 1}};

__device__ inline void ao_fill(const AoArgs &a, Q9 &G, const Q9 &Q, int sign, int qsel)
{
    for (int t = 0; t < AO_N; ++t) {
        const int qs = a.electric ? kQsE[qsel][t] : kQsH[qsel][t];
        G.q[t] = G.q[t] + (double)(sign * qs) * Q.q[t];
    }
}

// One thread per (lateral offset (i, j) in (-Nx, Nx) x (-Ny, Ny), transmitter layer): the four edge
// integrals of RunAlongX / RunAlongY (GreenTensorCalculator.cs:192-281) that FillResultAlongX/Y
// (:413-437) add into this offset, written at TransformI/J (AtoOLevelGreenTensorCalculator.cs:106-117).
__global__ __launch_bounds__(64) void atoo_lateral_kernel(AoArgs a)
{
    const int ncy = 2 * a.ny - 1;
    const int cell = blockIdx.x;
    const int i = cell / ncy - (a.nx - 1), j = cell % ncy - (a.ny - 1);
    const int tr = blockIdx.y * blockDim.x + threadIdx.x;
    if (tr >= a.nz) return;
    const double dx = a.dx, dy = a.dy;
    Q9 G;
    for (int t = 0; t < AO_N; ++t) G.q[t] = c_(0);
    for (int e = 0; e < 2; ++e) {     // edges along x at y = dy (jj - 1/2) - ys: jj = j (-1), j + 1 (+1)
        const int jj = j + e, sign = e == 0 ? -1 : 1;
        const double y = dy * (jj - 0.5) - a.ys;
        const double x1 = dx * (i - 0.5) - a.xs, x2 = dx * (i + 0.5) - a.xs;
        const Knot kl = make_knot(a.radii, a.nr, x1, y), kr = make_knot(a.radii, a.nr, x2, y);
        if (x1 > 0 && x2 > 0) {
            ao_fill(a, G, ao_integrate(a, tr, kl, kr, x1, x2, y, false), sign, 0);
        } else if (x1 < 0 && x2 < 0) {
            ao_fill(a, G, ao_integrate(a, tr, kr, kl, -x2, -x1, y, false), sign, 2);
        } else {
            const Knot kz = make_knot(a.radii, a.nr, 0, y);
            ao_fill(a, G, ao_integrate(a, tr, kz, kl, 0, -x1, y, false), sign, 2);
            ao_fill(a, G, ao_integrate(a, tr, kz, kr, 0, x2, y, false), sign, 0);
        }
    }
    for (int e = 0; e < 2; ++e) {     // edges along y at x = dx (ii - 1/2) - xs: ii = i (-1), i + 1 (+1)
        const int ii = i + e, sign = e == 0 ? -1 : 1;
        const double x = dx * (ii - 0.5) - a.xs;
        const double y1 = dy * (j - 0.5) - a.ys, y2 = dy * (j + 0.5) - a.ys;
        const Knot kl = make_knot(a.radii, a.nr, x, y1), kr = make_knot(a.radii, a.nr, x, y2);
        if (y1 > 0 && y2 > 0) {
            ao_fill(a, G, ao_integrate(a, tr, kl, kr, y1, y2, x, true), sign, 1);
        } else if (y1 < 0 && y2 < 0) {
            ao_fill(a, G, ao_integrate(a, tr, kr, kl, -y2, -y1, x, true), sign, 3);
        } else {
            const Knot kz = make_knot(a.radii, a.nr, x, 0);
            ao_fill(a, G, ao_integrate(a, tr, kz, kl, 0, -y1, x, true), sign, 3);
            ao_fill(a, G, ao_integrate(a, tr, kz, kr, 0, y2, x, true), sign, 1);
        }
    }
    const int64_t ti = i < 0 ? 2 * a.nx + i : i, tj = j < 0 ? 2 * a.ny + j : j;
    double2 *dst = a.G + (ti * 2 * a.ny + tj) * (int64_t)(a.ntc * a.nz) + tr;
    for (int t = 0; t < AO_N; ++t)
        if (a.slot[t] >= 0) dst[(int64_t)a.slot[t] * a.nz] = G.q[t];
}

}  // namespace

int atoo_components(int electric, int *slot)
{
    // compact order: electric xx yy zz xy xz yz zx zy; magnetic xx xy yx xz yz zx zy
    static const int e[AO_N] = {0, 1, 2, 3, -1, 4, 5, 6, 7};
    static const int h[AO_N] = {0, -1, -1, 1, 2, 3, 4, 5, 6};
    for (int t = 0; t < AO_N; ++t) slot[t] = electric ? e[t] : h[t];
    return electric ? 8 : 7;
}

void build_green_atoo_level(const GreenInput &in, double level_z, double shift_x, double shift_y, int electric,
                            double2 *G, const double2 *tw_x, const double2 *tw_y, hipStream_t s)
{
    const Geom &g = in.g;
    const int nz = g.nz;
    const int nl = (int)in.section.thickness.size();
    if (nl > kMaxLayers) throw hip_error("too many background layers for the device recursion");
    LayerModel lm{};
    lm.nl = nl;
    lm.omega = in.omega;
    const double eps0 = 1.0 / (kMu0 * 299792458.0 * 299792458.0);
    double d = in.section.zero_level;
    for (int k = 0; k < nl; ++k) {
        lm.zeta[k] = make_double2(in.section.sigma[k], -in.omega * eps0);
        lm.thick[k] = in.section.thickness[k];
        lm.depth[k] = d;
        d += in.section.thickness[k];
    }
    lm.depth[nl] = d;
    // the level's own scalar plan (AtoOCalculator.cs:129, 151)
    const HostPlan hp = make_level_plan(in, shift_x, shift_y);
    if (hp.zero_rho && hp.rho[0].size() < 3) throw hip_error("zero-rho plan needs three radii in its first item");
    const int nitems = (int)hp.rho.size();
    int nlam_max = 0;
    std::vector<int> nlam(nitems);
    for (int it = 0; it < nitems; ++it) {
        nlam[it] = (int)hp.lam[it].size();
        nlam_max = std::max(nlam_max, nlam[it]);
    }
    std::vector<double> lam_flat((size_t)nitems * nlam_max, 0.0);
    for (int it = 0; it < nitems; ++it) std::copy(hp.lam[it].begin(), hp.lam[it].end(), lam_flat.begin() + (size_t)it * nlam_max);
    const int nmerged = (int)hp.merged.size();
    Dev<int> d_nlam(nlam);
    Dev<double> d_lam(lam_flat);
    Dev<double> d_radii(hp.merged);
    Dev<int> d_mitem(hp.m_item), d_midx(hp.m_idx);
    Dev<double> d_h0(in.hank0), d_h1(in.hank1);
    const size_t auxn = (size_t)nitems * nlam_max * nl;
    Dev<double2> eta(auxn), ex(auxn), p1(auxn), q1(auxn), ps(auxn), qs(auxn);
    {
        dim3 grid((unsigned)ceil_div(nlam_max, 128), nitems);
        hipLaunchKernelGGL(layer_aux_kernel, grid, dim3(128), 0, s, lm, nitems, nlam_max, d_nlam.p, d_lam.p, eta.p,
                           ex.p, p1.p, q1.p, ps.p, qs.p);
        XTR_LAUNCH_CHECK("layer_aux_kernel");
    }
    // transmitters: the anomaly layers; receiver: the level point (AuxContainer.CreateContainer)
    const int r = layer_at_depth(in.section, level_z);
    std::vector<PointInfo> pts(nz);
    for (int tr = 0; tr < nz; ++tr) {
        PointInfo pi{};
        pi.zw = level_z;
        pi.z1 = in.depth[tr];
        pi.z2 = in.depth[tr] + in.thick[tr];
        pi.r = r;
        pi.s = layer_at_depth(in.section, pi.z1 + (pi.z2 - pi.z1) / 2);
        pi.t = std::min(pi.r, pi.s);
        pi.b = std::max(pi.r, pi.s);
        pi.dt = lm.depth[pi.t];
        pi.db1 = lm.depth[pi.b + 1];
        pi.kase = (pi.zw > pi.z1 && pi.zw < pi.z2) ? 0 : (pi.zw >= pi.z2 ? 1 : 2);   // CalcUGeneral (:54-62)
        pts[tr] = pi;
    }
    const int ncomp = electric ? 5 : 4;
    Dev<PointInfo> d_pts(pts);
    Dev<double2> U((size_t)nz * nitems * ncomp * nlam_max);
    {
        dim3 grid((unsigned)ceil_div(nlam_max, 64), nitems, nz);
        hipLaunchKernelGGL(point_u_kernel, grid, dim3(64), 0, s, lm, nz, d_pts.p, electric, nitems, nlam_max, d_nlam.p,
                           d_lam.p, eta.p, ex.p, p1.p, q1.p, ps.p, qs.p, U.p);
        XTR_LAUNCH_CHECK("point_u_kernel");
    }
    Dev<double2> I((size_t)ncomp * nz * nmerged);
    {
        dim3 grid((unsigned)ceil_div(nmerged, 128), nz, ncomp);
        hipLaunchKernelGGL(atoo_hankel_kernel, grid, dim3(128), 0, s, nz, nitems, ncomp, nlam_max, U.p, nmerged,
                           d_mitem.p, d_midx.p, d_radii.p, d_h0.p, d_h1.p, electric,
                           hp.zero_rho ? hp.rho[0][1] : 0.0, hp.zero_rho ? hp.rho[0][2] : 0.0,
                           log10_rho_step() * log10_rho_step(), I.p);
        XTR_LAUNCH_CHECK("atoo_hankel_kernel");
    }
    const int nseg = nmerged - 1;
    Dev<double2> segA((size_t)nseg * ncomp * nz), segB((size_t)nseg * ncomp * nz);
    {
        dim3 grid((unsigned)ceil_div(nz, 128), ncomp, nseg);
        hipLaunchKernelGGL(segments_kernel, grid, dim3(128), 0, s, nz, ncomp, nmerged, d_radii.p, I.p, segA.p, segB.p);
        XTR_LAUNCH_CHECK("segments_kernel");
    }
    AoArgs a{};
    a.nx = g.nx;
    a.ny = g.ny;
    a.nz = nz;
    a.nr = nmerged;
    a.electric = electric;
    a.ncomp = ncomp;
    a.ntc = atoo_components(electric, a.slot);
    a.dx = in.dx;
    a.dy = in.dy;
    a.xs = -shift_x;
    a.ys = -shift_y;
    a.segA = segA.p;
    a.segB = segB.p;
    a.radii = d_radii.p;
    a.G = G;
    XTR_CHECK(hipMemsetAsync(G, 0, (size_t)g.L() * a.ntc * nz * sizeof(double2), s));
    {
        dim3 grid((unsigned)((2 * g.nx - 1) * (2 * g.ny - 1)), (unsigned)ceil_div(nz, 64));
        hipLaunchKernelGGL(atoo_lateral_kernel, grid, dim3(64), 0, s, a);
        XTR_LAUNCH_CHECK("atoo_lateral_kernel");
    }
    launch_fft2_plain(G, 2 * g.nx, 2 * g.ny, (int64_t)a.ntc * nz, tw_x, tw_y, -1, s);
    XTR_CHECK(hipStreamSynchronize(s));   // the temporaries above are freed on return
}

}  // namespace xtr
