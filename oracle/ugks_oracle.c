/*
 * ugks_oracle.c -- FP64 CPU oracle for the pyUGKS time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or
 * shipped).  Each function below restates one reference OpenCL kernel or
 * host routine; the reference file:line is given beside it.  Arithmetic is
 * written in the reference expression order; velocity sums are sequential
 * in v, as with the reference's default CL_local_size = 1
 * (ugks_data.py:215).  OpenMP threads split only the independent outer
 * (cell / face) loops, so results do not depend on the thread count.
 *
 * Parity pins for each part are listed in the headers of
 * tests/test_oracle_pins.py, tests/test_oracle_flux_limits.py,
 * tests/test_oracle_boundaries.py and tests/test_oracle_joins.py.
 * Functions on the path not covered by one of them: none.
 */
#include "ugks_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_ 3.141592653589793   /* '%0.15e' % pi, ugks_CL.py:84 */
enum { SOUTH = 0, WEST = 1 };
enum { GNORTH = 0, GEAST = 1, GSOUTH = 2, GWEST = 3 };
#define ERR_NAN 1.0                      /* clError.c:4  */
#define ERR_CLAMPING_ADSORB_ISOTHERM 4.0 /* clError.c:13 */
enum { NO_FLIP = 0, FLIP_NS = 1, FLIP_D = 2, FLIP_EW = 3 };  /* clError.c:17-20 */

/* index macros, clKernel.c:4-26, clUGKS.c:106-107, :201-207, :444, :520 */
#define NI_ (P->NI)
#define NJ_ (P->NJ)
#define NV_ (P->NV)
#define G_ (P->ghost)
#define F3(a, i, j, v) (a)[(size_t)NV_ * NJ_ * (size_t)(i) + (size_t)NV_ * (size_t)(j) + (size_t)(v)]
#define C2(a, i, j) (a)[(size_t)(i) * NJ_ + (size_t)(j)]
#define M2(a, i, j) (a)[(size_t)(i) * P->nj + (size_t)(j)]
#define FACEA(a, i, j, face) (a)[(size_t)(i) * NJ_ * 2 + (size_t)(j) * 2 + (size_t)(face)]
#define XY(i, j) xy[(size_t)(i) * (NJ_ + 1) + (size_t)(j)]
#define WP(a, w, i) (a)[(size_t)(i) * 4 + (size_t)(w)]

static int g_threads = 1;
void or_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int or_get_threads(void) { return g_threads; }

/* ---- small vector helpers (OpenCL builtin semantics) ------------------- */
static inline od2 v2(double x, double y) { od2 r = {x, y}; return r; }
static inline od4 v4(double a, double b, double c, double d) { od4 r = {a, b, c, d}; return r; }
static inline od2 add2(od2 a, od2 b) { return v2(a.x + b.x, a.y + b.y); }
static inline od2 sub2(od2 a, od2 b) { return v2(a.x - b.x, a.y - b.y); }
static inline od2 mul2s(od2 a, double s) { return v2(a.x * s, a.y * s); }
static inline od2 smul2(double s, od2 a) { return v2(s * a.x, s * a.y); }
static inline od2 div2s(od2 a, double s) { return v2(a.x / s, a.y / s); }
static inline double dot2(od2 a, od2 b) { return a.x * b.x + a.y * b.y; }
static inline double len2(od2 a) { return sqrt(a.x * a.x + a.y * a.y); }
static inline od4 add4(od4 a, od4 b) { return v4(a.s0 + b.s0, a.s1 + b.s1, a.s2 + b.s2, a.s3 + b.s3); }
static inline od4 sub4(od4 a, od4 b) { return v4(a.s0 - b.s0, a.s1 - b.s1, a.s2 - b.s2, a.s3 - b.s3); }
static inline od4 smul4(double s, od4 a) { return v4(s * a.s0, s * a.s1, s * a.s2, s * a.s3); }
static inline od4 div4s(od4 a, double s) { return v4(a.s0 / s, a.s1 / s, a.s2 / s, a.s3 / s); }
static inline od4 zero4(void) { return v4(0.0, 0.0, 0.0, 0.0); }

/* OpenCL sign(): 1, -1, or the (signed) zero itself; NaN -> 0 */
static inline double osign(double x) {
    if (x > 0.0) return 1.0;
    if (x < 0.0) return -1.0;
    if (x == 0.0) return x;
    return 0.0;
}
/* int delta = (sign(u)+1)/2  (clUGKS.c:148, :743) */
static inline int upwind_delta(double u) { return (int)((osign(u) + 1.0) / 2.0); }
/* int delta = (sign(u)*rot + 1)/2  (clUGKS.c:852, clAdsorb.c:354) */
static inline int wall_delta(double u, int rot) { return (int)((osign(u) * (double)rot + 1.0) / 2.0); }

static inline od2 quad_at(const or_params *P, int v) { return v2(P->quad[2 * v], P->quad[2 * v + 1]); }

/* interfaceVelocity, clFunctions.c:191-209 */
static inline od2 interfaceVelocity(const or_params *P, int v, od2 n) {
    od2 e = quad_at(P, v);
    od2 t = v2(-n.y, n.x);
    return v2(dot2(e, n), dot2(e, t));
}
/* toGlobal, clFunctions.c:215-230 */
static inline od2 toGlobal(od2 in, od2 n) {
    od2 t = v2(-n.y, n.x);
    od2 a = mul2s(n, in.x), b = mul2s(t, in.y);
    return v2(dot2(a, v2(1.0, 0.0)) + dot2(b, v2(1.0, 0.0)), dot2(a, v2(0.0, 1.0)) + dot2(b, v2(0.0, 1.0)));
}
/* toLocal, clFunctions.c:236-250 */
static inline od2 toLocal(od2 in, od2 n) {
    od2 t = v2(-n.y, n.x);
    return v2(dot2(in, n), dot2(in, t));
}

/* relaxTime, clFunctions.c:8-32 */
double or_relaxTime(const or_params *P, od4 prim) {
    double tau = 0.0;
    switch (P->relax_type) {
    case 0: tau = (P->Kn_eff / prim.s0) * 0.79788456080286541 * pow(prim.s3, 1.0 - P->omega); break;
    case 1: tau = (5. / 8.) * (P->Kn_eff / prim.s0) * sqrt(PI_) * pow(prim.s3, 1.0 - P->omega); break;
    case 2: tau = (P->Kn_eff / prim.s0) * pow(prim.s3, 1.0 - P->omega); break;
    case 3: tau = (P->Kn_eff / prim.s0) * 2.2567583341910251 * pow(prim.s3, 1.0 - P->omega); break;
    }
    return tau;
}

/* fM, clFunctions.c:49-58 */
od2 or_fM(const or_params *P, od4 prim, od2 uv, int gv) {
    od2 c = sub2(uv, v2(prim.s1, prim.s2));
    od2 M;
    M.x = prim.s0 * (prim.s3 / PI_) * exp(-prim.s3 * dot2(c, c));
    M.y = (M.x * P->K) / (2.0 * prim.s3);
    return smul2(P->weight[gv], M);
}
/* fS, clFunctions.c:64-77 */
od2 or_fS(const or_params *P, od4 prim, od2 Q, od2 uv, od2 M) {
    od2 c = sub2(uv, v2(prim.s1, prim.s2));
    double s = 0.8 * (1 - P->Pr) * (prim.s3 * prim.s3) / prim.s0 * dot2(c, Q);
    double part = 2 * prim.s3 * (dot2(c, c));
    od2 S = v2(s * (part - 4), s * (part - 2));
    return v2(S.x * M.x, S.y * M.y);
}
/* fEQ, clFunctions.c:79-86 */
static inline od2 fEQ(const or_params *P, od4 prim, od2 Q, od2 uv, int gv) {
    od2 M = or_fM(P, prim, uv, gv);
    return add2(M, or_fS(P, prim, Q, uv, M));
}

/* vanLeer, clFunctions.c:256-260 (componentwise) */
static inline double vl1(double s1, double s2) {
    return (osign(s1) + osign(s2)) * (fabs(s1) * fabs(s2)) / (fabs(s1) + fabs(s2) + 1e-16);
}
od2 or_vanLeer(od2 s1, od2 s2) { return v2(vl1(s1.x, s2.x), vl1(s1.y, s2.y)); }

/* WENO5, clFunctions.c:268-312 (componentwise) */
static double weno1(const double *S, double dx) {
    const int J = 2;
    const double EPS = 1E-16;
    double temp1, temp2, B0, B1, B2, a0, a1, a2, w0, w1, w2, f0, f1, f2;
    temp1 = S[J - 2] - 2.0 * S[J - 1] + S[J];
    temp2 = S[J - 2] - 4.0 * S[J - 1] + 3.0 * S[J];
    B0 = (13.0 / 12.0) * temp1 * temp1 + 0.25 * temp2 * temp2;
    temp1 = S[J - 1] - 2.0 * S[J] + S[J + 1];
    temp2 = S[J - 1] - S[J + 1];
    B1 = (13.0 / 12.0) * temp1 * temp1 + 0.25 * temp2 * temp2;
    temp1 = S[J] - 2.0 * S[J + 1] + S[J + 2];
    temp2 = 3.0 * S[J] - 4.0 * S[J + 1] + S[J + 2];
    B2 = (13.0 / 12.0) * temp1 * temp1 + 0.25 * temp2 * temp2;
    temp1 = EPS + B0; a0 = 0.1 / (temp1 * temp1);
    temp1 = EPS + B1; a1 = 0.6 / (temp1 * temp1);
    temp1 = EPS + B2; a2 = 0.3 / (temp1 * temp1);
    temp1 = a0 + a1 + a2;
    w0 = a0 / temp1; w1 = a1 / temp1; w2 = a2 / temp1;
    f0 = (1. / 3.) * S[J - 2] - (7. / 6.) * S[J - 1] + (11. / 6.) * S[J];
    f1 = -(1. / 6.) * S[J - 1] + (5. / 6.) * S[J] + (1. / 3.) * S[J + 1];
    f2 = (1. / 3.) * S[J] + (5. / 6.) * S[J + 1] - (1. / 6.) * S[J + 2];
    double fL = w0 * f0 + w1 * f1 + w2 * f2;
    return (fL - S[J]) / dx;
}
od2 or_WENO5(const od2 *S, double dx) {
    double gx[5], gy[5];
    for (int k = 0; k < 5; k++) { gx[k] = S[k].x; gy[k] = S[k].y; }
    return v2(weno1(gx, dx), weno1(gy, dx));
}

/* minmod + NND, clFunctions.c:320-352 */
static double minmod(double a, double b) {
    double out = 0.0;  /* every branch of the reference assigns; NaN input -> 0 here */
    if ((fabs(a) < fabs(b)) && (a * b > 0)) out = a;
    else if ((fabs(b) < fabs(a)) && (a * b > 0)) out = b;
    else if (a == b) out = a;
    else if (a * b <= 0) out = 0;
    return out;
}
od2 or_NND(const od2 *S, double dx) {
    const int n = 1;
    od2 dIp12 = sub2(S[n + 1], S[n]);
    od2 dIm12 = sub2(S[n], S[n - 1]);
    od2 Ip12;
    Ip12.x = S[n].x + 0.5 * minmod(dIp12.x, dIm12.x);
    Ip12.y = S[n].y + 0.5 * minmod(dIp12.y, dIm12.y);
    return div2s(sub2(Ip12, S[n]), dx);
}

/* getPrimary / getConserved, clUGKS.c:16-28 */
od4 or_getPrimary(const or_params *P, od4 w) {
    w.s3 = 0.5 * w.s0 / (P->gam - 1.0) / (w.s3 - 0.5 * (w.s1 * w.s1 + w.s2 * w.s2) / w.s0);
    w.s1 /= w.s0; w.s2 /= w.s0;
    return w;
}
od4 or_getConserved(const or_params *P, od4 prim) {
    prim.s3 = 0.5 * prim.s0 / prim.s3 / (P->gam - 1.0) + 0.5 * prim.s0 * (prim.s1 * prim.s1 + prim.s2 * prim.s2);
    prim.s1 *= prim.s0; prim.s2 *= prim.s0;
    return prim;
}

/* microSlope, clUGKS.c:30-45 */
od4 or_microSlope(const or_params *P, od4 prim, od4 sw) {
    const double K = P->K;
    od4 m;
    m.s3 = 4.0 * (prim.s3 * prim.s3) / (K + 2) / prim.s0 *
           (2.0 * sw.s3 - 2.0 * prim.s1 * sw.s1 - 2.0 * prim.s2 * sw.s2 +
            sw.s0 * ((prim.s1 * prim.s1 + prim.s2 * prim.s2) - 0.5 * (K + 2) / prim.s3));
    m.s2 = 2.0 * prim.s3 / prim.s0 * (sw.s2 - prim.s2 * sw.s0) - prim.s2 * m.s3;
    m.s1 = 2.0 * prim.s3 / prim.s0 * (sw.s1 - prim.s1 * sw.s0) - prim.s1 * m.s3;
    m.s0 = sw.s0 / prim.s0 - prim.s1 * m.s1 - prim.s2 * m.s2 -
           0.5 * ((prim.s1 * prim.s1) + (prim.s2 * prim.s2) + 0.5 * (K + 2) / prim.s3) * m.s3;
    return m;
}

/* momentU, clUGKS.c:47-80 (MNUM = 7, MTUM = 5) */
void or_momentU(const or_params *P, od4 prim, double *Mu, double *Mv, double *Mxi,
                double *Mu_L, double *Mu_R) {
    Mu_L[0] = 0.5 * erfc(-sqrt(prim.s3) * prim.s1);
    Mu_L[1] = prim.s1 * Mu_L[0] + 0.5 * exp(-prim.s3 * (prim.s1 * prim.s1)) / sqrt(PI_ * prim.s3);
    Mu_R[0] = 0.5 * erfc(sqrt(prim.s3) * prim.s1);
    Mu_R[1] = prim.s1 * Mu_R[0] - 0.5 * exp(-prim.s3 * (prim.s1 * prim.s1)) / sqrt(PI_ * prim.s3);
    for (int i = 2; i < 7; ++i) {
        Mu_L[i] = prim.s1 * Mu_L[i - 1] + 0.5 * (i - 1) * Mu_L[i - 2] / prim.s3;
        Mu_R[i] = prim.s1 * Mu_R[i - 1] + 0.5 * (i - 1) * Mu_R[i - 2] / prim.s3;
    }
    for (int i = 0; i < 7; ++i) Mu[i] = Mu_L[i] + Mu_R[i];
    Mv[0] = 1.0;
    Mv[1] = prim.s2;
    for (int i = 2; i < 5; ++i) Mv[i] = prim.s2 * Mv[i - 1] + 0.5 * (i - 1) * Mv[i - 2] / prim.s3;
    Mxi[0] = 1.0;
    Mxi[1] = 0.5 * P->K / prim.s3;
    Mxi[2] = (P->K * P->K + 2.0 * P->K) / (4.0 * prim.s3 * prim.s3);
}
/* moment_uv, clUGKS.c:82-93 */
static od4 moment_uv(const double *Mu, const double *Mv, const double *Mxi, int a, int b, int d) {
    od4 m;
    m.s0 = Mu[a] * Mv[b] * Mxi[d / 2];
    m.s1 = Mu[a + 1] * Mv[b] * Mxi[d / 2];
    m.s2 = Mu[a] * Mv[b + 1] * Mxi[d / 2];
    m.s3 = 0.5 * (Mu[a + 2] * Mv[b] * Mxi[d / 2] + Mu[a] * Mv[b + 2] * Mxi[d / 2] + Mu[a] * Mv[b] * Mxi[(d + 2) / 2]);
    return m;
}
/* moment_au, clUGKS.c:95-104 */
static od4 moment_au(od4 a, const double *Mu, const double *Mv, const double *Mxi, int al, int be) {
    od4 r = smul4(a.s0, moment_uv(Mu, Mv, Mxi, al + 0, be + 0, 0));
    r = add4(r, smul4(a.s1, moment_uv(Mu, Mv, Mxi, al + 1, be + 0, 0)));
    r = add4(r, smul4(a.s2, moment_uv(Mu, Mv, Mxi, al + 0, be + 1, 0)));
    r = add4(r, smul4(0.5 * a.s3, moment_uv(Mu, Mv, Mxi, al + 2, be + 0, 0)));
    r = add4(r, smul4(0.5 * a.s3, moment_uv(Mu, Mv, Mxi, al + 0, be + 2, 0)));
    r = add4(r, smul4(0.5 * a.s3, moment_uv(Mu, Mv, Mxi, al + 0, be + 0, 2)));
    return r;
}

/* Mt[0..4], clUGKS.c:499-503 (same in :622-626, :725-729) */
void or_timeIntegrals(double tau, double dt, double *Mt) {
    Mt[3] = tau * (1.0 - exp(-dt / tau));
    Mt[4] = -tau * dt * exp(-dt / tau) + tau * Mt[3];
    Mt[0] = dt - Mt[3];
    Mt[1] = -tau * Mt[0] + Mt[4];
    Mt[2] = (dt * dt) / 2.0 - tau * Mt[0];
}

/* interior-face domain of the flux kernels (clUGKS.c:218-225, ugks_block.py:1105) */
static void face_range(const or_params *P, int face, int ob, int ot, int *i0, int *i1, int *j0, int *j1) {
    if (face == SOUTH) { *i0 = 0; *i1 = P->ni; *j0 = ob; *j1 = P->nj + 1 - ot; }
    else { *i0 = ob; *i1 = P->ni + 1 - ot; *j0 = 0; *j1 = P->nj; }
}

/* ==== kernels ============================================================ */

/* cellGeom, clKernel.c:31-135; launched over (Ni, Nj) (ugks_block.py:748) */
void or_cellGeom(const or_params *P, const od2 *xy, double *area, od2 *centre,
                 od2 *mid_side, od2 *normal, double *side_length) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 0; i < NI_; i++)
        for (int j = 0; j < NJ_; j++) {
            double xA = XY(i + 1, j).x, yA = XY(i + 1, j).y;
            double xB = XY(i + 1, j + 1).x, yB = XY(i + 1, j + 1).y;
            double xC = XY(i, j + 1).x, yC = XY(i, j + 1).y;
            double xD = XY(i, j).x, yD = XY(i, j).y;
            od2 tS = v2(xA - xD, yA - yD), tW = v2(xD - xC, yD - yC);
            od2 nS = v2(-tS.y, tS.x), nW = v2(-tW.y, tW.x);
            FACEA(normal, i, j, SOUTH) = div2s(nS, len2(nS));
            FACEA(normal, i, j, WEST) = div2s(nW, len2(nW));
            double xyarea = 0.5 * ((xB + xA) * (yB - yA) + (xC + xB) * (yC - yB) +
                                   (xD + xC) * (yD - yC) + (xA + xD) * (yA - yD));
            C2(area, i, j) = xyarea;
            double cx = 1.0 / (xyarea * 6.0) *
                        ((yB - yA) * (xA * xA + xA * xB + xB * xB) + (yC - yB) * (xB * xB + xB * xC + xC * xC) +
                         (yD - yC) * (xC * xC + xC * xD + xD * xD) + (yA - yD) * (xD * xD + xD * xA + xA * xA));
            if (fabs(cx) > 1.0e6) cx = NAN;
            double cy = -1.0 / (xyarea * 6.0) *
                        ((xB - xA) * (yA * yA + yA * yB + yB * yB) + (xC - xB) * (yB * yB + yB * yC + yC * yC) +
                         (xD - xC) * (yC * yC + yC * yD + yD * yD) + (xA - xD) * (yD * yD + yD * yA + yA * yA));
            if (fabs(cy) > 1.0e6) cy = NAN;
            C2(centre, i, j) = v2(cx, cy);
            FACEA(mid_side, i, j, SOUTH) = v2(0.5 * (xD + xA), 0.5 * (yD + yA));
            FACEA(mid_side, i, j, WEST) = v2(0.5 * (xD + xC), 0.5 * (yD + yC));
            FACEA(side_length, i, j, SOUTH) = sqrt((xD - xA) * (xD - xA) + (yD - yA) * (yD - yA));
            FACEA(side_length, i, j, WEST) = sqrt((xC - xD) * (xC - xD) + (yC - yD) * (yC - yD));
        }
}

/* paraBC with constant WALL_* values, clKernel.c:141-218 */
void or_paraBC(const or_params *P, od4 *wall_prop) {
    int wl = P->ni > P->nj ? P->ni : P->nj;
    for (int ci = 0; ci < wl; ci++)
        for (int w = 0; w < 4; w++) {
            int along = (w == GNORTH || w == GSOUTH) ? P->ni : P->nj;
            if (ci >= along) continue;
            od4 prim;
            prim.s0 = P->wall_D[w]; prim.s1 = P->wall_U[w]; prim.s2 = P->wall_V[w];
            if (P->wall_P[w] != -1) prim.s3 = P->wall_P[w] / P->wall_D[w];
            else prim.s3 = P->wall_T[w];
            WP(wall_prop, w, ci) = prim;
        }
}

/* clFindDT, clKernel.c:224-308 */
void or_clFindDT(const or_params *P, const od2 *xy, const double *area, const od4 *macro,
                 double *time_step, double Kn) {
    const int G = G_;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = 0; mi < P->ni; mi++)
        for (int mj = 0; mj < P->nj; mj++) {
            od2 a = XY(mi + 1 + G, mj + G), b = XY(mi + 1 + G, mj + 1 + G);
            od2 c = XY(mi + G, mj + 1 + G), d = XY(mi + G, mj + G);
            double dx = fabs(a.x - b.x), t;
            t = fabs(c.x - b.x); if (t > dx) dx = t;
            t = fabs(c.x - d.x); if (t > dx) dx = t;
            t = fabs(a.x - d.x); if (t > dx) dx = t;
            t = fabs(c.x - a.x); if (t > dx) dx = t;
            t = fabs(b.x - d.x); if (t > dx) dx = t;
            double dy = fabs(a.y - b.y);
            t = fabs(c.y - b.y); if (t > dy) dy = t;
            t = fabs(c.y - d.y); if (t > dy) dy = t;
            t = fabs(a.y - d.y); if (t > dy) dy = t;
            t = fabs(c.y - a.y); if (t > dy) dy = t;
            t = fabs(b.y - d.y); if (t > dy) dy = t;
            od4 m = M2(macro, mi, mj);
            double sos = sqrt(0.5 * P->gam / m.s3);
            double lo_u = fmin(P->umax, fabs(m.s1) + sos), lo_v = fmin(P->vmax, fabs(m.s2) + sos);
            double hi_u = fmax(P->umax, fabs(m.s1) + sos), hi_v = fmax(P->vmax, fabs(m.s2) + sos);
            double sigma = 10, offset = 1e-3;
            double scale = (erf(sigma * (Kn - offset)) + 1) / 2.0;
            double u = scale * (hi_u - lo_u) + lo_u;
            double v = scale * (hi_v - lo_v) + lo_v;
            M2(time_step, mi, mj) = (dx * u + dy * v) / C2(area, mi + G, mj + G);
        }
}

/* initFunctions, clKernel.c:382-412 */
void or_initFunctions(const or_params *P, od2 *Fin, const od4 *macro, const od2 *gQ) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = 0; mi < P->ni; mi++)
        for (int mj = 0; mj < P->nj; mj++)
            for (int gv = 0; gv < NV_; gv++)
                F3(Fin, mi + G_, mj + G_, gv) = fEQ(P, M2(macro, mi, mj), M2(gQ, mi, mj), quad_at(P, gv), gv);
}

/* moment accumulation used by getFaceCons/getAL/getAR/calcMacro
   (clUGKS.c:240-242, clKernel.c:441-443) */
static inline void acc_moments(od4 *S, od2 f, od2 uv) {
    S->s0 += f.x;
    S->s1 += uv.x * f.x;
    S->s2 += uv.y * f.x;
    S->s3 += 0.5 * (dot2(uv, uv) * f.x + f.y);
}

/* calcMacro, clKernel.c:414-478 */
void or_calcMacro(const or_params *P, const od2 *Fin, od4 *macro) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = 0; mi < P->ni; mi++)
        for (int mj = 0; mj < P->nj; mj++) {
            od4 S = zero4();
            for (int gv = 0; gv < NV_; gv++) acc_moments(&S, F3(Fin, mi + G_, mj + G_, gv), quad_at(P, gv));
            od4 prim = S;
            prim.s3 = 0.5 * prim.s0 / (P->gam - 1.0) / (prim.s3 - 0.5 * (prim.s1 * prim.s1 + prim.s2 * prim.s2) / prim.s0);
            prim.s1 /= prim.s0; prim.s2 /= prim.s0;
            M2(macro, mi, mj) = prim;
        }
}

/* heat-flux accumulation, clKernel.c:513-514 and clUGKS.c:555-556 */
static inline void acc_q(od2 *Qs, od2 f, od2 uv, od4 prim, double K) {
    od2 c = sub2(uv, v2(prim.s1, prim.s2));
    Qs->x += 0.5 * ((uv.x - prim.s1) * (dot2(c, c) * f.x + (f.y / K)));
    Qs->y += 0.5 * ((uv.y - prim.s2) * (dot2(c, c) * f.x + (f.y / K)));
}

/* calcQ, clKernel.c:480-543 */
void or_calcQ(const or_params *P, const od2 *Fin, const od4 *macro, od2 *gQ) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = 0; mi < P->ni; mi++)
        for (int mj = 0; mj < P->nj; mj++) {
            od2 Qs = v2(0.0, 0.0);
            od4 prim = M2(macro, mi, mj);
            for (int gv = 0; gv < NV_; gv++) acc_q(&Qs, F3(Fin, mi + G_, mj + G_, gv), quad_at(P, gv), prim, P->K);
            M2(gQ, mi, mj) = Qs;
        }
}

/* getInternal, clKernel.c:315-372 */
void or_getInternal(const or_params *P, const od2 *Fin, od4 *Txyz, od4 *Stress) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = 0; mi < P->ni; mi++)
        for (int mj = 0; mj < P->nj; mj++) {
            double D = 0.0, Su = 0.0, Sv = 0.0, Suu = 0.0, Svv = 0.0, Suv = 0.0, H = 0.0;
            for (int gv = 0; gv < NV_; gv++) {
                od2 uv = quad_at(P, gv), f = F3(Fin, mi + G_, mj + G_, gv);
                D += f.x;
                Su += uv.x * f.x;
                Sv += uv.y * f.x;
                Suu += uv.x * uv.x * f.x;
                Svv += uv.y * uv.y * f.x;
                Suv += uv.x * uv.y * f.x;
                H += f.y;
            }
            double U = Su / D, V = Sv / D;
            double Tx = 2 * (Suu / D - U * U), Ty = 2 * (Svv / D - V * V), Tz = 2 * (H / D);
            double T = (Tx + Ty + Tz) / P->B;
            M2(Txyz, mi, mj) = v4(Tx, Ty, Tz, T);
            M2(Stress, mi, mj) = v4(Suu - D * (U * U + T / 2.0), Suv - D * (U * V + T / 2.0), Svv - D * (V * V + T / 2.0), 0.0);
        }
}

/* iFace, clUGKS.c:110-199 */
void or_iFace(const or_params *P, const od2 *Fin, od2 *iface_f, od2 *fsigma,
              const od2 *centre, const od2 *mid_side, const od2 *normal, int face) {
    const int SL = P->flux_method == 1 ? 5 : 3, MID = P->flux_method == 1 ? 2 : 1;
    int ilim = face == SOUTH ? P->ni : P->ni + 1, jlim = face == SOUTH ? P->nj + 1 : P->nj;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = 0; mi < ilim; mi++)
        for (int mj = 0; mj < jlim; mj++) {
            int gi = mi + G_, gj = mj + G_;
            od2 n = FACEA(normal, gi, gj, face);
            for (int gv = 0; gv < NV_; gv++) {
                od2 uv = interfaceVelocity(P, gv, n);
                int delta = upwind_delta(uv.x);
                od2 cs[5], fs[5];
                for (int si = 0; si < SL; si++) {
                    int sx, sy;
                    if (face == SOUTH) {
                        sx = gi;
                        sy = delta * (gj - MID - 1 + si) + (1 - delta) * (gj + MID - si);
                    } else {
                        sx = delta * (gi - MID - 1 + si) + (1 - delta) * (gi + MID - si);
                        sy = gj;
                    }
                    fs[si] = F3(Fin, sx, sy, gv);
                    cs[si] = C2(centre, sx, sy);
                }
                double idist = len2(sub2(FACEA(mid_side, gi, gj, face), cs[MID]));
                od2 sigma;
                if (P->flux_method == 0) {
                    od2 s1 = div2s(sub2(fs[MID], fs[MID - 1]), len2(sub2(cs[MID], cs[MID - 1])));
                    od2 s2 = div2s(sub2(fs[MID + 1], fs[MID]), len2(sub2(cs[MID - 1], cs[MID])));
                    sigma = or_vanLeer(s1, s2);
                } else if (P->flux_method == 1) {
                    sigma = or_WENO5(fs, idist);
                } else {
                    sigma = or_NND(fs, idist);
                }
                F3(iface_f, gi, gj, gv) = add2(fs[MID], mul2s(sigma, idist));
                F3(fsigma, gi, gj, gv) = sigma;
            }
        }
}

/* getFaceCons, clUGKS.c:209-271 */
void or_getFaceCons(const or_params *P, const od2 *iface_f, const od2 *normal, int face,
                    od4 *primary, int ob, int ot) {
    int i0, i1, j0, j1;
    face_range(P, face, ob, ot, &i0, &i1, &j0, &j1);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = i0; mi < i1; mi++)
        for (int mj = j0; mj < j1; mj++) {
            int gi = mi + G_, gj = mj + G_;
            od4 S = zero4();
            for (int gv = 0; gv < NV_; gv++)
                acc_moments(&S, F3(iface_f, gi, gj, gv), interfaceVelocity(P, gv, FACEA(normal, gi, gj, face)));
            C2(primary, gi, gj) = S;
        }
}

/* getAL, clUGKS.c:273-335 ; getAR, clUGKS.c:337-399 */
static void cell_moments_face(const or_params *P, const od2 *Fin, const od2 *normal, int face,
                              od4 *out, int ob, int ot, int left) {
    int i0, i1, j0, j1;
    face_range(P, face, ob, ot, &i0, &i1, &j0, &j1);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = i0; mi < i1; mi++)
        for (int mj = j0; mj < j1; mj++) {
            int gi = mi + G_, gj = mj + G_;
            int ci = left ? gi - face : gi, cj = left ? gj - (1 - face) : gj;
            od4 S = zero4();
            for (int gv = 0; gv < NV_; gv++)
                acc_moments(&S, F3(Fin, ci, cj, gv), interfaceVelocity(P, gv, FACEA(normal, gi, gj, face)));
            C2(out, gi, gj) = S;
        }
}
void or_getAL(const or_params *P, const od2 *Fin, const od2 *normal, int face, od4 *gaL, int ob, int ot) {
    cell_moments_face(P, Fin, normal, face, gaL, ob, ot, 1);
}
void or_getAR(const or_params *P, const od2 *Fin, const od2 *normal, int face, od4 *gaR, int ob, int ot) {
    cell_moments_face(P, Fin, normal, face, gaR, ob, ot, 0);
}

/* getPLR, clUGKS.c:401-442 */
void or_getPLR(const or_params *P, const od2 *centre, const od2 *mid_side, int face,
               od4 *primary, od4 *gaL, od4 *gaR, int ob, int ot) {
    int i0, i1, j0, j1;
    face_range(P, face, ob, ot, &i0, &i1, &j0, &j1);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = i0; mi < i1; mi++)
        for (int mj = j0; mj < j1; mj++) {
            int gi = mi + G_, gj = mj + G_;
            od4 w = C2(primary, gi, gj);
            od4 prim = or_getPrimary(P, w);
            od2 midside = FACEA(mid_side, gi, gj, face);
            od4 sw = div4s(sub4(w, C2(gaL, gi, gj)), len2(sub2(midside, C2(centre, gi - face, gj - (1 - face)))));
            C2(gaL, gi, gj) = or_microSlope(P, prim, sw);
            sw = div4s(sub4(C2(gaR, gi, gj), w), len2(sub2(midside, C2(centre, gi, gj))));
            C2(gaR, gi, gj) = or_microSlope(P, prim, sw);
            C2(primary, gi, gj) = prim;
        }
}

/* initMacroFlux, clUGKS.c:446-518 */
void or_initMacroFlux(const or_params *P, od4 *flux_macro, int face, double dt,
                      const od4 *primary, const od4 *gaL, const od4 *gaR, od4 *gaT,
                      double *gMxi, int ob, int ot) {
    int i0, i1, j0, j1;
    face_range(P, face, ob, ot, &i0, &i1, &j0, &j1);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = i0; mi < i1; mi++)
        for (int mj = j0; mj < j1; mj++) {
            int gi = mi + G_, gj = mj + G_;
            od4 prim = C2(primary, gi, gj), aL = C2(gaL, gi, gj), aR = C2(gaR, gi, gj);
            double Mu[7], Mv[5], Mxi[3], Mu_L[7], Mu_R[7];
            or_momentU(P, prim, Mu, Mv, Mxi, Mu_L, Mu_R);
            for (int k = 0; k < 3; k++) gMxi[3 * (size_t)NJ_ * gi + 3 * (size_t)gj + k] = Mxi[k];
            od4 Mau_L = moment_au(aL, Mu_L, Mv, Mxi, 1, 0);
            od4 Mau_R = moment_au(aR, Mu_R, Mv, Mxi, 1, 0);
            od4 sw = smul4(-prim.s0, add4(Mau_L, Mau_R));
            od4 aT = or_microSlope(P, prim, sw);
            C2(gaT, gi, gj) = aT;
            double tau = or_relaxTime(P, prim);
            double Mt[5];
            or_timeIntegrals(tau, dt, Mt);
            od4 Mau_0 = moment_uv(Mu, Mv, Mxi, 1, 0, 0);
            Mau_L = moment_au(aL, Mu_L, Mv, Mxi, 2, 0);
            Mau_R = moment_au(aR, Mu_R, Mv, Mxi, 2, 0);
            od4 Mau_T = moment_au(aT, Mu, Mv, Mxi, 1, 0);
            od4 t = add4(add4(smul4(Mt[0], Mau_0), smul4(Mt[1], add4(Mau_L, Mau_R))), smul4(Mt[2], Mau_T));
            C2(flux_macro, gi, gj) = smul4(prim.s0, t);
        }
}

/* calcFaceQ, clUGKS.c:522-585 */
void or_calcFaceQ(const or_params *P, const od2 *iface_f, const od4 *primary,
                  const od2 *normal, od2 *faceQ, int face, int ob, int ot) {
    int i0, i1, j0, j1;
    face_range(P, face, ob, ot, &i0, &i1, &j0, &j1);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = i0; mi < i1; mi++)
        for (int mj = j0; mj < j1; mj++) {
            int gi = mi + G_, gj = mj + G_;
            od2 n = FACEA(normal, gi, gj, face);
            od4 prim = C2(primary, gi, gj);
            od2 Qs = v2(0.0, 0.0);
            for (int gv = 0; gv < NV_; gv++) acc_q(&Qs, F3(iface_f, gi, gj, gv), interfaceVelocity(P, gv, n), prim, P->K);
            C2(faceQ, gi, gj) = Qs;
        }
}

/* macroFlux, clUGKS.c:587-681 */
void or_macroFlux(const or_params *P, const od2 *flux_f, const od2 *fsigma, od4 *flux_macro,
                  const od2 *normal, const double *side_length, int face, double dt,
                  const od4 *primary, const od2 *faceQ, int ob, int ot) {
    int i0, i1, j0, j1;
    face_range(P, face, ob, ot, &i0, &i1, &j0, &j1);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = i0; mi < i1; mi++)
        for (int mj = j0; mj < j1; mj++) {
            int gi = mi + G_, gj = mj + G_;
            od2 n = FACEA(normal, gi, gj, face);
            od4 prim = C2(primary, gi, gj);
            double tau = or_relaxTime(P, prim), Mt[5];
            or_timeIntegrals(tau, dt, Mt);
            od4 S = zero4();
            od2 Q = C2(faceQ, gi, gj);
            for (int v = 0; v < NV_; v++) {
                od2 uv = interfaceVelocity(P, v, n);
                od2 F0 = or_fS(P, prim, Q, uv, or_fM(P, prim, uv, v));
                od2 fd = F3(flux_f, gi, gj, v), fsl = F3(fsigma, gi, gj, v);
                S.s0 += Mt[0] * uv.x * F0.x + Mt[3] * uv.x * fd.x - Mt[4] * uv.x * uv.x * fsl.x;
                S.s1 += Mt[0] * uv.x * uv.x * F0.x + Mt[3] * uv.x * uv.x * fd.x - Mt[4] * uv.x * uv.x * uv.x * fsl.x;
                S.s2 += Mt[0] * uv.y * uv.x * F0.x + Mt[3] * uv.y * uv.x * fd.x - Mt[4] * uv.y * uv.x * uv.x * fsl.x;
                S.s3 += Mt[0] * 0.5 * (uv.x * dot2(uv, uv) * F0.x + uv.x * F0.y) +
                        Mt[3] * 0.5 * (uv.x * dot2(uv, uv) * fd.x + uv.x * fd.y) -
                        Mt[4] * 0.5 * (uv.x * uv.x * dot2(uv, uv) * fsl.x + uv.x * uv.x * fsl.y);
            }
            od4 out = add4(S, C2(flux_macro, gi, gj));
            od2 g = toGlobal(v2(out.s1, out.s2), n);
            out.s1 = g.x; out.s2 = g.y;
            C2(flux_macro, gi, gj) = smul4(FACEA(side_length, gi, gj, face), out);
        }
}

/* distFlux, clUGKS.c:683-772 */
void or_distFlux(const or_params *P, od2 *flux_f, const od2 *fsigma, const od2 *normal,
                 const double *side_length, int face, double dt, const od4 *primary,
                 const od4 *gaL, const od4 *gaR, const od4 *gaT, const double *gMxi,
                 const od2 *faceQ, int ob, int ot) {
    int i0, i1, j0, j1;
    face_range(P, face, ob, ot, &i0, &i1, &j0, &j1);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = i0; mi < i1; mi++)
        for (int mj = j0; mj < j1; mj++) {
            int gi = mi + G_, gj = mj + G_;
            od4 prim = C2(primary, gi, gj), aL = C2(gaL, gi, gj), aR = C2(gaR, gi, gj), aT = C2(gaT, gi, gj);
            double Mxi2 = gMxi[3 * (size_t)NJ_ * gi + 3 * (size_t)gj + 2];
            double tau = or_relaxTime(P, prim), Mt[5];
            or_timeIntegrals(tau, dt, Mt);
            od2 n = FACEA(normal, gi, gj, face);
            od2 Q = C2(faceQ, gi, gj);
            double L = FACEA(side_length, gi, gj, face);
            for (int gv = 0; gv < NV_; gv++) {
                od2 uv = interfaceVelocity(P, gv, n);
                int delta = upwind_delta(uv.x);
                od2 f0 = or_fM(P, prim, uv, gv);
                od2 F0 = or_fS(P, prim, Q, uv, f0);
                od2 fd = F3(flux_f, gi, gj, gv), fsl = F3(fsigma, gi, gj, gv);
                double uu = dot2(uv, uv), ux2 = uv.x * uv.x;
                od2 ff;
                ff.x = Mt[0] * uv.x * (f0.x + F0.x) +
                       Mt[1] * ux2 * (aL.s0 * f0.x + aL.s1 * uv.x * f0.x + aL.s2 * uv.y * f0.x + 0.5 * aL.s3 * (uu * f0.x + f0.y)) * delta +
                       Mt[1] * ux2 * (aR.s0 * f0.x + aR.s1 * uv.x * f0.x + aR.s2 * uv.y * f0.x + 0.5 * aR.s3 * (uu * f0.x + f0.y)) * (1 - delta) +
                       Mt[2] * uv.x * (aT.s0 * f0.x + aT.s1 * uv.x * f0.x + aT.s2 * uv.y * f0.x + 0.5 * aT.s3 * (uu * f0.x + f0.y)) +
                       Mt[3] * uv.x * fd.x - Mt[4] * ux2 * fsl.x;
                ff.y = Mt[0] * uv.x * (f0.y + F0.y) +
                       Mt[1] * ux2 * (aL.s0 * f0.y + aL.s1 * uv.x * f0.y + aL.s2 * uv.y * f0.y + 0.5 * aL.s3 * (uu * f0.y + Mxi2 * f0.x)) * delta +
                       Mt[1] * ux2 * (aR.s0 * f0.y + aR.s1 * uv.x * f0.y + aR.s2 * uv.y * f0.y + 0.5 * aR.s3 * (uu * f0.y + Mxi2 * f0.x)) * (1 - delta) +
                       Mt[2] * uv.x * (aT.s0 * f0.y + aT.s1 * uv.x * f0.y + aT.s2 * uv.y * f0.y + 0.5 * aT.s3 * (uu * f0.y + Mxi2 * f0.x)) +
                       Mt[3] * uv.x * fd.y - Mt[4] * ux2 * fsl.y;
                F3(flux_f, gi, gj, gv) = smul2(L, ff);
            }
        }
}

/* the wall-face addressing shared by the wall kernels (clUGKS.c:795-824) */
static int wall_site(const or_params *P, int face, int ci, int *gi, int *gj, int *rot, int *face_id) {
    switch (face) {
    case GNORTH: *gi = ci + G_; *gj = NJ_ - G_; *rot = -1; *face_id = SOUTH; return ci < P->ni;
    case GEAST:  *gi = NI_ - G_; *gj = ci + G_; *rot = -1; *face_id = WEST;  return ci < P->nj;
    case GSOUTH: *gi = ci + G_; *gj = G_;       *rot = 1;  *face_id = SOUTH; return ci < P->ni;
    default:     *gi = G_;      *gj = ci + G_;  *rot = 1;  *face_id = WEST;  return ci < P->nj;
    }
}
static int wall_count(const or_params *P, int face) { return (face == GNORTH || face == GSOUTH) ? P->ni : P->nj; }

/* accommodatingWallDist, clUGKS.c:776-906 */
void or_accommodatingWallDist(const or_params *P, const od2 *normal, int face,
                              const od4 *wall_prop, od2 *flux_f, double dt) {
    (void)dt;
    int nw = wall_count(P, face);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int ci = 0; ci < nw; ci++) {
        int gi, gj, rot, fid;
        wall_site(P, face, ci, &gi, &gj, &rot, &fid);
        od2 n = FACEA(normal, gi, gj, fid);
        od4 wall = WP(wall_prop, face, ci);
        od2 wl = toLocal(v2(wall.s1, wall.s2), n);
        wall.s1 = wl.x; wall.s2 = wl.y;
        od2 data = v2(0.0, 0.0);
        for (int gv = 0; gv < NV_; gv++) {
            od2 uv = interfaceVelocity(P, gv, n);
            od2 fd = F3(flux_f, gi, gj, gv);
            int delta = wall_delta(uv.x, rot);
            data.x += uv.x * (1 - delta) * fd.x;
            od2 wd = or_fM(P, wall, uv, gv);
            data.y -= uv.x * delta * wd.x;
        }
        double ratio = data.x / data.y;
        for (int gv = 0; gv < NV_; gv++) {
            od2 uv = interfaceVelocity(P, gv, n);
            int delta = wall_delta(uv.x, rot);
            od2 fd = F3(flux_f, gi, gj, gv);
            od2 wd = add2(smul2(ratio * delta, or_fM(P, wall, uv, gv)), smul2((double)(1 - delta), fd));
            F3(flux_f, gi, gj, gv) = wd;
        }
    }
}

/* wallFlux, clUGKS.c:910-1011 */
void or_wallFlux(const or_params *P, const od2 *normal, const double *side_length, int face,
                 od2 *flux_f, od4 *flux_macro, double dt) {
    int nw = wall_count(P, face);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int ci = 0; ci < nw; ci++) {
        int gi, gj, rot, fid;
        wall_site(P, face, ci, &gi, &gj, &rot, &fid);
        od2 n = FACEA(normal, gi, gj, fid);
        double L = FACEA(side_length, gi, gj, fid);
        od4 S = zero4();
        for (int gv = 0; gv < NV_; gv++) {
            od2 uv = interfaceVelocity(P, gv, n);
            od2 wd = F3(flux_f, gi, gj, gv);
            F3(flux_f, gi, gj, gv) = smul2(dt, smul2(L, smul2(uv.x, wd)));
            S.s0 += uv.x * wd.x;
            S.s1 += uv.x * uv.x * wd.x;
            S.s2 += uv.x * uv.y * wd.x;
            S.s3 += 0.5 * uv.x * (dot2(uv, uv) * wd.x + wd.y);
        }
        od2 g = toGlobal(v2(S.s1, S.s2), n);
        S.s1 = g.x; S.s2 = g.y;
        C2(flux_macro, gi, gj) = smul4(dt * L, S);
    }
}

/* wallMassEnergyFluxes, clUGKS.c:1019-1147 ; WALL_FLUX(w,i) = wall_flux[i*4+w] (double8) */
void or_wallMassEnergyFluxes(const or_params *P, const od2 *normal, const double *side_length,
                             const double *area, int face, const od2 *flux_f,
                             double *wall_flux, double dt) {
    (void)side_length; (void)area; (void)dt;
    int nw = wall_count(P, face);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int ci = 0; ci < nw; ci++) {
        int gi, gj, rot, fid;
        wall_site(P, face, ci, &gi, &gj, &rot, &fid);
        od2 n = FACEA(normal, gi, gj, fid);
        od4 din = zero4(), dout = zero4();
        for (int gv = 0; gv < NV_; gv++) {
            od2 uv = interfaceVelocity(P, gv, n);
            int delta = wall_delta(uv.x, rot);
            od2 wd = F3(flux_f, gi, gj, gv);
            od2 fin = smul2((double)(1 - delta), wd), fout = smul2((double)delta, wd);
            din.s0 += fin.x; din.s1 += uv.x * fin.x; din.s2 += uv.y * fin.x;
            din.s3 += 0.5 * (dot2(uv, uv) * fin.x + fin.y);
            dout.s0 += fout.x; dout.s1 += uv.x * fout.x; dout.s2 += uv.y * fout.x;
            dout.s3 += 0.5 * (dot2(uv, uv) * fout.x + fout.y);
        }
        od4 mi = smul4((double)(-rot), din), mo = smul4((double)rot, dout);
        double *o = wall_flux + ((size_t)ci * 4 + (size_t)face) * 8;
        o[0] = mi.s0; o[1] = mo.s0; o[2] = mi.s1; o[3] = mo.s1;
        o[4] = mi.s2; o[5] = mo.s2; o[6] = mi.s3; o[7] = mo.s3;
    }
}

/* stickingProbability, clAdsorb.c:6-16 */
static double stickingProbability(const or_params *P, od2 uv, double T, int face) {
    double normal = 1.0 - P->beta_n[face] * exp(-uv.x * uv.x / T);
    double tangential = 1.0 - P->beta_t[face] * (1.0 - exp(-uv.y * uv.y / T));
    return normal * tangential;
}

/* intersect, clAdsorb.c:53-71: do segments ab and cd cross?  The reference
   forms the determinants in double and divides in single precision. */
static int seg_intersect(od2 a, od2 b, od2 c, od2 d) {
    float den = (float)((d.y - c.y) * (b.x - a.x) - (d.x - c.x) * (b.y - a.y));
    float num1 = (float)((d.x - c.x) * (a.y - c.y) - (d.y - c.y) * (a.x - c.x));
    float num2 = (float)((b.x - a.x) * (a.y - c.y) - (b.y - a.y) * (a.x - c.x));
    if (den == 0.0f) return 0;                /* parallel or coincident */
    float u1 = num1 / den, u2 = num2 / den;
    if (u1 < 0.0f || u1 > 1.0f || u2 < 0.0f || u2 > 1.0f) return 0;
    return 1;
}

/* distance_from_segment, clAdsorb.c:18-51: distance of c from segment ab */
static double seg_distance(od2 a, od2 b, od2 c) {
    double rn = (c.x - a.x) * (b.x - a.x) + (c.y - a.y) * (b.y - a.y);
    double rd = (b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y);
    double r = rn / rd;
    double s = ((a.y - c.y) * (b.x - a.x) - (a.x - c.x) * (b.y - a.y)) / rd;
    if (r >= 0 && r <= 1) return fabs(s) * sqrt(rd);
    double d1 = (c.x - a.x) * (c.x - a.x) + (c.y - a.y) * (c.y - a.y);
    double d2 = (c.x - b.x) * (c.x - b.x) + (c.y - b.y) * (c.y - b.y);
    return d1 < d2 ? sqrt(d1) : sqrt(d2);
}

/* vartheta_langmuir for a tabulated isotherm, clAdsorb.c:171-247 with
   get_tri (:73-157) and interpolate (:159-168): walk the triangulation from
   triangle 0 towards pt = (P, T), crossing the edge that the segment from the
   current triangle's centroid to pt hits (edge bc, then ca, then ab), for at
   most n_tri moves; stop inside, on an edge (1e-12), or at the hull (then
   *outside = 1); barycentric interpolation of theta in the final triangle. */
double or_langmuir_table(const or_params *P, int face, double Pr, double T, int *outside) {
    const double *iso = P->iso[face];
    const int32_t *tri = P->iso_tri[face], *nbr = P->iso_nbr[face];
    od2 pt = v2(Pr, T);
    int id = 0;
    *outside = 0;
    for (int counter = 0; counter < P->n_tri[face]; counter++) {
        const int32_t *t = tri + 3 * id;
        od2 a = v2(iso[3 * t[0]], iso[3 * t[0] + 1]);
        od2 b = v2(iso[3 * t[1]], iso[3 * t[1] + 1]);
        od2 c = v2(iso[3 * t[2]], iso[3 * t[2] + 1]);
        od2 start = smul2(1.0 / 3.0, add2(add2(a, b), c));
        int hz = seg_intersect(a, b, start, pt);
        int hx = seg_intersect(b, c, start, pt);
        int hy = seg_intersect(c, a, start, pt);
        if (hx + hy + hz == 0) break;
        const int32_t *nb = nbr + 3 * id;
        int next;
        od2 e0, e1;
        if (hx) { next = nb[0]; e0 = b; e1 = c; }
        else if (hy) { next = nb[1]; e0 = c; e1 = a; }
        else { next = nb[2]; e0 = a; e1 = b; }
        if (next == -1) { *outside = 1; break; }
        if (seg_distance(e0, e1, pt) < 1e-12) break;
        id = next;
    }
    const int32_t *t = tri + 3 * id;
    const double *A = iso + 3 * t[0], *B = iso + 3 * t[1], *Cc = iso + 3 * t[2];
    double den = (B[1] - Cc[1]) * (A[0] - Cc[0]) + (Cc[0] - B[0]) * (A[1] - Cc[1]);
    double L1 = ((B[1] - Cc[1]) * (pt.x - Cc[0]) + (Cc[0] - B[0]) * (pt.y - Cc[1])) / den;
    double L2 = ((Cc[1] - A[1]) * (pt.x - Cc[0]) + (A[0] - Cc[0]) * (pt.y - Cc[1])) / den;
    double L3 = 1 - L1 - L2;
    return L1 * A[2] + L2 * B[2] + L3 * Cc[2];
}

/* adsorbingWall_P1, clAdsorb.c:253-482, with the constant or the tabulated
   isotherm (clAdsorb.c:171-247) */
void or_adsorbingWall_P1(const or_params *P, const od2 *normal, int face, const od4 *wall_prop,
                         od4 *wall_cover, od2 *wall_dist, const od2 *flux_f, const od4 *macro,
                         double dt, double *flag) {
    (void)wall_prop;
    int nw = wall_count(P, face);
    int err = 0;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(| : err)
    for (int ci = 0; ci < nw; ci++) {
        int gi, gj, rot, fid;
        wall_site(P, face, ci, &gi, &gj, &rot, &fid);
        double T, D;
        switch (face) {
        case GNORTH: T = 1.0 / M2(macro, ci, gj - G_ - 1).s3; D = M2(macro, ci, gj - G_ - 1).s0; break;
        case GEAST:  T = 1.0 / M2(macro, gi - G_ - 1, ci).s3; D = M2(macro, gi - G_ - 1, ci).s0; break;
        case GSOUTH: T = 1.0 / M2(macro, ci, gj - G_).s3; D = M2(macro, ci, gj - G_).s0; break;
        default:     T = 1.0 / M2(macro, gi - G_, ci).s3; D = M2(macro, gi - G_, ci).s0; break;
        }
        od2 n = FACEA(normal, gi, gj, fid);
        double vartheta = WP(wall_cover, face, ci).s0;
        od2 data = v2(0.0, 0.0);
        double adjusted = 0.0;
        for (int gv = 0; gv < NV_; gv++) {
            od2 uv = interfaceVelocity(P, gv, n);
            od2 fd = F3(flux_f, gi, gj, gv);
            int delta = wall_delta(uv.x, rot);
            od2 total = smul2(-rot * uv.x * (1 - delta), fd);
            double beta = stickingProbability(P, uv, T, face);
            od2 ads = v2(beta * total.x * P->gamma_f[face] * (1.0 - vartheta),
                     beta * total.y * P->gamma_f[face] * (1.0 - vartheta));
            int flux_sign = (int)osign(total.x);
            if ((flux_sign * ads.x) > (flux_sign * total.x)) ads = total;
            od2 refl = sub2(total, ads);
            data.x += refl.x;
            data.y += ads.x;
            adjusted += beta * total.x;
            wall_dist[(size_t)ci * NV_ + gv] = div2s(refl, uv.x);
        }
        double correct = 1.0;
        double r_init = data.x, a_init = data.y;
        double r_final = r_init, a_final = a_init;
        double dvtheta, des = 0.0;
        int early = 0;
        if (P->gamma_f[face] > 0.0) {
            if (P->dynamic_vartheta[face] == 1) {
                dvtheta = dt * P->alpha_p[face] * a_init;
                if (dvtheta > (1 - vartheta)) {
                    dvtheta = 1 - vartheta;
                    a_final = dvtheta / dt * P->alpha_p[face];
                }
            } else {
                dvtheta = 0.0;
                a_final = a_init;
            }
            double veq_x, veq_y;
            if (P->constant_adsorb[face]) { veq_x = P->adsorb[face]; veq_y = 0.0; }
            else if (P->n_tri[face] > 0) {
                int out;
                veq_x = or_langmuir_table(P, face, 0.5 * D * T, T, &out);
                veq_y = (double)out;
            } else { veq_x = NAN; veq_y = 1.0; }   /* no isotherm at all: reported as out of range */
            if (veq_y == 1) {
                err = 1;
                early = 1;
            } else {
                double gamma_b = P->gamma_f[face] * (1.0 / veq_x - 1.0) * adjusted;
                des = gamma_b * vartheta;
                if (P->dynamic_vartheta[face] == 1) {
                    dvtheta -= dt * P->alpha_p[face] * des;
                    if ((vartheta + dvtheta) < 0.0) des = vartheta / (dt * P->alpha_p[face]);
                }
            }
        } else {
            dvtheta = 0.0;
        }
        if (!early) {
            r_final = r_init + a_init - a_final;
            correct = r_final / r_init;
            od4 cov = WP(wall_cover, face, ci);
            cov.s0 += dvtheta;
            cov.s1 = r_final; cov.s2 = a_final; cov.s3 = des;
            WP(wall_cover, face, ci) = cov;
        }
        if (correct != 1.0)
            for (int gv = 0; gv < NV_; gv++)
                wall_dist[(size_t)ci * NV_ + gv] = smul2(correct, wall_dist[(size_t)ci * NV_ + gv]);
    }
    if (err) flag[0] = ERR_CLAMPING_ADSORB_ISOTHERM;
}

/* adsorbingWallS_P2 (clAdsorb.c:656-816) and adsorbingWallD_P2 (:821-975) */
/* Modified Bessel function I0 by its integral I0(x) = (1/2pi) int_0^2pi
 * e^{x cos t} dt under the M-point trapezoidal rule, exact up to the aliased
 * 2 I_M(x) (< 1e-17 I0(x) for M = 256 and |x| < 800); a different route from
 * the reference's rational fit (clFunctions.c:88-143) on purpose. */
double or_bessel_i0(double x) {
    const int M = 256;
    double s = 0.0;
    for (int k = 0; k < M; k++) s += exp(x * cos((2.0 * PI_ * k) / M));
    return s / M;
}

/* CercignaniLampis with scatterNormal / scatterTangential,
 * clFunctions.c:146-185: weight of incoming (normal, tangential) velocity uin
 * in the outgoing distribution at uout, wall temperature Tw */
double or_cl_kernel(od2 uin, od2 uout, double alpha_n, double alpha_t, double Tw) {
    double Rn = (2.0 / (alpha_n * Tw)) * fabs(uout.x) *
                or_bessel_i0(((2 * sqrt(1 - alpha_n)) / (alpha_n * Tw)) * uin.x * uout.x) *
                exp(-(uout.x * uout.x + (1 - alpha_n) * uin.x * uin.x) / (alpha_n * Tw));
    double a = uout.y - (1 - alpha_t) * uin.y;
    double Rt = (1.0 / sqrt(PI_ * alpha_t * (2 - alpha_t) * Tw)) * exp(-(a * a) / (alpha_t * (2 - alpha_t) * Tw));
    return fabs(uin.x / uout.x) * Rn * Rt;
}

/* The re-emitted (g, h) of outgoing velocity gv under Cercignani-Lampis:
 * the sum over the velocities into the wall (u_n != 0) of the scattered
 * reflected distribution WALL_DIST, clAdsorb.c:557-574 */
static od2 cl_emission(const or_params *P, od2 n, int rot, int face, int ci, od4 wall, od2 uv, int gv,
                       const od2 *wall_dist) {
    double an = P->alpha_n[face], at = P->alpha_t[face];
    od2 s = v2(0.0, 0.0);
    for (int va = 0; va < NV_; va++) {
        od2 uin = interfaceVelocity(P, va, n);
        if (wall_delta(uin.x, rot) || uin.x == 0.0) continue;
        double k = or_cl_kernel(uin, uv, an, at, 1.0 / wall.s3);
        od2 fin = wall_dist[(size_t)ci * NV_ + va];
        od2 c = v2(k * fin.x, k * fin.y);
        c.y *= (1 - at) * (1 - at);
        c.y += 0.5 * at * (2 - at) / wall.s3 * c.x;
        s = add2(s, c);
    }
    return smul2(P->weight[gv], s);
}

/* adsorbingWallS_P2 / D_P2 / CL_P2 (clAdsorb.c:656-816, :821-975, :488-651);
 * mode 0 specular, 1 diffuse, 2 Cercignani-Lampis */
static void adsorb_P2(const or_params *P, const od2 *normal, int face, const od4 *wall_prop,
                      const od4 *wall_cover, const od2 *wall_dist, od2 *flux_f, int mode) {
    int nw = wall_count(P, face);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int ci = 0; ci < nw; ci++) {
        int gi, gj, rot, fid;
        wall_site(P, face, ci, &gi, &gj, &rot, &fid);
        od2 n = FACEA(normal, gi, gj, fid);
        od4 wall = WP(wall_prop, face, ci);
        od2 wl = toLocal(v2(wall.s1, wall.s2), n);
        wall.s1 = wl.x; wall.s2 = wl.y;
        double reflected = WP(wall_cover, face, ci).s1;
        od2 data = v2(0.0, 0.0);
        double ratio = 1.0;
        if (reflected > 0.0) {
            for (int gv = 0; gv < NV_; gv++) {
                od2 uv = interfaceVelocity(P, gv, n);
                int delta = wall_delta(uv.x, rot);
                od2 em = v2(0.0, 0.0);
                if (delta) {
                    if (mode == 2) {
                        em = cl_emission(P, n, rot, face, ci, wall, uv, gv, wall_dist);
                    } else if (mode == 1) {
                        em = or_fM(P, wall, uv, gv);
                    } else {
                        int64_t mv = fid == WEST ? P->mirror_EW[gv] : P->mirror_NS[gv];
                        em = wall_dist[(size_t)ci * NV_ + mv];
                    }
                }
                od2 fd = F3(flux_f, gi, gj, gv);
                F3(flux_f, gi, gj, gv) = add2(smul2((double)delta, em), smul2((double)(1 - delta), fd));
                data = add2(data, smul2(rot * uv.x, em));
            }
            data.x = reflected / data.x;
            ratio = data.x;
        }
        double des = WP(wall_cover, face, ci).s3;
        double ads = WP(wall_cover, face, ci).s2;
        int account = (((des == 0.0) && (ads > 0.0)) && (reflected == 0.0)) ? 1 : 0;
        wall.s0 = 2 * des * sqrt(PI_ * wall.s3);
        for (int gv = 0; gv < NV_; gv++) {
            od2 uv = interfaceVelocity(P, gv, n);
            int delta = wall_delta(uv.x, rot);
            od2 fd = F3(flux_f, gi, gj, gv);
            fd = smul2(1 - delta * (1 - ratio), fd);
            fd = add2(fd, smul2((double)delta, or_fM(P, wall, uv, gv)));
            fd = sub2(fd, smul2((double)(account * delta), fd));
            F3(flux_f, gi, gj, gv) = fd;
        }
    }
}
void or_adsorbingWallS_P2(const or_params *P, const od2 *normal, int face, const od4 *wall_prop,
                          const od4 *wall_cover, const od2 *wall_dist, od2 *flux_f,
                          const od4 *macro, double dt) {
    (void)macro; (void)dt;
    adsorb_P2(P, normal, face, wall_prop, wall_cover, wall_dist, flux_f, 0);
}
void or_adsorbingWallD_P2(const or_params *P, const od2 *normal, int face, const od4 *wall_prop,
                          const od4 *wall_cover, const od2 *wall_dist, od2 *flux_f,
                          const od4 *macro, double dt) {
    (void)macro; (void)dt;
    adsorb_P2(P, normal, face, wall_prop, wall_cover, wall_dist, flux_f, 1);
}
void or_adsorbingWallCL_P2(const or_params *P, const od2 *normal, int face, const od4 *wall_prop,
                           const od4 *wall_cover, const od2 *wall_dist, od2 *flux_f,
                           const od4 *macro, double dt) {
    (void)macro; (void)dt;
    adsorb_P2(P, normal, face, wall_prop, wall_cover, wall_dist, flux_f, 2);
}

/* updateMacro, clUGKS.c:1158-1187 */
void or_updateMacro(const or_params *P, const od4 *flux_macro_S, const od4 *flux_macro_W,
                    const double *area, od4 *macro, od4 *residual) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int mi = 0; mi < P->ni; mi++)
        for (int mj = 0; mj < P->nj; mj++) {
            int gi = mi + G_, gj = mj + G_;
            od4 prim_old = M2(macro, mi, mj);
            od4 w_old = or_getConserved(P, prim_old);
            double A = C2(area, gi, gj);
            /* reference order: ((FS(j) - FS(j+1)) + FW(i)) - FW(i+1) */
            od4 d = sub4(add4(sub4(C2(flux_macro_S, gi, gj), C2(flux_macro_S, gi, gj + 1)), C2(flux_macro_W, gi, gj)),
                     C2(flux_macro_W, gi + 1, gj));
            od4 w = add4(w_old, div4s(d, A));
            M2(macro, mi, mj) = or_getPrimary(P, w);
            M2(residual, mi, mj) = prim_old;
        }
}

/* UGKS_update, clUGKS.c:1189-1257 */
void or_UGKS_update(const or_params *P, od2 *Fin, const od2 *flux_f_S, const od2 *flux_f_W,
                    const double *area, const od4 *macro, const od2 *gQ, const od4 *residual,
                    double dt, double *flag) {
    int nan_found = 0;
#pragma omp parallel for num_threads(g_threads) schedule(static) reduction(| : nan_found)
    for (int mi = 0; mi < P->ni; mi++)
        for (int mj = 0; mj < P->nj; mj++) {
            int gi = mi + G_, gj = mj + G_;
            od4 prim_old = M2(residual, mi, mj);
            double tau_old = or_relaxTime(P, prim_old);
            od2 Q = M2(gQ, mi, mj);
            od4 prim = M2(macro, mi, mj);
            double tau = or_relaxTime(P, prim);
            double A = C2(area, gi, gj);
            for (int gv = 0; gv < NV_; gv++) {
                od2 q = quad_at(P, gv);
                od2 feq_old = fEQ(P, prim_old, Q, q, gv);
                od2 feq = fEQ(P, prim, Q, q, gv);
                od2 f_old = F3(Fin, gi, gj, gv);
                od2 fl = div2s(sub2(add2(sub2(F3(flux_f_S, gi, gj, gv), F3(flux_f_S, gi, gj + 1, gv)), F3(flux_f_W, gi, gj, gv)),
                                    F3(flux_f_W, gi + 1, gj, gv)), A);
                od2 relax = add2(div2s(feq, tau), div2s(sub2(feq_old, f_old), tau_old));
                od2 fn = div2s(add2(add2(f_old, fl), smul2(0.5 * dt, relax)), 1.0 + 0.5 * dt / tau);
                if (isnan(fn.x) || isnan(fn.y)) nan_found = 1;
                F3(Fin, gi, gj, gv) = fn;
            }
        }
    if (nan_found) flag[0] = ERR_NAN;
}

/* getResidual, clUGKS.c:1259-1277 */
void or_getResidual(const or_params *P, const od4 *macro, od4 *residual) {
    for (int mi = 0; mi < P->ni; mi++)
        for (int mj = 0; mj < P->nj; mj++) {
            od4 a = M2(macro, mi, mj), b = M2(residual, mi, mj), d = sub4(a, b);
            M2(residual, mi, mj) = v4(d.s0 * d.s0, d.s1 * d.s1, d.s2 * d.s2, d.s3 * d.s3);
        }
}

/* edgeMassFlow, clUGKS.c:1282-1337 */
void or_edgeMassFlow(const or_params *P, const od4 *macro, const od2 *normal,
                     const double *side_length, double *mdot, int this_face) {
    int n = wall_count(P, this_face);
    for (int ci = 0; ci < n; ci++) {
        int mi, mj, gi, gj, face;
        double dir;
        switch (this_face) {
        case GNORTH: mi = ci; mj = P->nj - 1; gi = G_ + ci; gj = NJ_ - G_; face = SOUTH; dir = 1; break;
        case GEAST:  mi = P->ni - 1; mj = ci; gi = NI_ - G_; gj = G_ + ci; face = WEST; dir = 1; break;
        case GSOUTH: mi = ci; mj = 0; gi = G_ + ci; gj = G_; face = SOUTH; dir = -1; break;
        default:     mi = 0; mj = ci; gi = G_; gj = G_ + ci; face = WEST; dir = -1; break;
        }
        od2 fn = smul2(dir, FACEA(normal, gi, gj, face));
        double L = FACEA(side_length, gi, gj, face);
        od4 cm = M2(macro, mi, mj);
        mdot[(size_t)ci * 4 + this_face] = dot2(fn, v2(cm.s1, cm.s2)) * cm.s0 * L;
    }
}

/* ---- ghost routines ----------------------------------------------------- */

/* rot90 / fliplr / flipud / getSwapIndex, clGhostRoutines.c:5-138 */
static void rot90(int *i, int *j, int n, int Ni, int Nj) {
    for (int a = 0; a < n; a++) {
        int temp = *i;
        *i = Nj - *j - 1;
        *j = temp;
        temp = Ni; Ni = Nj; Nj = temp;
    }
}
static void fliplr(int *i, int *j, int Ni, int Nj) { (void)j; (void)Nj; *i = Ni - *i; }
static void flipud(int *i, int *j, int Ni, int Nj) { (void)i; (void)Ni; *j = Nj - *j; }
static void getSwapIndex(const or_params *P, int this_face, int *iB, int *jB, int that_face, int NIB, int NJB, int v) {
    int af = v == 0 ? -1 : 0, ar = v == 0 ? 0 : 1;
    const int NI = NI_, NJ = NJ_, G = G_;
    if (this_face == GEAST) {
        if (that_face == GEAST) { fliplr(iB, jB, NI + af, NJ + af); *iB += NIB - 2 * G; }
        else if (that_face == GWEST) { *iB -= NI - 2 * G; }
        else if (that_face == GNORTH) { rot90(iB, jB, 3, NI + ar, NJ + ar); *jB += NJB - 2 * G; }
        else if (that_face == GSOUTH) { rot90(iB, jB, 1, NI + ar, NJ + ar); *jB -= NI - 2 * G; }
    } else if (this_face == GWEST) {
        if (that_face == GWEST) { fliplr(iB, jB, NI + af, NJ + af); *iB -= NI - 2 * G; }
        else if (that_face == GEAST) { *iB += NIB - 2 * G; }
        else if (that_face == GNORTH) { rot90(iB, jB, 1, NI + ar, NJ + ar); *jB += NJB - 2 * G; }
        else if (that_face == GSOUTH) { rot90(iB, jB, 3, NI + ar, NJ + ar); *jB -= NI - 2 * G; }
    } else if (this_face == GNORTH) {
        if (that_face == GNORTH) { flipud(iB, jB, NI + af, NJ + af); *jB += NJB - 2 * G; }
        else if (that_face == GSOUTH) { *jB -= NJ - 2 * G; }
        else if (that_face == GEAST) { rot90(iB, jB, 1, NI + ar, NJ + ar); *iB += NIB - 2 * G; }
        else if (that_face == GWEST) { rot90(iB, jB, 3, NI + ar, NJ + ar); *iB -= NJ - 2 * G; }
    } else if (this_face == GSOUTH) {
        if (that_face == GSOUTH) { flipud(iB, jB, NI + af, NJ + af); *jB -= NJ - 2 * G; }
        else if (that_face == GNORTH) { *jB += NJB - 2 * G; }
        else if (that_face == GEAST) { rot90(iB, jB, 3, NI + ar, NJ + ar); *iB += NIB - 2 * G; }
        else if (that_face == GWEST) { rot90(iB, jB, 1, NI + ar, NJ + ar); *iB -= NJ - 2 * G; }
    }
}

/* ghost-strip addressing shared by the edge* kernels: launch (ni, ghost) on
   N/S and (ghost, nj) on E/W (ugks_block.py:454-457) */
static void ghost_strip(const or_params *P, int face, int *ni_, int *nj_) {
    if (face == GNORTH || face == GSOUTH) { *ni_ = P->ni; *nj_ = G_; }
    else { *ni_ = G_; *nj_ = P->nj; }
}
static void ghost_cell(const or_params *P, int face, int li, int lj, int *gi, int *gj) {
    switch (face) {
    case GNORTH: *gi = li + G_; *gj = lj + NJ_ - G_; break;
    case GEAST:  *gi = li + NI_ - G_; *gj = lj + G_; break;
    case GSOUTH: *gi = li + G_; *gj = lj; break;
    default:     *gi = li; *gj = lj + G_; break;
    }
}

/* edgeExchange, clGhostRoutines.c:407-538 ; fB is indexed with NJB */
void or_edgeExchange(const or_params *PA, od2 *fA, const od2 *xyA, int this_face,
                     const od2 *fB, const od2 *xyB, int NIB, int NJB, int that_face,
                     int ori, int flip) {
    const or_params *P = PA;
    const int G = G_, NI = NI_, NJ = NJ_;
    int A1i, A1j, A2i, A2j, B1i, B1j, B2i, B2j;
    switch (this_face) {
    case GNORTH: A1i = G; A1j = NJ - G; A2i = NI - G; A2j = NJ - G; break;
    case GEAST:  A1i = NI - G; A1j = G; A2i = NI - G; A2j = NJ - G; break;
    case GSOUTH: A1i = G; A1j = G; A2i = NI - G; A2j = G; break;
    default:     A1i = G; A1j = G; A2i = G; A2j = NJ - G; break;
    }
    switch (that_face) {
    case GNORTH: B1i = G; B1j = NJB - G; B2i = NIB - G; B2j = NJB - G; break;
    case GEAST:  B1i = NIB - G; B1j = G; B2i = NIB - G; B2j = NJB - G; break;
    case GSOUTH: B1i = G; B1j = G; B2i = NIB - G; B2j = G; break;
    default:     B1i = G; B1j = G; B2i = G; B2j = NJB - G; break;
    }
    od2 A1 = xyA[(size_t)A1i * (NJ + 1) + A1j], A2 = xyA[(size_t)A2i * (NJ + 1) + A2j];
    od2 B1 = xyB[(size_t)B1i * (NJB + 1) + B1j], B2 = xyB[(size_t)B2i * (NJB + 1) + B2j];
    int misalign = !(fabs(A1.x - B1.x) <= 1e-6) || !(fabs(A1.y - B1.y) <= 1e-6) ||
                   !(fabs(A2.x - B2.x) <= 1e-6) || !(fabs(A2.y - B2.y) <= 1e-6);
    int si, sj;
    ghost_strip(P, this_face, &si, &sj);
    for (int li = 0; li < si; li++)
        for (int lj = 0; lj < sj; lj++) {
            int gi, gj;
            ghost_cell(P, this_face, li, lj, &gi, &gj);
            int iB = gi, jB = gj;
            if (misalign && ori) {
                if (this_face == GEAST || this_face == GWEST) flipud(&iB, &jB, NI - 1, NJ - 1);
                else fliplr(&iB, &jB, NI - 1, NJ - 1);
            }
            getSwapIndex(P, this_face, &iB, &jB, that_face, NIB, NJB, 0);
            for (int gv1 = 0; gv1 < NV_; gv1++) {
                int64_t gv2 = gv1;
                if (flip == FLIP_NS) gv2 = P->mirror_NS[gv1];
                else if (flip == FLIP_D) gv2 = P->mirror_D[gv1];
                else if (flip == FLIP_EW) gv2 = P->mirror_EW[gv1];
                fA[(size_t)NV_ * NJ * gi + (size_t)NV_ * gj + gv1] = fB[(size_t)NV_ * NJB * iB + (size_t)NV_ * jB + gv2];
            }
        }
}

/* edgeExtrapolate, clGhostRoutines.c:544-597 */
void or_edgeExtrapolate(const or_params *P, od2 *Fin, int this_face) {
    int si, sj;
    ghost_strip(P, this_face, &si, &sj);
    for (int li = 0; li < si; li++)
        for (int lj = 0; lj < sj; lj++) {
            int gi, gj, ei, ej;
            ghost_cell(P, this_face, li, lj, &gi, &gj);
            switch (this_face) {
            case GNORTH: ei = gi; ej = NJ_ - G_ - 1; break;
            case GEAST:  ei = NI_ - G_ - 1; ej = gj; break;
            case GSOUTH: ei = gi; ej = G_; break;
            default:     ei = G_; ej = gj; break;
            }
            for (int gv = 0; gv < NV_; gv++) F3(Fin, gi, gj, gv) = F3(Fin, ei, ej, gv);
        }
}

/* edgeConstant, clGhostRoutines.c:603-658 */
void or_edgeConstant(const or_params *P, od2 *Fin, int this_face, const od4 *wall_prop) {
    int si, sj;
    ghost_strip(P, this_face, &si, &sj);
    for (int li = 0; li < si; li++)
        for (int lj = 0; lj < sj; lj++) {
            int gi, gj;
            ghost_cell(P, this_face, li, lj, &gi, &gj);
            int ci = (this_face == GNORTH || this_face == GSOUTH) ? li : lj;
            od4 prim = WP(wall_prop, this_face, ci);
            for (int gv = 0; gv < NV_; gv++) F3(Fin, gi, gj, gv) = fEQ(P, prim, v2(0.0, 0.0), quad_at(P, gv), gv);
        }
}

/* edgeInflow (clGhostRoutines.c:664-749) and edgeOutflow (:755-839) */
static void edge_in_out(const or_params *P, od2 *Fin, int this_face, const od2 *normal,
                        const od4 *macro, const od2 *gQ, const od4 *wall_prop, int outflow) {
    int si, sj;
    ghost_strip(P, this_face, &si, &sj);
    for (int li = 0; li < si; li++)
        for (int lj = 0; lj < sj; lj++) {
            int gi, gj, ei, ej, noi, noj, face;
            ghost_cell(P, this_face, li, lj, &gi, &gj);
            int ci = (this_face == GNORTH || this_face == GSOUTH) ? li : lj;
            switch (this_face) {
            case GNORTH: ei = gi; ej = NJ_ - G_ - 1; noi = gi; noj = gj; face = SOUTH; break;
            case GEAST:  ei = NI_ - G_ - 1; ej = gj; noi = gi; noj = gj; face = WEST; break;
            case GSOUTH: ei = gi; ej = G_; noi = ei; noj = ej; face = SOUTH; break;
            default:     ei = G_; ej = gj; noi = ei; noj = ej; face = WEST; break;
            }
            od4 prim = WP(wall_prop, this_face, ci);
            od2 Q = M2(gQ, ei - G_, ej - G_);
            od4 gm = M2(macro, ei - G_, ej - G_);
            od2 wn = FACEA(normal, noi, noj, face);
            od2 l = toLocal(v2(gm.s1, gm.s2), wn);
            gm.s1 = l.x; gm.s2 = l.y;
            if (!outflow) { gm.s0 = prim.s0; gm.s3 = prim.s3; }
            else { gm.s3 = gm.s0 / (prim.s0 / prim.s3); }
            for (int gv = 0; gv < NV_; gv++) F3(Fin, gi, gj, gv) = fEQ(P, gm, Q, quad_at(P, gv), gv);
        }
}
void or_edgeInflow(const or_params *P, od2 *Fin, int this_face, const od2 *normal,
                   const od4 *macro, const od2 *gQ, const od4 *wall_prop) {
    edge_in_out(P, Fin, this_face, normal, macro, gQ, wall_prop, 0);
}
void or_edgeOutflow(const or_params *P, od2 *Fin, int this_face, const od2 *normal,
                    const od4 *macro, const od2 *gQ, const od4 *wall_prop) {
    edge_in_out(P, Fin, this_face, normal, macro, gQ, wall_prop, 1);
}

/* edgeMirror (clGhostRoutines.c:845-905) and edgeBounceBack (:911-961) */
static void edge_mirror(const or_params *P, od2 *Fin, int this_face, int bb) {
    int si, sj;
    ghost_strip(P, this_face, &si, &sj);
    for (int li = 0; li < si; li++)
        for (int lj = 0; lj < sj; lj++) {
            int gi, gj, mi, mj;
            ghost_cell(P, this_face, li, lj, &gi, &gj);
            switch (this_face) {
            case GNORTH: mi = gi; mj = 2 * (NJ_ - G_) - gj - 1; break;
            case GEAST:  mi = 2 * (NI_ - G_) - gi - 1; mj = gj; break;
            case GSOUTH: mi = gi; mj = 2 * G_ - gj - 1; break;
            default:     mi = 2 * G_ - gi - 1; mj = gj; break;
            }
            int ns = (this_face == GNORTH || this_face == GSOUTH);
            for (int gv = 0; gv < NV_; gv++) {
                int64_t mv = bb ? P->mirror_D[gv] : (ns ? P->mirror_NS[gv] : P->mirror_EW[gv]);
                F3(Fin, gi, gj, gv) = F3(Fin, mi, mj, mv);
            }
        }
}
void or_edgeMirror(const or_params *P, od2 *Fin, int this_face) { edge_mirror(P, Fin, this_face, 0); }
void or_edgeBounceBack(const or_params *P, od2 *Fin, int this_face) { edge_mirror(P, Fin, this_face, 1); }

/* edgeConstGrad, clGhostRoutines.c:967-1048 */
void or_edgeConstGrad(const or_params *P, od2 *Fin, int this_face) {
    int si, sj;
    ghost_strip(P, this_face, &si, &sj);
    for (int li = 0; li < si; li++)
        for (int lj = 0; lj < sj; lj++) {
            int gi, gj, ei0, ej0, ei1, ej1;
            double m;
            ghost_cell(P, this_face, li, lj, &gi, &gj);
            switch (this_face) {
            case GNORTH: ei1 = gi; ej1 = NJ_ - G_ - 1; ei0 = ei1; ej0 = ej1 - 1; m = (double)(gj - ej1); break;
            case GEAST:  ei1 = NI_ - G_ - 1; ej1 = gj; ei0 = ei1 - 1; ej0 = ej1; m = (double)(gi - ei1); break;
            case GSOUTH: ei1 = gi; ej1 = G_; ei0 = ei1; ej0 = ej1 + 1; m = (double)(ej1 - gj); break;
            default:     ei1 = G_; ej1 = gj; ei0 = ei1 + 1; ej0 = ej1; m = (double)(ei1 - gi); break;
            }
            for (int gv = 0; gv < NV_; gv++) {
                od2 f0 = F3(Fin, ei0, ej0, gv), f1 = F3(Fin, ei1, ej1, gv);
                od2 df = sub2(f1, f0);
                F3(Fin, gi, gj, gv) = add2(f1, smul2(m, df));
            }
        }
}

/* ==== host sequencing ==================================================== */

/* get_flag, ugks_block.py:840-876: returns 1 on the error codes */
static int get_flag(const or_block *b) {
    double e = b->s.flag[0];
    return e == 1.0 || e == 1.1 || e == 1.2 || e == 1.3 || e == 4.0;
}
static void reset_flag(or_block *b) { b->s.flag[0] = 0.0; b->s.flag[1] = 0.0; }

/* UGKSBlock.initFunctions device part, ugks_block.py:411-429 */
void or_block_init(or_block *b) {
    const or_params *P = &b->p;
    or_state *s = &b->s;
    or_initFunctions(P, s->f, s->macro, s->Q);
    or_calcMacro(P, s->f, s->macro);
    or_calcQ(P, s->f, s->macro, s->Q);
    or_paraBC(P, s->wall_prop);
}

/* UGKSBlock.wall_flux, ugks_block.py:930-983 */
static int wall_flux(or_block *b, int bci, int wall, od2 *flux, od4 *flux_macro, double dt) {
    const or_params *P = &b->p;
    or_state *s = &b->s;
    if (b->bc[bci].type == OR_DIFFUSE) {
        or_accommodatingWallDist(P, s->normal, wall, s->wall_prop, flux, dt);
    } else {
        reset_flag(b);
        or_adsorbingWall_P1(P, s->normal, wall, s->wall_prop, s->wall_cover, s->wall_dist, flux, s->macro, dt, s->flag);
        if (get_flag(b)) return 1;
        if (b->bc[bci].reflect_type == 1)
            or_adsorbingWallD_P2(P, s->normal, wall, s->wall_prop, s->wall_cover, s->wall_dist, flux, s->macro, dt);
        else if (b->bc[bci].reflect_type == 0)
            or_adsorbingWallS_P2(P, s->normal, wall, s->wall_prop, s->wall_cover, s->wall_dist, flux, s->macro, dt);
        else
            or_adsorbingWallCL_P2(P, s->normal, wall, s->wall_prop, s->wall_cover, s->wall_dist, flux, s->macro, dt);
    }
    or_wallFlux(P, s->normal, s->side_length, wall, flux, flux_macro, dt);
    or_wallMassEnergyFluxes(P, s->normal, s->side_length, s->area, wall, flux, s->wall_fluxes, dt);
    return 0;
}

static int is_wall(int t) { return t == OR_DIFFUSE || t == OR_ADSORBING; }

/* one face direction of UGKSBlock.UGKS_flux, ugks_block.py:1062-1156 (S) / :1167-1257 (W) */
static int flux_pass(or_block *b, int face, double dt) {
    const or_params *P = &b->p;
    or_state *s = &b->s;
    od2 *ff = face == SOUTH ? s->flux_f_S : s->flux_f_W;
    od4 *fm = face == SOUTH ? s->flux_macro_S : s->flux_macro_W;
    int top = face == SOUTH ? GNORTH : GEAST, bot = face == SOUTH ? GSOUTH : GWEST;
    or_iFace(P, s->f, ff, s->sigma, s->centre, s->mid_side, s->normal, face);
    int ot = 0, ob = 0;
    if (is_wall(b->bc[top].type)) { ot = 1; if (wall_flux(b, top, top, ff, fm, dt) == 2) return 2; }
    if (is_wall(b->bc[bot].type)) { ob = 1; if (wall_flux(b, bot, bot, ff, fm, dt) == 2) return 2; }
    or_getFaceCons(P, ff, s->normal, face, s->prim, ob, ot);
    or_getAL(P, s->f, s->normal, face, s->aL, ob, ot);
    or_getAR(P, s->f, s->normal, face, s->aR, ob, ot);
    or_getPLR(P, s->centre, s->mid_side, face, s->prim, s->aL, s->aR, ob, ot);
    or_initMacroFlux(P, fm, face, dt, s->prim, s->aL, s->aR, s->aT, s->Mxi, ob, ot);
    or_calcFaceQ(P, ff, s->prim, s->normal, s->faceQ, face, ob, ot);
    or_macroFlux(P, ff, s->sigma, fm, s->normal, s->side_length, face, dt, s->prim, s->faceQ, ob, ot);
    or_distFlux(P, ff, s->sigma, s->normal, s->side_length, face, dt, s->prim, s->aL, s->aR, s->aT, s->Mxi, s->faceQ, ob, ot);
    return 0;
}

/* UGKSBlock.UGKS_flux, ugks_block.py:1049-1260 (always returns False there) */
int or_block_flux(or_block *b, double dt) {
    int r = flux_pass(b, SOUTH, dt);
    if (r == 2) return r;
    return flux_pass(b, WEST, dt);
}

/* UGKSBlock.UGKS_update, ugks_block.py:1262-1329 (residual branch not restated) */
int or_block_update(or_block *b, double dt) {
    const or_params *P = &b->p;
    or_state *s = &b->s;
    or_calcQ(P, s->f, s->macro, s->Q);
    or_updateMacro(P, s->flux_macro_S, s->flux_macro_W, s->area, s->macro, s->residual);
    reset_flag(b);
    or_UGKS_update(P, s->f, s->flux_f_S, s->flux_f_W, s->area, s->macro, s->Q, s->residual, dt, s->flag);
    return get_flag(b);
}

/* UGKSBlock.updateBC, ugks_block.py:621-665 */
void or_block_updateBC(or_sim *sim, int bi) {
    or_block *b = &sim->blocks[bi];
    const or_params *P = &b->p;
    or_state *s = &b->s;
    for (int face = 0; face < 4; face++) {
        const or_bc *bc = &b->bc[face];
        switch (bc->type) {
        case OR_ADJACENT:
        case OR_PERIODIC: {
            or_block *o = &sim->blocks[bc->other_block];
            or_edgeExchange(P, s->f, s->xy, face, o->s.f, o->s.xy, o->p.NI, o->p.NJ, bc->other_face,
                            bc->orientation, bc->flip);
            break;
        }
        case OR_EXTRAPOLATE_OUT: or_edgeExtrapolate(P, s->f, face); break;
        case OR_CONSTANT: or_edgeConstant(P, s->f, face, s->wall_prop); break;
        case OR_REFLECT: or_edgeMirror(P, s->f, face); break;
        case OR_BOUNCE_BACK: or_edgeBounceBack(P, s->f, face); break;
        case OR_DIFFUSE:
        case OR_ADSORBING: or_edgeConstGrad(P, s->f, face); break;
        case OR_INFLOW: or_edgeInflow(P, s->f, face, s->normal, s->macro, s->Q, s->wall_prop); break;
        case OR_OUTFLOW: or_edgeOutflow(P, s->f, face, s->normal, s->macro, s->Q, s->wall_prop); break;
        }
    }
}

/* UGKSBlock.getDT, ugks_block.py:1415-1433 */
double or_block_getDT(or_sim *sim, int bi) {
    or_block *b = &sim->blocks[bi];
    const or_params *P = &b->p;
    or_clFindDT(P, b->s.xy, b->s.area, b->s.macro, b->s.time_step, sim->Kn_eff);
    double mx = -INFINITY;
    for (size_t k = 0; k < (size_t)P->ni * P->nj; k++)
        if (b->s.time_step[k] > mx) mx = b->s.time_step[k];
    double max_dt = sim->CFL / mx;
    if (sim->suggest_dt > 0.0 && sim->suggest_dt < max_dt) max_dt = sim->suggest_dt;
    sim->max_dt[bi] = max_dt;
    return max_dt;
}

void or_sim_updateAllBC(or_sim *sim) {
    for (int bi = 0; bi < sim->nblocks; bi++) or_block_updateBC(sim, bi);
}
void or_sim_calcTimeStep(or_sim *sim) {
    for (int bi = 0; bi < sim->nblocks; bi++) or_block_getDT(sim, bi);
}
/* UGKSim.updateTimeStep, ugks_sim.py:285-311 */
void or_sim_updateTimeStep(or_sim *sim, int limit) {
    if (sim->force_time) return;
    double dt_old = sim->dt, dt_new = sim->max_dt[0];
    for (int bi = 1; bi < sim->nblocks; bi++) if (sim->max_dt[bi] < dt_new) dt_new = sim->max_dt[bi];
    if (dt_old != dt_new) {
        if ((dt_new > (1.0 + sim->delta_dt) * dt_old) && limit) dt_new = (1.0 + sim->delta_dt) * dt_old;
        sim->dt = dt_new;
    }
}

/* UGKSim.__init__ from initFunctions onwards, ugks_sim.py:81-92 */
int or_sim_init(or_sim *sim) {
    for (int bi = 0; bi < sim->nblocks; bi++) {
        or_block *b = &sim->blocks[bi];
        or_cellGeom(&b->p, b->s.xy, b->s.area, b->s.centre, b->s.mid_side, b->s.normal, b->s.side_length);
    }
    for (int bi = 0; bi < sim->nblocks; bi++) or_block_init(&sim->blocks[bi]);
    or_sim_updateAllBC(sim);
    or_sim_calcTimeStep(sim);
    or_sim_updateTimeStep(sim, 0);
    return 0;
}

/* UGKSim.one_step, ugks_sim.py:407-501 */
int or_sim_one_step(or_sim *sim, int get_dt) {
    double dt_old = 0.0;
    int count = 0;
    while ((sim->dt != dt_old) && (count < 2)) {
        dt_old = sim->dt;
        for (int bi = 0; bi < sim->nblocks; bi++)
            if (or_block_flux(&sim->blocks[bi], sim->dt) == 2) return -2;
        or_sim_updateTimeStep(sim, 1);
        count += 1;
    }
    if (count > 1) {
        if (sim->step == 1) return -3;   /* "Overshot on first iteration" */
        get_dt = 1;
    }
    for (int bi = 0; bi < sim->nblocks; bi++)
        if (or_block_update(&sim->blocks[bi], sim->dt)) return -1;
    if (get_dt) {
        or_sim_calcTimeStep(sim);
        or_sim_updateTimeStep(sim, 1);
    }
    sim->time += sim->dt;
    sim->step += 1;
    or_sim_updateAllBC(sim);
    return 0;
}

/* the per-step part of UGKSim.run, ugks_sim.py:542-578 */
int or_sim_run(or_sim *sim, int nsteps) {
    for (int k = 0; k < nsteps; k++) {
        int get_dt;
        if (sim->dt_update_count < 0) get_dt = (sim->step == -sim->dt_update_count);
        else get_dt = ((sim->step + 1) % sim->dt_update_count) == 0;
        int r = or_sim_one_step(sim, get_dt);
        if (r) return r;
    }
    return 0;
}
