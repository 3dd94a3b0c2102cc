// ugks_device.hpp -- shared device types and the per-velocity physics of the
// UGKS step (gfx950).  The formulas follow clFunctions.c / clUGKS.c of the
// reference (cited per function); the code is written for wave64 kernels
// and keeps every operand in FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ugks {

constexpr double kPi = 3.141592653589793;   // '%0.15e' % pi (ugks_CL.py:84)
enum { SOUTH = 0, WEST = 1 };
enum { GNORTH = 0, GEAST = 1, GSOUTH = 2, GWEST = 3 };
enum { NO_FLIP = 0, FLIP_NS = 1, FLIP_D = 2, FLIP_EW = 3 };
constexpr double ERR_NAN = 1.0;
constexpr double ERR_CLAMPING_ADSORB_ISOTHERM = 4.0;

struct __align__(16) d2 { double x, y; };
struct __align__(16) d4 { double s0, s1, s2, s3; };

__host__ __device__ inline d2 mk2(double x, double y) { return d2{x, y}; }
__host__ __device__ inline d4 mk4(double a, double b, double c, double d) { return d4{a, b, c, d}; }
__host__ __device__ inline d2 operator+(d2 a, d2 b) { return d2{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline d2 operator-(d2 a, d2 b) { return d2{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline d2 operator*(double s, d2 a) { return d2{s * a.x, s * a.y}; }
__host__ __device__ inline d2 operator*(d2 a, double s) { return d2{a.x * s, a.y * s}; }
__host__ __device__ inline d2 operator/(d2 a, double s) { return d2{a.x / s, a.y / s}; }
__host__ __device__ inline d4 operator+(d4 a, d4 b) { return d4{a.s0 + b.s0, a.s1 + b.s1, a.s2 + b.s2, a.s3 + b.s3}; }
__host__ __device__ inline d4 operator-(d4 a, d4 b) { return d4{a.s0 - b.s0, a.s1 - b.s1, a.s2 - b.s2, a.s3 - b.s3}; }
__host__ __device__ inline d4 operator*(double s, d4 a) { return d4{s * a.s0, s * a.s1, s * a.s2, s * a.s3}; }
__host__ __device__ inline d4 operator/(d4 a, double s) { return d4{a.s0 / s, a.s1 / s, a.s2 / s, a.s3 / s}; }
__host__ __device__ inline double dot(d2 a, d2 b) { return a.x * b.x + a.y * b.y; }
__host__ __device__ inline double length(d2 a) { return sqrt(a.x * a.x + a.y * a.y); }

// Per-block constants passed by value to every kernel: the runtime form of
// the genHeader #defines (ugks_CL.py:58-395).
struct Consts {
    int ni, nj, NI, NJ, NV, G, method, relax;
    double Pr, Kn, gam, K, omega, umax, vmax;
    double B;                        // b = 2 / (gamma - 1), baked (ugks_CL.py:86-93)
    double wall_D[4], wall_U[4], wall_V[4], wall_T[4], wall_P[4];
    double beta_n[4], beta_t[4], gamma_f[4], alpha_p[4], adsorb[4];
    double alpha_n[4], alpha_t[4];   // Cercignani-Lampis accommodation (ALPHA_N/ALPHA_T, ugks_CL.py:134-148)
    int dyn[4], cads[4];
    // tabulated isotherms (ISO/TRI/NBR/N_TRI, ugks_CL.py:160-248): [.][3] points (P, T, theta),
    // vertex and neighbour triples; n_tri 0 where a wall has none
    const double *iso[4];
    const int *iso_tri[4], *iso_nbr[4];
    int n_tri[4];
    const d2 *quad;
    const double *weight;
    const int64_t *mNS, *mEW, *mD;
    // tensor structure of the velocity set: quad[v] = (qu[v / nvt], qv[v % nvt]),
    // so a Maxwellian factorises into per-axis exponentials (nut = 0 if not a tensor grid)
    int nut, nvt;
    const double *qu, *qv;
    // wall values given as expressions (UDF_*): slot 4*w + q, q = 0 D, 1 U, 2 V, 3 T;
    // udf_len 0 keeps the constant wall value; any_udf / time_udf summarise the block
    const int *udf_code;
    const double *udf_const;
    int udf_start[16], udf_len[16];
    int any_udf, time_udf;
};

// ---- wall-value expressions (paraBC with UDF text, clKernel.c:141-218) ----
// The stack program of one slot, in the wall cell's parametric position s,
// midpoint (x, y) and time t (opcodes: pyugks_amd/udf.py; every program is
// checked on the host to stay within 16 stack entries and its tables).
__device__ inline double udf_eval(const Consts &C, int slot, double s, double x, double y, double t) {
    const int *c = C.udf_code + C.udf_start[slot];
    const int n = C.udf_len[slot];
    double st[16];
    int sp = 0;
    for (int i = 0; i < n; i++) {
        const int op = c[i];
        if (op == 1) { st[sp++] = C.udf_const[c[++i]]; continue; }
        if (op >= 2 && op <= 5) { st[sp++] = op == 2 ? s : op == 3 ? x : op == 4 ? y : t; continue; }
        if ((op >= 10 && op <= 15 && op != 14) || op >= 40) {
            const double b = st[--sp], a = st[sp - 1];
            double r;
            switch (op) {
            case 10: r = a + b; break;
            case 11: r = a - b; break;
            case 12: r = a * b; break;
            case 13: r = a / b; break;
            case 15: r = pow(a, b); break;
            case 40: r = fmin(a, b); break;
            case 41: r = fmax(a, b); break;
            case 42: r = atan2(a, b); break;
            default: r = fmod(a, b); break;
            }
            st[sp - 1] = r;
            continue;
        }
        const double a = st[sp - 1];
        double r;
        switch (op) {
        case 14: r = -a; break;
        case 20: r = sin(a); break;
        case 21: r = cos(a); break;
        case 22: r = tan(a); break;
        case 23: r = exp(a); break;
        case 24: r = log(a); break;
        case 25: r = sqrt(a); break;
        case 26: r = fabs(a); break;
        case 27: r = tanh(a); break;
        case 28: r = atan(a); break;
        case 29: r = asin(a); break;
        case 30: r = acos(a); break;
        case 31: r = sinh(a); break;
        case 32: r = cosh(a); break;
        case 33: r = log10(a); break;
        case 34: r = floor(a); break;
        default: r = ceil(a); break;
        }
        st[sp - 1] = r;
    }
    return st[0];
}
// paraBC for wall w at one wall cell: the WALL_* values, an expression where
// the wall has one; s3 = P / D where a pressure is set, else 1 / T
__device__ inline d4 para_wall(const Consts &C, int w, d4 pa, double t) {
    auto val = [&](int q, double k) { return C.udf_len[4 * w + q] ? udf_eval(C, 4 * w + q, pa.s0, pa.s1, pa.s2, t) : k; };
    d4 p;
    p.s0 = val(0, C.wall_D[w]);
    p.s1 = val(1, C.wall_U[w]);
    p.s2 = val(2, C.wall_V[w]);
    if (C.wall_P[w] != -1) p.s3 = C.wall_P[w] / p.s0;
    else p.s3 = C.udf_len[4 * w + 3] ? 1.0 / udf_eval(C, 4 * w + 3, pa.s0, pa.s1, pa.s2, t) : C.wall_T[w];
    return p;
}

// ---- indexing of the reference buffers (clKernel.c:4-26) ------------------
__device__ __forceinline__ size_t fidx(const Consts &C, int i, int j, int v) {
    return ((size_t)i * C.NJ + (size_t)j) * (size_t)C.NV + (size_t)v;
}
__device__ __forceinline__ size_t cidx(const Consts &C, int i, int j) { return (size_t)i * C.NJ + j; }
__device__ __forceinline__ size_t midx(const Consts &C, int i, int j) { return (size_t)i * C.nj + j; }
__device__ __forceinline__ size_t sidx(const Consts &C, int i, int j, int face) {
    return ((size_t)i * C.NJ + (size_t)j) * 2 + face;
}

// OpenCL sign(): +-1, the signed zero itself, 0 for NaN
__device__ __forceinline__ double ocl_sign(double x) {
    return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : (x == 0.0 ? x : 0.0));
}
// int delta = (sign(u)+1)/2 : 1 for u > 0, else 0 (clUGKS.c:148, :743)
__device__ __forceinline__ int upwind(double u) { return u > 0.0 ? 1 : 0; }
// int delta = (sign(u)*rot+1)/2 : 1 when the velocity leaves the wall (clUGKS.c:852)
__device__ __forceinline__ int leaves_wall(double u, int rot) { return (rot > 0) ? (u > 0.0) : (u < 0.0); }

// interfaceVelocity (clFunctions.c:191-209): (normal, tangential) components
__device__ __forceinline__ d2 face_velocity(const Consts &C, int v, d2 n) {
    d2 e = C.quad[v];
    return d2{e.x * n.x + e.y * n.y, e.x * (-n.y) + e.y * n.x};
}
// toGlobal (clFunctions.c:215-230) and toLocal (:236-250)
__device__ __forceinline__ d2 to_global(d2 in, d2 n) {
    return d2{n.x * in.x + (-n.y) * in.y, n.y * in.x + n.x * in.y};
}
__device__ __forceinline__ d2 to_local(d2 in, d2 n) {
    return d2{in.x * n.x + in.y * n.y, in.x * (-n.y) + in.y * n.x};
}

// relaxTime (clFunctions.c:8-32)
__device__ __forceinline__ double relax_time(const Consts &C, d4 p) {
    double base = (C.Kn / p.s0);
    double pw = pow(p.s3, 1.0 - C.omega);
    switch (C.relax) {
    case 0: return base * 0.79788456080286541 * pw;
    case 1: return (5. / 8.) * base * sqrt(kPi) * pw;
    case 3: return base * 2.2567583341910251 * pw;
    default: return base * pw;
    }
}

// fM (clFunctions.c:49-58): weighted Maxwellian (g, h)
__device__ __forceinline__ d2 maxwell(const Consts &C, d4 p, d2 uv, int v) {
    double cx = uv.x - p.s1, cy = uv.y - p.s2;
    double g = p.s0 * (p.s3 / kPi) * exp(-p.s3 * (cx * cx + cy * cy));
    double h = (g * C.K) / (2.0 * p.s3);
    double w = C.weight[v];
    return d2{w * g, w * h};
}
// fS (clFunctions.c:64-77): Shakhov correction of a Maxwellian M
__device__ __forceinline__ d2 shakhov(const Consts &C, d4 p, d2 Q, d2 uv, d2 M) {
    double cx = uv.x - p.s1, cy = uv.y - p.s2;
    double s = 0.8 * (1 - C.Pr) * (p.s3 * p.s3) / p.s0 * (cx * Q.x + cy * Q.y);
    double part = 2 * p.s3 * (cx * cx + cy * cy);
    return d2{s * (part - 4) * M.x, s * (part - 2) * M.y};
}
// Modified Bessel function I0 of the normal scattering kernel.  (The
// reference evaluates a rational Chebyshev fit, clFunctions.c:88-143; any
// I0 accurate to a few ulp gives the same kernel to 1e-14.)  |x| <= 25: the
// power series sum_k (x^2/4)^k / (k!)^2, all terms positive; beyond: the
// asymptotic series e^x / sqrt(2 pi x) sum_k [(2k-1)!!]^2 / (k! (8x)^k)
// (Abramowitz & Stegun 9.6.10, 9.7.1), whose neglected part is O(e^-2x).
// e^x is formed as a square so that the result overflows only where I0 does.
__host__ __device__ inline double bessel_i0(double x) {
    x = fabs(x);
    double term = 1.0, sum = 1.0;
    if (x <= 25.0) {
        const double t = 0.25 * x * x;
        for (int k = 1; k < 64; k++) {
            term *= t / (double)(k * k);
            sum += term;
            if (term <= 1e-17 * sum) break;
        }
        return sum;
    }
    const double r = 0.125 / x;
    for (int k = 1; k < 40; k++) {
        const double a = (double)(2 * k - 1);
        term *= (a * a) * r / (double)k;
        sum += term;
        if (term <= 1e-17 * sum) break;
    }
    const double e = exp(0.5 * x);
    return ((sum / sqrt(2.0 * kPi * x)) * e) * e;
}
// Cercignani-Lampis scattering (clFunctions.c:146-185) for the wall at
// temperature Tw: the normal kernel (accommodation an) of incoming normal
// speed c_in into c_out, the tangential one (at), and their product with the
// flux ratio |u_in / u_out|, the weight of one incoming (g, h) in the outgoing
// distribution.  Its h part carries a further (1 - at)^2.
__host__ __device__ inline double cl_scatter_normal(double c_in, double c_out, double an, double Tw) {
    const double s = an * Tw;
    return (2.0 / s) * fabs(c_out) * bessel_i0(((2.0 * sqrt(1.0 - an)) / s) * c_in * c_out) *
           exp(-(c_out * c_out + (1.0 - an) * c_in * c_in) / s);
}
__host__ __device__ inline double cl_scatter_tangential(double c_in, double c_out, double at, double Tw) {
    const double s = at * (2.0 - at) * Tw, a = c_out - (1.0 - at) * c_in;
    return (1.0 / sqrt(kPi * s)) * exp(-(a * a) / s);
}
__host__ __device__ inline double cl_kernel(d2 uv_in, d2 uv_out, double an, double at, double Tw) {
    return fabs(uv_in.x / uv_out.x) * cl_scatter_normal(uv_in.x, uv_out.x, an, Tw) *
           cl_scatter_tangential(uv_in.y, uv_out.y, at, Tw);
}

// adsorbingWallCL_P2's re-emitted (g, h) at outgoing face-frame velocity uv,
// index v (clAdsorb.c:557-574): the Cercignani-Lampis scattering of the
// reflected distribution fin(va) over every velocity into the wall with
// u_n != 0, summed in index order, times the weight of v.  The h part scatters
// with a further (1 - at)^2 and gains at (2 - at) Tw / 2 times the g part.
template <typename Fin>
__device__ inline d2 cl_emission(const Consts &C, int face, d2 n, int rot, d4 wall, d2 uv, int v, Fin fin) {
    const double an = C.alpha_n[face], at = C.alpha_t[face], Tw = 1.0 / wall.s3;
    const double hk = (1.0 - at) * (1.0 - at), hc = 0.5 * at * (2.0 - at) / wall.s3;
    double sg = 0.0, sh = 0.0;
    for (int va = 0; va < C.NV; va++) {
        const d2 uin = face_velocity(C, va, n);
        if (leaves_wall(uin.x, rot) || uin.x == 0.0) continue;
        const double k = cl_kernel(uin, uv, an, at, Tw);
        const d2 f = fin(va);
        const double cg = k * f.x;
        sg += cg;
        sh += k * f.y * hk + hc * cg;
    }
    const double w = C.weight[v];
    return d2{w * sg, w * sh};
}

// ---- tabulated Langmuir isotherm (vartheta_langmuir, clAdsorb.c:171-247) ----
// Does segment ab cross segment cd?  (intersect, clAdsorb.c:53-71: the
// determinants in double, the two ratios in single precision.)  The walk's
// decisions hang on these roundings, so nothing here is fused into an FMA.
__host__ __device__ inline int segments_cross(d2 a, d2 b, d2 c, d2 d) {
#pragma clang fp contract(off)
    const float den = (float)((d.y - c.y) * (b.x - a.x) - (d.x - c.x) * (b.y - a.y));
    if (den == 0.0f) return 0;
    const float r1 = (float)((d.x - c.x) * (a.y - c.y) - (d.y - c.y) * (a.x - c.x)) / den;
    const float r2 = (float)((b.x - a.x) * (a.y - c.y) - (b.y - a.y) * (a.x - c.x)) / den;
    return (r1 >= 0.0f && r1 <= 1.0f && r2 >= 0.0f && r2 <= 1.0f) ? 1 : 0;
}
// distance of p from segment ab (distance_from_segment, clAdsorb.c:18-51)
__host__ __device__ inline double segment_distance(d2 a, d2 b, d2 p) {
#pragma clang fp contract(off)
    const double len2 = (b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y);
    const double r = ((p.x - a.x) * (b.x - a.x) + (p.y - a.y) * (b.y - a.y)) / len2;
    if (r >= 0 && r <= 1) {
        const double s = ((a.y - p.y) * (b.x - a.x) - (a.x - p.x) * (b.y - a.y)) / len2;
        return fabs(s) * sqrt(len2);
    }
    const double da = (p.x - a.x) * (p.x - a.x) + (p.y - a.y) * (p.y - a.y);
    const double db = (p.x - b.x) * (p.x - b.x) + (p.y - b.y) * (p.y - b.y);
    return sqrt(da < db ? da : db);
}
// Equilibrium coverage at (P, T) from the wall's table: walk the Delaunay
// triangulation from triangle 0, each move across the edge (bc, then ca,
// then ab) that the segment from the current triangle's centroid to (P, T)
// crosses, at most n_tri moves (get_tri, clAdsorb.c:73-157); stop inside, on
// the crossed edge (within 1e-12), or at the hull with outside = 1; then
// barycentric interpolation in that triangle (interpolate, :159-168).
__device__ inline double isotherm_coverage(const Consts &C, int face, double P, double T, int &outside) {
    const double *iso = C.iso[face];
    const int *tri = C.iso_tri[face], *nbr = C.iso_nbr[face];
    const d2 p = d2{P, T};
    int t = 0;
    outside = 0;
    for (int moves = 0; moves < C.n_tri[face]; moves++) {
        const int ia = tri[3 * t], ib = tri[3 * t + 1], ic = tri[3 * t + 2];
        const d2 a = d2{iso[3 * ia], iso[3 * ia + 1]}, b = d2{iso[3 * ib], iso[3 * ib + 1]};
        const d2 c = d2{iso[3 * ic], iso[3 * ic + 1]};
        const d2 mid = (1.0 / 3.0) * ((a + b) + c);
        const int cab = segments_cross(a, b, mid, p), cbc = segments_cross(b, c, mid, p);
        const int cca = segments_cross(c, a, mid, p);
        if (cab + cbc + cca == 0) break;
        const int k = cbc ? 0 : (cca ? 1 : 2);           // the crossed edge is opposite vertex k
        const d2 e0 = k == 0 ? b : (k == 1 ? c : a), e1 = k == 0 ? c : (k == 1 ? a : b);
        const int next = nbr[3 * t + k];
        if (next == -1) { outside = 1; break; }
        if (segment_distance(e0, e1, p) < 1e-12) break;
        t = next;
    }
    const double *A = iso + 3 * tri[3 * t], *B = iso + 3 * tri[3 * t + 1], *Cv = iso + 3 * tri[3 * t + 2];
    const double den = (B[1] - Cv[1]) * (A[0] - Cv[0]) + (Cv[0] - B[0]) * (A[1] - Cv[1]);
    const double l1 = ((B[1] - Cv[1]) * (P - Cv[0]) + (Cv[0] - B[0]) * (T - Cv[1])) / den;
    const double l2 = ((Cv[1] - A[1]) * (P - Cv[0]) + (A[0] - Cv[0]) * (T - Cv[1])) / den;
    const double l3 = 1 - l1 - l2;
    return l1 * A[2] + l2 * B[2] + l3 * Cv[2];
}
// the equilibrium coverage of an adsorbing wall, given the density and
// temperature of the cell beside it: the constant one, or the table's at
// (D T / 2, T); outside = 1 if (P, T) lies off the table or there is none
__device__ inline double equilibrium_coverage(const Consts &C, int face, double D, double T, int &outside) {
    outside = 0;
    if (C.cads[face]) return C.adsorb[face];
    if (C.n_tri[face] <= 0) { outside = 1; return 0.0; }
    return isotherm_coverage(C, face, 0.5 * D * T, T, outside);
}

// fEQ (clFunctions.c:79-86)
__device__ __forceinline__ d2 shakhov_eq(const Consts &C, d4 p, d2 Q, d2 uv, int v) {
    d2 M = maxwell(C, p, uv, v);
    return M + shakhov(C, p, Q, uv, M);
}

// vanLeer (clFunctions.c:256-260), one component
__device__ __forceinline__ double van_leer(double s1, double s2) {
    return (ocl_sign(s1) + ocl_sign(s2)) * (fabs(s1) * fabs(s2)) / (fabs(s1) + fabs(s2) + 1e-16);
}
// WENO5 slope of both components (g, h) (clFunctions.c:268-312), written
// in the first differences d_k = s_k - s_{k-1}:
//   smoothness  12 b_0 = 13 (d2 - d1)^2 + 3 (3 d2 - d1)^2
//               12 b_1 = 13 (d3 - d2)^2 + 3 (d2 + d3)^2
//               12 b_2 = 13 (d4 - d3)^2 + 3 (3 d3 - d4)^2
//   candidates  f_0 - s2 = -d1/3 + 5 d2/6,  f_1 - s2 = d2/6 + d3/3,
//               f_2 - s2 = 2 d3/3 - d4/6
// with eps scaled by 12 alongside (the weights only see the ratios), so that
// (f_L - s2) is formed without cancelling against s2.  The nonlinear weights
// a_k = d_k / (eps + b_k)^2 are written 1 / e_k^2 with e_k = 12 (eps + b_k)
// / sqrt(d_k): the linear weights d = (0.1, 0.6, 0.3) ride in the constants
// of e_k, which is one multiply and two FMAs from the differences.  They
// enter only through sum(a_k (f_k - s2)) / sum(a_k): in double both sums are
// scaled by prod_k e_k^2, leaving one quotient per component, and the two
// quotients share one reciprocal.  (In float that product can underflow:
// the FP32 form takes reciprocal estimates of the weights instead.)
// e_k = A_k u_k^2 + B_k v_k^2 + E_k with A_k = 13 / sqrt(d_k), B_k = 3 / sqrt(d_k), E_k = 1.2e-15 / sqrt(d_k)
constexpr double kWA0 = 41.10960958218893, kWB0 = 9.486832980505138, kWE0 = 3.794733192202055e-15;
constexpr double kWA1 = 16.782927833565473, kWB1 = 3.8729833462074166, kWE1 = 1.5491933384829667e-15;
constexpr double kWA2 = 23.7346441585572, kWB2 = 5.477225575051661, kWE2 = 2.1908902300206646e-15;
struct WenoParts { double num, den; };
__device__ __forceinline__ WenoParts weno5_parts(double s0, double s1, double s2, double s3, double s4) {
    const double d1 = s1 - s0, d2 = s2 - s1, d3 = s3 - s2, d4 = s4 - s3;
    const double u0 = d2 - d1, v0 = 3.0 * d2 - d1;
    const double u1 = d3 - d2, v1 = d2 + d3;
    const double u2 = d4 - d3, v2 = 3.0 * d3 - d4;
    const double e0 = fma(kWA0 * u0, u0, fma(kWB0 * v0, v0, kWE0));
    const double e1 = fma(kWA1 * u1, u1, fma(kWB1 * v1, v1, kWE1));
    const double e2 = fma(kWA2 * u2, u2, fma(kWB2 * v2, v2, kWE2));
    const double q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
    const double p0 = q1 * q2, p1 = q0 * q2, p2 = q0 * q1;
    // (f_k - s2)
    const double c0 = (-1.0 / 3.0) * d1 + (5.0 / 6.0) * d2;
    const double c1 = (1.0 / 6.0) * d2 + (1.0 / 3.0) * d3;
    const double c2 = (2.0 / 3.0) * d3 - (1.0 / 6.0) * d4;
    return WenoParts{p0 * c0 + p1 * c1 + p2 * c2, (p0 + p1) + p2};
}
// the two slopes (f_L - s2) / dist of a (g, h) stencil
__device__ __forceinline__ void weno5_slope_pair(double g0, double g1, double g2, double g3, double g4, double h0,
                                                 double h1, double h2, double h3, double h4, double rdx, double &sgx,
                                                 double &sgy) {
    const WenoParts a = weno5_parts(g0, g1, g2, g3, g4);
    const WenoParts b = weno5_parts(h0, h1, h2, h3, h4);
    // 1 / (den_a den_b) from the hardware estimate and two Newton steps; the
    // dens are >= (1.2e-15)^4 and far below overflow, so is their product
    const double D = a.den * b.den;
    double r = __builtin_amdgcn_rcp(D);
    double e = fma(-D, r, 1.0);
    r = fma(r, e, r);
    e = fma(-D, r, 1.0);
    r = fma(r, e, r);
    const double ra = r * b.den, rb = r * a.den;     // 1 / den_a, 1 / den_b
    double qa = a.num * ra, qb = b.num * rb;
    qa = fma(ra, fma(-a.den, qa, a.num), qa);        // one residual correction each
    qb = fma(rb, fma(-b.den, qb, b.num), qb);
    sgx = qa * rdx;
    sgy = qb * rdx;
}
// minmod + NND (clFunctions.c:320-352), one component
__device__ __forceinline__ double minmod(double a, double b) {
    if ((fabs(a) < fabs(b)) && (a * b > 0)) return a;
    if ((fabs(b) < fabs(a)) && (a * b > 0)) return b;
    if (a == b) return a;
    return 0.0;
}
__device__ __forceinline__ double nnd(double s0, double s1, double s2, double dx) {
    double ip = s1 + 0.5 * minmod(s2 - s1, s1 - s0);
    return (ip - s1) / dx;
}

// getPrimary / getConserved (clUGKS.c:16-28)
__device__ __forceinline__ d4 to_primary(const Consts &C, d4 w) {
    d4 p;
    p.s3 = 0.5 * w.s0 / (C.gam - 1.0) / (w.s3 - 0.5 * (w.s1 * w.s1 + w.s2 * w.s2) / w.s0);
    p.s0 = w.s0;
    p.s1 = w.s1 / w.s0;
    p.s2 = w.s2 / w.s0;
    return p;
}
__device__ __forceinline__ d4 to_conserved(const Consts &C, d4 p) {
    d4 w;
    w.s3 = 0.5 * p.s0 / p.s3 / (C.gam - 1.0) + 0.5 * p.s0 * (p.s1 * p.s1 + p.s2 * p.s2);
    w.s0 = p.s0;
    w.s1 = p.s1 * p.s0;
    w.s2 = p.s2 * p.s0;
    return w;
}

// microSlope (clUGKS.c:30-45): Maxwellian slope coefficients from dW
__device__ __forceinline__ d4 micro_slope(const Consts &C, d4 p, d4 sw) {
    const double K = C.K;
    d4 m;
    m.s3 = 4.0 * (p.s3 * p.s3) / (K + 2) / p.s0 *
           (2.0 * sw.s3 - 2.0 * p.s1 * sw.s1 - 2.0 * p.s2 * sw.s2 +
            sw.s0 * ((p.s1 * p.s1 + p.s2 * p.s2) - 0.5 * (K + 2) / p.s3));
    m.s2 = 2.0 * p.s3 / p.s0 * (sw.s2 - p.s2 * sw.s0) - p.s2 * m.s3;
    m.s1 = 2.0 * p.s3 / p.s0 * (sw.s1 - p.s1 * sw.s0) - p.s1 * m.s3;
    m.s0 = sw.s0 / p.s0 - p.s1 * m.s1 - p.s2 * m.s2 -
           0.5 * ((p.s1 * p.s1) + (p.s2 * p.s2) + 0.5 * (K + 2) / p.s3) * m.s3;
    return m;
}

// Half/full-range Maxwellian moments (momentU, clUGKS.c:47-80)
struct Moments {
    double Mu[7], MuL[7], MuR[7], Mv[5], Mxi[3];
};
__device__ inline void velocity_moments(const Consts &C, d4 p, Moments &m) {
    double sl = sqrt(p.s3);
    double ex = 0.5 * exp(-p.s3 * (p.s1 * p.s1)) / sqrt(kPi * p.s3);
    m.MuL[0] = 0.5 * erfc(-sl * p.s1);
    m.MuL[1] = p.s1 * m.MuL[0] + ex;
    m.MuR[0] = 0.5 * erfc(sl * p.s1);
    m.MuR[1] = p.s1 * m.MuR[0] - ex;
#pragma unroll
    for (int i = 2; i < 7; ++i) {
        m.MuL[i] = p.s1 * m.MuL[i - 1] + 0.5 * (i - 1) * m.MuL[i - 2] / p.s3;
        m.MuR[i] = p.s1 * m.MuR[i - 1] + 0.5 * (i - 1) * m.MuR[i - 2] / p.s3;
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) m.Mu[i] = m.MuL[i] + m.MuR[i];
    m.Mv[0] = 1.0;
    m.Mv[1] = p.s2;
#pragma unroll
    for (int i = 2; i < 5; ++i) m.Mv[i] = p.s2 * m.Mv[i - 1] + 0.5 * (i - 1) * m.Mv[i - 2] / p.s3;
    m.Mxi[0] = 1.0;
    m.Mxi[1] = 0.5 * C.K / p.s3;
    m.Mxi[2] = (C.K * C.K + 2.0 * C.K) / (4.0 * p.s3 * p.s3);
}
// moment_uv (clUGKS.c:82-93): <u^a v^b xi^d psi>
__device__ __forceinline__ d4 moment_uv(const double *Mu, const double *Mv, const double *Mxi, int a, int b, int d) {
    d4 r;
    r.s0 = Mu[a] * Mv[b] * Mxi[d / 2];
    r.s1 = Mu[a + 1] * Mv[b] * Mxi[d / 2];
    r.s2 = Mu[a] * Mv[b + 1] * Mxi[d / 2];
    r.s3 = 0.5 * (Mu[a + 2] * Mv[b] * Mxi[d / 2] + Mu[a] * Mv[b + 2] * Mxi[d / 2] + Mu[a] * Mv[b] * Mxi[(d + 2) / 2]);
    return r;
}
// moment_au (clUGKS.c:95-104): <a u^a v^b psi>
__device__ __forceinline__ d4 moment_au(d4 A, const double *Mu, const double *Mv, const double *Mxi, int a, int b) {
    return A.s0 * moment_uv(Mu, Mv, Mxi, a, b, 0) + A.s1 * moment_uv(Mu, Mv, Mxi, a + 1, b, 0) +
           A.s2 * moment_uv(Mu, Mv, Mxi, a, b + 1, 0) + (0.5 * A.s3) * moment_uv(Mu, Mv, Mxi, a + 2, b, 0) +
           (0.5 * A.s3) * moment_uv(Mu, Mv, Mxi, a, b + 2, 0) + (0.5 * A.s3) * moment_uv(Mu, Mv, Mxi, a, b, 2);
}

// Time integrals Mt[0..4] of the UGKS flux (clUGKS.c:499-503)
struct TimeInt { double m0, m1, m2, m3, m4; };
__device__ __forceinline__ TimeInt time_integrals(double tau, double dt) {
    TimeInt t;
    double e = exp(-dt / tau);
    t.m3 = tau * (1.0 - e);
    t.m4 = -tau * dt * e + tau * t.m3;
    t.m0 = dt - t.m3;
    t.m1 = -tau * t.m0 + t.m4;
    t.m2 = (dt * dt) / 2.0 - tau * t.m0;
    return t;
}

// distFlux per-velocity flux (clUGKS.c:741-763), before the length factor
__device__ __forceinline__ d2 ugks_face_flux(const Consts &C, d4 prim, d4 aL, d4 aR, d4 aT, double Mxi2,
                                             const TimeInt &mt, d2 Q, d2 uv, int v, d2 fface, d2 fslope) {
    d2 f0 = maxwell(C, prim, uv, v);
    d2 F0 = shakhov(C, prim, Q, uv, f0);
    double uu = uv.x * uv.x + uv.y * uv.y, ux2 = uv.x * uv.x;
    d4 a = upwind(uv.x) ? aL : aR;
    double pg = aT.s0 * f0.x + aT.s1 * uv.x * f0.x + aT.s2 * uv.y * f0.x + 0.5 * aT.s3 * (uu * f0.x + f0.y);
    double ph = aT.s0 * f0.y + aT.s1 * uv.x * f0.y + aT.s2 * uv.y * f0.y + 0.5 * aT.s3 * (uu * f0.y + Mxi2 * f0.x);
    double sg = a.s0 * f0.x + a.s1 * uv.x * f0.x + a.s2 * uv.y * f0.x + 0.5 * a.s3 * (uu * f0.x + f0.y);
    double sh = a.s0 * f0.y + a.s1 * uv.x * f0.y + a.s2 * uv.y * f0.y + 0.5 * a.s3 * (uu * f0.y + Mxi2 * f0.x);
    d2 r;
    r.x = mt.m0 * uv.x * (f0.x + F0.x) + mt.m1 * ux2 * sg + mt.m2 * uv.x * pg + mt.m3 * uv.x * fface.x - mt.m4 * ux2 * fslope.x;
    r.y = mt.m0 * uv.x * (f0.y + F0.y) + mt.m1 * ux2 * sh + mt.m2 * uv.x * ph + mt.m3 * uv.x * fface.y - mt.m4 * ux2 * fslope.y;
    return r;
}

// ---- wave64 helpers ---------------------------------------------------------
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// Multi-value wavefront sums without LDS traffic.  A "row" is 16 lanes.
// v_permlane32_swap exchanges the upper half of its first operand with the
// lower half of its second, v_permlane16_swap the odd rows of the first with
// the even rows of the second (gfx950); after either, the lane-wise sum of the
// two operands holds, in each half (row pair), one of the two inputs summed
// over both halves (rows): a reduce-scatter step that moves no data through
// LDS.  The last four steps add the row rotated by 8, 4, 2 and 1 lanes (DPP);
// every lane of a row then holds the row total, bit for bit the same in each
// lane (each step adds the same two operands in both lanes of a pair).
__device__ __forceinline__ double pack_d(unsigned lo, unsigned hi) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ unsigned lo_d(double x) { return (unsigned)(unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ unsigned hi_d(double x) { return (unsigned)((unsigned long long)__double_as_longlong(x) >> 32); }
// x + y after swapping x's upper 32 lanes with y's lower 32: lanes 0-31 hold
// x summed over both halves, lanes 32-63 y summed over both halves
__device__ __forceinline__ double fold32(double x, double y) {
    const auto l = __builtin_amdgcn_permlane32_swap(lo_d(x), lo_d(y), false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi_d(x), hi_d(y), false, false);
    return pack_d(l[0], h[0]) + pack_d(l[1], h[1]);
}
// the same over rows: rows 0 and 2 hold x summed over rows (0, 1) and (2, 3),
// rows 1 and 3 hold y likewise
__device__ __forceinline__ double fold16(double x, double y) {
    const auto l = __builtin_amdgcn_permlane16_swap(lo_d(x), lo_d(y), false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi_d(x), hi_d(y), false, false);
    return pack_d(l[0], h[0]) + pack_d(l[1], h[1]);
}
template <int CTRL> __device__ __forceinline__ double dpp_d(double x) {
    return pack_d((unsigned)__builtin_amdgcn_mov_dpp((int)lo_d(x), CTRL, 0xf, 0xf, false),
                  (unsigned)__builtin_amdgcn_mov_dpp((int)hi_d(x), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ double row_sum(double x) {
    x += dpp_d<0x128>(x);   // row_ror:8
    x += dpp_d<0x124>(x);   // row_ror:4
    x += dpp_d<0x122>(x);   // row_ror:2
    x += dpp_d<0x121>(x);   // row_ror:1
    return x;
}
// the double in lane l (l wavefront-uniform), to every lane
__device__ __forceinline__ double lane_d(double x, int l) {
    return pack_d((unsigned)__builtin_amdgcn_readlane((int)lo_d(x), l), (unsigned)__builtin_amdgcn_readlane((int)hi_d(x), l));
}
// Totals of a[0..3] over the wavefront: every lane of row r gets the total of a[r].
__device__ __forceinline__ double wave_sum4_rows(const double (&a)[4]) {
    return row_sum(fold16(fold32(a[0], a[2]), fold32(a[1], a[3])));
}
// Totals of a[0..3], to every lane (gathered from the rows).
__device__ __forceinline__ void wave_sum4(double (&a)[4]) {
    const double r = wave_sum4_rows(a);
#pragma unroll
    for (int q = 0; q < 4; q++) a[q] = lane_d(r, 16 * q);
}
// Totals of a[0..15] over the wavefront: every lane of row r gets the totals
// of a[4r .. 4r+3] in t[0..3].
__device__ __forceinline__ void wave_sum16_rows(const double (&a)[16], double (&t)[4]) {
    double s[8];
#pragma unroll
    for (int k = 0; k < 8; k++) s[k] = fold32(a[k], a[k + 8]);      // value k (lanes 0-31), k + 8 (32-63)
#pragma unroll
    for (int k = 0; k < 4; k++) t[k] = row_sum(fold16(s[k], s[k + 4]));   // value k + 4 r in row r
}

}  // namespace ugks
