// ugks_ref_kernels.hip -- one gfx950 kernel per OpenCL kernel of the pyUGKS
// time step, behind the reference's kernel names and argument order.
//
// Mapping: the reference runs one work-item per (cell or face) with a serial
// loop over velocities.  Here a 64-lane wavefront owns one cell or face and
// its lanes stride over the velocities: the heavy per-velocity kernels
// (iFace, distFlux, UGKS_update) form the cell's or face's set-up once and
// loop, the velocity reductions end in a shuffle tree; the light ones (ghost
// fills, initial state) take one (cell, 64-velocity chunk) per wavefront.
// The f layout [i][j][v] of (g,h) pairs makes every wave access a contiguous
// 1 KiB run, one dwordx4 per lane.
#include "ugks_device.hpp"
#include "ugks_runtime.hpp"

using namespace ugks;

namespace {

constexpr int kWaves = 4;                 // waves per 256-thread block
constexpr int kBlock = 64 * kWaves;

inline int nvchunks(const Consts &C) { return (C.NV + 63) / 64; }
inline unsigned blocks_for_waves(long long nwaves) { return (unsigned)((nwaves + kWaves - 1) / kWaves); }
inline unsigned blocks_for_threads(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ long long wave_id() { return (long long)blockIdx.x * kWaves + (threadIdx.x >> 6); }
// the same, known to the compiler to be uniform over the wavefront
__device__ __forceinline__ long long wave_id_u() {
    return (long long)blockIdx.x * kWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ long long thread_id() { return (long long)blockIdx.x * kBlock + threadIdx.x; }

// streaming (non-temporal) store of one (g, h) pair that nothing in the same
// launch reads again
__device__ __forceinline__ void stream_d2(d2 *p, d2 x) {
    typedef double v2d __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(v2d{x.x, x.y}, reinterpret_cast<v2d *>(p));
}

// interior-face domain of the flux kernels (clUGKS.c:218-225, ugks_block.py:1105)
struct FaceRange { int i0, ni, j0, nj; };
__host__ __device__ inline FaceRange face_range(int ni, int nj, int face, int ob, int ot) {
    if (face == SOUTH) return FaceRange{0, ni, ob, nj + 1 - ot - ob};
    return FaceRange{ob, ni + 1 - ot - ob, 0, nj};
}

// ---------------------------------------------------------------- geometry
__global__ void k_cellGeom(Consts C, const d2 *xy, double *area, d2 *centre, d2 *mid_side, d2 *normal,
                           double *side_length) {
    long long t = thread_id();
    if (t >= (long long)C.NI * C.NJ) return;
    int i = (int)(t / C.NJ), j = (int)(t % C.NJ);
    const int NY = C.NJ + 1;
    d2 A = xy[(size_t)(i + 1) * NY + j], B = xy[(size_t)(i + 1) * NY + j + 1];
    d2 Cc = xy[(size_t)i * NY + j + 1], D = xy[(size_t)i * NY + j];
    d2 tS = A - D, tW = D - Cc;
    d2 nS = mk2(-tS.y, tS.x), nW = mk2(-tW.y, tW.x);
    normal[sidx(C, i, j, SOUTH)] = nS / length(nS);
    normal[sidx(C, i, j, WEST)] = nW / length(nW);
    double ar = 0.5 * ((B.x + A.x) * (B.y - A.y) + (Cc.x + B.x) * (Cc.y - B.y) + (D.x + Cc.x) * (D.y - Cc.y) +
                       (A.x + D.x) * (A.y - D.y));
    area[cidx(C, i, j)] = ar;
    double cx = 1.0 / (ar * 6.0) *
                ((B.y - A.y) * (A.x * A.x + A.x * B.x + B.x * B.x) + (Cc.y - B.y) * (B.x * B.x + B.x * Cc.x + Cc.x * Cc.x) +
                 (D.y - Cc.y) * (Cc.x * Cc.x + Cc.x * D.x + D.x * D.x) + (A.y - D.y) * (D.x * D.x + D.x * A.x + A.x * A.x));
    double cy = -1.0 / (ar * 6.0) *
                ((B.x - A.x) * (A.y * A.y + A.y * B.y + B.y * B.y) + (Cc.x - B.x) * (B.y * B.y + B.y * Cc.y + Cc.y * Cc.y) +
                 (D.x - Cc.x) * (Cc.y * Cc.y + Cc.y * D.y + D.y * D.y) + (A.x - D.x) * (D.y * D.y + D.y * A.y + A.y * A.y));
    if (fabs(cx) > 1.0e6) cx = __builtin_nan("");
    if (fabs(cy) > 1.0e6) cy = __builtin_nan("");
    centre[cidx(C, i, j)] = mk2(cx, cy);
    mid_side[sidx(C, i, j, SOUTH)] = mk2(0.5 * (D.x + A.x), 0.5 * (D.y + A.y));
    mid_side[sidx(C, i, j, WEST)] = mk2(0.5 * (D.x + Cc.x), 0.5 * (D.y + Cc.y));
    side_length[sidx(C, i, j, SOUTH)] = sqrt((D.x - A.x) * (D.x - A.x) + (D.y - A.y) * (D.y - A.y));
    side_length[sidx(C, i, j, WEST)] = sqrt((Cc.x - D.x) * (Cc.x - D.x) + (Cc.y - D.y) * (Cc.y - D.y));
}

// paraBC (clKernel.c:141-218): constant wall values, or expressions in the
// wall cell's (s, x, y) = para_def and the time t
__global__ void k_paraBC(Consts C, const d4 *para_def, d4 *wall_prop, double t) {
    long long ci = thread_id();
    int wl = C.ni > C.nj ? C.ni : C.nj;
    if (ci >= wl) return;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        int along = (w == GNORTH || w == GSOUTH) ? C.ni : C.nj;
        if (ci >= along) continue;
        const d4 pa = C.any_udf ? para_def[(size_t)ci * 4 + w] : d4{0.0, 0.0, 0.0, 0.0};
        wall_prop[(size_t)ci * 4 + w] = para_wall(C, w, pa, t);
    }
}

// clFindDT (clKernel.c:224-308)
__global__ void k_clFindDT(Consts C, const d2 *xy, const double *area, const d4 *macro, double *time_step, double Kn) {
    long long t = thread_id();
    if (t >= (long long)C.ni * C.nj) return;
    int mi = (int)(t / C.nj), mj = (int)(t % C.nj);
    const int NY = C.NJ + 1, G = C.G;
    d2 a = xy[(size_t)(mi + 1 + G) * NY + mj + G], b = xy[(size_t)(mi + 1 + G) * NY + mj + 1 + G];
    d2 c = xy[(size_t)(mi + G) * NY + mj + 1 + G], d = xy[(size_t)(mi + G) * NY + mj + G];
    double dx = fmax(fmax(fmax(fabs(a.x - b.x), fabs(c.x - b.x)), fmax(fabs(c.x - d.x), fabs(a.x - d.x))),
                     fmax(fabs(c.x - a.x), fabs(b.x - d.x)));
    double dy = fmax(fmax(fmax(fabs(a.y - b.y), fabs(c.y - b.y)), fmax(fabs(c.y - d.y), fabs(a.y - d.y))),
                     fmax(fabs(c.y - a.y), fabs(b.y - d.y)));
    d4 m = macro[midx(C, mi, mj)];
    double sos = sqrt(0.5 * C.gam / m.s3);
    double lo_u = fmin(C.umax, fabs(m.s1) + sos), lo_v = fmin(C.vmax, fabs(m.s2) + sos);
    double hi_u = fmax(C.umax, fabs(m.s1) + sos), hi_v = fmax(C.vmax, fabs(m.s2) + sos);
    double scale = (erf(10.0 * (Kn - 1e-3)) + 1) / 2.0;
    double u = scale * (hi_u - lo_u) + lo_u;
    double v = scale * (hi_v - lo_v) + lo_v;
    time_step[midx(C, mi, mj)] = (dx * u + dy * v) / area[cidx(C, mi + G, mj + G)];
}

// maximum of time_step, one 1024-thread block
__global__ void __launch_bounds__(1024) k_max(const double *x, long long n, double *out) {
    __shared__ double red[16];
    double m = -__builtin_inf();
    for (long long k = threadIdx.x; k < n; k += 1024) m = fmax(m, x[k]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x < 64) {
        m = threadIdx.x < 16 ? red[threadIdx.x] : -__builtin_inf();
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if (threadIdx.x == 0) out[0] = m;
    }
}

// ---------------------------------------------------------------- cells
// initFunctions (clKernel.c:382-412): f = fEQ(macro, Q) in the interior
__global__ void k_initFunctions(Consts C, d2 *F, const d4 *macro, const d2 *gQ) {
    long long w = wave_id();
    int nvc = (C.NV + 63) / 64;
    if (w >= (long long)C.ni * C.nj * nvc) return;
    int cell = (int)(w / nvc), v = (int)(w % nvc) * 64 + lane_id();
    if (v >= C.NV) return;
    int mi = cell / C.nj, mj = cell % C.nj;
    F[fidx(C, mi + C.G, mj + C.G, v)] = shakhov_eq(C, macro[midx(C, mi, mj)], gQ[midx(C, mi, mj)], C.quad[v], v);
}

// calcMacro (clKernel.c:414-478)
__global__ void k_calcMacro(Consts C, const d2 *F, d4 *macro) {
    long long w = wave_id();
    if (w >= (long long)C.ni * C.nj) return;
    int mi = (int)(w / C.nj), mj = (int)(w % C.nj), lane = lane_id();
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int v = lane; v < C.NV; v += 64) {
        d2 f = F[fidx(C, mi + C.G, mj + C.G, v)];
        d2 e = C.quad[v];
        s0 += f.x;
        s1 += e.x * f.x;
        s2 += e.y * f.x;
        s3 += 0.5 * ((e.x * e.x + e.y * e.y) * f.x + f.y);
    }
    s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
    if (lane == 0) {
        d4 p;
        p.s3 = 0.5 * s0 / (C.gam - 1.0) / (s3 - 0.5 * (s1 * s1 + s2 * s2) / s0);
        p.s0 = s0; p.s1 = s1 / s0; p.s2 = s2 / s0;
        macro[midx(C, mi, mj)] = p;
    }
}

// heat flux about a mean velocity (clKernel.c:513-514, clUGKS.c:555-556)
__device__ __forceinline__ void acc_heat(double &qx, double &qy, d2 f, d2 uv, d4 p, double K) {
    double cx = uv.x - p.s1, cy = uv.y - p.s2;
    double t = (cx * cx + cy * cy) * f.x + (f.y / K);
    qx += 0.5 * (cx * t);
    qy += 0.5 * (cy * t);
}

// calcQ (clKernel.c:480-543)
__global__ void k_calcQ(Consts C, const d2 *F, const d4 *macro, d2 *gQ) {
    long long w = wave_id();
    if (w >= (long long)C.ni * C.nj) return;
    int mi = (int)(w / C.nj), mj = (int)(w % C.nj), lane = lane_id();
    d4 p = macro[midx(C, mi, mj)];
    double qx = 0, qy = 0;
    for (int v = lane; v < C.NV; v += 64) acc_heat(qx, qy, F[fidx(C, mi + C.G, mj + C.G, v)], C.quad[v], p, C.K);
    qx = wave_sum(qx); qy = wave_sum(qy);
    if (lane == 0) gQ[midx(C, mi, mj)] = mk2(qx, qy);
}

// getInternal (clKernel.c:315-372): the translational temperatures Tx, Ty, Tz
// and T = (Tx + Ty + Tz) / B, and the stress (tau_xx, tau_xy, tau_yy, 0), from
// the moments of f; a diagnostic the HDF5 output writes (ugks_block.py:1335-1348)
__global__ void k_getInternal(Consts C, const d2 *F, d4 *Txyz, d4 *Stress) {
    long long w = wave_id();
    if (w >= (long long)C.ni * C.nj) return;
    int mi = (int)(w / C.nj), mj = (int)(w % C.nj), lane = lane_id();
    double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};      // D, M1, M2, M4, M5, M6, H
    for (int v = lane; v < C.NV; v += 64) {
        const d2 f = F[fidx(C, mi + C.G, mj + C.G, v)], uv = C.quad[v];
        m[0] += f.x;
        m[1] += uv.x * f.x;
        m[2] += uv.y * f.x;
        m[3] += uv.x * uv.x * f.x;
        m[4] += uv.y * uv.y * f.x;
        m[5] += uv.x * uv.y * f.x;
        m[6] += f.y;
    }
#pragma unroll
    for (int q = 0; q < 7; q++) m[q] = wave_sum(m[q]);
    if (lane != 0) return;
    const double D = m[0], U = m[1] / D, V = m[2] / D;
    const double Tx = 2 * (m[3] / D - U * U), Ty = 2 * (m[4] / D - V * V), Tz = 2 * (m[6] / D);
    const double T = (Tx + Ty + Tz) / C.B;
    Txyz[midx(C, mi, mj)] = mk4(Tx, Ty, Tz, T);
    Stress[midx(C, mi, mj)] = mk4(m[3] - D * (U * U + T / 2.0), m[5] - D * (U * V + T / 2.0),
                                  m[4] - D * (V * V + T / 2.0), 0.0);
}

// ---------------------------------------------------------------- faces
// iFace (clUGKS.c:110-199): upwind reconstruction of f and its slope at a face.
// One wavefront per face strides over the velocities; the face's geometry
// (the stencil distances of both upwind choices) is formed once, and each
// lane walks its own upwind stencil from a per-lane start and step.
template <int METHOD>
__device__ __forceinline__ void iface_face(const Consts &C, const d2 *__restrict__ F, d2 *__restrict__ iface_f,
                                           d2 *__restrict__ fsigma, const d2 *__restrict__ centre, d2 mside, d2 n,
                                           int gi, int gj, int face) {
    constexpr int SL = METHOD == 1 ? 5 : 3, MID = METHOD == 1 ? 2 : 1;
    // stencil cell si of upwind choice delta lies `off` cells along the face
    // normal: delta = 1 at -MID-1+si, delta = 0 at MID-si
    const long long step = face == SOUTH ? (long long)C.NV : (long long)C.NJ * C.NV;
    const int di = face == SOUTH ? 0 : 1, dj = 1 - di;
    double dist[2], dlen[2], rdist[2];
#pragma unroll
    for (int delta = 0; delta < 2; ++delta) {
        const int o0 = delta ? -MID - 1 : MID, os = delta ? 1 : -1;
        const d2 c0 = centre[cidx(C, gi + di * o0, gj + dj * o0)];
        const d2 c1 = centre[cidx(C, gi + di * (o0 + os), gj + dj * (o0 + os))];
        const d2 cm = centre[cidx(C, gi + di * (o0 + MID * os), gj + dj * (o0 + MID * os))];
        dist[delta] = length(mside - cm);
        dlen[delta] = length(c1 - c0);
        rdist[delta] = 1.0 / dist[delta];
    }
    const d2 *fc = F + fidx(C, gi, gj, 0);
    const size_t out0 = fidx(C, gi, gj, 0);
    for (int v = lane_id(); v < C.NV; v += 64) {
        const d2 uv = face_velocity(C, v, n);
        const int delta = upwind(uv.x);
        const long long o0 = delta ? -MID - 1 : MID, os = delta ? step : -step;
        const d2 *p = fc + v + o0 * step;
        d2 fs[SL];
#pragma unroll
        for (int si = 0; si < SL; ++si) fs[si] = p[si * os];
        const double dst = delta ? dist[1] : dist[0];
        d2 sig;
        if constexpr (METHOD == 0) {
            // clUGKS.c:179-180: both differences over the distance between cells 0 and 1
            const double d = delta ? dlen[1] : dlen[0];
            sig.x = van_leer((fs[1].x - fs[0].x) / d, (fs[2].x - fs[1].x) / d);
            sig.y = van_leer((fs[1].y - fs[0].y) / d, (fs[2].y - fs[1].y) / d);
        } else if constexpr (METHOD == 1) {
            // the reference's WENO5 (clFunctions.c:268-312) rearranged to one
            // quotient per component; it agrees with the literal form to rounding
            weno5_slope_pair(fs[0].x, fs[1].x, fs[2].x, fs[3].x, fs[4].x, fs[0].y, fs[1].y, fs[2].y, fs[3].y, fs[4].y,
                             delta ? rdist[1] : rdist[0], sig.x, sig.y);
        } else {
            sig.x = nnd(fs[0].x, fs[1].x, fs[2].x, dst);
            sig.y = nnd(fs[0].y, fs[1].y, fs[2].y, dst);
        }
        stream_d2(iface_f + out0 + v, fs[MID] + sig * dst);
        stream_d2(fsigma + out0 + v, sig);
    }
}

__global__ void __launch_bounds__(kBlock) k_iFace(Consts C, const d2 *__restrict__ F, d2 *__restrict__ iface_f,
                                                  d2 *__restrict__ fsigma, const d2 *__restrict__ centre,
                                                  const d2 *__restrict__ mid_side, const d2 *__restrict__ normal,
                                                  int face) {
    const long long w = wave_id_u();
    const int ilim = face == SOUTH ? C.ni : C.ni + 1, jlim = face == SOUTH ? C.nj + 1 : C.nj;
    if (w >= (long long)ilim * jlim) return;
    const int gi = (int)(w / jlim) + C.G, gj = (int)(w % jlim) + C.G;
    const d2 n = normal[sidx(C, gi, gj, face)], ms = mid_side[sidx(C, gi, gj, face)];
    if (C.method == 0) iface_face<0>(C, F, iface_f, fsigma, centre, ms, n, gi, gj, face);
    else if (C.method == 1) iface_face<1>(C, F, iface_f, fsigma, centre, ms, n, gi, gj, face);
    else iface_face<2>(C, F, iface_f, fsigma, centre, ms, n, gi, gj, face);
}

// getFaceCons / getAL / getAR (clUGKS.c:209-399): sum psi f in face coordinates
// over the face distribution (src = 0), the left cell (1) or the right cell (2)
__global__ void k_faceMoments(Consts C, const d2 *F, const d2 *normal, int face, d4 *out, int ob, int ot, int src) {
    FaceRange r = face_range(C.ni, C.nj, face, ob, ot);
    long long w = wave_id();
    if (w >= (long long)r.ni * r.nj) return;
    int gi = (int)(w / r.nj) + r.i0 + C.G, gj = (int)(w % r.nj) + r.j0 + C.G, lane = lane_id();
    int ci = src == 1 ? gi - face : gi, cj = src == 1 ? gj - (1 - face) : gj;
    d2 n = normal[sidx(C, gi, gj, face)];
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int v = lane; v < C.NV; v += 64) {
        d2 f = F[fidx(C, ci, cj, v)];
        d2 uv = face_velocity(C, v, n);
        s0 += f.x;
        s1 += uv.x * f.x;
        s2 += uv.y * f.x;
        s3 += 0.5 * ((uv.x * uv.x + uv.y * uv.y) * f.x + f.y);
    }
    s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
    if (lane == 0) out[cidx(C, gi, gj)] = mk4(s0, s1, s2, s3);
}

// getPLR (clUGKS.c:401-442)
__global__ void k_getPLR(Consts C, const d2 *centre, const d2 *mid_side, int face, d4 *primary, d4 *gaL, d4 *gaR,
                         int ob, int ot) {
    FaceRange r = face_range(C.ni, C.nj, face, ob, ot);
    long long t = thread_id();
    if (t >= (long long)r.ni * r.nj) return;
    int gi = (int)(t / r.nj) + r.i0 + C.G, gj = (int)(t % r.nj) + r.j0 + C.G;
    d4 w = primary[cidx(C, gi, gj)];
    d4 prim = to_primary(C, w);
    d2 ms = mid_side[sidx(C, gi, gj, face)];
    d4 sw = (w - gaL[cidx(C, gi, gj)]) / length(ms - centre[cidx(C, gi - face, gj - (1 - face))]);
    gaL[cidx(C, gi, gj)] = micro_slope(C, prim, sw);
    sw = (gaR[cidx(C, gi, gj)] - w) / length(ms - centre[cidx(C, gi, gj)]);
    gaR[cidx(C, gi, gj)] = micro_slope(C, prim, sw);
    primary[cidx(C, gi, gj)] = prim;
}

// initMacroFlux (clUGKS.c:446-518)
__global__ void k_initMacroFlux(Consts C, d4 *flux_macro, int face, double dt, const d4 *primary, const d4 *gaL,
                                const d4 *gaR, d4 *gaT, double *gMxi, int ob, int ot) {
    FaceRange r = face_range(C.ni, C.nj, face, ob, ot);
    long long t = thread_id();
    if (t >= (long long)r.ni * r.nj) return;
    int gi = (int)(t / r.nj) + r.i0 + C.G, gj = (int)(t % r.nj) + r.j0 + C.G;
    size_t k = cidx(C, gi, gj);
    d4 prim = primary[k], aL = gaL[k], aR = gaR[k];
    Moments m;
    velocity_moments(C, prim, m);
    gMxi[3 * k + 0] = m.Mxi[0];
    gMxi[3 * k + 1] = m.Mxi[1];
    gMxi[3 * k + 2] = m.Mxi[2];
    d4 sw = (-prim.s0) * (moment_au(aL, m.MuL, m.Mv, m.Mxi, 1, 0) + moment_au(aR, m.MuR, m.Mv, m.Mxi, 1, 0));
    d4 aT = micro_slope(C, prim, sw);
    gaT[k] = aT;
    TimeInt mt = time_integrals(relax_time(C, prim), dt);
    d4 Mau0 = moment_uv(m.Mu, m.Mv, m.Mxi, 1, 0, 0);
    d4 MauL = moment_au(aL, m.MuL, m.Mv, m.Mxi, 2, 0);
    d4 MauR = moment_au(aR, m.MuR, m.Mv, m.Mxi, 2, 0);
    d4 MauT = moment_au(aT, m.Mu, m.Mv, m.Mxi, 1, 0);
    flux_macro[k] = prim.s0 * (mt.m0 * Mau0 + mt.m1 * (MauL + MauR) + mt.m2 * MauT);
}

// calcFaceQ (clUGKS.c:522-585)
__global__ void k_calcFaceQ(Consts C, const d2 *iface_f, const d4 *primary, const d2 *normal, d2 *faceQ, int face,
                            int ob, int ot) {
    FaceRange r = face_range(C.ni, C.nj, face, ob, ot);
    long long w = wave_id();
    if (w >= (long long)r.ni * r.nj) return;
    int gi = (int)(w / r.nj) + r.i0 + C.G, gj = (int)(w % r.nj) + r.j0 + C.G, lane = lane_id();
    d2 n = normal[sidx(C, gi, gj, face)];
    d4 p = primary[cidx(C, gi, gj)];
    double qx = 0, qy = 0;
    for (int v = lane; v < C.NV; v += 64) acc_heat(qx, qy, iface_f[fidx(C, gi, gj, v)], face_velocity(C, v, n), p, C.K);
    qx = wave_sum(qx); qy = wave_sum(qy);
    if (lane == 0) faceQ[cidx(C, gi, gj)] = mk2(qx, qy);
}

// macroFlux (clUGKS.c:587-681)
__global__ void k_macroFlux(Consts C, const d2 *flux_f, const d2 *fsigma, d4 *flux_macro, const d2 *normal,
                            const double *side_length, int face, double dt, const d4 *primary, const d2 *faceQ, int ob,
                            int ot) {
    FaceRange r = face_range(C.ni, C.nj, face, ob, ot);
    long long w = wave_id();
    if (w >= (long long)r.ni * r.nj) return;
    int gi = (int)(w / r.nj) + r.i0 + C.G, gj = (int)(w % r.nj) + r.j0 + C.G, lane = lane_id();
    size_t k = cidx(C, gi, gj);
    d2 n = normal[sidx(C, gi, gj, face)];
    d4 prim = primary[k];
    TimeInt mt = time_integrals(relax_time(C, prim), dt);
    d2 Q = faceQ[k];
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int v = lane; v < C.NV; v += 64) {
        d2 uv = face_velocity(C, v, n);
        d2 F0 = shakhov(C, prim, Q, uv, maxwell(C, prim, uv, v));
        d2 fd = flux_f[fidx(C, gi, gj, v)], sl = fsigma[fidx(C, gi, gj, v)];
        double uu = uv.x * uv.x + uv.y * uv.y;
        s0 += mt.m0 * uv.x * F0.x + mt.m3 * uv.x * fd.x - mt.m4 * uv.x * uv.x * sl.x;
        s1 += mt.m0 * uv.x * uv.x * F0.x + mt.m3 * uv.x * uv.x * fd.x - mt.m4 * uv.x * uv.x * uv.x * sl.x;
        s2 += mt.m0 * uv.y * uv.x * F0.x + mt.m3 * uv.y * uv.x * fd.x - mt.m4 * uv.y * uv.x * uv.x * sl.x;
        s3 += mt.m0 * 0.5 * (uv.x * uu * F0.x + uv.x * F0.y) + mt.m3 * 0.5 * (uv.x * uu * fd.x + uv.x * fd.y) -
              mt.m4 * 0.5 * (uv.x * uv.x * uu * sl.x + uv.x * uv.x * sl.y);
    }
    s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
    if (lane == 0) {
        d4 o = mk4(s0, s1, s2, s3) + flux_macro[k];
        d2 g = to_global(mk2(o.s1, o.s2), n);
        o.s1 = g.x; o.s2 = g.y;
        flux_macro[k] = side_length[sidx(C, gi, gj, face)] * o;
    }
}

// distFlux (clUGKS.c:683-772): per-velocity interface flux, in place.  One
// wavefront per face: the relaxation time, time integrals and face state are
// formed once, then the wavefront strides over the velocities.
__global__ void __launch_bounds__(kBlock) k_distFlux(Consts C, d2 *__restrict__ flux_f, const d2 *__restrict__ fsigma,
                                                     const d2 *__restrict__ normal,
                                                     const double *__restrict__ side_length, int face, double dt,
                                                     const d4 *__restrict__ primary, const d4 *__restrict__ gaL,
                                                     const d4 *__restrict__ gaR, const d4 *__restrict__ gaT,
                                                     const double *__restrict__ gMxi, const d2 *__restrict__ faceQ,
                                                     int ob, int ot) {
    FaceRange r = face_range(C.ni, C.nj, face, ob, ot);
    const long long w = wave_id_u();
    if (w >= (long long)r.ni * r.nj) return;
    const int gi = (int)(w / r.nj) + r.i0 + C.G, gj = (int)(w % r.nj) + r.j0 + C.G;
    const size_t k = cidx(C, gi, gj);
    const d4 prim = primary[k], aL = gaL[k], aR = gaR[k], aT = gaT[k];
    const TimeInt mt = time_integrals(relax_time(C, prim), dt);
    const d2 n = normal[sidx(C, gi, gj, face)], Q = faceQ[k];
    const double Mxi2 = gMxi[3 * k + 2], len = side_length[sidx(C, gi, gj, face)];
    const size_t f0 = fidx(C, gi, gj, 0);
    for (int v = lane_id(); v < C.NV; v += 64) {
        const d2 uv = face_velocity(C, v, n);
        const d2 ff = ugks_face_flux(C, prim, aL, aR, aT, Mxi2, mt, Q, uv, v, flux_f[f0 + v], fsigma[f0 + v]);
        stream_d2(flux_f + f0 + v, len * ff);
    }
}

// ---------------------------------------------------------------- walls
// wall-face addressing of the wall kernels (clUGKS.c:795-824)
struct WallSite { int gi, gj, rot, fid, ci; };
__device__ __forceinline__ WallSite wall_site(const Consts &C, int face, int ci) {
    switch (face) {
    case GNORTH: return WallSite{ci + C.G, C.NJ - C.G, -1, SOUTH, ci};
    case GEAST: return WallSite{C.NI - C.G, ci + C.G, -1, WEST, ci};
    case GSOUTH: return WallSite{ci + C.G, C.G, 1, SOUTH, ci};
    default: return WallSite{C.G, ci + C.G, 1, WEST, ci};
    }
}
__host__ __device__ inline int wall_count(int ni, int nj, int face) { return (face == GNORTH || face == GSOUTH) ? ni : nj; }

// The wall kernels take one 256-thread workgroup per wall face (a launch has
// only as many faces as the wall has cells, so one wavefront each would leave
// most of the chip idle); its threads stride over the velocities, and these
// sums end over the four wavefronts: every thread gets the totals.
template <int N> __device__ __forceinline__ void wall_face_sum(double (&a)[N], double (*red)[N]) {
#pragma unroll
    for (int q = 0; q < N; ++q) a[q] = wave_sum(a[q]);
    if (lane_id() == 0) {
#pragma unroll
        for (int q = 0; q < N; ++q) red[threadIdx.x >> 6][q] = a[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; ++q) a[q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
}

// accommodatingWallDist (clUGKS.c:776-906): diffuse re-emission with zero net mass flux
__global__ void __launch_bounds__(kBlock) k_accommodatingWallDist(Consts C, const d2 *normal, int face,
                                                                  const d4 *wall_prop, d2 *flux_f) {
    __shared__ double red[kWaves][2];
    WallSite s = wall_site(C, face, (int)blockIdx.x);
    d2 n = normal[sidx(C, s.gi, s.gj, s.fid)];
    d4 wall = wall_prop[(size_t)s.ci * 4 + face];
    d2 wl = to_local(mk2(wall.s1, wall.s2), n);
    wall.s1 = wl.x; wall.s2 = wl.y;
    double io[2] = {0.0, 0.0};   // outgoing (towards the wall) and re-emitted mass flux
    for (int v = threadIdx.x; v < C.NV; v += kBlock) {
        d2 uv = face_velocity(C, v, n);
        if (leaves_wall(uv.x, s.rot)) io[1] -= uv.x * maxwell(C, wall, uv, v).x;
        else io[0] += uv.x * flux_f[fidx(C, s.gi, s.gj, v)].x;
    }
    wall_face_sum<2>(io, red);
    const double ratio = io[0] / io[1];
    for (int v = threadIdx.x; v < C.NV; v += kBlock) {
        d2 uv = face_velocity(C, v, n);
        if (leaves_wall(uv.x, s.rot)) flux_f[fidx(C, s.gi, s.gj, v)] = ratio * maxwell(C, wall, uv, v);
    }
}

// wallFlux (clUGKS.c:910-1011): first-order wall flux u f len dt and its moments
__global__ void __launch_bounds__(kBlock) k_wallFlux(Consts C, const d2 *normal, const double *side_length, int face,
                                                     d2 *flux_f, d4 *flux_macro, double dt) {
    __shared__ double red[kWaves][4];
    WallSite s = wall_site(C, face, (int)blockIdx.x);
    d2 n = normal[sidx(C, s.gi, s.gj, s.fid)];
    double L = side_length[sidx(C, s.gi, s.gj, s.fid)];
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int v = threadIdx.x; v < C.NV; v += kBlock) {
        d2 uv = face_velocity(C, v, n);
        size_t fi = fidx(C, s.gi, s.gj, v);
        d2 wd = flux_f[fi];
        flux_f[fi] = ((uv.x * wd) * L) * dt;
        m[0] += uv.x * wd.x;
        m[1] += uv.x * uv.x * wd.x;
        m[2] += uv.x * uv.y * wd.x;
        m[3] += 0.5 * uv.x * ((uv.x * uv.x + uv.y * uv.y) * wd.x + wd.y);
    }
    wall_face_sum<4>(m, red);
    if (threadIdx.x == 0) {
        d2 g = to_global(mk2(m[1], m[2]), n);
        flux_macro[cidx(C, s.gi, s.gj)] = (dt * L) * mk4(m[0], g.x, g.y, m[3]);
    }
}

// wallMassEnergyFluxes (clUGKS.c:1019-1147): in/out mass, momentum, energy per wall face
__global__ void __launch_bounds__(kBlock) k_wallMassEnergyFluxes(Consts C, const d2 *normal, int face,
                                                                 const d2 *flux_f, double *wall_flux) {
    __shared__ double red[kWaves][8];
    WallSite s = wall_site(C, face, (int)blockIdx.x);
    d2 n = normal[sidx(C, s.gi, s.gj, s.fid)];
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // in0, out0, in1, out1, ...
    for (int v = threadIdx.x; v < C.NV; v += kBlock) {
        d2 uv = face_velocity(C, v, n);
        d2 wd = flux_f[fidx(C, s.gi, s.gj, v)];
        int o = leaves_wall(uv.x, s.rot);
        double e = 0.5 * ((uv.x * uv.x + uv.y * uv.y) * wd.x + wd.y);
        a[0 + o] += wd.x;
        a[2 + o] += uv.x * wd.x;
        a[4 + o] += uv.y * wd.x;
        a[6 + o] += e;
    }
    wall_face_sum<8>(a, red);
    if (threadIdx.x == 0) {
        double *o = wall_flux + ((size_t)s.ci * 4 + face) * 8;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            o[2 * q] = (double)(-s.rot) * a[2 * q];
            o[2 * q + 1] = (double)s.rot * a[2 * q + 1];
        }
    }
}

// stickingProbability (clAdsorb.c:6-16)
__device__ __forceinline__ double sticking(const Consts &C, d2 uv, double T, int face) {
    double nrm = 1.0 - C.beta_n[face] * exp(-uv.x * uv.x / T);
    double tan_ = 1.0 - C.beta_t[face] * (1.0 - exp(-uv.y * uv.y / T));
    return nrm * tan_;
}

// adsorbingWall_P1 (clAdsorb.c:253-482), constant or tabulated isotherm
__global__ void k_adsorbingWall_P1(Consts C, const d2 *normal, int face, d4 *wall_cover, d2 *wall_dist,
                                   const d2 *flux_f, const d4 *macro, double dt, double *flag) {
    long long w = wave_id();
    if (w >= wall_count(C.ni, C.nj, face)) return;
    WallSite s = wall_site(C, face, (int)w);
    int lane = lane_id();
    d4 cell;
    switch (face) {
    case GNORTH: cell = macro[midx(C, s.ci, s.gj - C.G - 1)]; break;
    case GEAST: cell = macro[midx(C, s.gi - C.G - 1, s.ci)]; break;
    case GSOUTH: cell = macro[midx(C, s.ci, s.gj - C.G)]; break;
    default: cell = macro[midx(C, s.gi - C.G, s.ci)]; break;
    }
    double T = 1.0 / cell.s3;
    d2 n = normal[sidx(C, s.gi, s.gj, s.fid)];
    size_t ck = (size_t)s.ci * 4 + face;
    double vartheta = wall_cover[ck].s0;
    double gf = C.gamma_f[face];
    double refl_sum = 0, ads_sum = 0, adj_sum = 0;
    for (int v = lane; v < C.NV; v += 64) {
        d2 uv = face_velocity(C, v, n);
        d2 fd = flux_f[fidx(C, s.gi, s.gj, v)];
        int delta = leaves_wall(uv.x, s.rot);
        d2 total = ((-s.rot * uv.x) * (1 - delta)) * fd;
        double beta = sticking(C, uv, T, face);
        d2 ads = mk2(beta * total.x * gf * (1.0 - vartheta), beta * total.y * gf * (1.0 - vartheta));
        double sgn = ocl_sign(total.x);
        int fs = (int)sgn;
        if ((fs * ads.x) > (fs * total.x)) ads = total;
        d2 refl = total - ads;
        refl_sum += refl.x;
        ads_sum += ads.x;
        adj_sum += beta * total.x;
        wall_dist[(size_t)s.ci * C.NV + v] = refl / uv.x;
    }
    double r_init = wave_sum(refl_sum), a_init = wave_sum(ads_sum), adjusted = wave_sum(adj_sum);
    double a_final = a_init, dvtheta = 0.0, des = 0.0, correct = 1.0;
    bool early = false;
    if (gf > 0.0) {
        if (C.dyn[face] == 1) {
            dvtheta = dt * C.alpha_p[face] * a_init;
            if (dvtheta > (1 - vartheta)) {
                dvtheta = 1 - vartheta;
                a_final = dvtheta / dt * C.alpha_p[face];
            }
        }
        int outside;
        const double veq = equilibrium_coverage(C, face, cell.s0, T, outside);
        if (outside) {
            early = true;   // off the isotherm table: reported, the coverage is left as it was
        } else {
            double gamma_b = gf * (1.0 / veq - 1.0) * adjusted;
            des = gamma_b * vartheta;
            if (C.dyn[face] == 1) {
                dvtheta -= dt * C.alpha_p[face] * des;
                if ((vartheta + dvtheta) < 0.0) des = vartheta / (dt * C.alpha_p[face]);
            }
        }
    }
    if (early) {
        if (lane == 0) flag[0] = ERR_CLAMPING_ADSORB_ISOTHERM;
        return;
    }
    double r_final = r_init + a_init - a_final;
    correct = r_final / r_init;
    if (lane == 0) {
        d4 cv = wall_cover[ck];
        cv.s0 += dvtheta;
        cv.s1 = r_final; cv.s2 = a_final; cv.s3 = des;
        wall_cover[ck] = cv;
    }
    if (correct != 1.0)
        for (int v = lane; v < C.NV; v += 64) {
            size_t k = (size_t)s.ci * C.NV + v;
            wall_dist[k] = correct * wall_dist[k];
        }
}

// adsorbingWallS_P2 (clAdsorb.c:656-816) / adsorbingWallD_P2 (:821-975) /
// adsorbingWallCL_P2 (:488-651); mode 0 specular, 1 diffuse, 2 Cercignani-Lampis
__global__ void k_adsorbingWall_P2(Consts C, const d2 *normal, int face, const d4 *wall_prop, const d4 *wall_cover,
                                   const d2 *wall_dist, d2 *flux_f, int mode) {
    long long w = wave_id();
    if (w >= wall_count(C.ni, C.nj, face)) return;
    WallSite s = wall_site(C, face, (int)w);
    int lane = lane_id();
    d2 n = normal[sidx(C, s.gi, s.gj, s.fid)];
    d4 wall = wall_prop[(size_t)s.ci * 4 + face];
    d2 wl = to_local(mk2(wall.s1, wall.s2), n);
    wall.s1 = wl.x; wall.s2 = wl.y;
    d4 cov = wall_cover[(size_t)s.ci * 4 + face];
    double reflected = cov.s1;
    double ratio = 1.0;
    if (reflected > 0.0) {
        double out = 0;
        for (int v = lane; v < C.NV; v += 64) {
            d2 uv = face_velocity(C, v, n);
            if (leaves_wall(uv.x, s.rot)) {
                d2 em;
                if (mode == 2) {
                    const d2 *wrow = wall_dist + (size_t)s.ci * C.NV;
                    em = cl_emission(C, face, n, s.rot, wall, uv, v, [&](int va) { return wrow[va]; });
                } else if (mode == 1) {
                    em = maxwell(C, wall, uv, v);
                } else {
                    em = wall_dist[(size_t)s.ci * C.NV + (s.fid == WEST ? C.mEW[v] : C.mNS[v])];
                }
                flux_f[fidx(C, s.gi, s.gj, v)] = em;
                out += s.rot * uv.x * em.x;
            }
        }
        ratio = reflected / wave_sum(out);
    }
    double des = cov.s3, ads = cov.s2;
    int account = ((des == 0.0) && (ads > 0.0) && (reflected == 0.0)) ? 1 : 0;
    wall.s0 = 2 * des * sqrt(kPi * wall.s3);
    const double keep = 1 - (1 - ratio);   // the reference's factor 1 - delta*(1 - ratio)
    for (int v = lane; v < C.NV; v += 64) {
        d2 uv = face_velocity(C, v, n);
        if (!leaves_wall(uv.x, s.rot)) continue;
        size_t fi = fidx(C, s.gi, s.gj, v);
        d2 fd = keep * flux_f[fi];
        fd = fd + maxwell(C, wall, uv, v);
        if (account) fd = fd - fd;
        flux_f[fi] = fd;
    }
}

// ---------------------------------------------------------------- update
// updateMacro (clUGKS.c:1158-1187)
__global__ void k_updateMacro(Consts C, const d4 *fmS, const d4 *fmW, const double *area, d4 *macro, d4 *residual) {
    long long t = thread_id();
    if (t >= (long long)C.ni * C.nj) return;
    int mi = (int)(t / C.nj), mj = (int)(t % C.nj), gi = mi + C.G, gj = mj + C.G;
    d4 old = macro[midx(C, mi, mj)];
    d4 wv = to_conserved(C, old);
    d4 d = ((fmS[cidx(C, gi, gj)] - fmS[cidx(C, gi, gj + 1)]) + fmW[cidx(C, gi, gj)]) - fmW[cidx(C, gi + 1, gj)];
    macro[midx(C, mi, mj)] = to_primary(C, wv + d / area[cidx(C, gi, gj)]);
    residual[midx(C, mi, mj)] = old;
}

// UGKS_update (clUGKS.c:1189-1257).  One wavefront per cell: the two
// relaxation times and the cell's state are formed once, then the wavefront
// strides over the velocities.
__global__ void __launch_bounds__(kBlock) k_UGKS_update(Consts C, d2 *__restrict__ F, const d2 *__restrict__ ffS,
                                                        const d2 *__restrict__ ffW, const double *__restrict__ area,
                                                        const d4 *__restrict__ macro, const d2 *__restrict__ gQ,
                                                        const d4 *__restrict__ residual, double dt, double *flag) {
    const long long cell = wave_id_u();
    if (cell >= (long long)C.ni * C.nj) return;
    const int mi = (int)(cell / C.nj), mj = (int)(cell % C.nj), gi = mi + C.G, gj = mj + C.G;
    const d4 pold = residual[midx(C, mi, mj)], p = macro[midx(C, mi, mj)];
    const double tau_old = relax_time(C, pold), tau = relax_time(C, p);
    const d2 Q = gQ[midx(C, mi, mj)];
    const double A = area[cidx(C, gi, gj)], den = 1.0 + 0.5 * dt / tau;
    const size_t k0 = fidx(C, gi, gj, 0), kN = fidx(C, gi, gj + 1, 0), kE = fidx(C, gi + 1, gj, 0);
    bool bad = false;
    for (int v = lane_id(); v < C.NV; v += 64) {
        const d2 e = C.quad[v];
        const d2 feq_old = shakhov_eq(C, pold, Q, e, v);
        const d2 feq = shakhov_eq(C, p, Q, e, v);
        const d2 fo = F[k0 + v];
        const d2 fl = (((ffS[k0 + v] - ffS[kN + v]) + ffW[k0 + v]) - ffW[kE + v]) / A;
        const d2 relax = feq / tau + (feq_old - fo) / tau_old;
        const d2 fn = ((fo + fl) + (0.5 * dt) * relax) / den;
        bad |= isnan(fn.x) || isnan(fn.y);
        stream_d2(F + k0 + v, fn);
    }
    if (bad) flag[0] = ERR_NAN;
}

// getResidual (clUGKS.c:1259-1277)
__global__ void k_getResidual(Consts C, const d4 *macro, d4 *residual) {
    long long t = thread_id();
    if (t >= (long long)C.ni * C.nj) return;
    d4 d = macro[t] - residual[t];
    residual[t] = mk4(d.s0 * d.s0, d.s1 * d.s1, d.s2 * d.s2, d.s3 * d.s3);
}

// edgeMassFlow (clUGKS.c:1282-1337)
__global__ void k_edgeMassFlow(Consts C, const d4 *macro, const d2 *normal, const double *side_length, double *mdot,
                               int this_face) {
    long long t = thread_id();
    if (t >= wall_count(C.ni, C.nj, this_face)) return;
    int ci = (int)t, mi, mj, gi, gj, face;
    double dir;
    switch (this_face) {
    case GNORTH: mi = ci; mj = C.nj - 1; gi = C.G + ci; gj = C.NJ - C.G; face = SOUTH; dir = 1; break;
    case GEAST: mi = C.ni - 1; mj = ci; gi = C.NI - C.G; gj = C.G + ci; face = WEST; dir = 1; break;
    case GSOUTH: mi = ci; mj = 0; gi = C.G + ci; gj = C.G; face = SOUTH; dir = -1; break;
    default: mi = 0; mj = ci; gi = C.G; gj = C.G + ci; face = WEST; dir = -1; break;
    }
    d2 fn = dir * normal[sidx(C, gi, gj, face)];
    d4 cm = macro[midx(C, mi, mj)];
    mdot[(size_t)ci * 4 + this_face] = (fn.x * cm.s1 + fn.y * cm.s2) * cm.s0 * side_length[sidx(C, gi, gj, face)];
}

// ---------------------------------------------------------------- ghosts
// ghost strip of one face: (ni x G) on N/S, (G x nj) on E/W (ugks_block.py:454-457)
__host__ __device__ inline void ghost_strip(int ni, int nj, int G, int face, int &si, int &sj) {
    if (face == GNORTH || face == GSOUTH) { si = ni; sj = G; } else { si = G; sj = nj; }
}
__device__ __forceinline__ void ghost_cell(const Consts &C, int face, int li, int lj, int &gi, int &gj) {
    switch (face) {
    case GNORTH: gi = li + C.G; gj = lj + C.NJ - C.G; break;
    case GEAST: gi = li + C.NI - C.G; gj = lj + C.G; break;
    case GSOUTH: gi = li + C.G; gj = lj; break;
    default: gi = li; gj = lj + C.G; break;
    }
}

enum GhostKind { GK_EXTRAPOLATE, GK_CONSTANT, GK_INFLOW, GK_OUTFLOW, GK_MIRROR, GK_BOUNCE, GK_CONSTGRAD };

// edgeExtrapolate / edgeConstant / edgeInflow / edgeOutflow / edgeMirror /
// edgeBounceBack / edgeConstGrad (clGhostRoutines.c:544-1048)
__global__ void k_ghost(Consts C, d2 *F, int face, int kind, const d2 *normal, const d4 *macro, const d2 *gQ,
                        const d4 *wall_prop) {
    int si, sj;
    ghost_strip(C.ni, C.nj, C.G, face, si, sj);
    long long w = wave_id();
    int nvc = (C.NV + 63) / 64;
    if (w >= (long long)si * sj * nvc) return;
    int item = (int)(w / nvc), v = (int)(w % nvc) * 64 + lane_id();
    if (v >= C.NV) return;
    int li = item / sj, lj = item % sj, gi, gj;
    ghost_cell(C, face, li, lj, gi, gj);
    bool ns = (face == GNORTH || face == GSOUTH);
    int ci = ns ? li : lj;
    // the interior cell next to the boundary, clGhostRoutines.c:558-587
    int ei, ej;
    switch (face) {
    case GNORTH: ei = gi; ej = C.NJ - C.G - 1; break;
    case GEAST: ei = C.NI - C.G - 1; ej = gj; break;
    case GSOUTH: ei = gi; ej = C.G; break;
    default: ei = C.G; ej = gj; break;
    }
    d2 out;
    switch (kind) {
    case GK_EXTRAPOLATE: out = F[fidx(C, ei, ej, v)]; break;
    case GK_CONSTANT: out = shakhov_eq(C, wall_prop[(size_t)ci * 4 + face], mk2(0.0, 0.0), C.quad[v], v); break;
    case GK_INFLOW:
    case GK_OUTFLOW: {
        int noi = (face == GNORTH || face == GEAST) ? gi : ei, noj = (face == GNORTH || face == GEAST) ? gj : ej;
        d4 prim = wall_prop[(size_t)ci * 4 + face];
        d2 Q = gQ[midx(C, ei - C.G, ej - C.G)];
        d4 gm = macro[midx(C, ei - C.G, ej - C.G)];
        d2 l = to_local(mk2(gm.s1, gm.s2), normal[sidx(C, noi, noj, ns ? SOUTH : WEST)]);
        gm.s1 = l.x; gm.s2 = l.y;
        if (kind == GK_INFLOW) { gm.s0 = prim.s0; gm.s3 = prim.s3; }
        else gm.s3 = gm.s0 / (prim.s0 / prim.s3);
        out = shakhov_eq(C, gm, Q, C.quad[v], v);
        break;
    }
    case GK_MIRROR:
    case GK_BOUNCE: {
        int mi, mj;
        switch (face) {
        case GNORTH: mi = gi; mj = 2 * (C.NJ - C.G) - gj - 1; break;
        case GEAST: mi = 2 * (C.NI - C.G) - gi - 1; mj = gj; break;
        case GSOUTH: mi = gi; mj = 2 * C.G - gj - 1; break;
        default: mi = 2 * C.G - gi - 1; mj = gj; break;
        }
        int64_t mv = kind == GK_BOUNCE ? C.mD[v] : (ns ? C.mNS[v] : C.mEW[v]);
        out = F[fidx(C, mi, mj, (int)mv)];
        break;
    }
    default: {   // GK_CONSTGRAD, clGhostRoutines.c:967-1048
        int ei0 = ei, ej0 = ej;
        double m;
        switch (face) {
        case GNORTH: ej0 = ej - 1; m = (double)(gj - ej); break;
        case GEAST: ei0 = ei - 1; m = (double)(gi - ei); break;
        case GSOUTH: ej0 = ej + 1; m = (double)(ej - gj); break;
        default: ei0 = ei + 1; m = (double)(ei - gi); break;
        }
        d2 f0 = F[fidx(C, ei0, ej0, v)], f1 = F[fidx(C, ei, ej, v)];
        out = f1 + m * (f1 - f0);
        break;
    }
    }
    F[fidx(C, gi, gj, v)] = out;
}

// edgeExchange (clGhostRoutines.c:407-538): copy the neighbour's interior strip
__device__ __forceinline__ void rot90(int &i, int &j, int n, int Ni, int Nj) {
    for (int a = 0; a < n; a++) {
        int t = i;
        i = Nj - j - 1;
        j = t;
        t = Ni; Ni = Nj; Nj = t;
    }
}
// getSwapIndex for cell indices (v = 0), clGhostRoutines.c:47-138
__device__ inline void swap_index(const Consts &C, int tf, int &iB, int &jB, int of, int NIB, int NJB) {
    const int NI = C.NI, NJ = C.NJ, G = C.G;
    if (tf == GEAST) {
        if (of == GEAST) { iB = (NI - 1) - iB; iB += NIB - 2 * G; }
        else if (of == GWEST) iB -= NI - 2 * G;
        else if (of == GNORTH) { rot90(iB, jB, 3, NI, NJ); jB += NJB - 2 * G; }
        else { rot90(iB, jB, 1, NI, NJ); jB -= NI - 2 * G; }
    } else if (tf == GWEST) {
        if (of == GWEST) { iB = (NI - 1) - iB; iB -= NI - 2 * G; }
        else if (of == GEAST) iB += NIB - 2 * G;
        else if (of == GNORTH) { rot90(iB, jB, 1, NI, NJ); jB += NJB - 2 * G; }
        else { rot90(iB, jB, 3, NI, NJ); jB -= NI - 2 * G; }
    } else if (tf == GNORTH) {
        if (of == GNORTH) { jB = (NJ - 1) - jB; jB += NJB - 2 * G; }
        else if (of == GSOUTH) jB -= NJ - 2 * G;
        else if (of == GEAST) { rot90(iB, jB, 1, NI, NJ); iB += NIB - 2 * G; }
        else { rot90(iB, jB, 3, NI, NJ); iB -= NJ - 2 * G; }
    } else {
        if (of == GSOUTH) { jB = (NJ - 1) - jB; jB -= NJ - 2 * G; }
        else if (of == GNORTH) jB += NJB - 2 * G;
        else if (of == GEAST) { rot90(iB, jB, 3, NI, NJ); iB += NIB - 2 * G; }
        else { rot90(iB, jB, 1, NI, NJ); iB -= NJ - 2 * G; }
    }
}

__global__ void k_edgeExchange(Consts C, d2 *fA, int this_face, const d2 *fB, int NIB, int NJB, int that_face,
                               int flipidx, int flip) {
    int si, sj;
    ghost_strip(C.ni, C.nj, C.G, this_face, si, sj);
    long long w = wave_id();
    int nvc = (C.NV + 63) / 64;
    if (w >= (long long)si * sj * nvc) return;
    int item = (int)(w / nvc), v = (int)(w % nvc) * 64 + lane_id();
    if (v >= C.NV) return;
    int gi, gj;
    ghost_cell(C, this_face, item / sj, item % sj, gi, gj);
    int iB = gi, jB = gj;
    if (flipidx) {   // misaligned corner vertices with ori set, clGhostRoutines.c:505-512
        if (this_face == GEAST || this_face == GWEST) jB = (C.NJ - 1) - jB;
        else iB = (C.NI - 1) - iB;
    }
    swap_index(C, this_face, iB, jB, that_face, NIB, NJB);
    if (iB < 0 || iB >= NIB || jB < 0 || jB >= NJB) return;   // incompatible block pair: leave the ghost
    int64_t v2_ = v;
    if (flip == FLIP_NS) v2_ = C.mNS[v];
    else if (flip == FLIP_D) v2_ = C.mD[v];
    else if (flip == FLIP_EW) v2_ = C.mEW[v];
    fA[fidx(C, gi, gj, v)] = fB[((size_t)iB * NJB + (size_t)jB) * (size_t)C.NV + (size_t)v2_];
}

}  // namespace

// ============================================================== C ABI ====
#define UGKS_BLK_CHECK(b)                              \
    do {                                               \
        if (!(b)) return ugks_rt::fail(-1, "null block"); \
    } while (0)
#define UGKS_FACE_CHECK(f)                                                                  \
    do {                                                                                    \
        if ((f) != SOUTH && (f) != WEST) return ugks_rt::fail(-1, "face must be 0 (SOUTH) or 1 (WEST)"); \
    } while (0)
#define UGKS_WALL_CHECK(f)                                                              \
    do {                                                                                \
        if ((f) < 0 || (f) > 3) return ugks_rt::fail(-1, "wall face must be 0..3 (N,E,S,W)"); \
    } while (0)
#define UGKS_OFFS_CHECK(ob, ot)                                                             \
    do {                                                                                    \
        if ((ob) < 0 || (ob) > 1 || (ot) < 0 || (ot) > 1) return ugks_rt::fail(-1, "offsets must be 0 or 1"); \
    } while (0)

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

extern "C" {

int ugks_cellGeom(ugks_block *blk, void *stream, const void *xy, void *area, void *centre, void *mid_side, void *normal,
                  void *side_length) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    long long n = (long long)C.NI * C.NJ;
    hipLaunchKernelGGL(k_cellGeom, dim3(blocks_for_threads(n)), dim3(kBlock), 0, S(stream), C, (const d2 *)xy,
                       (double *)area, (d2 *)centre, (d2 *)mid_side, (d2 *)normal, (double *)side_length);
    return ugks_rt::check_launch("cellGeom");
}

int ugks_paraBC(ugks_block *blk, void *stream, const void *para_def, void *wall_prop, double t) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    if (C.any_udf && !para_def) return ugks_rt::fail(-1, "paraBC: wall-value expressions need para_def");
    long long n = C.ni > C.nj ? C.ni : C.nj;
    hipLaunchKernelGGL(k_paraBC, dim3(blocks_for_threads(n)), dim3(kBlock), 0, S(stream), C, (const d4 *)para_def,
                       (d4 *)wall_prop, t);
    return ugks_rt::check_launch("paraBC");
}

int ugks_clFindDT(ugks_block *blk, void *stream, const void *xy, const void *area, const void *macro, void *time_step,
                  double Kn) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    long long n = (long long)C.ni * C.nj;
    hipLaunchKernelGGL(k_clFindDT, dim3(blocks_for_threads(n)), dim3(kBlock), 0, S(stream), C, (const d2 *)xy,
                       (const double *)area, (const d4 *)macro, (double *)time_step, Kn);
    return ugks_rt::check_launch("clFindDT");
}

int ugks_time_step_max(ugks_block *blk, void *stream, const void *time_step, void *out) {
    UGKS_BLK_CHECK(blk);
    hipLaunchKernelGGL(k_max, dim3(1), dim3(1024), 0, S(stream), (const double *)time_step,
                       (long long)blk->c.ni * blk->c.nj, (double *)out);
    return ugks_rt::check_launch("time_step_max");
}

int ugks_initFunctions(ugks_block *blk, void *stream, void *Fin, const void *macro, const void *gQ) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    long long nw = (long long)C.ni * C.nj * nvchunks(C);
    hipLaunchKernelGGL(k_initFunctions, dim3(blocks_for_waves(nw)), dim3(kBlock), 0, S(stream), C, (d2 *)Fin,
                       (const d4 *)macro, (const d2 *)gQ);
    return ugks_rt::check_launch("initFunctions");
}

int ugks_calcMacro(ugks_block *blk, void *stream, const void *Fin, void *macro) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_calcMacro, dim3(blocks_for_waves((long long)C.ni * C.nj)), dim3(kBlock), 0, S(stream), C,
                       (const d2 *)Fin, (d4 *)macro);
    return ugks_rt::check_launch("calcMacro");
}

int ugks_calcQ(ugks_block *blk, void *stream, const void *Fin, const void *macro, void *gQ) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_calcQ, dim3(blocks_for_waves((long long)C.ni * C.nj)), dim3(kBlock), 0, S(stream), C,
                       (const d2 *)Fin, (const d4 *)macro, (d2 *)gQ);
    return ugks_rt::check_launch("calcQ");
}

int ugks_getInternal(ugks_block *blk, void *stream, const void *Fin, void *Txyz, void *Stress) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_getInternal, dim3(blocks_for_waves((long long)C.ni * C.nj)), dim3(kBlock), 0, S(stream), C,
                       (const d2 *)Fin, (d4 *)Txyz, (d4 *)Stress);
    return ugks_rt::check_launch("getInternal");
}

int ugks_iFace(ugks_block *blk, void *stream, const void *Fin, void *iface_f, void *fsigma, const void *centre,
               const void *mid_side, const void *normal, int32_t face) {
    UGKS_BLK_CHECK(blk);
    UGKS_FACE_CHECK(face);
    const Consts &C = blk->c;
    long long nf = face == SOUTH ? (long long)C.ni * (C.nj + 1) : (long long)(C.ni + 1) * C.nj;
    hipLaunchKernelGGL(k_iFace, dim3(blocks_for_waves(nf)), dim3(kBlock), 0, S(stream), C,
                       (const d2 *)Fin, (d2 *)iface_f, (d2 *)fsigma, (const d2 *)centre, (const d2 *)mid_side,
                       (const d2 *)normal, (int)face);
    return ugks_rt::check_launch("iFace");
}

static int face_moments(ugks_block *blk, void *stream, const void *src, const void *normal, int32_t face, void *out,
                        int32_t ob, int32_t ot, int which, const char *name) {
    UGKS_BLK_CHECK(blk);
    UGKS_FACE_CHECK(face);
    UGKS_OFFS_CHECK(ob, ot);
    const Consts &C = blk->c;
    FaceRange r = face_range(C.ni, C.nj, face, ob, ot);
    long long n = (long long)r.ni * r.nj;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_faceMoments, dim3(blocks_for_waves(n)), dim3(kBlock), 0, S(stream), C, (const d2 *)src,
                       (const d2 *)normal, (int)face, (d4 *)out, (int)ob, (int)ot, which);
    return ugks_rt::check_launch(name);
}

int ugks_getFaceCons(ugks_block *blk, void *stream, const void *iface_f, const void *normal, int32_t face,
                     void *primary, int32_t offset_bottom, int32_t offset_top) {
    return face_moments(blk, stream, iface_f, normal, face, primary, offset_bottom, offset_top, 0, "getFaceCons");
}
int ugks_getAL(ugks_block *blk, void *stream, const void *Fin, const void *normal, int32_t face, void *gaL,
               int32_t offset_bottom, int32_t offset_top) {
    return face_moments(blk, stream, Fin, normal, face, gaL, offset_bottom, offset_top, 1, "getAL");
}
int ugks_getAR(ugks_block *blk, void *stream, const void *Fin, const void *normal, int32_t face, void *gaR,
               int32_t offset_bottom, int32_t offset_top) {
    return face_moments(blk, stream, Fin, normal, face, gaR, offset_bottom, offset_top, 2, "getAR");
}

int ugks_getPLR(ugks_block *blk, void *stream, const void *centre, const void *mid_side, int32_t face, void *primary,
                void *gaL, void *gaR, int32_t offset_bottom, int32_t offset_top) {
    UGKS_BLK_CHECK(blk);
    UGKS_FACE_CHECK(face);
    UGKS_OFFS_CHECK(offset_bottom, offset_top);
    const Consts &C = blk->c;
    FaceRange r = face_range(C.ni, C.nj, face, offset_bottom, offset_top);
    long long n = (long long)r.ni * r.nj;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_getPLR, dim3(blocks_for_threads(n)), dim3(kBlock), 0, S(stream), C, (const d2 *)centre,
                       (const d2 *)mid_side, (int)face, (d4 *)primary, (d4 *)gaL, (d4 *)gaR, (int)offset_bottom,
                       (int)offset_top);
    return ugks_rt::check_launch("getPLR");
}

int ugks_initMacroFlux(ugks_block *blk, void *stream, void *flux_macro, int32_t face, double dt, const void *primary,
                       const void *gaL, const void *gaR, void *gaT, void *gMxi, int32_t offset_bottom,
                       int32_t offset_top) {
    UGKS_BLK_CHECK(blk);
    UGKS_FACE_CHECK(face);
    UGKS_OFFS_CHECK(offset_bottom, offset_top);
    const Consts &C = blk->c;
    FaceRange r = face_range(C.ni, C.nj, face, offset_bottom, offset_top);
    long long n = (long long)r.ni * r.nj;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_initMacroFlux, dim3(blocks_for_threads(n)), dim3(kBlock), 0, S(stream), C, (d4 *)flux_macro,
                       (int)face, dt, (const d4 *)primary, (const d4 *)gaL, (const d4 *)gaR, (d4 *)gaT,
                       (double *)gMxi, (int)offset_bottom, (int)offset_top);
    return ugks_rt::check_launch("initMacroFlux");
}

int ugks_calcFaceQ(ugks_block *blk, void *stream, const void *iface_f, const void *primary, const void *normal,
                   void *faceQ, int32_t face, int32_t offset_bottom, int32_t offset_top) {
    UGKS_BLK_CHECK(blk);
    UGKS_FACE_CHECK(face);
    UGKS_OFFS_CHECK(offset_bottom, offset_top);
    const Consts &C = blk->c;
    FaceRange r = face_range(C.ni, C.nj, face, offset_bottom, offset_top);
    long long n = (long long)r.ni * r.nj;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_calcFaceQ, dim3(blocks_for_waves(n)), dim3(kBlock), 0, S(stream), C, (const d2 *)iface_f,
                       (const d4 *)primary, (const d2 *)normal, (d2 *)faceQ, (int)face, (int)offset_bottom,
                       (int)offset_top);
    return ugks_rt::check_launch("calcFaceQ");
}

int ugks_macroFlux(ugks_block *blk, void *stream, const void *flux_f, const void *fsigma, void *flux_macro,
                   const void *normal, const void *side_length, int32_t face, double dt, const void *primary,
                   const void *faceQ, int32_t offset_bottom, int32_t offset_top) {
    UGKS_BLK_CHECK(blk);
    UGKS_FACE_CHECK(face);
    UGKS_OFFS_CHECK(offset_bottom, offset_top);
    const Consts &C = blk->c;
    FaceRange r = face_range(C.ni, C.nj, face, offset_bottom, offset_top);
    long long n = (long long)r.ni * r.nj;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_macroFlux, dim3(blocks_for_waves(n)), dim3(kBlock), 0, S(stream), C, (const d2 *)flux_f,
                       (const d2 *)fsigma, (d4 *)flux_macro, (const d2 *)normal, (const double *)side_length, (int)face,
                       dt, (const d4 *)primary, (const d2 *)faceQ, (int)offset_bottom, (int)offset_top);
    return ugks_rt::check_launch("macroFlux");
}

int ugks_distFlux(ugks_block *blk, void *stream, void *flux_f, const void *fsigma, const void *normal,
                  const void *side_length, int32_t face, double dt, const void *primary, const void *gaL,
                  const void *gaR, const void *gaT, const void *gMxi, const void *faceQ, int32_t offset_bottom,
                  int32_t offset_top) {
    UGKS_BLK_CHECK(blk);
    UGKS_FACE_CHECK(face);
    UGKS_OFFS_CHECK(offset_bottom, offset_top);
    const Consts &C = blk->c;
    FaceRange r = face_range(C.ni, C.nj, face, offset_bottom, offset_top);
    long long n = (long long)r.ni * r.nj;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_distFlux, dim3(blocks_for_waves(n)), dim3(kBlock), 0, S(stream), C,
                       (d2 *)flux_f, (const d2 *)fsigma, (const d2 *)normal, (const double *)side_length, (int)face,
                       dt, (const d4 *)primary, (const d4 *)gaL, (const d4 *)gaR, (const d4 *)gaT,
                       (const double *)gMxi, (const d2 *)faceQ, (int)offset_bottom, (int)offset_top);
    return ugks_rt::check_launch("distFlux");
}

int ugks_accommodatingWallDist(ugks_block *blk, void *stream, const void *normal, int32_t face, const void *wall_prop,
                               void *flux_f, double dt) {
    (void)dt;
    UGKS_BLK_CHECK(blk);
    UGKS_WALL_CHECK(face);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_accommodatingWallDist, dim3((unsigned)wall_count(C.ni, C.nj, face)), dim3(kBlock), 0,
                       S(stream), C, (const d2 *)normal, (int)face, (const d4 *)wall_prop, (d2 *)flux_f);
    return ugks_rt::check_launch("accommodatingWallDist");
}

int ugks_wallFlux(ugks_block *blk, void *stream, const void *normal, const void *side_length, int32_t face,
                  void *flux_f, void *flux_macro, double dt) {
    UGKS_BLK_CHECK(blk);
    UGKS_WALL_CHECK(face);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_wallFlux, dim3((unsigned)wall_count(C.ni, C.nj, face)), dim3(kBlock), 0, S(stream), C,
                       (const d2 *)normal, (const double *)side_length, (int)face, (d2 *)flux_f, (d4 *)flux_macro, dt);
    return ugks_rt::check_launch("wallFlux");
}

int ugks_wallMassEnergyFluxes(ugks_block *blk, void *stream, const void *normal, const void *side_length,
                              const void *area, int32_t face, const void *flux_f, void *wall_flux, double dt) {
    (void)side_length; (void)area; (void)dt;
    UGKS_BLK_CHECK(blk);
    UGKS_WALL_CHECK(face);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_wallMassEnergyFluxes, dim3((unsigned)wall_count(C.ni, C.nj, face)), dim3(kBlock), 0,
                       S(stream), C, (const d2 *)normal, (int)face, (const d2 *)flux_f, (double *)wall_flux);
    return ugks_rt::check_launch("wallMassEnergyFluxes");
}

int ugks_adsorbingWall_P1(ugks_block *blk, void *stream, const void *normal, int32_t face, const void *wall_prop,
                          void *wall_cover, void *wall_dist, const void *flux_f, const void *macro, double dt,
                          void *flag) {
    (void)wall_prop;
    UGKS_BLK_CHECK(blk);
    UGKS_WALL_CHECK(face);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_adsorbingWall_P1, dim3(blocks_for_waves(wall_count(C.ni, C.nj, face))), dim3(kBlock), 0,
                       S(stream), C, (const d2 *)normal, (int)face, (d4 *)wall_cover, (d2 *)wall_dist,
                       (const d2 *)flux_f, (const d4 *)macro, dt, (double *)flag);
    return ugks_rt::check_launch("adsorbingWall_P1");
}

static int adsorb_p2(ugks_block *blk, void *stream, const void *normal, int32_t face, const void *wall_prop,
                     const void *wall_cover, const void *wall_dist, void *flux_f, int mode, const char *name) {
    UGKS_BLK_CHECK(blk);
    UGKS_WALL_CHECK(face);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_adsorbingWall_P2, dim3(blocks_for_waves(wall_count(C.ni, C.nj, face))), dim3(kBlock), 0,
                       S(stream), C, (const d2 *)normal, (int)face, (const d4 *)wall_prop, (const d4 *)wall_cover,
                       (const d2 *)wall_dist, (d2 *)flux_f, mode);
    return ugks_rt::check_launch(name);
}
int ugks_adsorbingWallS_P2(ugks_block *blk, void *stream, const void *normal, int32_t face, const void *wall_prop,
                           const void *wall_cover, const void *wall_dist, void *flux_f, const void *macro, double dt) {
    (void)macro; (void)dt;
    return adsorb_p2(blk, stream, normal, face, wall_prop, wall_cover, wall_dist, flux_f, 0, "adsorbingWallS_P2");
}
int ugks_adsorbingWallD_P2(ugks_block *blk, void *stream, const void *normal, int32_t face, const void *wall_prop,
                           const void *wall_cover, const void *wall_dist, void *flux_f, const void *macro, double dt) {
    (void)macro; (void)dt;
    return adsorb_p2(blk, stream, normal, face, wall_prop, wall_cover, wall_dist, flux_f, 1, "adsorbingWallD_P2");
}
int ugks_adsorbingWallCL_P2(ugks_block *blk, void *stream, const void *normal, int32_t face, const void *wall_prop,
                            const void *wall_cover, const void *wall_dist, void *flux_f, const void *macro, double dt) {
    (void)macro; (void)dt;
    return adsorb_p2(blk, stream, normal, face, wall_prop, wall_cover, wall_dist, flux_f, 2, "adsorbingWallCL_P2");
}

int ugks_updateMacro(ugks_block *blk, void *stream, const void *flux_macro_S, const void *flux_macro_W,
                     const void *area, void *macro, void *residual) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_updateMacro, dim3(blocks_for_threads((long long)C.ni * C.nj)), dim3(kBlock), 0, S(stream), C,
                       (const d4 *)flux_macro_S, (const d4 *)flux_macro_W, (const double *)area, (d4 *)macro,
                       (d4 *)residual);
    return ugks_rt::check_launch("updateMacro");
}

int ugks_UGKS_update(ugks_block *blk, void *stream, void *Fin, const void *flux_f_S, const void *flux_f_W,
                     const void *area, const void *macro, const void *gQ, const void *residual, double dt, void *flag) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_UGKS_update, dim3(blocks_for_waves((long long)C.ni * C.nj)), dim3(kBlock), 0, S(stream), C, (d2 *)Fin,
                       (const d2 *)flux_f_S, (const d2 *)flux_f_W, (const double *)area, (const d4 *)macro,
                       (const d2 *)gQ, (const d4 *)residual, dt, (double *)flag);
    return ugks_rt::check_launch("UGKS_update");
}

int ugks_getResidual(ugks_block *blk, void *stream, const void *macro, void *residual) {
    UGKS_BLK_CHECK(blk);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_getResidual, dim3(blocks_for_threads((long long)C.ni * C.nj)), dim3(kBlock), 0, S(stream), C,
                       (const d4 *)macro, (d4 *)residual);
    return ugks_rt::check_launch("getResidual");
}

int ugks_edgeMassFlow(ugks_block *blk, void *stream, const void *macro, const void *normal, const void *side_length,
                      void *mdot, int32_t this_face) {
    UGKS_BLK_CHECK(blk);
    UGKS_WALL_CHECK(this_face);
    const Consts &C = blk->c;
    hipLaunchKernelGGL(k_edgeMassFlow, dim3(blocks_for_threads(wall_count(C.ni, C.nj, this_face))), dim3(kBlock), 0,
                       S(stream), C, (const d4 *)macro, (const d2 *)normal, (const double *)side_length, (double *)mdot,
                       (int)this_face);
    return ugks_rt::check_launch("edgeMassFlow");
}

static int ghost_launch(ugks_block *blk, void *stream, void *Fin, int32_t face, int kind, const void *normal,
                        const void *macro, const void *gQ, const void *wall_prop, const char *name) {
    UGKS_BLK_CHECK(blk);
    UGKS_WALL_CHECK(face);
    const Consts &C = blk->c;
    int si, sj;
    ghost_strip(C.ni, C.nj, C.G, face, si, sj);
    long long nw = (long long)si * sj * nvchunks(C);
    hipLaunchKernelGGL(k_ghost, dim3(blocks_for_waves(nw)), dim3(kBlock), 0, S(stream), C, (d2 *)Fin, (int)face, kind,
                       (const d2 *)normal, (const d4 *)macro, (const d2 *)gQ, (const d4 *)wall_prop);
    return ugks_rt::check_launch(name);
}

int ugks_edgeExtrapolate(ugks_block *blk, void *stream, void *Fin, int32_t this_face) {
    return ghost_launch(blk, stream, Fin, this_face, GK_EXTRAPOLATE, nullptr, nullptr, nullptr, nullptr, "edgeExtrapolate");
}
int ugks_edgeConstant(ugks_block *blk, void *stream, void *Fin, int32_t this_face, const void *wall_prop) {
    return ghost_launch(blk, stream, Fin, this_face, GK_CONSTANT, nullptr, nullptr, nullptr, wall_prop, "edgeConstant");
}
int ugks_edgeInflow(ugks_block *blk, void *stream, void *Fin, int32_t this_face, const void *normal, const void *macro,
                    const void *gQ, const void *wall_prop) {
    return ghost_launch(blk, stream, Fin, this_face, GK_INFLOW, normal, macro, gQ, wall_prop, "edgeInflow");
}
int ugks_edgeOutflow(ugks_block *blk, void *stream, void *Fin, int32_t this_face, const void *normal, const void *macro,
                     const void *gQ, const void *wall_prop) {
    return ghost_launch(blk, stream, Fin, this_face, GK_OUTFLOW, normal, macro, gQ, wall_prop, "edgeOutflow");
}
int ugks_edgeMirror(ugks_block *blk, void *stream, void *Fin, int32_t this_face) {
    return ghost_launch(blk, stream, Fin, this_face, GK_MIRROR, nullptr, nullptr, nullptr, nullptr, "edgeMirror");
}
int ugks_edgeBounceBack(ugks_block *blk, void *stream, void *Fin, int32_t this_face) {
    return ghost_launch(blk, stream, Fin, this_face, GK_BOUNCE, nullptr, nullptr, nullptr, nullptr, "edgeBounceBack");
}
int ugks_edgeConstGrad(ugks_block *blk, void *stream, void *Fin, int32_t this_face) {
    return ghost_launch(blk, stream, Fin, this_face, GK_CONSTGRAD, nullptr, nullptr, nullptr, nullptr, "edgeConstGrad");
}

int ugks_edgeExchange(ugks_block *blk, void *stream, void *fA, const void *xyA, int32_t this_face, const void *fB,
                      const void *xyB, int32_t NIB, int32_t NJB, int32_t that_face, int32_t ori, int32_t flip) {
    UGKS_BLK_CHECK(blk);
    UGKS_WALL_CHECK(this_face);
    UGKS_WALL_CHECK(that_face);
    const Consts &C = blk->c;
    if (NIB < 2 * C.G + 1 || NJB < 2 * C.G + 1) return ugks_rt::fail(-1, "edgeExchange: bad NIB/NJB");
    if (flip < 0 || flip > 3) return ugks_rt::fail(-1, "edgeExchange: bad flip");
    bool ew = (this_face == GEAST || this_face == GWEST), ew_o = (that_face == GEAST || that_face == GWEST);
    if (ew && ew_o && NJB != C.NJ) return ugks_rt::fail(-1, "edgeExchange: E/W join needs equal NJ");
    if (!ew && !ew_o && NIB != C.NI) return ugks_rt::fail(-1, "edgeExchange: N/S join needs equal NI");
    // face-corner vertex check (clGhostRoutines.c:491-499), evaluated once per call
    const int G = C.G, NI = C.NI, NJ = C.NJ;
    int A[4], B[4];
    switch (this_face) {
    case GNORTH: A[0] = G; A[1] = NJ - G; A[2] = NI - G; A[3] = NJ - G; break;
    case GEAST: A[0] = NI - G; A[1] = G; A[2] = NI - G; A[3] = NJ - G; break;
    case GSOUTH: A[0] = G; A[1] = G; A[2] = NI - G; A[3] = G; break;
    default: A[0] = G; A[1] = G; A[2] = G; A[3] = NJ - G; break;
    }
    switch (that_face) {
    case GNORTH: B[0] = G; B[1] = NJB - G; B[2] = NIB - G; B[3] = NJB - G; break;
    case GEAST: B[0] = NIB - G; B[1] = G; B[2] = NIB - G; B[3] = NJB - G; break;
    case GSOUTH: B[0] = G; B[1] = G; B[2] = NIB - G; B[3] = G; break;
    default: B[0] = G; B[1] = G; B[2] = G; B[3] = NJB - G; break;
    }
    int flipidx = 0;
    if (ori) {
        d2 va[2], vb[2];
        hipStream_t st = S(stream);
        const d2 *xa = (const d2 *)xyA, *xb = (const d2 *)xyB;
        int e = 0;
        e |= ugks_rt::copy_small(&va[0], xa + (size_t)A[0] * (NJ + 1) + A[1], sizeof(d2), st);
        e |= ugks_rt::copy_small(&va[1], xa + (size_t)A[2] * (NJ + 1) + A[3], sizeof(d2), st);
        e |= ugks_rt::copy_small(&vb[0], xb + (size_t)B[0] * (NJB + 1) + B[1], sizeof(d2), st);
        e |= ugks_rt::copy_small(&vb[1], xb + (size_t)B[2] * (NJB + 1) + B[3], sizeof(d2), st);
        if (e) return ugks_rt::fail(-1, "edgeExchange: vertex read failed");
        auto ne = [](double a, double b) { return !(fabs(a - b) <= 1e-6); };
        flipidx = ne(va[0].x, vb[0].x) || ne(va[0].y, vb[0].y) || ne(va[1].x, vb[1].x) || ne(va[1].y, vb[1].y);
    }
    int si, sj;
    ghost_strip(C.ni, C.nj, C.G, this_face, si, sj);
    long long nw = (long long)si * sj * nvchunks(C);
    hipLaunchKernelGGL(k_edgeExchange, dim3(blocks_for_waves(nw)), dim3(kBlock), 0, S(stream), C, (d2 *)fA,
                       (int)this_face, (const d2 *)fB, (int)NIB, (int)NJB, (int)that_face, flipidx, (int)flip);
    return ugks_rt::check_launch("edgeExchange");
}

}  // extern "C"
