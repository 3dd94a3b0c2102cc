// ugks_fused.hip -- the batched, fused UGKS time step for gfx950.
//
// One step of every block of every case in eight launches, with no host
// round trip (the reference issues ~30 launches per block per step plus two
// blocking readbacks, ugks_block.py:1049-1329, :1415-1433):
//
//   k_faceA   one wavefront per interior face (S and W faces of all blocks):
//             reconstruction (iFace) and the raw velocity moments of the face
//             values (getFaceCons and the sums calcFaceQ needs), of the left
//             and right cells (getAL/getAR) and of the owner cell (calcQ, on
//             S faces).
//   k_faceS   one thread per interior face: primary variables, slopes and
//             time slope (getPLR/initMacroFlux), face heat flux (calcFaceQ)
//             and the equilibrium part of the macro flux, folded into the
//             per-face constants of the flux pass.
//   k_faceC   one wavefront per interior face: reconstruction again, macro
//             flux (macroFlux) and the per-velocity flux (distFlux).
//   k_wall    one workgroup per diffuse/adsorbing wall face: reconstruction,
//             accommodatingWallDist or adsorbingWall_P1 + S/D_P2, wallFlux,
//             wallMassEnergyFluxes.
//   k_update  one wavefront per cell: updateMacro + UGKS_update and the
//             cell's moments for the next step's left/right face states.
//   k_dt      one thread per cell: clFindDT and the per-block maximum.
//   k_dtfin   one workgroup: the global time step with the reference's growth
//             limit (ugks_sim.py:285-311, applied as one_step does, :430-446).
//   k_ghost   one wavefront per ghost cell: every BC of every block
//             (updateBC, ugks_block.py:621-665).
//
// Items of all blocks are laid out back to back; a wavefront finds its block
// by bisection of a prefix table.  f is stored as T2 = (g, h) pairs of T
// (double, or float for the mixed-precision step); moment sums, the
// macroscopic update, heat fluxes, equilibrium moments and the time step are
// FP64 in both.  Lane l of a wavefront owns velocities v = l + 64 k (thread t
// owns t + 256 k in the workgroup-per-face wall pass): every wavefront load or
// store touches 64 consecutive velocities of one cell.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "../../include/ugks_hip.h"
#include "ugks_device.hpp"
#include "ugks_runtime.hpp"

using namespace ugks;

namespace {

constexpr int kT = 256;           // threads per workgroup
constexpr int kWavesT = kT / 64;

// wavefront-per-face kernels come in workgroups of FW = 1, 2 or 4 wavefronts
// (faces); the plan's autotuner picks one per kernel (ugks_step_autotune)

template <typename T> struct __align__(2 * sizeof(T)) P2 { T x, y; };

// device record of one block for the fused kernels
struct FBlock {
    Consts c;
    void *f, *fluxS, *fluxW, *wdist;       // P2<T> arrays ([NI][NJ][NV]; wdist [4][wl][NV])
    d4 *fmS, *fmW, *macro, *macro_new, *residual, *wall_cover;
    d2 *Q;
    const d2 *xy, *centre, *mid_side, *normal;
    const double *area, *side_length;
    const d4 *wall_prop;
    double *wall_fluxes, *time_step, *flag;
    const d4 *para;                         // [wl][4] (s, x, y, 0) of each wall cell
    int bc[4], other[4], oface[4], ori[4], flip[4], reflect[4];
    int nS, nSj, ob_S, ob_W, nWi, wl;       // interior-face counts and offsets
    int flipidx[4];                         // exchange row flip from the corner check
    int cs;                                 // case of this block (its time-step state)
    d4 *cmom;                               // [NI][NJ] moments of f per cell (all cells, ghosts included)
};

// A case's time-step state.  dt is gdata.dt as the host sees it after a step;
// dt_flux and dt_up are the time steps the next step's flux and update use.
// They differ from dt when the growth limit binds as the step re-evaluates
// dt: the reference computes the flux with dt, re-evaluates, and if dt moved
// computes the flux again with the new value and re-evaluates once more
// before the update (ugks_sim.py:430-446).
struct DtState {
    double dt, dt_flux, dt_up, time, CFL, delta_dt, suggest_dt, Kn;
    int step, overshoot, force_time, pad;
};
// updateTimeStep (ugks_sim.py:285-311) towards the block minimum M
__host__ __device__ inline double limited_dt(double cur, double M, double lim) {
    return cur != M ? (M > lim * cur ? lim * cur : M) : cur;
}

// Workgroups are dealt to the 8 XCDs round robin (workgroup id % 8), and
// each XCD has its own L2.  Neighbouring faces and cells share f stencils and
// fluxes, so runs of UGKS_XCD_GROUP consecutive work items go to one XCD, and
// the runs are dealt to the XCDs in turn: every XCD takes an equal share of
// every region of the work, however the work per item varies over a batch.
// (One contiguous eighth of the items per XCD keeps more neighbours in one
// L2, but on the 64-case batch, whose cases have 289 to 1681 velocities, the
// busiest XCD then had 1.48 times the mean work and the step took 35% longer.)
// The map is a bijection on [0, n) for any n; a tail shorter than eight runs
// keeps the hardware order.
#ifndef UGKS_XCD_GROUP
#define UGKS_XCD_GROUP 32
#endif
__device__ __forceinline__ long long xcd_remap(long long id, long long n) {
    constexpr long long G = UGKS_XCD_GROUP;
    const long long full = n / (8 * G) * (8 * G);
    if (id >= full) return id;
    const long long x = id & 7, k = id >> 3;
    return ((k / G) * 8 + x) * G + (k % G);
}

// the last block whose first item is at or before item (blocks without items
// are passed over), by bisection: a batch of many blocks costs each wavefront
// log2(nb) dependent loads, not nb
__device__ __forceinline__ int find_block(const long long *off, int nb, long long item) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// the same over a list of items ordered by their start
template <typename Item> __device__ __forceinline__ int find_item(const Item *items, int nitems, long long item) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].start <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T> __device__ __forceinline__ P2<T> ld(const void *p, size_t k) { return ((const P2<T> *)p)[k]; }
template <typename T> __device__ __forceinline__ void st(void *p, size_t k, T x, T y) { ((P2<T> *)p)[k] = P2<T>{x, y}; }

// ---- block-wide sum of N doubles (every thread receives the totals) --------
template <int N> __device__ __forceinline__ void block_sum(double (&a)[N], double (*red)[16]) {
#pragma unroll
    for (int q = 0; q < N; q++) a[q] = wave_sum(a[q]);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int q = 0; q < N; q++) red[w][q] = a[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; q++) a[q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
    __syncthreads();
}

// ---- reconstruction (iFace, clUGKS.c:134-195) in precision T ---------------
__device__ __forceinline__ float rcp_T(float x) { return __builtin_amdgcn_rcpf(x); }

// WENO5 slopes of a (g, h) stencil: FP64 in weno5_slope_pair (ugks_device.hpp)
__device__ __forceinline__ void weno5_pair(P2<double> s0, P2<double> s1, P2<double> s2, P2<double> s3, P2<double> s4,
                                           double rdx, double &sgx, double &sgy) {
    weno5_slope_pair(s0.x, s1.x, s2.x, s3.x, s4.x, s0.y, s1.y, s2.y, s3.y, s4.y, rdx, sgx, sgy);
}
__device__ __forceinline__ float weno5_f(float s0, float s1, float s2, float s3, float s4) {
    const float d1 = s1 - s0, d2 = s2 - s1, d3 = s3 - s2, d4 = s4 - s3;
    const float u0 = d2 - d1, v0 = 3.f * d2 - d1;
    const float u1 = d3 - d2, v1 = d2 + d3;
    const float u2 = d4 - d3, v2 = 3.f * d3 - d4;
    const float e0 = fmaf((float)kWA0 * u0, u0, fmaf((float)kWB0 * v0, v0, (float)kWE0));
    const float e1 = fmaf((float)kWA1 * u1, u1, fmaf((float)kWB1 * v1, v1, (float)kWE1));
    const float e2 = fmaf((float)kWA2 * u2, u2, fmaf((float)kWB2 * v2, v2, (float)kWE2));
    const float a0 = rcp_T(e0 * e0), a1 = rcp_T(e1 * e1), a2 = rcp_T(e2 * e2);
    const float c0 = (-1.f / 3.f) * d1 + (5.f / 6.f) * d2;
    const float c1 = (1.f / 6.f) * d2 + (1.f / 3.f) * d3;
    const float c2 = (2.f / 3.f) * d3 - (1.f / 6.f) * d4;
    return (a0 * c0 + a1 * c1 + a2 * c2) * rcp_T((a0 + a1) + a2);
}
__device__ __forceinline__ void weno5_pair(P2<float> s0, P2<float> s1, P2<float> s2, P2<float> s3, P2<float> s4,
                                           float rdx, float &sgx, float &sgy) {
    sgx = weno5_f(s0.x, s1.x, s2.x, s3.x, s4.x) * rdx;
    sgy = weno5_f(s0.y, s1.y, s2.y, s3.y, s4.y) * rdx;
}
template <typename T> __device__ __forceinline__ T vanleer_T(T s1, T s2) {
    T a1 = s1 < T(0) ? -s1 : s1, a2 = s2 < T(0) ? -s2 : s2;
    T sg = (s1 > T(0) ? T(1) : (s1 < T(0) ? T(-1) : T(0))) + (s2 > T(0) ? T(1) : (s2 < T(0) ? T(-1) : T(0)));
    return sg * (a1 * a2) / (a1 + a2 + T(1e-16));
}
template <typename T> __device__ __forceinline__ T minmod_T(T a, T b) {
    T fa = a < T(0) ? -a : a, fb = b < T(0) ? -b : b;
    if ((fa < fb) && (a * b > T(0))) return a;
    if ((fb < fa) && (a * b > T(0))) return b;
    if (a == b) return a;
    return T(0);
}

#define AS1 __attribute__((address_space(1)))

// per-face geometry of both upwind choices (delta = 1: upwind cell below the
// face, stencil runs upwards; delta = 0: from above, downwards)
struct FaceGeo {
    size_t own;         // owner cell (gi, gj) linear index
    long long step;     // cells per unit step along the normal (1 for S, NJ for W)
    double dist[2];     // |midside - c[MID]|
    double rdist[2];    // 1 / dist
    double rd1[2];      // 1 / |c[MID] - c[MID-1]| (vanLeer divisor)
    float fdist[2], frdist[2], frd1[2];   // the same in float, for the FP32 reconstruction
};

__device__ inline FaceGeo face_geo(const FBlock &B, int gi, int gj, int face) {
    const Consts &C = B.c;
    FaceGeo g;
    g.own = (size_t)gi * C.NJ + gj;
    g.step = face == SOUTH ? 1 : C.NJ;
    d2 ms = B.mid_side[g.own * 2 + face];
#pragma unroll
    for (int delta = 0; delta < 2; delta++) {
        long long cm = delta ? (long long)g.own - g.step : (long long)g.own;             // stencil[MID]
        long long cm1 = delta ? (long long)g.own - 2 * g.step : (long long)g.own + g.step; // stencil[MID-1]
        d2 c0 = B.centre[cm], c1 = B.centre[cm1];
        g.dist[delta] = length(ms - c0);
        g.rdist[delta] = 1.0 / g.dist[delta];
        g.rd1[delta] = 1.0 / length(c0 - c1);
        g.fdist[delta] = (float)g.dist[delta];
        g.frdist[delta] = (float)g.rdist[delta];
        g.frd1[delta] = (float)g.rd1[delta];
    }
    return g;
}

// Stencil addressing: one (wave-uniform) base pointer at the lowest stencil
// cell, own - R*step (R = 3 for WENO5, 2 otherwise), and 32-bit byte offsets
// rel*sb + v*sizeof(P2) per lane; delta = 1 takes rel = k, delta = 0 rel =
// 2R-1-k for stencil index k.
template <typename T> struct Stencil {
    const AS1 char *base;
    unsigned sb;
};
// M: the reconstruction method when the kernel is specialised for one
// (0 vanLeer, 1 WENO5, 2 NND), -1 to read it from the block
template <int M> __device__ __forceinline__ int method_of(const Consts &C) { return M >= 0 ? M : C.method; }

template <typename T, int M = -1>
__device__ __forceinline__ Stencil<T> stencil(const Consts &C, const void *F, const FaceGeo &g) {
    const int R = method_of<M>(C) == 1 ? 3 : 2;
    Stencil<T> st;
    st.base = (const AS1 char *)F + ((long long)g.own - R * g.step) * (long long)C.NV * (long long)sizeof(P2<T>);
    st.sb = (unsigned)(g.step * C.NV * (long long)sizeof(P2<T>));
    return st;
}
// The lane offsets rel*sb + vb above are 32-bit: the farthest stencil cell of
// a W face is 2R-1 rows (NJ cells each) above the base, plus the lane's
// velocity.  Plans refuse blocks whose span does not fit.
inline bool stencil_span_fits(long long NJ, long long NV, int method, int precision) {
    const long long pair = precision == UGKS_FP64 ? 16 : 8;
    const long long R = method == 1 ? 3 : 2;
    const long long span = ((2 * R - 1) * NJ + 1) * NV * pair;
    return span < (1LL << 32);
}
template <typename T> using V2 = T __attribute__((ext_vector_type(2)));
// a store that is not read again soon (per-velocity fluxes, the new f): it
// streams past the caches instead of evicting stencil data still in use
template <typename T> __device__ __forceinline__ void stream_store(V2<T> *p, V2<T> x) { __builtin_nontemporal_store(x, p); }
template <typename T> __device__ __forceinline__ void stream_store(__attribute__((address_space(1))) V2<T> *p, V2<T> x) {
    __builtin_nontemporal_store(x, p);
}

// velocity table reads through global (not generic) pointers: they count only
// against the vector-memory counter, so waiting for them does not also wait on LDS
__device__ __forceinline__ d2 quad_at(const Consts &C, int v) {
    const V2<double> x = ((const AS1 V2<double> *)C.quad)[v];
    return d2{x.x, x.y};
}
__device__ __forceinline__ double weight_at(const Consts &C, int v) { return ((const AS1 double *)C.weight)[v]; }
// a wavefront-uniform double held in scalar registers (keeps per-face
// constants out of VGPRs, so they are neither spilled nor recomputed in loops)
__device__ __forceinline__ double uniform_d(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b), hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// the reciprocals go through an empty asm so the compiler cannot turn the
// per-lane select of two of them back into a division inside the velocity loop
__device__ __forceinline__ void uniform_geo(FaceGeo &g) {
#pragma unroll
    for (int k = 0; k < 2; k++) {
        g.dist[k] = uniform_d(g.dist[k]);
        g.rdist[k] = uniform_d(g.rdist[k]);
        g.rd1[k] = uniform_d(g.rd1[k]);
        __asm__ volatile("" : "+v"(g.rdist[k]));
        __asm__ volatile("" : "+v"(g.rd1[k]));
        // float copies held in VGPRs: the per-lane pick of one of two is then a
        // single select, not two moves, a select and a conversion per velocity
        g.fdist[k] = (float)g.dist[k];
        g.frdist[k] = (float)g.rdist[k];
        g.frd1[k] = (float)g.rd1[k];
        __asm__ volatile("" : "+v"(g.fdist[k]));
        __asm__ volatile("" : "+v"(g.frdist[k]));
        __asm__ volatile("" : "+v"(g.frd1[k]));
    }
}
// the per-face distance of upwind choice delta in the reconstruction's precision
template <typename T> __device__ __forceinline__ T geo_pick(const double (&d)[2], const float (&f)[2], int delta) {
    if constexpr (sizeof(T) == 4) return delta ? f[1] : f[0];
    else return delta ? d[1] : d[0];
}
template <typename T> __device__ __forceinline__ P2<T> sld(const Stencil<T> &st, unsigned rel, unsigned vb) {
    const V2<T> x = *(const AS1 V2<T> *)(st.base + (rel * st.sb + vb));
    return P2<T>{x.x, x.y};
}

// the stencil values of velocity v for upwind choice delta (5 cells for
// WENO5, 3 otherwise), stencil index order; named members keep them in
// registers across the load / compute split
template <typename T> struct StenVals { P2<T> s0, s1, s2, s3, s4; };
template <typename T, int M = -1>
__device__ __forceinline__ void stencil_load(const Consts &C, const Stencil<T> &st, int v, int delta, StenVals<T> &x) {
    const unsigned vb = (unsigned)v * (unsigned)sizeof(P2<T>);
    if (method_of<M>(C) == 1) {
        x.s0 = sld<T>(st, delta ? 0 : 5, vb);
        x.s1 = sld<T>(st, delta ? 1 : 4, vb);
        x.s2 = sld<T>(st, delta ? 2 : 3, vb);
        x.s3 = sld<T>(st, delta ? 3 : 2, vb);
        x.s4 = sld<T>(st, delta ? 4 : 1, vb);
    } else {
        x.s0 = sld<T>(st, delta ? 0 : 3, vb);
        x.s1 = sld<T>(st, delta ? 1 : 2, vb);
        x.s2 = sld<T>(st, delta ? 2 : 1, vb);
        x.s3 = x.s4 = x.s2;
    }
}

// reconstructed face value and slope (iFace) from the stencil values; fl /
// fr, the cells below / above the face, on request
template <typename T, int M = -1>
__device__ __forceinline__ void reconstruct_from(const Consts &C, const FaceGeo &g, int delta, const StenVals<T> &x,
                                                 P2<T> &ff, P2<T> &sg, P2<T> *fl = nullptr, P2<T> *fr = nullptr) {
    if (method_of<M>(C) == 1) {
        const T dist = geo_pick<T>(g.dist, g.fdist, delta), rdx = geo_pick<T>(g.rdist, g.frdist, delta);
        weno5_pair(x.s0, x.s1, x.s2, x.s3, x.s4, rdx, sg.x, sg.y);
        ff.x = x.s2.x + sg.x * dist;
        ff.y = x.s2.y + sg.y * dist;
        if (fl) *fl = P2<T>{delta ? x.s2.x : x.s3.x, delta ? x.s2.y : x.s3.y};
        if (fr) *fr = P2<T>{delta ? x.s3.x : x.s2.x, delta ? x.s3.y : x.s2.y};
    } else {
        const T dist = geo_pick<T>(g.dist, g.fdist, delta);
        if (method_of<M>(C) == 0) {
            const T r1 = geo_pick<T>(g.rd1, g.frd1, delta);
            sg.x = vanleer_T<T>((x.s1.x - x.s0.x) * r1, (x.s2.x - x.s1.x) * r1);
            sg.y = vanleer_T<T>((x.s1.y - x.s0.y) * r1, (x.s2.y - x.s1.y) * r1);
        } else {
            const T r = geo_pick<T>(g.rdist, g.frdist, delta);
            sg.x = (x.s1.x + T(0.5) * minmod_T<T>(x.s2.x - x.s1.x, x.s1.x - x.s0.x) - x.s1.x) * r;
            sg.y = (x.s1.y + T(0.5) * minmod_T<T>(x.s2.y - x.s1.y, x.s1.y - x.s0.y) - x.s1.y) * r;
        }
        ff.x = x.s1.x + sg.x * dist;
        ff.y = x.s1.y + sg.y * dist;
        if (fl) *fl = P2<T>{delta ? x.s1.x : x.s2.x, delta ? x.s1.y : x.s2.y};
        if (fr) *fr = P2<T>{delta ? x.s2.x : x.s1.x, delta ? x.s2.y : x.s1.y};
    }
}

template <typename T>
__device__ __forceinline__ void reconstruct(const Consts &C, const Stencil<T> &st, const FaceGeo &g, int v, int delta,
                                            P2<T> &ff, P2<T> &sg, P2<T> *fl = nullptr, P2<T> *fr = nullptr) {
    StenVals<T> x;
    stencil_load<T>(C, st, v, delta, x);
    reconstruct_from<T>(C, g, delta, x, ff, sg, fl, fr);
}

#define AS3 __attribute__((address_space(3)))
// LDS bytes of one wavefront's stencil staging: all 2R stencil cells (both
// upwind choices) of 64 velocities.  FP64 only: a lane's (g, h) pair of floats
// is 8 bytes, a size the LDS loads lack.  The flux pass stages its stencils
// this way; the moment pass, which has no store in its loop and runs at a
// higher occupancy, measured faster with plain loads.
template <typename T> constexpr int kStenRow = 64 * (int)sizeof(P2<T>);
template <typename T> constexpr bool kStenViaLds = sizeof(T) == 8;
template <typename T> constexpr int kStenLds = kStenViaLds<T> ? 6 * kStenRow<T> : 16;

// Loads of the 2R stencil cells of the velocities at byte offset vb, issued
// straight into the wavefront's LDS buffer lb (no registers held while they
// are in flight): cell r of lane l lands at lb + r * kStenRow + l * sizeof(P2).
template <typename T, int M>
__device__ __forceinline__ void stencil_dma(const Consts &C, const Stencil<T> &st, unsigned vb, AS3 char *lb) {
    constexpr int sz = 16;   // one (g, h) pair of doubles per lane
    const int ncell = method_of<M>(C) == 1 ? 6 : 4;
    // The row addresses are formed again at every call, by scalar adds.
    // Hoisted out of the velocity loop, the six 64-bit global and six LDS
    // row addresses took 18 scalar registers the flux pass does not have:
    // they were spilled to vector lanes and read back with v_readlane
    // (vector instructions, each followed by wait states) in every iteration.
    unsigned sb = st.sb, lds = (unsigned)(size_t)lb;
    __asm__ volatile("" : "+s"(sb), "+s"(lds));
#pragma unroll
    for (int r = 0; r < 6; r++) {
        if (r < ncell) {
            const AS1 char *src = st.base + r * sb;
            __builtin_amdgcn_global_load_lds((const AS1 void *)(src + vb), (AS3 void *)(size_t)(lds + r * kStenRow<T>), sz,
                                             0, 0);
        }
    }
}
template <typename T> __device__ __forceinline__ P2<T> lld(const AS3 char *p) {
    const V2<T> x = *(const AS3 V2<T> *)p;
    return P2<T>{x.x, x.y};
}
// the stencil of this lane's velocity from the staging buffer, in stencil
// index order for upwind choice delta (cells 0..4 for delta = 1, 5..1 for 0)
template <typename T, int M>
__device__ __forceinline__ void stencil_read(const Consts &C, const AS3 char *lb, int lane, int delta, StenVals<T> &x) {
    constexpr int row = kStenRow<T>;
    const int last = method_of<M>(C) == 1 ? 5 : 3;
    const AS3 char *p = lb + lane * (int)sizeof(P2<T>) + (delta ? 0 : last * row);
    const int s = delta ? row : -row;
    x.s0 = lld<T>(p);
    x.s1 = lld<T>(p + s);
    x.s2 = lld<T>(p + 2 * s);
    if (method_of<M>(C) == 1) {
        x.s3 = lld<T>(p + 3 * s);
        x.s4 = lld<T>(p + 4 * s);
    } else {
        x.s3 = x.s4 = x.s2;
    }
}

// Velocity loop of a wavefront-per-face pass, v = lane, lane + 64, ...: the
// body gets the global velocity e, its weight w, the face-frame velocity uv,
// the upwind choice and the stencil.  The stencil of the next velocity is
// loaded into the wavefront's LDS buffer lb while the body of this one runs
// (both upwind choices, so the loads need not wait for the next velocity);
// the velocity and weight are loaded one iteration ahead into registers.
template <typename T, int M, bool LDS, typename Body>
__device__ __forceinline__ void face_velocity_loop(const Consts &C, const Stencil<T> &st, d2 n, int lane, AS3 char *lb,
                                                   Body body) {
    static_assert(!LDS || kStenViaLds<T>, "LDS stencil staging needs FP64 pairs");
    if (lane >= C.NV) return;
    d2 e = quad_at(C, lane);
    double w = weight_at(C, lane);
    unsigned vb = (unsigned)lane * (unsigned)sizeof(P2<T>);
    if constexpr (LDS) stencil_dma<T, M>(C, st, vb, lb);
    for (int v = lane; v < C.NV; v += 64) {
        const d2 ec = e;
        const double wc = w;
        const d2 uv = d2{ec.x * n.x + ec.y * n.y, ec.x * (-n.y) + ec.y * n.x};     // face_velocity
        const int delta = upwind(uv.x);
        StenVals<T> cur;
        // the next velocity's table entries, or this one's again on the last
        // round: the loads then need no branch (and no register copies around
        // one).  The stencil loads into LDS stay conditional, so none is still
        // in flight when the wavefront leaves the loop.
        const int vn = v + 64 < C.NV ? v + 64 : v;
        if constexpr (LDS) {
            stencil_read<T, M>(C, lb, lane, delta, cur);
            // the reads have returned before the next loads overwrite the buffer
            __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
            e = quad_at(C, vn);
            w = weight_at(C, vn);
            if (v + 64 < C.NV) {
                vb += 64u * (unsigned)sizeof(P2<T>);
                stencil_dma<T, M>(C, st, vb, lb);
            }
        } else {
            e = quad_at(C, vn);
            w = weight_at(C, vn);
            stencil_load<T, M>(C, st, v, delta, cur);
        }
        body(v, ec, wc, uv, delta, cur);
    }
}

// the cells below (rel R-1) and above (rel R, the owner) the face
template <typename T>
__device__ __forceinline__ void face_cells(const Consts &C, const Stencil<T> &st, int v, P2<T> &fl, P2<T> &fr) {
    const unsigned vb = (unsigned)v * (unsigned)sizeof(P2<T>);
    const unsigned R = C.method == 1 ? 3 : 2;
    fl = sld<T>(st, R - 1, vb);
    fr = sld<T>(st, R, vb);
}

// ============================================================ face passes ==
// Faces are numbered so that neighbouring ids share stencil cells: S faces
// column by column (j fastest), W faces row by row (i fastest).
// Per interior face: (A) one wavefront reconstructs f at the face for every
// velocity and sums the raw face moments (the conserved moments of
// getFaceCons, the third-order ones the face heat flux of calcFaceQ needs,
// clUGKS.c:522-585), the left/right cell moments (getAL/getAR) and the owner
// cell's heat flux (calcQ); (S) one thread turns those sums into the face
// state (getPLR, initMacroFlux, calcFaceQ); (C) one wavefront rebuilds the
// face values and forms the per-velocity flux (distFlux) and the macro flux
// (macroFlux).  Face ids: per block, S faces then W faces (walls excluded).
constexpr int kNS = 15;   // face sums: 13 raw moments of the face values, the owner's heat flux
constexpr int kFS = 16;   // face sums per face (padded)
constexpr int kFT = 32;   // face state per face

struct FaceLoc { int face, mi, mj, gi, gj; };
__device__ __forceinline__ FaceLoc face_loc(const FBlock &B, long long loc) {
    const Consts &C = B.c;
    FaceLoc f;
    if (loc < B.nS) { f.face = SOUTH; f.mi = (int)(loc / B.nSj); f.mj = (int)(loc % B.nSj) + B.ob_S; }
    else { long long r = loc - B.nS; f.face = WEST; f.mj = (int)(r / B.nWi); f.mi = (int)(r % B.nWi) + B.ob_W; }
    f.gi = f.mi + C.G;
    f.gj = f.mj + C.G;
    return f;
}
// the face_loc index of a face
__device__ __forceinline__ long long face_index(const FBlock &B, const FaceLoc &f) {
    return f.face == SOUTH ? (long long)f.mi * B.nSj + (f.mj - B.ob_S)
                           : B.nS + (long long)f.mj * B.nWi + (f.mi - B.ob_W);
}

// The moment pass walks a block's faces in tiles of kTileI x kTileJ cells
// (tiles j-fastest): each tile's S faces (j fastest), then its W faces (i
// fastest).  The S faces form an ni x nSj rectangle (mi, mj - ob_S), the W faces
// an nWi x nj one (mi - ob_W, mj); the same tile grid cuts both, so the two
// kinds of face that read a cell's stencil data run at about the same time
// and share it in the caches.  (The flux pass measured faster in the plain
// order, and it and the state pass keep it; the moment pass writes each
// face's sums at the face's plain index.)  loc < 2^31 (checked on the host).
constexpr int kTileI = 4, kTileJ = 16;
__device__ __forceinline__ FaceLoc face_loc_tiled(const FBlock &B, long long loc64) {
    const Consts &C = B.c;
    const int loc = (int)loc64;
    const int ni = C.ni, nj = C.nj, nSj = B.nSj, nWi = B.nWi;
    const int nti = (max(ni, nWi) + kTileI - 1) / kTileI, ntj = (max(nSj, nj) + kTileJ - 1) / kTileJ;
    auto rows_before = [&](int ti) { return min(ti * kTileI, ni) * nSj + min(ti * kTileI, nWi) * nj; };
    int ti = min(loc / (kTileI * (nSj + nj)), nti - 1);
    while (ti > 0 && rows_before(ti) > loc) ti--;
    while (ti + 1 < nti && rows_before(ti + 1) <= loc) ti++;
    const int rS = max(0, min(kTileI, ni - ti * kTileI)), rW = max(0, min(kTileI, nWi - ti * kTileI));
    const int local = loc - rows_before(ti);
    auto cols_before = [&](int tj) { return rS * min(tj * kTileJ, nSj) + rW * min(tj * kTileJ, nj); };
    const int per = kTileJ * (rS + rW);
    int tj = per > 0 ? min(local / per, ntj - 1) : 0;
    while (tj > 0 && cols_before(tj) > local) tj--;
    while (tj + 1 < ntj && cols_before(tj + 1) <= local) tj++;
    const int cS = max(0, min(kTileJ, nSj - tj * kTileJ));
    const int l2 = local - cols_before(tj);
    FaceLoc f;
    if (l2 < rS * cS) {
        f.face = SOUTH; f.mi = ti * kTileI + l2 / cS; f.mj = tj * kTileJ + l2 % cS + B.ob_S;
    } else {
        const int r = l2 - rS * cS;
        f.face = WEST; f.mi = ti * kTileI + r % rW + B.ob_W; f.mj = tj * kTileJ + r / rW;
    }
    f.gi = f.mi + C.G;
    f.gj = f.mj + C.G;
    return f;
}

// One interior face as the face passes see it (64 bytes, one scalar load):
// its block, side (SOUTH / WEST), owner cell and the reconstruction
// distances of both upwind choices.  The plan fills the table from the mesh
// (k_facerec) at every run call, so a geometry update under a live plan is
// picked up; the face passes then neither search the block offsets nor
// recompute distances per face.
struct __align__(64) FaceRec {
    int b, face, gi, gj;
    double dist[2], rdist[2], rd1[2];
};
static_assert(sizeof(FaceRec) == 64, "face record is one 64-byte line");

__device__ __forceinline__ FaceGeo face_geo_of(const FaceRec &r, const Consts &C) {
    FaceGeo g;
    g.own = (size_t)r.gi * C.NJ + r.gj;
    g.step = r.face == SOUTH ? 1 : C.NJ;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        g.dist[k] = r.dist[k];
        g.rdist[k] = r.rdist[k];
        g.rd1[k] = r.rd1[k];
    }
    return g;
}

// rec[fid]: the face of plain index fid (per block, S faces then W faces);
// tmap[fid]: the plain index of the moment pass's item fid (tiled order)
__global__ void __launch_bounds__(kT) k_facerec(const FBlock *__restrict__ blocks, const long long *__restrict__ off,
                                                int nb, long long nfaces, FaceRec *__restrict__ rec,
                                                int *__restrict__ tmap) {
    const long long fid = (long long)blockIdx.x * kT + threadIdx.x;
    if (fid >= nfaces) return;
    const int b = find_block(off, nb, fid);
    const FBlock &B = blocks[b];
    const long long loc = fid - off[b];
    const FaceLoc L = face_loc(B, loc);
    const FaceGeo g = face_geo(B, L.gi, L.gj, L.face);
    FaceRec r;
    r.b = b; r.face = L.face; r.gi = L.gi; r.gj = L.gj;
#pragma unroll
    for (int k = 0; k < 2; k++) { r.dist[k] = g.dist[k]; r.rdist[k] = g.rdist[k]; r.rd1[k] = g.rd1[k]; }
    rec[fid] = r;
    tmap[fid] = (int)(off[b] + face_index(B, face_loc_tiled(B, loc)));
}

template <typename T, int FW, int M>
__global__ void __launch_bounds__(64 * FW) k_faceA(const FBlock *__restrict__ blocks, const FaceRec *__restrict__ recs,
                                                   const int *__restrict__ tmap, long long nfaces,
                                                   double *__restrict__ fsum) {
    const long long fid = xcd_remap(blockIdx.x, gridDim.x) * FW + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (fid >= nfaces) return;
    const int lane = threadIdx.x & 63;
    const long long sid = tmap[fid];                          // where the state pass looks for this face
    const FaceRec R = recs[sid];
    const FBlock &B = blocks[R.b];
    const Consts &C = B.c;
    const FaceLoc L{R.face, R.gi - C.G, R.gj - C.G, R.gi, R.gj};
    FaceGeo g = face_geo_of(R, C);
    uniform_geo(g);
    const Stencil<T> stn = stencil<T, M>(C, B.f, g);
    const d2 n = B.normal[g.own * 2 + L.face];
    const bool wantQ = (L.face == SOUTH) && (L.mj < C.nj);
    const d4 cm = wantQ ? B.macro[(size_t)L.mi * C.nj + L.mj] : mk4(0, 0, 0, 1);
    const double rK = 1.0 / C.K;
    double a[kFS];            // kNS sums and one pad, so the reduction folds 16
#pragma unroll
    for (int q = 0; q < kFS; q++) a[q] = 0.0;
    face_velocity_loop<T, M, false>(C, stn, n, lane, nullptr, [&](int v, d2 e, double, d2 uv, int delta, const StenVals<T> &sv) __attribute__((always_inline)) {
        P2<T> ff, sg, fl, fr;
        reconstruct_from<T, M>(C, g, delta, sv, ff, sg, &fl, &fr);
        const double u = uv.x, w = uv.y, gx = (double)ff.x, hx = (double)ff.y;
        const double ug = u * gx, wg = w * gx, uug = u * ug, wwg = w * wg;
        a[0] += gx; a[1] += ug; a[2] += wg;
        a[3] += uug; a[4] += wwg; a[5] += u * wg;
        a[6] += u * uug; a[7] += w * wwg; a[8] += u * wwg; a[9] += w * uug;
        a[10] += hx; a[11] += u * hx; a[12] += w * hx;
        if (wantQ) {
            const double x = (double)fr.x, y = (double)fr.y;
            const double cx = e.x - cm.s1, cy = e.y - cm.s2, t = (cx * cx + cy * cy) * x + y * rK;
            a[13] += 0.5 * (cx * t);
            a[14] += 0.5 * (cy * t);
        }
    });
    // row r of the wavefront gets the totals of sums 4r .. 4r+3; its first
    // lane stores them (the pad, sum 15, is zero)
    double t[4];
    wave_sum16_rows(a, t);
    if ((lane & 15) == 0) {
        d2 *o = (d2 *)(fsum + sid * kFS + 4 * (lane >> 4));
        o[0] = mk2(t[0], t[1]);
        o[1] = mk2(t[2], t[3]);
        if (wantQ && lane == 48) B.Q[(size_t)L.mi * C.nj + L.mj] = mk2(t[1], t[2]);
    }
}

// face state: primary variables, slopes, time slope, time integrals, the
// equilibrium part of the macro flux and the face heat flux (from the raw
// moments about the face velocity) -- getPLR, initMacroFlux, calcFaceQ
__global__ void __launch_bounds__(kT) k_faceS(const FBlock *__restrict__ blocks, const FaceRec *__restrict__ recs,
                                              long long nfaces, const double *__restrict__ fsum,
                                              double *__restrict__ fstate, const DtState *__restrict__ ds) {
    const long long fid = (long long)blockIdx.x * kT + threadIdx.x;
    if (fid >= nfaces) return;
    const FaceRec R = recs[fid];
    const FBlock &B = blocks[R.b];
    const Consts &C = B.c;
    const FaceLoc L{R.face, R.gi - C.G, R.gj - C.G, R.gi, R.gj};
    const size_t own = (size_t)L.gi * C.NJ + L.gj;
    const long long step = L.face == SOUTH ? 1 : C.NJ;
    const double *S = fsum + fid * kFS;
    const d4 W = mk4(S[0], S[1], S[2], 0.5 * ((S[3] + S[4]) + S[10]));
    // getAL / getAR: the moments of the cells below and above the face in face coordinates
    const d2 nf = B.normal[own * 2 + L.face];
    const d4 ML = B.cmom[(long long)own - step], MR = B.cmom[own];
    const d2 uL = to_local(mk2(ML.s1, ML.s2), nf), uR = to_local(mk2(MR.s1, MR.s2), nf);
    const d4 WL = mk4(ML.s0, uL.x, uL.y, ML.s3), WR = mk4(MR.s0, uR.x, uR.y, MR.s3);
    const d4 prim = to_primary(C, W);
    // |midside - centre| of the cells below (upwind choice 1) and above (0) the face
    const d4 aL = micro_slope(C, prim, (W - WL) / R.dist[1]);
    const d4 aR = micro_slope(C, prim, (WR - W) / R.dist[0]);
    Moments m;
    velocity_moments(C, prim, m);
    const d4 sw = (-prim.s0) * (moment_au(aL, m.MuL, m.Mv, m.Mxi, 1, 0) + moment_au(aR, m.MuR, m.Mv, m.Mxi, 1, 0));
    const d4 aT = micro_slope(C, prim, sw);
    const TimeInt mt = time_integrals(relax_time(C, prim), ds[B.cs].dt_flux);
    const d4 g0 = prim.s0 * (mt.m0 * moment_uv(m.Mu, m.Mv, m.Mxi, 1, 0, 0) +
                             mt.m1 * (moment_au(aL, m.MuL, m.Mv, m.Mxi, 2, 0) + moment_au(aR, m.MuR, m.Mv, m.Mxi, 2, 0)) +
                             mt.m2 * moment_au(aT, m.Mu, m.Mv, m.Mxi, 1, 0));
    // face heat flux 1/2 sum (c-U)(|c-U|^2 g + h/K) from the raw moments
    const double U = prim.s1, V = prim.s2;
    const double qx = 0.5 * ((S[6] - 3.0 * U * S[3] + 3.0 * U * U * S[1] - U * U * U * S[0]) +
                             (S[8] - 2.0 * V * S[5] + V * V * S[1] - U * S[4] + 2.0 * U * V * S[2] - U * V * V * S[0]) +
                             (S[11] - U * S[10]) / C.K);
    const double qy = 0.5 * ((S[7] - 3.0 * V * S[4] + 3.0 * V * V * S[2] - V * V * V * S[0]) +
                             (S[9] - 2.0 * U * S[5] + U * U * S[2] - V * S[3] + 2.0 * U * V * S[1] - V * U * U * S[0]) +
                             (S[12] - V * S[10]) / C.K);
    const d2 n = B.normal[own * 2 + L.face];
    // The per-velocity flux (distFlux) with its per-face factors folded, so the
    // flux pass evaluates, for P_a = a0 + a1 u + a2 w + a3 |u|^2 / 2,
    //   Q_side = m1 (P_side + hk a3 / 2)        (side = L or R, by the upwind choice)
    //   Q_T    = m2 (P_T + hk aT3 / 2) + m0
    //   D_a    = m1 a3 (Mxi2 - hk^2) / 2        (D_T likewise with m2, aT)
    //   S      = m0 0.8 (1 - Pr) lam^2 / rho ((c - U) . q)
    //   g: Lf u (f0 (Q_T + u Q_side + S (2 lam c^2 - 4)) + m3 f - m4 u sigma)
    //   h: Lf u (f0 (hk (Q_T + u Q_side + S (2 lam c^2 - 2)) + u D_side + D_T) + m3 f_h - m4 u sigma_h)
    // with Lf the face length carried in f0 = w Lf rho lam / pi e^{-lam c^2} and
    // in m3, m4; the macro flux sums take the same factors.
    const double Lf = B.side_length[own * 2 + L.face];
    const double lam = prim.s3, hk = C.K / (2.0 * lam), dk = m.Mxi[2] - hk * hk;
    const double hL = 0.5 * aL.s3, hR = 0.5 * aR.s3, hT = 0.5 * aT.s3;
    const double sm = mt.m0 * (0.8 * (1 - C.Pr) * (lam * lam) / prim.s0);
    double *o = fstate + fid * kFT;
    o[0] = prim.s1; o[1] = prim.s2; o[2] = lam; o[3] = hk;
    o[4] = mt.m1 * (aL.s0 + hL * hk); o[5] = mt.m1 * aL.s1; o[6] = mt.m1 * aL.s2; o[7] = mt.m1 * hL;
    o[8] = mt.m1 * (aR.s0 + hR * hk); o[9] = mt.m1 * aR.s1; o[10] = mt.m1 * aR.s2; o[11] = mt.m1 * hR;
    o[12] = mt.m2 * (aT.s0 + hT * hk) + mt.m0; o[13] = mt.m2 * aT.s1; o[14] = mt.m2 * aT.s2; o[15] = mt.m2 * hT;
    o[16] = mt.m1 * hL * dk; o[17] = mt.m1 * hR * dk; o[18] = mt.m2 * hT * dk;
    o[19] = Lf * mt.m3; o[20] = Lf * mt.m4;
    o[21] = sm * qx; o[22] = sm * qy;
    o[23] = Lf * (prim.s0 * (lam / kPi));
    o[24] = g0.s0; o[25] = g0.s1; o[26] = g0.s2; o[27] = g0.s3;
    o[28] = Lf; o[29] = n.x; o[30] = n.y; o[31] = 0.0;
}

// minimum wavefronts per SIMD asked of the flux pass: the FP32 variants with a
// fixed reconstruction run six (80 VGPRs; WENO5 spills 4 registers there and
// is still faster than at five); the others are left to the compiler
template <typename T, int M> constexpr int kFaceCWaves = (sizeof(T) == 4 && M >= 0) ? 6 : 1;
// the FP64 flux stores one iteration late (the FP32 variant would spill more at six waves)
template <typename T> constexpr bool kDeferStore = sizeof(T) == 8;
template <typename T, int FW, int M>
__global__ void __launch_bounds__(64 * FW) __attribute__((amdgpu_waves_per_eu(kFaceCWaves<T, M>))) k_faceC(const FBlock *__restrict__ blocks, const FaceRec *__restrict__ recs,
                                                   long long nfaces, const double *__restrict__ fstate) {
    __shared__ double tab[FW][128];
    __shared__ __attribute__((aligned(16))) char sbuf[FW][kStenLds<T>];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long long fid = xcd_remap(blockIdx.x, gridDim.x) * FW + wv;
    if (fid >= nfaces) return;
    const int lane = threadIdx.x & 63;
    const FaceRec R = recs[fid];
    const FBlock &B = blocks[R.b];
    const Consts &C = B.c;
    const FaceLoc L{R.face, R.gi - C.G, R.gj - C.G, R.gi, R.gj};
    FaceGeo g = face_geo_of(R, C);
    uniform_geo(g);
    const Stencil<T> stn = stencil<T, M>(C, B.f, g);
    // the face's folded flux constants (k_faceS)
    const double *sv = fstate + fid * kFT;
    const double U = sv[0], V = sv[1], lam = sv[2], hk = sv[3];
    const d4 QL = mk4(sv[4], sv[5], sv[6], sv[7]), QR = mk4(sv[8], sv[9], sv[10], sv[11]);
    const d4 QT = mk4(sv[12], sv[13], sv[14], sv[15]);
    const double DL = sv[16], DR = sv[17], DT = sv[18], m3 = sv[19], m4 = sv[20];
    const d2 SQ = mk2(sv[21], sv[22]);
    const double rl = sv[23], Lf = sv[28];
    const d2 n = mk2(sv[29], sv[30]);
    // separable Maxwellian on a tensor grid with an axis-aligned normal:
    // exp(-lam |uv - (U,V)|^2) = Ei[v / nvt] * Ej[v % nvt], one table pair per wave in LDS
    const bool sep = C.nut > 0 && (n.x == 0.0 || n.y == 0.0);
    double *Ei = tab[wv], *Ej = tab[wv] + 64;
    if (sep) {
        const bool wtype = (n.y == 0.0);
        if (lane < C.nut) {
            const double a = wtype ? n.x * C.qu[lane] : -n.y * C.qu[lane], c = wtype ? U : V;
            Ei[lane] = exp(-lam * ((a - c) * (a - c)));
        }
        if (lane < C.nvt) {
            const double a = wtype ? n.x * C.qv[lane] : n.y * C.qv[lane], c = wtype ? V : U;
            Ej[lane] = exp(-lam * ((a - c) * (a - c)));
        }
    }
    // the tables are private to this wavefront: order its LDS writes before its reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    AS1 V2<T> *fout = (AS1 V2<T> *)(L.face == SOUTH ? B.fluxS : B.fluxW) + g.own * (size_t)C.NV;
    const double lam2 = 2.0 * lam;
    double mf[4] = {0.0, 0.0, 0.0, 0.0};
    // tensor indices of v = lane + 64 k, advanced without integer division
    int iu = 0, jv = 0, diu = 0, djv = 0;
    // the grid width in a register: the flux stores could alias the block
    // record, so C.nvt would otherwise be loaded again in every iteration
    const int nvt = __builtin_amdgcn_readfirstlane(C.nvt);
    if (sep) { iu = lane / nvt; jv = lane - iu * nvt; diu = 64 / nvt; djv = 64 - diu * nvt; }
    // each velocity's flux is stored one iteration late, at the start of the
    // next, so the wait for the loads issued there does not also wait for it
    V2<T> pend = V2<T>{T(0), T(0)};
    int last = -1;
    face_velocity_loop<T, M, kStenViaLds<T>>(C, stn, n, lane, (AS3 char *)sbuf[wv], [&](int v, d2, double wv, d2 uv, int delta, const StenVals<T> &stv) __attribute__((always_inline)) {
        if constexpr (kDeferStore<T>) {
            if (v >= 64) stream_store<T>(fout + (v - 64), pend);
            last = v;
        }
        P2<T> ff, sg;
        reconstruct_from<T, M>(C, g, delta, stv, ff, sg);
        const double ux = uv.x, uy = uv.y, cx = ux - U, cy = uy - V, c2 = cx * cx + cy * cy;
        const double E = sep ? Ei[iu] * Ej[jv] : exp(-lam * c2);
        if (sep) {
            iu += diu; jv += djv;
            if (jv >= nvt) { jv -= nvt; iu++; }
        }
        const double f0 = wv * (rl * E);                       // Lf f0 of g; hk f0 is h's
        const double uu = ux * ux + uy * uy;
        double PL = QL.s0 + QL.s1 * ux + QL.s2 * uy + QL.s3 * uu, PR = QR.s0 + QR.s1 * ux + QR.s2 * uy + QR.s3 * uu;
        // FP32 storage: both sides' polynomials are evaluated and one is picked
        // per lane.  Left to itself the compiler picks the coefficients instead,
        // which costs two moves per select (a select takes one scalar operand):
        // 19 more VALU per velocity.  The FP64 variant keeps the compiler's
        // form: with fewer VGPRs it runs four waves per SIMD, and measured
        // slower there (1.43 -> 1.50 ms on 2x2 blocks of 128^2 cells, 33^2 velocities).
        if constexpr (sizeof(T) == 4) __asm__ volatile("" : "+v"(PL), "+v"(PR));
        const double Qa = delta ? PL : PR, Da = delta ? DL : DR;
        const double Bg = fma(ux, Qa, QT.s0 + QT.s1 * ux + QT.s2 * uy + QT.s3 * uu);
        const double S = cx * SQ.x + cy * SQ.y, SP4 = S * fma(lam2, c2, -4.0), SP2 = fma(2.0, S, SP4);
        const double fdg = (double)ff.x, fdh = (double)ff.y, slg = (double)sg.x, slh = (double)sg.y;
        const double m4u = m4 * ux;
        const double tg = m3 * fdg - m4u * slg, th = m3 * fdh - m4u * slh;
        const double rg = ux * fma(f0, Bg + SP4, tg);
        const double rh = ux * fma(f0, fma(hk, Bg + SP2, fma(ux, Da, DT)), th);
        if constexpr (kDeferStore<T>) pend = V2<T>{(T)rg, (T)rh};
        else stream_store<T>(fout + v, V2<T>{(T)rg, (T)rh});
        // macroFlux (clUGKS.c:607-661): the Shakhov and the non-equilibrium parts
        const double X = ux * fma(SP4, f0, tg);
        mf[0] += X;
        mf[1] += ux * X;
        mf[2] += uy * X;
        mf[3] += 0.5 * (uu * X + ux * fma(SP2, hk * f0, th));
    });
    if (last >= 0) stream_store<T>(fout + last, pend);
    wave_sum4(mf);
    if (lane == 0) {
        d4 o = mk4(mf[0], mf[1], mf[2], mf[3]) + Lf * mk4(sv[24], sv[25], sv[26], sv[27]);
        const d2 gl = to_global(mk2(o.s1, o.s2), n);
        o.s1 = gl.x; o.s2 = gl.y;
        (L.face == SOUTH ? B.fmS : B.fmW)[g.own] = o;
    }
}

// ============================================================== k_wall ====
struct WallItem { int block, wall; long long start; };

// Three waves per SIMD asked of the wall pass up to 7 velocities per thread
// (NV <= 1792): the compiler's own choice, 170-184 VGPRs, leaves two, and the
// pass, a chain of loads and workgroup reductions per wall face, then waits
// with nothing to switch to.  Measured on the 64-case batch (41^2 velocities
// at most): 1.10 -> 0.86 ms mixed, 1.20 -> 0.97 ms FP64 (36 bytes spilled).
// Beyond 7 the spills grow and the pass is a negligible share of the step.
template <int NPT> constexpr int kWallWaves = NPT <= 7 ? 3 : 1;
template <typename T, int NPT>
__global__ void __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(kWallWaves<NPT>))) k_wall(const FBlock *__restrict__ blocks, const WallItem *__restrict__ items,
                                             int nitems, const DtState *__restrict__ ds) {
    __shared__ double red[kWavesT][16];
    __shared__ P2<T> wd[NPT * kT];   // reflected distribution (clAdsorb.c:377) for the specular mirror lookup
    __shared__ double sc[8];
    const long long item = blockIdx.x;
    const WallItem W = items[find_item(items, nitems, item)];
    const FBlock &B = blocks[W.block];
    const Consts &C = B.c;
    const int face = W.wall;
    const int ci = (int)(item - W.start);
    int gi, gj, rot, fid;
    switch (face) {
    case GNORTH: gi = ci + C.G; gj = C.NJ - C.G; rot = -1; fid = SOUTH; break;
    case GEAST: gi = C.NI - C.G; gj = ci + C.G; rot = -1; fid = WEST; break;
    case GSOUTH: gi = ci + C.G; gj = C.G; rot = 1; fid = SOUTH; break;
    default: gi = C.G; gj = ci + C.G; rot = 1; fid = WEST; break;
    }
    const double dt = ds[B.cs].dt_flux;
    const FaceGeo g = face_geo(B, gi, gj, fid);
    const d2 n = B.normal[g.own * 2 + fid];
    const double L = B.side_length[g.own * 2 + fid];
    const int mi = gi - C.G, mj = gj - C.G;
    const bool wantQ = (fid == SOUTH) && (mj < C.nj);
    const d4 cm = wantQ ? B.macro[(size_t)mi * C.nj + mj] : mk4(0, 0, 0, 1);

    // reconstruction at the wall face (iFace) and the owner heat flux
    P2<T> ff[NPT];
    double acc[2] = {0.0, 0.0};
    const Stencil<T> stn = stencil<T>(C, B.f, g);
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int v = threadIdx.x + kT * k;
        if (v < C.NV) {
            d2 uv = face_velocity(C, v, n);
            P2<T> sgk;
            reconstruct<T>(C, stn, g, v, upwind(uv.x), ff[k], sgk);
            if (wantQ) {
                P2<T> fl, fr;
                face_cells<T>(C, stn, v, fl, fr);
                d2 e = C.quad[v];
                double a = (double)fr.x, bq = (double)fr.y;
                double cx = e.x - cm.s1, cy = e.y - cm.s2, t = (cx * cx + cy * cy) * a + bq / C.K;
                acc[0] += 0.5 * (cx * t);
                acc[1] += 0.5 * (cy * t);
            }
        }
    }
    if (wantQ) {
        block_sum<2>(acc, red);
        if (threadIdx.x == 0) B.Q[(size_t)mi * C.nj + mj] = mk2(acc[0], acc[1]);
    }
    d4 wall = B.wall_prop[(size_t)ci * 4 + face];
    {
        d2 wl = to_local(mk2(wall.s1, wall.s2), n);
        wall.s1 = wl.x; wall.s2 = wl.y;
    }
    if (B.bc[face] == 4) {
        // accommodatingWallDist (clUGKS.c:776-906)
        double s[2] = {0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            const int v = threadIdx.x + kT * k;
            if (v < C.NV) {
                d2 uv = face_velocity(C, v, n);
                if (leaves_wall(uv.x, rot)) s[1] -= uv.x * maxwell(C, wall, uv, v).x;
                else s[0] += uv.x * (double)ff[k].x;
            }
        }
        block_sum<2>(s, red);
        const double ratio = s[0] / s[1];
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            const int v = threadIdx.x + kT * k;
            if (v < C.NV) {
                d2 uv = face_velocity(C, v, n);
                if (leaves_wall(uv.x, rot)) {
                    d2 m = ratio * maxwell(C, wall, uv, v);
                    ff[k] = P2<T>{(T)m.x, (T)m.y};
                }
            }
        }
    } else {
        // adsorbingWall_P1 (clAdsorb.c:253-482), constant or tabulated isotherm
        d4 cell;
        switch (face) {
        case GNORTH: cell = B.macro[(size_t)ci * C.nj + (gj - C.G - 1)]; break;
        case GEAST: cell = B.macro[(size_t)(gi - C.G - 1) * C.nj + ci]; break;
        case GSOUTH: cell = B.macro[(size_t)ci * C.nj + (gj - C.G)]; break;
        default: cell = B.macro[(size_t)(gi - C.G) * C.nj + ci]; break;
        }
        const double Tc = 1.0 / cell.s3;
        const size_t ck = (size_t)ci * 4 + face;
        const d4 cov0 = B.wall_cover[ck];
        const double vartheta = cov0.s0, gf = C.gamma_f[face];
        double s3[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            const int v = threadIdx.x + kT * k;
            if (v < C.NV) {
                d2 uv = face_velocity(C, v, n);
                int delta = leaves_wall(uv.x, rot);
                double tg = ((-rot * uv.x) * (1 - delta)) * (double)ff[k].x;
                double th = ((-rot * uv.x) * (1 - delta)) * (double)ff[k].y;
                double beta = (1.0 - C.beta_n[face] * exp(-uv.x * uv.x / Tc)) *
                              (1.0 - C.beta_t[face] * (1.0 - exp(-uv.y * uv.y / Tc)));
                double ag = beta * tg * gf * (1.0 - vartheta), ah = beta * th * gf * (1.0 - vartheta);
                int fs = (int)ocl_sign(tg);
                if ((fs * ag) > (fs * tg)) { ag = tg; ah = th; }
                s3[0] += tg - ag;
                s3[1] += ag;
                s3[2] += beta * tg;
                wd[threadIdx.x + kT * k] = P2<T>{(T)((tg - ag) / uv.x), (T)((th - ah) / uv.x)};
            }
        }
        block_sum<3>(s3, red);
        const double r_init = s3[0], a_init = s3[1], adjusted = s3[2];
        double a_final = a_init, dvtheta = 0.0, des = 0.0;
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
            const double veq = equilibrium_coverage(C, face, cell.s0, Tc, outside);
            if (outside) {
                early = true;
            } else {
                des = gf * (1.0 / veq - 1.0) * adjusted * vartheta;
                if (C.dyn[face] == 1) {
                    dvtheta -= dt * C.alpha_p[face] * des;
                    if ((vartheta + dvtheta) < 0.0) des = vartheta / (dt * C.alpha_p[face]);
                }
            }
        }
        double reflected = cov0.s1, ads_cov = cov0.s2, des_cov = cov0.s3;
        if (early) {
            if (threadIdx.x == 0) B.flag[0] = ERR_CLAMPING_ADSORB_ISOTHERM;
        } else {
            const double r_final = r_init + a_init - a_final;
            const double correct = r_final / r_init;
            if (correct != 1.0) {
#pragma unroll
                for (int k = 0; k < NPT; k++) {
                    P2<T> x = wd[threadIdx.x + kT * k];
                    wd[threadIdx.x + kT * k] = P2<T>{(T)(correct * (double)x.x), (T)(correct * (double)x.y)};
                }
            }
            reflected = r_final; ads_cov = a_final; des_cov = des;
            if (threadIdx.x == 0) B.wall_cover[ck] = mk4(vartheta + dvtheta, r_final, a_final, des);
        }
        __syncthreads();   // wd complete before the mirror lookups
        // adsorbingWallS_P2 / D_P2 / CL_P2 (clAdsorb.c:488-975): 0 specular, 1 diffuse, 2 Cercignani-Lampis
        const int mode = B.reflect[face];
        double ratio = 1.0;
        if (reflected > 0.0) {
            double o[1] = {0.0};
#pragma unroll
            for (int k = 0; k < NPT; k++) {
                const int v = threadIdx.x + kT * k;
                if (v < C.NV) {
                    d2 uv = face_velocity(C, v, n);
                    if (leaves_wall(uv.x, rot)) {
                        double eg, eh;
                        if (mode == 2) {
                            const d2 m = cl_emission(C, face, n, rot, wall, uv, v, [&](int va) {
                                const P2<T> x = wd[va];
                                return d2{(double)x.x, (double)x.y};
                            });
                            eg = m.x; eh = m.y;
                        } else if (mode == 1) {
                            d2 m = maxwell(C, wall, uv, v);
                            eg = m.x; eh = m.y;
                        } else {
                            int mv = (int)(fid == WEST ? C.mEW[v] : C.mNS[v]);
                            P2<T> x = wd[mv];
                            eg = (double)x.x; eh = (double)x.y;
                        }
                        ff[k] = P2<T>{(T)eg, (T)eh};
                        o[0] += rot * uv.x * eg;
                    }
                }
            }
            block_sum<1>(o, red);
            ratio = reflected / o[0];
        }
        const int account = ((des_cov == 0.0) && (ads_cov > 0.0) && (reflected == 0.0)) ? 1 : 0;
        d4 wdes = wall;
        wdes.s0 = 2 * des_cov * sqrt(kPi * wall.s3);
        const double keep = 1 - (1 - ratio);
#pragma unroll
        for (int k = 0; k < NPT; k++) {
            const int v = threadIdx.x + kT * k;
            if (v < C.NV) {
                d2 uv = face_velocity(C, v, n);
                if (leaves_wall(uv.x, rot)) {
                    d2 m = maxwell(C, wdes, uv, v);
                    double xg = keep * (double)ff[k].x + m.x, xh = keep * (double)ff[k].y + m.y;
                    if (account) { xg = xg - xg; xh = xh - xh; }
                    ff[k] = P2<T>{(T)xg, (T)xh};
                }
            }
        }
    }
    // wallFlux (clUGKS.c:910-1011) and wallMassEnergyFluxes (:1019-1147)
    void *fbuf = fid == SOUTH ? B.fluxS : B.fluxW;
    double m4[12];
#pragma unroll
    for (int q = 0; q < 12; q++) m4[q] = 0.0;
#pragma unroll
    for (int k = 0; k < NPT; k++) {
        const int v = threadIdx.x + kT * k;
        if (v < C.NV) {
            d2 uv = face_velocity(C, v, n);
            double wg = (double)ff[k].x, wh = (double)ff[k].y;
            double Fg = ((uv.x * wg) * L) * dt, Fh = ((uv.x * wh) * L) * dt;
            st<T>(fbuf, g.own * (size_t)C.NV + v, (T)Fg, (T)Fh);
            double u2 = uv.x * uv.x + uv.y * uv.y;
            m4[0] += uv.x * wg;
            m4[1] += uv.x * uv.x * wg;
            m4[2] += uv.x * uv.y * wg;
            m4[3] += 0.5 * uv.x * (u2 * wg + wh);
            int o = leaves_wall(uv.x, rot);
            double fgs = (double)(T)Fg, fhs = (double)(T)Fh;
            m4[4 + o] += fgs;
            m4[6 + o] += uv.x * fgs;
            m4[8 + o] += uv.y * fgs;
            m4[10 + o] += 0.5 * (u2 * fgs + fhs);
        }
    }
    block_sum<12>(m4, red);
    if (threadIdx.x == 0) {
        d2 gl = to_global(mk2(m4[1], m4[2]), n);
        (fid == SOUTH ? B.fmS : B.fmW)[g.own] = (dt * L) * mk4(m4[0], gl.x, gl.y, m4[3]);
        double *o = B.wall_fluxes + ((size_t)ci * 4 + face) * 8;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            o[2 * q] = (double)(-rot) * m4[4 + 2 * q];
            o[2 * q + 1] = (double)rot * m4[5 + 2 * q];
        }
    }
    (void)sc;
}

// ======================================================= cell moments ====
// M = sum_v (g, e_x g, e_y g, (|e|^2 g + h) / 2) of one cell in the global
// frame: the getAL/getAR moments of its faces after rotation into the face
// frame.  Kept for every cell (ghosts included) by the update and the ghost
// fill; k_cellmom recomputes them all at the start of each run.
template <typename T>
__device__ __forceinline__ void add_moments(d2 e, T gx, T hx, double (&m)[4]) {
    // The update (where g and h are products) and the refresh (where they are
    // loads) must round alike, or a step would depend on whether the moments
    // came from the last update or from a refresh: no contraction here (it
    // would fuse the sum into the multiply that made g), explicit FMAs only.
#pragma clang fp contract(off)
    const double g = (double)gx, h = (double)hx;
    const double c2 = fma(e.x, e.x, e.y * e.y);
    m[0] += g;
    m[1] = fma(e.x, g, m[1]);
    m[2] = fma(e.y, g, m[2]);
    m[3] = fma(0.5, fma(c2, g, h), m[3]);
}
template <typename T>
__device__ __forceinline__ void add_moments(const Consts &C, int v, T gx, T hx, double (&m)[4]) {
    add_moments<T>(quad_at(C, v), gx, hx, m);
}

// A cell's four moments summed over the wavefront (all 64 lanes active) and
// stored: the first lane of row r writes component r.  The update, the ghost
// fills and the per-call refresh all store through here, so a cell's
// moments come out bit for bit the same whichever of them wrote them.
__device__ __forceinline__ void store_moments(d4 *dst, const double (&m)[4], int lane) {
    const double r = wave_sum4_rows(m);
    if ((lane & 15) == 0) ((double *)dst)[lane >> 4] = r;
}

template <typename T>
__global__ void __launch_bounds__(kT) k_cellmom(const FBlock *__restrict__ blocks, const long long *__restrict__ off,
                                                int nb, long long nall) {
    const long long item = xcd_remap(blockIdx.x, gridDim.x) * kWavesT + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= nall) return;
    const int lane = threadIdx.x & 63;
    const int b = find_block(off, nb, item);
    const FBlock &B = blocks[b];
    const Consts &C = B.c;
    const long long cell = item - off[b];
    const P2<T> *F = (const P2<T> *)B.f + cell * C.NV;
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int v = lane; v < C.NV; v += 64) {
        const P2<T> x = F[v];
        add_moments<T>(C, v, x.x, x.y, m);
    }
    store_moments(B.cmom + cell, m, lane);
}

// ============================================================ k_update ====
// one wavefront per cell: updateMacro, UGKS_update and the cell's moments.
// A workgroup takes a 2 x 2 tile of cells (tiles j-fastest), so the north
// S flux and the east W flux of half its cells are the other waves' own.
template <typename T>
__global__ void __launch_bounds__(kT) k_update(const FBlock *__restrict__ blocks, const long long *__restrict__ off,
                                               int nb, long long ntiles, const DtState *__restrict__ ds) {
    __shared__ double tabs[kWavesT][4][64];
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long long tile = xcd_remap(blockIdx.x, gridDim.x);
    if (tile >= ntiles) return;
    const int lane = threadIdx.x & 63;
    const int b = find_block(off, nb, tile);
    const FBlock &B = blocks[b];
    const Consts &C = B.c;
    const long long t = tile - off[b];
    const int nrp = (C.nj + 1) >> 1;
    const int mi = 2 * (int)(t / nrp) + (wv >> 1), mj = 2 * (int)(t % nrp) + (wv & 1);
    if (mi >= C.ni || mj >= C.nj) return;
    const long long cell = (long long)mi * C.nj + mj;
    const int gi = mi + C.G, gj = mj + C.G;
    const size_t c0 = (size_t)gi * C.NJ + gj;
    const double dt = ds[B.cs].dt_up;
    // updateMacro (clUGKS.c:1158-1187)
    const d4 pold = B.macro[cell];
    const double A = B.area[c0];
    const d4 dW = ((B.fmS[c0] - B.fmS[c0 + 1]) + B.fmW[c0]) - B.fmW[c0 + C.NJ];
    const d4 p = to_primary(C, to_conserved(C, pold) + dW / A);
    if (lane == 0) {
        B.macro_new[cell] = p;
        B.residual[cell] = pold;
    }
    // UGKS_update (clUGKS.c:1189-1257)
    const double tau_old = relax_time(C, pold), tau = relax_time(C, p);
    const d2 Q = B.Q[cell];
    const bool sep = C.nut > 0;
    double (*tab)[64] = tabs[wv];
    if (sep) {
        for (int k = lane; k < 2 * (C.nut + C.nvt); k += 64) {
            const int which = k >= C.nut + C.nvt, r = which ? k - (C.nut + C.nvt) : k;
            const d4 &pp = which ? p : pold;
            const double a = r < C.nut ? C.qu[r] - pp.s1 : C.qv[r - C.nut] - pp.s2;
            tab[2 * which + (r < C.nut ? 0 : 1)][r < C.nut ? r : r - C.nut] = exp(-pp.s3 * (a * a));
        }
    }
    // the tables are private to this wavefront: order its LDS writes before its reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const double rlo = pold.s0 * (pold.s3 / kPi), rln = p.s0 * (p.s3 / kPi);
    const double so = 0.8 * (1 - C.Pr) * (pold.s3 * pold.s3) / pold.s0, sn = 0.8 * (1 - C.Pr) * (p.s3 * p.s3) / p.s0;
    const double hfo = C.K / (2.0 * pold.s3), hfn = C.K / (2.0 * p.s3);
    const double rA = 1.0 / A, rden = 1.0 / (1.0 + 0.5 * dt / tau);
    const double htn = (0.5 * dt) / tau, hto = (0.5 * dt) / tau_old;
    // the per-cell constants of the velocity loop in scalar registers: the FP32
    // update then needs 74 VGPRs (six waves per SIMD) instead of 104, the FP64 one 84
    // (five); held to 72 for seven waves, the FP32 one spills and runs slower
    const double po1 = uniform_d(pold.s1), po2 = uniform_d(pold.s2), po3 = uniform_d(pold.s3);
    const double pn1 = uniform_d(p.s1), pn2 = uniform_d(p.s2), pn3 = uniform_d(p.s3);
    const double Qx = uniform_d(Q.x), Qy = uniform_d(Q.y);
    const double urlo = uniform_d(rlo), urln = uniform_d(rln), uso = uniform_d(so), usn = uniform_d(sn);
    const double uhfo = uniform_d(hfo), uhfn = uniform_d(hfn), urA = uniform_d(rA), urden = uniform_d(rden);
    const double uhtn = uniform_d(htn), uhto = uniform_d(hto);
    // this cell's velocities in f and in the four face fluxes: uniform bases, 32-bit lane offsets
    const long long NV = C.NV;
    AS1 V2<T> *fc = (AS1 V2<T> *)B.f + c0 * NV;
    const AS1 V2<T> *fS = (const AS1 V2<T> *)B.fluxS + c0 * NV;
    const AS1 V2<T> *fW = (const AS1 V2<T> *)B.fluxW + c0 * NV;
    const unsigned nS = (unsigned)NV, nE = (unsigned)(C.NJ * NV);
    // tensor indices of v = lane + 64 k, advanced without integer division
    int iu = 0, jv = 0, diu = 0, djv = 0;
    // the grid width in a register: the flux stores could alias the block
    // record, so C.nvt would otherwise be loaded again in every iteration
    const int nvt = __builtin_amdgcn_readfirstlane(C.nvt);
    if (sep) { iu = lane / nvt; jv = lane - iu * nvt; diu = 64 / nvt; djv = 64 - diu * nvt; }
    bool nan_seen = false;
    double mom[4] = {0.0, 0.0, 0.0, 0.0};
    // the velocity and weight of the next iteration are loaded one iteration
    // ahead, and this velocity's f and face fluxes before its Maxwellians, so
    // the loads of an iteration overlap its arithmetic instead of stalling it
    d2 enext = lane < C.NV ? quad_at(C, lane) : d2{0.0, 0.0};
    double wnext = lane < C.NV ? weight_at(C, lane) : 0.0;
    for (int v = lane; v < C.NV; v += 64) {
        const V2<T> fo = fc[v], sS = fS[v], nS_ = fS[v + nS], wW = fW[v], eW = fW[v + nE];
        const d2 e = enext;
        const double w = wnext;
        {   // the next velocity's table entries (this one's again on the last round: no branch)
            const int vn = v + 64 < C.NV ? v + 64 : v;
            enext = quad_at(C, vn);
            wnext = weight_at(C, vn);
        }
        const double ox = e.x - po1, oy = e.y - po2, nx = e.x - pn1, ny = e.y - pn2;
        const double co = ox * ox + oy * oy, cn = nx * nx + ny * ny;
        double Eo, En;
        if (sep) {
            Eo = tab[0][iu] * tab[1][jv];
            En = tab[2][iu] * tab[3][jv];
            iu += diu; jv += djv;
            if (jv >= nvt) { jv -= nvt; iu++; }
        } else {
            Eo = exp(-po3 * co);
            En = exp(-pn3 * cn);
        }
        const double gO = w * (urlo * Eo), hO = gO * uhfo;
        const double shO = uso * (ox * Qx + oy * Qy), ptO = 2 * po3 * co;
        const double feqOg = gO + shO * (ptO - 4) * gO, feqOh = hO + shO * (ptO - 2) * hO;
        const double gN = w * (urln * En), hN = gN * uhfn;
        const double shN = usn * (nx * Qx + ny * Qy), ptN = 2 * pn3 * cn;
        const double feqNg = gN + shN * (ptN - 4) * gN, feqNh = hN + shN * (ptN - 2) * hN;
        const double flg = (((double)sS.x - (double)nS_.x) + (double)wW.x - (double)eW.x) * urA;
        const double flh = (((double)sS.y - (double)nS_.y) + (double)wW.y - (double)eW.y) * urA;
        const double fg = (double)fo.x, fh = (double)fo.y;
        const double ng = (fg + flg + (uhtn * feqNg + uhto * (feqOg - fg))) * urden;
        const double nh = (fh + flh + (uhtn * feqNh + uhto * (feqOh - fh))) * urden;
        const T tg = (T)ng, th = (T)nh;
        if (isnan(tg) || isnan(th)) nan_seen = true;
        stream_store<T>(fc + v, V2<T>{tg, th});
        // the new cell moments (the getAL / getAR inputs of the next step)
        add_moments<T>(e, tg, th, mom);
    }
    if (nan_seen) B.flag[0] = ERR_NAN;
    store_moments(B.cmom + c0, mom, lane);
}

// ============================================================ k_dt =========
__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

__global__ void __launch_bounds__(kT) k_dt(const FBlock *__restrict__ blocks, const long long *__restrict__ off, int nb,
                                           const DtState *__restrict__ ds, unsigned long long *maxbits) {
    const long long t = (long long)blockIdx.x * kT + threadIdx.x;
    const long long total = off[nb];
    double val = 0.0;
    int b = -1;
    if (t < total) {
        b = find_block(off, nb, t);
        const FBlock &B = blocks[b];
        const Consts &C = B.c;
        const long long cell = t - off[b];
        const int mi = (int)(cell / C.nj), mj = (int)(cell % C.nj);
        const d4 m = B.macro_new[cell];
        B.macro[cell] = m;
        const int NY = C.NJ + 1, G = C.G;
        d2 a = B.xy[(size_t)(mi + 1 + G) * NY + mj + G], bq = B.xy[(size_t)(mi + 1 + G) * NY + mj + 1 + G];
        d2 c = B.xy[(size_t)(mi + G) * NY + mj + 1 + G], d = B.xy[(size_t)(mi + G) * NY + mj + G];
        double dx = fmax(fmax(fmax(fabs(a.x - bq.x), fabs(c.x - bq.x)), fmax(fabs(c.x - d.x), fabs(a.x - d.x))),
                         fmax(fabs(c.x - a.x), fabs(bq.x - d.x)));
        double dy = fmax(fmax(fmax(fabs(a.y - bq.y), fabs(c.y - bq.y)), fmax(fabs(c.y - d.y), fabs(a.y - d.y))),
                         fmax(fabs(c.y - a.y), fabs(bq.y - d.y)));
        double sos = sqrt(0.5 * C.gam / m.s3);
        double lo_u = fmin(C.umax, fabs(m.s1) + sos), lo_v = fmin(C.vmax, fabs(m.s2) + sos);
        double hi_u = fmax(C.umax, fabs(m.s1) + sos), hi_v = fmax(C.vmax, fabs(m.s2) + sos);
        double scale = (erf(10.0 * (ds[B.cs].Kn - 1e-3)) + 1) / 2.0;
        double u = scale * (hi_u - lo_u) + lo_u, v = scale * (hi_v - lo_v) + lo_v;
        val = (dx * u + dy * v) / B.area[(size_t)(mi + G) * C.NJ + mj + G];
        B.time_step[cell] = val;
    }
    // positive doubles order like their bit patterns: max via integer max
    unsigned long long bits = dbits(val);
    const int b0 = __shfl(b, 0, 64);
    bool uniform = true;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) uniform &= (__shfl_xor(b, o, 64) == b);
    uniform = __all(uniform);
    if (uniform) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            unsigned long long x = __shfl_xor(bits, o, 64);
            bits = x > bits ? x : bits;
        }
        if ((threadIdx.x & 63) == 0 && b0 >= 0) atomicMax(&maxbits[b0], bits);
    } else if (b >= 0) {
        atomicMax(&maxbits[b], bits);
    }
}

// the global time step of each case (UGKSBlock.getDT + UGKSim.updateTimeStep),
// applied once at the end of the step and once more when the next step's flux
// loop re-evaluates it (ugks_sim.py:430-441); time advances by the first
// value.  One thread per case; the blocks of case c are first[c] .. first[c+1]-1.
__global__ void k_dtfin(DtState *dss, unsigned long long *maxbits, double *max_dt, const int *first, int ncases) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncases) return;
    DtState *ds = dss + c;
    double M = __builtin_inf();
    for (int b = first[c]; b < first[c + 1]; b++) {
        double mx = __longlong_as_double((long long)maxbits[b]);
        double md = ds->CFL / mx;
        if (ds->suggest_dt > 0.0 && ds->suggest_dt < md) md = ds->suggest_dt;
        max_dt[b] = md;
        M = fmin(M, md);
        maxbits[b] = 0ull;
    }
    if (ds->force_time) {       // dt held at the value that lands on max_time (ugks_sim.py:292, :673-675)
        ds->time += ds->dt;
        ds->step += 1;
        return;
    }
    const double lim = 1.0 + ds->delta_dt;
    // updateTimeStep after the end-of-step calcTimeStep (ugks_sim.py:457-459),
    // from the dt the update used; time advances by the result (:491)
    const double d1 = limited_dt(ds->dt_up, M, lim);
    ds->time += d1;
    ds->step += 1;
    ds->dt = d1;
    // the next step's flux loop (ugks_sim.py:430-446): the flux with d1, dt
    // re-evaluated; if it moved, the flux again with d2 and dt re-evaluated
    // once more for the update ("Overshot" if that happens on gdata.step 1)
    const double d2 = limited_dt(d1, M, lim);
    if (d2 == d1) {
        ds->dt_flux = ds->dt_up = d1;
    } else {
        ds->dt_flux = d2;
        ds->dt_up = limited_dt(d2, M, lim);
        if (ds->step == 1) ds->overshoot = 1;
    }
}

// paraBC (clKernel.c:141-218) at the start of a step, for the blocks whose
// wall values are expressions in t (ugks_sim.py:425-427): every wall cell of
// block list[blockIdx.x] at its case's time
__global__ void __launch_bounds__(kT) k_parabc(const FBlock *__restrict__ blocks, const int *__restrict__ list,
                                               const DtState *__restrict__ ds) {
    const FBlock &B = blocks[list[blockIdx.x]];
    const Consts &C = B.c;
    const int ci = blockIdx.y * kT + threadIdx.x;
    const int wl = C.ni > C.nj ? C.ni : C.nj;
    if (ci >= wl) return;
    const double t = ds[B.cs].time;
    d4 *wp = (d4 *)B.wall_prop;
    for (int w = 0; w < 4; ++w) {
        const int along = (w == GNORTH || w == GSOUTH) ? C.ni : C.nj;
        if (ci < along) wp[(size_t)ci * 4 + w] = para_wall(C, w, B.para[(size_t)ci * 4 + w], t);
    }
}

// edge-exchange corner check (clGhostRoutines.c:491-512): a joined face whose
// end vertices do not coincide with the partner face's (1e-6) has its ghost row
// flipped, if its orientation flag asks for it.  Thread 4 b + k checks face k
// of block b from the vertices as they stand, at every run call.
__global__ void k_flipidx(FBlock *__restrict__ blocks, int nb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * nb) return;
    FBlock &F = blocks[t >> 2];
    const int k = t & 3;
    int flip = 0;
    if ((F.bc[k] == 0 || F.bc[k] == 5) && F.ori[k]) {
        const FBlock &O = blocks[F.other[k]];
        const int G = F.c.G, NI = F.c.NI, NJ = F.c.NJ, NIB = O.c.NI, NJB = O.c.NJ;
        int A[4], Bv[4];
        switch (k) {
        case GNORTH: A[0] = G; A[1] = NJ - G; A[2] = NI - G; A[3] = NJ - G; break;
        case GEAST: A[0] = NI - G; A[1] = G; A[2] = NI - G; A[3] = NJ - G; break;
        case GSOUTH: A[0] = G; A[1] = G; A[2] = NI - G; A[3] = G; break;
        default: A[0] = G; A[1] = G; A[2] = G; A[3] = NJ - G; break;
        }
        switch (F.oface[k]) {
        case GNORTH: Bv[0] = G; Bv[1] = NJB - G; Bv[2] = NIB - G; Bv[3] = NJB - G; break;
        case GEAST: Bv[0] = NIB - G; Bv[1] = G; Bv[2] = NIB - G; Bv[3] = NJB - G; break;
        case GSOUTH: Bv[0] = G; Bv[1] = G; Bv[2] = NIB - G; Bv[3] = G; break;
        default: Bv[0] = G; Bv[1] = G; Bv[2] = G; Bv[3] = NJB - G; break;
        }
        const d2 a0 = F.xy[(size_t)A[0] * (NJ + 1) + A[1]], a1 = F.xy[(size_t)A[2] * (NJ + 1) + A[3]];
        const d2 b0 = O.xy[(size_t)Bv[0] * (NJB + 1) + Bv[1]], b1 = O.xy[(size_t)Bv[2] * (NJB + 1) + Bv[3]];
        auto ne = [](double a, double c) { return !(fabs(a - c) <= 1e-6); };
        flip = ne(a0.x, b0.x) || ne(a0.y, b0.y) || ne(a1.x, b1.x) || ne(a1.y, b1.y);
    }
    F.flipidx[k] = flip;
}

// ============================================================ k_ghost ======
struct GhostItem { int block, face; long long start; };

template <typename T>
__global__ void __launch_bounds__(kT) k_ghost(const FBlock *__restrict__ blocks, const GhostItem *__restrict__ items,
                                              int nitems, long long nghosts) {
    const long long item = (long long)blockIdx.x * kWavesT + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (item >= nghosts) return;
    const int lane = threadIdx.x & 63;
    const GhostItem GI = items[find_item(items, nitems, item)];
    const FBlock &B = blocks[GI.block];
    const Consts &C = B.c;
    const int face = GI.face;
    const long long loc = item - GI.start;
    const bool ns = (face == GNORTH || face == GSOUTH);
    const int sj = ns ? C.G : C.nj;
    const int li = (int)(loc / sj), lj = (int)(loc % sj);
    int gi, gj;
    switch (face) {
    case GNORTH: gi = li + C.G; gj = lj + C.NJ - C.G; break;
    case GEAST: gi = li + C.NI - C.G; gj = lj + C.G; break;
    case GSOUTH: gi = li + C.G; gj = lj; break;
    default: gi = li; gj = lj + C.G; break;
    }
    int ei, ej;
    switch (face) {
    case GNORTH: ei = gi; ej = C.NJ - C.G - 1; break;
    case GEAST: ei = C.NI - C.G - 1; ej = gj; break;
    case GSOUTH: ei = gi; ej = C.G; break;
    default: ei = C.G; ej = gj; break;
    }
    const int ci = ns ? li : lj;
    const size_t NV = C.NV;
    const size_t dst = ((size_t)gi * C.NJ + gj) * NV;
    const int type = B.bc[face];
    // the ghost cell's moments for the faces it borders (uniform exits: the whole workgroup reaches these)
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    auto put = [&](int v, P2<T> x) {
        ((P2<T> *)B.f)[dst + v] = x;
        add_moments<T>(C, v, x.x, x.y, m);
    };
    auto finish = [&]() { store_moments(B.cmom + (size_t)gi * C.NJ + gj, m, lane); };
    if (type == 0 || type == 5) {           // edgeExchange (clGhostRoutines.c:407-538)
        const FBlock &O = blocks[B.other[face]];
        int iB = gi, jB = gj;
        if (B.flipidx[face]) {
            if (!ns) jB = (C.NJ - 1) - jB;
            else iB = (C.NI - 1) - iB;
        }
        const int of = B.oface[face], NIB = O.c.NI, NJB = O.c.NJ, G = C.G, NI = C.NI, NJ = C.NJ;
        // getSwapIndex for the pure shifts (E-W, N-S) and the rotations (clGhostRoutines.c:47-138)
        auto r90 = [&](int nrot) {
            int Ni = NI, Nj = NJ;
            for (int a = 0; a < nrot; a++) { int t = iB; iB = Nj - jB - 1; jB = t; t = Ni; Ni = Nj; Nj = t; }
        };
        if (face == GEAST) {
            if (of == GEAST) { iB = (NI - 1) - iB; iB += NIB - 2 * G; } else if (of == GWEST) iB -= NI - 2 * G;
            else if (of == GNORTH) { r90(3); jB += NJB - 2 * G; } else { r90(1); jB -= NI - 2 * G; }
        } else if (face == GWEST) {
            if (of == GWEST) { iB = (NI - 1) - iB; iB -= NI - 2 * G; } else if (of == GEAST) iB += NIB - 2 * G;
            else if (of == GNORTH) { r90(1); jB += NJB - 2 * G; } else { r90(3); jB -= NI - 2 * G; }
        } else if (face == GNORTH) {
            if (of == GNORTH) { jB = (NJ - 1) - jB; jB += NJB - 2 * G; } else if (of == GSOUTH) jB -= NJ - 2 * G;
            else if (of == GEAST) { r90(1); iB += NIB - 2 * G; } else { r90(3); iB -= NJ - 2 * G; }
        } else {
            if (of == GSOUTH) { jB = (NJ - 1) - jB; jB -= NJ - 2 * G; } else if (of == GNORTH) jB += NJB - 2 * G;
            else if (of == GEAST) { r90(3); iB += NIB - 2 * G; } else { r90(1); iB -= NJ - 2 * G; }
        }
        if (iB < 0 || iB >= NIB || jB < 0 || jB >= NJB) return;
        const size_t src = ((size_t)iB * NJB + jB) * NV;
        const int fl = B.flip[face];
        for (int v = lane; v < C.NV; v += 64) {
            int v2 = v;
            if (fl == FLIP_NS) v2 = (int)C.mNS[v];
            else if (fl == FLIP_D) v2 = (int)C.mD[v];
            else if (fl == FLIP_EW) v2 = (int)C.mEW[v];
            put(v, ld<T>(O.f, src + v2));
        }
        finish();
        return;
    }
    if (type == 1) {                         // edgeExtrapolate
        const size_t src = ((size_t)ei * C.NJ + ej) * NV;
        for (int v = lane; v < C.NV; v += 64) put(v, ld<T>(B.f, src + v));
        finish();
        return;
    }
    if (type == 3 || type == 9) {            // edgeMirror / edgeBounceBack
        int mi, mj;
        switch (face) {
        case GNORTH: mi = gi; mj = 2 * (C.NJ - C.G) - gj - 1; break;
        case GEAST: mi = 2 * (C.NI - C.G) - gi - 1; mj = gj; break;
        case GSOUTH: mi = gi; mj = 2 * C.G - gj - 1; break;
        default: mi = 2 * C.G - gi - 1; mj = gj; break;
        }
        const size_t src = ((size_t)mi * C.NJ + mj) * NV;
        for (int v = lane; v < C.NV; v += 64) {
            int mv = (int)(type == 9 ? C.mD[v] : (ns ? C.mNS[v] : C.mEW[v]));
            put(v, ld<T>(B.f, src + mv));
        }
        finish();
        return;
    }
    if (type == 4 || type == 8) {            // edgeConstGrad
        int ei0 = ei, ej0 = ej;
        double m;
        switch (face) {
        case GNORTH: ej0 = ej - 1; m = (double)(gj - ej); break;
        case GEAST: ei0 = ei - 1; m = (double)(gi - ei); break;
        case GSOUTH: ej0 = ej + 1; m = (double)(ej - gj); break;
        default: ei0 = ei + 1; m = (double)(ei - gi); break;
        }
        const size_t s0 = ((size_t)ei0 * C.NJ + ej0) * NV, s1 = ((size_t)ei * C.NJ + ej) * NV;
        const T mm = (T)m;
        for (int v = lane; v < C.NV; v += 64) {
            P2<T> f0 = ld<T>(B.f, s0 + v), f1 = ld<T>(B.f, s1 + v);
            put(v, P2<T>{f1.x + mm * (f1.x - f0.x), f1.y + mm * (f1.y - f0.y)});
        }
        finish();
        return;
    }
    // edgeConstant (2), edgeInflow (6), edgeOutflow (7): a Shakhov equilibrium
    d4 pr;
    d2 Q = mk2(0.0, 0.0);
    const d4 wp = B.wall_prop[(size_t)ci * 4 + face];
    if (type == 2) {
        pr = wp;
    } else {
        const int noi = (face == GNORTH || face == GEAST) ? gi : ei, noj = (face == GNORTH || face == GEAST) ? gj : ej;
        Q = B.Q[(size_t)(ei - C.G) * C.nj + (ej - C.G)];
        pr = B.macro[(size_t)(ei - C.G) * C.nj + (ej - C.G)];
        d2 l = to_local(mk2(pr.s1, pr.s2), B.normal[((size_t)noi * C.NJ + noj) * 2 + (ns ? SOUTH : WEST)]);
        pr.s1 = l.x; pr.s2 = l.y;
        if (type == 6) { pr.s0 = wp.s0; pr.s3 = wp.s3; }
        else pr.s3 = pr.s0 / (wp.s0 / wp.s3);
    }
    for (int v = lane; v < C.NV; v += 64) {
        d2 fe = shakhov_eq(C, pr, Q, C.quad[v], v);
        put(v, P2<T>{(T)fe.x, (T)fe.y});
    }
    finish();
}

// FP64 -> FP32 conversion of f (the mixed-precision state starts from the FP64 initial state)
__global__ void k_to_f32(const double *src, float *dst, long long n) {
    long long t = (long long)blockIdx.x * kT + threadIdx.x;
    if (t < n) dst[t] = (float)src[t];
}
__global__ void k_to_f64(const float *src, double *dst, long long n) {
    long long t = (long long)blockIdx.x * kT + threadIdx.x;
    if (t < n) dst[t] = (double)src[t];
}

// plain streaming copy (16 B per lane per iteration, grid-stride)
__global__ void __launch_bounds__(kT) k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, long long n) {
    for (long long t = (long long)blockIdx.x * kT + threadIdx.x; t < n; t += (long long)gridDim.x * kT) dst[t] = src[t];
}

}  // namespace

// ============================================================ host plan ====
struct ugks_step_plan {
    int nb, precision, npt;
    int device;
    FBlock *d_blocks;
    long long *d_face_off, *d_cell_off, *d_all_off;   // prefix offsets: interior faces, cells, all cells
    long long *d_tile_off;                              // prefix offsets of the update's 2 x 2 cell tiles
    long long nall, ntiles;
    int fa_waves, fc_waves;               // workgroup size of the face passes, in wavefronts (1, 2 or 4)
    int method;                           // the blocks' common reconstruction method, -1 if they differ
    d4 *d_cmom;                                         // cell moments of every block, back to back
    WallItem *d_walls;
    GhostItem *d_ghosts;
    int nwall_items, nghost_items;
    long long nfaces, nwalls, ncells, nghosts;
    int ncases;
    DtState *d_state;                   // [ncases]
    int *d_case_first;                  // [ncases + 1] first block of each case
    double *d_fsum, *d_fstate;          // per interior face: raw sums, face state
    FaceRec *d_frec;                    // per interior face: block, side, owner, distances (plain order)
    int *d_udf_blocks;                  // blocks whose wall values are re-evaluated every step
    int nudf, udf_wl;                   // their count and largest wall length
    int *d_tmap;                        // moment-pass item -> plain face index (tiled order)
    hipStream_t cap_stream;             // stream the step graph is captured on
    hipGraph_t graph;
    hipGraphExec_t graph_exec;
    unsigned long long *d_maxbits;
    double *d_max_dt;
    std::vector<FBlock> host_blocks;
};

namespace {
template <typename T, int FW, int M> void launch_faceA(ugks_step_plan *p, hipStream_t s) {
    hipLaunchKernelGGL((k_faceA<T, FW, M>), dim3((unsigned)((p->nfaces + FW - 1) / FW)), dim3(64 * FW), 0, s,
                       p->d_blocks, (const FaceRec *)p->d_frec, (const int *)p->d_tmap, p->nfaces, p->d_fsum);
}
template <typename T, int FW, int M> void launch_faceC(ugks_step_plan *p, hipStream_t s) {
    hipLaunchKernelGGL((k_faceC<T, FW, M>), dim3((unsigned)((p->nfaces + FW - 1) / FW)), dim3(64 * FW), 0, s,
                       p->d_blocks, (const FaceRec *)p->d_frec, p->nfaces, (const double *)p->d_fstate);
}
// the face passes specialised for the plan's one reconstruction method (or generic)
template <typename T, int FW> void faceA_m(ugks_step_plan *p, hipStream_t s) {
    switch (p->method) {
    case 0: launch_faceA<T, FW, 0>(p, s); break;
    case 1: launch_faceA<T, FW, 1>(p, s); break;
    case 2: launch_faceA<T, FW, 2>(p, s); break;
    default: launch_faceA<T, FW, -1>(p, s); break;
    }
}
template <typename T, int FW> void faceC_m(ugks_step_plan *p, hipStream_t s) {
    switch (p->method) {
    case 0: launch_faceC<T, FW, 0>(p, s); break;
    case 1: launch_faceC<T, FW, 1>(p, s); break;
    case 2: launch_faceC<T, FW, 2>(p, s); break;
    default: launch_faceC<T, FW, -1>(p, s); break;
    }
}
template <typename T> void faceA(ugks_step_plan *p, hipStream_t s, int fw) {
    if (fw == 1) faceA_m<T, 1>(p, s); else if (fw == 2) faceA_m<T, 2>(p, s); else faceA_m<T, 4>(p, s);
}
template <typename T> void faceC(ugks_step_plan *p, hipStream_t s, int fw) {
    if (fw == 1) faceC_m<T, 1>(p, s); else if (fw == 2) faceC_m<T, 2>(p, s); else faceC_m<T, 4>(p, s);
}

// ev != nullptr: record ev[0..kLaunches] around the launches (per-kernel timing)
constexpr int kLaunches = UGKS_STEP_NKERNELS;
template <typename T, int NPT> int launch_step(ugks_step_plan *p, hipStream_t s, hipEvent_t *ev) {
    const unsigned gt = (unsigned)((p->nfaces + kT - 1) / kT);             // one thread per face
    if (p->nudf > 0)
        hipLaunchKernelGGL(k_parabc, dim3((unsigned)p->nudf, (unsigned)((p->udf_wl + kT - 1) / kT)), dim3(kT), 0, s,
                           p->d_blocks, (const int *)p->d_udf_blocks, (const DtState *)p->d_state);
    if (ev) (void)hipEventRecord(ev[0], s);
    if (p->nfaces > 0)
        faceA<T>(p, s, p->fa_waves);
    if (ev) (void)hipEventRecord(ev[1], s);
    if (p->nfaces > 0)
        hipLaunchKernelGGL(k_faceS, dim3(gt), dim3(kT), 0, s, p->d_blocks, (const FaceRec *)p->d_frec, p->nfaces,
                           (const double *)p->d_fsum, p->d_fstate, (const DtState *)p->d_state);
    if (ev) (void)hipEventRecord(ev[2], s);
    if (p->nfaces > 0)
        faceC<T>(p, s, p->fc_waves);
    if (ev) (void)hipEventRecord(ev[3], s);
    if (p->nwalls > 0)
        hipLaunchKernelGGL((k_wall<T, NPT>), dim3((unsigned)p->nwalls), dim3(kT), 0, s, p->d_blocks, p->d_walls,
                           p->nwall_items, p->d_state);
    if (ev) (void)hipEventRecord(ev[4], s);
    hipLaunchKernelGGL((k_update<T>), dim3((unsigned)p->ntiles), dim3(kT), 0, s, p->d_blocks, p->d_tile_off, p->nb,
                       p->ntiles, p->d_state);
    if (ev) (void)hipEventRecord(ev[5], s);
    hipLaunchKernelGGL(k_dt, dim3((unsigned)((p->ncells + kT - 1) / kT)), dim3(kT), 0, s, p->d_blocks, p->d_cell_off,
                       p->nb, p->d_state, p->d_maxbits);
    if (ev) (void)hipEventRecord(ev[6], s);
    hipLaunchKernelGGL(k_dtfin, dim3((unsigned)((p->ncases + 63) / 64)), dim3(64), 0, s, p->d_state, p->d_maxbits,
                       p->d_max_dt, (const int *)p->d_case_first, p->ncases);
    if (ev) (void)hipEventRecord(ev[7], s);
    if (p->nghosts > 0)
        hipLaunchKernelGGL((k_ghost<T>), dim3((unsigned)((p->nghosts + kWavesT - 1) / kWavesT)), dim3(kT), 0, s,
                           p->d_blocks, p->d_ghosts, p->nghost_items, p->nghosts);
    if (ev) (void)hipEventRecord(ev[8], s);
    return ugks_rt::check_launch("ugks_step_run");
}

// moments of every cell from f as it stands (f may have been changed between
// runs), and the face records from the geometry as it stands
template <typename T> int refresh_moments(ugks_step_plan *p, hipStream_t s) {
    hipLaunchKernelGGL((k_cellmom<T>), dim3((unsigned)((p->nall + kWavesT - 1) / kWavesT)), dim3(kT), 0, s, p->d_blocks,
                       p->d_all_off, p->nb, p->nall);
    if (p->nfaces > 0)
        hipLaunchKernelGGL(k_facerec, dim3((unsigned)((p->nfaces + kT - 1) / kT)), dim3(kT), 0, s, p->d_blocks,
                           p->d_face_off, p->nb, p->nfaces, p->d_frec, p->d_tmap);
    hipLaunchKernelGGL(k_flipidx, dim3((unsigned)((4 * p->nb + kT - 1) / kT)), dim3(kT), 0, s, p->d_blocks, p->nb);
    return ugks_rt::check_launch("ugks_step: cell moments, face records and join corners");
}
int refresh(ugks_step_plan *p, hipStream_t s) {
    return p->precision == UGKS_FP64 ? refresh_moments<double>(p, s) : refresh_moments<float>(p, s);
}

template <typename T> int dispatch_npt(ugks_step_plan *p, hipStream_t s, hipEvent_t *ev) {
    switch (p->npt) {
    case 1: return launch_step<T, 1>(p, s, ev);
    case 2: return launch_step<T, 2>(p, s, ev);
    case 3: return launch_step<T, 3>(p, s, ev);
    case 4: return launch_step<T, 4>(p, s, ev);
    case 5: return launch_step<T, 5>(p, s, ev);
    case 6: return launch_step<T, 6>(p, s, ev);
    case 7: return launch_step<T, 7>(p, s, ev);
    case 8: return launch_step<T, 8>(p, s, ev);
    default: return launch_step<T, 10>(p, s, ev);
    }
}
}  // namespace

namespace {
// The first flux loop of a step whose dt was set from outside (a new plan, or
// set_dt): M is the smallest block maximum time step from the last
// calcTimeStep (0 if unknown: dt is then taken as is).
__host__ __device__ void flux_loop_dts(DtState &d, double M) {
    d.dt_flux = d.dt_up = d.dt;
    if (d.force_time || !(M > 0.0)) return;
    const double lim = 1.0 + d.delta_dt;
    const double d1 = limited_dt(d.dt, M, lim);
    if (d1 == d.dt) return;
    d.dt_flux = d1;
    d.dt_up = limited_dt(d1, M, lim);
    if (d.step == 1) d.overshoot = 1;
}

template <typename T> int autotune(ugks_step_plan *p, hipStream_t s, int reps, int32_t *choice) {
    // The face passes only read f and the step state and rewrite per-step scratch
    // (face sums, face states, fluxes, Q from the same f), so timing them
    // repeatedly leaves the simulation state as it was.
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return ugks_rt::fail(-1, "autotune: event create");
    const int cand[3] = {1, 2, 4};
    float best[2] = {1e30f, 1e30f};
    int pick[2] = {4, 4};
    int rc = 0;
    for (int which = 0; which < 2 && !rc; which++)
        for (int c = 0; c < 3 && !rc; c++) {
            const int fw = cand[c];
            // kC reads kA's results through kS: run the chain once, then time the kernel alone
            faceA<T>(p, s, which == 0 ? fw : p->fa_waves);
            hipLaunchKernelGGL(k_faceS, dim3((unsigned)((p->nfaces + kT - 1) / kT)), dim3(kT), 0, s, p->d_blocks,
                               (const FaceRec *)p->d_frec, p->nfaces, (const double *)p->d_fsum, p->d_fstate,
                               (const DtState *)p->d_state);
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps; r++) {
                if (which == 0) faceA<T>(p, s, fw);
                else faceC<T>(p, s, fw);
            }
            (void)hipEventRecord(e1, s);
            if ((rc = ugks_rt::check_launch("autotune"))) break;
            if (hipEventSynchronize(e1) != hipSuccess) { rc = ugks_rt::fail(-1, "autotune: sync"); break; }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best[which]) { best[which] = ms; pick[which] = fw; }
            if (which == 0 && c == 2) p->fa_waves = pick[0];
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    p->fa_waves = pick[0];
    p->fc_waves = pick[1];
    if (choice) { choice[0] = pick[0]; choice[1] = pick[1]; }
    // a captured graph holds the old launch shapes
    if (p->graph_exec) { (void)hipGraphExecDestroy(p->graph_exec); p->graph_exec = nullptr; }
    if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
    return 0;
}
}  // namespace

namespace {
// One block's device record from its descriptor, buffers and boundary
// conditions, with the checks every plan makes; wallf[k] tells which faces
// are walls (diffuse or adsorbing: their faces leave the interior face list).
int fill_block(FBlock &F, const ugks_block *blk, const ugks_block_buffers &u, const ugks_bc_desc *bc4, int cs,
               int nblocks, const int32_t *case_of_block, int precision, bool (&wallf)[4]) {
    const Consts &C = blk->c;
    memset(&F, 0, sizeof F);
    F.c = C;
    F.cs = cs;
    F.f = u.f; F.fluxS = u.flux_f_S; F.fluxW = u.flux_f_W; F.wdist = u.wall_dist;
    F.fmS = (d4 *)u.flux_macro_S; F.fmW = (d4 *)u.flux_macro_W;
    F.macro = (d4 *)u.macro; F.macro_new = (d4 *)u.macro_new; F.residual = (d4 *)u.residual;
    F.wall_cover = (d4 *)u.wall_cover; F.Q = (d2 *)u.Q;
    F.xy = (const d2 *)u.xy; F.centre = (const d2 *)u.centre; F.mid_side = (const d2 *)u.mid_side;
    F.normal = (const d2 *)u.normal; F.area = (const double *)u.area; F.side_length = (const double *)u.side_length;
    F.wall_prop = (const d4 *)u.wall_prop; F.wall_fluxes = (double *)u.wall_fluxes;
    F.time_step = (double *)u.time_step; F.flag = (double *)u.flag;
    F.para = (const d4 *)u.para;
    if (C.time_udf && !u.para) return ugks_rt::fail(-1, "step plan: time-dependent wall values need para");
    const void *need[] = {u.f, u.flux_f_S, u.flux_f_W, u.flux_macro_S, u.flux_macro_W, u.macro, u.macro_new,
                          u.residual, u.Q, u.xy, u.centre, u.mid_side, u.normal, u.area, u.side_length, u.wall_prop,
                          u.wall_cover, u.wall_fluxes, u.time_step, u.flag};
    for (const void *q : need)
        if (!q) return ugks_rt::fail(-1, "step plan: a required buffer is null");
    for (int k = 0; k < 4; k++) {
        const ugks_bc_desc &d = bc4[k];
        F.bc[k] = d.type; F.other[k] = d.other_block; F.oface[k] = d.other_face; F.ori[k] = d.orientation;
        F.flip[k] = d.flip; F.reflect[k] = d.reflect_type;
        if (d.type < 0 || d.type > 9) return ugks_rt::fail(-1, "step plan: bad BC type");
        if ((d.type == 0 || d.type == 5) && (d.other_block < 0 || d.other_block >= nblocks))
            return ugks_rt::fail(-1, "step plan: exchange partner out of range");
        if ((d.type == 0 || d.type == 5) && case_of_block && case_of_block[d.other_block] != F.cs)
            return ugks_rt::fail(-1, "step plan: exchange partner belongs to another case");
        if (d.type == 8 && !C.cads[k] && C.n_tri[k] <= 0)
            return ugks_rt::fail(-1, "step plan: adsorbing wall without an isotherm");
        wallf[k] = (d.type == 4 || d.type == 8);
    }
    F.ob_S = wallf[GSOUTH] ? 1 : 0;
    const int ot_S = wallf[GNORTH] ? 1 : 0;
    F.ob_W = wallf[GWEST] ? 1 : 0;
    const int ot_W = wallf[GEAST] ? 1 : 0;
    F.nSj = C.nj + 1 - F.ob_S - ot_S;
    F.nS = (int)((long long)C.ni * F.nSj);
    F.nWi = C.ni + 1 - F.ob_W - ot_W;
    F.wl = C.ni > C.nj ? C.ni : C.nj;
    if ((long long)F.nS + (long long)F.nWi * C.nj >= (1LL << 31) - (long long)kTileI * kTileJ * 2)
        return ugks_rt::fail(-2, "step plan: too many faces in one block");
    return ugks_step_block_fits(C.NJ, C.NV, C.method, precision);
}
}  // namespace

extern "C" {

int ugks_step_plan_create_batch(ugks_block *const *blocks, const ugks_block_buffers *bufs, const ugks_bc_desc *bcs,
                                int32_t nblocks, const int32_t *case_of_block, int32_t ncases, int32_t precision,
                                const ugks_dt_desc *dtd, ugks_step_plan **out) {
    if (!blocks || !bufs || !bcs || !dtd || !out || nblocks < 1 || ncases < 1)
        return ugks_rt::fail(-1, "step plan: bad arguments");
    if (precision != UGKS_FP64 && precision != UGKS_MIXED) return ugks_rt::fail(-1, "step plan: bad precision");
    *out = nullptr;
    // blocks of one case are contiguous and cases are numbered in block order
    std::vector<int> case_first(ncases + 1, 0);
    {
        int prev = 0;
        for (int b = 0; b < nblocks; b++) {
            const int c = case_of_block ? case_of_block[b] : 0;
            if (c < 0 || c >= ncases || c < prev || c > prev + 1 || (b == 0 && c != 0))
                return ugks_rt::fail(-1, "step plan: case_of_block must run 0,0,..,1,1,.. up to ncases-1");
            if (c != prev) case_first[c] = b;
            prev = c;
        }
        if (prev != ncases - 1) return ugks_rt::fail(-1, "step plan: a case has no blocks");
        case_first[ncases] = nblocks;
    }
    ugks_step_plan *p = new ugks_step_plan();
    p->nb = nblocks;
    p->ncases = ncases;
    p->fa_waves = p->fc_waves = 4;
    p->method = blocks[0] ? blocks[0]->c.method : -1;
    for (int b = 1; b < nblocks; b++)
        if (blocks[b] && blocks[b]->c.method != p->method) p->method = -1;
    p->precision = precision;
    int maxnv = 0;
    std::vector<long long> face_off(nblocks + 1, 0), cell_off(nblocks + 1, 0), all_off(nblocks + 1, 0),
        tile_off(nblocks + 1, 0);
    std::vector<WallItem> walls;
    std::vector<GhostItem> ghosts;
    std::vector<int> udf_blocks;          // blocks whose wall values depend on time
    long long nw = 0, ng = 0;
    p->host_blocks.resize(nblocks);
    for (int b = 0; b < nblocks; b++) {
        if (!blocks[b]) { delete p; return ugks_rt::fail(-1, "step plan: null block"); }
        const Consts &C = blocks[b]->c;
        FBlock &F = p->host_blocks[b];
        bool wallf[4];
        if (int rc = fill_block(F, blocks[b], bufs[b], bcs + 4 * b, case_of_block ? case_of_block[b] : 0, nblocks,
                                case_of_block, precision, wallf)) {
            delete p;
            return rc;
        }
        if (C.time_udf) udf_blocks.push_back(b);
        face_off[b + 1] = face_off[b] + F.nS + (long long)F.nWi * C.nj;
        cell_off[b + 1] = cell_off[b] + (long long)C.ni * C.nj;
        all_off[b + 1] = all_off[b] + (long long)C.NI * C.NJ;
        tile_off[b + 1] = tile_off[b] + (long long)((C.ni + 1) / 2) * ((C.nj + 1) / 2);
        for (int k = 0; k < 4; k++) {
            int len = (k == GNORTH || k == GSOUTH) ? C.ni : C.nj;
            if (wallf[k]) { walls.push_back(WallItem{b, k, nw}); nw += len; }
            ghosts.push_back(GhostItem{b, k, ng});
            ng += (long long)len * C.G;
        }
        if (C.NV > maxnv) maxnv = C.NV;
    }
    // joined blocks must agree in velocities and ghost width (the corner check itself runs on the device)
    for (int b = 0; b < nblocks; b++) {
        const FBlock &F = p->host_blocks[b];
        for (int k = 0; k < 4; k++) {
            if (!(F.bc[k] == 0 || F.bc[k] == 5)) continue;
            const FBlock &O = p->host_blocks[F.other[k]];
            if (O.c.NV != F.c.NV || O.c.G != F.c.G) {
                delete p; return ugks_rt::fail(-1, "step plan: joined blocks differ in NV or ghost width");
            }
        }
    }
    int npt = (maxnv + kT - 1) / kT;
    if (npt > 10) { delete p; return ugks_rt::fail(-2, "step plan: NV above 2560 is not supported by the fused step"); }
    if (npt == 9) npt = 10;
    p->npt = npt;
    for (int b = 0; b < nblocks; b++)
        if (p->host_blocks[b].c.nut > 64 || p->host_blocks[b].c.nvt > 64) {
            delete p; return ugks_rt::fail(-2, "step plan: more than 64 velocities per axis");
        }
    p->nfaces = face_off[nblocks];
    if (p->nfaces >= (1LL << 31)) { delete p; return ugks_rt::fail(-2, "step plan: 2^31 or more interior faces"); }
    p->ncells = cell_off[nblocks];
    p->nall = all_off[nblocks];
    p->ntiles = tile_off[nblocks];
    p->nwalls = nw;
    p->nghosts = ng;
    p->nwall_items = (int)walls.size();
    p->nghost_items = (int)ghosts.size();
    hipError_t e = hipGetDevice(&p->device);
    size_t bytes_blocks = sizeof(FBlock) * nblocks;
    auto alloc = [&](void **ptr, size_t n) { return e == hipSuccess ? (e = hipMalloc(ptr, n ? n : 8)) : e; };
    (void)alloc((void **)&p->d_blocks, bytes_blocks);
    (void)alloc((void **)&p->d_face_off, sizeof(long long) * (nblocks + 1));
    (void)alloc((void **)&p->d_cell_off, sizeof(long long) * (nblocks + 1));
    (void)alloc((void **)&p->d_all_off, sizeof(long long) * (nblocks + 1));
    (void)alloc((void **)&p->d_tile_off, sizeof(long long) * (nblocks + 1));
    (void)alloc((void **)&p->d_cmom, sizeof(d4) * (size_t)p->nall);
    (void)alloc((void **)&p->d_walls, sizeof(WallItem) * (walls.size() ? walls.size() : 1));
    (void)alloc((void **)&p->d_ghosts, sizeof(GhostItem) * ghosts.size());
    (void)alloc((void **)&p->d_state, sizeof(DtState) * ncases);
    (void)alloc((void **)&p->d_case_first, sizeof(int) * (ncases + 1));
    (void)alloc((void **)&p->d_maxbits, sizeof(unsigned long long) * nblocks);
    (void)alloc((void **)&p->d_max_dt, sizeof(double) * nblocks);
    (void)alloc((void **)&p->d_fsum, sizeof(double) * kFS * (size_t)p->nfaces);
    (void)alloc((void **)&p->d_fstate, sizeof(double) * kFT * (size_t)p->nfaces);
    (void)alloc((void **)&p->d_frec, sizeof(FaceRec) * (size_t)p->nfaces);
    (void)alloc((void **)&p->d_tmap, sizeof(int) * (size_t)p->nfaces);
    (void)alloc((void **)&p->d_udf_blocks, sizeof(int) * (udf_blocks.size() ? udf_blocks.size() : 1));
    p->nudf = (int)udf_blocks.size();
    p->udf_wl = 0;
    for (int b : udf_blocks) p->udf_wl = std::max(p->udf_wl, std::max(p->host_blocks[b].c.ni, p->host_blocks[b].c.nj));
    if (e != hipSuccess) { ugks_step_plan_destroy(p); return ugks_rt::hip_fail(e, "step plan: hipMalloc"); }
    std::vector<DtState> st(ncases);
    for (int c = 0; c < ncases; c++) {
        st[c] = DtState{dtd[c].dt, dtd[c].dt, dtd[c].dt, dtd[c].time, dtd[c].CFL, dtd[c].delta_dt, dtd[c].suggest_dt,
                        dtd[c].Kn_eff, dtd[c].step, 0, dtd[c].force_time, 0};
        flux_loop_dts(st[c], dtd[c].max_dt);
    }
    std::vector<unsigned long long> zeros(nblocks, 0ull);
    if (e == hipSuccess)
        for (int b = 0; b < nblocks; b++) p->host_blocks[b].cmom = p->d_cmom + all_off[b];
    if (e == hipSuccess) e = hipMemcpy(p->d_blocks, p->host_blocks.data(), bytes_blocks, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_face_off, face_off.data(), sizeof(long long) * (nblocks + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_cell_off, cell_off.data(), sizeof(long long) * (nblocks + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_all_off, all_off.data(), sizeof(long long) * (nblocks + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_tile_off, tile_off.data(), sizeof(long long) * (nblocks + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && p->nudf)
        e = hipMemcpy(p->d_udf_blocks, udf_blocks.data(), sizeof(int) * udf_blocks.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && walls.size())
        e = hipMemcpy(p->d_walls, walls.data(), sizeof(WallItem) * walls.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_ghosts, ghosts.data(), sizeof(GhostItem) * ghosts.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_state, st.data(), sizeof(DtState) * ncases, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->d_case_first, case_first.data(), sizeof(int) * (ncases + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_maxbits, zeros.data(), sizeof(unsigned long long) * nblocks, hipMemcpyHostToDevice);
    std::vector<double> md0(nblocks);
    for (int b = 0; b < nblocks; b++) md0[b] = dtd[case_of_block ? case_of_block[b] : 0].max_dt;
    if (e == hipSuccess) e = hipMemcpy(p->d_max_dt, md0.data(), sizeof(double) * nblocks, hipMemcpyHostToDevice);
    if (e != hipSuccess) { ugks_step_plan_destroy(p); return ugks_rt::hip_fail(e, "step plan: upload"); }
    // face records and cell moments before anything (the autotuner included) runs a face pass
    if (int rc = refresh(p, nullptr)) { ugks_step_plan_destroy(p); return rc; }
    e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) { ugks_step_plan_destroy(p); return ugks_rt::hip_fail(e, "step plan: face records"); }
    *out = p;
    return 0;
}

int ugks_step_plan_create(ugks_block *const *blocks, const ugks_block_buffers *bufs, const ugks_bc_desc *bcs,
                          int32_t nblocks, int32_t precision, const ugks_dt_desc *dtd, ugks_step_plan **out) {
    return ugks_step_plan_create_batch(blocks, bufs, bcs, nblocks, nullptr, 1, precision, dtd, out);
}

int ugks_step_plan_destroy(ugks_step_plan *p) {
    if (!p) return 0;
    if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    void *ptrs[] = {p->d_blocks, p->d_face_off, p->d_cell_off, p->d_all_off, p->d_tile_off, p->d_cmom, p->d_walls, p->d_ghosts,
                    p->d_state, p->d_case_first, p->d_maxbits, p->d_max_dt, p->d_fsum, p->d_fstate, p->d_frec,
                    p->d_tmap, p->d_udf_blocks};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    delete p;
    return 0;
}

int ugks_step_run(ugks_step_plan *p, void *stream, int32_t nsteps) {
    if (!p || nsteps < 0) return ugks_rt::fail(-1, "step run: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    if (nsteps == 0) return 0;
    int rc = refresh(p, s);
    for (int k = 0; k < nsteps && !rc; k++)
        rc = p->precision == UGKS_FP64 ? dispatch_npt<double>(p, s, nullptr) : dispatch_npt<float>(p, s, nullptr);
    return rc;
}

int ugks_step_run_timed(ugks_step_plan *p, void *stream, int32_t nsteps, float *kernel_ms) {
    if (!p || nsteps < 0 || !kernel_ms) return ugks_rt::fail(-1, "step run timed: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t ev[kLaunches + 1];
    for (int k = 0; k <= kLaunches; k++)
        if (hipEventCreate(&ev[k]) != hipSuccess) {
            for (int j = 0; j < k; j++) (void)hipEventDestroy(ev[j]);
            return ugks_rt::fail(-1, "step run timed: event create");
        }
    for (int k = 0; k < kLaunches; k++) kernel_ms[k] = 0.0f;
    int rc = nsteps > 0 ? refresh(p, s) : 0;
    for (int it = 0; it < nsteps && !rc; it++) {
        rc = p->precision == UGKS_FP64 ? dispatch_npt<double>(p, s, ev) : dispatch_npt<float>(p, s, ev);
        if (rc) break;
        if (hipEventSynchronize(ev[kLaunches]) != hipSuccess) { rc = ugks_rt::fail(-1, "step run timed: sync"); break; }
        for (int k = 0; k < kLaunches; k++) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
            kernel_ms[k] += ms;
        }
    }
    for (int k = 0; k <= kLaunches; k++) (void)hipEventDestroy(ev[k]);
    return rc;
}

int ugks_step_states(ugks_step_plan *p, void *stream, ugks_dt_desc *out, double *max_dt) {
    if (!p || !out) return ugks_rt::fail(-1, "step state: bad arguments");
    std::vector<DtState> st(p->ncases);
    std::vector<double> md(p->nb);
    std::vector<int> first(p->ncases + 1);
    hipError_t e = hipMemcpyAsync(st.data(), p->d_state, sizeof(DtState) * p->ncases, hipMemcpyDeviceToHost,
                                  (hipStream_t)stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(md.data(), p->d_max_dt, sizeof(double) * p->nb, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(first.data(), p->d_case_first, sizeof(int) * (p->ncases + 1), hipMemcpyDeviceToHost,
                           (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return ugks_rt::hip_fail(e, "step state");
    if (max_dt) memcpy(max_dt, md.data(), sizeof(double) * p->nb);
    for (int c = 0; c < p->ncases; c++) {
        out[c].dt = st[c].dt; out[c].time = st[c].time; out[c].CFL = st[c].CFL; out[c].delta_dt = st[c].delta_dt;
        out[c].suggest_dt = st[c].suggest_dt; out[c].Kn_eff = st[c].Kn; out[c].step = st[c].step;
        out[c].overshoot = st[c].overshoot; out[c].force_time = st[c].force_time; out[c].pad = 0;
        double M = 0.0;
        for (int b = first[c]; b < first[c + 1]; b++) M = (b == first[c] || md[b] < M) ? md[b] : M;
        out[c].max_dt = M;
    }
    return 0;
}

int ugks_step_state(ugks_step_plan *p, void *stream, ugks_dt_desc *out, double *max_dt) {
    if (!p || !out) return ugks_rt::fail(-1, "step state: bad arguments");
    std::vector<ugks_dt_desc> all(p->ncases);
    std::vector<double> md(max_dt ? p->nb : 0);
    int rc = ugks_step_states(p, stream, all.data(), max_dt ? md.data() : nullptr);
    if (rc) return rc;
    *out = all[0];
    if (max_dt) memcpy(max_dt, md.data(), sizeof(double) * p->nb);
    return 0;
}

int ugks_step_set_dt_case(ugks_step_plan *p, void *stream, int32_t case_index, double dt) {
    if (!p || case_index < 0 || case_index >= p->ncases) return ugks_rt::fail(-1, "set dt: bad arguments");
    // the next step re-evaluates dt against the blocks' last maximum time steps
    DtState d;
    std::vector<int> first(p->ncases + 1);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(&d, p->d_state + case_index, sizeof d, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(first.data(), p->d_case_first, sizeof(int) * (p->ncases + 1),
                                            hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ugks_rt::hip_fail(e, "set dt");
    const int b0 = first[case_index], nb = first[case_index + 1] - b0;
    std::vector<double> md(nb > 0 ? nb : 1, 0.0);
    if (nb > 0) e = hipMemcpyAsync(md.data(), p->d_max_dt + b0, sizeof(double) * nb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return ugks_rt::hip_fail(e, "set dt");
    double M = 0.0;
    for (int k = 0; k < nb; k++) M = (k == 0 || md[k] < M) ? md[k] : M;
    d.dt = d.dt_flux = d.dt_up = dt;
    flux_loop_dts(d, M);
    e = hipMemcpyAsync(p->d_state + case_index, &d, sizeof d, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : ugks_rt::hip_fail(e, "set dt");
}

int ugks_step_set_dt(ugks_step_plan *p, void *stream, double dt) { return ugks_step_set_dt_case(p, stream, 0, dt); }

int ugks_step_force_dt_case(ugks_step_plan *p, void *stream, int32_t case_index, double dt) {
    if (!p || case_index < 0 || case_index >= p->ncases) return ugks_rt::fail(-1, "force dt: bad arguments");
    DtState *d = p->d_state + case_index;
    const int one = 1;
    const double dts[3] = {dt, dt, dt};     // dt, dt_flux, dt_up: held, nothing re-evaluates them
    hipError_t e = hipMemcpyAsync(&d->dt, dts, sizeof dts, hipMemcpyHostToDevice, (hipStream_t)stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&d->force_time, &one, sizeof(int), hipMemcpyHostToDevice, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? 0 : ugks_rt::hip_fail(e, "force dt");
}

int ugks_step_run_graph(ugks_step_plan *p, void *stream, int32_t nsteps) {
    if (!p || nsteps < 0) return ugks_rt::fail(-1, "step run graph: bad arguments");
    if (!p->graph_exec) {
        // one step captured on a stream of our own; the graph then runs on the caller's stream
        hipError_t e = hipSuccess;
        if (!p->cap_stream) e = hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return ugks_rt::hip_fail(e, "step graph: begin capture");
        int rc = p->precision == UGKS_FP64 ? dispatch_npt<double>(p, p->cap_stream, nullptr)
                                           : dispatch_npt<float>(p, p->cap_stream, nullptr);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(p->cap_stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return ugks_rt::hip_fail(e, "step graph: end capture");
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (e != hipSuccess) { (void)hipGraphDestroy(g); return ugks_rt::hip_fail(e, "step graph: instantiate"); }
        p->graph = g;
        p->graph_exec = x;
    }
    if (nsteps == 0) return 0;
    int rc = refresh(p, (hipStream_t)stream);
    if (rc) return rc;
    for (int k = 0; k < nsteps; k++) {
        hipError_t e = hipGraphLaunch(p->graph_exec, (hipStream_t)stream);
        if (e != hipSuccess) return ugks_rt::hip_fail(e, "step graph: launch");
    }
    return 0;
}


int ugks_step_autotune(ugks_step_plan *p, void *stream, int32_t reps, int32_t *choice) {
    if (!p || reps < 1) return ugks_rt::fail(-1, "autotune: bad arguments");
    if (p->nfaces == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    return p->precision == UGKS_FP64 ? autotune<double>(p, s, reps, choice) : autotune<float>(p, s, reps, choice);
}

int ugks_step_block_fits(int32_t NJ, int32_t NV, int32_t flux_method, int32_t precision) {
    if (NJ < 1 || NV < 1) return ugks_rt::fail(-1, "block fits: NJ and NV must be positive");
    if (!stencil_span_fits(NJ, NV, flux_method, precision))
        return ugks_rt::fail(-2, "step plan: a block's stencil span (2R-1)*NJ*NV pairs exceeds the 32-bit byte "
                                 "offsets of the face passes; split the block along i");
    return 0;
}

int ugks_copy_kernel(void *stream, const void *src, void *dst, int64_t bytes) {
    if (bytes <= 0 || bytes % 16) return ugks_rt::fail(-1, "copy: size must be a positive multiple of 16");
    long long n = bytes / 16;
    unsigned grid = (unsigned)((n + kT - 1) / kT);
    if (grid > 8192) grid = 8192;
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kT), 0, (hipStream_t)stream, (const double2 *)src, (double2 *)dst, n);
    return ugks_rt::check_launch("copy");
}

int ugks_f64_to_f32(void *stream, const void *src, void *dst, int64_t n) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, (hipStream_t)stream,
                       (const double *)src, (float *)dst, (long long)n);
    return ugks_rt::check_launch("f64_to_f32");
}

int ugks_f32_to_f64(void *stream, const void *src, void *dst, int64_t n) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_to_f64, dim3((unsigned)((n + kT - 1) / kT)), dim3(kT), 0, (hipStream_t)stream,
                       (const float *)src, (double *)dst, (long long)n);
    return ugks_rt::check_launch("f32_to_f64");
}

}  // extern "C"

// ================================================== allocation-free step ==
// ugks_step_run_noalloc: the plan's step without a plan, for callers that
// cannot allocate device memory.  Its working memory is carved from one
// file-scope device array per GPU; the host refers to that array only through
// the address hipGetSymbolAddress returns for the current device.
namespace {
constexpr size_t kNoAllocBytes = (size_t)64 << 20;   // working memory per GPU
constexpr int kNoAllocMaxBlocks = 16;                 // blocks of the one case a call advances
constexpr int kNoAllocMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kNoAllocBytes];
std::atomic<unsigned char *> g_scratch[kNoAllocMaxDevices];   // the array's address on each device

// the small tables of a call, carried to the device as one kernel argument
struct NoAllocTables {
    long long off[4][kNoAllocMaxBlocks + 1];          // face, cell, all-cell and tile prefix offsets
    WallItem walls[4 * kNoAllocMaxBlocks];
    GhostItem ghosts[4 * kNoAllocMaxBlocks];
    int udf[kNoAllocMaxBlocks];
    int nb, nwall, nghost, nudf;
};
static_assert(sizeof(NoAllocTables) < 4000, "the tables travel as a kernel argument");

__global__ void k_put_block(FBlock F, FBlock *dst) {
    if (threadIdx.x == 0) *dst = F;
}
__global__ void k_put_tables(NoAllocTables T, long long *off, WallItem *walls, GhostItem *ghosts, int *udf,
                             int *case_first) {
    const int t = threadIdx.x;
    if (t < T.nb + 1)
        for (int k = 0; k < 4; k++) off[k * (T.nb + 1) + t] = T.off[k][t];
    if (t < T.nwall) walls[t] = T.walls[t];
    if (t < T.nghost) ghosts[t] = T.ghosts[t];
    if (t < T.nudf) udf[t] = T.udf[t];
    if (t == 0) { case_first[0] = 0; case_first[1] = T.nb; }
}
// the case's time-step state from the caller's descriptor (device memory), and
// the per-block time-step words cleared (the last call left its own there)
__global__ void k_dt_in(const ugks_dt_desc *in, DtState *ds, double *max_dt, unsigned long long *maxbits, int nb) {
    const int t = threadIdx.x;
    if (t < nb) { max_dt[t] = in->max_dt; maxbits[t] = 0ull; }
    if (t == 0) {
        DtState d{in->dt, in->dt, in->dt, in->time, in->CFL, in->delta_dt, in->suggest_dt, in->Kn_eff,
                  in->step, 0, in->force_time, 0};
        flux_loop_dts(d, in->max_dt);
        *ds = d;
    }
}
__global__ void k_dt_out(const DtState *ds, const double *max_dt, int nb, ugks_dt_desc *out) {
    if (threadIdx.x != 0) return;
    const DtState d = *ds;
    double M = 0.0;
    for (int b = 0; b < nb; b++) M = (b == 0 || max_dt[b] < M) ? max_dt[b] : M;
    ugks_dt_desc o;
    o.dt = d.dt; o.time = d.time; o.CFL = d.CFL; o.delta_dt = d.delta_dt; o.suggest_dt = d.suggest_dt;
    o.Kn_eff = d.Kn; o.step = d.step; o.overshoot = d.overshoot; o.force_time = d.force_time; o.pad = 0;
    o.max_dt = M;
    *out = o;
}

// the array's address on the current device, looked up once per device
int scratch_base(unsigned char **base) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return ugks_rt::hip_fail(e, "step noalloc: hipGetDevice");
    if (dev < 0 || dev >= kNoAllocMaxDevices) return ugks_rt::fail(-1, "step noalloc: device index out of range");
    unsigned char *known = g_scratch[dev].load(std::memory_order_acquire);
    if (!known) {
        // threads that race here look up the same address; either store is right
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return ugks_rt::hip_fail(e, "step noalloc: hipGetSymbolAddress(scratch)");
        known = (unsigned char *)ptr;
        g_scratch[dev].store(known, std::memory_order_release);
    }
    *base = known;
    return 0;
}
}  // namespace

extern "C" {

size_t ugks_step_noalloc_capacity(void) { return kNoAllocBytes; }

int ugks_step_run_noalloc(ugks_block *const *blocks, const ugks_block_buffers *bufs, const ugks_bc_desc *bcs,
                          int32_t nblocks, int32_t precision, ugks_dt_desc *dt, int32_t nsteps, void *stream) {
    if (!blocks || !bufs || !bcs || !dt || nsteps < 0 || nblocks < 1)
        return ugks_rt::fail(-1, "step noalloc: bad arguments");
    if (nblocks > kNoAllocMaxBlocks)
        return ugks_rt::fail(-2, "step noalloc: more than 16 blocks in the case");
    if (precision != UGKS_FP64 && precision != UGKS_MIXED) return ugks_rt::fail(-1, "step noalloc: bad precision");
    hipStream_t s = (hipStream_t)stream;
    ugks_step_plan P{};                    // on the stack: every device pointer below points into scratch
    P.nb = nblocks;
    P.precision = precision;
    P.ncases = 1;
    P.fa_waves = P.fc_waves = 4;
    P.method = blocks[0] ? blocks[0]->c.method : -1;
    FBlock host[kNoAllocMaxBlocks];
    NoAllocTables T;
    memset(&T, 0, sizeof T);
    T.nb = nblocks;
    int maxnv = 0, udf_wl = 0;
    long long nw = 0, ng = 0;
    for (int b = 0; b < nblocks; b++) {
        if (!blocks[b]) return ugks_rt::fail(-1, "step noalloc: null block");
        const Consts &C = blocks[b]->c;
        if (C.method != P.method) P.method = -1;
        bool wallf[4];
        if (int rc = fill_block(host[b], blocks[b], bufs[b], bcs + 4 * b, 0, nblocks, nullptr, precision, wallf))
            return rc;
        FBlock &F = host[b];
        T.off[0][b + 1] = T.off[0][b] + F.nS + (long long)F.nWi * C.nj;
        T.off[1][b + 1] = T.off[1][b] + (long long)C.ni * C.nj;
        T.off[2][b + 1] = T.off[2][b] + (long long)C.NI * C.NJ;
        T.off[3][b + 1] = T.off[3][b] + (long long)((C.ni + 1) / 2) * ((C.nj + 1) / 2);
        for (int k = 0; k < 4; k++) {
            const int len = (k == GNORTH || k == GSOUTH) ? C.ni : C.nj;
            if (wallf[k]) { T.walls[T.nwall++] = WallItem{b, k, nw}; nw += len; }
            T.ghosts[T.nghost++] = GhostItem{b, k, ng};
            ng += (long long)len * C.G;
        }
        if (C.time_udf) { T.udf[T.nudf++] = b; udf_wl = std::max(udf_wl, std::max(C.ni, C.nj)); }
        if (C.NV > maxnv) maxnv = C.NV;
    }
    for (int b = 0; b < nblocks; b++)
        for (int k = 0; k < 4; k++) {
            const FBlock &F = host[b];
            if ((F.bc[k] == 0 || F.bc[k] == 5) && (host[F.other[k]].c.NV != F.c.NV || host[F.other[k]].c.G != F.c.G))
                return ugks_rt::fail(-1, "step noalloc: joined blocks differ in NV or ghost width");
        }
    int npt = (maxnv + kT - 1) / kT;
    if (npt > 10) return ugks_rt::fail(-2, "step noalloc: NV above 2560 is not supported by the fused step");
    if (npt == 9) npt = 10;
    for (int b = 0; b < nblocks; b++)
        if (host[b].c.nut > 64 || host[b].c.nvt > 64)
            return ugks_rt::fail(-2, "step noalloc: more than 64 velocities per axis");
    if (T.off[0][nblocks] >= (1LL << 31)) return ugks_rt::fail(-2, "step noalloc: 2^31 or more interior faces");
    P.npt = npt;
    P.nfaces = T.off[0][nblocks];
    P.ncells = T.off[1][nblocks];
    P.nall = T.off[2][nblocks];
    P.ntiles = T.off[3][nblocks];
    P.nwalls = nw;
    P.nghosts = ng;
    P.nwall_items = T.nwall;
    P.nghost_items = T.nghost;
    P.nudf = T.nudf;
    P.udf_wl = udf_wl;
    // carve the working memory: 256-byte aligned regions, refused if they do not fit
    size_t at = 0;
    bool fits = true;
    auto carve = [&](size_t bytes) {
        const size_t here = at;
        at = (at + bytes + 255) & ~(size_t)255;
        if (at > kNoAllocBytes) fits = false;
        return here;
    };
    const size_t o_blocks = carve(sizeof(FBlock) * nblocks), o_off = carve(sizeof(long long) * 4 * (nblocks + 1)),
                 o_walls = carve(sizeof(WallItem) * 4 * nblocks), o_ghosts = carve(sizeof(GhostItem) * 4 * nblocks),
                 o_udf = carve(sizeof(int) * nblocks), o_case = carve(sizeof(int) * 2), o_state = carve(sizeof(DtState)),
                 o_maxbits = carve(sizeof(unsigned long long) * nblocks), o_maxdt = carve(sizeof(double) * nblocks),
                 o_cmom = carve(sizeof(d4) * (size_t)P.nall), o_fsum = carve(sizeof(double) * kFS * (size_t)P.nfaces),
                 o_fstate = carve(sizeof(double) * kFT * (size_t)P.nfaces),
                 o_frec = carve(sizeof(FaceRec) * (size_t)P.nfaces), o_tmap = carve(sizeof(int) * (size_t)P.nfaces);
    if (!fits) {
        char msg[200];
        snprintf(msg, sizeof msg, "step noalloc: the case needs %zu bytes of working memory; the allocation-free "
                                  "step has %zu per GPU", at, kNoAllocBytes);
        return ugks_rt::fail(-2, msg);
    }
    unsigned char *base = nullptr;
    if (int rc = scratch_base(&base)) return rc;
    P.d_blocks = (FBlock *)(base + o_blocks);
    long long *off = (long long *)(base + o_off);
    P.d_face_off = off;
    P.d_cell_off = off + (nblocks + 1);
    P.d_all_off = off + 2 * (nblocks + 1);
    P.d_tile_off = off + 3 * (nblocks + 1);
    P.d_walls = (WallItem *)(base + o_walls);
    P.d_ghosts = (GhostItem *)(base + o_ghosts);
    P.d_udf_blocks = (int *)(base + o_udf);
    P.d_case_first = (int *)(base + o_case);
    P.d_state = (DtState *)(base + o_state);
    P.d_maxbits = (unsigned long long *)(base + o_maxbits);
    P.d_max_dt = (double *)(base + o_maxdt);
    P.d_cmom = (d4 *)(base + o_cmom);
    P.d_fsum = (double *)(base + o_fsum);
    P.d_fstate = (double *)(base + o_fstate);
    P.d_frec = (FaceRec *)(base + o_frec);
    P.d_tmap = (int *)(base + o_tmap);
    for (int b = 0; b < nblocks; b++) {
        host[b].cmom = P.d_cmom + T.off[2][b];
        hipLaunchKernelGGL(k_put_block, dim3(1), dim3(64), 0, s, host[b], P.d_blocks + b);
    }
    hipLaunchKernelGGL(k_put_tables, dim3(1), dim3(kT), 0, s, T, off, P.d_walls, P.d_ghosts, P.d_udf_blocks,
                       P.d_case_first);
    hipLaunchKernelGGL(k_dt_in, dim3(1), dim3(64), 0, s, (const ugks_dt_desc *)dt, P.d_state, P.d_max_dt, P.d_maxbits,
                       nblocks);
    int rc = ugks_rt::check_launch("step noalloc: set-up");
    if (!rc) rc = refresh(&P, s);
    for (int k = 0; k < nsteps && !rc; k++)
        rc = precision == UGKS_FP64 ? dispatch_npt<double>(&P, s, nullptr) : dispatch_npt<float>(&P, s, nullptr);
    if (!rc) {
        hipLaunchKernelGGL(k_dt_out, dim3(1), dim3(64), 0, s, (const DtState *)P.d_state, (const double *)P.d_max_dt,
                           nblocks, dt);
        rc = ugks_rt::check_launch("step noalloc: time-step state");
    }
    return rc;
}

}  // extern "C"
