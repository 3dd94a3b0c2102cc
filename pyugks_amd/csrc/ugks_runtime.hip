// ugks_runtime.hip -- block creation, memory helpers and error reporting of
// libugks_hip.so.  Replaces the per-block OpenCL program build of the
// reference (ugks_block.py:151-189): the constants genHeader bakes into
// source text are kept in a Consts struct and the velocity tables are copied
// to device memory once per block.
#include <stdio.h>
#include <string.h>

#include <string>

#include "../../include/ugks_hip.h"
#include "ugks_runtime.hpp"

namespace {
thread_local std::string g_err;
}

namespace ugks_rt {
int fail(int code, const char *msg) {
    g_err = msg ? msg : "";
    return code;
}
int hip_fail(hipError_t e, const char *where) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s: %s (%d)", where, hipGetErrorString(e), (int)e);
    g_err = buf;
    return -(int)e;
}
int check_launch(const char *name) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, name);
    return 0;
}
int copy_small(void *dst_host, const void *src_dev, size_t bytes, hipStream_t s) {
    hipError_t e = hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? 0 : 1;
}
}  // namespace ugks_rt

using ugks_rt::fail;
using ugks_rt::hip_fail;

extern "C" {

const char *ugks_last_error(void) { return g_err.c_str(); }

int ugks_get_device_count(int *count) {
    if (!count) return fail(-1, "null count");
    hipError_t e = hipGetDeviceCount(count);
    return e == hipSuccess ? 0 : hip_fail(e, "hipGetDeviceCount");
}

int ugks_set_device(int device) {
    hipError_t e = hipSetDevice(device);
    return e == hipSuccess ? 0 : hip_fail(e, "hipSetDevice");
}

int ugks_block_create(const ugks_block_desc *d, ugks_block **out) {
    if (!d || !out) return fail(-1, "null descriptor");
    *out = nullptr;
    if (d->ni < 1 || d->nj < 1 || d->NV < 1) return fail(-1, "ni, nj and NV must be positive");
    int G = d->flux_method == 1 ? 3 : 2;
    if (d->flux_method != 0 && d->flux_method != 1 && d->flux_method != 3)
        return fail(-2, "flux_method must be 0 (vanLeer), 1 (WENO5) or 3 (minmod)");
    if (d->ghost != G || d->NI != d->ni + 2 * G || d->NJ != d->nj + 2 * G)
        return fail(-1, "ghost / NI / NJ inconsistent with ni, nj and the flux method");
    if (d->relax_type < 0 || d->relax_type > 3) return fail(-2, "relax_type must be 0..3");
    if (!d->quad || !d->weight) return fail(-1, "quad and weight tables are required");
    const size_t nv = (size_t)d->NV;
    // tensor structure of QUAD: v = i*nvt + j with quad[v] = (qu[i], qv[j])
    size_t nvt = 1;
    while (nvt < nv && d->quad[2 * nvt] == d->quad[0]) nvt++;
    size_t nut = nv / nvt;
    bool tensor = (nut * nvt == nv);
    for (size_t v = 0; tensor && v < nv; v++)
        tensor = d->quad[2 * v] == d->quad[2 * (v - v % nvt)] && d->quad[2 * v + 1] == d->quad[2 * (v % nvt) + 1];
    if (!tensor) nut = nvt = 0;
    // tabulated isotherms: every index the device walk follows must stay inside its table
    size_t iso_doubles = 0, iso_ints = 0;
    for (int k = 0; k < 4; k++) {
        const int32_t ni_ = d->n_iso[k], nt = d->n_tri[k];
        if (nt == 0) continue;
        if (nt < 0 || ni_ < 3 || !d->iso[k] || !d->iso_tri[k] || !d->iso_nbr[k])
            return fail(-1, "isotherm table: need n_iso >= 3 points, triangles and neighbours");
        for (int32_t t = 0; t < 3 * nt; t++) {
            if (d->iso_tri[k][t] < 0 || d->iso_tri[k][t] >= ni_) return fail(-1, "isotherm table: vertex index out of range");
            if (d->iso_nbr[k][t] < -1 || d->iso_nbr[k][t] >= nt) return fail(-1, "isotherm table: neighbour index out of range");
        }
        iso_doubles += 3 * (size_t)ni_;
        iso_ints += 6 * (size_t)nt;
    }
    // wall-value programs: opcodes known, operands in range, the stack within 16 entries and
    // ending with one value, so the device interpreter never leaves its arrays
    bool any_udf = false;
    for (int k = 0; k < 16; k++) {
        const int32_t st = d->udf_start[k], n = d->udf_len[k];
        if (n == 0) continue;
        any_udf = true;
        if (!d->udf_code || n < 0 || st < 0 || (long long)st + n > d->udf_ncode)
            return fail(-1, "wall-value program out of its code table");
        int sp = 0;
        for (int i = 0; i < n; i++) {
            const int32_t op = d->udf_code[st + i];
            if (op == 1) {
                if (i + 1 >= n || !d->udf_const || d->udf_code[st + i + 1] < 0 || d->udf_code[st + i + 1] >= d->udf_nconst)
                    return fail(-1, "wall-value program: constant index out of range");
                i++; sp++;
            } else if (op >= 2 && op <= 5) {
                sp++;
            } else if ((op >= 10 && op <= 13) || op == 15 || (op >= 40 && op <= 43)) {
                if (sp < 2) return fail(-1, "wall-value program: stack underflow");
                sp--;
            } else if (op == 14 || (op >= 20 && op <= 35)) {
                if (sp < 1) return fail(-1, "wall-value program: stack underflow");
            } else {
                return fail(-1, "wall-value program: unknown opcode");
            }
            if (sp > 16) return fail(-1, "wall-value program: deeper than 16 stack entries");
        }
        if (sp != 1) return fail(-1, "wall-value program must leave exactly one value");
    }
    const size_t udf_ints = any_udf ? (size_t)d->udf_ncode : 0, udf_doubles = any_udf ? (size_t)d->udf_nconst : 0;
    const size_t bytes = nv * (2 * sizeof(double) + sizeof(double) + 3 * sizeof(int64_t)) + (nut + nvt) * sizeof(double) +
                         iso_doubles * sizeof(double) + udf_doubles * sizeof(double) + iso_ints * sizeof(int32_t) +
                         udf_ints * sizeof(int32_t);
    ugks_block *b = new ugks_block();
    hipError_t e = hipGetDevice(&b->device);
    if (e != hipSuccess) { delete b; return hip_fail(e, "hipGetDevice"); }
    e = hipMalloc(&b->dev_tables, bytes);
    if (e != hipSuccess) { delete b; return hip_fail(e, "hipMalloc(tables)"); }
    char *host = (char *)malloc(bytes);
    if (!host) { (void)hipFree(b->dev_tables); delete b; return fail(-1, "out of host memory"); }
    char *p = host;
    memcpy(p, d->quad, nv * 2 * sizeof(double)); p += nv * 2 * sizeof(double);
    memcpy(p, d->weight, nv * sizeof(double)); p += nv * sizeof(double);
    const int64_t *mir[3] = {d->mirror_NS, d->mirror_EW, d->mirror_D};
    for (int k = 0; k < 3; k++) {
        if (mir[k]) {
            for (size_t v = 0; v < nv; v++)
                if (mir[k][v] < 0 || mir[k][v] >= (int64_t)nv) {
                    free(host); (void)hipFree(b->dev_tables); delete b;
                    return fail(-1, "mirror table entry out of range");
                }
            memcpy(p, mir[k], nv * sizeof(int64_t));
        } else {
            memset(p, 0, nv * sizeof(int64_t));
        }
        p += nv * sizeof(int64_t);
    }
    for (size_t i = 0; i < nut; i++) ((double *)p)[i] = d->quad[2 * i * nvt];
    p += nut * sizeof(double);
    for (size_t j = 0; j < nvt; j++) ((double *)p)[j] = d->quad[2 * j + 1];
    p += nvt * sizeof(double);
    for (int k = 0; k < 4; k++)
        if (d->n_tri[k] > 0) { memcpy(p, d->iso[k], 3 * (size_t)d->n_iso[k] * sizeof(double)); p += 3 * (size_t)d->n_iso[k] * sizeof(double); }
    if (udf_doubles) { memcpy(p, d->udf_const, udf_doubles * sizeof(double)); p += udf_doubles * sizeof(double); }
    for (int k = 0; k < 4; k++)
        if (d->n_tri[k] > 0) {
            memcpy(p, d->iso_tri[k], 3 * (size_t)d->n_tri[k] * sizeof(int32_t)); p += 3 * (size_t)d->n_tri[k] * sizeof(int32_t);
            memcpy(p, d->iso_nbr[k], 3 * (size_t)d->n_tri[k] * sizeof(int32_t)); p += 3 * (size_t)d->n_tri[k] * sizeof(int32_t);
        }
    if (udf_ints) { memcpy(p, d->udf_code, udf_ints * sizeof(int32_t)); p += udf_ints * sizeof(int32_t); }
    e = hipMemcpy(b->dev_tables, host, bytes, hipMemcpyHostToDevice);
    free(host);
    if (e != hipSuccess) { (void)hipFree(b->dev_tables); delete b; return hip_fail(e, "hipMemcpy(tables)"); }

    ugks::Consts &c = b->c;
    c.ni = d->ni; c.nj = d->nj; c.NI = d->NI; c.NJ = d->NJ; c.NV = d->NV; c.G = d->ghost;
    c.method = d->flux_method; c.relax = d->relax_type;
    c.Pr = d->Pr; c.Kn = d->Kn_eff; c.gam = d->gam; c.K = d->K; c.omega = d->omega;
    c.umax = d->umax; c.vmax = d->vmax;
    c.B = d->B;
    for (int k = 0; k < 4; k++) {
        c.wall_D[k] = d->wall_D[k]; c.wall_U[k] = d->wall_U[k]; c.wall_V[k] = d->wall_V[k];
        c.wall_T[k] = d->wall_T[k]; c.wall_P[k] = d->wall_P[k];
        c.beta_n[k] = d->beta_n[k]; c.beta_t[k] = d->beta_t[k]; c.gamma_f[k] = d->gamma_f[k];
        c.alpha_p[k] = d->alpha_p[k]; c.adsorb[k] = d->adsorb[k];
        c.alpha_n[k] = d->alpha_n[k]; c.alpha_t[k] = d->alpha_t[k];
        c.dyn[k] = d->dynamic_vartheta[k]; c.cads[k] = d->constant_adsorb[k];
    }
    char *dp = (char *)b->dev_tables;
    c.quad = (const ugks::d2 *)dp; dp += nv * 2 * sizeof(double);
    c.weight = (const double *)dp; dp += nv * sizeof(double);
    c.mNS = (const int64_t *)dp; dp += nv * sizeof(int64_t);
    c.mEW = (const int64_t *)dp; dp += nv * sizeof(int64_t);
    c.mD = (const int64_t *)dp; dp += nv * sizeof(int64_t);
    c.nut = (int)nut; c.nvt = (int)nvt;
    c.qu = nut ? (const double *)dp : nullptr; dp += nut * sizeof(double);
    c.qv = nvt ? (const double *)dp : nullptr; dp += nvt * sizeof(double);
    for (int k = 0; k < 4; k++) {
        c.n_tri[k] = d->n_tri[k];
        c.iso[k] = nullptr; c.iso_tri[k] = c.iso_nbr[k] = nullptr;
        if (d->n_tri[k] > 0) { c.iso[k] = (const double *)dp; dp += 3 * (size_t)d->n_iso[k] * sizeof(double); }
    }
    c.udf_const = udf_doubles ? (const double *)dp : nullptr;
    dp += udf_doubles * sizeof(double);
    for (int k = 0; k < 4; k++)
        if (d->n_tri[k] > 0) {
            c.iso_tri[k] = (const int *)dp; dp += 3 * (size_t)d->n_tri[k] * sizeof(int32_t);
            c.iso_nbr[k] = (const int *)dp; dp += 3 * (size_t)d->n_tri[k] * sizeof(int32_t);
        }
    c.udf_code = udf_ints ? (const int *)dp : nullptr;
    c.any_udf = any_udf ? 1 : 0;
    c.time_udf = 0;
    for (int k = 0; k < 16; k++) {
        c.udf_start[k] = d->udf_start[k];
        c.udf_len[k] = any_udf ? d->udf_len[k] : 0;
        // a program that reads t makes the block's wall values change every step
        for (int i = 0; c.udf_len[k] && i < d->udf_len[k]; i++) {
            const int32_t op = d->udf_code[d->udf_start[k] + i];
            if (op == 1) { i++; continue; }
            if (op == 5) c.time_udf = 1;
        }
    }
    *out = b;
    return 0;
}

int ugks_block_destroy(ugks_block *blk) {
    if (!blk) return 0;
    hipError_t e = hipFree(blk->dev_tables);
    delete blk;
    return e == hipSuccess ? 0 : hip_fail(e, "hipFree(tables)");
}

int ugks_malloc(void **ptr, size_t bytes) {
    if (!ptr) return fail(-1, "null ptr");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    return e == hipSuccess ? 0 : hip_fail(e, "hipMalloc");
}
int ugks_free(void *ptr) {
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? 0 : hip_fail(e, "hipFree");
}
int ugks_memset(void *ptr, int value, size_t bytes, void *stream) {
    hipError_t e = hipMemsetAsync(ptr, value, bytes, (hipStream_t)stream);
    return e == hipSuccess ? 0 : hip_fail(e, "hipMemsetAsync");
}
static int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind k, void *stream, const char *w) {
    if (bytes == 0) return 0;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream);
    if (e == hipSuccess && k != hipMemcpyDeviceToDevice) e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? 0 : hip_fail(e, w);
}
int ugks_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream) {
    return copy(dst, src, bytes, hipMemcpyHostToDevice, stream, "memcpy htod");
}
int ugks_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream) {
    return copy(dst, src, bytes, hipMemcpyDeviceToHost, stream, "memcpy dtoh");
}
int ugks_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream) {
    return copy(dst, src, bytes, hipMemcpyDeviceToDevice, stream, "memcpy dtod");
}
int ugks_stream_create(void **stream) {
    if (!stream) return fail(-1, "null stream");
    hipError_t e = hipStreamCreate((hipStream_t *)stream);
    return e == hipSuccess ? 0 : hip_fail(e, "hipStreamCreate");
}
int ugks_stream_destroy(void *stream) {
    hipError_t e = hipStreamDestroy((hipStream_t)stream);
    return e == hipSuccess ? 0 : hip_fail(e, "hipStreamDestroy");
}
int ugks_stream_synchronize(void *stream) {
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? 0 : hip_fail(e, "hipStreamSynchronize");
}
int ugks_event_create(void **event) {
    if (!event) return fail(-1, "null event");
    hipError_t e = hipEventCreate((hipEvent_t *)event);
    return e == hipSuccess ? 0 : hip_fail(e, "hipEventCreate");
}
int ugks_event_destroy(void *event) {
    hipError_t e = hipEventDestroy((hipEvent_t)event);
    return e == hipSuccess ? 0 : hip_fail(e, "hipEventDestroy");
}
int ugks_event_record(void *event, void *stream) {
    hipError_t e = hipEventRecord((hipEvent_t)event, (hipStream_t)stream);
    return e == hipSuccess ? 0 : hip_fail(e, "hipEventRecord");
}
int ugks_event_elapsed_ms(float *ms, void *start, void *stop) {
    if (!ms) return fail(-1, "null ms");
    hipError_t e = hipEventSynchronize((hipEvent_t)stop);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
    return e == hipSuccess ? 0 : hip_fail(e, "hipEventElapsedTime");
}

}  // extern "C"
