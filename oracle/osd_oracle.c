/*
 * CPU oracle for the Osd evaluator hot path.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, the smoke check and the benchmark's CPU-baseline leg,
 * never by the shipped evaluator.  See osd_oracle.h for the reference
 * functions restated here and how each one is pinned.
 */
#include "osd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define R float
#define SFX f32
#include "osd_oracle_basis.inc"
#undef R
#undef SFX

#define R double
#define SFX f64
#include "osd_oracle_basis.inc"
#undef R
#undef SFX

/* ----------------------------------------------------------------------- */
/* Stencils: osd/cpuKernel.cpp:71-240.  Every output row is cleared, then
 * accumulates w * src[index] in table order in the working precision, then
 * is copied to dst row (i - start). */

static int stencil_rows_f32(const float *src, oracle_desc sd,
                            float *const *outs, const oracle_desc *ods,
                            const float *const *ws, int nout,
                            const int *sizes, const int *offsets,
                            const int *indices, int start, int end)
{
    float acc[6][64];
    int L = sd.length;
    if (L > 64) return 0;
    src += sd.offset;
    for (int i = start; i < end; ++i) {
        int off = offsets[i], n = sizes[i];
        for (int o = 0; o < nout; ++o) memset(acc[o], 0, sizeof(float) * (size_t)L);
        for (int j = 0; j < n; ++j) {
            const float *p = src + (size_t)indices[off + j] * (size_t)sd.stride;
            for (int o = 0; o < nout; ++o) {
                float w = ws[o][off + j];
                for (int k = 0; k < L; ++k) acc[o][k] += p[k] * w;
            }
        }
        for (int o = 0; o < nout; ++o) {
            float *d = outs[o] + ods[o].offset + (size_t)(i - start) * (size_t)ods[o].stride;
            memcpy(d, acc[o], sizeof(float) * (size_t)L);
        }
    }
    return 1;
}

int oracle_eval_stencils_f32(const float *src, oracle_desc srcDesc,
                             float *dst, oracle_desc dstDesc,
                             const int *sizes, const int *offsets,
                             const int *indices, const float *weights,
                             int start, int end)
{
    if (end <= start) return 1;
    if (srcDesc.length != dstDesc.length) return 0;
    float *outs[1] = { dst };
    oracle_desc ods[1] = { dstDesc };
    const float *ws[1] = { weights };
    return stencil_rows_f32(src, srcDesc, outs, ods, ws, 1, sizes, offsets, indices, start, end);
}

int oracle_eval_stencils_d1_f32(const float *src, oracle_desc srcDesc,
                                float *dst, oracle_desc dstDesc,
                                float *du, oracle_desc duDesc,
                                float *dv, oracle_desc dvDesc,
                                const int *sizes, const int *offsets,
                                const int *indices, const float *weights,
                                const float *duWeights, const float *dvWeights,
                                int start, int end)
{
    if (end <= start) return 1;
    if (srcDesc.length != dstDesc.length || srcDesc.length != duDesc.length ||
        srcDesc.length != dvDesc.length) return 0;
    float *outs[3] = { dst, du, dv };
    oracle_desc ods[3] = { dstDesc, duDesc, dvDesc };
    const float *ws[3] = { weights, duWeights, dvWeights };
    return stencil_rows_f32(src, srcDesc, outs, ods, ws, 3, sizes, offsets, indices, start, end);
}

int oracle_eval_stencils_d2_f32(const float *src, oracle_desc srcDesc,
                                float *dst, oracle_desc dstDesc,
                                float *du, oracle_desc duDesc,
                                float *dv, oracle_desc dvDesc,
                                float *duu, oracle_desc duuDesc,
                                float *duv, oracle_desc duvDesc,
                                float *dvv, oracle_desc dvvDesc,
                                const int *sizes, const int *offsets,
                                const int *indices, const float *weights,
                                const float *duWeights, const float *dvWeights,
                                const float *duuWeights, const float *duvWeights,
                                const float *dvvWeights,
                                int start, int end)
{
    if (end <= start) return 1;
    int L = srcDesc.length;
    if (L != dstDesc.length || L != duDesc.length || L != dvDesc.length ||
        L != duuDesc.length || L != duvDesc.length || L != dvvDesc.length) return 0;
    float *outs[6] = { dst, du, dv, duu, duv, dvv };
    oracle_desc ods[6] = { dstDesc, duDesc, dvDesc, duuDesc, duvDesc, dvvDesc };
    const float *ws[6] = { weights, duWeights, dvWeights, duuWeights, duvWeights, dvvWeights };
    return stencil_rows_f32(src, srcDesc, outs, ods, ws, 6, sizes, offsets, indices, start, end);
}

int oracle_eval_stencils_f64(const double *src, oracle_desc srcDesc,
                             double *dst, oracle_desc dstDesc,
                             const int *sizes, const int *offsets,
                             const int *indices, const double *weights,
                             int start, int end)
{
    if (end <= start) return 1;
    if (srcDesc.length != dstDesc.length) return 0;
    int L = srcDesc.length;
    if (L > 64) return 0;
    src += srcDesc.offset;
    dst += dstDesc.offset;
    for (int i = start; i < end; ++i) {
        double acc[64];
        memset(acc, 0, sizeof(double) * (size_t)L);
        int off = offsets[i];
        for (int j = 0; j < sizes[i]; ++j) {
            const double *p = src + (size_t)indices[off + j] * (size_t)srcDesc.stride;
            double w = weights[off + j];
            for (int k = 0; k < L; ++k) acc[k] += p[k] * w;
        }
        memcpy(dst + (size_t)(i - start) * (size_t)dstDesc.stride, acc, sizeof(double) * (size_t)L);
    }
    return 1;
}

int oracle_eval_stencils_omp_f32(const float *src, oracle_desc srcDesc,
                                 float *dst, oracle_desc dstDesc,
                                 const int *sizes, const int *offsets,
                                 const int *indices, const float *weights,
                                 int start, int end, int numThreads)
{
    if (end <= start) return 1;
    if (srcDesc.length != dstDesc.length) return 0;
    int L = srcDesc.length;
    if (L > 64) return 0;
    const float *s0 = src + srcDesc.offset;
    float *d0 = dst + dstDesc.offset;
#ifdef _OPENMP
    if (numThreads > 0) omp_set_num_threads(numThreads);
#else
    (void)numThreads;
#endif
    /* osd/ompKernel.cpp:93 - default schedule, one stencil per iteration */
#pragma omp parallel for
    for (int i = start; i < end; ++i) {
        float acc[64];
        memset(acc, 0, sizeof(float) * (size_t)L);
        int off = offsets[i];
        for (int j = 0; j < sizes[i]; ++j) {
            const float *p = s0 + (size_t)indices[off + j] * (size_t)srcDesc.stride;
            float w = weights[off + j];
            for (int k = 0; k < L; ++k) acc[k] += p[k] * w;
        }
        memcpy(d0 + (size_t)(i - start) * (size_t)dstDesc.stride, acc, sizeof(float) * (size_t)L);
    }
    return 1;
}

/* ----------------------------------------------------------------------- */
/* Patches: osd/cpuEvaluator.cpp:157-381.  For each coord: pick the regular
 * or irregular descriptor of its array by the PatchParam "regular" bit
 * (cpuEvaluator.cpp:257-259), evaluate the basis, locate the patch's
 * control points at indexBase + stride * (patchIndex - primitiveIdBase),
 * and accumulate each requested output. */

static int patch_checks(const void *src, oracle_desc sd, const void *const *outs,
                        const oracle_desc *ods, int nout)
{
    if (!src) return 0;
    for (int o = 0; o < nout; ++o)
        if (outs[o] && ods[o].length != sd.length) return 0;
    return 1;
}

int oracle_eval_patches_f32(const float *src, oracle_desc srcDesc,
                            float *dst, oracle_desc dstDesc,
                            float *du, oracle_desc duDesc,
                            float *dv, oracle_desc dvDesc,
                            float *duu, oracle_desc duuDesc,
                            float *duv, oracle_desc duvDesc,
                            float *dvv, oracle_desc dvvDesc,
                            int numPatchCoords,
                            const oracle_patch_coord *coords,
                            const oracle_patch_array *arrays,
                            const int *patchIndices,
                            const oracle_patch_param *params)
{
    float *outs[6] = { dst, du, dv, duu, duv, dvv };
    oracle_desc ods[6] = { dstDesc, duDesc, dvDesc, duuDesc, duvDesc, dvvDesc };
    if (!patch_checks(src, srcDesc, (const void *const *)outs, ods, 6)) return 0;
    int L = srcDesc.length;
    if (L > 64) return 0;
    int want1 = du || dv, want2 = duu || duv || dvv;
    const float *s0 = src + srcDesc.offset;
    for (int i = 0; i < numPatchCoords; ++i) {
        const oracle_patch_coord *c = &coords[i];
        const oracle_patch_array *a = &arrays[c->arrayIndex];
        const oracle_patch_param *pp = &params[c->patchIndex];
        int type = ((pp->field1 >> 5) & 1u) ? a->regDesc : a->desc;
        float w[6][20];
        int n = oracle_patch_basis_f32(type, pp->field0, pp->field1, c->s, c->t,
                                       w[0], want1 || want2 ? w[1] : 0, want1 || want2 ? w[2] : 0,
                                       want2 ? w[3] : 0, want2 ? w[4] : 0, want2 ? w[5] : 0);
        const int *cvs = patchIndices + a->indexBase + a->stride * (c->patchIndex - a->primitiveIdBase);
        for (int o = 0; o < 6; ++o) {
            if (!outs[o]) continue;
            float acc[64];
            memset(acc, 0, sizeof(float) * (size_t)L);
            for (int j = 0; j < n; ++j) {
                const float *p = s0 + (size_t)cvs[j] * (size_t)srcDesc.stride;
                for (int k = 0; k < L; ++k) acc[k] += p[k] * w[o][j];
            }
            memcpy(outs[o] + ods[o].offset + (size_t)i * (size_t)ods[o].stride, acc,
                   sizeof(float) * (size_t)L);
        }
    }
    return 1;
}

int oracle_eval_patches_f64(const double *src, oracle_desc srcDesc,
                            double *dst, oracle_desc dstDesc,
                            double *du, oracle_desc duDesc,
                            double *dv, oracle_desc dvDesc,
                            int numPatchCoords,
                            const oracle_patch_coord *coords,
                            const oracle_patch_array *arrays,
                            const int *patchIndices,
                            const oracle_patch_param *params)
{
    double *outs[3] = { dst, du, dv };
    oracle_desc ods[3] = { dstDesc, duDesc, dvDesc };
    if (!patch_checks(src, srcDesc, (const void *const *)outs, ods, 3)) return 0;
    int L = srcDesc.length;
    if (L > 64) return 0;
    int want1 = du || dv;
    const double *s0 = src + srcDesc.offset;
    for (int i = 0; i < numPatchCoords; ++i) {
        const oracle_patch_coord *c = &coords[i];
        const oracle_patch_array *a = &arrays[c->arrayIndex];
        const oracle_patch_param *pp = &params[c->patchIndex];
        int type = ((pp->field1 >> 5) & 1u) ? a->regDesc : a->desc;
        double w[3][20];
        int n = oracle_patch_basis_f64(type, pp->field0, pp->field1, (double)c->s, (double)c->t,
                                       w[0], want1 ? w[1] : 0, want1 ? w[2] : 0, 0, 0, 0);
        const int *cvs = patchIndices + a->indexBase + a->stride * (c->patchIndex - a->primitiveIdBase);
        for (int o = 0; o < 3; ++o) {
            if (!outs[o]) continue;
            double acc[64];
            memset(acc, 0, sizeof(double) * (size_t)L);
            for (int j = 0; j < n; ++j) {
                const double *p = s0 + (size_t)cvs[j] * (size_t)srcDesc.stride;
                for (int k = 0; k < L; ++k) acc[k] += p[k] * w[o][j];
            }
            memcpy(outs[o] + ods[o].offset + (size_t)i * (size_t)ods[o].stride, acc,
                   sizeof(double) * (size_t)L);
        }
    }
    return 1;
}

int oracle_eval_patches_omp_f32(const float *src, oracle_desc srcDesc,
                                float *dst, oracle_desc dstDesc,
                                float *du, oracle_desc duDesc,
                                float *dv, oracle_desc dvDesc,
                                int numPatchCoords,
                                const oracle_patch_coord *coords,
                                const oracle_patch_array *arrays,
                                const int *patchIndices,
                                const oracle_patch_param *params,
                                int numThreads)
{
    float *outs[3] = { dst, du, dv };
    oracle_desc ods[3] = { dstDesc, duDesc, dvDesc };
    if (!patch_checks(src, srcDesc, (const void *const *)outs, ods, 3)) return 0;
    int L = srcDesc.length;
    if (L > 64) return 0;
    int want1 = du || dv;
    const float *s0 = src + srcDesc.offset;
#ifdef _OPENMP
    if (numThreads > 0) omp_set_num_threads(numThreads);
#else
    (void)numThreads;
#endif
#pragma omp parallel for
    for (int i = 0; i < numPatchCoords; ++i) {
        const oracle_patch_coord *c = &coords[i];
        const oracle_patch_array *a = &arrays[c->arrayIndex];
        const oracle_patch_param *pp = &params[c->patchIndex];
        int type = ((pp->field1 >> 5) & 1u) ? a->regDesc : a->desc;
        float w[3][20];
        int n = oracle_patch_basis_f32(type, pp->field0, pp->field1, c->s, c->t,
                                       w[0], want1 ? w[1] : 0, want1 ? w[2] : 0, 0, 0, 0);
        const int *cvs = patchIndices + a->indexBase + a->stride * (c->patchIndex - a->primitiveIdBase);
        for (int o = 0; o < 3; ++o) {
            if (!outs[o]) continue;
            float acc[64];
            memset(acc, 0, sizeof(float) * (size_t)L);
            for (int j = 0; j < n; ++j) {
                const float *p = s0 + (size_t)cvs[j] * (size_t)srcDesc.stride;
                for (int k = 0; k < L; ++k) acc[k] += p[k] * w[o][j];
            }
            memcpy(outs[o] + ods[o].offset + (size_t)i * (size_t)ods[o].stride, acc,
                   sizeof(float) * (size_t)L);
        }
    }
    return 1;
}

/* ---- limit normals (tutorials/far/tutorial_2_3/far_tutorial_2_3.cpp:73-87) -- */

int oracle_limit_normals_f32(const float *du, oracle_desc duDesc, const float *dv,
                             oracle_desc dvDesc, float *n, oracle_desc nDesc, int count,
                             int normalize)
{
    if (duDesc.length < 3 || dvDesc.length < 3 || nDesc.length < 3) return 0;
    for (int i = 0; i < count; ++i) {
        const float *a = du + duDesc.offset + (size_t)i * (size_t)duDesc.stride;
        const float *b = dv + dvDesc.offset + (size_t)i * (size_t)dvDesc.stride;
        float *o = n + nDesc.offset + (size_t)i * (size_t)nDesc.stride;
        float c[3];
        c[0] = a[1] * b[2] - a[2] * b[1];   /* cross(v1, v2): lines 82-87 */
        c[1] = a[2] * b[0] - a[0] * b[2];
        c[2] = a[0] * b[1] - a[1] * b[0];
        if (normalize) {                    /* normalize(n): lines 73-79 */
            const float rn = 1.0f / sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            c[0] *= rn;
            c[1] *= rn;
            c[2] *= rn;
        }
        o[0] = c[0];
        o[1] = c[1];
        o[2] = c[2];
    }
    return 1;
}

/* Bfr::Tessellation::GetCoords<float> of a quad at a uniform rate
 * (bfr/tessellation.cpp:158-174 iso-lines whose coordinate accumulates its
 * step point by point; 929-945 the boundary, counter-clockwise from (0,0);
 * 947-990 the interior rings from the outside in, the centre last for an
 * even rate).  Writes (rate + 1)^2 (u, v) pairs; returns the count, or -1
 * for rate < 1. */
static int tess_iso_v(float *uv, int n, float u, float v, float du)
{
    for (int i = 0; i < n; ++i, u += du) {
        uv[2 * i] = u;
        uv[2 * i + 1] = v;
    }
    return n;
}

static int tess_iso_u(float *uv, int n, float u, float v, float dv)
{
    for (int i = 0; i < n; ++i, v += dv) {
        uv[2 * i] = u;
        uv[2 * i + 1] = v;
    }
    return n;
}

static int tess_uv_line(float *uv, int n, float u, float v, float du, float dv)
{
    for (int i = 0; i < n; ++i, u += du, v += dv) {
        uv[2 * i] = u;
        uv[2 * i + 1] = v;
    }
    return n;
}

static int tess_point(float *uv, float u, float v)
{
    uv[0] = u;
    uv[1] = v;
    return 1;
}

/* Bfr::Tessellation::GetCoords<float> for any parameterization and rates
 * (bfr/tessellation.cpp): scheme 1 Catmark / 0 bilinear (quads: QUAD,
 * other sizes: QUAD_SUBFACES) or 2 Loop (triangles: TRI); rates as
 * Tessellation(param, numRates, rates) takes them -- 1 uniform rate, 2 inner
 * rates of a quad, N outer rates, N + 1 (N + 2 for quads) outer and inner
 * rates.  Rates initialised as initializeRates (:2077-2146, rates capped at
 * 32767), interior point count as initializeInventoryFor* (:2165-2275),
 * boundary coords per parameterization (quad :929-945, tri :1368-1381, qsub
 * :1745-1760 via getRingEdgeCoords :1687-1733), interior coords (:2345-2371;
 * quad :945-995, tri :1383-1422, qsub :1762-1822).  Writes the (u, v)
 * pairs when uv is not NULL; returns the coordinate count, -1 for invalid
 * arguments. */
typedef struct {
    int type;           /* 0 quad, 1 tri, 2 quad sub-faces */
    int N, uDim;
} tess_param;

static void tess_vertex(const tess_param *P, int k, float *u, float *v)
{
    if (P->type == 0) { *u = (float)(k && k < 3); *v = (float)(k > 1); }
    else if (P->type == 1) { *u = (float)(k == 1); *v = (float)(k == 2); }
    else { *u = (float)(k % P->uDim); *v = (float)(k / P->uDim); }
}

static int tess_qsub_edge(const tess_param *P, float *uv, int edge, int edgeRes, int incFirst,
                          float t0, float dt)
{
    const int n0 = (edgeRes - 1) / 2, n1 = (edgeRes - 1) - n0;
    int n = 0;
    if (incFirst || n0) {
        float u0, v0;
        tess_vertex(P, edge, &u0, &v0);
        if (incFirst) n += tess_point(uv + 2 * n, u0 + t0, v0 + t0);
        if (n0) n += tess_iso_v(uv + 2 * n, n0, u0 + t0 + dt, v0 + t0, dt);
    }
    if (n1) {
        float u1, v1;
        tess_vertex(P, (edge + 1) % P->N, &u1, &v1);
        n += tess_iso_u(uv + 2 * n, n1, u1 + t0,
                        v1 + ((edgeRes & 1) ? (0.5f - 0.5f * dt) : 0.5f), -dt);
    }
    return n;
}

int oracle_tess_coords_f32(int scheme, int faceSize, int numRates, const int *rates, float *uv)
{
    if (faceSize < 3 || numRates < 1 || !rates) return -1;
    if (scheme == 2 && faceSize != 3) return -1;
    for (int i = 0; i < numRates; ++i)
        if (rates[i] < 1) return -1;
    tess_param P;
    P.N = faceSize;
    P.type = scheme == 2 ? 1 : (faceSize == 4 ? 0 : 2);
    P.uDim = faceSize < 10 ? 2 + (faceSize > 4) : 1 + (int)sqrtf((float)(faceSize - 1));
    const int N = faceSize, isQuad = P.type == 0, maxRate = 32767;
    int outer[64], inner[2], uniform = 1, total = 0;
    int *out = outer;
    int *heap = NULL;
    if (N > 64) {
        heap = (int *)malloc(sizeof(int) * (size_t)N);
        if (!heap) return -1;
        out = heap;
    }
#define CAP(x) ((x) < maxRate ? (x) : maxRate)
    if (numRates < N) {
        if (numRates == 2 && isQuad) {
            inner[0] = CAP(rates[0]);
            inner[1] = CAP(rates[1]);
            out[0] = out[2] = inner[0];
            out[1] = out[3] = inner[1];
            uniform = inner[0] == inner[1];
            total = 2 * (inner[0] + inner[1]);
        } else {
            inner[0] = inner[1] = CAP(rates[0]);
            for (int i = 0; i < N; ++i) out[i] = inner[0];
            total = inner[0] * N;
        }
    } else {
        for (int i = 0; i < N; ++i) {
            out[i] = CAP(rates[i]);
            uniform = uniform && out[i] == out[0];
            total += out[i];
        }
        if (numRates > N) {
            inner[0] = CAP(rates[N]);
            inner[1] = (numRates == 6 && isQuad) ? CAP(rates[5]) : inner[0];
            uniform = uniform && inner[0] == out[0] && inner[1] == out[0];
        } else if (isQuad) {
            inner[0] = (out[0] + out[2]) / 2;
            inner[1] = (out[1] + out[3]) / 2;
        } else {
            inner[0] = inner[1] = total / N;
        }
    }
#undef CAP
    /* interior point count (initializeInventoryFor*) */
    int nInterior = 0;
    if (P.type == 0) {
        if (uniform) nInterior = inner[0] > 1 ? (inner[0] - 1) * (inner[0] - 1) : 0;
        else if (inner[0] > 1 && inner[1] > 1) nInterior = (inner[0] - 1) * (inner[1] - 1);
        else if (out[0] == inner[0] && inner[0] == out[2] && out[1] == inner[1] && inner[1] == out[3])
            nInterior = 0;
        else nInterior = 1;
    } else if (P.type == 1) {
        const int r = inner[0];
        if (uniform) nInterior = r > 1 ? (r - 2) * (r - 1) / 2 : 0;
        else nInterior = r > 2 ? (r - 2) * (r - 1) / 2 : 1;
    } else {
        const int r = inner[0];
        if (r > 1) {
            const int e = r - 2, H = e / 2;
            nInterior = (e & 1) ? (H + 1) * (H + 1) * N + (N == 3 ? 0 : 1) : H * (H + 1) * N + 1;
        } else {
            nInterior = (uniform && N == 3) ? 0 : 1;
        }
    }
    const int count = total + nInterior;
    if (!uv) {
        free(heap);
        return count;
    }
    int n = 0;
    /* boundary */
    if (P.type == 0) {
        n += tess_iso_v(uv + 2 * n, out[0], 0.0f, 0.0f, 1.0f / (float)out[0]);
        n += tess_iso_u(uv + 2 * n, out[1], 1.0f, 0.0f, 1.0f / (float)out[1]);
        n += tess_iso_v(uv + 2 * n, out[2], 1.0f, 1.0f, -1.0f / (float)out[2]);
        n += tess_iso_u(uv + 2 * n, out[3], 0.0f, 1.0f, -1.0f / (float)out[3]);
    } else if (P.type == 1) {
        n += tess_iso_v(uv + 2 * n, out[0], 0.0f, 0.0f, 1.0f / (float)out[0]);
        n += tess_uv_line(uv + 2 * n, out[1], 1.0f, 0.0f, -1.0f / (float)out[1],
                          1.0f / (float)out[1]);
        n += tess_iso_u(uv + 2 * n, out[2], 0.0f, 1.0f, -1.0f / (float)out[2]);
    } else {
        for (int i = 0; i < N; ++i)
            n += tess_qsub_edge(&P, uv + 2 * n, i, out[i], 1, 0.0f, 1.0f / (float)out[i]);
    }
    /* interior */
    if (nInterior == 1) {
        n += tess_point(uv + 2 * n, P.type == 1 ? 1.0f / 3.0f : 0.5f,
                        P.type == 1 ? 1.0f / 3.0f : 0.5f);
    } else if (nInterior > 1 && P.type == 0) {
        const int rings = inner[0] / 2 < inner[1] / 2 ? inner[0] / 2 : inner[1] / 2;
        const float du = 1.0f / (float)inner[0], dv = 1.0f / (float)inner[1];
        float u = du, v = dv;
        int ur = inner[0] - 2, vr = inner[1] - 2;
        for (int i = 0; i < rings; ++i, ur -= 2, vr -= 2, u += du, v += dv) {
            if (ur > 0 && vr > 0) {
                const float u1 = 1.0f - u, v1 = 1.0f - v;
                n += tess_iso_v(uv + 2 * n, ur, u, v, du);
                n += tess_iso_u(uv + 2 * n, vr, u1, v, dv);
                n += tess_iso_v(uv + 2 * n, ur, u1, v1, -du);
                n += tess_iso_u(uv + 2 * n, vr, u, v1, -dv);
            } else if (ur > 0) {
                n += tess_iso_v(uv + 2 * n, ur + 1, u, v, du);
            } else if (vr > 0) {
                n += tess_iso_u(uv + 2 * n, vr + 1, u, v, dv);
            } else {
                n += tess_point(uv + 2 * n, 0.5f, 0.5f);
            }
        }
    } else if (nInterior > 1 && P.type == 1) {
        const int r = inner[0], rings = r / 3;
        const float dt = 1.0f / (float)r;
        float u = dt, v = dt;
        int rr = r - 3;
        for (int i = 0; i < rings; ++i, rr -= 3, u += dt, v += dt) {
            if (rr == 0) {
                n += tess_point(uv + 2 * n, 1.0f / 3.0f, 1.0f / 3.0f);
            } else {
                const float u1 = 1.0f - u * 2.0f, v1 = 1.0f - v * 2.0f;
                n += tess_iso_v(uv + 2 * n, rr, u, v, dt);
                n += tess_uv_line(uv + 2 * n, rr, u1, v, -dt, dt);
                n += tess_iso_u(uv + 2 * n, rr, u, v1, -dt);
            }
        }
    } else if (nInterior > 1) {
        const int r = inner[0], rings = r / 2;
        const float dt = 1.0f / (float)r;
        float t = dt;
        int rr = r - 2;
        for (int i = 0; i < rings; ++i, rr -= 2, t += dt) {
            if (rr == 0) {
                n += tess_point(uv + 2 * n, 0.5f, 0.5f);
            } else if (rr == 1) {
                for (int k = 0; k < N; ++k) {
                    float u0, v0;
                    tess_vertex(&P, k, &u0, &v0);
                    n += tess_point(uv + 2 * n, u0 + t, v0 + t);
                }
                if (N != 3) n += tess_point(uv + 2 * n, 0.5f, 0.5f);
            } else {
                for (int k = 0; k < N; ++k) n += tess_qsub_edge(&P, uv + 2 * n, k, rr, 1, t, dt);
            }
        }
    }
    free(heap);
    return n == count ? n : -1;
}

int oracle_tess_uniform_quad_f32(int rate, float *uv)
{
    if (rate < 1) return -1;
    int n = 0;
    n += tess_iso_v(uv + 2 * n, rate, 0.0f, 0.0f, 1.0f / (float)rate);
    n += tess_iso_u(uv + 2 * n, rate, 1.0f, 0.0f, 1.0f / (float)rate);
    n += tess_iso_v(uv + 2 * n, rate, 1.0f, 1.0f, -1.0f / (float)rate);
    n += tess_iso_u(uv + 2 * n, rate, 0.0f, 1.0f, -1.0f / (float)rate);
    const int rings = rate / 2;
    const float du = 1.0f / (float)rate, dv = 1.0f / (float)rate;
    float u = du, v = dv;
    int ur = rate - 2, vr = rate - 2;
    for (int i = 0; i < rings; ++i, ur -= 2, vr -= 2, u += du, v += dv) {
        if (ur > 0 && vr > 0) {
            const float u1 = 1.0f - u, v1 = 1.0f - v;
            n += tess_iso_v(uv + 2 * n, ur, u, v, du);
            n += tess_iso_u(uv + 2 * n, vr, u1, v, dv);
            n += tess_iso_v(uv + 2 * n, ur, u1, v1, -du);
            n += tess_iso_u(uv + 2 * n, vr, u, v1, -dv);
        } else if (ur > 0) {
            n += tess_iso_v(uv + 2 * n, ur + 1, u, v, du);
        } else if (vr > 0) {
            n += tess_iso_u(uv + 2 * n, vr + 1, u, v, dv);
        } else {
            uv[2 * n] = 0.5f;
            uv[2 * n + 1] = 0.5f;
            n += 1;
        }
    }
    return n;
}
