/*
 * CPU oracle for the Osd evaluator hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This library restates, in plain C, what OpenSubdiv's CpuEvaluator computes
 * for EvalStencils and EvalPatches, in float (the values CpuEvaluator
 * returns) and in double (the FP64 oracle of Far::StencilTableReal<double>
 * and Far::PatchTable::EvaluateBasis<double>).  It exists to check the HIP
 * evaluator; nothing in the shipped library links it.  Only the tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * Reference files restated (paths relative to the OpenSubdiv source tree):
 *   osd/cpuKernel.cpp:71-240        CpuEvalStencils (generic, du/dv, 2nd-deriv)
 *   osd/cpuEvaluator.cpp:37-125     EvalStencils argument checks
 *   osd/cpuEvaluator.cpp:127-381    EvalPatches (P / +d1 / +d2)
 *   osd/patchBasis.h:55-97          bilinear basis
 *   osd/patchBasis.h:101-245        cubic B-spline basis + boundary fold
 *   osd/patchBasis.h:250-330        cubic Bezier curve basis
 *   osd/patchBasis.h:334-490        Gregory (20-point) basis, approximate derivs
 *   osd/patchBasis.h:1447-1610      dispatch, normalization, 2^depth scaling
 *   osd/patchBasisTypes.h:298-397   PatchParam bitfield decode and Normalize
 *   far/stencilTable.h:616-674      StencilTableReal<double>::UpdateValues
 *   tutorials/far/tutorial_2_3/far_tutorial_2_3.cpp:73-87  limit normal (cross, normalize)
 *
 * Parity pins (each function is checked by tests/ against at least one of):
 *   - golden vectors written by the real reference (tests/golden/*.npz),
 *   - closed forms (textbook cubic B-spline / Bezier bases, bilinear),
 *   - invariants (partition of unity, zero-sum derivatives, affine
 *     reproduction, finite differences, boundary-fold identities),
 *   - brute force on tiny inputs (stencils).
 */
#ifndef OSD_ORACLE_H
#define OSD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int offset, length, stride; } oracle_desc;

/* Osd::PatchCoord / Osd::PatchArray / Osd::PatchParam binary layouts
 * (osd/types.h:42-130, osd/patchBasisTypes.h:249-315). */
typedef struct { int arrayIndex, patchIndex, vertIndex; float s, t; } oracle_patch_coord;
typedef struct { int regDesc, desc, numPatches, indexBase, stride, primitiveIdBase; } oracle_patch_array;
typedef struct { unsigned int field0, field1; float sharpness; } oracle_patch_param;

/* --- stencils ----------------------------------------------------------- */

/* CpuEvalStencils generic path (osd/cpuKernel.cpp:71-122): stencils
 * [start,end) are read, row (i - start) of dst is written.  Returns 0 on the
 * argument errors CpuEvaluator::EvalStencils rejects (length mismatch),
 * 1 otherwise (including end <= start, which is a no-op). */
int oracle_eval_stencils_f32(const float *src, oracle_desc srcDesc,
                             float *dst, oracle_desc dstDesc,
                             const int *sizes, const int *offsets,
                             const int *indices, const float *weights,
                             int start, int end);

/* du/dv variant (osd/cpuKernel.cpp:124-170). */
int oracle_eval_stencils_d1_f32(const float *src, oracle_desc srcDesc,
                                float *dst, oracle_desc dstDesc,
                                float *du, oracle_desc duDesc,
                                float *dv, oracle_desc dvDesc,
                                const int *sizes, const int *offsets,
                                const int *indices, const float *weights,
                                const float *duWeights, const float *dvWeights,
                                int start, int end);

/* du/dv/duu/duv/dvv variant (osd/cpuKernel.cpp:172-240). */
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
                                int start, int end);

/* FP64 twin (far/stencilTable.h:616-674 with REAL=double): double weights,
 * double source and destination values, same row convention as above. */
int oracle_eval_stencils_f64(const double *src, oracle_desc srcDesc,
                             double *dst, oracle_desc dstDesc,
                             const int *sizes, const int *offsets,
                             const int *indices, const double *weights,
                             int start, int end);

/* Multi-threaded (OpenMP) float variant, the OmpEvaluator baseline
 * (osd/ompKernel.cpp:70-120): same results, rows written at (i - start). */
int oracle_eval_stencils_omp_f32(const float *src, oracle_desc srcDesc,
                                 float *dst, oracle_desc dstDesc,
                                 const int *sizes, const int *offsets,
                                 const int *indices, const float *weights,
                                 int start, int end, int numThreads);

/* --- patch basis ---------------------------------------------------------- */

/* OsdEvaluatePatchBasis (osd/patchBasis.h:1558-1610): normalizes (s,t) with
 * the PatchParam, evaluates the basis of patchType, scales derivatives by
 * 2^depth (first) and 4^depth (second).  Any of wDs..wDtt may be NULL
 * (wDs and wDt together; wDss, wDst and wDtt together).  Returns the number
 * of points (16, 20, 4) or 0 for an unsupported type. */
int oracle_patch_basis_f32(int patchType, unsigned field0, unsigned field1,
                           float s, float t,
                           float *wP, float *wDs, float *wDt,
                           float *wDss, float *wDst, float *wDtt);
int oracle_patch_basis_f64(int patchType, unsigned field0, unsigned field1,
                           double s, double t,
                           double *wP, double *wDs, double *wDt,
                           double *wDss, double *wDst, double *wDtt);

/* --- patches --------------------------------------------------------------- */

/* CpuEvaluator::EvalPatches (osd/cpuEvaluator.cpp:157-381).  dst may be NULL
 * only when derivatives are requested (the reference crashes on a NULL dst;
 * here a NULL output is skipped).  Returns 0 for a NULL src or a length
 * mismatch, as the reference does. */
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
                            const oracle_patch_param *params);

/* FP64 twin: double source/outputs, double basis (Far::PatchTable
 * EvaluateBasis<double>, far/patchTable.cpp:581-649). */
int oracle_eval_patches_f64(const double *src, oracle_desc srcDesc,
                            double *dst, oracle_desc dstDesc,
                            double *du, oracle_desc duDesc,
                            double *dv, oracle_desc dvDesc,
                            int numPatchCoords,
                            const oracle_patch_coord *coords,
                            const oracle_patch_array *arrays,
                            const int *patchIndices,
                            const oracle_patch_param *params);

/* OpenMP float variant, the OmpEvaluator::EvalPatches baseline
 * (osd/ompEvaluator.cpp:163-350). */
int oracle_eval_patches_omp_f32(const float *src, oracle_desc srcDesc,
                                float *dst, oracle_desc dstDesc,
                                float *du, oracle_desc duDesc,
                                float *dv, oracle_desc dvDesc,
                                int numPatchCoords,
                                const oracle_patch_coord *coords,
                                const oracle_patch_array *arrays,
                                const int *patchIndices,
                                const oracle_patch_param *params,
                                int numThreads);

/* --- the step after EvalPatches: limit normals ------------------------- */

/* N = cross(dPdu, dPdv) over xyz (the first three elements), optionally
 * scaled to unit length as 1 / sqrtf(dot) -- the reference tutorial's
 * `cross` and `normalize` (tutorials/far/tutorial_2_3/far_tutorial_2_3.cpp:73-87).
 * Returns 0 when a descriptor has fewer than 3 elements, 1 otherwise. */
int oracle_limit_normals_f32(const float *du, oracle_desc duDesc, const float *dv,
                             oracle_desc dvDesc, float *n, oracle_desc nDesc, int count,
                             int normalize);

#ifdef __cplusplus
}
#endif
/* Bfr::Tessellation::GetCoords<float> of a quad at a uniform rate
 * (bfr/tessellation.cpp:158-174, 929-990): (rate + 1)^2 (u, v) pairs. */
int oracle_tess_uniform_quad_f32(int rate, float *uv);
/* Bfr::Tessellation::GetCoords<float> for any parameterization (scheme 0
 * bilinear / 1 Catmark / 2 Loop, faceSize) and rates as Tessellation takes
 * them (bfr/tessellation.cpp:2023-2371); uv NULL returns the count only;
 * -1 for invalid arguments. */
int oracle_tess_coords_f32(int scheme, int faceSize, int numRates, const int *rates, float *uv);

#endif
