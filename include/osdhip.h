/*
 * osdhip.h -- C ABI of the MI355X (gfx950) Osd evaluator library
 * (opensubdiv_amd/lib/libosdhip.so).
 *
 * Plain C: pointers, sizes and small POD structs only.  Every float / int
 * pointer named src, dst, du, dv, duu, duv, dvv, sizes, offsets, indices,
 * *weights, coords, arrays, params is a DEVICE pointer.  `stream` is a
 * hipStream_t (NULL = the null stream).  Launches are asynchronous with
 * respect to the host; call osdhip_stream_synchronize() before reading
 * results.  No entry point throws.
 *
 * Return codes: OSDHIP_OK (0) on success; OSDHIP_ERR_ARGS (-1) when the
 * arguments are rejected the way Osd::CpuEvaluator rejects them (length
 * mismatch, NULL src/dst); OSDHIP_ERR_UNSUPPORTED (-2) for a request the
 * library does not implement; OSDHIP_ERR_TOO_LARGE (-3) for a request
 * beyond a fixed-size entry point's limit; a positive value is a
 * hipError_t.
 *
 * Interfaces replaced (OpenSubdiv v3.6.0, paths relative to its tree):
 *   osd/cudaEvaluator.cpp:33-68     extern "C" CudaEvalStencils /
 *                                   CudaEvalPatches / CudaEvalPatchesWithDerivatives
 *   osd/cudaEvaluator.h:159-178     CudaEvaluator::EvalStencils (raw)
 *   osd/cudaEvaluator.h:272-296     ... with du/dv
 *   osd/cudaEvaluator.h:437-470     ... with du/dv/duu/duv/dvv
 *   osd/cudaEvaluator.h:700-830     CudaEvaluator::EvalPatches (raw, 3 arities)
 *   osd/cudaEvaluator.cpp:377-380   CudaEvaluator::Synchronize
 *   osd/cudaVertexBuffer.cpp:34-94  device buffer allocation / UpdateData
 */
#ifndef OSDHIP_H
#define OSDHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OSDHIP_OK 0
#define OSDHIP_ERR_ARGS (-1)
#define OSDHIP_ERR_UNSUPPORTED (-2)
#define OSDHIP_ERR_TOO_LARGE (-3)

/* Osd::BufferDescriptor (osd/bufferDescriptor.h:61-103), in floats. */
typedef struct { int offset, length, stride; } osdhip_buffer_desc;

/* Osd::PatchCoord (osd/types.h:42-64): 20 bytes, s/t in ptex-face space. */
typedef struct { int arrayIndex, patchIndex, vertIndex; float s, t; } osdhip_patch_coord;

/* Osd::PatchArray (osd/types.h:66-125): 6 ints. */
typedef struct {
    int regDesc, desc, numPatches, indexBase, stride, primitiveIdBase;
} osdhip_patch_array;

/* Osd::PatchParam (osd/types.h:127-130, far/patchParam.h:80-97): field0
 * faceId[0:28] transition[28:32]; field1 depth[0:4] nonquad[4] regular[5]
 * boundary[7:12] v[12:22] u[22:32]; then the crease sharpness. */
typedef struct { unsigned int field0, field1; float sharpness; } osdhip_patch_param;

/* Patch type ids (osd/patchBasisTypes.h:241-246). */
#define OSDHIP_PATCH_QUADS 3
#define OSDHIP_PATCH_TRIANGLES 4
#define OSDHIP_PATCH_LOOP 5
#define OSDHIP_PATCH_REGULAR 6
#define OSDHIP_PATCH_GREGORY_BASIS 9
#define OSDHIP_PATCH_GREGORY_TRIANGLE 10

/* ---- stencils ----------------------------------------------------------
 * dst row r (r = 0 .. end-start-1, relative to dstDesc.offset) =
 *   sum_{j < sizes[i]} weights[offsets[i]+j] * src[indices[offsets[i]+j]]
 * for i = start + r, accumulated in fp32.  The row convention follows
 * CpuEvalStencils' generic path (osd/cpuKernel.cpp:104-121); every templated
 * reference entry point passes start = 0, where all backends agree.
 * end <= start is a no-op returning OSDHIP_OK; srcDesc.length != any output
 * length returns OSDHIP_ERR_ARGS (osd/cpuEvaluator.cpp:46-47). */

/* replaces CudaEvaluator::EvalStencils raw (osd/cudaEvaluator.h:159-178) and
 * extern "C" CudaEvalStencils (osd/cudaEvaluator.cpp:33-40) */
int osdhip_eval_stencils(const float *src, osdhip_buffer_desc srcDesc,
                         float *dst, osdhip_buffer_desc dstDesc,
                         const int *sizes, const int *offsets,
                         const int *indices, const float *weights,
                         int start, int end, void *stream);

/* replaces CudaEvaluator::EvalStencils with du/dv (osd/cudaEvaluator.h:
 * 272-296; osd/cudaEvaluator.cpp:172-214 makes 3 launches, this makes one).
 * A NULL output (dst, du or dv) is skipped, as the CUDA backend skips it. */
int osdhip_eval_stencils_d1(const float *src, osdhip_buffer_desc srcDesc,
                            float *dst, osdhip_buffer_desc dstDesc,
                            float *du, osdhip_buffer_desc duDesc,
                            float *dv, osdhip_buffer_desc dvDesc,
                            const int *sizes, const int *offsets,
                            const int *indices, const float *weights,
                            const float *duWeights, const float *dvWeights,
                            int start, int end, void *stream);

/* replaces CudaEvaluator::EvalStencils with 1st+2nd derivatives
 * (osd/cudaEvaluator.h:437-470; osd/cudaEvaluator.cpp:217-292). */
int osdhip_eval_stencils_d2(const float *src, osdhip_buffer_desc srcDesc,
                            float *dst, osdhip_buffer_desc dstDesc,
                            float *du, osdhip_buffer_desc duDesc,
                            float *dv, osdhip_buffer_desc dvDesc,
                            float *duu, osdhip_buffer_desc duuDesc,
                            float *duv, osdhip_buffer_desc duvDesc,
                            float *dvv, osdhip_buffer_desc dvvDesc,
                            const int *sizes, const int *offsets,
                            const int *indices, const float *weights,
                            const float *duWeights, const float *dvWeights,
                            const float *duuWeights, const float *duvWeights,
                            const float *dvvWeights,
                            int start, int end, void *stream);

/* ---- device stencil tables (HipStencilTable) ------------------------------
 * replaces CudaStencilTable (osd/cudaEvaluator.h:52-91,
 * osd/cudaEvaluator.cpp:75-145): owns the device copy of a Far
 * (Limit)StencilTable.  Besides the Far CSR arrays (returned by
 * osdhip_stencil_table_csr for the raw entry points) it holds a 64-row
 * sliced layout (column-major per slice) that osdhip_stencil_table_eval
 * streams with fully coalesced loads. */
typedef struct osdhip_stencil_table osdhip_stencil_table;

/* HOST arrays in; numWeightArrays = 1 (P), 3 (P, du, dv) or 6 (+ duu, duv,
 * dvv); weights[k] = host pointer of array k; offsets may be NULL (then the
 * exclusive prefix sum of sizes).  Returns NULL on invalid input or when a
 * device allocation fails.  Default options (osdhip_table_options_init). */
osdhip_stencil_table *osdhip_stencil_table_create(int numStencils, const int *sizes,
                                                  const int *offsets, const int *indices,
                                                  int numWeightArrays,
                                                  const float *const *weights, void *stream);

/* How a table is laid out on the device and evaluated.  Everything is fixed
 * per table (nothing process-wide), so tables built with different options
 * can be used concurrently.  Results are the same bits whatever the options.
 *   layout:  0 = one interleaved record {index, weights} per entry in 64-row
 *            slices (default); 1 = index and weights in separate streams;
 *            2 = locality-ordered slices: rows grouped by their dominant
 *            control vertex (see order), each slice's distinct control
 *            vertices staged in LDS (for multi-frame batches and large
 *            refinement tables);
 *            3 = paired records for tables of one weight array: a row's
 *            entries 2p and 2p + 1 side by side, read with one 16-byte load
 *            (NULL for tables of 3 or 6 weight arrays; not for batches).
 *   keepCsr: 1 (default) = the Far CSR arrays stay on the device
 *            (osdhip_stencil_table_csr, the CSR kernel); 0 = only the slice
 *            layout is kept (the CSR getters return NULL).
 *   group:   entries per gather group of the slice kernels: 4, 8 (default)
 *            or 16 (layout 2 uses 4, or 8 for 8 and 16).
 *   kernel:  osdhip_stencil_table_eval kernel: 0 = slice layout (default),
 *            1 = CSR with LDS staging (needs keepCsr).
 *   order:   layout 2's row order: 0 = by dominant control vertex, then
 *            by row size within windows of 128 or 512 rows when that
 *            leaves fewer padded record entries and no fewer slices
 *            listing their control vertices (default: the fewest entries
 *            streamed per evaluation); 2 = by dominant control vertex only
 *            (the fewest control vertices per slice: for many-frame
 *            batches, which re-read records from cache but stage control
 *            vertices per frame chunk); 1 = Far's row order (a slice's
 *            outputs are 64 consecutive rows; slices with more distinct
 *            control vertices than the LDS list holds gather them from
 *            memory). */
typedef struct {
    int layout;
    int keepCsr;
    int group;
    int kernel;
    int order;
} osdhip_table_options;

void osdhip_table_options_init(osdhip_table_options *options);
/* osdhip_stencil_table_create with explicit options (NULL = defaults);
 * NULL also for options out of range. */
osdhip_stencil_table *osdhip_stencil_table_create_ex(int numStencils, const int *sizes,
                                                     const int *offsets, const int *indices,
                                                     int numWeightArrays,
                                                     const float *const *weights,
                                                     const osdhip_table_options *options,
                                                     void *stream);
/* Changes an evaluation option of one table, "group" or "kernel" (the
 * layout options are fixed at creation); returns the previous value,
 * OSDHIP_ERR_ARGS for a value out of range, OSDHIP_ERR_UNSUPPORTED for
 * another key.  Not to be called while the table is being evaluated. */
int osdhip_stencil_table_set(osdhip_stencil_table *table, const char *key, int value);
/* The table's options as built. */
int osdhip_stencil_table_options(const osdhip_stencil_table *table,
                                 osdhip_table_options *options);
void osdhip_stencil_table_destroy(osdhip_stencil_table *table);
int osdhip_stencil_table_info(const osdhip_stencil_table *table, int *numStencils,
                              long long *numEntries, int *numWeightArrays,
                              long long *deviceBytes);
/* The table's device layout: layout (see osdhip_table_options), number of
 * 64-row slices and, for layout 2, the slices whose control vertices are
 * staged in LDS and the total length of their control lists.  No
 * counterpart in the reference: a diagnostic of this library's layouts. */
int osdhip_stencil_table_layout(const osdhip_stencil_table *table, int *layout, int *numSlices,
                                int *listedSlices, long long *listEntries);
/* Record entries the table's slices hold (64 x the widest row of every
 * slice: the entries each evaluation streams, padding included) and, for
 * layout 2 with rows by dominant vertex, the window (0, 128 or 512 rows)
 * within which rows were regrouped by size to make a slice's rows alike.
 * No counterpart in the reference: a diagnostic of this library's layouts. */
int osdhip_stencil_table_slices(const osdhip_stencil_table *table, long long *recordEntries,
                                int *sizeWindow);
/* DEVICE pointers of the Far-layout arrays; weights receives 6 pointers
 * (NULL for arrays the table does not have). */
int osdhip_stencil_table_csr(const osdhip_stencil_table *table, const int **sizes,
                             const int **offsets, const int **indices, const float **weights);
/* EvalStencils over stencils [start, end) of a device table (templated
 * CudaEvaluator::EvalStencils, osd/cudaEvaluator.h:125-143, 217-241,
 * 352-385): numOutputs = 1, 3 or 6 (at most the table's weight arrays);
 * outputs[k] / outputDescs[k] as the raw entry points; NULL outputs are
 * skipped.  One launch for all outputs. */
int osdhip_stencil_table_eval(const osdhip_stencil_table *table, const float *src,
                              osdhip_buffer_desc srcDesc, int numOutputs, float *const *outputs,
                              const osdhip_buffer_desc *outputDescs, int start, int end,
                              void *stream);

/* ---- patches -------------------------------------------------------------
 * replaces CudaEvaluator::EvalPatches raw, all three arities
 * (osd/cudaEvaluator.h:700-830) and extern "C" CudaEvalPatches /
 * CudaEvalPatchesWithDerivatives (osd/cudaEvaluator.cpp:42-68).
 * Output i (relative to each desc offset, with each desc stride) is the
 * limit value / derivative at coords[i], as CpuEvaluator::EvalPatches
 * (osd/cpuEvaluator.cpp:157-381).  Any output pointer may be NULL (skipped);
 * src == NULL or an output length != srcDesc.length returns OSDHIP_ERR_ARGS.
 * Patch types: REGULAR (B-spline, with boundary fold), GREGORY_BASIS,
 * QUADS (bilinear), LOOP (box spline, with boundary phantom points),
 * GREGORY_TRIANGLE and TRIANGLES (linear), every type
 * OsdEvaluatePatchBasis evaluates (osd/patchBasis.h:1447-1610); any other
 * type produces zeros, as CpuEvaluator's Clear() with no point added.
 * Second derivatives on rotated triangles keep the reference's sign
 * (-4^depth, osd/patchBasis.h:1596). */
int osdhip_eval_patches(const float *src, osdhip_buffer_desc srcDesc,
                        float *dst, osdhip_buffer_desc dstDesc,
                        float *du, osdhip_buffer_desc duDesc,
                        float *dv, osdhip_buffer_desc dvDesc,
                        float *duu, osdhip_buffer_desc duuDesc,
                        float *duv, osdhip_buffer_desc duvDesc,
                        float *dvv, osdhip_buffer_desc dvvDesc,
                        int numPatchCoords,
                        const osdhip_patch_coord *coords,
                        const osdhip_patch_array *arrays,
                        const int *patchIndices,
                        const osdhip_patch_param *params,
                        void *stream);
/* osdhip_eval_patches evaluated in patch order: the coords are binned by
 * patch (1024 consecutive patches per bin; up to 4 MB + 4 bytes per coord
 * of stream-ordered scratch memory, hipMallocAsync; caller order, with the
 * same results, when that allocation fails), then each result still goes to
 * its coord's own row.  Same results, bit for bit.  Pays off when many
 * coords share patches in no particular order.  No counterpart in the
 * reference. */
int osdhip_eval_patches_sorted(const float *src, osdhip_buffer_desc srcDesc,
                               float *dst, osdhip_buffer_desc dstDesc,
                               float *du, osdhip_buffer_desc duDesc,
                               float *dv, osdhip_buffer_desc dvDesc,
                               float *duu, osdhip_buffer_desc duuDesc,
                               float *duv, osdhip_buffer_desc duvDesc,
                               float *dvv, osdhip_buffer_desc dvvDesc,
                               int numPatchCoords,
                               const osdhip_patch_coord *coords,
                               const osdhip_patch_array *arrays,
                               const int *patchIndices,
                               const osdhip_patch_param *params,
                               void *stream);

/* osdhip_eval_patches_sorted without any allocation: the same kernels and
 * results, the sort's workspace in a fixed device array the library holds
 * on each GPU it runs on (OSDHIP_NOALLOC_SCRATCH_BYTES, kept until the
 * process ends).  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  At most OSDHIP_NOALLOC_MAX_PATCH_COORDS
 * coords per call (OSDHIP_ERR_TOO_LARGE beyond).  Calls share their GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race;
 * calls ordered on one stream do not. */
#define OSDHIP_NOALLOC_MAX_PATCH_COORDS (1 << 22)
#define OSDHIP_NOALLOC_SCRATCH_BYTES (25u << 20)
int osdhip_eval_patches_sorted_noalloc(const float *src, osdhip_buffer_desc srcDesc,
                                       float *dst, osdhip_buffer_desc dstDesc,
                                       float *du, osdhip_buffer_desc duDesc,
                                       float *dv, osdhip_buffer_desc dvDesc,
                                       float *duu, osdhip_buffer_desc duuDesc,
                                       float *duv, osdhip_buffer_desc duvDesc,
                                       float *dvv, osdhip_buffer_desc dvvDesc,
                                       int numPatchCoords,
                                       const osdhip_patch_coord *coords,
                                       const osdhip_patch_array *arrays,
                                       const int *patchIndices,
                                       const osdhip_patch_param *params,
                                       void *stream);

/* ---- PatchCoords on the device: Far::PatchMap ------------------------------
 * Far::PatchMap (far/patchMap.h:58-120, far/patchMap.cpp:88-200) held on the
 * device.  osdhip_patch_map_create builds, on the host, the quadtree Far's
 * constructor builds from a patch table -- here from the Osd flattening of
 * its PatchArrays and PatchParams (HOST pointers; osd/cpuPatchTable.cpp:
 * 35-158) -- and uploads it with the patch handles; NULL for an empty or
 * inconsistent table.  osdhip_patch_map_find is PatchMap::FindPatch
 * (far/patchMap.h:193-231) for n (ptex face, u, v) queries in one launch
 * (DEVICE pointers): coords[i] = {handle, s = (float)u, t = (float)v} with
 * Far's handle {arrayIndex, patchIndex, vertIndex = first point within the
 * array}; a query outside the map (no such face, u or v outside [0, 1])
 * gets {-1, -1, -1}.  Same records as Far's PatchMap, including the
 * upper-half rule at every median (u >= median goes to the upper half). */
typedef struct osdhip_patch_map osdhip_patch_map;
osdhip_patch_map *osdhip_patch_map_create(const osdhip_patch_array *arrays, int numArrays,
                                          const osdhip_patch_param *params, int numPatches,
                                          void *stream);
int osdhip_patch_map_find(const osdhip_patch_map *map, int n, const int *ptexFaces,
                          const double *u, const double *v, osdhip_patch_coord *coords,
                          void *stream);
void osdhip_patch_map_destroy(osdhip_patch_map *map);
/* Bfr-style per-face queries (Bfr::Surface::Evaluate, bfr/surface.h:
 * 227-266): query i is base face faces[i] at (u[i], v[i]) in Bfr's
 * parameterization of that face (bfr/parameterization.h:60-200) -- a
 * quad's unit square, or for an n-gon its n quad sub-faces tiled on a grid
 * (sub-face k on tile (k % uDim, k / uDim), uDim = 2, 3 for n < 10, else
 * 1 + floor(sqrt(n - 1)); normalized coordinates twice the offset into the
 * tile).  Each is mapped to its ptex face (ptexBase[f] for a quad,
 * ptexBase[f] + k for sub-face k; faceSizes and ptexBase: numFaces device
 * ints per base face) and located as osdhip_patch_map_find does; coords[i]
 * carries the ptex (s, t).  On a patch table built with Bfr's approximation
 * options (osdfar_patch_options), osdhip_eval_patches at these coords
 * gives Bfr::Surface::Evaluate's P, Du, Dv (derivatives with respect to
 * the quad's (u, v) or the sub-face's normalized coordinates, as Bfr's). */
int osdhip_patch_map_find_faces(const osdhip_patch_map *map, int n, const int *faces,
                                const double *u, const double *v, const int *faceSizes,
                                const int *ptexBase, int numFaces, osdhip_patch_coord *coords,
                                void *stream);

/* Bfr::Tessellation (bfr/tessellation.h:51-240) of quad parameterizations
 * at a uniform rate, on the device, for numFaces ptex faces at once: each
 * face's (rate + 1)^2 coordinates in Bfr's order -- the boundary
 * counter-clockwise from (0,0), then the interior rings from the outside
 * in -- with the same float values as Tessellation::GetCoords<float>.
 * Written face-major as (u, v) pairs to `uv` (may be NULL) and, with a
 * patch map, as the PatchCoords of (ptexFaces[f], u, v) to `coords`, ready
 * for osdhip_eval_patches (no host round trip).  rate in [1, 4096];
 * OSDHIP_ERR_ARGS when nothing is to be written, coords without a map or a
 * map without coords / faces, or more than 2^31 - 1 coordinates. */
int osdhip_tess_uniform_quads(int numFaces, const int *ptexFaces, int rate, float *uv,
                              const osdhip_patch_map *map, osdhip_patch_coord *coords,
                              void *stream);

/* Bfr::Tessellation coordinates (bfr/tessellation.h:51-240) for any
 * parameterization and rates, on the device.  scheme: 0 bilinear or 1
 * Catmark (a quad's unit square; other face sizes: Bfr's quad sub-faces),
 * 2 Loop (triangles).  rates as Tessellation(param, numRates, rates) takes
 * them: 1 uniform rate, 2 inner rates of a quad, N outer (edge) rates, or N
 * outer plus 1 (2 for a quad) inner rates.  osdhip_tess_count returns the
 * coordinates per face (GetNumCoords), OSDHIP_ERR_ARGS when invalid.
 * osdhip_tess_coords writes the pattern for numFaces faces: face-major
 * (u, v) float pairs to `uv` (may be NULL), in GetCoords order with the
 * same float values (boundary first, then the interior); with a patch map,
 * also the PatchCoords of base face faces[f] at each coordinate, located
 * as osdhip_patch_map_find_faces does (faceSizes / ptexBase / numBaseFaces
 * as there).  OSDHIP_ERR_ARGS when nothing is to be written, for coords
 * without a map (or a map without faces / coords), or more than 2^31 - 1
 * coordinates. */
int osdhip_tess_count(int scheme, int faceSize, int numRates, const int *rates);
int osdhip_tess_coords(int scheme, int faceSize, int numRates, const int *rates, int numFaces,
                       float *uv, const int *faces, const osdhip_patch_map *map,
                       const int *faceSizes, const int *ptexBase, int numBaseFaces,
                       osdhip_patch_coord *coords, void *stream);

/* ---- the step after EvalPatches: limit normals ------------------------------
 * normals[i] = cross(dPdu[i], dPdv[i]) over xyz (the first three elements of
 * each descriptor), scaled to unit length by 1 / sqrtf(dot) when `normalize`
 * is nonzero: the steps the reference's tutorial uses for limit normals
 * (tutorials/far/tutorial_2_3/far_tutorial_2_3.cpp:73-87, 404-412); Osd has no
 * evaluator entry point for them.  Same results as those steps, bit for bit.
 * OSDHIP_ERR_ARGS when a descriptor has fewer than 3 elements or a buffer
 * is NULL with count > 0. */
int osdhip_limit_normals(const float *dPdu, osdhip_buffer_desc duDesc, const float *dPdv,
                         osdhip_buffer_desc dvDesc, float *normals, osdhip_buffer_desc nDesc,
                         int count, int normalize, void *stream);

/* ---- device memory and streams -------------------------------------------
 * Used by the C++ HipVertexBuffer / HipStencilTable / HipPatchTable
 * (mirroring osd/cudaVertexBuffer.cpp:34-94 and cudaEvaluator.cpp:75-145). */
int osdhip_malloc(void **ptr, size_t bytes);
int osdhip_free(void *ptr);
int osdhip_memcpy_htod(void *dst, const void *src, size_t bytes, void *stream);
int osdhip_memcpy_dtoh(void *dst, const void *src, size_t bytes, void *stream);
int osdhip_memcpy_dtod(void *dst, const void *src, size_t bytes, void *stream);
/* replaces CudaEvaluator::Synchronize (osd/cudaEvaluator.cpp:377-380) */
int osdhip_stream_synchronize(void *stream);
int osdhip_device_synchronize(void);
/* Writes the device name (NUL-terminated) into buf; returns the device
 * count, or a negative value when no device is visible. */
int osdhip_device_info(char *buf, int buflen);
/* Library version string, e.g. "osdhip 0.1 gfx950". */
const char *osdhip_version(void);

/* ---- batched evaluation -----------------------------------------------
 * No counterpart in the reference (its evaluators take one mesh and one
 * frame per call).  For callers that evaluate many meshes, each with many
 * frames of control points, per step: the batch captures all of it into
 * one HIP graph (phase A: every stencil segment; phase B: every patch
 * segment, which may read the vertex buffers phase A writes), replayed by
 * osdhip_batch_launch.  Element (v, f, k) of a multi-frame buffer is at
 * base + v * vertexStride + f * frameStride + k (floats); `length` floats
 * (1..8) per element.  Results equal osdhip_stencil_table_eval /
 * osdhip_eval_patches of each (segment, frame) bit for bit.  The batch
 * keeps the pointers it is given: buffers and tables must outlive it. */
typedef struct osdhip_batch osdhip_batch;

typedef struct {
    const osdhip_stencil_table *table;   /* one weight array; layout 0 or 2 (all segments alike) */
    const float *src;
    long long srcVertexStride, srcFrameStride;
    float *dst;                          /* row r of the table -> element r */
    long long dstVertexStride, dstFrameStride;
    int frames;
    int stage;                           /* see below */
} osdhip_batch_stencils;

typedef struct {
    int numCoords;
    const osdhip_patch_coord *coords;
    const osdhip_patch_array *arrays;
    const int *patchIndices;
    const osdhip_patch_param *params;
    const float *src;                    /* patch points [control | refined | end cap] */
    long long srcVertexStride, srcFrameStride;
    float *P, *dPdu, *dPdv;              /* NULL = not written; coord c -> element c */
    long long outVertexStride, outFrameStride;
    int frames;
    int stage;
} osdhip_batch_patches;

/* Stages: a launch runs stage 0, 1, ... in order, and within a stage every
 * stencil segment before any patch segment, so segments of different
 * stages may share output memory (a batch whose outputs exceed the device
 * can stream them through one arena, each output still written once per
 * launch).  osdhip_batch_epoch ignores `stage` (an epoch is one stage). */

/* NULL on invalid segments (tables in the split-stream layout, tables of
 * both slice layouts, more than one weight array, strides below `length`)
 * or allocation failure.  For every patch segment the batch lists, per 64
 * coords, the distinct points of its regular patches (staged phase B). */
osdhip_batch *osdhip_batch_create(int length, int numStencilSegments,
                                  const osdhip_batch_stencils *stencils,
                                  int numPatchSegments, const osdhip_batch_patches *patches,
                                  void *stream);
void osdhip_batch_destroy(osdhip_batch *batch);
/* One graph launch (reset, phase A, phase B) on `stream`. */
int osdhip_batch_launch(osdhip_batch *batch, void *stream);
/* The same kernels without the graph: phases 1 = A, 2 = B, 3 = both; every
 * stage in order, or one stage (osdhip_batch_run_stage). */
int osdhip_batch_run(osdhip_batch *batch, int phases, void *stream);
int osdhip_batch_run_stage(osdhip_batch *batch, int stage, int phases, void *stream);
int osdhip_batch_num_stages(const osdhip_batch *batch);
/* The same work one mesh and one frame per launch through
 * osdhip_stencil_table_eval and osdhip_eval_patches (the plain evaluator
 * path), for comparison: every stencil segment, then every patch segment. */
int osdhip_batch_run_plain(osdhip_batch *batch, void *stream);
int osdhip_batch_run_plain_stage(osdhip_batch *batch, int stage, void *stream);
/* Variant knobs: "a_frames" x "a_group" (frames per chunk, gather group;
 * layout-0 tables: {1x4, 1x8, 2x4, 4x2} or a_frames 0 = a slice's records
 * stay in registers for all frames; layout-2 tables: {1, 2, 4} x {4, 8};
 * other pairs make launches return OSDHIP_ERR_UNSUPPORTED),
 * "a_item_slices" (64-row slices per work item), "a_persistent" /
 * "b_persistent" (0 = one block per item, 1 = fixed grid claiming items
 * from a device counter), "b_frames" in {1, 2, 4}, "b_stage" (1 = waves
 * whose regular-patch coords use at most 256 distinct points read them
 * from LDS, 0 = gathers only).  The graph is re-captured on the next
 * launch.  osdhip_batch_get also reads "staged_coords": coords in waves
 * that can be staged. */
int osdhip_batch_set(osdhip_batch *batch, const char *key, int value);
int osdhip_batch_get(const osdhip_batch *batch, const char *key);
int osdhip_batch_info(const osdhip_batch *batch, long long *stencilRows, long long *frameRows,
                      long long *patchCoords, long long *patchFrames, long long *tableBytes);
/* Inspection of a patch segment's point staging lists, built at batch
 * creation (no counterpart in the reference): per wave of 64 coords the
 * number of distinct points its regular-interior coords use (-1: more than
 * 256, not staged) and the offset of their ascending list in `pts`, and per
 * coord 4 words of slot bytes (byte j of a regular-interior coord: its
 * point j's position in the wave's list).  `numPts` (required) receives the
 * length of `pts`, -1 when the segment has no lists; null outputs are
 * skipped.  Host buffers of nw = ceil(ncoords / 64) (nuniq, ptsOff),
 * numPts (pts) and 4 * ncoords (slots) elements; synchronous. */
int osdhip_batch_patch_staging(const osdhip_batch *batch, int segment, int *nuniq,
                               long long *ptsOff, int *pts, unsigned *slots, long long *numPts);
/* Autotuner: times every variant of each phase (median of `reps` runs with
 * HIP events on `stream`), keeps the fastest and re-captures the graph.
 * `report` (optional) receives the timings as JSON. */
int osdhip_batch_tune(osdhip_batch *batch, int reps, void *stream, char *report, int reportBytes);

/* One-shot epoch of batch work, for serving loops whose segments change
 * every epoch: no graph, no allocation.  The segment and work-item tables
 * are written into `staging` (pinned host memory, hipHostMalloc), copied to
 * `scratch` (device memory) on `stream`, then phase A and phase B run on
 * `stream` (static scheduling, one frame per chunk, gather group 8).  Both
 * buffers need osdhip_batch_epoch_bytes() bytes and must not be reused
 * until the work queued on `stream` has completed.  Results equal
 * osdhip_stencil_table_eval / osdhip_eval_patches per segment and frame.
 * No counterpart in the reference. */
long long osdhip_batch_epoch_bytes(int length, int numStencilSegments,
                                   const osdhip_batch_stencils *stencils, int numPatchSegments,
                                   const osdhip_batch_patches *patches, int itemSlices);
int osdhip_batch_epoch(int length, int numStencilSegments, const osdhip_batch_stencils *stencils,
                       int numPatchSegments, const osdhip_batch_patches *patches, int itemSlices,
                       void *staging, long long stagingBytes, void *scratch,
                       long long scratchBytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* OSDHIP_H */
