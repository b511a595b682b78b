/*
 * osdfar.h -- C ABI of the host table compiler (opensubdiv_amd/lib/libosdfar.so).
 *
 * Builds, on the host, the tables the HIP evaluator consumes, for Catmark
 * meshes: uniform refinement with Far's vertex numbering, factorized stencil
 * tables, limit stencil tables with derivatives, patch tables
 * (B-spline patches + Gregory-basis end caps) and a patch map.  All pointers
 * are HOST pointers.  Functions returning int return 0 on success and a
 * negative value on invalid input (OSDFAR_ERR_HOST when host memory ran
 * out); no function throws.
 *
 * Interfaces restated (OpenSubdiv v3.6.0, paths relative to its tree):
 *   far/topologyDescriptor.h:40-80      TopologyDescriptor (input mesh)
 *   far/topologyRefiner.h:118-193       TopologyRefiner::RefineUniform
 *   far/stencilTableFactory.h:55-130    StencilTableFactoryReal<REAL>::Create
 *   far/stencilTable.h:141-297          StencilTableReal<REAL> (sizes/offsets/indices/weights)
 */
#ifndef OSDFAR_H
#define OSDFAR_H

#ifdef __cplusplus
extern "C" {
#endif

#define OSDFAR_ERR_HOST (-9)   /* host allocation failed */

typedef struct osdfar_topology osdfar_topology;
typedef struct osdfar_stencil_table osdfar_stencil_table;

/* Host threads the calling thread's builds use (0 = all available); a
 * worker pool that compiles several meshes at once sets 1 in each worker.
 * Returns 0.  No counterpart in the reference (Far is single-threaded). */
int osdfar_set_threads(int n);

/* TopologyDescriptor + Sdc::Options::VtxBoundaryInterpolation
 * (0 none, 1 edge only, 2 edge and corner); creasing is uniform unless
 * osdfar_topology_set_creasing selects Chaikin. Returns
 * NULL for input the compiler does not handle (non-manifold edges). */
osdfar_topology *osdfar_topology_create(int numVertices, int numFaces,
                                        const int *numVertsPerFace,
                                        const int *vertIndicesPerFace,
                                        int numCreases, const int *creaseVertexIndexPairs,
                                        const float *creaseWeights,
                                        int numCorners, const int *cornerVertexIndices,
                                        const float *cornerWeights,
                                        int boundaryInterpolation);
void osdfar_topology_destroy(osdfar_topology *topo);

/* Sdc::Options::SetCreasingMethod (sdc/options.h:66-72): 0 uniform
 * (default), 1 Chaikin.  Before the first refinement only: -2 afterwards,
 * -1 for an unknown method. */
int osdfar_topology_set_creasing(osdfar_topology *topo, int method);

/* Sdc::SchemeType (sdc/types.h:37-41; Far::TopologyRefinerFactory options):
 * 0 bilinear, 1 Catmark (default), 2 Loop (triangles only).  Uniform
 * stencil tables are built for all three (the bilinear scheme's vertex
 * tables are its linear, varying-mode tables); patch tables (and their
 * limit stencils) for Catmark and bilinear -- QUADS patches, n-gons refined
 * once -- and face-varying channels for Catmark; vertex locations for the
 * quad-split schemes.  Before the first refinement only: -2 afterwards, -1
 * for an unknown scheme or a Loop mesh with non-triangular faces. */
int osdfar_topology_set_scheme(osdfar_topology *topo, int scheme);

/* Refines uniformly to `level` and keeps every level's factorized vertex
 * stencils with the topology, so a later osdfar_stencil_table_create
 * (vertex mode) and osdfar_patch_table_create (isoLevel == level) reuse
 * them instead of recomputing.  osdfar_topology_drop_stencils frees them.
 * No counterpart in the reference. */
int osdfar_topology_cache_stencils(osdfar_topology *topo, int level);
int osdfar_topology_drop_stencils(osdfar_topology *topo);

/* TopologyRefiner::RefineUniform(level). */
int osdfar_refine_uniform(osdfar_topology *topo, int maxLevel);
/* Counts of a refinement level (level 0 = control mesh). */
int osdfar_level_info(const osdfar_topology *topo, int level, int *numVertices,
                      int *numFaces, int *numEdges);
/* Face-vertex indices of a level (4 per face above level 0). */
int osdfar_level_face_vertices(const osdfar_topology *topo, int level, int *out);

/* StencilTableFactory::Create: mode 0 vertex / 1 varying; intermediate
 * levels on/off; control-vertex identity stencils on/off; double weights
 * when useDouble != 0. */
osdfar_stencil_table *osdfar_stencil_table_create(const osdfar_topology *topo, int mode,
                                                  int intermediateLevels, int controlVerts,
                                                  int useDouble);
/* Pointers to a float table's own sizes / indices / weights (valid until
 * the table is destroyed), for uploads without a copy.  -1 for double
 * tables. */
int osdfar_stencil_table_data(const osdfar_stencil_table *t, const int **sizes,
                              const int **indices, const float **weights);

/* numStencils, nnz (64-bit), control vertex count, weight arrays (1, 3, 6). */
int osdfar_stencil_table_info(const osdfar_stencil_table *t, int *numStencils,
                              long long *nnz, int *numControlVertices, int *numWeightArrays);
/* Copies the arrays (each pointer may be NULL).  offsets32 requires
 * nnz < 2^31.  which = 0 P, 1 du, 2 dv, 3 duu, 4 duv, 5 dvv. */
int osdfar_stencil_table_copy(const osdfar_stencil_table *t, int *sizes, int *offsets32,
                              long long *offsets64, int *indices, int which,
                              float *weights, double *weights64);
void osdfar_stencil_table_destroy(osdfar_stencil_table *t);

/* ---- feature-adaptive patch tables --------------------------------------
 * PatchTableFactory::Create(refiner, Options(isoLevel)) with
 * ENDCAP_GREGORY_BASIS (far/patchTableFactory.h:40-160) for manifold
 * Catmark meshes with semi-sharp creases, quad or n-gon base faces (n-gons
 * need isoLevel >= 1; their patches carry PatchParam's non-quad bit),
 * uniform or Chaikin creasing, and boundaries (boundary patches carry Far's
 * boundary mask and fill their missing points with the face's first vertex;
 * a smooth single-face corner is isolated, then patched as a sharp corner,
 * Far's default; irregular boundary points get end caps with the boundary
 * limit masks), and interior infinitely
 * sharp creases and corners (isolated to isoLevel, as Far does with
 * inf-sharp patches off).  Returns NULL for non-manifold input.
 * PatchParams carry Far's transition masks.  Refines the topology uniformly to isoLevel.
 * The patch points live in a vertex buffer laid out as
 *   [control vertices | refined vertices | end-cap points]
 * whose non-control rows are given by osdfar_patch_table_point_stencils. */
typedef struct osdfar_patch_table osdfar_patch_table;
osdfar_patch_table *osdfar_patch_table_create(osdfar_topology *topo, int isoLevel);
/* Options of Far::PatchTableFactory / TopologyRefiner::RefineAdaptive the
 * compiler follows (osdfar_patch_table_create uses the defaults):
 *   secondaryLevel      smooth extraordinary vertices are isolated to this
 *                       level only, sharp features to isoLevel; irregular
 *                       faces left at a shallower level get Gregory end caps
 *                       there (AdaptiveOptions::secondaryLevel,
 *                       far/topologyRefiner.h:156-182; default 15 = isoLevel)
 *   legacySharpCorners  1 (Far's default): smooth single-face boundary
 *                       corners are patched as sharp corners; 0: end caps
 *   shareEndCapPoints   1 (Far's default): end-cap points shared between
 *                       neighbouring caps; 0: 20 points per cap
 * With secondaryLevel 2, isoLevel 6, legacySharpCorners 0 and
 * shareEndCapPoints 0 the table is the one Bfr's surfaces are compared with
 * (regression/bfr_evaluate/farPatchEvaluator.cpp:52-84): evaluated through
 * the patch map it gives Bfr::Surface::Evaluate's limit surface (for meshes
 * without interior infinitely sharp features). */
typedef struct {
    int secondaryLevel;
    int legacySharpCorners;
    int shareEndCapPoints;
} osdfar_patch_options;
osdfar_patch_table *osdfar_patch_table_create_ex(osdfar_topology *topo, int isoLevel,
                                                 const osdfar_patch_options *options);
void osdfar_patch_table_destroy(osdfar_patch_table *t);
int osdfar_patch_table_info(const osdfar_patch_table *t, int *numControl, int *numRefined,
                            int *numLocal, int *numArrays, int *numIndices, int *numPatches,
                            int *maxDepth);
/* Osd::CpuPatchTable layout (osd/cpuPatchTable.cpp:35-158): PatchArray
 * (6 ints each), patch vertex indices, PatchParam (field0, field1, sharpness). */
int osdfar_patch_table_copy(const osdfar_patch_table *t, int *arrays, int *indices, void *params);
/* Stencils of the refined and end-cap points (Far's appended local-point
 * stencil table, far/stencilTableFactory.cpp:256-408). */
osdfar_stencil_table *osdfar_patch_table_point_stencils(const osdfar_patch_table *t, int useDouble);
/* A face-varying channel on the topology (Far::TopologyDescriptor::
 * FVarChannel, far/topologyDescriptor.h:73-84): numValues values and one
 * value index per base face-vertex (numValues 0 removes it).  Later patch
 * tables carry the channel with Far's default options: linear refinement
 * everywhere (Sdc FVAR_LINEAR_ALL) and a linear QUADS face-varying patch per
 * patch over the four values of its face (generateFVarLegacyLinearPatches,
 * far/patchTableFactory.cpp:1255-1266), PatchParams equal to the vertex
 * patches'.  Returns -1 for bad arguments. */
int osdfar_topology_set_fvar(osdfar_topology *topo, int numValues, const int *valueIndices);
/* The channel's buffer is [base values | refined values]; info gives their
 * counts and the flattened sizes (Osd::CpuPatchTable's face-varying buffers,
 * osd/cpuPatchTable.cpp:54-115), copy the arrays (6 ints each), indices and
 * params, and the stencil table the refined values from the base values
 * (Far's INTERPOLATE_FACE_VARYING table, far/stencilTableFactory.h:55-80). */
int osdfar_patch_table_fvar_info(const osdfar_patch_table *t, int channel, int *numBase,
                                 int *numRefined, int *numArrays, int *numIndices);
int osdfar_patch_table_fvar_copy(const osdfar_patch_table *t, int channel, int *arrays,
                                 int *indices, void *params);
osdfar_stencil_table *osdfar_patch_table_fvar_stencils(const osdfar_patch_table *t, int channel,
                                                       int useDouble);
/* PatchMap::FindPatch (far/patchMap.h:193-231) for n (ptex face, u, v)
 * queries; writes Osd::PatchCoord records {arrayIndex, patchIndex,
 * vertIndex, s, t} (arrayIndex = -1 when no patch covers the location).
 * Returns the number of locations found. */
int osdfar_patch_map_find(const osdfar_patch_table *t, int n, const int *ptex, const double *u,
                          const double *v, void *patchCoords);
/* LimitStencilTableFactory::Create (far/stencilTableFactory.cpp:411-662):
 * one stencil per location with P, du, dv (and duu, duv, dvv when
 * secondDerivatives) weights, factorized to control vertices. */
osdfar_stencil_table *osdfar_limit_stencil_table_create(const osdfar_patch_table *t, int n,
                                                        const int *ptex, const double *u,
                                                        const double *v, int secondDerivatives,
                                                        int useDouble);
/* One (ptex face, u, v) location per vertex of uniform level `level` of a
 * refined topology (the vertex's corner of an incident refined face). */
int osdfar_vertex_locations(const osdfar_topology *topo, int level, int *ptex, double *u,
                            double *v);

/* ---- table files and topology keys --------------------------------------
 * No counterpart in the reference: Far rebuilds every table from the
 * topology on each run.  A table file holds named, typed arrays (sections)
 * and the 64-bit key of what they were built from, so a process that keeps
 * many topologies resident loads their tables instead of recompiling them.
 *
 * osdfar_topology_key hashes a TopologyDescriptor's arrays (the arguments
 * of osdfar_topology_create), the boundary rule and a caller's recipe
 * string naming what was built (levels, options, sampled locations); 0 for
 * invalid input.  osdfar_hash64 is XXH64 (the published xxHash 64-bit
 * algorithm) of nbytes bytes.
 *
 * osdfar_table_file_write writes the sections (names unique, 1-31 bytes)
 * under a temporary name and renames it onto `path`, so readers see no
 * file or a whole one; 0, -1 for invalid arguments, -2 for I/O errors.
 * osdfar_table_file_open maps the file read-only and returns NULL when it
 * is missing, truncated, of another format version or key, or has a
 * corrupt directory -- or, when verify != 0, a section whose payload hash
 * does not match.  Section data pointers stay valid until
 * osdfar_table_file_close; each section is 256-byte aligned in the map. */
enum { OSDFAR_I32 = 1, OSDFAR_I64 = 2, OSDFAR_F32 = 3, OSDFAR_F64 = 4, OSDFAR_BYTES = 5 };
typedef struct osdfar_section {
    const char *name;
    int dtype;              /* OSDFAR_* */
    long long count;        /* elements */
    const void *data;
} osdfar_section;
typedef struct osdfar_table_file osdfar_table_file;

unsigned long long osdfar_hash64(const void *data, long long nbytes, unsigned long long seed);
unsigned long long osdfar_topology_key(int numVertices, int numFaces, const int *numVertsPerFace,
                                       const int *vertIndicesPerFace, int numCreases,
                                       const int *creaseVertexIndexPairs,
                                       const float *creaseWeights, int numCorners,
                                       const int *cornerVertexIndices, const float *cornerWeights,
                                       int boundaryInterpolation, const char *recipe);
int osdfar_table_file_write(const char *path, unsigned long long key, int numSections,
                            const osdfar_section *sections);
osdfar_table_file *osdfar_table_file_open(const char *path, unsigned long long key, int verify);
/* Number of sections (-1 for NULL); section i's name, dtype, count, data. */
int osdfar_table_file_sections(const osdfar_table_file *f);
int osdfar_table_file_section(const osdfar_table_file *f, int i, const char **name, int *dtype,
                              long long *count, const void **data);
/* The section called `name`: 0, or -2 when there is none. */
int osdfar_table_file_find(const osdfar_table_file *f, const char *name, int *dtype,
                           long long *count, const void **data);
void osdfar_table_file_close(osdfar_table_file *f);

#ifdef __cplusplus
}
#endif
#endif
