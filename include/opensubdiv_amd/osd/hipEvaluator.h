// Osd::HipStencilTable and Osd::HipEvaluator -- the MI355X evaluator with the
// static interface of Osd::CudaEvaluator (osd/cudaEvaluator.h:52-1265), so
// it drops in beside CpuEvaluator and under Osd::Mesh<HipVertexBuffer,
// HipStencilTable, HipEvaluator, HipPatchTable>.  Every method forwards to
// the C ABI of libosdhip (include/osdhip.h); all buffer pointers are device
// pointers, deviceContext may carry a hipStream_t (NULL: null stream), and
// results are bool as in the reference.
#pragma once

#include <cstddef>
#include <type_traits>
#include <vector>

#include "../../osdhip.h"
#include "hipPatchTable.h"
#include "hipVertexBuffer.h"
#include "types.h"

namespace OpenSubdiv {
namespace OPENSUBDIV_VERSION {
namespace Osd {

namespace hip_detail {
template <class T, class = void> struct HasDu : std::false_type {};
template <class T>
struct HasDu<T, decltype((void)std::declval<T const &>().GetDuWeights(), void())> : std::true_type {};
template <class T, class = void> struct HasDuu : std::false_type {};
template <class T>
struct HasDuu<T, decltype((void)std::declval<T const &>().GetDuuWeights(), void())> : std::true_type {};

inline osdhip_buffer_desc D(BufferDescriptor const &d)
{
    osdhip_buffer_desc r;
    r.offset = d.offset;
    r.length = d.length;
    r.stride = d.stride;
    return r;
}
}  // namespace hip_detail

/// Device copy of a Far::StencilTable or Far::LimitStencilTable (replaces
/// CudaStencilTable, osd/cudaEvaluator.h:52-91).  Holds the Far CSR arrays
/// (Get*Buffer) and the 64-row sliced layout the evaluator streams.
class HipStencilTable {
public:
    /// Any Far (Limit)StencilTable-like object: GetNumStencils, GetSizes,
    /// GetOffsets, GetControlIndices, GetWeights [, GetDu/DvWeights
    /// [, GetDuu/Duv/DvvWeights]].  NULL on allocation failure.
    template <class FAR_STENCIL_TABLE>
    static HipStencilTable *Create(FAR_STENCIL_TABLE const *table, void *deviceContext = NULL)
    {
        return Create(table, (const osdhip_table_options *)NULL, deviceContext);
    }

    /// The same with explicit device options (layout, CSR arrays kept,
    /// kernel; osdhip_table_options in osdhip.h).  NULL options = defaults;
    /// options out of range give NULL.
    template <class FAR_STENCIL_TABLE>
    static HipStencilTable *Create(FAR_STENCIL_TABLE const *table,
                                   const osdhip_table_options *options, void *deviceContext)
    {
        std::vector<const float *> w;
        w.push_back(data(table->GetWeights()));
        collectDerivatives(table, w, hip_detail::HasDu<FAR_STENCIL_TABLE>(),
                           hip_detail::HasDuu<FAR_STENCIL_TABLE>());
        const int n = table->GetNumStencils();
        osdhip_stencil_table *t = osdhip_stencil_table_create_ex(
            n, data(table->GetSizes()), table->GetOffsets().empty() ? NULL : data(table->GetOffsets()),
            data(table->GetControlIndices()), (int)w.size(), w.data(), options, deviceContext);
        if (!t && n > 0) return NULL;
        return new HipStencilTable(t, n);
    }

    ~HipStencilTable() { osdhip_stencil_table_destroy(_table); }

    void *GetSizesBuffer() const { return (void *)_sizes; }
    void *GetOffsetsBuffer() const { return (void *)_offsets; }
    void *GetIndicesBuffer() const { return (void *)_indices; }
    void *GetWeightsBuffer() const { return (void *)_w[0]; }
    void *GetDuWeightsBuffer() const { return (void *)_w[1]; }
    void *GetDvWeightsBuffer() const { return (void *)_w[2]; }
    void *GetDuuWeightsBuffer() const { return (void *)_w[3]; }
    void *GetDuvWeightsBuffer() const { return (void *)_w[4]; }
    void *GetDvvWeightsBuffer() const { return (void *)_w[5]; }
    int GetNumStencils() const { return _numStencils; }
    osdhip_stencil_table *GetDeviceTable() const { return _table; }

private:
    HipStencilTable(osdhip_stencil_table *t, int n) : _table(t), _numStencils(n)
    {
        _sizes = _offsets = _indices = NULL;
        for (int k = 0; k < 6; ++k) _w[k] = NULL;
        if (t) osdhip_stencil_table_csr(t, &_sizes, &_offsets, &_indices, _w);
    }
    HipStencilTable(const HipStencilTable &);
    HipStencilTable &operator=(const HipStencilTable &);

    template <class V> static const typename V::value_type *data(V const &v)
    {
        return v.empty() ? NULL : &v[0];
    }
    template <class T>
    static void collectDerivatives(T const *, std::vector<const float *> &, std::false_type,
                                   std::false_type) {}
    template <class T, class D2>
    static void collectDerivatives(T const *t, std::vector<const float *> &w, std::true_type, D2)
    {
        if (t->GetDuWeights().empty() || t->GetDvWeights().empty()) return;
        w.push_back(data(t->GetDuWeights()));
        w.push_back(data(t->GetDvWeights()));
        collectSecond(t, w, D2());
    }
    template <class T>
    static void collectSecond(T const *, std::vector<const float *> &, std::false_type) {}
    template <class T>
    static void collectSecond(T const *t, std::vector<const float *> &w, std::true_type)
    {
        if (t->GetDuuWeights().empty() || t->GetDuvWeights().empty() || t->GetDvvWeights().empty())
            return;
        w.push_back(data(t->GetDuuWeights()));
        w.push_back(data(t->GetDuvWeights()));
        w.push_back(data(t->GetDvvWeights()));
    }

    osdhip_stencil_table *_table;
    int _numStencils;
    const int *_sizes, *_offsets, *_indices;
    const float *_w[6];
};

/// Static evaluator (replaces CudaEvaluator, osd/cudaEvaluator.h:94-1265).
/// There is no instance state: GetEvaluator<> returns NULL for it, as for
/// CudaEvaluator (osd/mesh.h:304-410).
class HipEvaluator {
public:
    // ---- EvalStencils, templated (osd/cudaEvaluator.h:125-143, 217-241, 352-385)

    template <typename SRC_BUFFER, typename DST_BUFFER, typename STENCIL_TABLE>
    static bool EvalStencils(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                             DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                             STENCIL_TABLE const *stencilTable,
                             const void *instance = NULL, void *deviceContext = NULL)
    {
        (void)instance;
        float *outs[1] = { dstBuffer->BindHipBuffer() };
        BufferDescriptor descs[1] = { dstDesc };
        return evalTable(srcBuffer->BindHipBuffer(), srcDesc, 1, outs, descs, stencilTable,
                         deviceContext);
    }

    template <typename SRC_BUFFER, typename DST_BUFFER, typename STENCIL_TABLE>
    static bool EvalStencils(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                             DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                             DST_BUFFER *duBuffer, BufferDescriptor const &duDesc,
                             DST_BUFFER *dvBuffer, BufferDescriptor const &dvDesc,
                             STENCIL_TABLE const *stencilTable,
                             const void *instance = NULL, void *deviceContext = NULL)
    {
        (void)instance;
        float *outs[3] = { bind(dstBuffer), bind(duBuffer), bind(dvBuffer) };
        BufferDescriptor descs[3] = { dstDesc, duDesc, dvDesc };
        return evalTable(srcBuffer->BindHipBuffer(), srcDesc, 3, outs, descs, stencilTable,
                         deviceContext);
    }

    template <typename SRC_BUFFER, typename DST_BUFFER, typename STENCIL_TABLE>
    static bool EvalStencils(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                             DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                             DST_BUFFER *duBuffer, BufferDescriptor const &duDesc,
                             DST_BUFFER *dvBuffer, BufferDescriptor const &dvDesc,
                             DST_BUFFER *duuBuffer, BufferDescriptor const &duuDesc,
                             DST_BUFFER *duvBuffer, BufferDescriptor const &duvDesc,
                             DST_BUFFER *dvvBuffer, BufferDescriptor const &dvvDesc,
                             STENCIL_TABLE const *stencilTable,
                             const void *instance = NULL, void *deviceContext = NULL)
    {
        (void)instance;
        float *outs[6] = { bind(dstBuffer), bind(duBuffer), bind(dvBuffer),
                           bind(duuBuffer), bind(duvBuffer), bind(dvvBuffer) };
        BufferDescriptor descs[6] = { dstDesc, duDesc, dvDesc, duuDesc, duvDesc, dvvDesc };
        return evalTable(srcBuffer->BindHipBuffer(), srcDesc, 6, outs, descs, stencilTable,
                         deviceContext);
    }

    // ---- EvalStencils, raw device pointers (osd/cudaEvaluator.h:159-178,
    // 272-296, 437-470).  Stencils [start, end); dst row r is stencil start+r.

    static bool EvalStencils(const float *src, BufferDescriptor const &srcDesc,
                             float *dst, BufferDescriptor const &dstDesc,
                             const int *sizes, const int *offsets, const int *indices,
                             const float *weights, int start, int end, void *stream = NULL)
    {
        return osdhip_eval_stencils(src, hip_detail::D(srcDesc), dst, hip_detail::D(dstDesc),
                                    sizes, offsets, indices, weights, start, end,
                                    stream) == OSDHIP_OK;
    }

    static bool EvalStencils(const float *src, BufferDescriptor const &srcDesc,
                             float *dst, BufferDescriptor const &dstDesc,
                             float *du, BufferDescriptor const &duDesc,
                             float *dv, BufferDescriptor const &dvDesc,
                             const int *sizes, const int *offsets, const int *indices,
                             const float *weights, const float *duWeights,
                             const float *dvWeights, int start, int end, void *stream = NULL)
    {
        using hip_detail::D;
        return osdhip_eval_stencils_d1(src, D(srcDesc), dst, D(dstDesc), du, D(duDesc), dv,
                                       D(dvDesc), sizes, offsets, indices, weights, duWeights,
                                       dvWeights, start, end, stream) == OSDHIP_OK;
    }

    static bool EvalStencils(const float *src, BufferDescriptor const &srcDesc,
                             float *dst, BufferDescriptor const &dstDesc,
                             float *du, BufferDescriptor const &duDesc,
                             float *dv, BufferDescriptor const &dvDesc,
                             float *duu, BufferDescriptor const &duuDesc,
                             float *duv, BufferDescriptor const &duvDesc,
                             float *dvv, BufferDescriptor const &dvvDesc,
                             const int *sizes, const int *offsets, const int *indices,
                             const float *weights, const float *duWeights,
                             const float *dvWeights, const float *duuWeights,
                             const float *duvWeights, const float *dvvWeights,
                             int start, int end, void *stream = NULL)
    {
        using hip_detail::D;
        return osdhip_eval_stencils_d2(src, D(srcDesc), dst, D(dstDesc), du, D(duDesc), dv,
                                       D(dvDesc), duu, D(duuDesc), duv, D(duvDesc), dvv,
                                       D(dvvDesc), sizes, offsets, indices, weights, duWeights,
                                       dvWeights, duuWeights, duvWeights, dvvWeights, start,
                                       end, stream) == OSDHIP_OK;
    }

    // ---- EvalPatches, templated (osd/cudaEvaluator.h:502-690)

    template <typename SRC_BUFFER, typename DST_BUFFER, typename PATCHCOORD_BUFFER,
              typename PATCH_TABLE>
    static bool EvalPatches(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                            DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                            int numPatchCoords, PATCHCOORD_BUFFER *patchCoords,
                            PATCH_TABLE *patchTable, HipEvaluator const *instance = NULL,
                            void *deviceContext = NULL)
    {
        (void)instance;
        return EvalPatches(srcBuffer->BindHipBuffer(), srcDesc, dstBuffer->BindHipBuffer(),
                           dstDesc, numPatchCoords,
                           (const PatchCoord *)patchCoords->BindHipBuffer(),
                           (const PatchArray *)patchTable->GetPatchArrayBuffer(),
                           (const int *)patchTable->GetPatchIndexBuffer(),
                           (const PatchParam *)patchTable->GetPatchParamBuffer(), deviceContext);
    }

    template <typename SRC_BUFFER, typename DST_BUFFER, typename PATCHCOORD_BUFFER,
              typename PATCH_TABLE>
    static bool EvalPatches(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                            DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                            DST_BUFFER *duBuffer, BufferDescriptor const &duDesc,
                            DST_BUFFER *dvBuffer, BufferDescriptor const &dvDesc,
                            int numPatchCoords, PATCHCOORD_BUFFER *patchCoords,
                            PATCH_TABLE *patchTable, HipEvaluator const *instance = NULL,
                            void *deviceContext = NULL)
    {
        (void)instance;
        return EvalPatches(srcBuffer->BindHipBuffer(), srcDesc, bind(dstBuffer), dstDesc,
                           bind(duBuffer), duDesc, bind(dvBuffer), dvDesc, numPatchCoords,
                           (const PatchCoord *)patchCoords->BindHipBuffer(),
                           (const PatchArray *)patchTable->GetPatchArrayBuffer(),
                           (const int *)patchTable->GetPatchIndexBuffer(),
                           (const PatchParam *)patchTable->GetPatchParamBuffer(), deviceContext);
    }

    template <typename SRC_BUFFER, typename DST_BUFFER, typename PATCHCOORD_BUFFER,
              typename PATCH_TABLE>
    static bool EvalPatches(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                            DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                            DST_BUFFER *duBuffer, BufferDescriptor const &duDesc,
                            DST_BUFFER *dvBuffer, BufferDescriptor const &dvDesc,
                            DST_BUFFER *duuBuffer, BufferDescriptor const &duuDesc,
                            DST_BUFFER *duvBuffer, BufferDescriptor const &duvDesc,
                            DST_BUFFER *dvvBuffer, BufferDescriptor const &dvvDesc,
                            int numPatchCoords, PATCHCOORD_BUFFER *patchCoords,
                            PATCH_TABLE *patchTable, HipEvaluator const *instance = NULL,
                            void *deviceContext = NULL)
    {
        (void)instance;
        return EvalPatches(srcBuffer->BindHipBuffer(), srcDesc, bind(dstBuffer), dstDesc,
                           bind(duBuffer), duDesc, bind(dvBuffer), dvDesc, bind(duuBuffer),
                           duuDesc, bind(duvBuffer), duvDesc, bind(dvvBuffer), dvvDesc,
                           numPatchCoords, (const PatchCoord *)patchCoords->BindHipBuffer(),
                           (const PatchArray *)patchTable->GetPatchArrayBuffer(),
                           (const int *)patchTable->GetPatchIndexBuffer(),
                           (const PatchParam *)patchTable->GetPatchParamBuffer(), deviceContext);
    }

    // ---- EvalPatches, raw device pointers (osd/cudaEvaluator.h:700-830)

    static bool EvalPatches(const float *src, BufferDescriptor const &srcDesc,
                            float *dst, BufferDescriptor const &dstDesc,
                            int numPatchCoords, const PatchCoord *patchCoords,
                            const PatchArray *patchArrays, const int *patchIndices,
                            const PatchParam *patchParams, void *stream = NULL)
    {
        return EvalPatches(src, srcDesc, dst, dstDesc, NULL, BufferDescriptor(), NULL,
                           BufferDescriptor(), numPatchCoords, patchCoords, patchArrays,
                           patchIndices, patchParams, stream);
    }

    static bool EvalPatches(const float *src, BufferDescriptor const &srcDesc,
                            float *dst, BufferDescriptor const &dstDesc,
                            float *du, BufferDescriptor const &duDesc,
                            float *dv, BufferDescriptor const &dvDesc,
                            int numPatchCoords, const PatchCoord *patchCoords,
                            const PatchArray *patchArrays, const int *patchIndices,
                            const PatchParam *patchParams, void *stream = NULL)
    {
        return EvalPatches(src, srcDesc, dst, dstDesc, du, duDesc, dv, dvDesc, NULL,
                           BufferDescriptor(), NULL, BufferDescriptor(), NULL, BufferDescriptor(),
                           numPatchCoords, patchCoords, patchArrays, patchIndices, patchParams,
                           stream);
    }

    static bool EvalPatches(const float *src, BufferDescriptor const &srcDesc,
                            float *dst, BufferDescriptor const &dstDesc,
                            float *du, BufferDescriptor const &duDesc,
                            float *dv, BufferDescriptor const &dvDesc,
                            float *duu, BufferDescriptor const &duuDesc,
                            float *duv, BufferDescriptor const &duvDesc,
                            float *dvv, BufferDescriptor const &dvvDesc,
                            int numPatchCoords, const PatchCoord *patchCoords,
                            const PatchArray *patchArrays, const int *patchIndices,
                            const PatchParam *patchParams, void *stream = NULL)
    {
        using hip_detail::D;
        // the reference dereferences dst unconditionally (osd/cpuEvaluator.cpp:
        // 131-133): a NULL dst is rejected here instead
        if (!dst) return false;
        return osdhip_eval_patches(src, D(srcDesc), dst, D(dstDesc), du, D(duDesc), dv, D(dvDesc),
                                   duu, D(duuDesc), duv, D(duvDesc), dvv, D(dvvDesc),
                                   numPatchCoords, (const osdhip_patch_coord *)patchCoords,
                                   (const osdhip_patch_array *)patchArrays, patchIndices,
                                   (const osdhip_patch_param *)patchParams,
                                   stream) == OSDHIP_OK;
    }

    // ---- EvalPatchesVarying / EvalPatchesFaceVarying (osd/cudaEvaluator.h:857-1253)

    template <typename SRC_BUFFER, typename DST_BUFFER, typename PATCHCOORD_BUFFER,
              typename PATCH_TABLE>
    static bool EvalPatchesVarying(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                                   DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                                   int numPatchCoords, PATCHCOORD_BUFFER *patchCoords,
                                   PATCH_TABLE *patchTable, HipEvaluator const *instance = NULL,
                                   void *deviceContext = NULL)
    {
        (void)instance;
        return EvalPatches(srcBuffer->BindHipBuffer(), srcDesc, dstBuffer->BindHipBuffer(),
                           dstDesc, numPatchCoords,
                           (const PatchCoord *)patchCoords->BindHipBuffer(),
                           (const PatchArray *)patchTable->GetVaryingPatchArrayBuffer(),
                           (const int *)patchTable->GetVaryingPatchIndexBuffer(),
                           (const PatchParam *)patchTable->GetPatchParamBuffer(), deviceContext);
    }

    template <typename SRC_BUFFER, typename DST_BUFFER, typename PATCHCOORD_BUFFER,
              typename PATCH_TABLE>
    static bool EvalPatchesVarying(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                                   DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                                   DST_BUFFER *duBuffer, BufferDescriptor const &duDesc,
                                   DST_BUFFER *dvBuffer, BufferDescriptor const &dvDesc,
                                   int numPatchCoords, PATCHCOORD_BUFFER *patchCoords,
                                   PATCH_TABLE *patchTable, HipEvaluator const *instance = NULL,
                                   void *deviceContext = NULL)
    {
        (void)instance;
        return EvalPatches(srcBuffer->BindHipBuffer(), srcDesc, bind(dstBuffer), dstDesc,
                           bind(duBuffer), duDesc, bind(dvBuffer), dvDesc, numPatchCoords,
                           (const PatchCoord *)patchCoords->BindHipBuffer(),
                           (const PatchArray *)patchTable->GetVaryingPatchArrayBuffer(),
                           (const int *)patchTable->GetVaryingPatchIndexBuffer(),
                           (const PatchParam *)patchTable->GetPatchParamBuffer(), deviceContext);
    }

    template <typename SRC_BUFFER, typename DST_BUFFER, typename PATCHCOORD_BUFFER,
              typename PATCH_TABLE>
    static bool EvalPatchesFaceVarying(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                                       DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                                       int numPatchCoords, PATCHCOORD_BUFFER *patchCoords,
                                       PATCH_TABLE *patchTable, int fvarChannel,
                                       HipEvaluator const *instance = NULL,
                                       void *deviceContext = NULL)
    {
        (void)instance;
        return EvalPatches(srcBuffer->BindHipBuffer(), srcDesc, dstBuffer->BindHipBuffer(),
                           dstDesc, numPatchCoords,
                           (const PatchCoord *)patchCoords->BindHipBuffer(),
                           (const PatchArray *)patchTable->GetFVarPatchArrayBuffer(fvarChannel),
                           (const int *)patchTable->GetFVarPatchIndexBuffer(fvarChannel),
                           (const PatchParam *)patchTable->GetFVarPatchParamBuffer(fvarChannel),
                           deviceContext);
    }

    template <typename SRC_BUFFER, typename DST_BUFFER, typename PATCHCOORD_BUFFER,
              typename PATCH_TABLE>
    static bool EvalPatchesFaceVarying(SRC_BUFFER *srcBuffer, BufferDescriptor const &srcDesc,
                                       DST_BUFFER *dstBuffer, BufferDescriptor const &dstDesc,
                                       DST_BUFFER *duBuffer, BufferDescriptor const &duDesc,
                                       DST_BUFFER *dvBuffer, BufferDescriptor const &dvDesc,
                                       int numPatchCoords, PATCHCOORD_BUFFER *patchCoords,
                                       PATCH_TABLE *patchTable, int fvarChannel,
                                       HipEvaluator const *instance = NULL,
                                       void *deviceContext = NULL)
    {
        (void)instance;
        return EvalPatches(srcBuffer->BindHipBuffer(), srcDesc, bind(dstBuffer), dstDesc,
                           bind(duBuffer), duDesc, bind(dvBuffer), dvDesc, numPatchCoords,
                           (const PatchCoord *)patchCoords->BindHipBuffer(),
                           (const PatchArray *)patchTable->GetFVarPatchArrayBuffer(fvarChannel),
                           (const int *)patchTable->GetFVarPatchIndexBuffer(fvarChannel),
                           (const PatchParam *)patchTable->GetFVarPatchParamBuffer(fvarChannel),
                           deviceContext);
    }

    /// Waits for the work queued on the stream in deviceContext (NULL: the
    /// null stream); replaces CudaEvaluator::Synchronize
    /// (osd/cudaEvaluator.cpp:377-380).
    static void Synchronize(void *deviceContext = NULL) { osdhip_stream_synchronize(deviceContext); }

private:
    template <class B> static float *bind(B *b) { return b ? b->BindHipBuffer() : NULL; }

    // Device-table path: one fused launch for all outputs.
    static bool evalTable(const float *src, BufferDescriptor const &srcDesc, int nout,
                          float *const *outs, BufferDescriptor const *descs,
                          HipStencilTable const *table, void *stream)
    {
        if (!table || table->GetNumStencils() == 0) return false;   // osd/cpuEvaluator.h:82-83
        osdhip_buffer_desc d[6];
        for (int k = 0; k < nout; ++k) d[k] = hip_detail::D(descs[k]);
        return osdhip_stencil_table_eval(table->GetDeviceTable(), src, hip_detail::D(srcDesc), nout,
                                         outs, d, 0, table->GetNumStencils(),
                                         stream) == OSDHIP_OK;
    }

    // Any other table type exposing device buffers: raw CSR path.
    template <class STENCIL_TABLE>
    static bool evalTable(const float *src, BufferDescriptor const &srcDesc, int nout,
                          float *const *outs, BufferDescriptor const *descs,
                          STENCIL_TABLE const *t, void *stream)
    {
        if (!t || t->GetNumStencils() == 0) return false;
        const int *s = (const int *)t->GetSizesBuffer(), *o = (const int *)t->GetOffsetsBuffer();
        const int *i = (const int *)t->GetIndicesBuffer();
        const float *w = (const float *)t->GetWeightsBuffer();
        const int n = t->GetNumStencils();
        if (nout == 1)
            return EvalStencils(src, srcDesc, outs[0], descs[0], s, o, i, w, 0, n, stream);
        const float *du = (const float *)t->GetDuWeightsBuffer();
        const float *dv = (const float *)t->GetDvWeightsBuffer();
        if (nout == 3)
            return EvalStencils(src, srcDesc, outs[0], descs[0], outs[1], descs[1], outs[2],
                                descs[2], s, o, i, w, du, dv, 0, n, stream);
        return EvalStencils(src, srcDesc, outs[0], descs[0], outs[1], descs[1], outs[2], descs[2],
                            outs[3], descs[3], outs[4], descs[4], outs[5], descs[5], s, o, i, w,
                            du, dv, (const float *)t->GetDuuWeightsBuffer(),
                            (const float *)t->GetDuvWeightsBuffer(),
                            (const float *)t->GetDvvWeightsBuffer(), 0, n, stream);
    }
};

}  // namespace Osd
}  // namespace OPENSUBDIV_VERSION
using namespace OPENSUBDIV_VERSION;
}  // namespace OpenSubdiv
