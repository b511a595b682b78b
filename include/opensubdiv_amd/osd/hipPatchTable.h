// Osd::HipPatchTable -- device copy of a patch table for HipEvaluator's
// EvalPatches.  Replaces Osd::CudaPatchTable (osd/cudaPatchTable.h:45-100,
// osd/cudaPatchTable.cpp:58-163); the flattening into PatchArray / index /
// PatchParam buffers follows Osd::CpuPatchTable (osd/cpuPatchTable.cpp:
// 35-158): one PatchArray per Far patch array, index buffers concatenated,
// PatchParams with the crease sharpness of their patch, varying and
// face-varying arrays alongside.
#pragma once

#include <cstddef>
#include <vector>

#include "../../osdhip.h"
#include "types.h"

namespace OpenSubdiv {
namespace OPENSUBDIV_VERSION {
namespace Osd {

class HipPatchTable {
public:
    /// What Osd::Mesh::BindVertexBuffer returns for this table
    /// (osd/mesh.h:71, 562-568): the device pointer of the vertex buffer.
    typedef float *VertexBufferBinding;

    /// Flattens any Far::PatchTable-like object (GetNumPatchArrays,
    /// GetPatchArrayDescriptor, GetNumPatches, GetPatchArrayVertices,
    /// GetPatchParamTable, GetSharpnessIndexTable, GetSharpnessValues,
    /// GetVaryingPatchDescriptor, GetPatchArrayVaryingVertices and the
    /// face-varying accessors) and uploads it.  NULL on allocation failure.
    template <class FAR_PATCH_TABLE>
    static HipPatchTable *Create(FAR_PATCH_TABLE const *farPatchTable, void *deviceContext = NULL)
    {
        std::vector<PatchArray> arrays, varyingArrays;
        std::vector<int> indices, varyingIndices;
        std::vector<PatchParam> params;
        const int nArrays = farPatchTable->GetNumPatchArrays();
        const int nfvc = farPatchTable->GetNumFVarChannels();
        std::vector<std::vector<PatchArray> > fvArrays(nfvc);
        std::vector<std::vector<int> > fvIndices(nfvc);
        std::vector<std::vector<PatchParam> > fvParams(nfvc);
        for (int j = 0; j < nArrays; ++j) {
            const int np = farPatchTable->GetNumPatches(j);
            arrays.push_back(PatchArray((int)farPatchTable->GetPatchArrayDescriptor(j).GetType(), np,
                                        (int)indices.size(), (int)params.size()));
            auto cvs = farPatchTable->GetPatchArrayVertices(j);
            indices.insert(indices.end(), cvs.begin(), cvs.end());
            varyingArrays.push_back(
                PatchArray((int)farPatchTable->GetVaryingPatchDescriptor().GetType(), np,
                           (int)varyingIndices.size(), (int)params.size()));
            auto vcvs = farPatchTable->GetPatchArrayVaryingVertices(j);
            varyingIndices.insert(varyingIndices.end(), vcvs.begin(), vcvs.end());
            for (int c = 0; c < nfvc; ++c) {
                fvArrays[c].push_back(
                    PatchArray((int)farPatchTable->GetFVarPatchDescriptorRegular(c).GetType(),
                               (int)farPatchTable->GetFVarPatchDescriptorIrregular(c).GetType(),
                               np, (int)fvIndices[c].size(), (int)fvParams[c].size()));
                auto fv = farPatchTable->GetPatchArrayFVarValues(j, c);
                fvIndices[c].insert(fvIndices[c].end(), fv.begin(), fv.end());
                auto fp = farPatchTable->GetPatchArrayFVarPatchParams(j, c);
                for (int k = 0; k < np; ++k) {
                    PatchParam p;
                    p.field0 = fp[k].field0;
                    p.field1 = fp[k].field1;
                    p.sharpness = 0.0f;
                    fvParams[c].push_back(p);
                }
            }
            const auto &ptab = farPatchTable->GetPatchParamTable();
            const auto &sharpIdx = farPatchTable->GetSharpnessIndexTable();
            for (int k = 0; k < np; ++k) {
                const int pi = (int)params.size();
                float sharp = 0.0f;
                if (pi < (int)sharpIdx.size() && sharpIdx[pi] >= 0)
                    sharp = farPatchTable->GetSharpnessValues()[sharpIdx[pi]];
                PatchParam p;
                p.field0 = ptab[pi].field0;
                p.field1 = ptab[pi].field1;
                p.sharpness = sharp;
                params.push_back(p);
            }
        }
        HipPatchTable *t = new HipPatchTable();
        bool ok = t->upload(arrays, indices, params, deviceContext) &&
                  upload(varyingArrays, &t->_varyingArrays, deviceContext) &&
                  upload(varyingIndices, &t->_varyingIndices, deviceContext);
        for (int c = 0; ok && c < nfvc; ++c) {
            void *a = NULL, *i = NULL, *p = NULL;
            ok = upload(fvArrays[c], &a, deviceContext) && upload(fvIndices[c], &i, deviceContext) &&
                 upload(fvParams[c], &p, deviceContext);
            t->_fvarArrays.push_back(a);
            t->_fvarIndices.push_back(i);
            t->_fvarParams.push_back(p);
        }
        osdhip_stream_synchronize(deviceContext);
        if (!ok) { delete t; return NULL; }
        return t;
    }

    /// From already flattened host buffers (e.g. the project's own table
    /// compiler, libosdfar): arrays / indices / params as above.
    static HipPatchTable *Create(const PatchArray *arrays, int numArrays, const int *indices,
                                 int numIndices, const PatchParam *params, int numParams,
                                 void *deviceContext = NULL)
    {
        HipPatchTable *t = new HipPatchTable();
        std::vector<PatchArray> a(arrays, arrays + numArrays);
        std::vector<int> i(indices, indices + numIndices);
        std::vector<PatchParam> p(params, params + numParams);
        bool ok = t->upload(a, i, p, deviceContext);
        osdhip_stream_synchronize(deviceContext);
        if (!ok) { delete t; return NULL; }
        return t;
    }

    ~HipPatchTable()
    {
        osdhip_free(_arrays);
        osdhip_free(_indices);
        osdhip_free(_params);
        osdhip_free(_varyingArrays);
        osdhip_free(_varyingIndices);
        for (size_t c = 0; c < _fvarArrays.size(); ++c) {
            osdhip_free(_fvarArrays[c]);
            osdhip_free(_fvarIndices[c]);
            osdhip_free(_fvarParams[c]);
        }
    }

    void *GetPatchArrayBuffer() const { return _arrays; }
    void *GetPatchIndexBuffer() const { return _indices; }
    void *GetPatchParamBuffer() const { return _params; }
    void *GetVaryingPatchArrayBuffer() const { return _varyingArrays; }
    void *GetVaryingPatchIndexBuffer() const { return _varyingIndices; }
    int GetNumFVarChannels() const { return (int)_fvarArrays.size(); }
    void *GetFVarPatchArrayBuffer(int fvarChannel) const { return _fvarArrays[fvarChannel]; }
    void *GetFVarPatchIndexBuffer(int fvarChannel = 0) const { return _fvarIndices[fvarChannel]; }
    void *GetFVarPatchParamBuffer(int fvarChannel = 0) const { return _fvarParams[fvarChannel]; }

private:
    HipPatchTable()
        : _arrays(NULL), _indices(NULL), _params(NULL), _varyingArrays(NULL),
          _varyingIndices(NULL) {}
    HipPatchTable(const HipPatchTable &);
    HipPatchTable &operator=(const HipPatchTable &);

    template <class T>
    static bool upload(const std::vector<T> &v, void **dst, void *stream)
    {
        *dst = NULL;
        if (v.empty()) return true;
        if (osdhip_malloc(dst, v.size() * sizeof(T)) != OSDHIP_OK) return false;
        return osdhip_memcpy_htod(*dst, v.data(), v.size() * sizeof(T), stream) == OSDHIP_OK;
    }
    bool upload(const std::vector<PatchArray> &a, const std::vector<int> &i,
                const std::vector<PatchParam> &p, void *stream)
    {
        return upload(a, &_arrays, stream) && upload(i, &_indices, stream) &&
               upload(p, &_params, stream);
    }

    void *_arrays, *_indices, *_params, *_varyingArrays, *_varyingIndices;
    std::vector<void *> _fvarArrays, _fvarIndices, _fvarParams;
};

}  // namespace Osd
}  // namespace OPENSUBDIV_VERSION
using namespace OPENSUBDIV_VERSION;
}  // namespace OpenSubdiv
