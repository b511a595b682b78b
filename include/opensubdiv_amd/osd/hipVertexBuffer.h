// Osd::HipVertexBuffer -- device primvar buffer for the HIP evaluator.
// Replaces Osd::CudaVertexBuffer (osd/cudaVertexBuffer.h:37-68,
// osd/cudaVertexBuffer.cpp:34-94): Create / UpdateData / BindHipBuffer.
// Calls only the C ABI of libosdhip (include/osdhip.h).
#pragma once

#include <cstddef>

#include "../../osdhip.h"
#include "types.h"

namespace OpenSubdiv {
namespace OPENSUBDIV_VERSION {
namespace Osd {

class HipVertexBuffer {
public:
    /// numElements floats per vertex; NULL when the device allocation fails.
    static HipVertexBuffer *Create(int numElements, int numVertices, void *deviceContext = NULL)
    {
        (void)deviceContext;
        HipVertexBuffer *b = new HipVertexBuffer(numElements, numVertices);
        if (numElements > 0 && numVertices > 0 && !b->_data) {
            delete b;
            return NULL;
        }
        return b;
    }

    ~HipVertexBuffer() { osdhip_free(_data); }

    /// Copies numVertices vertices from host memory to vertex startVertex.
    /// deviceContext may carry a hipStream_t (NULL: null stream); the copy
    /// is ordered on that stream.
    void UpdateData(const float *src, int startVertex, int numVertices, void *deviceContext = NULL)
    {
        osdhip_memcpy_htod(_data + (size_t)startVertex * _numElements, src,
                           (size_t)numVertices * _numElements * sizeof(float), deviceContext);
    }

    int GetNumElements() const { return _numElements; }
    int GetNumVertices() const { return _numVertices; }
    float *BindHipBuffer() { return _data; }
    /// Binding used by Osd::Mesh (osd/mesh.h:562-568); same device pointer.
    float *BindVBO(void *deviceContext = NULL) { (void)deviceContext; return _data; }

private:
    HipVertexBuffer(int numElements, int numVertices)
        : _numElements(numElements), _numVertices(numVertices), _data(NULL)
    {
        void *p = NULL;
        if (osdhip_malloc(&p, (size_t)numElements * numVertices * sizeof(float)) == OSDHIP_OK)
            _data = (float *)p;
    }
    HipVertexBuffer(const HipVertexBuffer &);
    HipVertexBuffer &operator=(const HipVertexBuffer &);

    int _numElements, _numVertices;
    float *_data;
};

}  // namespace Osd
}  // namespace OPENSUBDIV_VERSION
using namespace OPENSUBDIV_VERSION;
}  // namespace OpenSubdiv
