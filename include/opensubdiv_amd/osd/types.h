// Osd value types used by the HIP evaluator: BufferDescriptor, PatchCoord,
// PatchArray, PatchParam.
//
// When OpenSubdiv's own osd/bufferDescriptor.h / osd/types.h were included
// first, their definitions are used (these headers only add the HIP
// classes).  Otherwise layout- and API-compatible definitions are provided
// here (osd/bufferDescriptor.h:61-103, osd/types.h:42-130), so the HIP
// evaluator also builds without an OpenSubdiv installation.
#pragma once

#include <algorithm>

#ifndef OPENSUBDIV_VERSION
#define OPENSUBDIV_VERSION v3_6_0
#endif

namespace OpenSubdiv {
namespace OPENSUBDIV_VERSION {
namespace Osd {

#ifndef OPENSUBDIV3_OSD_BUFFER_DESCRIPTOR_H
#define OPENSUBDIV3_OSD_BUFFER_DESCRIPTOR_H
/// Element addressing in an interleaved float buffer: element i starts at
/// offset + i * stride and has `length` floats.
struct BufferDescriptor {
    BufferDescriptor() : offset(0), length(0), stride(0) {}
    BufferDescriptor(int o, int l, int s) : offset(o), length(l), stride(s) {}
    int GetLocalOffset() const { return stride > 0 ? offset % stride : 0; }
    bool IsValid() const { return length > 0 && length <= stride - GetLocalOffset(); }
    void Reset() { offset = length = stride = 0; }
    bool operator==(BufferDescriptor const &o) const {
        return offset == o.offset && length == o.length && stride == o.stride;
    }
    bool operator!=(BufferDescriptor const &o) const { return !(*this == o); }
    int offset;
    int length;
    int stride;
};
#endif

#ifndef OPENSUBDIV3_OSD_TYPES_H
#define OPENSUBDIV3_OSD_TYPES_H
/// Location on a patch: a patch handle and (s, t) in ptex-face space.
struct PatchCoord {
    struct Handle {
        int arrayIndex, patchIndex, vertIndex;
    };
    PatchCoord() : s(0), t(0) { handle.arrayIndex = handle.patchIndex = handle.vertIndex = 0; }
    PatchCoord(Handle h, float s_, float t_) : handle(h), s(s_), t(t_) {}
    Handle handle;
    float s, t;
};

/// A run of patches of one regular / irregular descriptor pair.  Patch type
/// ids: QUADS 3, TRIANGLES 4, LOOP 5, REGULAR 6, GREGORY_BASIS 9,
/// GREGORY_TRIANGLE 10; control-vertex counts 4, 3, 12, 16, 20, 18.
struct PatchArray {
    static int numControlVertices(int type) {
        switch (type) {
        case 3: return 4;
        case 4: return 3;
        case 5: return 12;
        case 6: return 16;
        case 9: return 20;
        case 10: return 18;
        default: return 0;
        }
    }
    PatchArray(int desc_in, int numPatches_in, int indexBase_in, int primitiveIdBase_in)
        : regDesc(desc_in), desc(desc_in), numPatches(numPatches_in), indexBase(indexBase_in),
          stride(numControlVertices(desc_in)), primitiveIdBase(primitiveIdBase_in) {}
    PatchArray(int regDesc_in, int irregDesc_in, int numPatches_in, int indexBase_in,
               int primitiveIdBase_in)
        : regDesc(regDesc_in), desc(irregDesc_in), numPatches(numPatches_in),
          indexBase(indexBase_in),
          stride(std::max(numControlVertices(regDesc_in), numControlVertices(irregDesc_in))),
          primitiveIdBase(primitiveIdBase_in) {}
    int GetPatchType() const { return desc; }
    int GetPatchTypeRegular() const { return regDesc; }
    int GetPatchTypeIrregular() const { return desc; }
    int GetNumPatches() const { return numPatches; }
    int GetIndexBase() const { return indexBase; }
    int GetStride() const { return stride; }
    int GetPrimitiveIdBase() const { return primitiveIdBase; }
    int regDesc, desc, numPatches, indexBase, stride, primitiveIdBase;
};

/// Far::PatchParam bitfields (far/patchParam.h:80-97) plus the crease
/// sharpness of single-crease patches.
struct PatchParam {
    unsigned int field0, field1;
    float sharpness;
    int GetFaceId() const { return (int)(field0 & 0xfffffff); }
    int GetU() const { return (int)((field1 >> 22) & 0x3ff); }
    int GetV() const { return (int)((field1 >> 12) & 0x3ff); }
    int GetDepth() const { return (int)(field1 & 0xf); }
    bool IsRegular() const { return ((field1 >> 5) & 1) != 0; }
    int GetBoundary() const { return (int)((field1 >> 7) & 0x1f); }
};
#endif

}  // namespace Osd
}  // namespace OPENSUBDIV_VERSION
using namespace OPENSUBDIV_VERSION;
}  // namespace OpenSubdiv
