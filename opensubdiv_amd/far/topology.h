// Host-side topology levels and uniform refinement for the table compiler
// (libosdfar).  Scope: manifold meshes -- Catmark and bilinear with n-gon
// faces at level 0 (quad split), Loop with triangles (triangle split) --
// with semi-sharp and infinitely sharp creases, corner sharpness, uniform
// or Chaikin creasing and boundary edges (sharpened).
//
// Numbering follows OpenSubdiv's Vtr uniform refinement so refined vertex i
// here is refined vertex i of Far (vtr/refinement.cpp:253-270: children of
// vertices, then of faces, then of edges; vtr/quadRefinement.cpp:125-345 and
// vtr/triRefinement.cpp:141-345 for child faces / edges; vtr/level.cpp:
// 1640-1790 for level-0 edges).
#pragma once
#include <cstdint>
#include <vector>

namespace osdfar {

constexpr float kSharpInfinite = 10.0f;   // Sdc::Crease::SHARPNESS_INFINITE

enum Rule { RULE_SMOOTH = 1, RULE_DART = 2, RULE_CREASE = 4, RULE_CORNER = 8 };
enum Creasing { CREASE_UNIFORM = 0, CREASE_CHAIKIN = 1 };   // Sdc::Options::CreasingMethod
// Sdc::SchemeType.  Bilinear refines as Catmark does (quad split, same
// numbering) with linear masks; Loop splits triangles (vtr/triRefinement.cpp).
enum Scheme { SCHEME_BILINEAR = 0, SCHEME_CATMARK = 1, SCHEME_LOOP = 2 };

struct Level {
    int nv = 0, nf = 0, ne = 0;
    std::vector<int> fvOff;     // nf + 1
    std::vector<int> fv;        // face vertices (counter-clockwise)
    std::vector<int> fe;        // face edges, fe[k] joins fv[k] and fv[k+1]
    std::vector<int> ev;        // 2 per edge
    std::vector<int> efOff, ef; // edge -> faces
    std::vector<int> veOff, ve; // vertex -> edges
    std::vector<int> vfOff, vf; // vertex -> faces
    std::vector<float> esharp, vsharp;

    int faceSize(int f) const { return fvOff[f + 1] - fvOff[f]; }
    int valence(int v) const { return veOff[v + 1] - veOff[v]; }
    int otherVertex(int e, int v) const { return ev[2 * e] == v ? ev[2 * e + 1] : ev[2 * e]; }
    Rule vertexRule(int v) const;
    void buildIncidence();      // ef / ve / vf from fv / fe / ev
};

struct TopologyDesc {
    int numVertices = 0, numFaces = 0;
    const int *numVertsPerFace = nullptr;
    const int *vertIndicesPerFace = nullptr;
    int numCreases = 0;
    const int *creaseVertexIndexPairs = nullptr;
    const float *creaseWeights = nullptr;
    int numCorners = 0;
    const int *cornerVertexIndices = nullptr;
    const float *cornerWeights = nullptr;
    int boundaryInterpolation = 1;   // 0 none, 1 edge only, 2 edge and corner
    int creasing = CREASE_UNIFORM;
};

// Where a refined face came from: the base (ptex) face, depth and integer
// (u,v) of its parametric square inside the ptex face (quads only).
struct FaceOrigin {
    int ptex;
    uint16_t u, v;
};

class Refiner {
public:
    // Returns false for input the compiler does not handle (non-manifold
    // edges, out-of-range indices).
    bool create(const TopologyDesc &d);
    void refineUniform(int maxLevel);

    int maxLevel() const { return (int)levels.size() - 1; }
    const Level &level(int l) const { return levels[l]; }
    int numVerticesTotal() const;
    int ptexFaceCount() const { return ptexCount; }

    std::vector<Level> levels;
    // per level >= 1: origin of every face (empty for level 0)
    std::vector<std::vector<FaceOrigin>> origins;
    // ptex index of each base face (one per quad; n-gons have n ptex faces)
    std::vector<int> ptexBase;
    int ptexCount = 0;
    // 1 for the ptex faces of n-gon base faces (one per corner), whose
    // patches carry PatchParam's non-quad bit
    std::vector<uint8_t> ptexNonQuad;
    int creasing = CREASE_UNIFORM;
    int boundaryInterpolation = 1;
    int scheme = SCHEME_CATMARK;

    // Before the first refinement; Loop needs triangles only.
    bool setScheme(int s);

private:
    void refineOnce();
    void refineTriangles();
    void childSharpness(const Level &P, Level &C, int eFromE) const;
};

inline float decrementSharpness(float s)
{
    if (s <= 0.0f) return 0.0f;
    if (s >= kSharpInfinite) return kSharpInfinite;
    if (s > 1.0f) return s - 1.0f;
    return 0.0f;
}

// Child sharpness of the n edges around one vertex, given the parents'
// (Sdc::Crease::SubdivideEdgeSharpnessesAroundVertex, sdc/crease.cpp:
// 138-198).  Uniform: each decremented by 1.  Chaikin: a semi-sharp edge
// takes 3/4 of its own sharpness plus 1/4 of the mean of the other
// semi-sharp edges at the vertex, minus 1; with no other semi-sharp edge it
// is decremented; smooth and infinitely sharp edges keep their state.
inline void subdivideSharpnessAroundVertex(int creasing, int n, const float *parent,
                                           float *child)
{
    int count = 0;
    float sum = 0.0f;
    if (creasing == CREASE_CHAIKIN && n >= 2)
        for (int i = 0; i < n; ++i)
            if (parent[i] > 0.0f && parent[i] < kSharpInfinite) {
                ++count;
                sum += parent[i];
            }
    for (int i = 0; i < n; ++i) {
        const float p = parent[i];
        if (count < 2 || p <= 0.0f || p >= kSharpInfinite) {
            child[i] = decrementSharpness(p);
        } else {
            const float others = (sum - p) / (float)(count - 1);
            const float c = (0.75f * p + 0.25f * others) - 1.0f;
            child[i] = c > 0.0f ? c : 0.0f;
        }
    }
}

}  // namespace osdfar
