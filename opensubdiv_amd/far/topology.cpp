// Topology levels and uniform refinement (see topology.h).
#include "topology.h"

#include <algorithm>
#include <unordered_map>

namespace osdfar {

Rule Level::vertexRule(int v) const
{
    // Sdc::Crease::DetermineVertexVertexRule (sdc/crease.cpp:44-60)
    if (vsharp[v] > 0.0f) return RULE_CORNER;
    int sharp = 0;
    for (int k = veOff[v]; k < veOff[v + 1]; ++k) sharp += esharp[ve[k]] > 0.0f;
    if (sharp > 2) return RULE_CORNER;
    return (Rule)(1 << sharp);
}

static void invert(int nOwners, const std::vector<int> &ownerOf, const std::vector<int> &members,
                   int nTargets, std::vector<int> &off, std::vector<int> &out)
{
    // members[k] is a target; ownerOf[k] the owner listing it.  Produces,
    // per target, the owners in increasing k order (stable).
    (void)nOwners;
    off.assign(nTargets + 1, 0);
    for (int t : members) off[t + 1]++;
    for (int i = 0; i < nTargets; ++i) off[i + 1] += off[i];
    out.resize(members.size());
    std::vector<int> cur(off.begin(), off.end() - 1);
    for (size_t k = 0; k < members.size(); ++k) out[cur[members[k]]++] = ownerOf[k];
}

void Level::buildIncidence()
{
    // edge -> faces, vertex -> faces: in face order
    std::vector<int> owner(fv.size());
    for (int f = 0; f < nf; ++f)
        for (int k = fvOff[f]; k < fvOff[f + 1]; ++k) owner[k] = f;
    invert(nf, owner, fe, ne, efOff, ef);
    invert(nf, owner, fv, nv, vfOff, vf);
    // vertex -> edges: in edge order
    std::vector<int> eowner(2 * (size_t)ne);
    for (int e = 0; e < ne; ++e) eowner[2 * e] = eowner[2 * e + 1] = e;
    invert(ne, eowner, ev, nv, veOff, ve);
    // an edge listed twice by a vertex (degenerate) is not expected here
}

bool Refiner::create(const TopologyDesc &d)
{
    levels.assign(1, Level());
    origins.assign(1, {});
    Level &L = levels[0];
    L.nv = d.numVertices;
    L.nf = d.numFaces;
    L.fvOff.assign(L.nf + 1, 0);
    for (int f = 0; f < L.nf; ++f) {
        if (d.numVertsPerFace[f] < 3) return false;
        L.fvOff[f + 1] = L.fvOff[f] + d.numVertsPerFace[f];
    }
    L.fv.assign(d.vertIndicesPerFace, d.vertIndicesPerFace + L.fvOff[L.nf]);
    for (int v : L.fv)
        if (v < 0 || v >= L.nv) return false;

    // Level-0 edges in order of first appearance walking faces and their
    // edges (vtr/level.cpp:1640-1720); vertices in first-seen order.
    std::unordered_map<uint64_t, int> edgeOf;
    edgeOf.reserve(L.fv.size());
    L.fe.resize(L.fv.size());
    std::vector<int> edgeFaceCount;
    for (int f = 0; f < L.nf; ++f) {
        int n = L.faceSize(f), b = L.fvOff[f];
        for (int k = 0; k < n; ++k) {
            int v0 = L.fv[b + k], v1 = L.fv[b + (k + 1) % n];
            if (v0 == v1) return false;
            uint64_t key = ((uint64_t)(uint32_t)std::min(v0, v1) << 32) | (uint32_t)std::max(v0, v1);
            auto it = edgeOf.find(key);
            int e;
            if (it == edgeOf.end()) {
                e = L.ne++;
                edgeOf.emplace(key, e);
                L.ev.push_back(v0);
                L.ev.push_back(v1);
                edgeFaceCount.push_back(0);
            } else {
                e = it->second;
            }
            if (++edgeFaceCount[e] > 2) return false;   // non-manifold
            L.fe[b + k] = e;
        }
    }
    L.esharp.assign(L.ne, 0.0f);
    L.vsharp.assign(L.nv, 0.0f);
    for (int c = 0; c < d.numCreases; ++c) {
        int a = d.creaseVertexIndexPairs[2 * c], b = d.creaseVertexIndexPairs[2 * c + 1];
        uint64_t key = ((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b);
        auto it = edgeOf.find(key);
        if (it != edgeOf.end())
            L.esharp[it->second] = std::min(std::max(d.creaseWeights[c], 0.0f), kSharpInfinite);
    }
    for (int c = 0; c < d.numCorners; ++c) {
        int v = d.cornerVertexIndices[c];
        if (v >= 0 && v < L.nv) L.vsharp[v] = std::min(std::max(d.cornerWeights[c], 0.0f), kSharpInfinite);
    }
    L.buildIncidence();
    // boundary edges are infinitely sharp; with edge-and-corner, boundary
    // vertices of a single face become sharp corners
    for (int e = 0; e < L.ne; ++e)
        if (L.efOff[e + 1] - L.efOff[e] == 1) L.esharp[e] = kSharpInfinite;
    if (d.boundaryInterpolation == 2)
        for (int v = 0; v < L.nv; ++v)
            if (L.vfOff[v + 1] - L.vfOff[v] == 1) L.vsharp[v] = kSharpInfinite;

    creasing = d.creasing == CREASE_CHAIKIN ? CREASE_CHAIKIN : CREASE_UNIFORM;
    boundaryInterpolation = d.boundaryInterpolation;

    // ptex faces (far/ptexIndices.cpp): one per quad, one per corner of an n-gon
    ptexBase.resize(L.nf);
    ptexCount = 0;
    for (int f = 0; f < L.nf; ++f) {
        ptexBase[f] = ptexCount;
        ptexCount += L.faceSize(f) == 4 ? 1 : L.faceSize(f);
    }
    ptexNonQuad.assign(ptexCount, 0);
    for (int f = 0; f < L.nf; ++f)
        if (L.faceSize(f) != 4)
            for (int j = 0; j < L.faceSize(f); ++j) ptexNonQuad[ptexBase[f] + j] = 1;
    return true;
}

void Refiner::refineUniform(int maxLevel)
{
    while (this->maxLevel() < maxLevel) refineOnce();
}

int Refiner::numVerticesTotal() const
{
    int n = 0;
    for (const Level &l : levels) n += l.nv;
    return n;
}

bool Refiner::setScheme(int s)
{
    if (s != SCHEME_BILINEAR && s != SCHEME_CATMARK && s != SCHEME_LOOP) return false;
    if (maxLevel() > 0 || levels.empty()) return false;
    if (s == SCHEME_LOOP)
        for (int f = 0; f < levels[0].nf; ++f)
            if (levels[0].faceSize(f) != 3) return false;
    scheme = s;
    return true;
}

// Sharpness of a new level: the two child edges of parent edge e are child
// edges eFromE + 2e + {0, 1} (at the parent's ev[0], ev[1]); vertex children
// come first, so child vertex v is parent vertex v.
void Refiner::childSharpness(const Level &P, Level &C, int eFromE) const
{
    auto edgeChild = [&](int pe, int corner) {
        return eFromE + 2 * pe + (P.ev[2 * pe] == corner ? 0 : 1);
    };
    C.esharp.assign(C.ne, 0.0f);
    if (creasing == CREASE_CHAIKIN) {
        // the child edge at parent vertex v subdivides the parent's sharpness
        // with the edges around v (vtr/refinement.cpp:870-930)
        std::vector<float> ps, cs;
        for (int v = 0; v < P.nv; ++v) {
            const int n = P.valence(v);
            ps.resize(n);
            cs.resize(n);
            for (int i = 0; i < n; ++i) ps[i] = P.esharp[P.ve[P.veOff[v] + i]];
            subdivideSharpnessAroundVertex(creasing, n, ps.data(), cs.data());
            for (int i = 0; i < n; ++i) C.esharp[edgeChild(P.ve[P.veOff[v] + i], v)] = cs[i];
        }
    } else {
        for (int e = 0; e < P.ne; ++e)
            C.esharp[eFromE + 2 * e] = C.esharp[eFromE + 2 * e + 1] = decrementSharpness(P.esharp[e]);
    }
    C.vsharp.assign(C.nv, 0.0f);
    for (int v = 0; v < P.nv; ++v) C.vsharp[v] = decrementSharpness(P.vsharp[v]);
}

// Loop's triangle split (vtr/triRefinement.cpp:141-345).  Child vertices:
// [children of vertices | of edges]; child faces: four per triangle, the
// corner triangles at face vertices 0, 1, 2 and the middle one; child
// edges: [three per triangle (the middle triangle's sides) | two per edge].
// Face-child edge i joins the edge points of face edges i and i - 1; the
// middle triangle starts at the edge point of face edge 1.
void Refiner::refineTriangles()
{
    const int pl = maxLevel();
    levels.emplace_back();
    origins.emplace_back();
    const Level &P = levels[pl];
    Level &C = levels.back();
    const int vFromV = 0, vFromE = P.nv;
    C.nv = P.nv + P.ne;
    C.nf = 4 * P.nf;
    const int eFromF = 0, eFromE = 3 * P.nf;
    C.ne = 3 * P.nf + 2 * P.ne;
    C.fvOff.resize(C.nf + 1);
    for (int i = 0; i <= C.nf; ++i) C.fvOff[i] = 3 * i;
    C.fv.resize(3 * (size_t)C.nf);
    C.fe.resize(3 * (size_t)C.nf);
    C.ev.resize(2 * (size_t)C.ne);
    auto edgeChild = [&](int pe, int corner) {
        return eFromE + 2 * pe + (P.ev[2 * pe] == corner ? 0 : 1);
    };
    std::vector<FaceOrigin> &org = origins.back();
    org.assign(C.nf, FaceOrigin{ 0, 0, 0 });
    for (int f = 0; f < P.nf; ++f) {
        const int b = P.fvOff[f];
        int fv[3], ecv[3], atStart[3], atEnd[3], fce[3];
        for (int i = 0; i < 3; ++i) {
            fv[i] = P.fv[b + i];
            const int pe = P.fe[b + i];                  // joins fv[i] and fv[i + 1]
            ecv[i] = vFromE + pe;
            atStart[i] = edgeChild(pe, P.fv[b + i]);
            atEnd[i] = edgeChild(pe, P.fv[b + (i + 1) % 3]);
            fce[i] = eFromF + 3 * f + i;
        }
        const int cfv[4][3] = { { vFromV + fv[0], ecv[0], ecv[2] },
                                { ecv[0], vFromV + fv[1], ecv[1] },
                                { ecv[2], ecv[1], vFromV + fv[2] },
                                { ecv[1], ecv[2], ecv[0] } };
        const int cfe[4][3] = { { atStart[0], fce[0], atEnd[2] },
                                { atEnd[0], atStart[1], fce[1] },
                                { fce[2], atEnd[1], atStart[2] },
                                { fce[2], fce[0], fce[1] } };
        for (int k = 0; k < 4; ++k)
            for (int j = 0; j < 3; ++j) {
                C.fv[3 * (4 * f + k) + j] = cfv[k][j];
                C.fe[3 * (4 * f + k) + j] = cfe[k][j];
            }
        const int cev[3][2] = { { ecv[0], ecv[2] }, { ecv[1], ecv[0] }, { ecv[2], ecv[1] } };
        for (int i = 0; i < 3; ++i) {
            C.ev[2 * fce[i]] = cev[i][0];
            C.ev[2 * fce[i] + 1] = cev[i][1];
        }
    }
    for (int e = 0; e < P.ne; ++e)
        for (int j = 0; j < 2; ++j) {
            C.ev[2 * (eFromE + 2 * e + j)] = vFromE + e;
            C.ev[2 * (eFromE + 2 * e + j) + 1] = vFromV + P.ev[2 * e + j];
        }
    childSharpness(P, C, eFromE);
    C.buildIncidence();
}

void Refiner::refineOnce()
{
    if (scheme == SCHEME_LOOP) {
        refineTriangles();
        return;
    }
    const int pl = maxLevel();
    levels.emplace_back();
    origins.emplace_back();
    const Level &P = levels[pl];
    Level &C = levels.back();
    const int nFaceVerts = (int)P.fv.size();

    // child vertices: [children of vertices | of faces | of edges]
    const int vFromV = 0, vFromF = P.nv, vFromE = P.nv + P.nf;
    C.nv = P.nv + P.nf + P.ne;
    // child faces: face f owns children fvOff[f] .. fvOff[f+1]-1
    C.nf = nFaceVerts;
    // child edges: [from faces: one per face-vertex | from edges: two per edge]
    const int eFromF = 0, eFromE = nFaceVerts;
    C.ne = nFaceVerts + 2 * P.ne;

    C.fvOff.resize(C.nf + 1);
    for (int i = 0; i <= C.nf; ++i) C.fvOff[i] = 4 * i;
    C.fv.resize(4 * (size_t)C.nf);
    C.fe.resize(4 * (size_t)C.nf);
    C.ev.resize(2 * (size_t)C.ne);

    auto edgeChild = [&](int pe, int corner) {
        // child edge of parent edge pe touching parent vertex `corner`
        return eFromE + 2 * pe + (P.ev[2 * pe] == corner ? 0 : 1);
    };

    std::vector<FaceOrigin> &org = origins.back();
    org.resize(C.nf);
    for (int f = 0; f < P.nf; ++f) {
        const int n = P.faceSize(f), b = P.fvOff[f];
        for (int j = 0; j < n; ++j) {
            const int cf = b + j;
            const int jPrev = j ? j - 1 : n - 1;
            const int vert = P.fv[b + j];
            const int eNext = P.fe[b + j], ePrev = P.fe[b + jPrev];
            int cv[4] = { vFromV + vert, vFromE + eNext, vFromF + f, vFromE + ePrev };
            int ce[4] = { edgeChild(eNext, vert), eFromF + b + j, eFromF + b + jPrev,
                          edgeChild(ePrev, vert) };
            // quads keep the parent's orientation: child j starts at index j
            // (vtr/quadRefinement.cpp:150-168); n-gon children start at 0
            const int rot = (n == 4) ? j : 0;
            for (int k = 0; k < 4; ++k) {
                C.fv[4 * cf + (rot + k) % 4] = cv[k];
                C.fe[4 * cf + (rot + k) % 4] = ce[k];
            }
            FaceOrigin o;
            if (pl == 0) {
                if (n == 4) {
                    o.ptex = ptexBase[f];
                    o.u = (uint16_t)(j == 1 || j == 2);
                    o.v = (uint16_t)(j >= 2);
                } else {
                    o.ptex = ptexBase[f] + j;
                    o.u = o.v = 0;
                }
            } else {
                const FaceOrigin &po = origins[pl][f];
                o.ptex = po.ptex;
                o.u = (uint16_t)(2 * po.u + (j == 1 || j == 2));
                o.v = (uint16_t)(2 * po.v + (j >= 2));
            }
            org[cf] = o;
        }
        for (int j = 0; j < n; ++j) {
            // child edge from face: (face point, edge point of edge j)
            C.ev[2 * (eFromF + b + j)] = vFromF + f;
            C.ev[2 * (eFromF + b + j) + 1] = vFromE + P.fe[b + j];
        }
    }
    for (int e = 0; e < P.ne; ++e)
        for (int j = 0; j < 2; ++j) {
            C.ev[2 * (eFromE + 2 * e + j)] = vFromE + e;
            C.ev[2 * (eFromE + 2 * e + j) + 1] = vFromV + P.ev[2 * e + j];
        }
    childSharpness(P, C, eFromE);
    C.buildIncidence();
}

}  // namespace osdfar
