// Feature-adaptive patch tables, patch-point stencils, patch map and limit
// stencils (see patches.h for the scope and the reference functions).
#include "patches.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "basis.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace osdfar {

namespace {

constexpr int kQuads = 3, kRegular = 6, kGregory = 9;

int otherFace(const Level &L, int e, int f)
{
    for (int k = L.efOff[e]; k < L.efOff[e + 1]; ++k)
        if (L.ef[k] != f) return L.ef[k];
    return -1;
}

int indexInFace(const Level &L, int f, int v)
{
    for (int k = L.fvOff[f]; k < L.fvOff[f + 1]; ++k)
        if (L.fv[k] == v) return k - L.fvOff[f];
    return -1;
}

// Counter-clockwise one-ring of interior vertex v starting at face f:
// ring[2k] = vertex across edge k, ring[2k+1] = vertex of face k opposite v
// (quads), face k lies between edges k and k+1, face 0 = f, edge 0 = the edge
// from v to the next vertex of f.  faces[k] receives face k.  Returns the
// valence, or -1 at a boundary.
int orderedRing(const Level &L, int v, int f, std::vector<int> &ring, std::vector<int> *faces = nullptr)
{
    ring.clear();
    if (faces) faces->clear();
    int cur = f;
    for (int guard = 0; guard < 256; ++guard) {
        const int b = L.fvOff[cur], n = L.faceSize(cur);
        const int j = indexInFace(L, cur, v);
        ring.push_back(L.fv[b + (j + 1) % n]);
        ring.push_back(L.fv[b + (j + 2) % n]);
        if (faces) faces->push_back(cur);
        const int ePrev = L.fe[b + (j + n - 1) % n];     // edge (v, next edge neighbour)
        const int nxt = otherFace(L, ePrev, cur);
        if (nxt < 0) return -1;
        cur = nxt;
        if (cur == f) return (int)ring.size() / 2;
    }
    return -1;
}

// Sparse row over level vertices.
struct Row {
    std::vector<std::pair<int, double>> e;
    void add(int v, double w)
    {
        for (auto &x : e)
            if (x.first == v) { x.second += w; return; }
        e.push_back({ v, w });
    }
    void addRow(const Row &r, double s)
    {
        for (auto &x : r.e) add(x.first, x.second * s);
    }
};

// Tangent scale of the interior limit tangent masks:
// 16 / (n (cos(2pi/n) + 5 + sqrt((cos(2pi/n) + 9)(cos(2pi/n) + 1)))).
double limitTangentCoefficient(int n)
{
    const double c = std::cos(2.0 * M_PI / n);
    return 16.0 / (n * ((c + 5.0) + std::sqrt((c + 9.0) * (c + 1.0))));
}

// Limit point and the two edge tangent points of an interior corner of
// valence n whose patch face is face 0 of its ring (far/catmarkPatchBuilder.cpp
// CatmarkLimits::ComputeInteriorPointWeights with faceInRing = 0):
// weights[0] for the corner, weights[1 + 2k] edge k, weights[2 + 2k] face k.
void interiorPointWeights(int n, std::vector<double> &p, std::vector<double> &ep,
                          std::vector<double> &em)
{
    const int W = 1 + 2 * n;
    p.assign(W, 0.0);
    ep.assign(W, 0.0);
    em.assign(W, 0.0);
    std::vector<double> tw(W, 0.0);
    const double inv = 1.0 / n, invP5 = 1.0 / (n + 5.0);
    const double pc = inv * invP5;
    const double tc = limitTangentCoefficient(n) * 0.5 * invP5;
    p[0] = n * invP5;
    for (int i = 0; i < n; ++i) {
        p[1 + 2 * i] = 4.0 * pc;
        p[2 + 2 * i] = pc;
        const int ip = (i + n - 1) % n, in = (i + 1) % n;
        const double ci = tc * std::cos(2.0 * M_PI * i / n);
        tw[1 + 2 * ip] += 2.0 * ci;
        tw[2 + 2 * ip] += ci;
        tw[1 + 2 * i] += 4.0 * ci;
        tw[2 + 2 * i] += ci;
        tw[1 + 2 * in] += 2.0 * ci;
    }
    // Ep: tangent along edge 0; Em: along edge 1 (rotate by one edge)
    ep[0] = em[0] = p[0];
    for (int i = 1; i < W; ++i) {
        ep[i] = p[i] + tw[i];
        int im = i + 2 * (n - 1);
        if (im >= W) im -= W - 1;
        em[i] = p[i] + tw[im];
    }
}

bool isBoundaryEdge(const Level &L, int e) { return L.efOff[e + 1] - L.efOff[e] == 1; }

// Corner k (0..3) of face g at vertex v.
int cornerOf(const Level &L, int g, int v)
{
    for (int k = 0; k < 4; ++k)
        if (L.fv[L.fvOff[g] + k] == v) return k;
    return -1;
}

// Limit point and edge-tangent points of a smooth boundary corner with
// valence (edges) n >= 2 whose patch face is face `faceInRing` of its ring
// (the boundary limit masks of far/catmarkPatchBuilder.cpp:201-309):
// weights[0] for the corner, weights[i] for ring point i - 1 (edges 0..n-1
// at the even positions, faces between them at the odd ones).
void boundaryPointWeights(int n, int faceInRing, std::vector<double> &p, std::vector<double> &ep,
                          std::vector<double> &em)
{
    const int W = 2 * n, N = W - 1, k = n - 1;
    const double theta = M_PI / k;
    p.assign(W, 0.0);
    ep.assign(W, 0.0);
    em.assign(W, 0.0);
    p[0] = 4.0 / 6.0;
    p[1] = 1.0 / 6.0;
    p[N] = 1.0 / 6.0;
    // tangent across the boundary (interior direction)
    std::vector<double> t(W, 0.0);
    const double c = std::cos(theta), sn = std::sin(theta);
    const double d3kc = 1.0 / (3.0 * k + c);
    const double gamma = -4.0 * sn * d3kc;
    const double a0 = -((1.0 + 2.0 * c) * std::sqrt(1.0 + c)) * d3kc / std::sqrt(1.0 - c);
    const double b0 = sn * d3kc;
    t[0] = gamma / 3.0;
    t[1] = a0 / 3.0;
    t[2] = b0 / 3.0;
    t[N] = a0 / 3.0;
    for (int i = 1; i < n - 1; ++i) {
        const double si = std::sin(theta * i), si1 = std::sin(theta * (i + 1));
        t[1 + 2 * i] = 4.0 * si * d3kc / 3.0;
        t[2 + 2 * i] = (si + si1) * d3kc / 3.0;
    }
    // each edge point mixes the tangent along the boundary with the one
    // across it by the angle of its edge; on a boundary edge it lies on it
    auto edgePoint = [&](std::vector<double> &e, bool onBoundary, int lead, int angleIndex) {
        if (onBoundary) {
            e[0] = 2.0 / 3.0;
            e[lead] = 1.0 / 3.0;
            return;
        }
        const double ca = std::cos(theta * angleIndex), sa = std::sin(theta * angleIndex);
        for (int i = 0; i < W; ++i) e[i] = t[i] * sa;
        e[0] += p[0];
        e[1] += p[1] + ca / 6.0;
        e[N] += p[N] - ca / 6.0;
    };
    edgePoint(ep, faceInRing == 0, 1, faceInRing);
    edgePoint(em, faceInRing == k - 1, N, (faceInRing + 1) % n);
}

// The 20 Gregory points of an end-cap face (Loop, Schaefer, Ni, Castano,
// "Approximating subdivision surfaces with Gregory patches for hardware
// tessellation", as converted in far/catmarkPatchBuilder.cpp:727-1416):
// rows over level vertices, point 5c+{0,1,2,3,4} = P, E+, E-, F+, F- of
// corner c.  Corners are interior (regular or not), on a boundary (two
// faces: regular; more: the boundary limit masks), a single-face smooth
// corner, or sharp.
bool gregoryPoints(const Level &L, int f, Row out[20])
{
    struct Corner {
        int v, numFaces, valence, faceInRing;
        bool boundary, sharp, regular, epOnB, emOnB, fpReg, fmReg, fpCopy, fmCopy;
        double cosA;
        std::vector<int> ring;
    } C[4];
    const int b = L.fvOff[f];
    for (int c = 0; c < 4; ++c) {
        Corner &K = C[c];
        K.v = L.fv[b + c];
        int nb = 0;
        for (int i = L.veOff[K.v]; i < L.veOff[K.v + 1]; ++i) nb += isBoundaryEdge(L, L.ve[i]);
        K.boundary = nb > 0;
        if (!K.boundary) {
            K.numFaces = orderedRing(L, K.v, f, K.ring);
            if (K.numFaces < 3) return false;
            K.valence = K.numFaces;
            K.faceInRing = 0;
            K.sharp = false;
        } else {
            if (nb != 2) return false;
            // ring from the face whose edge (v, next) is on the boundary to the
            // face whose edge (v, prev) is
            int s = f;
            for (int guard = 0; guard < 256; ++guard) {
                const int j = cornerOf(L, s, K.v);
                const int eNext = L.fe[L.fvOff[s] + j];
                if (isBoundaryEdge(L, eNext)) break;
                s = otherFace(L, eNext, s);
                if (s < 0 || L.faceSize(s) != 4) return false;
            }
            K.ring.clear();
            K.faceInRing = -1;
            int g = s, k = 0;
            for (;; ++k) {
                if (L.faceSize(g) != 4 || k > 255) return false;
                const int j = cornerOf(L, g, K.v), gb = L.fvOff[g];
                if (g == f) K.faceInRing = k;
                K.ring.push_back(L.fv[gb + ((j + 1) & 3)]);
                K.ring.push_back(L.fv[gb + ((j + 2) & 3)]);
                const int ePrev = L.fe[gb + ((j + 3) & 3)];
                if (isBoundaryEdge(L, ePrev)) {
                    K.ring.push_back(L.fv[gb + ((j + 3) & 3)]);
                    break;
                }
                g = otherFace(L, ePrev, g);
                if (g < 0) return false;
            }
            K.numFaces = k + 1;
            K.valence = K.numFaces + 1;
            if (K.faceInRing < 0) return false;
            K.sharp = K.numFaces == 1 && L.vsharp[K.v] >= kSharpInfinite;
        }
        K.regular = ((K.numFaces << (K.boundary ? 1 : 0)) == 4) && !K.sharp;
        K.cosA = K.regular ? 0.0
                           : std::cos((K.boundary ? M_PI : 2.0 * M_PI) / (double)K.numFaces);
    }
    for (int c = 0; c < 4; ++c) {
        Corner &K = C[c];
        K.epOnB = K.emOnB = K.fpCopy = K.fmCopy = false;
        K.fpReg = K.regular && C[(c + 1) & 3].regular;
        K.fmReg = K.regular && C[(c + 3) & 3].regular;
        if (K.boundary) {
            K.epOnB = K.faceInRing == 0;
            K.emOnB = K.faceInRing == K.numFaces - 1;
            if (K.numFaces > 1) {
                if (K.epOnB) { K.fpReg = K.fmReg; K.fpCopy = !K.fpReg; }
                if (K.emOnB) { K.fmReg = K.fpReg; K.fmCopy = !K.fmReg; }
            } else {
                K.fpReg = K.fmReg = true;
            }
        }
    }
    // corner and edge points
    for (int c = 0; c < 4; ++c) {
        const Corner &K = C[c];
        const int cv = K.v, vn = C[(c + 1) & 3].v, vp = C[(c + 3) & 3].v;
        Row &P = out[5 * c], &Ep = out[5 * c + 1], &Em = out[5 * c + 2];
        const std::vector<int> &r = K.ring;
        if (K.sharp) {
            P.add(cv, 1.0);
            Ep.add(cv, 2.0 / 3.0);
            Ep.add(vn, 1.0 / 3.0);
            Em.add(cv, 2.0 / 3.0);
            Em.add(vp, 1.0 / 3.0);
        } else if (!K.boundary && K.regular) {
            P.add(cv, 4.0 / 9.0);
            for (int k = 0; k < 4; ++k) {
                P.add(r[2 * k], 1.0 / 9.0);
                P.add(r[2 * k + 1], 1.0 / 36.0);
            }
            Ep.add(cv, 4.0 / 9.0);
            Ep.add(r[0], 2.0 / 9.0);
            Ep.add(r[2], 1.0 / 9.0);
            Ep.add(r[6], 1.0 / 9.0);
            Ep.add(r[1], 1.0 / 18.0);
            Ep.add(r[7], 1.0 / 18.0);
            Em.add(cv, 4.0 / 9.0);
            Em.add(r[2], 2.0 / 9.0);
            Em.add(r[0], 1.0 / 9.0);
            Em.add(r[4], 1.0 / 9.0);
            Em.add(r[1], 1.0 / 18.0);
            Em.add(r[3], 1.0 / 18.0);
        } else if (!K.boundary) {
            std::vector<double> pw, epw, emw;
            interiorPointWeights(K.valence, pw, epw, emw);
            P.add(cv, pw[0]);
            Ep.add(cv, epw[0]);
            Em.add(cv, emw[0]);
            for (int i = 1; i < (int)pw.size(); ++i) {
                P.add(r[i - 1], pw[i]);
                Ep.add(r[i - 1], epw[i]);
                Em.add(r[i - 1], emw[i]);
            }
        } else if (K.regular) {
            // two faces: the boundary crease limit and its tangents
            P.add(cv, 2.0 / 3.0);
            P.add(r[0], 1.0 / 6.0);
            P.add(r[4], 1.0 / 6.0);
            Row &eB = K.epOnB ? Ep : Em, &eI = K.epOnB ? Em : Ep;
            eB.add(cv, 2.0 / 3.0);
            eB.add(r[K.epOnB ? 0 : 4], 1.0 / 3.0);
            eI.add(cv, 4.0 / 9.0);
            eI.add(r[2], 2.0 / 9.0);
            eI.add(r[0], 1.0 / 9.0);
            eI.add(r[4], 1.0 / 9.0);
            eI.add(r[1], 1.0 / 18.0);
            eI.add(r[3], 1.0 / 18.0);
        } else if (K.numFaces > 1) {
            std::vector<double> pw, epw, emw;
            boundaryPointWeights(K.valence, K.faceInRing, pw, epw, emw);
            const int N = (int)pw.size() - 1;
            P.add(cv, pw[0]);
            P.add(r[0], pw[1]);
            P.add(r[2 * (K.valence - 1)], pw[N]);
            Ep.add(cv, epw[0]);
            Em.add(cv, emw[0]);
            if (K.epOnB) Ep.add(r[0], epw[1]);
            else for (int i = 1; i <= N; ++i) Ep.add(r[i - 1], epw[i]);
            if (K.emOnB) Em.add(r[2 * (K.valence - 1)], emw[N]);
            else for (int i = 1; i <= N; ++i) Em.add(r[i - 1], emw[i]);
        } else {
            // a smooth single-face corner
            P.add(cv, 4.0 / 6.0);
            P.add(vn, 1.0 / 6.0);
            P.add(vp, 1.0 / 6.0);
            Ep.add(cv, 2.0 / 3.0);
            Ep.add(vn, 1.0 / 3.0);
            Em.add(cv, 2.0 / 3.0);
            Em.add(vp, 1.0 / 3.0);
        }
    }
    // face points
    for (int c = 0; c < 4; ++c) {
        const Corner &K = C[c];
        const int cn = (c + 1) & 3, co = (c + 2) & 3, cp = (c + 3) & 3;
        const std::vector<int> &r = K.ring;
        const int n = K.valence;
        for (int side = 0; side < 2; ++side) {
            Row &F = out[5 * c + 3 + side];
            const bool reg = side == 0 ? K.fpReg : K.fmReg;
            const bool copied = side == 0 ? K.fpCopy : K.fmCopy;
            if (reg) {
                F.add(K.v, 4.0 / 9.0);
                F.add(C[cp].v, 2.0 / 9.0);
                F.add(C[cn].v, 2.0 / 9.0);
                F.add(C[co].v, 1.0 / 9.0);
                continue;
            }
            if (copied) continue;
            const int far = side == 0 ? cn : cp;
            const double cN = K.cosA, cF = C[far].cosA;
            const Row &eNear = out[5 * c + 1 + side];                 // E+ or E- of c
            const Row &eFar = out[5 * far + (side == 0 ? 2 : 1)];    // E- of next / E+ of prev
            F.addRow(out[5 * c], cF / 3.0);
            F.addRow(eNear, (3.0 - 2.0 * cN - cF) / 3.0);
            F.addRow(eFar, 2.0 * cN / 3.0);
            const double sgn = side == 0 ? 1.0 : -1.0;
            const int eInt = side == 0 ? K.faceInRing : (K.faceInRing + 1) % n;
            const int ePrev = (eInt + n - 1) % n, eNext = (eInt + 1) % n;
            F.add(r[2 * ePrev], -sgn / 9.0);
            F.add(r[2 * ePrev + 1], -sgn / 18.0);
            F.add(r[2 * eInt + 1], sgn / 18.0);
            F.add(r[2 * eNext], sgn / 9.0);
        }
        if (K.fpCopy) out[5 * c + 3] = out[5 * c + 4];
        if (K.fmCopy) out[5 * c + 4] = out[5 * c + 3];
    }
    return true;
}

// Far's feature test for refining face f of a level (scope: manifold,
// interior; far/topologyRefiner.cpp:627-690): non-quad, a base-level corner
// shared with a non-quad face, an extraordinary corner, or a corner with
// vertex sharpness or an incident sharp edge.  `reduced`: the feature set
// of levels at or past the secondary level (FeatureMask::ReduceFeatures,
// far/topologyRefiner.cpp:367-380): extraordinary corners (interior or on a
// boundary) no longer select the face, sharp features still do.

// A corner a B-spline patch reproduces (Far's regular tags, far/patchBuilder.cpp:
// 581-700 for this scope): interior with 4 faces; on a boundary with 2 faces
// (the boundary crease); or a boundary corner of 1 face made sharp by
// edge-and-corner interpolation.
bool regularCorner(const Level &L, int v)
{
    const int nf = L.vfOff[v + 1] - L.vfOff[v], ne = L.valence(v);
    int nb = 0;
    for (int i = L.veOff[v]; i < L.veOff[v + 1]; ++i) nb += isBoundaryEdge(L, L.ve[i]);
    if (nb == 0) return nf == 4 && ne == 4;
    if (nf == 2) return ne == 3 && nb == 2 && L.vsharp[v] <= 0.0f;
    return nf == 1 && ne == 2 && L.vsharp[v] >= kSharpInfinite;
}

bool faceHasFeatures(const Level &L, int f, bool baseLevel, bool reduced)
{
    if (L.faceSize(f) != 4) return true;
    for (int k = L.fvOff[f]; k < L.fvOff[f + 1]; ++k) {
        const int v = L.fv[k];
        // a base-level corner shared with an n-gon (Far's _incidIrregFace tag)
        if (baseLevel)
            for (int i = L.vfOff[v]; i < L.vfOff[v + 1]; ++i)
                if (L.faceSize(L.vf[i]) != 4) return true;
        if (!reduced && !regularCorner(L, v)) return true;
        // semi-sharp or interior infinitely sharp vertex or edge (Far selects
        // inf-sharp creases and corners when inf-sharp patches are off,
        // far/topologyRefiner.cpp:340-360); the infinitely sharp boundary
        // and its corners are regular
        const bool corner1 = L.vfOff[v + 1] - L.vfOff[v] == 1;
        if (L.vsharp[v] > 0.0f && !(corner1 && L.vsharp[v] >= kSharpInfinite)) return true;
        for (int i = L.veOff[v]; i < L.veOff[v + 1]; ++i) {
            const int e = L.ve[i];
            if (L.esharp[e] > 0.0f && !isBoundaryEdge(L, e)) return true;
        }
    }
    return false;
}

// Whether a leaf face needs an end cap.  A smooth single-face boundary
// corner is isolated like an irregular point but then, with legacy
// sharp-corner patches (Far's default, generateLegacySharpCornerPatches,
// far/patchBuilder.cpp:676-686), patched as a sharp corner; without them it
// gets an end cap too.
bool isExtraordinary(const Level &L, int f, bool legacySharpCorners)
{
    for (int k = L.fvOff[f]; k < L.fvOff[f + 1]; ++k) {
        const int v = L.fv[k];
        if (legacySharpCorners && L.vfOff[v + 1] - L.vfOff[v] == 1 && L.valence(v) == 2) continue;
        if (!regularCorner(L, v)) return true;
    }
    return false;
}

// 16 B-spline points of an interior regular face (far/patchBuilder.cpp:
// 878-968): corner c of the face sits at kCornerPoint[c][0], the vertices of
// the face diagonally opposite at that corner fill the other three slots.
const int kCornerPoint[4][4] = { { 5, 4, 0, 1 }, { 6, 2, 3, 7 }, { 10, 11, 15, 14 }, { 9, 13, 12, 8 } };

// Boundary edges of face f as PatchParam's boundary mask: bit c for the
// face's edge c, from corner c to corner c + 1 (far/patchBuilder.cpp:718-762).
int boundaryMask(const Level &L, int f)
{
    int m = 0;
    for (int c = 0; c < 4; ++c) m |= isBoundaryEdge(L, L.fe[L.fvOff[f] + c]) << c;
    return m;
}


bool regularPoints(const Level &L, int f, int mask, int pts[16])
{
    // far/patchBuilder.cpp:876-966: a corner without boundary edges takes the
    // three points of the face diagonally opposite; with a boundary edge the
    // missing points are the face's first vertex (fillMissingBoundaryPoints)
    const int b = L.fvOff[f];
    const int fill = L.fv[b];
    std::vector<int> ring;
    for (int c = 0; c < 4; ++c) {
        const int v = L.fv[b + c];
        const bool lead = (mask >> c) & 1, trail = (mask >> ((c + 3) & 3)) & 1;
        pts[kCornerPoint[c][0]] = v;
        if (!lead && !trail) {
            if (orderedRing(L, v, f, ring) != 4) return false;
            pts[kCornerPoint[c][1]] = ring[4];
            pts[kCornerPoint[c][2]] = ring[5];
            pts[kCornerPoint[c][3]] = ring[6];
        } else if (lead && trail) {
            pts[kCornerPoint[c][1]] = pts[kCornerPoint[c][2]] = pts[kCornerPoint[c][3]] = fill;
        } else if (lead) {
            // the other face at v lies across the trailing edge; its corner
            // before v
            const int g = otherFace(L, L.fe[b + ((c + 3) & 3)], f);
            const int k = g < 0 ? -1 : cornerOf(L, g, v);
            if (k < 0) return false;
            pts[kCornerPoint[c][1]] = L.fv[L.fvOff[g] + ((k + 3) & 3)];
            pts[kCornerPoint[c][2]] = pts[kCornerPoint[c][3]] = fill;
        } else {
            // across the leading edge; its corner after v
            const int g = otherFace(L, L.fe[b + c], f);
            const int k = g < 0 ? -1 : cornerOf(L, g, v);
            if (k < 0) return false;
            pts[kCornerPoint[c][1]] = pts[kCornerPoint[c][2]] = fill;
            pts[kCornerPoint[c][3]] = L.fv[L.fvOff[g] + ((k + 1) & 3)];
        }
    }
    return true;
}

}  // namespace

bool PatchTable::build(Refiner &refiner, int isoLevel, const std::vector<LevelStencils> *levels,
                       const PatchOptions &opt)
{
    // The bilinear scheme (linear QUADS patches): Far refines only its
    // irregular faces, one level, whatever the isolation level
    // (far/topologyRefiner.cpp:400-420, selectLinearIrregularFaces); every
    // leaf is a QUADS patch of its face's 4 vertices without boundary bits
    // (far/patchBuilder.cpp:720-724).
    const bool linear = refiner.scheme == SCHEME_BILINEAR;
    if (refiner.scheme != SCHEME_CATMARK && !linear) return false;
    bool hasNgon = false;
    {
        const Level &L0 = refiner.level(0);
        for (int f = 0; f < L0.nf; ++f) hasNgon = hasNgon || L0.faceSize(f) != 4;
        if (hasNgon && isoLevel < 1 && !linear) return false;   // n-gons need a level
        // boundaries need interpolated boundary edges; interior edges and
        // vertices must not be infinitely sharp (only boundary corners are)
        bool boundary = false;
        for (int e = 0; e < L0.ne; ++e) {
            const int nf = L0.efOff[e + 1] - L0.efOff[e];
            if (nf == 1) boundary = true;
            else if (nf != 2) return false;
        }
        if (boundary && refiner.boundaryInterpolation == 0) return false;
    }
    const int maxL = linear ? (hasNgon ? 1 : 0) : isoLevel;
    refiner.refineUniform(maxL);                               // may reallocate the levels
    std::vector<LevelStencils> ownLS;
    if (!levels || (int)levels->size() <= maxL) {
        computeLevelStencils(refiner, INTERPOLATE_VERTEX, ownLS);
        levels = &ownLS;
    }
    const std::vector<LevelStencils> &LS = *levels;

    numControl = refiner.level(0).nv;
    maxDepth = 0;
    struct Leaf { int level, face; bool regular; int transition; };
    std::vector<Leaf> leaves;
    std::vector<char> complete(refiner.level(0).nf, 1), next, selected;
    for (int d = 0; d <= maxL; ++d) {
        const Level &L = refiner.level(d);
        const size_t first = leaves.size();
        if (d < maxL) next.assign(refiner.level(d + 1).nf, 0);
        selected.assign(L.nf, 0);
        for (int f = 0; f < L.nf; ++f) {
            if (!complete[f]) continue;
            // levels at or past the secondary level use the reduced
            // feature set (far/topologyRefiner.cpp:412-470)
            const bool reduced = d >= std::min(opt.secondaryLevel, isoLevel);
            const bool feature = linear ? d == 0 && L.faceSize(f) != 4
                                        : faceHasFeatures(L, f, d == 0, reduced);
            if (d < maxL && feature) {
                selected[f] = 1;
                for (int j = 0; j < L.faceSize(f); ++j) next[L.fvOff[f] + j] = 1;
            } else {
                leaves.push_back({ d, f, linear || !isExtraordinary(L, f, opt.legacySharpCorners),
                                   0 });
                maxDepth = std::max(maxDepth, d);
            }
        }
        // PatchParam's transition mask (far/patchBuilder.cpp:1374-1379): the
        // edges of a leaf that the next level splits, i.e. that a selected
        // face shares
        if (d < maxL)
            for (size_t i = first; i < leaves.size(); ++i) {
                const int f = leaves[i].face;
                for (int k = 0; k < 4; ++k) {
                    const int e = L.fe[L.fvOff[f] + k];
                    for (int j = L.efOff[e]; j < L.efOff[e + 1]; ++j)
                        if (selected[L.ef[j]]) leaves[i].transition |= 1 << k;
                }
            }
        if (d < maxL) complete.swap(next);
    }

    // refined vertices referenced by B-spline patches
    std::vector<std::vector<int>> bufIndex(maxL + 1);
    for (int d = 1; d <= maxL; ++d) bufIndex[d].assign(refiner.level(d).nv, -1);
    std::vector<int> regPts;
    std::vector<Row> greg;
    std::vector<const Leaf *> regLeaves, gregLeaves;
    const int stride = linear ? 4 : 16;                        // points per regular patch
    for (const Leaf &lf : leaves) {
        if (lf.regular) {
            int pts[16];
            const Level &L = refiner.level(lf.level);
            if (linear) {
                for (int k = 0; k < 4; ++k) pts[k] = L.fv[L.fvOff[lf.face] + k];
            } else if (!regularPoints(L, lf.face, boundaryMask(L, lf.face), pts)) {
                return false;
            }
            if (lf.level > 0)
                for (int k = 0; k < stride; ++k) bufIndex[lf.level][pts[k]] = 0;
            regPts.insert(regPts.end(), pts, pts + stride);
            regLeaves.push_back(&lf);
        } else {
            Row rows[20];
            if (!gregoryPoints(refiner.level(lf.level), lf.face, rows)) return false;
            greg.insert(greg.end(), rows, rows + 20);
            gregLeaves.push_back(&lf);
        }
    }
    int next_index = numControl;
    for (int d = 1; d <= maxL; ++d)
        for (int &b : bufIndex[d])
            if (b == 0) b = next_index++;
    numRefined = next_index - numControl;
    // End-cap points: corner points shared by every cap at the vertex, edge
    // points by the two caps on an interior edge, face points never -- Far's
    // shareEndCapPatchPoints (far/patchTableFactory.cpp:1686-1860): a point
    // reuses the first cap's local point, numbered in order of appearance.
    std::vector<int> gregLocal(greg.size(), -1), newRows;
    {
        std::vector<std::vector<int>> cornerPt(maxL + 1), edgePt(maxL + 1);
        for (size_t i = 0; i < gregLeaves.size(); ++i) {
            const Leaf &lf = *gregLeaves[i];
            const Level &L = refiner.level(lf.level);
            std::vector<int> &cp = cornerPt[lf.level], &ep = edgePt[lf.level];
            if (cp.empty()) cp.assign(L.nv, -1);
            if (ep.empty()) ep.assign(2 * (size_t)L.ne, -1);
            const int b = L.fvOff[lf.face];
            for (int k = 0; k < 20; ++k) {
                const int c = k / 5, kind = k % 5;
                int *slot = nullptr;
                if (kind == 0) {
                    slot = &cp[L.fv[b + c]];
                } else if (kind <= 2) {
                    // E+ near corner c of edge c; E- near corner c of edge c - 1
                    const int fe = kind == 1 ? c : (c + 3) & 3;
                    const int e = L.fe[b + fe];
                    if (!isBoundaryEdge(L, e)) {
                        const bool reversed = L.ev[2 * e] != L.fv[b + fe];
                        const int end = (kind == 2) != reversed;
                        slot = &ep[2 * (size_t)e + end];
                    }
                }
                if (!opt.shareEndCapPoints) slot = nullptr;
                int id;
                if (slot && *slot >= 0) {
                    id = *slot;
                } else {
                    id = (int)newRows.size();
                    newRows.push_back((int)(20 * i + k));
                    if (slot) *slot = id;
                }
                gregLocal[20 * i + k] = id;
            }
        }
    }
    numLocal = (int)newRows.size();

    // patch arrays / indices / params, and point stencils
    arrays.clear();
    indices.clear();
    params.clear();
    patchDepth_.clear();
    patchLeaf_.clear();
    fvar.clear();
    auto paramOf = [&](const Leaf &lf, bool regular) {
        PatchParamRec p;
        int ptex, u = 0, v = 0;
        if (lf.level == 0) {
            ptex = refiner.ptexBase[lf.face];
        } else {
            const FaceOrigin &o = refiner.origins[lf.level][lf.face];
            ptex = o.ptex;
            u = o.u;
            v = o.v;
        }
        p.field0 = ((uint32_t)ptex & 0xfffffffu) | ((uint32_t)lf.transition << 28);
        // far/patchBuilder.cpp:1272-1380: depth = level; under an n-gon the
        // (u,v) bits count from level 1, flagged by the non-quad bit
        const uint32_t bmask = regular && !linear
                                   ? (uint32_t)boundaryMask(refiner.level(lf.level), lf.face) : 0u;
        p.field1 = (uint32_t)lf.level | ((uint32_t)refiner.ptexNonQuad[ptex] << 4) |
                   ((uint32_t)(regular ? 1 : 0) << 5) | (bmask << 7) | ((uint32_t)v << 12) |
                   ((uint32_t)u << 22);
        p.sharpness = 0.0f;
        return p;
    };
    const int nReg = (int)regLeaves.size(), nGreg = (int)gregLeaves.size();
    const int regType = linear ? kQuads : kRegular;
    if (nReg) arrays.push_back({ regType, regType, nReg, 0, stride, 0 });
    for (int i = 0; i < nReg; ++i) {
        const Leaf &lf = *regLeaves[i];
        for (int k = 0; k < stride; ++k) {
            const int p = regPts[stride * i + k];
            indices.push_back(lf.level == 0 ? p : bufIndex[lf.level][p]);
        }
        params.push_back(paramOf(lf, true));
        patchDepth_.push_back(lf.level);
        patchLeaf_.push_back({ lf.level, lf.face });
    }
    if (nGreg) arrays.push_back({ kGregory, kGregory, nGreg, (int)indices.size(), 20, nReg });
    for (int i = 0; i < nGreg; ++i) {
        for (int k = 0; k < 20; ++k) indices.push_back(numControl + numRefined + gregLocal[20 * i + k]);
        params.push_back(paramOf(*gregLeaves[i], false));
        patchDepth_.push_back(gregLeaves[i]->level);
        patchLeaf_.push_back({ gregLeaves[i]->level, gregLeaves[i]->face });
    }

    // point stencils: refined vertices (level order), then end-cap points
    LevelStencils &S = pointStencils;
    S.off.assign(1, 0);
    S.idx.clear();
    S.w.clear();
    for (int d = 1; d <= maxL; ++d)
        for (int v = 0; v < refiner.level(d).nv; ++v)
            if (bufIndex[d][v] >= 0) {
                const LevelStencils &src = LS[d];
                S.idx.insert(S.idx.end(), src.idx.begin() + src.off[v], src.idx.begin() + src.off[v + 1]);
                S.w.insert(S.w.end(), src.w.begin() + src.off[v], src.w.begin() + src.off[v + 1]);
                S.off.push_back((int64_t)S.idx.size());
            }
    {
        // an end cap's rows are over the vertices of its own level (the last
        // one, or a shallower one under a secondary level); level 0 is the
        // control vertices themselves
        std::vector<int> pos(numControl, -1), list;
        std::vector<double> val;
        auto add = [&](int c, double w) {
            if (pos[c] < 0) { pos[c] = (int)list.size(); list.push_back(c); val.push_back(0.0); }
            val[pos[c]] += w;
        };
        for (int rowIndex : newRows) {
            const Row &row = greg[(size_t)rowIndex];
            const int lvl = gregLeaves[(size_t)rowIndex / 20]->level;
            if (lvl == 0) {
                for (auto &x : row.e) add(x.first, x.second);
            } else {
                const LevelStencils &src = LS[lvl];
                for (auto &x : row.e)
                    for (int64_t k = src.off[x.first]; k < src.off[x.first + 1]; ++k)
                        add(src.idx[k], x.second * src.w[k]);
            }
            std::vector<int> ord(list.size());
            for (size_t k = 0; k < ord.size(); ++k) ord[k] = (int)k;
            std::sort(ord.begin(), ord.end(), [&](int a, int b) { return list[a] < list[b]; });
            for (int k : ord) { S.idx.push_back(list[k]); S.w.push_back(val[k]); pos[list[k]] = -1; }
            S.off.push_back((int64_t)S.idx.size());
            list.clear();
            val.clear();
        }
    }

    // patch map: one quadtree per ptex face
    ptexRoot.assign(refiner.ptexCount, -1);
    nodes.clear();
    auto rootOf = [&](int ptex) {
        if (ptexRoot[ptex] < 0) {
            ptexRoot[ptex] = (int)nodes.size();
            nodes.push_back({ { -INT_MAX, -INT_MAX, -INT_MAX, -INT_MAX } });
        }
        return ptexRoot[ptex];
    };
    const int nPatches = (int)params.size();
    for (int p = 0; p < nPatches; ++p) {
        const uint32_t f1 = params[p].field1;
        const int ptex = (int)(params[p].field0 & 0xfffffffu);
        const int depth = (int)(f1 & 0xf), u = (int)((f1 >> 22) & 0x3ff), v = (int)((f1 >> 12) & 0x3ff);
        // a patch at the root depth (0, or 1 under an n-gon) covers its
        // whole ptex face (far/patchMap.cpp:155-180)
        const int rootDepth = (int)((f1 >> 4) & 1);
        int node = rootOf(ptex);
        if (depth == rootDepth) {
            for (int q = 0; q < 4; ++q) nodes[node].child[q] = -(p + 1);
            continue;
        }
        for (int k = depth - rootDepth - 1; k >= 0; --k) {
            const int q = (((v >> k) & 1) << 1) | ((u >> k) & 1);
            if (k == 0) {
                nodes[node].child[q] = -(p + 1);
            } else {
                int c = nodes[node].child[q];
                if (c < 0) {
                    c = (int)nodes.size();
                    nodes[node].child[q] = c;
                    nodes.push_back({ { -INT_MAX, -INT_MAX, -INT_MAX, -INT_MAX } });
                }
                node = c;
            }
        }
    }
    return true;
}

bool PatchTable::findPatch(int ptex, double u, double v, PatchHandleRec &h) const
{
    if (ptex < 0 || ptex >= (int)ptexRoot.size() || ptexRoot[ptex] < 0) return false;
    if (!(u >= 0.0 && u <= 1.0 && v >= 0.0 && v <= 1.0)) return false;
    int node = ptexRoot[ptex];
    double median = 0.5;
    for (int depth = 0; depth <= maxDepth; ++depth, median *= 0.5) {
        // far/patchMap.h:136-143: the upper half includes the median
        const int uh = u >= median, vh = v >= median;
        if (uh) u -= median;
        if (vh) v -= median;
        const int c = nodes[node].child[(vh << 1) | uh];
        if (c == -INT_MAX) return false;
        if (c < 0) {
            const int p = -c - 1;
            h.arrayIndex = (arrays.size() > 1 && p >= arrays[0].numPatches) ? 1 : 0;
            h.patchIndex = p;
            // far/patchMap.cpp:100-125: the first point of the patch
            // counted within its array (j * patch size), not in the index buffer
            h.vertIndex = arrays[h.arrayIndex].stride * (p - arrays[h.arrayIndex].primitiveIdBase);
            return true;
        }
        node = c;
    }
    return false;
}

const int *PatchTable::patchPoints(int p, int &n) const
{
    const int a = (arrays.size() > 1 && p >= arrays[0].numPatches) ? 1 : 0;
    n = arrays[a].stride;
    return indices.data() + arrays[a].indexBase + arrays[a].stride * (p - arrays[a].primitiveIdBase);
}

int PatchTable::evalBasis(int p, double u, double v, double *wP, double *wDu, double *wDv,
                          double *wDuu, double *wDuv, double *wDvv) const
{
    const int a = (arrays.size() > 1 && p >= arrays[0].numPatches) ? 1 : 0;
    return patchBasis<double>(arrays[a].desc, params[p].field1, u, v, wP, wDu, wDv, wDuu, wDuv, wDvv);
}

template <typename REAL>
void PatchTable::pointStencilTable(StencilTableT<REAL> &out) const
{
    out = StencilTableT<REAL>();
    out.numControlVertices = numControl;
    const int n = (int)pointStencils.off.size() - 1;
    out.sizes.resize(n);
    out.offsets.resize(n);
    for (int i = 0; i < n; ++i) {
        out.sizes[i] = pointStencils.size(i);
        out.offsets[i] = pointStencils.off[i];
    }
    out.indices.assign(pointStencils.idx.begin(), pointStencils.idx.end());
    out.weights.assign(pointStencils.w.begin(), pointStencils.w.end());
}

template <typename REAL>
int PatchTable::limitStencils(int n, const int *ptex, const double *u, const double *v,
                              bool second, StencilTableT<REAL> &out) const
{
    out = StencilTableT<REAL>();
    out.numControlVertices = numControl;
    const int nw = second ? 6 : 3;
    int T = 1;
#ifdef _OPENMP
    T = std::max(1, std::min(omp_get_max_threads(), (n + 4095) / 4096));
#endif
    std::vector<std::vector<int>> tIdx(T), tSz(T);
    std::vector<std::vector<double>> tW(T);
    int missing = 0;
#pragma omp parallel num_threads(T) reduction(+ : missing)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        const int lo = (int)((int64_t)n * t / T), hi = (int)((int64_t)n * (t + 1) / T);
        std::vector<int> pos(numControl, -1), list;
        std::vector<double> acc;    // nw per entry
        double w[6][20];
        for (int i = lo; i < hi; ++i) {
            PatchHandleRec h;
            if (!findPatch(ptex[i], u[i], v[i], h)) {
                ++missing;                   // no stencil, as Far drops the location
                continue;
            }
            const int np = evalBasis(h.patchIndex, u[i], v[i], w[0], w[1], w[2],
                                     second ? w[3] : nullptr, second ? w[4] : nullptr,
                                     second ? w[5] : nullptr);
            int cnt;
            const int *pts = patchPoints(h.patchIndex, cnt);
            for (int j = 0; j < np; ++j) {
                const int cvi = pts[j];
                auto addc = [&](int c, double s) {
                    if (pos[c] < 0) {
                        pos[c] = (int)list.size();
                        list.push_back(c);
                        acc.insert(acc.end(), (size_t)nw, 0.0);
                    }
                    double *a = &acc[(size_t)pos[c] * nw];
                    for (int k = 0; k < nw; ++k) a[k] += s * w[k][j];
                };
                if (cvi < numControl) {
                    addc(cvi, 1.0);
                } else {
                    const int r = cvi - numControl;
                    for (int64_t k = pointStencils.off[r]; k < pointStencils.off[r + 1]; ++k)
                        addc(pointStencils.idx[k], pointStencils.w[k]);
                }
            }
            std::vector<int> ord(list.size());
            for (size_t k = 0; k < ord.size(); ++k) ord[k] = (int)k;
            std::sort(ord.begin(), ord.end(), [&](int a, int b) { return list[a] < list[b]; });
            for (int k : ord) {
                tIdx[t].push_back(list[k]);
                for (int q = 0; q < nw; ++q) tW[t].push_back(acc[(size_t)k * nw + q]);
                pos[list[k]] = -1;
            }
            tSz[t].push_back((int)list.size());
            list.clear();
            acc.clear();
        }
    }
    out.sizes.reserve(n);
    out.offsets.reserve(n);
    int64_t off = 0;
    for (int t = 0; t < T; ++t)
        for (int s : tSz[t]) { out.sizes.push_back(s); out.offsets.push_back(off); off += s; }
    out.indices.reserve(off);
    hvec<REAL> *dst[6] = { &out.weights, &out.du, &out.dv, &out.duu, &out.duv, &out.dvv };
    for (int k = 0; k < nw; ++k) dst[k]->reserve(off);
    for (int t = 0; t < T; ++t) {
        out.indices.insert(out.indices.end(), tIdx[t].begin(), tIdx[t].end());
        const size_t m = tIdx[t].size();
        for (size_t e = 0; e < m; ++e)
            for (int k = 0; k < nw; ++k) dst[k]->push_back((REAL)tW[t][e * nw + k]);
    }
    return missing;
}

template void PatchTable::pointStencilTable<float>(StencilTable &) const;
template void PatchTable::pointStencilTable<double>(StencilTableD &) const;
template int PatchTable::limitStencils<float>(int, const int *, const double *, const double *,
                                              bool, StencilTable &) const;
template int PatchTable::limitStencils<double>(int, const int *, const double *, const double *,
                                               bool, StencilTableD &) const;

void uniformVertexLocations(const Refiner &r, int level, std::vector<int> &ptex,
                            std::vector<double> &u, std::vector<double> &v)
{
    const Level &L = r.level(level);
    ptex.assign(L.nv, -1);
    u.assign(L.nv, 0.0);
    v.assign(L.nv, 0.0);
    const int du[4] = { 0, 1, 1, 0 }, dv[4] = { 0, 0, 1, 1 };
    for (int f = 0; f < L.nf; ++f) {
        int p = 0, fu = 0, fv = 0;
        if (level == 0) {
            p = r.ptexBase[f];
            if (L.faceSize(f) != 4) {
                // corner k of an n-gon is the origin of its ptex face k
                for (int k = 0; k < L.faceSize(f); ++k) {
                    const int vert = L.fv[L.fvOff[f] + k];
                    if (ptex[vert] >= 0) continue;
                    ptex[vert] = p + k;
                }
                continue;
            }
        } else {
            const FaceOrigin &o = r.origins[level][f];
            p = o.ptex;
            fu = o.u;
            fv = o.v;
        }
        const double scale = 1.0 / (double)(1 << (level - (level ? r.ptexNonQuad[p] : 0)));
        for (int k = 0; k < 4; ++k) {
            const int vert = L.fv[L.fvOff[f] + k];
            if (ptex[vert] >= 0) continue;
            ptex[vert] = p;
            u[vert] = (fu + du[k]) * scale;
            v[vert] = (fv + dv[k]) * scale;
        }
    }
}

}  // namespace osdfar
