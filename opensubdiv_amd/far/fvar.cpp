// Face-varying channels of feature-adaptive patch tables, as Far builds them
// with its default options: the channel refines linearly everywhere
// (Sdc::Options::FVAR_LINEAR_ALL, sdc/options.h:60-67, the Sdc default) and
// every patch gets a linear QUADS face-varying patch over the four values
// of its face (generateFVarLegacyLinearPatches, the PatchTableFactory
// default: far/patchTableFactory.cpp:1255-1266 assigns the face's values
// with assignFacePoints, :655-675, and copies the vertex PatchParam).
//
// Linear refinement of a value seen from a parent face F (corners c_0 ..
// c_{n-1}) gives the child values of far/primvarRefiner.h:475-660 with the
// linear masks: a vertex child keeps c_i (weight 1), an edge child is
// (c_i + c_{i+1}) / 2 on the side of F, the face child is the mean of F's
// corners.  Values are computed per face-vertex, so they split along seams
// (edges whose two faces disagree on a value) exactly as Far's siblings do;
// equal rows at one vertex are one value.  Rows are factorized to the base
// values, accumulated in double and sorted by base index.
//
// Numbering of the face-varying buffer the channel indexes:
//   [ base values | refined values used by the patches ]
// with refined values in level order, then vertex order, then first use.
// Far numbers every value of its sparse levels (vtr/fvarRefinement.cpp:
// 290-340), so indices may differ from Far's while every patch evaluates to
// the same surface.
#include <algorithm>
#include <map>

#include "patches.h"

namespace osdfar {

namespace {

constexpr int kFVarLinear = 3;   // Far::PatchDescriptor::QUADS

using Entry = std::pair<int, double>;
using SRow = std::vector<Entry>;

// s * a + t * b (rows sorted by index) into out, sorted.
void blend(const SRow &a, double s, const SRow &b, double t, SRow &out)
{
    out.clear();
    size_t i = 0, j = 0;
    while (i < a.size() || j < b.size()) {
        if (j == b.size() || (i < a.size() && a[i].first < b[j].first)) {
            out.push_back({ a[i].first, s * a[i].second });
            ++i;
        } else if (i == a.size() || b[j].first < a[i].first) {
            out.push_back({ b[j].first, t * b[j].second });
            ++j;
        } else {
            out.push_back({ a[i].first, s * a[i].second + t * b[j].second });
            ++i;
            ++j;
        }
    }
}

void mean(const std::vector<const SRow *> &rows, SRow &out)
{
    std::map<int, double> acc;
    const double w = 1.0 / (double)rows.size();
    for (const SRow *r : rows)
        for (const Entry &e : *r) acc[e.first] += w * e.second;
    out.assign(acc.begin(), acc.end());
}

}  // namespace

bool PatchTable::buildFaceVarying(const Refiner &r, int numValues, const int *valueIndices)
{
    const Level &L0 = r.level(0);
    for (size_t k = 0; k < L0.fv.size(); ++k)
        if (valueIndices[k] < 0 || valueIndices[k] >= numValues) return false;
    const int maxL = patchLeaf_.empty() ? 0 : maxDepth;
    if (r.maxLevel() < maxL) return false;

    // faces whose rows are needed: the leaves and their ancestors
    std::vector<std::vector<char>> need(maxL + 1);
    for (int d = 0; d <= maxL; ++d) need[d].assign(r.level(d).nf, 0);
    for (const auto &lf : patchLeaf_) need[lf.first][lf.second] = 1;
    auto parentOf = [&](int d, int cf, int &corner) {
        // child face cf of level d + 1 is face-vertex cf of level d
        const std::vector<int> &off = r.level(d).fvOff;
        const int f = (int)(std::upper_bound(off.begin(), off.end(), cf) - off.begin()) - 1;
        corner = cf - off[f];
        return f;
    };
    for (int d = maxL; d >= 1; --d)
        for (int cf = 0; cf < r.level(d).nf; ++cf)
            if (need[d][cf]) {
                int j;
                need[d - 1][parentOf(d - 1, cf, j)] = 1;
            }

    // rows of every face-vertex of a needed face, level by level; first[d][f]
    // is face f's first row in rows[d] (-1: not needed)
    std::vector<std::vector<int>> first(maxL + 1);
    std::vector<std::vector<SRow>> rows(maxL + 1);
    for (int d = 0; d <= maxL; ++d) {
        const Level &L = r.level(d);
        first[d].assign(L.nf, -1);
        int count = 0;
        for (int f = 0; f < L.nf; ++f)
            if (need[d][f]) {
                first[d][f] = count;
                count += L.faceSize(f);
            }
        rows[d].resize(count);
    }
    for (int f = 0; f < L0.nf; ++f)
        if (need[0][f])
            for (int k = 0; k < L0.faceSize(f); ++k)
                rows[0][first[0][f] + k] = { { valueIndices[L0.fvOff[f] + k], 1.0 } };
    std::vector<const SRow *> corners;
    for (int d = 1; d <= maxL; ++d) {
        const Level &P = r.level(d - 1), &Cl = r.level(d);
        for (int cf = 0; cf < Cl.nf; ++cf) {
            if (!need[d][cf]) continue;
            int j;
            const int f = parentOf(d - 1, cf, j);
            const int n = P.faceSize(f);
            const SRow *Pr = &rows[d - 1][first[d - 1][f]];
            const SRow &Rj = Pr[j], &Rn = Pr[(j + 1) % n], &Rp = Pr[(j + n - 1) % n];
            // child corners before rotation (the refiner's order): vertex
            // child of corner j, edge child of edge j, face child, edge
            // child of edge j - 1; quads keep the parent's orientation
            const int rot = (n == 4) ? j : 0;
            SRow *out = &rows[d][first[d][cf]];
            out[(rot + 0) % 4] = Rj;
            blend(Rj, 0.5, Rn, 0.5, out[(rot + 1) % 4]);
            corners.clear();
            for (int i = 0; i < n; ++i) corners.push_back(&Pr[i]);
            mean(corners, out[(rot + 2) % 4]);
            blend(Rp, 0.5, Rj, 0.5, out[(rot + 3) % 4]);
        }
    }
    auto rowOf = [&](int d, int f, int k) -> const SRow & { return rows[d][first[d][f] + k]; };

    // values: distinct rows per (level, vertex), numbered level by level
    FVarChannel ch;
    ch.numBase = numValues;
    std::vector<std::map<int, std::vector<int>>> atVertex(maxL + 1);   // vertex -> value ids
    std::vector<SRow> valueRows;
    std::vector<int> valueLevel, valueVertex;
    for (int d = 1; d <= maxL; ++d)
        for (const auto &lf : patchLeaf_) {
            if (lf.first != d) continue;
            const Level &L = r.level(d);
            for (int k = 0; k < L.faceSize(lf.second); ++k) {
                const int v = L.fv[L.fvOff[lf.second] + k];
                const SRow &row = rowOf(d, lf.second, k);
                std::vector<int> &ids = atVertex[d][v];
                bool found = false;
                for (int id : ids)
                    if (valueRows[id] == row) { found = true; break; }
                if (!found) {
                    ids.push_back((int)valueRows.size());
                    valueRows.push_back(row);
                    valueLevel.push_back(d);
                    valueVertex.push_back(v);
                }
            }
        }
    std::vector<int> order(valueRows.size()), rank(valueRows.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return valueLevel[a] != valueLevel[b] ? valueLevel[a] < valueLevel[b]
                                              : valueVertex[a] < valueVertex[b];
    });
    for (size_t i = 0; i < order.size(); ++i) rank[order[i]] = (int)i;
    ch.numRefined = (int)valueRows.size();
    ch.stencils.off.assign(1, 0);
    ch.stencils.idx.clear();
    ch.stencils.w.clear();
    for (int id : order) {
        for (const Entry &e : valueRows[id]) {
            ch.stencils.idx.push_back(e.first);
            ch.stencils.w.push_back(e.second);
        }
        ch.stencils.off.push_back((int64_t)ch.stencils.idx.size());
    }

    // one QUADS array per vertex array, the leaf face's values per patch
    int base = 0;
    for (const PatchArrayRec &a : arrays) {
        ch.arrays.push_back({ kFVarLinear, kFVarLinear, a.numPatches, base, 4, a.primitiveIdBase });
        base += 4 * a.numPatches;
    }
    ch.indices.reserve(4 * patchLeaf_.size());
    for (const auto &lf : patchLeaf_) {
        const Level &L = r.level(lf.first);
        if (L.faceSize(lf.second) != 4) return false;
        for (int k = 0; k < 4; ++k) {
            const int fk = L.fvOff[lf.second] + k;
            if (lf.first == 0) {
                ch.indices.push_back(valueIndices[fk]);
                continue;
            }
            const SRow &row = rowOf(lf.first, lf.second, k);
            for (int id : atVertex[lf.first][L.fv[fk]])
                if (valueRows[id] == row) {
                    ch.indices.push_back(numValues + rank[id]);
                    break;
                }
        }
    }
    ch.params = params;
    fvar.push_back(std::move(ch));
    return true;
}

template <typename REAL>
void PatchTable::fvarStencilTable(int channel, StencilTableT<REAL> &out) const
{
    const FVarChannel &ch = fvar[channel];
    out = StencilTableT<REAL>();
    out.numControlVertices = ch.numBase;
    const int n = (int)ch.stencils.off.size() - 1;
    out.sizes.resize(n);
    out.offsets.resize(n);
    for (int i = 0; i < n; ++i) {
        out.sizes[i] = ch.stencils.size(i);
        out.offsets[i] = ch.stencils.off[i];
    }
    out.indices.assign(ch.stencils.idx.begin(), ch.stencils.idx.end());
    out.weights.assign(ch.stencils.w.begin(), ch.stencils.w.end());
}

template void PatchTable::fvarStencilTable<float>(int, StencilTable &) const;
template void PatchTable::fvarStencilTable<double>(int, StencilTableD &) const;

}  // namespace osdfar
