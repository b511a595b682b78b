// Factorized stencils of uniformly refined Catmark levels.
//
// Masks restate Sdc's Catmark rules (sdc/catmarkScheme.h:71-210) with
// uniform creasing and fractional blending (sdc/scheme.h:392-600,
// sdc/crease.cpp:44-100), applied level by level as Far's PrimvarRefiner
// does (far/primvarRefiner.h:475-660): face points, then edge points (which
// use the new face points), then vertex points.  Each new vertex's row is the
// mask-weighted sum of its parents' rows, so every row refers to control
// vertices only (far/stencilTableFactory.cpp:78-151).  Rows are accumulated
// in double with a sparse accumulator and stored sorted by control index.
#include "stencils.h"

#include <algorithm>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace osdfar {

namespace {

struct Accum {
    std::vector<int> pos, list;
    std::vector<double> val;
    std::vector<std::pair<int, double>> tmp;
    explicit Accum(int nControl) : pos(nControl, -1) {}

    void add(int i, double w)
    {
        int p = pos[i];
        if (p < 0) {
            pos[i] = (int)list.size();
            list.push_back(i);
            val.push_back(w);
        } else {
            val[p] += w;
        }
    }
    // parent row v (identity at level 0) scaled by w
    void addRow(const LevelStencils *S, int v, double w)
    {
        if (!S) { add(v, w); return; }
        for (int64_t k = S->off[v]; k < S->off[v + 1]; ++k) add(S->idx[k], w * S->w[k]);
    }
    void flush(std::vector<int> &idx, std::vector<double> &w)
    {
        tmp.resize(list.size());
        for (size_t k = 0; k < list.size(); ++k) tmp[k] = { list[k], val[k] };
        std::sort(tmp.begin(), tmp.end(),
                  [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
                      return a.first < b.first;
                  });
        for (auto &e : tmp) {
            idx.push_back(e.first);
            w.push_back(e.second);
            pos[e.first] = -1;
        }
        list.clear();
        val.clear();
    }
};

// Runs row(i, acc) for i in [0, n) in parallel contiguous chunks and
// gathers the rows, in order, into out.
template <class F>
void parallelRows(int n, int nControl, F row, LevelStencils &out)
{
    int T = 1;
#ifdef _OPENMP
    T = std::max(1, std::min(omp_get_max_threads(), (n + 1023) / 1024));
#endif
    std::vector<std::vector<int>> idx(T), sz(T);
    std::vector<std::vector<double>> w(T);
#pragma omp parallel num_threads(T)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int lo = (int)((int64_t)n * t / T), hi = (int)((int64_t)n * (t + 1) / T);
        Accum acc(nControl);
        sz[t].reserve(hi - lo);
        for (int i = lo; i < hi; ++i) {
            row(i, acc);
            size_t before = idx[t].size();
            acc.flush(idx[t], w[t]);
            sz[t].push_back((int)(idx[t].size() - before));
        }
    }
    out.off.assign(n + 1, 0);
    int r = 0;
    for (int t = 0; t < T; ++t)
        for (int s : sz[t]) { out.off[r + 1] = out.off[r] + s; ++r; }
    out.idx.resize(out.off[n]);
    out.w.resize(out.off[n]);
    int64_t o = 0;
    for (int t = 0; t < T; ++t) {
        std::memcpy(out.idx.data() + o, idx[t].data(), idx[t].size() * sizeof(int));
        std::memcpy(out.w.data() + o, w[t].data(), w[t].size() * sizeof(double));
        o += (int64_t)idx[t].size();
    }
}

void append(LevelStencils &dst, const LevelStencils &src)
{
    int64_t base = dst.off.empty() ? 0 : dst.off.back();
    if (dst.off.empty()) dst.off.push_back(0);
    for (size_t i = 1; i < src.off.size(); ++i) dst.off.push_back(base + src.off[i]);
    dst.idx.insert(dst.idx.end(), src.idx.begin(), src.idx.end());
    dst.w.insert(dst.w.end(), src.w.begin(), src.w.end());
}

Rule ruleFrom(float vs, int n, const float *es)
{
    if (vs > 0.0f) return RULE_CORNER;
    int sharp = 0;
    for (int i = 0; i < n; ++i) sharp += es[i] > 0.0f;
    return sharp > 2 ? RULE_CORNER : (Rule)(1 << sharp);
}

// Vertex-vertex mask weights: vertex, per incident edge (opposite vertex),
// per incident face (face point).  Returns whether face weights are used.
struct VMask {
    double v;
    std::vector<double> e, f;
};

void smoothVertexMask(int n, VMask &m)
{
    m.v = (double)(n - 2) / (double)n;
    double fw = 1.0 / (double)(n * n);
    std::fill(m.e.begin(), m.e.end(), fw);
    std::fill(m.f.begin(), m.f.end(), fw);
}

void creaseVertexMask(int n, const float *es, VMask &m)
{
    m.v = 0.75;
    std::fill(m.e.begin(), m.e.end(), 0.0);
    std::fill(m.f.begin(), m.f.end(), 0.0);
    int found = 0;
    for (int i = 0; i < n && found < 2; ++i)
        if (es[i] > 0.0f) { m.e[i] = 0.125; ++found; }
}

void cornerVertexMask(VMask &m)
{
    m.v = 1.0;
    std::fill(m.e.begin(), m.e.end(), 0.0);
    std::fill(m.f.begin(), m.f.end(), 0.0);
}

void ruleMask(Rule r, int n, const float *es, VMask &m)
{
    if (r == RULE_SMOOTH || r == RULE_DART) smoothVertexMask(n, m);
    else if (r == RULE_CREASE) creaseVertexMask(n, es, m);
    else cornerVertexMask(m);
}

}  // namespace

void computeLevelStencils(const Refiner &r, int mode, std::vector<LevelStencils> &LS)
{
    const int nControl = r.level(0).nv;
    LS.assign(r.maxLevel() + 1, LevelStencils());
    for (int l = 1; l <= r.maxLevel(); ++l) {
        const Level &P = r.level(l - 1);
        const LevelStencils *PS = (l == 1) ? nullptr : &LS[l - 1];
        LevelStencils FS, ES, VS;

        // face points: average of the face's vertices
        parallelRows(P.nf, nControl, [&](int f, Accum &acc) {
            const int n = P.faceSize(f);
            const double w = 1.0 / (double)n;
            for (int k = P.fvOff[f]; k < P.fvOff[f + 1]; ++k) acc.addRow(PS, P.fv[k], w);
        }, FS);

        // edge points
        parallelRows(P.ne, nControl, [&](int e, Accum &acc) {
            const int a = P.ev[2 * e], b = P.ev[2 * e + 1];
            const int nf = P.efOff[e + 1] - P.efOff[e];
            const float s = P.esharp[e];
            if (mode == INTERPOLATE_VARYING || (s > 0.0f && (decrementSharpness(s) > 0.0f || nf != 2))) {
                // crease (child vertex is a crease) or linear varying rule
                acc.addRow(PS, a, 0.5);
                acc.addRow(PS, b, 0.5);
                return;
            }
            double vw = 0.25, fw = nf == 2 ? 0.25 : 0.5 / (double)nf;
            if (s > 0.0f) {
                // parent crease, child smooth: blend by the edge sharpness
                const double p = (double)s, c = 1.0 - p;
                vw = p * 0.5 + c * vw;
                fw *= c;
            }
            acc.addRow(PS, a, vw);
            acc.addRow(PS, b, vw);
            for (int k = P.efOff[e]; k < P.efOff[e + 1]; ++k) acc.addRow(&FS, P.ef[k], fw);
        }, ES);

        // vertex points
        parallelRows(P.nv, nControl, [&](int v, Accum &acc) {
            if (mode == INTERPOLATE_VARYING) { acc.addRow(PS, v, 1.0); return; }
            const int n = P.valence(v);
            const int nf = P.vfOff[v + 1] - P.vfOff[v];
            float pes[64], ces[64];
            const int nn = std::min(n, 64);
            for (int i = 0; i < nn; ++i) {
                pes[i] = P.esharp[P.ve[P.veOff[v] + i]];
                ces[i] = decrementSharpness(pes[i]);
            }
            const float pvs = P.vsharp[v], cvs = decrementSharpness(pvs);
            const Rule pr = ruleFrom(pvs, nn, pes);
            VMask m;
            m.e.assign(n, 0.0);
            m.f.assign(nf, 0.0);
            if (pr == RULE_SMOOTH || pr == RULE_DART) {
                smoothVertexMask(n, m);
            } else {
                ruleMask(pr, nn, pes, m);
                const Rule cr = ruleFrom(cvs, nn, ces);
                if (cr != pr) {
                    // fractional blend (sdc/crease.cpp:62-95, uniform creasing)
                    int cnt = 0;
                    double sum = 0.0;
                    if (pvs > 0.0f && cvs <= 0.0f) { cnt = 1; sum = pvs; }
                    for (int i = 0; i < nn; ++i)
                        if (pes[i] > 0.0f && pes[i] <= 1.0f) { sum += pes[i]; ++cnt; }
                    double pw = cnt ? std::min(1.0, sum / cnt) : 0.0;
                    VMask cm;
                    cm.e.assign(n, 0.0);
                    cm.f.assign(nf, 0.0);
                    ruleMask(cr, nn, ces, cm);
                    const double cw = 1.0 - pw;
                    m.v = pw * m.v + cw * cm.v;
                    for (int i = 0; i < n; ++i) m.e[i] = pw * m.e[i] + cw * cm.e[i];
                    for (int i = 0; i < nf; ++i) m.f[i] = pw * m.f[i] + cw * cm.f[i];
                }
            }
            for (int i = 0; i < nf; ++i)
                if (m.f[i] != 0.0) acc.addRow(&FS, P.vf[P.vfOff[v] + i], m.f[i]);
            for (int i = 0; i < n; ++i)
                if (m.e[i] != 0.0) acc.addRow(PS, P.otherVertex(P.ve[P.veOff[v] + i], v), m.e[i]);
            acc.addRow(PS, v, m.v);
        }, VS);

        LevelStencils &C = LS[l];
        append(C, VS);
        append(C, FS);
        append(C, ES);
    }
}

template <typename REAL>
void buildStencilTable(const Refiner &r, const StencilOptions &opt, StencilTableT<REAL> &out)
{
    std::vector<LevelStencils> LS;
    computeLevelStencils(r, opt.interpolationMode, LS);
    out = StencilTableT<REAL>();
    out.numControlVertices = r.level(0).nv;
    size_t total = 0;
    int rows = 0;
    const int first = opt.generateIntermediateLevels ? 1 : r.maxLevel();
    for (int l = std::max(first, 1); l <= r.maxLevel(); ++l) {
        total += LS[l].idx.size();
        rows += r.level(l).nv;
    }
    if (opt.generateControlVerts) { total += out.numControlVertices; rows += out.numControlVertices; }
    out.sizes.reserve(rows);
    out.offsets.reserve(rows);
    out.indices.reserve(total);
    out.weights.reserve(total);
    int64_t off = 0;
    if (opt.generateControlVerts) {
        for (int v = 0; v < out.numControlVertices; ++v) {
            out.sizes.push_back(1);
            out.offsets.push_back(off++);
            out.indices.push_back(v);
            out.weights.push_back((REAL)1);
        }
    }
    for (int l = std::max(first, 1); l <= r.maxLevel(); ++l) {
        const LevelStencils &S = LS[l];
        for (int v = 0; v < r.level(l).nv; ++v) {
            out.sizes.push_back(S.size(v));
            out.offsets.push_back(off);
            off += S.size(v);
        }
        out.indices.insert(out.indices.end(), S.idx.begin(), S.idx.end());
        for (double w : S.w) out.weights.push_back((REAL)w);
    }
}

template void buildStencilTable<float>(const Refiner &, const StencilOptions &, StencilTable &);
template void buildStencilTable<double>(const Refiner &, const StencilOptions &, StencilTableD &);

}  // namespace osdfar
