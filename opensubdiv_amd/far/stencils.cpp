// Factorized stencils of uniformly refined levels (Catmark, Loop, and the
// bilinear scheme as Catmark topology with linear masks).
//
// Masks restate Sdc's Catmark and Loop rules (sdc/catmarkScheme.h:71-210,
// sdc/loopScheme.h:70-215) with
// uniform or Chaikin creasing and fractional blending (sdc/scheme.h:392-600,
// sdc/crease.cpp:44-198), applied level by level as Far's PrimvarRefiner
// does (far/primvarRefiner.h:475-660): face points, then edge points (which
// use the new face points), then vertex points.  Each new vertex's row is the
// mask-weighted sum of its parents' rows, so every row refers to control
// vertices only (far/stencilTableFactory.cpp:78-151).  Rows are accumulated
// in double with a sparse accumulator and stored sorted by control index.
//
// Rows of one level are computed in dynamically scheduled chunks into
// per-thread buffers; the level's arrays are then assembled by the same
// threads (parallel first touch, no serial copies).  The last level of a
// table is written straight into the table's arrays.
#include "stencils.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace osdfar {

static int g_threads = 0;
static thread_local int t_threads = -1;

void setCompilerThreads(int n) { t_threads = n < 0 ? -1 : n; }

int compilerThreads()
{
    int n = t_threads >= 0 ? t_threads : g_threads;
#ifdef _OPENMP
    if (n <= 0) n = omp_get_max_threads();
#else
    n = 1;
#endif
    return std::max(1, n);
}

namespace {

// Sparse accumulator over control vertices: pos[c] is c's slot in the
// current row or -1.
struct Accum {
    std::vector<int> pos, list;
    std::vector<double> val;
    explicit Accum(int nControl) : pos(nControl, -1)
    {
        list.reserve(64);
        val.reserve(64);
    }

    inline void add(int i, double w)
    {
        const int p = pos[i];
        if (p < 0) {
            pos[i] = (int)list.size();
            list.push_back(i);
            val.push_back(w);
        } else {
            val[p] += w;
        }
    }
    // parent row v (identity at level 0) scaled by w
    inline void addRow(const LevelStencils *S, int v, double w)
    {
        if (!S) { add(v, w); return; }
        const int64_t k1 = S->off[v + 1];
        const int *ix = S->idx.data();
        const double *ww = S->w.data();
        for (int64_t k = S->off[v]; k < k1; ++k) add(ix[k], w * ww[k]);
    }
    // Appends the row sorted by control index and resets the accumulator.
    int flush(std::vector<int> &idx, std::vector<double> &w)
    {
        const int n = (int)list.size();
        int *L = list.data();
        double *V = val.data();
        if (n <= 48) {
            for (int a = 1; a < n; ++a) {              // insertion sort
                const int ki = L[a];
                const double kv = V[a];
                int b = a - 1;
                while (b >= 0 && L[b] > ki) { L[b + 1] = L[b]; V[b + 1] = V[b]; --b; }
                L[b + 1] = ki;
                V[b + 1] = kv;
            }
            for (int a = 0; a < n; ++a) {
                idx.push_back(L[a]);
                w.push_back(V[a]);
                pos[L[a]] = -1;
            }
        } else {
            std::vector<int> ord(n);
            for (int a = 0; a < n; ++a) ord[a] = a;
            std::sort(ord.begin(), ord.end(), [&](int x, int y) { return L[x] < L[y]; });
            for (int a : ord) {
                idx.push_back(L[a]);
                w.push_back(V[a]);
                pos[L[a]] = -1;
            }
        }
        list.clear();
        val.clear();
        return n;
    }
};

// Per-thread row buffers, kept across levels and calls so their pages are
// faulted in once per thread.
struct RowBuffers {
    std::vector<int> idx, sz;
    std::vector<double> w;
};
static thread_local RowBuffers tl_rows;
constexpr size_t kKeepBytes = size_t(256) << 20;

// Destination of parallelRows: begin(n, total) once, then row(i, offset,
// size) for every row and entries(at, idx, w, count) for every chunk.
struct LevelSink {
    LevelStencils &S;
    void begin(int n, int64_t total)
    {
        S.off.resize((size_t)n + 1);
        S.off[n] = total;
        S.idx.resize((size_t)total);
        S.w.resize((size_t)total);
    }
    void row(int i, int64_t off, int) { S.off[i] = off; }
    void entries(int64_t at, const int *ix, const double *w, size_t cnt)
    {
        std::memcpy(S.idx.data() + at, ix, cnt * sizeof(int));
        std::memcpy(S.w.data() + at, w, cnt * sizeof(double));
    }
};

template <typename REAL>
struct TableSink {           // rows [rb, rb + n) of a table, entries from eb
    StencilTableT<REAL> &T;
    int64_t rb, eb;
    void begin(int, int64_t total)
    {
        T.indices.resize((size_t)(eb + total));
        T.weights.resize((size_t)(eb + total));
    }
    void row(int i, int64_t off, int size)
    {
        T.sizes[rb + i] = size;
        T.offsets[rb + i] = eb + off;
    }
    void entries(int64_t at, const int *ix, const double *w, size_t cnt)
    {
        std::memcpy(T.indices.data() + eb + at, ix, cnt * sizeof(int));
        REAL *d = T.weights.data() + eb + at;
        for (size_t k = 0; k < cnt; ++k) d[k] = (REAL)w[k];
    }
};

// Runs row(i, acc) for i in [0, n) over chunks of rows scheduled
// dynamically, then assembles the rows, in order, into the sink (each
// thread copies the chunks it owns).  avgHint: expected entries per row.
template <class F, class Sink>
void parallelRows(int n, int nControl, double avgHint, F row, Sink &sink)
{
    constexpr int kChunk = 2048;
    const int nch = (n + kChunk - 1) / kChunk;
    const int T = std::max(1, std::min(compilerThreads(), nch));
    struct Chunk { int owner; int64_t at, count, sizeAt, dst; };
    std::vector<Chunk> ch((size_t)nch);
    std::vector<const RowBuffers *> owner((size_t)T, nullptr);
    int64_t total = 0;
#pragma omp parallel num_threads(T)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        RowBuffers &B = tl_rows;
        owner[t] = &B;
        B.idx.clear();
        B.w.clear();
        B.sz.clear();
        const size_t est = (size_t)((double)n / T * avgHint * 1.15) + 1024;
        if (B.idx.capacity() < est) { B.idx.reserve(est); B.w.reserve(est); }
        Accum acc(nControl);
#pragma omp for schedule(dynamic, 1)
        for (int c = 0; c < nch; ++c) {
            Chunk &k = ch[c];
            k.owner = t;
            k.at = (int64_t)B.idx.size();
            k.sizeAt = (int64_t)B.sz.size();
            const int hi = std::min(n, (c + 1) * kChunk);
            for (int i = c * kChunk; i < hi; ++i) {
                row(i, acc);
                B.sz.push_back(acc.flush(B.idx, B.w));
            }
            k.count = (int64_t)B.idx.size() - k.at;
        }
#pragma omp single
        {
            for (int c = 0; c < nch; ++c) { ch[c].dst = total; total += ch[c].count; }
            sink.begin(n, total);
        }
        // copy out (any thread may copy any chunk from its owner's buffers)
#pragma omp for schedule(dynamic, 1)
        for (int c = 0; c < nch; ++c) {
            const Chunk &k = ch[c];
            const RowBuffers &O = *owner[k.owner];
            int64_t o = k.dst;
            const int hi = std::min(n, (c + 1) * kChunk);
            for (int i = c * kChunk; i < hi; ++i) {
                const int z = O.sz[(size_t)(k.sizeAt + (i - c * kChunk))];
                sink.row(i, o, z);
                o += z;
            }
            sink.entries(k.dst, O.idx.data() + k.at, O.w.data() + k.at, (size_t)k.count);
        }
#pragma omp barrier
        // keep small buffers for the next level / call; give large ones back
        if (B.idx.capacity() * (sizeof(int) + sizeof(double)) > kKeepBytes) {
            std::vector<int>().swap(B.idx);
            std::vector<double>().swap(B.w);
            std::vector<int>().swap(B.sz);
        }
    }
}

Rule ruleFrom(float vs, int n, const float *es)
{
    if (vs > 0.0f) return RULE_CORNER;
    int sharp = 0;
    for (int i = 0; i < n; ++i) sharp += es[i] > 0.0f;
    return sharp > 2 ? RULE_CORNER : (Rule)(1 << sharp);
}

// Vertex-vertex mask: weight of the vertex, per incident edge (opposite
// vertex) and per incident face (face point).  Fixed capacity: valences
// above kMaxValence use heap storage.
constexpr int kMaxValence = 64;
struct VMask {
    double v;
    double *e, *f;
    double eb[kMaxValence], fb[kMaxValence];
    std::vector<double> eh, fh;
    void init(int n, int nf)
    {
        if (n <= kMaxValence && nf <= kMaxValence) {
            e = eb;
            f = fb;
        } else {
            eh.assign(n, 0.0);
            fh.assign(nf, 0.0);
            e = eh.data();
            f = fh.data();
        }
        v = 0.0;
        std::fill(e, e + n, 0.0);
        std::fill(f, f + nf, 0.0);
    }
};

void smoothVertexMask(int n, int nf, VMask &m)
{
    m.v = (double)(n - 2) / (double)n;
    const double fw = 1.0 / (double)(n * n);
    std::fill(m.e, m.e + n, fw);
    std::fill(m.f, m.f + nf, fw);
}

void creaseVertexMask(int n, int nf, const float *es, VMask &m)
{
    m.v = 0.75;
    std::fill(m.e, m.e + n, 0.0);
    std::fill(m.f, m.f + nf, 0.0);
    int found = 0;
    for (int i = 0; i < n && found < 2; ++i)
        if (es[i] > 0.0f) { m.e[i] = 0.125; ++found; }
}

void cornerVertexMask(int n, int nf, VMask &m)
{
    m.v = 1.0;
    std::fill(m.e, m.e + n, 0.0);
    std::fill(m.f, m.f + nf, 0.0);
}

void ruleMask(Rule r, int n, int nf, const float *es, VMask &m)
{
    if (r == RULE_SMOOTH || r == RULE_DART) smoothVertexMask(n, nf, m);
    else if (r == RULE_CREASE) creaseVertexMask(n, nf, es, m);
    else cornerVertexMask(n, nf, m);
}

}  // namespace

// Rows of a Catmark (or, with the varying mode, bilinear) level l (parent
// rows PS, NULL at l == 1) into the sink; returns the level's row count.
template <class Sink>
static int catmarkLevelRows(const Refiner &r, int mode, int l, const LevelStencils *PS,
                            double avgPrev, Sink &sink)
{
    const int nControl = r.level(0).nv;
    const Level &P = r.level(l - 1);
    LevelStencils FS;
    LevelSink fsink{ FS };

    // face points: average of the face's vertices
    parallelRows(P.nf, nControl, 2.0 * avgPrev + 4.0, [&](int f, Accum &acc) {
        const int n = P.faceSize(f);
        const double w = 1.0 / (double)n;
        for (int k = P.fvOff[f]; k < P.fvOff[f + 1]; ++k) acc.addRow(PS, P.fv[k], w);
    }, fsink);

    // one pass over the level's rows in Far order:
    // [vertex points | face points (copied) | edge points]
    const int nv = P.nv, nf = P.nf;
    auto vertexRow = [&](int v, Accum &acc) {
        if (mode == INTERPOLATE_VARYING) { acc.addRow(PS, v, 1.0); return; }
        const int n = P.valence(v);
        const int nfv = P.vfOff[v + 1] - P.vfOff[v];
        float pesb[kMaxValence], cesb[kMaxValence];
        const int nn = std::min(n, kMaxValence);
        for (int i = 0; i < nn; ++i) pesb[i] = P.esharp[P.ve[P.veOff[v] + i]];
        subdivideSharpnessAroundVertex(r.creasing, nn, pesb, cesb);
        const float pvs = P.vsharp[v], cvs = decrementSharpness(pvs);
        const Rule pr = ruleFrom(pvs, nn, pesb);
        VMask m;
        m.init(n, nfv);
        if (pr == RULE_SMOOTH || pr == RULE_DART) {
            smoothVertexMask(n, nfv, m);
        } else {
            ruleMask(pr, nn, nfv, pesb, m);
            const Rule cr = ruleFrom(cvs, nn, cesb);
            if (cr != pr) {
                // fractional blend (sdc/crease.cpp:62-101): the parent
                // sharpness of the vertex and of every edge that turns smooth
                int cnt = 0;
                double sum = 0.0;
                if (pvs > 0.0f && cvs <= 0.0f) { cnt = 1; sum = pvs; }
                for (int i = 0; i < nn; ++i)
                    if (pesb[i] > 0.0f && cesb[i] <= 0.0f) { sum += pesb[i]; ++cnt; }
                const double pw = cnt ? std::min(1.0, sum / cnt) : 0.0;
                VMask cm;
                cm.init(n, nfv);
                ruleMask(cr, nn, nfv, cesb, cm);
                const double cw = 1.0 - pw;
                m.v = pw * m.v + cw * cm.v;
                for (int i = 0; i < n; ++i) m.e[i] = pw * m.e[i] + cw * cm.e[i];
                for (int i = 0; i < nfv; ++i) m.f[i] = pw * m.f[i] + cw * cm.f[i];
            }
        }
        for (int i = 0; i < nfv; ++i)
            if (m.f[i] != 0.0) acc.addRow(&FS, P.vf[P.vfOff[v] + i], m.f[i]);
        for (int i = 0; i < n; ++i)
            if (m.e[i] != 0.0) acc.addRow(PS, P.otherVertex(P.ve[P.veOff[v] + i], v), m.e[i]);
        acc.addRow(PS, v, m.v);
    };
    // the two child edges of parent edge e are child edges eFromE + 2e + {0,1}
    const Level &C = r.level(l);
    const size_t eFromE = P.fv.size();
    auto edgeRow = [&](int e, Accum &acc) {
        const int a = P.ev[2 * e], b = P.ev[2 * e + 1];
        const int nfe = P.efOff[e + 1] - P.efOff[e];
        const float s = P.esharp[e];
        // the child vertex is a crease when both child edges stay sharp
        // (vtr/refinement.cpp:960-993); uniform creasing: when s > 1
        const bool childCrease = C.esharp[eFromE + 2 * (size_t)e] > 0.0f &&
                                 C.esharp[eFromE + 2 * (size_t)e + 1] > 0.0f;
        if (mode == INTERPOLATE_VARYING || (s > 0.0f && (childCrease || nfe != 2))) {
            // crease (child vertex is a crease) or linear varying rule
            acc.addRow(PS, a, 0.5);
            acc.addRow(PS, b, 0.5);
            return;
        }
        double vw = 0.25, fw = nfe == 2 ? 0.25 : 0.5 / (double)nfe;
        if (s > 0.0f) {
            // parent crease, child smooth or dart: blend by the edge's
            // sharpness, unclamped (sdc/scheme.h:451-470); above 1 only
            // with Chaikin creasing, where one child edge can turn smooth
            const double p = (double)s, c = 1.0 - p;
            vw = p * 0.5 + c * vw;
            fw *= c;
        }
        acc.addRow(PS, a, vw);
        acc.addRow(PS, b, vw);
        for (int k = P.efOff[e]; k < P.efOff[e + 1]; ++k) acc.addRow(&FS, P.ef[k], fw);
    };
    const int nrows = nv + nf + P.ne;
    parallelRows(nrows, nControl, 3.0 * avgPrev + 6.0, [&](int i, Accum &acc) {
        if (i < nv) {
            vertexRow(i, acc);
        } else if (i < nv + nf) {
            const int f = i - nv;           // copy (already sorted)
            for (int64_t k = FS.off[f]; k < FS.off[f + 1]; ++k) acc.add(FS.idx[k], FS.w[k]);
        } else {
            edgeRow(i - nv - nf, acc);
        }
    }, sink);
    return nrows;
}

// Loop vertex-vertex masks (sdc/loopScheme.h:140-215): smooth -- the
// vertex and, per incident edge, its other end, with Loop's weights for n
// incident faces (5/8 and 1/16 at n = 6; otherwise from beta, evaluated in
// double and rounded to float as the reference's float tables are);
// crease 3/4 and 1/8 at the two crease ends; corner the vertex alone.
static void loopSmoothVertexMask(int n, int nf, VMask &m)
{
    float ew = 0.0625f, vw = 0.625f;
    if (nf != 6) {
        const double dn = (double)nf, inv = 1.0 / dn;
        const double c = std::cos(M_PI * 2.0 * inv);
        const double beta = 0.25 * c + 0.375;
        ew = (float)((0.625 - beta * beta) * inv);
        vw = (float)(1.0 - (double)ew * dn);
    }
    m.v = vw;
    std::fill(m.e, m.e + n, (double)ew);
    std::fill(m.f, m.f + nf, 0.0);
}

static void loopRuleMask(Rule r, int n, int nf, const float *es, VMask &m)
{
    if (r == RULE_SMOOTH || r == RULE_DART) loopSmoothVertexMask(n, nf, m);
    else if (r == RULE_CREASE) creaseVertexMask(n, nf, es, m);
    else cornerVertexMask(n, nf, m);
}

// The vertex of triangle f opposite edge e.
static int oppositeVertex(const Level &P, int f, int e)
{
    const int b = P.fvOff[f];
    for (int j = 0; j < 3; ++j)
        if (P.fe[b + j] == e) return P.fv[b + (j + 2) % 3];
    return P.fv[b];
}

// Rows of a Loop level l: [vertex points | edge points] (no face points).
// Edge points (sdc/loopScheme.h:70-130): 3/8 per end and 1/8 per vertex
// opposite the edge in its two triangles, 1/2 per end on a crease;
// fractional sharpness blends as for Catmark (sdc/scheme.h:392-600).
template <class Sink>
static int loopLevelRows(const Refiner &r, int mode, int l, const LevelStencils *PS,
                         double avgPrev, Sink &sink)
{
    const int nControl = r.level(0).nv;
    const Level &P = r.level(l - 1);
    const Level &C = r.level(l);
    const size_t eFromE = 3 * (size_t)P.nf;
    const int nv = P.nv;
    auto vertexRow = [&](int v, Accum &acc) {
        if (mode == INTERPOLATE_VARYING) { acc.addRow(PS, v, 1.0); return; }
        const int n = P.valence(v);
        const int nfv = P.vfOff[v + 1] - P.vfOff[v];
        float pesb[kMaxValence], cesb[kMaxValence];
        const int nn = std::min(n, kMaxValence);
        for (int i = 0; i < nn; ++i) pesb[i] = P.esharp[P.ve[P.veOff[v] + i]];
        subdivideSharpnessAroundVertex(r.creasing, nn, pesb, cesb);
        const float pvs = P.vsharp[v], cvs = decrementSharpness(pvs);
        const Rule pr = ruleFrom(pvs, nn, pesb);
        VMask m;
        m.init(n, nfv);
        loopRuleMask(pr, nn, nfv, pesb, m);
        if (pr != RULE_SMOOTH && pr != RULE_DART) {
            const Rule cr = ruleFrom(cvs, nn, cesb);
            if (cr != pr) {
                int cnt = 0;
                double sum = 0.0;
                if (pvs > 0.0f && cvs <= 0.0f) { cnt = 1; sum = pvs; }
                for (int i = 0; i < nn; ++i)
                    if (pesb[i] > 0.0f && cesb[i] <= 0.0f) { sum += pesb[i]; ++cnt; }
                const double pw = cnt ? std::min(1.0, sum / cnt) : 0.0;
                VMask cm;
                cm.init(n, nfv);
                loopRuleMask(cr, nn, nfv, cesb, cm);
                const double cw = 1.0 - pw;
                m.v = pw * m.v + cw * cm.v;
                for (int i = 0; i < n; ++i) m.e[i] = pw * m.e[i] + cw * cm.e[i];
            }
        }
        for (int i = 0; i < n; ++i)
            if (m.e[i] != 0.0) acc.addRow(PS, P.otherVertex(P.ve[P.veOff[v] + i], v), m.e[i]);
        acc.addRow(PS, v, m.v);
    };
    auto edgeRow = [&](int e, Accum &acc) {
        const int a = P.ev[2 * e], b = P.ev[2 * e + 1];
        const int nfe = P.efOff[e + 1] - P.efOff[e];
        const float s = P.esharp[e];
        const bool childCrease = C.esharp[eFromE + 2 * (size_t)e] > 0.0f &&
                                 C.esharp[eFromE + 2 * (size_t)e + 1] > 0.0f;
        if (mode == INTERPOLATE_VARYING || (s > 0.0f && (childCrease || nfe != 2))) {
            acc.addRow(PS, a, 0.5);
            acc.addRow(PS, b, 0.5);
            return;
        }
        double vw = 0.375, fw = nfe == 2 ? 0.125 : 0.25 / (double)nfe;
        if (s > 0.0f) {
            const double p = (double)s, c = 1.0 - p;
            vw = p * 0.5 + c * vw;
            fw *= c;
        }
        acc.addRow(PS, a, vw);
        acc.addRow(PS, b, vw);
        for (int k = P.efOff[e]; k < P.efOff[e + 1]; ++k)
            acc.addRow(PS, oppositeVertex(P, P.ef[k], e), fw);
    };
    const int nrows = nv + P.ne;
    parallelRows(nrows, nControl, 3.0 * avgPrev + 6.0, [&](int i, Accum &acc) {
        if (i < nv) vertexRow(i, acc);
        else edgeRow(i - nv, acc);
    }, sink);
    return nrows;
}

// The level's rows in the refiner's scheme; the bilinear scheme refines as
// Catmark does with the linear (varying) masks.
template <class Sink>
static int levelRows(const Refiner &r, int mode, int l, const LevelStencils *PS, double avgPrev,
                     Sink &sink)
{
    if (r.scheme == SCHEME_LOOP) return loopLevelRows(r, mode, l, PS, avgPrev, sink);
    return catmarkLevelRows(r, r.scheme == SCHEME_BILINEAR ? INTERPOLATE_VARYING : mode, l, PS,
                            avgPrev, sink);
}

void computeLevelStencils(const Refiner &r, int mode, std::vector<LevelStencils> &LS)
{
    LS.assign(r.maxLevel() + 1, LevelStencils());
    double avgPrev = 1.0;                       // entries per row of the parent level
    for (int l = 1; l <= r.maxLevel(); ++l) {
        LevelSink sink{ LS[l] };
        const int nrows = levelRows(r, mode, l, l == 1 ? nullptr : &LS[l - 1], avgPrev, sink);
        avgPrev = nrows ? (double)LS[l].idx.size() / nrows : 1.0;
    }
}

template <typename REAL>
void buildStencilTable(const Refiner &r, const StencilOptions &opt, StencilTableT<REAL> &out)
{
    out = StencilTableT<REAL>();
    out.numControlVertices = r.level(0).nv;
    const int maxL = r.maxLevel();
    const int first = std::max(opt.generateIntermediateLevels ? 1 : maxL, 1);
    const int nc = opt.generateControlVerts ? out.numControlVertices : 0;
    int rows = nc;
    for (int l = first; l <= maxL; ++l) rows += r.level(l).nv;
    out.sizes.resize(rows);
    out.offsets.resize(rows);
    out.indices.resize(nc);
    out.weights.resize(nc);
    for (int v = 0; v < nc; ++v) {
        out.sizes[v] = 1;
        out.offsets[v] = v;
        out.indices[v] = v;
        out.weights[v] = (REAL)1;
    }
    if (maxL < 1) return;
    // levels 1..maxL: kept levels are appended to the table as they are
    // produced; a parent level is dropped once its children exist
    std::vector<LevelStencils> LS(maxL + 1);
    double avgPrev = 1.0;
    int64_t rb = nc;
    for (int l = 1; l <= maxL; ++l) {
        const LevelStencils *PS = l == 1 ? nullptr : &LS[l - 1];
        int nrows;
        if (l == maxL) {
            TableSink<REAL> sink{ out, rb, (int64_t)out.indices.size() };
            nrows = levelRows(r, opt.interpolationMode, l, PS, avgPrev, sink);
        } else {
            LevelSink sink{ LS[l] };
            nrows = levelRows(r, opt.interpolationMode, l, PS, avgPrev, sink);
            avgPrev = nrows ? (double)LS[l].idx.size() / nrows : 1.0;
            if (l >= first) {
                TableSink<REAL> ts{ out, rb, (int64_t)out.indices.size() };
                const LevelStencils &S = LS[l];
                ts.begin(nrows, (int64_t)S.idx.size());
                const int T = compilerThreads();
#pragma omp parallel for num_threads(T) schedule(static)
                for (int v = 0; v < nrows; ++v) ts.row(v, S.off[v], S.size(v));
                ts.entries(0, S.idx.data(), S.w.data(), S.idx.size());
            }
        }
        if (l >= first) rb += nrows;                    // rows of kept levels only
        if (l >= 2) LS[l - 1] = LevelStencils();        // parent no longer needed
    }
}

template <typename REAL>
void stencilTableFromLevels(const Refiner &r, const std::vector<LevelStencils> &LS,
                            const StencilOptions &opt, StencilTableT<REAL> &out)
{
    out = StencilTableT<REAL>();
    out.numControlVertices = r.level(0).nv;
    const int maxL = r.maxLevel();
    const int first = std::max(opt.generateIntermediateLevels ? 1 : maxL, 1);
    const int nc = opt.generateControlVerts ? out.numControlVertices : 0;
    int rows = nc;
    for (int l = first; l <= maxL; ++l) rows += r.level(l).nv;
    out.sizes.resize(rows);
    out.offsets.resize(rows);
    out.indices.resize(nc);
    out.weights.resize(nc);
    for (int v = 0; v < nc; ++v) {
        out.sizes[v] = 1;
        out.offsets[v] = v;
        out.indices[v] = v;
        out.weights[v] = (REAL)1;
    }
    int64_t rb = nc;
    const int T = compilerThreads();
    for (int l = first; l <= maxL; ++l) {
        const LevelStencils &S = LS[l];
        const int nrows = r.level(l).nv;
        TableSink<REAL> ts{ out, rb, (int64_t)out.indices.size() };
        ts.begin(nrows, (int64_t)S.idx.size());
#pragma omp parallel num_threads(T)
        {
#pragma omp for schedule(static)
            for (int v = 0; v < nrows; ++v) ts.row(v, S.off[v], S.size(v));
            // entries in parallel contiguous chunks
            int t = 0, nt = 1;
#ifdef _OPENMP
            t = omp_get_thread_num();
            nt = omp_get_num_threads();
#endif
            const size_t n = S.idx.size();
            const size_t lo = n * t / nt, hi = n * (t + 1) / nt;
            ts.entries((int64_t)lo, S.idx.data() + lo, S.w.data() + lo, hi - lo);
        }
        rb += nrows;
    }
}

template void stencilTableFromLevels<float>(const Refiner &, const std::vector<LevelStencils> &,
                                            const StencilOptions &, StencilTable &);
template void stencilTableFromLevels<double>(const Refiner &, const std::vector<LevelStencils> &,
                                             const StencilOptions &, StencilTableD &);
template void buildStencilTable<float>(const Refiner &, const StencilOptions &, StencilTable &);
template void buildStencilTable<double>(const Refiner &, const StencilOptions &, StencilTableD &);

}  // namespace osdfar
