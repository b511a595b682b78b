// Table files and topology keys (include/osdfar.h).  The reference has no
// on-disk form for its tables: Far rebuilds them from the topology on every
// run, which for a few hundred resident topologies dominates start-up.
//
// File layout (little-endian, every offset from the start of the file):
//   header      64 bytes: magic "OSDFARTB", u32 version, u32 section count,
//               u64 key, u64 file size, u64 hash of the directory, 24 bytes 0
//   directory   64 bytes per section: name[32] (NUL-terminated), i32 dtype,
//               i32 element size, i64 count, i64 offset, u64 payload hash
//   payload     each section at a 256-byte aligned offset
// Hashes are XXH64 (seed 0), the published xxHash 64-bit algorithm.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/osdfar.h"

namespace {

// ---- XXH64 (xxHash specification, 64-bit variant) ----------------------
constexpr uint64_t kP1 = 0x9E3779B185EBCA87ull;
constexpr uint64_t kP2 = 0xC2B2AE3D27D4EB4Full;
constexpr uint64_t kP3 = 0x165667B19E3779F9ull;
constexpr uint64_t kP4 = 0x85EBCA77C2B2AE63ull;
constexpr uint64_t kP5 = 0x27D4EB2F165667C5ull;

inline uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
inline uint64_t rd64(const unsigned char *p)
{
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}
inline uint32_t rd32(const unsigned char *p)
{
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}
inline uint64_t xround(uint64_t acc, uint64_t in)
{
    acc += in * kP2;
    return rotl(acc, 31) * kP1;
}
inline uint64_t xmerge(uint64_t h, uint64_t v)
{
    h ^= xround(0, v);
    return h * kP1 + kP4;
}

uint64_t xxh64(const void *data, size_t n, uint64_t seed)
{
    static const unsigned char kNone[1] = {0};
    const unsigned char *p = n ? (const unsigned char *)data : kNone, *end = p + n;
    uint64_t h;
    if (n >= 32) {
        uint64_t v1 = seed + kP1 + kP2, v2 = seed + kP2, v3 = seed, v4 = seed - kP1;
        const unsigned char *last = end - 32;
        do {
            v1 = xround(v1, rd64(p));
            v2 = xround(v2, rd64(p + 8));
            v3 = xround(v3, rd64(p + 16));
            v4 = xround(v4, rd64(p + 24));
            p += 32;
        } while (p <= last);
        h = rotl(v1, 1) + rotl(v2, 7) + rotl(v3, 12) + rotl(v4, 18);
        h = xmerge(h, v1);
        h = xmerge(h, v2);
        h = xmerge(h, v3);
        h = xmerge(h, v4);
    } else {
        h = seed + kP5;
    }
    h += (uint64_t)n;
    size_t rem = (size_t)(end - p);
    for (; rem >= 8; rem -= 8, p += 8) h = rotl(h ^ xround(0, rd64(p)), 27) * kP1 + kP4;
    if (rem >= 4) {
        h = rotl(h ^ ((uint64_t)rd32(p) * kP1), 23) * kP2 + kP3;
        rem -= 4;
        p += 4;
    }
    for (; rem; --rem, ++p) h = rotl(h ^ ((uint64_t)*p * kP5), 11) * kP1;
    h ^= h >> 33;
    h *= kP2;
    h ^= h >> 29;
    h *= kP3;
    h ^= h >> 32;
    return h;
}

// ---- file layout ------------------------------------------------------
constexpr char kMagic[8] = {'O', 'S', 'D', 'F', 'A', 'R', 'T', 'B'};
constexpr uint32_t kVersion = 1;
constexpr size_t kAlign = 256;
constexpr int kMaxSections = 4096;
constexpr int kNameBytes = 32;

struct Header {
    char magic[8];
    uint32_t version;
    uint32_t numSections;
    uint64_t key;
    uint64_t fileBytes;
    uint64_t dirHash;
    unsigned char pad[24];
};
struct Entry {
    char name[kNameBytes];
    int32_t dtype;
    int32_t elemSize;
    int64_t count;
    int64_t offset;
    uint64_t hash;
};
static_assert(sizeof(Header) == 64 && sizeof(Entry) == 64, "on-disk records are 64 bytes");

int elem_size(int dtype)
{
    switch (dtype) {
    case OSDFAR_I32: return 4;
    case OSDFAR_I64: return 8;
    case OSDFAR_F32: return 4;
    case OSDFAR_F64: return 8;
    case OSDFAR_BYTES: return 1;
    default: return 0;
    }
}

size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

bool write_all(FILE *f, const void *p, size_t n)
{
    return n == 0 || std::fwrite(p, 1, n, f) == n;
}

std::atomic<unsigned> g_tmpCounter{0};

}  // namespace

struct osdfar_table_file {
    void *map = nullptr;
    size_t bytes = 0;
    uint64_t key = 0;
    std::vector<Entry> dir;
};

extern "C" {

unsigned long long osdfar_hash64(const void *data, long long nbytes, unsigned long long seed)
{
    if (nbytes < 0 || (!data && nbytes > 0)) return 0;
    return xxh64(data, (size_t)nbytes, seed);
}

unsigned long long osdfar_topology_key(int numVertices, int numFaces, const int *numVertsPerFace,
                                       const int *vertIndicesPerFace, int numCreases,
                                       const int *creaseVertexIndexPairs,
                                       const float *creaseWeights, int numCorners,
                                       const int *cornerVertexIndices, const float *cornerWeights,
                                       int boundaryInterpolation, const char *recipe)
{
    try {
        if (numVertices < 0 || numFaces < 0 || numCreases < 0 || numCorners < 0) return 0;
        if ((numFaces > 0 && !numVertsPerFace) || (numCreases > 0 && !(creaseVertexIndexPairs &&
                                                                      creaseWeights)) ||
            (numCorners > 0 && !(cornerVertexIndices && cornerWeights)))
            return 0;
        long long nfv = 0;
        for (int f = 0; f < numFaces; ++f) {
            if (numVertsPerFace[f] < 0) return 0;
            nfv += numVertsPerFace[f];
        }
        if (nfv > 0 && !vertIndicesPerFace) return 0;
        // each field is hashed on its own and chained, with its length, so that
        // moving bytes between fields changes the key
        const int counts[5] = {numVertices, numFaces, numCreases, numCorners, boundaryInterpolation};
        uint64_t h = xxh64(counts, sizeof(counts), 0);
        auto chain = [&h](const void *p, size_t n) { h = xxh64(p, n, h ^ (uint64_t)n); };
        chain(numVertsPerFace, (size_t)numFaces * sizeof(int));
        chain(vertIndicesPerFace, (size_t)nfv * sizeof(int));
        chain(creaseVertexIndexPairs, (size_t)numCreases * 2 * sizeof(int));
        chain(creaseWeights, (size_t)numCreases * sizeof(float));
        chain(cornerVertexIndices, (size_t)numCorners * sizeof(int));
        chain(cornerWeights, (size_t)numCorners * sizeof(float));
        const char *r = recipe ? recipe : "";
        chain(r, std::strlen(r));
        return h;
    } catch (...) {
        return 0;
    }
}

int osdfar_table_file_write(const char *path, unsigned long long key, int numSections,
                            const osdfar_section *sections)
{
    try {
        if (!path || numSections < 0 || numSections > kMaxSections || (numSections && !sections))
            return -1;
        std::vector<Entry> dir((size_t)numSections);
        size_t off = align_up(sizeof(Header) + (size_t)numSections * sizeof(Entry));
        for (int i = 0; i < numSections; ++i) {
            const osdfar_section &s = sections[i];
            const int es = elem_size(s.dtype);
            if (!s.name || !es || s.count < 0 || s.count > (1LL << 48) || (s.count > 0 && !s.data))
                return -1;
            const size_t len = std::strlen(s.name);
            if (len == 0 || len >= (size_t)kNameBytes) return -1;
            for (int j = 0; j < i; ++j)
                if (std::strcmp(dir[j].name, s.name) == 0) return -1;   // names are unique
            Entry &e = dir[i];
            std::memset(&e, 0, sizeof(e));
            std::memcpy(e.name, s.name, len);
            e.dtype = s.dtype;
            e.elemSize = es;
            e.count = s.count;
            e.offset = (int64_t)off;
            e.hash = xxh64(s.data, (size_t)s.count * (size_t)es, 0);
            off = align_up(off + (size_t)s.count * (size_t)es);
        }
        Header hd;
        std::memset(&hd, 0, sizeof(hd));
        std::memcpy(hd.magic, kMagic, sizeof(kMagic));
        hd.version = kVersion;
        hd.numSections = (uint32_t)numSections;
        hd.key = key;
        hd.fileBytes = off;
        hd.dirHash = xxh64(dir.data(), dir.size() * sizeof(Entry), 0);

        // write under a name of this process and call, then rename into place:
        // a reader sees either no file or a whole one
        const std::string tmp = std::string(path) + ".tmp." + std::to_string((long)getpid()) + "." +
                                std::to_string(g_tmpCounter.fetch_add(1));
        FILE *f = std::fopen(tmp.c_str(), "wb");
        if (!f) return -2;
        static const unsigned char zeros[kAlign] = {};
        bool ok = write_all(f, &hd, sizeof(hd)) && write_all(f, dir.data(), dir.size() * sizeof(Entry));
        size_t pos = sizeof(hd) + dir.size() * sizeof(Entry);
        for (int i = 0; ok && i < numSections; ++i) {
            ok = write_all(f, zeros, (size_t)dir[i].offset - pos);
            const size_t n = (size_t)dir[i].count * (size_t)dir[i].elemSize;
            ok = ok && write_all(f, sections[i].data, n);
            pos = (size_t)dir[i].offset + n;
        }
        ok = ok && write_all(f, zeros, off - pos);
        ok = (std::fclose(f) == 0) && ok;
        if (!ok || std::rename(tmp.c_str(), path) != 0) {
            std::remove(tmp.c_str());
            return -2;
        }
        return 0;
    } catch (...) {
        return OSDFAR_ERR_HOST;
    }
}

osdfar_table_file *osdfar_table_file_open(const char *path, unsigned long long key, int verify)
{
    try {
        if (!path) return nullptr;
        const int fd = open(path, O_RDONLY);
        if (fd < 0) return nullptr;
        struct stat st;
        if (fstat(fd, &st) != 0 || (size_t)st.st_size < sizeof(Header)) {
            close(fd);
            return nullptr;
        }
        const size_t bytes = (size_t)st.st_size;
        void *map = mmap(nullptr, bytes, PROT_READ, MAP_PRIVATE, fd, 0);
        close(fd);
        if (map == MAP_FAILED) return nullptr;
        auto fail = [&]() -> osdfar_table_file * {
            munmap(map, bytes);
            return nullptr;
        };
        const unsigned char *base = (const unsigned char *)map;
        Header hd;
        std::memcpy(&hd, base, sizeof(hd));
        if (std::memcmp(hd.magic, kMagic, sizeof(kMagic)) != 0 || hd.version != kVersion ||
            hd.key != key || hd.fileBytes != bytes || hd.numSections > (uint32_t)kMaxSections ||
            sizeof(Header) + (size_t)hd.numSections * sizeof(Entry) > bytes)
            return fail();
        std::vector<Entry> dir(hd.numSections);
        if (hd.numSections) std::memcpy(dir.data(), base + sizeof(Header), dir.size() * sizeof(Entry));
        if (xxh64(dir.data(), dir.size() * sizeof(Entry), 0) != hd.dirHash) return fail();
        for (const Entry &e : dir) {
            if (std::memchr(e.name, 0, kNameBytes) == nullptr || e.elemSize != elem_size(e.dtype) ||
                e.count < 0 || e.offset < 0 || (size_t)e.offset % kAlign != 0 ||
                (size_t)e.offset > bytes || (size_t)e.count > (bytes - (size_t)e.offset) / (size_t)e.elemSize)
                return fail();
            if (verify && xxh64(base + e.offset, (size_t)e.count * (size_t)e.elemSize, 0) != e.hash)
                return fail();
        }
        osdfar_table_file *t = new (std::nothrow) osdfar_table_file;
        if (!t) return fail();
        t->map = map;
        t->bytes = bytes;
        t->key = hd.key;
        t->dir = std::move(dir);
        return t;
    } catch (...) {
        return nullptr;
    }
}

int osdfar_table_file_sections(const osdfar_table_file *f)
{
    return f ? (int)f->dir.size() : -1;
}

int osdfar_table_file_section(const osdfar_table_file *f, int i, const char **name, int *dtype,
                              long long *count, const void **data)
{
    if (!f || i < 0 || i >= (int)f->dir.size()) return -1;
    const Entry &e = f->dir[(size_t)i];
    if (name) *name = e.name;
    if (dtype) *dtype = e.dtype;
    if (count) *count = e.count;
    if (data) *data = (const unsigned char *)f->map + e.offset;
    return 0;
}

int osdfar_table_file_find(const osdfar_table_file *f, const char *name, int *dtype,
                           long long *count, const void **data)
{
    if (!f || !name) return -1;
    for (size_t i = 0; i < f->dir.size(); ++i)
        if (std::strncmp(f->dir[i].name, name, kNameBytes) == 0)
            return osdfar_table_file_section(f, (int)i, nullptr, dtype, count, data);
    return -2;
}

void osdfar_table_file_close(osdfar_table_file *f)
{
    if (!f) return;
    if (f->map) munmap(f->map, f->bytes);
    delete f;
}

}  // extern "C"
