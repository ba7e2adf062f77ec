// checkpoint.hip -- checkpoint files of device-resident state (host code).
//
// Layout (all integers little-endian):
//   bytes 0-7      "CTSMCKP1"
//   bytes 8-11     field count N
//   bytes 12-15    0
//   N entries of   char name[64], uint64 bytes, uint64 offset, uint64 checksum
//                  (FNV-1a, 64-bit, over the field's little-endian 64-bit words,
//                  a last partial word byte by byte)
//   data           each field's bytes at its offset, offsets aligned to 4 KiB
// Data moves through two pinned staging buffers: while the host writes (or
// reads and hashes) one chunk, the copy engine moves the next one, so a
// checkpoint runs at the slower of PCIe and the file system.
// A write goes to "<path>.tmp", is flushed and fsync'ed, and only then
// renamed over <path>: an interrupted write (a kill, a full disk) leaves the
// previous checkpoint at <path> whole.  A read first hashes every requested
// range from the file and compares the checksums; only a file that passes
// is copied into the caller's memory, so a damaged file leaves the
// destination untouched.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <unistd.h>

#include "ctsm_gpu.h"

namespace ctsm {
namespace {

constexpr char kMagic[8] = {'C', 'T', 'S', 'M', 'C', 'K', 'P', '1'};
constexpr size_t kNameLen = 64;
constexpr size_t kEntryBytes = kNameLen + 3 * sizeof(uint64_t);
constexpr uint64_t kAlign = 4096;
constexpr size_t kChunk = 32u << 20;

// Chunks of one field are multiples of 8 bytes except the field's last, so
// word steps and the byte tail fall at the same places on write and read.
uint64_t fnv1a(uint64_t h, const unsigned char *p, size_t n)
{
    const size_t nw = n / 8;
    for (size_t i = 0; i < nw; ++i) {
        uint64_t w;
        memcpy(&w, p + 8 * i, 8);
        h ^= w;
        h *= 0x100000001b3ull;
    }
    for (size_t i = 8 * nw; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull;

uint64_t align_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct Staging {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t init()
    {
        for (int i = 0; i < 2; ++i) {
            hipError_t e = hipHostMalloc(&buf[i], kChunk, hipHostMallocDefault);
            if (e != hipSuccess) return e;
            if ((e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
    ~Staging()
    {
        for (int i = 0; i < 2; ++i) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
    }
};

struct File {
    FILE *f = nullptr;
    ~File() { if (f) fclose(f); }
};

// The temporary file of a write: removed unless the write reached the rename.
struct TmpFile {
    std::string path;
    bool keep = false;
    ~TmpFile() { if (!keep) (void)remove(path.c_str()); }
};

// Declared after the staging buffers, so it is destroyed first: no copy is
// still in flight when the buffers are freed, whichever way a call returns.
struct StreamDrain {
    hipStream_t st;
    ~StreamDrain() { (void)hipStreamSynchronize(st); }
};

bool seek(FILE *f, uint64_t at) { return fseeko(f, (off_t)at, SEEK_SET) == 0; }

bool valid_fields(const ctsm_ckpt_field *fields, int n)
{
    for (int i = 0; i < n; ++i) {
        if (!fields[i].name || strlen(fields[i].name) >= kNameLen) return false;
        if (fields[i].bytes && !fields[i].ptr) return false;
        for (int j = 0; j < i; ++j)
            if (!strcmp(fields[i].name, fields[j].name)) return false;
    }
    return true;
}

}  // namespace

// Returns a CTSM code; *he receives the failing HIP call's error for CTSM_E_HIP.
int ckpt_write(hipStream_t st, const char *path, const ctsm_ckpt_field *fields, int n, hipError_t *he)
{
    *he = hipSuccess;
    if (!path || n < 0 || (n && !fields) || !valid_fields(fields, n)) return CTSM_E_ARG;
    std::vector<uint64_t> off(n), sum(n);
    uint64_t pos = align_up(16 + (uint64_t)n * kEntryBytes);
    for (int i = 0; i < n; ++i) {
        off[i] = pos;
        pos = align_up(pos + fields[i].bytes);
    }
    TmpFile tmp{std::string(path) + ".tmp"};
    File fp;
    fp.f = fopen(tmp.path.c_str(), "wb");
    if (!fp.f) return CTSM_E_IO;
    Staging S;
    StreamDrain drain{st};
    if ((*he = S.init()) != hipSuccess) return CTSM_E_HIP;
    // every chunk of every field, in order; chunk c uses buffer c % 2
    struct Chunk { int field; uint64_t at, len; };
    std::vector<Chunk> chunks;
    for (int i = 0; i < n; ++i)
        for (uint64_t a = 0; a < fields[i].bytes; a += kChunk)
            chunks.push_back({i, a, std::min<uint64_t>(kChunk, fields[i].bytes - a)});
    auto issue = [&](size_t c) -> hipError_t {
        const Chunk &k = chunks[c];
        hipError_t e = hipMemcpyAsync(S.buf[c % 2], (const char *)fields[k.field].ptr + k.at, k.len,
                                      hipMemcpyDefault, st);
        if (e == hipSuccess) e = hipEventRecord(S.ev[c % 2], st);
        return e;
    };
    for (int i = 0; i < n; ++i) sum[i] = kFnvBasis;
    if (!chunks.empty() && (*he = issue(0)) != hipSuccess) return CTSM_E_HIP;
    for (size_t c = 0; c < chunks.size(); ++c) {
        if (c + 1 < chunks.size() && (*he = issue(c + 1)) != hipSuccess) return CTSM_E_HIP;
        if ((*he = hipEventSynchronize(S.ev[c % 2])) != hipSuccess) return CTSM_E_HIP;
        const Chunk &k = chunks[c];
        const unsigned char *p = (const unsigned char *)S.buf[c % 2];
        sum[k.field] = fnv1a(sum[k.field], p, k.len);
        if (!seek(fp.f, off[k.field] + k.at) || fwrite(p, 1, k.len, fp.f) != k.len)
            return CTSM_E_IO;
    }
    // header last, once the checksums are known
    std::vector<unsigned char> hdr(align_up(16 + (uint64_t)n * kEntryBytes), 0);
    memcpy(hdr.data(), kMagic, 8);
    const uint32_t nf = (uint32_t)n;
    memcpy(hdr.data() + 8, &nf, 4);
    for (int i = 0; i < n; ++i) {
        unsigned char *e = hdr.data() + 16 + (size_t)i * kEntryBytes;
        strncpy((char *)e, fields[i].name, kNameLen - 1);
        memcpy(e + kNameLen, &fields[i].bytes, 8);
        memcpy(e + kNameLen + 8, &off[i], 8);
        memcpy(e + kNameLen + 16, &sum[i], 8);
    }
    if (!seek(fp.f, 0) || fwrite(hdr.data(), 1, hdr.size(), fp.f) != hdr.size()) return CTSM_E_IO;
    if (fflush(fp.f) != 0 || fsync(fileno(fp.f)) != 0) return CTSM_E_IO;
    const int closed = fclose(fp.f);
    fp.f = nullptr;
    if (closed != 0 || rename(tmp.path.c_str(), path) != 0) return CTSM_E_IO;
    tmp.keep = true;
    return CTSM_OK;
}

int ckpt_read(hipStream_t st, const char *path, const ctsm_ckpt_field *fields, int n, hipError_t *he)
{
    *he = hipSuccess;
    if (!path || n < 0 || (n && !fields) || !valid_fields(fields, n)) return CTSM_E_ARG;
    File fp;
    fp.f = fopen(path, "rb");
    if (!fp.f) return CTSM_E_IO;
    unsigned char head[16];
    if (fread(head, 1, 16, fp.f) != 16 || memcmp(head, kMagic, 8) != 0) return CTSM_E_IO;
    uint32_t nf;
    memcpy(&nf, head + 8, 4);
    if (nf > (1u << 20)) return CTSM_E_IO;
    std::vector<unsigned char> tab((size_t)nf * kEntryBytes);
    if (nf && fread(tab.data(), 1, tab.size(), fp.f) != tab.size()) return CTSM_E_IO;
    struct Entry { uint64_t bytes, off, sum; };
    std::vector<Entry> want(n);
    for (int i = 0; i < n; ++i) {
        bool found = false;
        for (uint32_t k = 0; k < nf && !found; ++k) {
            const unsigned char *e = tab.data() + (size_t)k * kEntryBytes;
            char name[kNameLen];
            memcpy(name, e, kNameLen);
            name[kNameLen - 1] = 0;
            if (strcmp(name, fields[i].name) != 0) continue;
            memcpy(&want[i].bytes, e + kNameLen, 8);
            memcpy(&want[i].off, e + kNameLen + 8, 8);
            memcpy(&want[i].sum, e + kNameLen + 16, 8);
            found = true;
        }
        if (!found || want[i].bytes != fields[i].bytes) return CTSM_E_IO;
    }
    Staging S;
    StreamDrain drain{st};
    if ((*he = S.init()) != hipSuccess) return CTSM_E_HIP;
    // first pass: hash every requested range from the file, touching nothing
    for (int i = 0; i < n; ++i) {
        uint64_t h = kFnvBasis;
        for (uint64_t a = 0; a < fields[i].bytes; a += kChunk) {
            const uint64_t len = std::min<uint64_t>(kChunk, fields[i].bytes - a);
            if (!seek(fp.f, want[i].off + a) || fread(S.buf[0], 1, len, fp.f) != len) return CTSM_E_IO;
            h = fnv1a(h, (const unsigned char *)S.buf[0], len);
        }
        if (h != want[i].sum) return CTSM_E_IO;
    }
    // second pass: the verified ranges into the caller's memory; a range that
    // reads differently now (the file changed in between) still fails
    bool used[2] = {false, false};
    size_t c = 0;
    bool ok = true;
    for (int i = 0; i < n; ++i) {
        uint64_t h = kFnvBasis;
        for (uint64_t a = 0; a < fields[i].bytes; a += kChunk, ++c) {
            const uint64_t len = std::min<uint64_t>(kChunk, fields[i].bytes - a);
            const int b = (int)(c % 2);
            // the buffer's previous copy must have left before it is refilled
            if (used[b] && (*he = hipEventSynchronize(S.ev[b])) != hipSuccess) return CTSM_E_HIP;
            if (!seek(fp.f, want[i].off + a) || fread(S.buf[b], 1, len, fp.f) != len)
                return CTSM_E_IO;
            h = fnv1a(h, (const unsigned char *)S.buf[b], len);
            if ((*he = hipMemcpyAsync((char *)fields[i].ptr + a, S.buf[b], len, hipMemcpyDefault, st)) != hipSuccess ||
                (*he = hipEventRecord(S.ev[b], st)) != hipSuccess)
                return CTSM_E_HIP;
            used[b] = true;
        }
        if (h != want[i].sum) ok = false;
    }
    if ((*he = hipStreamSynchronize(st)) != hipSuccess) return CTSM_E_HIP;
    return ok ? CTSM_OK : CTSM_E_IO;
}

}  // namespace ctsm
