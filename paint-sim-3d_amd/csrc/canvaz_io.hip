// canvaz_io.hip - painting files (.canvaz) and their LZ4 framing, host side.
//
// File layout (src/session.rs:248-317): "PSIM", version 1 (u32), w, h (u32),
// tilt_deg, tilt_dir_deg, cure_rate, h_max, s_max (f32), total_deposited
// (f64), all little endian (44 bytes), then pack() of the eight state parts
// of read_state (src/film/mod.rs:1277-1300) in their order.
//
// pack (src/session.rs:59-72): u32 part count; per part u64 byte length,
// u32 chunk count, and per 32 MiB chunk u64 compressed length followed by
// the chunk as lz4_flex::compress_prepend_size writes it: the u32 length of
// the uncompressed chunk, then one LZ4 block.
//
// The LZ4 block format is the published one (lz4 "Block format" description:
// sequences of a token, literal-length extension bytes, literals, a 2-byte
// little-endian match offset and match-length extension bytes; the last
// sequence holds only literals; a block's last 5 bytes are literals and its
// last match starts at least 12 bytes before its end).  The compressor here
// is a greedy single-probe hash matcher of our own; any decoder of the format
// reads its blocks, and the decoder reads any valid block, whoever wrote it.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <cmath>

#include <algorithm>
#include <thread>
#include <vector>

#include "film_internal.h"

namespace {

constexpr uint32_t kMagic = 0x4D495350u;  // "PSIM" little endian
constexpr uint32_t kVersion = 1u;
constexpr uint64_t kChunk = 32ull << 20;
constexpr int kHashLog = 16;
constexpr uint32_t kMinMatch = 4, kLastLiterals = 5, kMfLimit = 12;

inline uint32_t rd32(const uint8_t* p) {
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

uint64_t bound(uint64_t n) { return n + n / 255u + 16u; }

void put_len(std::vector<uint8_t>& o, uint64_t v) {  // the extension bytes of a length >= 15
    while (v >= 255u) {
        o.push_back(255u);
        v -= 255u;
    }
    o.push_back((uint8_t)v);
}

void emit(std::vector<uint8_t>& o, const uint8_t* lit, uint64_t nlit, uint32_t off, uint64_t mlen) {
    const size_t tok = o.size();
    o.push_back(0u);
    uint8_t t = (uint8_t)((nlit >= 15u ? 15u : nlit) << 4);
    if (nlit >= 15u) put_len(o, nlit - 15u);
    o.insert(o.end(), lit, lit + nlit);
    if (mlen > 0u) {
        o.push_back((uint8_t)(off & 0xffu));
        o.push_back((uint8_t)(off >> 8));
        const uint64_t m = mlen - kMinMatch;
        t |= (uint8_t)(m >= 15u ? 15u : m);
        if (m >= 15u) put_len(o, m - 15u);
    }
    o[tok] = t;
}

// One LZ4 block of src[0, n).
void compress_block(const uint8_t* src, uint64_t n, std::vector<uint8_t>& o) {
    o.clear();
    o.reserve(bound(n));
    if (n < kMfLimit + 1u) {
        emit(o, src, n, 0u, 0u);
        return;
    }
    std::vector<uint32_t> table((size_t)1 << kHashLog, 0xffffffffu);
    const uint64_t limit = n - kMfLimit, mlimit = n - kLastLiterals;
    uint64_t ip = 0, anchor = 0;
    while (ip < limit) {
        const uint32_t seq = rd32(src + ip);
        const uint32_t h = (seq * 2654435761u) >> (32 - kHashLog);
        const uint32_t ref = table[h];
        table[h] = (uint32_t)ip;
        if (ref != 0xffffffffu && ip - ref <= 65535u && rd32(src + ref) == seq) {
            uint64_t len = kMinMatch;
            while (ip + len < mlimit && src[ref + len] == src[ip + len]) len++;
            emit(o, src + anchor, ip - anchor, (uint32_t)(ip - ref), len);
            ip += len;
            anchor = ip;
        } else {
            ip++;
        }
    }
    emit(o, src + anchor, n - anchor, 0u, 0u);
}

// Decode one block into dst[0, cap); the decoded length goes to *got.
bool decompress_block(const uint8_t* src, uint64_t n, uint8_t* dst, uint64_t cap, uint64_t* got) {
    uint64_t ip = 0, op = 0;
    while (ip < n) {
        const uint8_t t = src[ip++];
        uint64_t nlit = t >> 4;
        if (nlit == 15u) {
            uint8_t b;
            do {
                if (ip >= n) return false;
                b = src[ip++];
                nlit += b;
            } while (b == 255u);
        }
        if (nlit > n - ip || nlit > cap - op) return false;
        memcpy(dst + op, src + ip, nlit);
        ip += nlit;
        op += nlit;
        if (ip == n) break;  // the last sequence: literals only
        if (n - ip < 2u) return false;
        const uint32_t off = (uint32_t)src[ip] | ((uint32_t)src[ip + 1] << 8);
        ip += 2;
        if (off == 0u || off > op) return false;
        uint64_t m = t & 15u;
        if (m == 15u) {
            uint8_t b;
            do {
                if (ip >= n) return false;
                b = src[ip++];
                m += b;
            } while (b == 255u);
        }
        m += kMinMatch;
        if (m > cap - op) return false;
        for (uint64_t k = 0; k < m; k++) dst[op + k] = dst[op - off + k];  // overlapping copies repeat
        op += m;
    }
    *got = op;
    return true;
}

void put_u32(std::vector<uint8_t>& o, uint32_t v) { o.insert(o.end(), (uint8_t*)&v, (uint8_t*)&v + 4); }
void put_u64(std::vector<uint8_t>& o, uint64_t v) { o.insert(o.end(), (uint8_t*)&v, (uint8_t*)&v + 8); }
void put_f32(std::vector<uint8_t>& o, float v) { o.insert(o.end(), (uint8_t*)&v, (uint8_t*)&v + 4); }

// pack: the chunks of every part are compressed on worker threads.
void pack(const uint8_t* const* parts, const uint64_t* sizes, uint32_t nparts, std::vector<uint8_t>& out) {
    struct Job { const uint8_t* p; uint64_t n; std::vector<uint8_t> z; };
    std::vector<Job> jobs;
    std::vector<uint32_t> nch(nparts);
    for (uint32_t k = 0; k < nparts; k++) {
        nch[k] = (uint32_t)((sizes[k] + kChunk - 1u) / kChunk);
        for (uint32_t c = 0; c < nch[k]; c++)
            jobs.push_back({parts[k] + (uint64_t)c * kChunk, std::min<uint64_t>(kChunk, sizes[k] - (uint64_t)c * kChunk), {}});
    }
    const unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; w++)
        pool.emplace_back([&, w]() {
            std::vector<uint8_t> blk;
            for (size_t j = w; j < jobs.size(); j += nt) {
                compress_block(jobs[j].p, jobs[j].n, blk);
                jobs[j].z.clear();
                put_u32(jobs[j].z, (uint32_t)jobs[j].n);
                jobs[j].z.insert(jobs[j].z.end(), blk.begin(), blk.end());
            }
        });
    for (auto& t : pool) t.join();
    put_u32(out, nparts);
    size_t j = 0;
    for (uint32_t k = 0; k < nparts; k++) {
        put_u64(out, sizes[k]);
        put_u32(out, nch[k]);
        for (uint32_t c = 0; c < nch[k]; c++, j++) {
            put_u64(out, jobs[j].z.size());
            out.insert(out.end(), jobs[j].z.begin(), jobs[j].z.end());
        }
    }
}

struct Reader {
    const uint8_t* p;
    uint64_t n, off = 0;
    bool take(void* dst, uint64_t k) {
        if (k > n - off) return false;
        memcpy(dst, p + off, k);
        off += k;
        return true;
    }
};

// unpack into parts of the expected sizes (nparts of them): the chunk table
// is read in file order, then the chunks are decoded on worker threads; the
// first failing chunk in file order is the one reported.
int unpack(const uint8_t* data, uint64_t n, uint8_t* const* parts, const uint64_t* sizes, uint32_t nparts) {
    Reader r{data, n};
    uint32_t count = 0;
    if (!r.take(&count, 4)) return set_err(CZ_EINVAL, "canvaz: truncated");
    if (count != nparts) return set_err(CZ_EINVAL, "canvaz: %u parts, expected %u", count, nparts);
    struct Job { uint32_t part, chunk; const uint8_t* src; uint64_t n; uint8_t* dst; uint32_t raw; };
    std::vector<Job> jobs;
    for (uint32_t k = 0; k < nparts; k++) {
        uint64_t len = 0;
        uint32_t nch = 0;
        if (!r.take(&len, 8) || !r.take(&nch, 4)) return set_err(CZ_EINVAL, "canvaz: truncated part %u", k);
        if (len != sizes[k]) return set_err(CZ_ESIZE, "canvaz: part %u has %llu bytes, expected %llu", k,
                                            (unsigned long long)len, (unsigned long long)sizes[k]);
        uint64_t op = 0;
        for (uint32_t c = 0; c < nch; c++) {
            uint64_t clen = 0;
            if (!r.take(&clen, 8) || clen < 4u || clen > r.n - r.off) return set_err(CZ_EINVAL, "canvaz: bad chunk");
            uint32_t raw = 0;
            memcpy(&raw, r.p + r.off, 4);
            if (raw > len - op) return set_err(CZ_EINVAL, "canvaz: part %u chunk %u does not decode", k, c);
            jobs.push_back({k, c, r.p + r.off + 4, clen - 4u, parts[k] + op, raw});
            r.off += clen;
            op += raw;
        }
        if (op != len) return set_err(CZ_EINVAL, "canvaz: part %u decodes to %llu bytes", k, (unsigned long long)op);
    }
    std::vector<uint8_t> ok(jobs.size(), 0);
    const unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nt; w++)
        pool.emplace_back([&, w]() {
            for (size_t j = w; j < jobs.size(); j += nt) {
                uint64_t got = 0;
                ok[j] = decompress_block(jobs[j].src, jobs[j].n, jobs[j].dst, jobs[j].raw, &got) && got == jobs[j].raw;
            }
        });
    for (auto& t : pool) t.join();
    for (size_t j = 0; j < jobs.size(); j++)
        if (!ok[j]) return set_err(CZ_EINVAL, "canvaz: part %u chunk %u does not decode", jobs[j].part, jobs[j].chunk);
    return CZ_OK;
}

}  // namespace

extern "C" {

int cz_lz4_compress(const void* src, uint64_t n, void* dst, uint64_t cap, uint64_t* out_len) {
    if ((!src && n) || !dst || !out_len || n > 0xffffffffull) return set_err(CZ_EINVAL, "lz4_compress: bad arguments");
    std::vector<uint8_t> blk;
    compress_block((const uint8_t*)src, n, blk);
    if (blk.size() + 4u > cap) return set_err(CZ_ESIZE, "lz4_compress: output needs %zu bytes", blk.size() + 4u);
    const uint32_t raw = (uint32_t)n;
    memcpy(dst, &raw, 4);
    memcpy((uint8_t*)dst + 4, blk.data(), blk.size());
    *out_len = blk.size() + 4u;
    return CZ_OK;
}

uint64_t cz_lz4_bound(uint64_t n) { return bound(n) + 4u; }

int cz_lz4_decompress(const void* src, uint64_t n, void* dst, uint64_t cap, uint64_t* out_len) {
    if (!src || n < 4u || (!dst && cap) || !out_len) return set_err(CZ_EINVAL, "lz4_decompress: bad arguments");
    uint32_t raw = 0;
    memcpy(&raw, src, 4);
    if (raw > cap) return set_err(CZ_ESIZE, "lz4_decompress: %u bytes do not fit in %llu", raw, (unsigned long long)cap);
    uint64_t got = 0;
    if (!decompress_block((const uint8_t*)src + 4, n - 4u, (uint8_t*)dst, raw, &got) || got != raw)
        return set_err(CZ_EINVAL, "lz4_decompress: corrupt block");
    *out_len = got;
    return CZ_OK;
}

int cz_film_save_canvaz(cz_film* f, const char* path) {
    if (!f || !path) return set_err(CZ_EINVAL, "save_canvaz: null argument");
    if (f->batch) return set_err(CZ_EINVAL, "save_canvaz: the film is stepped by a batch (cz_batch_remove it first)");
    int rc = cz_film_sync(f);
    if (rc) return rc;
    uint64_t sizes[CZ_NPARTS];
    rc = cz_film_state_sizes(f, sizes);
    if (rc) return rc;
    std::vector<std::vector<uint8_t>> bufs(CZ_NPARTS);
    void* ptrs[CZ_NPARTS];
    const uint8_t* cptrs[CZ_NPARTS];
    for (int k = 0; k < CZ_NPARTS; k++) {
        bufs[k].resize(sizes[k]);
        ptrs[k] = bufs[k].data();
        cptrs[k] = bufs[k].data();
    }
    rc = cz_film_read_state(f, ptrs);
    if (rc) return rc;
    // s_max as the last poll left it: the largest finite surface value of the
    // polled frame's pool (poll_stats, src/film/mod.rs:956-962)
    const float s_max = f->s_max;
    std::vector<uint8_t> out;
    put_u32(out, kMagic);
    put_u32(out, kVersion);
    put_u32(out, f->W);
    put_u32(out, f->H);
    put_f32(out, f->s.tilt_deg);
    put_f32(out, f->s.tilt_dir_deg);
    put_f32(out, f->s.cure_rate);
    put_f32(out, f->h_max);
    put_f32(out, s_max);
    put_u64(out, 0);
    memcpy(out.data() + 36, &f->total_deposited, 8);
    pack(cptrs, sizes, CZ_NPARTS, out);
    FILE* fp = fopen(path, "wb");
    if (!fp) return set_err(CZ_EINVAL, "save_canvaz: cannot open %s", path);
    const size_t w = fwrite(out.data(), 1, out.size(), fp);
    const int cl = fclose(fp);
    if (w != out.size() || cl != 0) return set_err(CZ_EINVAL, "save_canvaz: writing %s failed", path);
    return CZ_OK;
}

int cz_film_load_canvaz(cz_film* f, const char* path) {
    if (!f || !path) return set_err(CZ_EINVAL, "load_canvaz: null argument");
    if (f->batch) return set_err(CZ_EINVAL, "load_canvaz: the film is stepped by a batch (cz_batch_remove it first)");
    FILE* fp = fopen(path, "rb");
    if (!fp) return set_err(CZ_EINVAL, "load_canvaz: cannot open %s", path);
    std::vector<uint8_t> data;
    long fsize = -1;
    if (fseek(fp, 0, SEEK_END) == 0) fsize = ftell(fp);
    if (fsize < 0 || fseek(fp, 0, SEEK_SET) != 0) {
        fclose(fp);
        return set_err(CZ_EINVAL, "load_canvaz: cannot read %s", path);
    }
    data.resize((size_t)fsize);
    const size_t got = fsize > 0 ? fread(data.data(), 1, data.size(), fp) : 0;
    fclose(fp);
    if (got != data.size()) return set_err(CZ_EINVAL, "load_canvaz: reading %s failed", path);
    uint32_t magic = 0, version = 0, w = 0, h = 0;
    if (data.size() < 44u) return set_err(CZ_EINVAL, "load_canvaz: not a painting file");
    memcpy(&magic, data.data(), 4);
    memcpy(&version, data.data() + 4, 4);
    memcpy(&w, data.data() + 8, 4);
    memcpy(&h, data.data() + 12, 4);
    if (magic != kMagic) return set_err(CZ_EINVAL, "load_canvaz: not a painting file");
    if (version != kVersion) return set_err(CZ_EINVAL, "load_canvaz: version %u, expected %u", version, kVersion);
    if (w != f->W || h != f->H)
        return set_err(CZ_ESIZE, "load_canvaz: the painting is %ux%u cells, the film %ux%u", w, h, f->W, f->H);
    uint64_t sizes[CZ_NPARTS];
    int rc = cz_film_state_sizes(f, sizes);
    if (rc) return rc;
    std::vector<std::vector<uint8_t>> bufs(CZ_NPARTS);
    uint8_t* ptrs[CZ_NPARTS];
    const void* cptrs[CZ_NPARTS];
    for (int k = 0; k < CZ_NPARTS; k++) {
        bufs[k].resize(sizes[k]);
        ptrs[k] = bufs[k].data();
        cptrs[k] = bufs[k].data();
    }
    rc = unpack(data.data() + 44, data.size() - 44u, ptrs, sizes, CZ_NPARTS);
    if (rc) return rc;
    rc = cz_film_write_state(f, cptrs, sizes);
    if (rc) return rc;
    // open_painting (src/session.rs:296-303)
    memcpy(&f->s.tilt_deg, data.data() + 16, 4);
    memcpy(&f->s.tilt_dir_deg, data.data() + 20, 4);
    memcpy(&f->s.cure_rate, data.data() + 24, 4);
    memcpy(&f->h_max, data.data() + 28, 4);
    memcpy(&f->s_max, data.data() + 32, 4);
    memcpy(&f->total_deposited, data.data() + 36, 8);
    return CZ_OK;
}

}  // extern "C"
