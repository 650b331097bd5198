// Host-side input formats and seeded synthetic inputs (no GPU needed).
//
// Format semantics follow the reference loaders:
//   graph file:   src/utils/parser.cpp:296-406 (ReadGraphMarket_bigdata)
//   update file:  include/framework/Loader.h:64-131 (ReadWeightList / ReadList)
//   stream sizes: include/framework/Loader.h:27-62 (Loader ctor)
// The parsers here are single-pass byte scanners over the whole file instead of
// std::stringstream per line; a line that does not carry the expected integers
// is skipped.
#include "sgraph_io.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sg_rng.h"

namespace {

bool read_file(const char *path, std::string &out) {
    if (!path) return false;
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) { std::fclose(f); return false; }
    out.resize((size_t)n);
    size_t got = n ? std::fread(&out[0], 1, (size_t)n, f) : 0;
    std::fclose(f);
    return got == (size_t)n;
}

// Cursor over one text line.
struct Line {
    const char *p, *e;
    void skip_ws() { while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p; }
    bool u64(uint64_t &v) {
        skip_ws();
        if (p >= e || *p < '0' || *p > '9') return false;
        uint64_t x = 0;
        while (p < e && *p >= '0' && *p <= '9') { x = x * 10 + (uint64_t)(*p - '0'); ++p; }
        v = x;
        return true;
    }
    bool u32(uint32_t &v) {
        uint64_t x;
        if (!u64(x) || x > 0xFFFFFFFFull) return false;
        v = (uint32_t)x;
        return true;
    }
    bool word(const char *&s, size_t &n) {
        skip_ws();
        s = p;
        while (p < e && *p != ' ' && *p != '\t' && *p != '\r') ++p;
        n = (size_t)(p - s);
        return n > 0;
    }
};

template <class F>
void for_each_line(const std::string &buf, F &&f) {
    const char *p = buf.data(), *end = p + buf.size();
    while (p < end) {
        const char *nl = (const char *)std::memchr(p, '\n', (size_t)(end - p));
        const char *le = nl ? nl : end;
        Line l{p, le};
        f(l);
        p = nl ? nl + 1 : end;
    }
}

template <class T>
T *dup_array(const std::vector<T> &v) {
    T *p = (T *)std::malloc(std::max<size_t>(1, v.size()) * sizeof(T));
    if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
}

unsigned n_threads() {
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0) t = 1;
    return std::min(t, 16u);
}

// Sort 64-bit (src << 32 | dst) keys: bucket by the top bits of src, then
// std::sort each bucket on a pool of threads.
void parallel_sort_keys(std::vector<uint64_t> &keys, uint32_t nv_bits) {
    const size_t n = keys.size();
    if (n < (1u << 16)) { std::sort(keys.begin(), keys.end()); return; }
    const uint32_t bbits = std::min<uint32_t>(nv_bits, 10);
    const uint32_t nb = 1u << bbits;
    const uint32_t shift = 32 + nv_bits - bbits;
    std::vector<uint64_t> cnt(nb + 1, 0);
    for (uint64_t k : keys) cnt[(k >> shift) + 1]++;
    for (uint32_t i = 0; i < nb; ++i) cnt[i + 1] += cnt[i];
    std::vector<uint64_t> out(n);
    {
        std::vector<uint64_t> pos(cnt.begin(), cnt.end() - 1);
        for (uint64_t k : keys) out[pos[k >> shift]++] = k;
    }
    std::atomic<uint32_t> next{0};
    auto worker = [&]() {
        for (;;) {
            uint32_t b = next.fetch_add(1);
            if (b >= nb) break;
            std::sort(out.begin() + (ptrdiff_t)cnt[b], out.begin() + (ptrdiff_t)cnt[b + 1]);
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 0; t < n_threads(); ++t) th.emplace_back(worker);
    for (auto &t : th) t.join();
    keys.swap(out);
}

// R-MAT thresholds on a 32-bit uniform (the draw itself is sg_rmat_draw).
struct RmatThresh : SgRmatThresh {
    RmatThresh(double A, double B, double C) {
        auto q = [](double x) { return (uint64_t)std::llround(std::min(1.0, std::max(0.0, x)) * 4294967296.0); };
        a = q(A); ab = q(A + B); abc = q(A + B + C);
    }
};

inline void rmat_draw(SgRng &rng, uint32_t scale, const RmatThresh &t, uint32_t &src, uint32_t &dst) {
    sg_rmat_draw(rng, scale, t, src, dst);
}

// Open-addressing set/map of 64-bit keys -> 32-bit values (value UINT32_MAX = empty slot marker not used).
struct KeyMap {
    std::vector<uint64_t> k;
    std::vector<uint32_t> v;
    std::vector<uint8_t> used;
    size_t mask = 0, count = 0;
    static constexpr uint32_t DEAD = 0xFFFFFFFFu;
    void init(size_t expect) {
        size_t cap = 16;
        while (cap < expect * 2 + 16) cap <<= 1;
        k.assign(cap, 0); v.assign(cap, DEAD); used.assign(cap, 0);
        mask = cap - 1; count = 0;
    }
    size_t slot(uint64_t key) const {
        size_t h = (size_t)(sg_splitmix64(key) & mask);
        while (used[h] && k[h] != key) h = (h + 1) & mask;
        return h;
    }
    // value or DEAD
    uint32_t get(uint64_t key) const {
        size_t h = slot(key);
        return used[h] ? v[h] : DEAD;
    }
    void put(uint64_t key, uint32_t val) {
        if ((count + 1) * 2 > k.size()) grow();
        size_t h = slot(key);
        if (!used[h]) { used[h] = 1; k[h] = key; ++count; }
        v[h] = val;
    }
    void grow() {
        std::vector<uint64_t> ok; ok.swap(k);
        std::vector<uint32_t> ov; ov.swap(v);
        std::vector<uint8_t> ou; ou.swap(used);
        size_t cap = ok.size() * 2;
        k.assign(cap, 0); v.assign(cap, DEAD); used.assign(cap, 0);
        mask = cap - 1; count = 0;
        for (size_t i = 0; i < ok.size(); ++i)
            if (ou[i]) { size_t h = slot(ok[i]); used[h] = 1; k[h] = ok[i]; v[h] = ov[i]; ++count; }
    }
};

// Snapshot file: magic, then u32 nv, u32 nsnap, u64 ne, u64 total inserts,
// u64 total deletes, then row[nv + 1] (u64), col[ne] (u32), sizes[2 nsnap]
// (u64), inserts and deletes as (u, v) u32 pairs; all little-endian.
constexpr char kSnapMagic[8] = {'S', 'G', 'S', 'N', 'A', 'P', '0', '1'};

struct FileCloser {
    FILE *f;
    ~FileCloser() {
        if (f) std::fclose(f);
    }
};

bool write_all(FILE *f, const void *p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes; }
bool read_all(FILE *f, void *p, size_t bytes) { return bytes == 0 || std::fread(p, 1, bytes, f) == bytes; }

}  // namespace

extern "C" {

void sg_free(void *p) { std::free(p); }

sg_status sg_write_snapshot_file(const char *path, uint32_t nv, uint64_t ne, const uint64_t *row, const uint32_t *col,
                                 uint32_t nsnap, const uint64_t *sizes, const uint32_t *ins_uv,
                                 const uint32_t *del_uv) {
    if (!path || !row || (ne && !col) || (nsnap && !sizes)) return SG_EINVAL;
    if (row[0] != 0 || row[nv] != ne) return SG_EINVAL;
    uint64_t ti = 0, td = 0;
    for (uint32_t i = 0; i < nsnap; ++i) {
        ti += sizes[2 * i];
        td += sizes[2 * i + 1];
    }
    if ((ti && !ins_uv) || (td && !del_uv)) return SG_EINVAL;
    FileCloser fc{std::fopen(path, "wb")};
    if (!fc.f) return SG_EIO;
    const uint32_t h32[2] = {nv, nsnap};
    const uint64_t h64[3] = {ne, ti, td};
    const bool ok = write_all(fc.f, kSnapMagic, sizeof kSnapMagic) && write_all(fc.f, h32, sizeof h32) &&
                    write_all(fc.f, h64, sizeof h64) && write_all(fc.f, row, ((size_t)nv + 1) * 8) &&
                    write_all(fc.f, col, (size_t)ne * 4) && write_all(fc.f, sizes, (size_t)nsnap * 16) &&
                    write_all(fc.f, ins_uv, (size_t)ti * 8) && write_all(fc.f, del_uv, (size_t)td * 8);
    if (!ok) return SG_EIO;
    const int rc = std::fclose(fc.f);
    fc.f = nullptr;
    return rc == 0 ? SG_OK : SG_EIO;
}

sg_status sg_read_snapshot_file(const char *path, uint32_t *nv_out, uint64_t *ne_out, uint64_t **row_out,
                                uint32_t **col_out, uint32_t *nsnap_out, uint64_t **sizes_out, uint32_t **ins_out,
                                uint32_t **del_out) {
    if (!path || !nv_out || !ne_out || !row_out || !col_out || !nsnap_out || !sizes_out || !ins_out || !del_out)
        return SG_EINVAL;
    FileCloser fc{std::fopen(path, "rb")};
    if (!fc.f) return SG_EIO;
    char magic[8];
    uint32_t h32[2];
    uint64_t h64[3];
    if (!read_all(fc.f, magic, 8) || std::memcmp(magic, kSnapMagic, 8) != 0 || !read_all(fc.f, h32, sizeof h32) ||
        !read_all(fc.f, h64, sizeof h64))
        return SG_EIO;
    const uint32_t nv = h32[0], nsnap = h32[1];
    const uint64_t ne = h64[0], ti = h64[1], td = h64[2];
    // the header must agree with the file's length before anything is allocated;
    // every count is bounded by the length first (by division), so no product
    // or sum below can wrap
    if (std::fseek(fc.f, 0, SEEK_END) != 0) return SG_EIO;
    const long long slen = std::ftell(fc.f);
    if (slen < 0) return SG_EIO;
    const uint64_t len = (uint64_t)slen;
    if (ne > len / 4 || ti > len / 8 || td > len / 8 || (uint64_t)nsnap > len / 16 || (uint64_t)nv + 1 > len / 8)
        return SG_EIO;
    const uint64_t want = 8 + sizeof h32 + sizeof h64 + ((uint64_t)nv + 1) * 8 + ne * 4 + (uint64_t)nsnap * 16 +
                          ti * 8 + td * 8;  // each term <= len: the sum fits in 64 bits
    if (len != want || std::fseek(fc.f, 8 + sizeof h32 + sizeof h64, SEEK_SET) != 0) return SG_EIO;
    uint64_t *row = (uint64_t *)std::malloc(((size_t)nv + 1) * 8);
    uint32_t *col = (uint32_t *)std::malloc(std::max<size_t>((size_t)ne * 4, 4));
    uint64_t *sizes = (uint64_t *)std::malloc(std::max<size_t>((size_t)nsnap * 16, 16));
    uint32_t *ins = (uint32_t *)std::malloc(std::max<size_t>((size_t)ti * 8, 8));
    uint32_t *del = (uint32_t *)std::malloc(std::max<size_t>((size_t)td * 8, 8));
    auto drop = [&](sg_status st) {
        std::free(row); std::free(col); std::free(sizes); std::free(ins); std::free(del);
        return st;
    };
    if (!row || !col || !sizes || !ins || !del) return drop(SG_ENOMEM);
    if (!read_all(fc.f, row, ((size_t)nv + 1) * 8) || !read_all(fc.f, col, (size_t)ne * 4) ||
        !read_all(fc.f, sizes, (size_t)nsnap * 16) || !read_all(fc.f, ins, (size_t)ti * 8) ||
        !read_all(fc.f, del, (size_t)td * 8))
        return drop(SG_EIO);
    // structure: offsets from 0 to ne, non-decreasing; stream sizes that sum to the totals
    if (row[0] != 0 || row[nv] != ne) return drop(SG_EIO);
    for (uint32_t v = 0; v < nv; ++v)
        if (row[v + 1] < row[v]) return drop(SG_EIO);
    uint64_t si = 0, sd = 0;
    for (uint32_t i = 0; i < nsnap; ++i) {  // a size beyond the total (or a wrapping sum) is refused
        if (sizes[2 * i] > ti - si || sizes[2 * i + 1] > td - sd) return drop(SG_EIO);
        si += sizes[2 * i];
        sd += sizes[2 * i + 1];
    }
    if (si != ti || sd != td) return drop(SG_EIO);
    *nv_out = nv;
    *ne_out = ne;
    *row_out = row;
    *col_out = col;
    *nsnap_out = nsnap;
    *sizes_out = sizes;
    *ins_out = ins;
    *del_out = del;
    return SG_OK;
}

uint32_t sg_edge_weight(uint32_t mode, uint64_t seed, uint32_t u, uint32_t v) {
    return sg_weight(mode, seed, u, v);
}

sg_status sg_read_edge_list(const char *path, uint32_t *nv_out, uint64_t *ne_out,
                            uint64_t **row_out, uint32_t **col_out) {
    if (!nv_out || !ne_out || !row_out || !col_out) return SG_EINVAL;
    std::string buf;
    if (!read_file(path, buf)) return SG_EIO;
    std::vector<uint32_t> us, vs;
    uint64_t max_id = 0;
    bool any = false;
    for_each_line(buf, [&](Line &l) {
        uint32_t u, v;
        if (!l.u32(u) || !l.u32(v)) return;
        us.push_back(u); vs.push_back(v);
        max_id = std::max<uint64_t>(max_id, std::max(u, v));
        any = true;
    });
    if (max_id >= 0xFFFFFFFFull) return SG_EINVAL;
    const uint32_t nv = any ? (uint32_t)(max_id + 1) : 0;
    const uint64_t ne = us.size();
    std::vector<uint64_t> row((size_t)nv + 1, 0);
    for (uint32_t u : us) row[(size_t)u + 1]++;
    for (uint32_t i = 0; i < nv; ++i) row[i + 1] += row[i];
    std::vector<uint64_t> pos(row.begin(), row.end() - 1);
    std::vector<uint32_t> col(ne);
    for (uint64_t i = 0; i < ne; ++i) col[pos[us[i]]++] = vs[i];  // file order per source
    uint64_t *r = dup_array(row);
    uint32_t *c = dup_array(col);
    if (!r || !c) { std::free(r); std::free(c); return SG_ENOMEM; }
    *nv_out = nv; *ne_out = ne; *row_out = r; *col_out = c;
    return SG_OK;
}

sg_status sg_write_edge_list(const char *path, uint32_t nv, const uint64_t *row, const uint32_t *col) {
    if (!path || !row || (!col && row[nv] > 0)) return SG_EINVAL;
    FILE *f = std::fopen(path, "wb");
    if (!f) return SG_EIO;
    std::vector<char> out;
    out.reserve(1 << 20);
    char tmp[32];
    for (uint32_t u = 0; u < nv; ++u) {
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            int n = std::snprintf(tmp, sizeof tmp, "%u %u\n", u, col[e]);
            out.insert(out.end(), tmp, tmp + n);
        }
        if (out.size() > (1u << 20)) { std::fwrite(out.data(), 1, out.size(), f); out.clear(); }
    }
    std::fwrite(out.data(), 1, out.size(), f);
    return std::fclose(f) == 0 ? SG_OK : SG_EIO;
}

sg_status sg_read_stream_sizes(const char *path, uint32_t *nsnap, uint64_t **sizes) {
    if (!nsnap || !sizes) return SG_EINVAL;
    std::string buf;
    if (!read_file(path, buf)) return SG_EIO;
    std::vector<uint64_t> s;
    for_each_line(buf, [&](Line &l) {
        uint64_t a, d;
        if (!l.u64(a) || !l.u64(d)) return;  // trailing text after the two counts is ignored
        s.push_back(a); s.push_back(d);
    });
    uint64_t *p = dup_array(s);
    if (!p) return SG_ENOMEM;
    *nsnap = (uint32_t)(s.size() / 2);
    *sizes = p;
    return SG_OK;
}

sg_status sg_read_updates(const char *path, uint64_t *nins, uint32_t **ins_uv,
                          uint64_t *ndel, uint32_t **del_uv) {
    if (!nins || !ins_uv || !ndel || !del_uv) return SG_EINVAL;
    std::string buf;
    if (!read_file(path, buf)) return SG_EIO;
    std::vector<uint32_t> ins, del;
    for_each_line(buf, [&](Line &l) {
        const char *w; size_t n;
        if (!l.word(w, n) || n != 1) return;  // tag must be exactly "a" or "d"
        uint32_t u, v;
        if (!l.u32(u) || !l.u32(v)) return;
        if (w[0] == 'a') { ins.push_back(u); ins.push_back(v); }
        else if (w[0] == 'd') { del.push_back(u); del.push_back(v); }
    });
    uint32_t *pi = dup_array(ins), *pd = dup_array(del);
    if (!pi || !pd) { std::free(pi); std::free(pd); return SG_ENOMEM; }
    *nins = ins.size() / 2; *ins_uv = pi;
    *ndel = del.size() / 2; *del_uv = pd;
    return SG_OK;
}

sg_status sg_write_updates(const char *update_path, const char *size_path, uint32_t nsnap,
                           const uint64_t *sizes, const uint32_t *ins_uv, const uint32_t *del_uv) {
    if (!update_path || !size_path || (nsnap && !sizes)) return SG_EINVAL;
    FILE *fu = std::fopen(update_path, "wb");
    if (!fu) return SG_EIO;
    FILE *fs = std::fopen(size_path, "wb");
    if (!fs) { std::fclose(fu); return SG_EIO; }
    uint64_t io = 0, dof = 0;
    for (uint32_t s = 0; s < nsnap; ++s) {
        std::fprintf(fs, "%llu %llu\n", (unsigned long long)sizes[2 * s], (unsigned long long)sizes[2 * s + 1]);
        for (uint64_t i = 0; i < sizes[2 * s]; ++i, ++io)
            std::fprintf(fu, "a %u %u\n", ins_uv[2 * io], ins_uv[2 * io + 1]);
        for (uint64_t i = 0; i < sizes[2 * s + 1]; ++i, ++dof)
            std::fprintf(fu, "d %u %u\n", del_uv[2 * dof], del_uv[2 * dof + 1]);
    }
    int a = std::fclose(fu), b = std::fclose(fs);
    return (a == 0 && b == 0) ? SG_OK : SG_EIO;
}

sg_status sg_gen_rmat(uint32_t scale, uint32_t edge_factor, uint64_t seed,
                      double a, double b, double c, int permute,
                      uint32_t *nv_out, uint64_t *ne_out, uint64_t **row_out, uint32_t **col_out) {
    if (!nv_out || !ne_out || !row_out || !col_out) return SG_EINVAL;
    if (scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return SG_EINVAL;
    const uint32_t nv = 1u << scale;
    const uint64_t ndraw = (uint64_t)edge_factor * nv;
    const RmatThresh th(a, b, c);
    std::vector<uint64_t> keys(ndraw);
    const unsigned T = n_threads();
    {
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < T; ++t) {
            pool.emplace_back([&, t]() {
                const uint64_t lo = ndraw * t / T, hi = ndraw * (t + 1) / T;
                for (uint64_t i = lo; i < hi; ++i) {
                    SgRng rng(seed, i);  // stream keyed by (seed, draw index)
                    uint32_t s, d;
                    rmat_draw(rng, scale, th, s, d);
                    keys[i] = ((uint64_t)s << 32) | d;
                }
            });
        }
        for (auto &t : pool) t.join();
    }
    if (permute) {
        std::vector<uint32_t> perm(nv);
        for (uint32_t i = 0; i < nv; ++i) perm[i] = i;
        SgRng rng(seed ^ 0x5045524d55544531ull, 0);
        for (uint32_t i = nv - 1; i > 0; --i) {
            uint32_t j = (uint32_t)(rng.next() % ((uint64_t)i + 1));
            std::swap(perm[i], perm[j]);
        }
        for (auto &k : keys) k = ((uint64_t)perm[k >> 32] << 32) | perm[k & 0xFFFFFFFFu];
    }
    // drop self-loops, sort, drop duplicates
    keys.erase(std::remove_if(keys.begin(), keys.end(),
                              [](uint64_t k) { return (k >> 32) == (k & 0xFFFFFFFFu); }),
               keys.end());
    parallel_sort_keys(keys, scale);
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    const uint64_t ne = keys.size();
    uint64_t *row = (uint64_t *)std::malloc(((size_t)nv + 1) * sizeof(uint64_t));
    uint32_t *col = (uint32_t *)std::malloc(std::max<uint64_t>(1, ne) * sizeof(uint32_t));
    if (!row || !col) { std::free(row); std::free(col); return SG_ENOMEM; }
    std::memset(row, 0, ((size_t)nv + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < ne; ++i) {
        row[(keys[i] >> 32) + 1]++;
        col[i] = (uint32_t)(keys[i] & 0xFFFFFFFFu);
    }
    for (uint32_t i = 0; i < nv; ++i) row[i + 1] += row[i];
    *nv_out = nv; *ne_out = ne; *row_out = row; *col_out = col;
    return SG_OK;
}

sg_status sg_gen_updates(uint32_t nv, const uint64_t *row, const uint32_t *col,
                         uint32_t scale, double a, double b, double c,
                         uint32_t nsnap, uint64_t n_ins, uint64_t n_del,
                         double hub_frac, double hub_top, uint64_t seed,
                         uint32_t **ins_out, uint32_t **del_out) {
    if (!row || !ins_out || !del_out || nv == 0 || scale > 31 || (1ull << scale) > nv)
        return SG_EINVAL;
    const uint64_t ne0 = row[nv];
    if (ne0 && !col) return SG_EINVAL;
    // G_0 adjacency sorted per vertex for membership tests; ids index that copy.
    std::vector<uint32_t> scol(col, col + ne0);
    for (uint32_t u = 0; u < nv; ++u) std::sort(scol.begin() + (ptrdiff_t)row[u], scol.begin() + (ptrdiff_t)row[u + 1]);
    std::vector<uint64_t> g0_dead((ne0 + 63) / 64, 0);
    auto g0_find = [&](uint32_t u, uint32_t v) -> int64_t {
        auto b0 = scol.begin() + (ptrdiff_t)row[u], e0 = scol.begin() + (ptrdiff_t)row[u + 1];
        auto it = std::lower_bound(b0, e0, v);
        if (it == e0 || *it != v) return -1;
        // first live copy
        for (; it != e0 && *it == v; ++it) {
            uint64_t idx = (uint64_t)(it - scol.begin());
            if (!((g0_dead[idx >> 6] >> (idx & 63)) & 1)) return (int64_t)idx;
        }
        return -2;  // present but every copy deleted
    };
    // inserted edges: key -> index into ins_keys (live), DEAD if deleted
    std::vector<uint64_t> ins_keys;
    std::vector<uint8_t> ins_dead;
    KeyMap live;
    live.init((size_t)std::min<uint64_t>(nsnap * n_ins, 1ull << 28));
    // hub set: top hub_top fraction by G_0 out-degree
    std::vector<uint32_t> hubs;
    if (hub_frac > 0) {
        uint32_t k = (uint32_t)std::max<double>(1.0, std::floor(hub_top * nv));
        hubs.resize(nv);
        for (uint32_t i = 0; i < nv; ++i) hubs[i] = i;
        std::nth_element(hubs.begin(), hubs.begin() + (k - 1), hubs.end(), [&](uint32_t x, uint32_t y) {
            uint64_t dx = row[x + 1] - row[x], dy = row[y + 1] - row[y];
            return dx != dy ? dx > dy : x < y;
        });
        hubs.resize(k);
        std::sort(hubs.begin(), hubs.end());
    }
    const RmatThresh th(a, b, c);
    SgRng rng(seed ^ 0x5550444154455321ull, 0);
    const uint64_t hub_cut = (uint64_t)std::llround(std::min(1.0, std::max(0.0, hub_frac)) * 4294967296.0);
    std::vector<uint32_t> ins_uv, del_uv;
    ins_uv.reserve((size_t)(2 * nsnap * n_ins));
    del_uv.reserve((size_t)(2 * nsnap * n_del));
    auto exists = [&](uint32_t u, uint32_t v) {
        if (g0_find(u, v) >= 0) return true;
        return live.get(((uint64_t)u << 32) | v) != KeyMap::DEAD;
    };
    for (uint32_t s = 0; s < nsnap; ++s) {
        const uint64_t ins_before = ins_keys.size();  // G_{t-1} inserted edges are [0, ins_before)
        for (uint64_t i = 0; i < n_ins; ++i) {
            uint32_t u, v;
            uint64_t tries = 0;
            for (;;) {
                if (++tries > 1000000) return SG_EINVAL;  // graph saturated
                rmat_draw(rng, scale, th, u, v);
                if (!hubs.empty() && (rng.next() & 0xFFFFFFFFull) < hub_cut)
                    u = hubs[(size_t)(rng.next() % hubs.size())];
                if (u == v || exists(u, v)) continue;
                break;
            }
            live.put(((uint64_t)u << 32) | v, (uint32_t)ins_keys.size());
            ins_keys.push_back(((uint64_t)u << 32) | v);
            ins_dead.push_back(0);
            ins_uv.push_back(u); ins_uv.push_back(v);
        }
        for (uint64_t i = 0; i < n_del; ++i) {
            const uint64_t pool = ne0 + ins_before;
            uint64_t tries = 0;
            for (;;) {
                if (pool == 0 || ++tries > 100000000ull) return SG_EINVAL;  // nothing left to delete
                uint64_t r = rng.next() % pool;
                if (r < ne0) {
                    if ((g0_dead[r >> 6] >> (r & 63)) & 1) continue;
                    uint32_t u = (uint32_t)(std::upper_bound(row, row + nv + 1, r) - row - 1);
                    uint32_t v = scol[r];
                    // remove every live copy of (u, v): all G_0 copies and an inserted one
                    for (int64_t idx; (idx = g0_find(u, v)) >= 0;) g0_dead[(uint64_t)idx >> 6] |= 1ull << ((uint64_t)idx & 63);
                    uint32_t li = live.get(((uint64_t)u << 32) | v);
                    if (li != KeyMap::DEAD) { ins_dead[li] = 1; live.put(((uint64_t)u << 32) | v, KeyMap::DEAD); }
                    del_uv.push_back(u); del_uv.push_back(v);
                } else {
                    uint64_t j = r - ne0;
                    if (ins_dead[j]) continue;
                    uint32_t u = (uint32_t)(ins_keys[j] >> 32), v = (uint32_t)ins_keys[j];
                    ins_dead[j] = 1;
                    live.put(ins_keys[j], KeyMap::DEAD);
                    for (int64_t idx; (idx = g0_find(u, v)) >= 0;) g0_dead[(uint64_t)idx >> 6] |= 1ull << ((uint64_t)idx & 63);
                    del_uv.push_back(u); del_uv.push_back(v);
                }
                break;
            }
        }
    }
    uint32_t *pi = dup_array(ins_uv), *pd = dup_array(del_uv);
    if (!pi || !pd) { std::free(pi); std::free(pd); return SG_ENOMEM; }
    *ins_out = pi; *del_out = pd;
    return SG_OK;
}

}  // extern "C"
