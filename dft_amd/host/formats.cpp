// formats.cpp -- on-disk formats (include/dft_io.h): field files, patch
// shards and the result document.  Host code only; no GPU call.
//
// The reference keeps every field in memory and prints its results
// (SelfConsistentFieldMolecule.cpp:343-499); MolecularResult (DFTData.h:62-91)
// is what these formats serialise: the scalars and orbital list as JSON, the
// N-vectors as field files in the device layout.
#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "capi_util.hpp"
#include "dft_io.h"

namespace {

// ---- CRC-32, ISO-HDLC polynomial (reflected 0xEDB88320, zlib's), computed
// eight bytes per step with eight 256-entry tables ("slicing-by-8",
// Kounavis & Berry, IEEE Trans. Computers 57 (2008) 1550).
struct Crc32Tables {
    uint32_t t[8][256];
    Crc32Tables() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int b = 0; b < 8; ++b) c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
            t[0][i] = c;
        }
        for (int k = 1; k < 8; ++k)
            for (uint32_t i = 0; i < 256; ++i) t[k][i] = (t[k - 1][i] >> 8) ^ t[0][t[k - 1][i] & 0xFFu];
    }
};

const Crc32Tables& crc_tables() {
    static const Crc32Tables tabs;
    return tabs;
}

uint32_t crc32_update(uint32_t crc, const unsigned char* p, size_t n) {
    const auto& T = crc_tables().t;
    uint32_t c = ~crc;
    while (n >= 8) {
        uint32_t lo, hi;
        std::memcpy(&lo, p, 4);
        std::memcpy(&hi, p + 4, 4);
        lo ^= c;
        c = T[7][lo & 0xFFu] ^ T[6][(lo >> 8) & 0xFFu] ^ T[5][(lo >> 16) & 0xFFu] ^ T[4][lo >> 24] ^
            T[3][hi & 0xFFu] ^ T[2][(hi >> 8) & 0xFFu] ^ T[1][(hi >> 16) & 0xFFu] ^ T[0][hi >> 24];
        p += 8;
        n -= 8;
    }
    while (n--) c = (c >> 8) ^ T[0][(c ^ *p++) & 0xFFu];
    return ~c;
}

constexpr char kFieldMagic[8] = {'D', 'F', 'T', 'F', 'L', 'D', '0', '1'};
constexpr char kPatchMagic[8] = {'D', 'F', 'T', 'P', 'A', 'T', '0', '1'};
constexpr uint32_t kFieldHeader = 256;
constexpr uint32_t kPatchHeader = 128;
constexpr size_t kChunk = size_t(64) << 20;  // bytes per fread/fwrite

void require_little_endian() {
    const uint16_t one = 1;
    unsigned char b;
    std::memcpy(&b, &one, 1);
    if (b != 1) throw std::invalid_argument("dft_io: the formats are little-endian; this host is not");
}

template <class T>
void put(unsigned char* h, size_t off, T v) {
    std::memcpy(h + off, &v, sizeof(T));
}
template <class T>
T get(const unsigned char* h, size_t off) {
    T v;
    std::memcpy(&v, h + off, sizeof(T));
    return v;
}

// product of non-negative factors with an overflow check against 2^63
uint64_t checked_mul(uint64_t a, uint64_t b, const char* what) {
    if (a != 0 && b > (uint64_t)std::numeric_limits<int64_t>::max() / a)
        throw std::invalid_argument(std::string(what) + ": size overflows");
    return a * b;
}

struct File {
    std::FILE* f = nullptr;
    File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
    int close() {
        int rc = f ? std::fclose(f) : 0;
        f = nullptr;
        return rc;
    }
};

// Writes header + payload parts to path through a temporary file in the same
// directory, renamed over path only when every byte is written, so a failed
// write never leaves a file with a valid header under path.
void write_atomically(const char* path, const unsigned char* header, size_t header_bytes,
                      std::initializer_list<std::pair<const void*, uint64_t>> parts, const char* what) {
    const std::string tmp = std::string(path) + ".tmp";
    {
        File out(tmp, "wb");
        if (!out.f) throw std::invalid_argument(std::string(what) + ": cannot open " + tmp + " for writing");
        bool ok = std::fwrite(header, 1, header_bytes, out.f) == header_bytes;
        for (const auto& part : parts) {
            const unsigned char* p = static_cast<const unsigned char*>(part.first);
            for (uint64_t off = 0; ok && off < part.second; off += kChunk) {
                const size_t nb = (size_t)std::min<uint64_t>(kChunk, part.second - off);
                ok = std::fwrite(p + off, 1, nb, out.f) == nb;
            }
        }
        if (out.close() != 0) ok = false;
        if (!ok) {
            std::remove(tmp.c_str());
            throw std::invalid_argument(std::string(what) + ": write to " + tmp + " failed");
        }
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        throw std::invalid_argument(std::string(what) + ": cannot rename " + tmp + " to " + path);
    }
}

// Reads nbytes at the current position into dst, returning their CRC-32.
uint32_t read_payload(std::FILE* f, void* dst, uint64_t nbytes, const char* what) {
    unsigned char* p = static_cast<unsigned char*>(dst);
    uint32_t crc = 0;
    for (uint64_t off = 0; off < nbytes; off += kChunk) {
        const size_t nb = (size_t)std::min<uint64_t>(kChunk, nbytes - off);
        if (std::fread(p + off, 1, nb, f) != nb) throw std::runtime_error(std::string(what) + ": file is truncated");
        crc = crc32_update(crc, p + off, nb);
    }
    return crc;
}

// Read one whole payload section of nbytes from the current position for its
// CRC-32, keeping bytes [first, first + count) of it in dst: a range read is
// verified like a whole read.
uint32_t read_range_verified(std::FILE* f, uint64_t nbytes, uint64_t first, uint64_t count, void* dst,
                             const char* what) {
    std::vector<unsigned char> buf((size_t)std::min<uint64_t>(kChunk, std::max<uint64_t>(nbytes, 1)));
    unsigned char* out = static_cast<unsigned char*>(dst);
    uint32_t crc = 0;
    for (uint64_t off = 0; off < nbytes; off += kChunk) {
        const size_t nb = (size_t)std::min<uint64_t>(kChunk, nbytes - off);
        if (std::fread(buf.data(), 1, nb, f) != nb) throw std::runtime_error(std::string(what) + ": file is truncated");
        crc = crc32_update(crc, buf.data(), nb);
        const uint64_t lo = std::max(off, first), hi = std::min(off + nb, first + count);
        if (lo < hi) std::memcpy(out + (lo - first), buf.data() + (lo - off), (size_t)(hi - lo));
    }
    return crc;
}

uint64_t file_size(std::FILE* f) {
    const long here = std::ftell(f);
    std::fseek(f, 0, SEEK_END);
    const long end = std::ftell(f);
    std::fseek(f, here, SEEK_SET);
    return end < 0 ? 0 : (uint64_t)end;
}

// ---- field files
void field_header_from_bytes(const unsigned char* h, dft_field_header* out, const char* path) {
    if (std::memcmp(h, kFieldMagic, 8) != 0) throw std::invalid_argument(std::string("not a field file: ") + path);
    if (get<uint32_t>(h, 8) != kFieldHeader) throw std::invalid_argument(std::string("field file: unknown header size: ") + path);
    if (get<uint32_t>(h, kFieldHeader - 4) != crc32_update(0, h, kFieldHeader - 4))
        throw std::runtime_error(std::string("field file: header checksum mismatch: ") + path);
    dft_field_header r{};
    r.dtype = (int)get<uint32_t>(h, 12);
    r.grid.nx = get<int64_t>(h, 16);
    r.grid.ny = get<int64_t>(h, 24);
    r.grid.nz = get<int64_t>(h, 32);
    r.ncols = get<int64_t>(h, 40);
    r.grid.xmin = get<double>(h, 48);
    r.grid.xmax = get<double>(h, 56);
    r.grid.ymin = get<double>(h, 64);
    r.grid.ymax = get<double>(h, 72);
    r.grid.zmin = get<double>(h, 80);
    r.grid.zmax = get<double>(h, 88);
    r.grid.dx = get<double>(h, 96);
    r.grid.dy = get<double>(h, 104);
    r.grid.dz = get<double>(h, 112);
    r.data_bytes = get<uint64_t>(h, 120);
    r.crc = get<uint32_t>(h, 128);
    std::memcpy(r.name, h + 132, 64);
    if (r.dtype != DFT_F64 && r.dtype != DFT_F32) throw std::invalid_argument(std::string("field file: unknown dtype: ") + path);
    if (r.grid.nx < 1 || r.grid.ny < 1 || r.grid.nz < 1 || r.ncols < 1 || r.name[63] != '\0')
        throw std::invalid_argument(std::string("field file: corrupt header: ") + path);
    const uint64_t es = r.dtype == DFT_F64 ? 8 : 4;
    const uint64_t want = checked_mul(
        checked_mul(checked_mul(checked_mul((uint64_t)r.grid.nx, (uint64_t)r.grid.ny, "field file"), (uint64_t)r.grid.nz,
                                "field file"),
                    (uint64_t)r.ncols, "field file"),
        es, "field file");
    if (r.data_bytes != want) throw std::invalid_argument(std::string("field file: data size disagrees with its shape: ") + path);
    *out = r;
}

dft_field_header read_field_header(std::FILE* f, const char* path) {
    unsigned char h[kFieldHeader];
    if (std::fread(h, 1, kFieldHeader, f) != kFieldHeader)
        throw std::invalid_argument(std::string("field file: no header: ") + path);
    dft_field_header r;
    field_header_from_bytes(h, &r, path);
    return r;
}

// ---- patch shards
struct PatchHeader {
    dft_patch_shard_header pub{};
    uint64_t in_bytes = 0, tgt_bytes = 0;
};

PatchHeader read_patch_header(std::FILE* f, const char* path) {
    unsigned char h[kPatchHeader];
    if (std::fread(h, 1, kPatchHeader, f) != kPatchHeader)
        throw std::invalid_argument(std::string("patch shard: no header: ") + path);
    if (std::memcmp(h, kPatchMagic, 8) != 0) throw std::invalid_argument(std::string("not a patch shard: ") + path);
    if (get<uint32_t>(h, 8) != kPatchHeader) throw std::invalid_argument(std::string("patch shard: unknown header size: ") + path);
    if (get<uint32_t>(h, kPatchHeader - 4) != crc32_update(0, h, kPatchHeader - 4))
        throw std::runtime_error(std::string("patch shard: header checksum mismatch: ") + path);
    PatchHeader r;
    r.pub.channels = (int)get<uint32_t>(h, 12);
    r.pub.patch = get<int64_t>(h, 16);
    r.pub.count = get<int64_t>(h, 24);
    r.in_bytes = get<uint64_t>(h, 32);
    r.tgt_bytes = get<uint64_t>(h, 40);
    r.pub.crc_in = get<uint32_t>(h, 48);
    r.pub.crc_tgt = get<uint32_t>(h, 52);
    r.pub.target_scale = get<int32_t>(h, 56);
    r.pub.h = get<double>(h, 64);
    if (r.pub.channels != 2 || r.pub.patch < 1 || r.pub.count < 0)
        throw std::invalid_argument(std::string("patch shard: corrupt header: ") + path);
    const uint64_t p3 = checked_mul(checked_mul((uint64_t)r.pub.patch, (uint64_t)r.pub.patch, "patch shard"),
                                    (uint64_t)r.pub.patch, "patch shard");
    const uint64_t tgt = checked_mul(checked_mul(p3, (uint64_t)r.pub.count, "patch shard"), 4, "patch shard");
    if (r.tgt_bytes != tgt || r.in_bytes != checked_mul(tgt, 2, "patch shard"))
        throw std::invalid_argument(std::string("patch shard: data size disagrees with its shape: ") + path);
    return r;
}

// ---- JSON
void jnum(std::string& s, double v) {
    if (!std::isfinite(v)) {
        s += "null";
        return;
    }
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    s += b;
}
void jint(std::string& s, int64_t v) { s += std::to_string(v); }
void jkey(std::string& s, const char* k) {
    s += '"';
    s += k;
    s += "\":";
}
void jstr(std::string& s, const char* v) {
    s += '"';
    s += v;
    s += '"';
}
// enum value as its name, or as the number when it has none
void jenum(std::string& s, int v, std::initializer_list<const char*> names) {
    if (v >= 0 && (size_t)v < names.size())
        jstr(s, names.begin()[v]);
    else
        jint(s, v);
}

template <class F>
int guarded(F&& f) {
    return dftc::guarded_call(std::function<void()>(std::forward<F>(f)));
}

}  // namespace

extern "C" {

uint32_t dft_crc32(uint32_t crc, const void* data, size_t n) {
    if (!data || n == 0) return crc;
    return crc32_update(crc, static_cast<const unsigned char*>(data), n);
}

int dft_field_write(const char* path, const dft_grid* g, int dtype, int64_t ncols, const char* name, const void* data) {
    return guarded([&] {
        require_little_endian();
        if (!path || !g || !data || ncols < 1 || (dtype != DFT_F64 && dtype != DFT_F32))
            throw std::invalid_argument("dft_field_write: bad arguments");
        if (g->nx < 1 || g->ny < 1 || g->nz < 1) throw std::invalid_argument("dft_field_write: empty grid");
        const char* nm = name ? name : "";
        if (std::strlen(nm) > 63) throw std::invalid_argument("dft_field_write: name longer than 63 bytes");
        const uint64_t es = dtype == DFT_F64 ? 8 : 4;
        const uint64_t bytes = checked_mul(
            checked_mul(checked_mul(checked_mul((uint64_t)g->nx, (uint64_t)g->ny, "dft_field_write"), (uint64_t)g->nz,
                                    "dft_field_write"),
                        (uint64_t)ncols, "dft_field_write"),
            es, "dft_field_write");
        unsigned char h[kFieldHeader] = {};
        std::memcpy(h, kFieldMagic, 8);
        put<uint32_t>(h, 8, kFieldHeader);
        put<uint32_t>(h, 12, (uint32_t)dtype);
        put<int64_t>(h, 16, g->nx);
        put<int64_t>(h, 24, g->ny);
        put<int64_t>(h, 32, g->nz);
        put<int64_t>(h, 40, ncols);
        const double geo[9] = {g->xmin, g->xmax, g->ymin, g->ymax, g->zmin, g->zmax, g->dx, g->dy, g->dz};
        std::memcpy(h + 48, geo, sizeof geo);
        put<uint64_t>(h, 120, bytes);
        put<uint32_t>(h, 128, crc32_update(0, static_cast<const unsigned char*>(data), (size_t)bytes));
        std::memcpy(h + 132, nm, std::strlen(nm));
        put<uint32_t>(h, kFieldHeader - 4, crc32_update(0, h, kFieldHeader - 4));
        write_atomically(path, h, kFieldHeader, {{data, bytes}}, "dft_field_write");
    });
}

int dft_field_read_header(const char* path, dft_field_header* hdr) {
    return guarded([&] {
        require_little_endian();
        if (!path || !hdr) throw std::invalid_argument("dft_field_read_header: null argument");
        File in(path, "rb");
        if (!in.f) throw std::invalid_argument(std::string("dft_field_read_header: cannot open ") + path);
        *hdr = read_field_header(in.f, path);
    });
}

int dft_field_read(const char* path, dft_field_header* hdr, void* data, uint64_t capacity_bytes) {
    return guarded([&] {
        require_little_endian();
        if (!path || !data) throw std::invalid_argument("dft_field_read: null argument");
        File in(path, "rb");
        if (!in.f) throw std::invalid_argument(std::string("dft_field_read: cannot open ") + path);
        const dft_field_header r = read_field_header(in.f, path);
        if (r.data_bytes > capacity_bytes) throw std::invalid_argument(std::string("dft_field_read: destination too small for ") + path);
        if (file_size(in.f) != kFieldHeader + r.data_bytes)
            throw std::runtime_error(std::string("dft_field_read: file length disagrees with its header: ") + path);
        const uint32_t crc = read_payload(in.f, data, r.data_bytes, "dft_field_read");
        if (crc != r.crc) throw std::runtime_error(std::string("dft_field_read: checksum mismatch: ") + path);
        if (hdr) *hdr = r;
    });
}

int dft_patch_shard_write(const char* path, int64_t patch, int64_t count, int target_scale, double h, const float* in,
                          const float* tgt) {
    return guarded([&] {
        require_little_endian();
        if (!path || patch < 1 || count < 0 || (count > 0 && (!in || !tgt)))
            throw std::invalid_argument("dft_patch_shard_write: bad arguments");
        const uint64_t p3 = checked_mul(checked_mul((uint64_t)patch, (uint64_t)patch, "dft_patch_shard_write"),
                                        (uint64_t)patch, "dft_patch_shard_write");
        const uint64_t tb = checked_mul(checked_mul(p3, (uint64_t)count, "dft_patch_shard_write"), 4, "dft_patch_shard_write");
        const uint64_t ib = checked_mul(tb, 2, "dft_patch_shard_write");
        unsigned char hd[kPatchHeader] = {};
        std::memcpy(hd, kPatchMagic, 8);
        put<uint32_t>(hd, 8, kPatchHeader);
        put<uint32_t>(hd, 12, 2u);
        put<int64_t>(hd, 16, patch);
        put<int64_t>(hd, 24, count);
        put<uint64_t>(hd, 32, ib);
        put<uint64_t>(hd, 40, tb);
        put<uint32_t>(hd, 48, count ? crc32_update(0, reinterpret_cast<const unsigned char*>(in), (size_t)ib) : 0u);
        put<uint32_t>(hd, 52, count ? crc32_update(0, reinterpret_cast<const unsigned char*>(tgt), (size_t)tb) : 0u);
        put<int32_t>(hd, 56, (int32_t)target_scale);
        put<double>(hd, 64, h);
        put<uint32_t>(hd, kPatchHeader - 4, crc32_update(0, hd, kPatchHeader - 4));
        write_atomically(path, hd, kPatchHeader, {{in, ib}, {tgt, tb}}, "dft_patch_shard_write");
    });
}

int dft_patch_shard_read_header(const char* path, dft_patch_shard_header* hdr) {
    return guarded([&] {
        require_little_endian();
        if (!path || !hdr) throw std::invalid_argument("dft_patch_shard_read_header: null argument");
        File f(path, "rb");
        if (!f.f) throw std::invalid_argument(std::string("dft_patch_shard_read_header: cannot open ") + path);
        *hdr = read_patch_header(f.f, path).pub;
    });
}

int dft_patch_shard_read(const char* path, int64_t first, int64_t count, float* in, float* tgt) {
    return guarded([&] {
        require_little_endian();
        if (!path || first < 0 || count < 0) throw std::invalid_argument("dft_patch_shard_read: bad arguments");
        File f(path, "rb");
        if (!f.f) throw std::invalid_argument(std::string("dft_patch_shard_read: cannot open ") + path);
        const PatchHeader r = read_patch_header(f.f, path);
        if (first > r.pub.count || count > r.pub.count - first)
            throw std::invalid_argument(std::string("dft_patch_shard_read: range beyond the shard: ") + path);
        if (file_size(f.f) != kPatchHeader + r.in_bytes + r.tgt_bytes)
            throw std::runtime_error(std::string("dft_patch_shard_read: file length disagrees with its header: ") + path);
        if (count == 0) return;
        // every read covers its whole section for the checksum, so a range read is verified too
        const uint64_t p3 = (uint64_t)r.pub.patch * r.pub.patch * r.pub.patch;
        if (in) {
            std::fseek(f.f, (long)kPatchHeader, SEEK_SET);
            const uint32_t crc = read_range_verified(f.f, r.in_bytes, 2 * 4 * p3 * (uint64_t)first,
                                                     2 * 4 * p3 * (uint64_t)count, in, "dft_patch_shard_read");
            if (crc != r.pub.crc_in) throw std::runtime_error(std::string("dft_patch_shard_read: input checksum mismatch: ") + path);
        }
        if (tgt) {
            std::fseek(f.f, (long)(kPatchHeader + r.in_bytes), SEEK_SET);
            const uint32_t crc = read_range_verified(f.f, r.tgt_bytes, 4 * p3 * (uint64_t)first,
                                                     4 * p3 * (uint64_t)count, tgt, "dft_patch_shard_read");
            if (crc != r.pub.crc_tgt) throw std::runtime_error(std::string("dft_patch_shard_read: target checksum mismatch: ") + path);
        }
    });
}

int dft_scf_result_json(const dft_grid* g, int natoms, const double* nuclei, const dft_scf_config* cfg,
                        const dft_scf_report* rep, const double* eig, int n_alpha, int n_beta, char* buf, size_t cap,
                        size_t* len) {
    return guarded([&] {
        if (!g || !cfg || !rep || !len || natoms < 0 || (natoms > 0 && !nuclei) || n_alpha < 0 || n_beta < 0 ||
            (n_alpha + n_beta > 0 && !eig))
            throw std::invalid_argument("dft_scf_result_json: bad arguments");
        std::string s;
        s.reserve(2048 + 64 * (size_t)(natoms + n_alpha + n_beta));
        double zsum = 0.0;
        for (int a = 0; a < natoms; ++a) zsum += nuclei[4 * a];
        s += "{\"format\":\"dft-scf-result\",\"version\":1,";
        // MolecularResult (DFTData.h:86-91)
        jkey(s, "electrons");
        jint(s, (int64_t)std::llround(zsum) - cfg->charge);
        s += ',';
        jkey(s, "charge");
        jint(s, cfg->charge);
        s += ",\"grid\":{\"n\":[";
        jint(s, g->nx), s += ',', jint(s, g->ny), s += ',', jint(s, g->nz);
        s += "],\"min\":[";
        jnum(s, g->xmin), s += ',', jnum(s, g->ymin), s += ',', jnum(s, g->zmin);
        s += "],\"max\":[";
        jnum(s, g->xmax), s += ',', jnum(s, g->ymax), s += ',', jnum(s, g->zmax);
        s += "],\"spacing\":[";
        jnum(s, g->dx), s += ',', jnum(s, g->dy), s += ',', jnum(s, g->dz);
        s += "]},\"nuclei\":[";
        for (int a = 0; a < natoms; ++a) {
            if (a) s += ',';
            s += "{\"Z\":";
            jnum(s, nuclei[4 * a]);
            s += ",\"position\":[";
            jnum(s, nuclei[4 * a + 1]), s += ',', jnum(s, nuclei[4 * a + 2]), s += ',', jnum(s, nuclei[4 * a + 3]);
            s += "]}";
        }
        s += "],\"config\":{";
        jkey(s, "functional");
        if (cfg->functional == DFT_XC_PZ81)
            jstr(s, "PZ81");
        else if (cfg->functional == DFT_XC_PBE96)
            jstr(s, "PBE96");
        else
            jint(s, cfg->functional);
        s += ',', jkey(s, "multiplicity"), jint(s, cfg->multiplicity);
        s += ',', jkey(s, "precision"), jenum(s, cfg->precision, {"fp64", "mixed"});
        s += ',', jkey(s, "nstates"), jint(s, cfg->nstates);
        s += ',', jkey(s, "poisson"), jenum(s, cfg->poisson_method, {"jacobi", "chebyshev", "direct"});
        s += ',', jkey(s, "poissonFixedSweeps"), jint(s, cfg->poisson_fixed);
        s += ',', jkey(s, "stopRule"), jenum(s, cfg->stop_rule, {"reference", "energy"});
        s += ',', jkey(s, "energyTolerance"), jnum(s, cfg->energy_tol);
        s += ',', jkey(s, "initialGuess"), jenum(s, cfg->init_guess, {"bare", "hydrogenic"});
        s += ',', jkey(s, "davidsonTolerance"), jnum(s, cfg->davidson_tol);
        s += ',', jkey(s, "preconditioner"),
            jenum(s, cfg->precond, {"diagonal", "cnn", "kinetic", "kinetic_chebyshev", "kinetic_exact"});
        s += ',', jkey(s, "mixer"), jenum(s, cfg->mixer, {"linear", "pulay"});
        s += ',', jkey(s, "smearing"), jnum(s, cfg->smearing);
        // the remaining knobs of dft_scf_config, so that the document re-runs as it ran
        s += ',', jkey(s, "maxIterations"), jint(s, cfg->max_iterations);
        s += ',', jkey(s, "maxSubspace"), jint(s, cfg->max_subspace);
        s += ',', jkey(s, "poissonBatch"), jint(s, cfg->poisson_batch);
        s += ',', jkey(s, "poissonTolerance"), jnum(s, cfg->poisson_tol);
        s += ',', jkey(s, "davidsonMaxIterations"), jint(s, cfg->davidson_max_iter);
        s += ',', jkey(s, "refineIterations"), jint(s, cfg->refine_iterations);
        s += ',', jkey(s, "precondSweeps"), jint(s, cfg->precond_sweeps);
        s += ',', jkey(s, "precondShift"), jnum(s, cfg->precond_shift);
        s += ',', jkey(s, "precondShiftScale"), jnum(s, cfg->precond_shift_scale);
        s += ',', jkey(s, "mixDepth"), jint(s, cfg->mix_depth);
        s += ',', jkey(s, "mixBeta"), jnum(s, cfg->mix_beta);
        s += ',', jkey(s, "mixKerker"), jnum(s, cfg->mix_kerker);
        s += ',', jkey(s, "orthoGram"), jenum(s, cfg->ortho_gram, {"default", "f32", "split_bf16"});
        s += ',', jkey(s, "xcFlux"), jenum(s, cfg->xc_flux, {"reference", "consistent"});
        s += "},\"scf\":{\"energy\":{";
        // EnergyComponents (DFTData.h:47-54)
        static const char* const ek[6] = {"kinetic", "external", "hartree", "exchangeCorrelation", "nuclearRepulsion", "total"};
        for (int i = 0; i < 6; ++i) {
            if (i) s += ',';
            jkey(s, ek[i]);
            jnum(s, rep->energy[i]);
        }
        // SCFResult scalars (DFTData.h:73-80)
        s += "},";
        jkey(s, "densityDifference"), jnum(s, rep->density_difference);
        s += ',', jkey(s, "energyDifference"), jnum(s, rep->energy_difference);
        s += ',', jkey(s, "maxKSResidual"), jnum(s, rep->ks_residual);
        s += ',', jkey(s, "iterations"), jint(s, rep->iterations);
        s += ',', jkey(s, "converged"), s += rep->converged ? "true" : "false";
        s += ",\"molecularOrbitals\":[";
        for (int i = 0; i < n_alpha + n_beta; ++i) {
            if (i) s += ',';
            s += "{\"spin\":";
            jstr(s, i < n_alpha ? "alpha" : "beta");
            s += ",\"electrons\":";
            // whole orbitals carry one electron each (buildSpinOccupations, MolecularSCFHelpers.cpp:185-199);
            // smeared occupations are fractional and not listed
            s += cfg->smearing > 0.0 ? "null" : "1";
            s += ",\"eigenvalue\":";
            jnum(s, eig[i]);
            s += '}';
        }
        s += "]}";
        if (cfg->smearing > 0.0) {
            s += ",\"smearing\":{";
            jkey(s, "kT"), jnum(s, cfg->smearing);
            s += ",\"fermiLevel\":[", jnum(s, rep->fermi_level[0]), s += ',', jnum(s, rep->fermi_level[1]), s += "],";
            jkey(s, "entropyEnergy"), jnum(s, rep->smearing_energy);
            s += '}';
        }
        s += ",\"timing\":{";
        jkey(s, "total"), jnum(s, rep->seconds_total);
        s += ',', jkey(s, "toFirstEnergyTolerance"), jnum(s, rep->seconds_to_first_de);
        s += ',', jkey(s, "poisson"), jnum(s, rep->seconds_poisson);
        s += ',', jkey(s, "xc"), jnum(s, rep->seconds_xc);
        s += ',', jkey(s, "davidson"), jnum(s, rep->seconds_davidson);
        s += ',', jkey(s, "other"), jnum(s, rep->seconds_other);
        s += "},\"counters\":{";
        jkey(s, "firstEnergyToleranceIteration"), jint(s, rep->first_de_iteration);
        s += ',', jkey(s, "refineIterations"), jint(s, rep->refine_iterations);
        s += ',', jkey(s, "hamiltonianColumns"), jint(s, rep->happly_columns);
        s += ',', jkey(s, "poissonSweeps"), jint(s, rep->poisson_sweeps);
        s += ',', jkey(s, "davidsonIterations"), jint(s, rep->davidson_iterations);
        s += ',', jkey(s, "cnnApplied"), jint(s, rep->cnn_applied);
        s += ',', jkey(s, "cnnFallbacks"), jint(s, rep->cnn_fallbacks);
        s += ',', jkey(s, "blockPasses"), jint(s, rep->block_passes);
        s += ',', jkey(s, "blockFallbacks"), jint(s, rep->block_fallbacks);
        s += "}}";
        *len = s.size();
        if (!buf) return;
        if (cap <= s.size()) throw std::invalid_argument("dft_scf_result_json: buffer too small");
        std::memcpy(buf, s.c_str(), s.size() + 1);
    });
}

}  // extern "C"
