// formats_check.cpp -- the file-format code (dft_amd/host/formats.cpp) under
// AddressSanitizer and UBSan: round trips, every truncation of a file, random
// byte corruption and random headers must end in a status code, never in a
// crash, an out-of-bounds access or undefined behaviour; the result document
// must be NUL-terminated and exactly as long as reported.  Built and run by
// tests/test_formats_sanitized.py (host only, no GPU).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "dft_io.h"

namespace dftc {
// the status translation of libdft_core.so (core_capi.cpp), without the rest of the library
int guarded_call(const std::function<void()>& f) {
    try {
        f();
        return DFT_OK;
    } catch (const std::invalid_argument&) {
        return DFT_EINVAL;
    } catch (const std::exception&) {
        return DFT_ENUMERIC;
    }
}
}  // namespace dftc

static int failures = 0;
#define CHECK(c)                                                      \
    do {                                                              \
        if (!(c)) {                                                   \
            std::fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                               \
        }                                                             \
    } while (0)

static std::vector<unsigned char> slurp(const std::string& p) {
    std::vector<unsigned char> b;
    if (std::FILE* f = std::fopen(p.c_str(), "rb")) {
        unsigned char buf[4096];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + n);
        std::fclose(f);
    }
    return b;
}
static void spit(const std::string& p, const std::vector<unsigned char>& b, size_t n) {
    std::FILE* f = std::fopen(p.c_str(), "wb");
    if (n) std::fwrite(b.data(), 1, n, f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    std::mt19937_64 rng(7);
    dft_grid g{};
    g.nx = 5; g.ny = 4; g.nz = 3;
    g.xmin = -1; g.xmax = 1; g.ymin = -1; g.ymax = 0.5; g.zmin = 0; g.zmax = 2;
    g.dx = 0.5; g.dy = 0.5; g.dz = 1.0;
    const size_t n = 60, ncols = 3;
    std::vector<double> data(n * ncols);
    for (auto& v : data) v = std::uniform_real_distribution<double>(-1, 1)(rng);
    const std::string fpath = dir + "/f.dftf", bad = dir + "/bad.dftf";
    CHECK(dft_field_write(fpath.c_str(), &g, DFT_F64, (int64_t)ncols, "psi", data.data()) == DFT_OK);
    std::vector<double> back(n * ncols);
    dft_field_header h;
    CHECK(dft_field_read(fpath.c_str(), &h, back.data(), back.size() * 8) == DFT_OK);
    CHECK(std::memcmp(back.data(), data.data(), back.size() * 8) == 0);
    CHECK(dft_field_read(fpath.c_str(), &h, back.data(), back.size() * 8 - 1) == DFT_EINVAL);
    const std::vector<unsigned char> good = slurp(fpath);
    // every truncation
    for (size_t len = 0; len < good.size(); ++len) {
        spit(bad, good, len);
        const int rc = dft_field_read(bad.c_str(), &h, back.data(), back.size() * 8);
        CHECK(rc == DFT_EINVAL || rc == DFT_ENUMERIC);
    }
    // random corruption of one to four bytes: a status, and the data stays within the destination
    for (int t = 0; t < 3000; ++t) {
        std::vector<unsigned char> b = good;
        const int k = 1 + (int)(rng() % 4);
        for (int i = 0; i < k; ++i) b[rng() % b.size()] ^= (unsigned char)(1 + rng() % 255);
        spit(bad, b, b.size());
        std::vector<double> dst(n * ncols);
        const int rc = dft_field_read(bad.c_str(), &h, dst.data(), dst.size() * 8);
        CHECK(rc == DFT_OK || rc == DFT_EINVAL || rc == DFT_ENUMERIC);
        if (rc == DFT_OK) CHECK(b == good);  // a changed byte is always detected
        dft_field_header h2;
        const int rh = dft_field_read_header(bad.c_str(), &h2);
        CHECK(rh == DFT_OK || rh == DFT_EINVAL || rh == DFT_ENUMERIC);
        // the header's own checksum: only an untouched header reads
        CHECK(rh != DFT_OK || std::memcmp(b.data(), good.data(), 256) == 0);
    }
    // random headers
    for (int t = 0; t < 2000; ++t) {
        std::vector<unsigned char> b(256 + (rng() % 64));
        for (auto& c : b) c = (unsigned char)rng();
        if (t % 2) std::memcpy(b.data(), "DFTFLD01\x00\x01\x00\x00", 12);
        spit(bad, b, b.size());
        std::vector<double> dst(64);
        const int rc = dft_field_read(bad.c_str(), &h, dst.data(), dst.size() * 8);
        CHECK(rc == DFT_EINVAL || rc == DFT_ENUMERIC);
    }
    // patch shards
    const int64_t P = 3, cnt = 4;
    std::vector<float> in(2 * 27 * cnt), tg(27 * cnt);
    for (auto& v : in) v = (float)std::uniform_real_distribution<double>(-1, 1)(rng);
    for (auto& v : tg) v = (float)std::uniform_real_distribution<double>(-1, 1)(rng);
    const std::string spath = dir + "/s.dftp", sbad = dir + "/bad.dftp";
    CHECK(dft_patch_shard_write(spath.c_str(), P, cnt, 1, 0.25, in.data(), tg.data()) == DFT_OK);
    std::vector<float> in2(in.size()), tg2(tg.size());
    CHECK(dft_patch_shard_read(spath.c_str(), 0, cnt, in2.data(), tg2.data()) == DFT_OK);
    CHECK(in2 == in && tg2 == tg);
    CHECK(dft_patch_shard_read(spath.c_str(), 2, 3, in2.data(), tg2.data()) == DFT_EINVAL);
    const std::vector<unsigned char> sgood = slurp(spath);
    for (size_t len = 0; len < sgood.size(); ++len) {
        spit(sbad, sgood, len);
        const int rc = dft_patch_shard_read(sbad.c_str(), 0, cnt, in2.data(), tg2.data());
        CHECK(rc == DFT_EINVAL || rc == DFT_ENUMERIC);
    }
    for (int t = 0; t < 3000; ++t) {
        std::vector<unsigned char> b = sgood;
        b[rng() % b.size()] ^= (unsigned char)(1 + rng() % 255);
        spit(sbad, b, b.size());
        dft_patch_shard_header ph;
        const int rh = dft_patch_shard_read_header(sbad.c_str(), &ph);
        CHECK(rh == DFT_OK || rh == DFT_EINVAL || rh == DFT_ENUMERIC);
        CHECK(rh != DFT_OK || std::memcmp(b.data(), sgood.data(), 128) == 0);
        if (rh == DFT_OK && ph.patch == P && ph.count == cnt) {
            const int rc = dft_patch_shard_read(sbad.c_str(), 0, cnt, in2.data(), tg2.data());
            CHECK(rc == DFT_OK || rc == DFT_EINVAL || rc == DFT_ENUMERIC);
        }
    }
    // ranges: huge first/count values are rejected (no overflow); a corrupted byte is detected by a
    // range read that does not cover it
    {
        const int64_t big = INT64_MAX - 1;
        CHECK(dft_patch_shard_read(spath.c_str(), big, 2, in2.data(), tg2.data()) == DFT_EINVAL);
        CHECK(dft_patch_shard_read(spath.c_str(), 1, big, in2.data(), tg2.data()) == DFT_EINVAL);
        CHECK(dft_patch_shard_read(spath.c_str(), 0, 1, in2.data(), tg2.data()) == DFT_OK);
        std::vector<unsigned char> b = sgood;
        b[b.size() - 1] ^= 0x40;  // the last target value
        spit(sbad, b, b.size());
        CHECK(dft_patch_shard_read(sbad.c_str(), 0, 1, nullptr, tg2.data()) == DFT_ENUMERIC);
    }
    // the result document: random values (non-finite ones included), enums out of range
    for (int t = 0; t < 500; ++t) {
        dft_scf_config cfg{};
        dft_scf_report rep{};
        const double odd[5] = {NAN, INFINITY, -INFINITY, 1e-310, -0.0};
        auto val = [&]() { return rng() % 5 == 0 ? odd[rng() % 5] : std::ldexp(std::uniform_real_distribution<double>(-1, 1)(rng), (int)(rng() % 2000) - 1000); };
        cfg.functional = (int)(rng() % 5) - 1;
        cfg.precision = (int)(rng() % 4) - 1;
        cfg.poisson_method = (int)(rng() % 5) - 1;
        cfg.precond = (int)(rng() % 8) - 1;
        cfg.mixer = (int)(rng() % 4) - 1;
        cfg.ortho_gram = (int)(rng() % 5) - 1;
        cfg.smearing = val();
        cfg.energy_tol = val();
        for (double& e : rep.energy) e = val();
        rep.ks_residual = val();
        rep.iterations = (int)rng();
        rep.happly_columns = (int64_t)rng();
        const int na = (int)(rng() % 4), nb = (int)(rng() % 4);
        std::vector<double> eig(na + nb + 1), nuc(4 * 3);
        for (auto& e : eig) e = val();
        for (auto& e : nuc) e = val();
        size_t len = 0;
        CHECK(dft_scf_result_json(&g, 3, nuc.data(), &cfg, &rep, eig.data(), na, nb, nullptr, 0, &len) == DFT_OK);
        std::vector<char> buf(len + 1, 'x');
        CHECK(dft_scf_result_json(&g, 3, nuc.data(), &cfg, &rep, eig.data(), na, nb, buf.data(), len, &len) == DFT_EINVAL);
        CHECK(dft_scf_result_json(&g, 3, nuc.data(), &cfg, &rep, eig.data(), na, nb, buf.data(), buf.size(), &len) == DFT_OK);
        CHECK(buf[len] == '\0' && std::strlen(buf.data()) == len);
        CHECK(std::strstr(buf.data(), "nan") == nullptr && std::strstr(buf.data(), "inf") == nullptr);
    }
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
