// upf_io.cpp -- particle state as text, in the reference's file formats
// (UnifiedIO, src/cpu/UnifiedIO.cpp): the liquid and rigid dumps, the
// all-particle position dump, and the PCISPH restart file that
// Scene::readFluidObjectFromFile loads.  Host code; the simulation downloads
// its slot arrays and hands them over (upf_save_text in upf_sim.hip).
//
// Lines are formatted with printf conversions that give the bytes the
// reference's iostream formatting gives: an ostream prints a float as the
// double it converts to, through "%.<precision>g", padded to setw() on the
// right when std::ios::left is set.  Large files are formatted by several
// threads, each into its own buffer, and written in particle order.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "upf.h"
#include "upf_internal.h"

namespace upf {
namespace {

bool kept(int which, int t) {
    if (which == UPF_TEXT_LIQUID) return t == UPF_LIQUID;
    if (which == UPF_TEXT_RIGID) return t == UPF_RIGID || t == UPF_FROZEN;
    return true;
}

struct Fields {
    const float* pos4;
    const int* type;
    const float *vel4, *p, *rho;
};

// One particle's line, appended to `out`.
void format_line(std::string& out, int which, const Fields& f, int i) {
    char buf[256];
    const double x = f.pos4[4 * i], y = f.pos4[4 * i + 1], z = f.pos4[4 * i + 2];
    int len = 0;
    switch (which) {
        case UPF_TEXT_LIQUID:
        case UPF_TEXT_RIGID:
            // SaveLiquidParticleInfoFull / SaveRigidParticleInfoFull
            // (src/cpu/UnifiedIO.cpp:62-166): precision 8, left-aligned in
            // widths 15 15 15 3 15 - - 2 2 2 2; velocity, rest density, mass,
            // density and gas constant are written as zeros.
            len = std::snprintf(buf, sizeof buf, "%-15.8g %-15.8g %-15.8g %-3d %-15d %d %d %-2d %-2d %-2d %-2d \n", x,
                                y, z, f.type[i], 0, 0, 0, 0, 0, 0, 0);
            break;
        case UPF_TEXT_POSITIONS:
            // SaveParticlePositions (src/cpu/UnifiedIO.cpp:13-60): precision
            // 20, then eight zero fields.
            len = std::snprintf(buf, sizeof buf, "%.20g %.20g %.20g 0 0 0 0 0 0 0 0 \n", x, y, z);
            break;
        default:
            // The nine fields ReadInFluidParticlesPCISPH reads
            // (src/cpu/UnifiedIO.cpp:196-220); nine significant digits give
            // every float back exactly.
            len = std::snprintf(buf, sizeof buf, "%.9g %.9g %.9g %d %.9g %.9g %.9g %.9g %.9g\n", x, y, z,
                                f.type[i], (double)f.vel4[4 * i], (double)f.vel4[4 * i + 1],
                                (double)f.vel4[4 * i + 2], (double)f.p[i], (double)f.rho[i]);
            break;
    }
    out.append(buf, (size_t)len);
}

}  // namespace

int write_text(const char* path, int which, int n, const float* pos4, const int* type, const float* vel4,
               const float* p, const float* rho) {
    if (!path || n < 0 || (n > 0 && (!pos4 || !type)) || which < UPF_TEXT_LIQUID || which > UPF_TEXT_RESTART)
        return UPF_ERR_ARG;
    if (which == UPF_TEXT_RESTART && n > 0 && (!vel4 || !p || !rho)) return UPF_ERR_ARG;
    const Fields f{pos4, type, vel4, p, rho};
    long long count = 0;
    for (int i = 0; i < n; ++i) count += kept(which, type[i]);

    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int parts = n < 65536 ? 1 : std::min(16, hw);
    std::vector<std::string> text(parts);
    auto work = [&](int t) {
        const int lo = (int)((long long)n * t / parts), hi = (int)((long long)n * (t + 1) / parts);
        text[t].reserve((size_t)(hi - lo) * 96);
        for (int i = lo; i < hi; ++i)
            if (kept(which, type[i])) format_line(text[t], which, f, i);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < parts; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();

    FILE* fp = std::fopen(path, "wb");
    if (!fp) return UPF_ERR_IO;
    bool ok = std::fprintf(fp, "%lld\n", count) > 0;
    for (const auto& s : text) ok = ok && std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? UPF_OK : UPF_ERR_IO;
}

// DumpGPUParamsToFile (src/cpu/UnifiedPhysics.cpp:411-474): raw native-endian
// uint32 / float records appended to the two files.
int write_cell_dump(const char* particles_path, const char* blocks_path, int n, const float* rec4,
                    std::vector<uint32_t>& blocks) {
    const size_t nb = blocks.size() / 3;
    std::vector<uint32_t> order(nb);
    for (size_t b = 0; b < nb; ++b) order[b] = (uint32_t)b;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return blocks[3 * a] < blocks[3 * b]; });
    std::vector<uint32_t> out(1 + 3 * nb);
    out[0] = (uint32_t)nb;
    for (size_t b = 0; b < nb; ++b)
        for (int f = 0; f < 3; ++f) out[1 + 3 * b + f] = blocks[3 * order[b] + f];

    FILE* fp = std::fopen(particles_path, "ab");
    if (!fp) return UPF_ERR_IO;
    const uint32_t count = (uint32_t)n;
    bool ok = std::fwrite(&count, 4, 1, fp) == 1 && std::fwrite(rec4, 16, (size_t)n, fp) == (size_t)n;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return UPF_ERR_IO;
    fp = std::fopen(blocks_path, "ab");
    if (!fp) return UPF_ERR_IO;
    ok = std::fwrite(out.data(), 4, out.size(), fp) == out.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? UPF_OK : UPF_ERR_IO;
}

}  // namespace upf

namespace {

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

bool all_of(const char* b, const char* e, const char* set) {
    for (const char* c = b; c < e; ++c)
        if (!std::strchr(set, *c)) return false;
    return true;
}

// strtof on the token: the general case of parse_float.
bool parse_float_strtof(const char* b, const char* e, float& v) {
    char* end = nullptr;
    errno = 0;
    v = std::strtof(b, &end);
    return end == e && !(errno == ERANGE && (v > 1.0f || v < -1.0f));
}

// A float in the decimal forms an istream accepts -- [+-] digits [. digits]
// [(e|E) [+-] digits], at least one mantissa digit; no hexadecimal, inf or
// nan, and no overflow (the istream fails there too); an underflow keeps the
// rounded value.  [b, e) is one token of a NUL-terminated text.  Correctly
// rounded: with at most 19 significant digits m <= 2^53 and a power of ten
// |p| <= 22 (both exact doubles), m * 10^p or m / 10^-p is the correctly
// rounded double (Clinger's fast path, 1990), and rounding that double to
// float is the correctly rounded float of the decimal unless the double is a
// midpoint between two floats; that case, subnormal or huge results, and
// all other forms go to strtof.
bool parse_float(const char* b, const char* e, float& v) {
    static const double pow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                     1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    const char* c = b;
    const bool neg = c < e && *c == '-';
    if (c < e && (*c == '+' || *c == '-')) ++c;
    uint64_t m = 0;
    int sig = 0, p = 0;
    bool digit = false, exact = true;
    for (; c < e && *c >= '0' && *c <= '9'; ++c) {
        digit = true;
        if (m == 0 && *c == '0') continue;
        if (sig < 19) { m = m * 10 + (uint64_t)(*c - '0'); ++sig; }
        else { ++p; exact = false; }
    }
    if (c < e && *c == '.') {
        for (++c; c < e && *c >= '0' && *c <= '9'; ++c) {
            digit = true;
            if (m == 0 && *c == '0') { --p; continue; }
            if (sig < 19) { m = m * 10 + (uint64_t)(*c - '0'); ++sig; --p; }
            else exact = false;
        }
    }
    if (!digit) return false;
    if (c < e && (*c == 'e' || *c == 'E')) {
        ++c;
        const bool eneg = c < e && *c == '-';
        if (c < e && (*c == '+' || *c == '-')) ++c;
        if (c == e || *c < '0' || *c > '9') return false;
        int x = 0;
        for (; c < e && *c >= '0' && *c <= '9'; ++c) x = x < 100000 ? x * 10 + (*c - '0') : x;
        p += eneg ? -x : x;
    }
    if (c != e) return false;
    if (m == 0) {
        v = neg ? -0.0f : 0.0f;
        return true;
    }
    if (exact && m <= (1ull << 53) && p >= -22 && p <= 22) {
        const double d = p < 0 ? (double)m / pow10[-p] : (double)m * pow10[p];
        uint64_t bits;
        std::memcpy(&bits, &d, 8);
        const bool midpoint = (bits & ((1ull << 29) - 1)) == (1ull << 28);
        if (!midpoint && d >= 0x1p-126 && d <= 0x1.fffffep127) {
            v = (float)(neg ? -d : d);
            return true;
        }
    }
    return parse_float_strtof(b, e, v);
}

bool parse_int(const char* b, const char* e, int& v) {
    if (b == e || !all_of(b, e, "0123456789+-")) return false;
    char* end = nullptr;
    errno = 0;
    const long x = std::strtol(b, &end, 10);
    if (end != e || errno == ERANGE || x < INT_MIN || x > INT_MAX) return false;
    v = (int)x;
    return true;
}

// The next token at or after *cur before end: [b, e); false at the end.
bool next_token(const char*& cur, const char* end, const char*& b, const char*& e) {
    while (cur < end && is_space(*cur)) ++cur;
    b = cur;
    while (cur < end && !is_space(*cur)) ++cur;
    e = cur;
    return e > b;
}

}  // namespace

extern "C" {

int upf_write_text(const char* path, int which, int n, const float* pos4, const int* type, const float* vel4,
                   const float* corr_pressure, const float* predicted_density) {
    return upf::write_text(path, which, n, pos4, type, vel4, corr_pressure, predicted_density);
}

// ReadInFluidParticleNum + ReadInFluidParticlesPCISPH (src/cpu/UnifiedIO.cpp:168-220).
// The file is read whole; large files are parsed by several threads: the text
// is cut at whitespace, each thread counts the tokens of its piece, a prefix
// sum gives every piece's first token number, and token k of the particle
// fields goes to particle k / 9, field k % 9.  Tokens after the last
// particle's are not read (as the reference's loop stops there).
int upf_read_restart(const char* path, int cap, int* n_out, float* pos4, float* vel4, int* type,
                     float* corr_pressure, float* predicted_density) {
    if (!path || !n_out || cap < 0) return UPF_ERR_ARG;
    const bool count_only = !pos4 && !vel4 && !type && !corr_pressure && !predicted_density;
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return UPF_ERR_IO;
    std::string text;
    if (!count_only && std::fseek(fp, 0, SEEK_END) == 0) {  // one read of the whole file
        const long size = std::ftell(fp);
        std::rewind(fp);
        if (size > 0) {
            text.resize((size_t)size);
            text.resize(std::fread(&text[0], 1, (size_t)size, fp));
        }
    }
    char chunk[1 << 12];
    size_t got;
    // the rest (all of it for a stream without a size); the count alone needs
    // only the text up to the end of its token
    while ((got = std::fread(chunk, 1, sizeof chunk, fp)) > 0) {
        text.append(chunk, got);
        if (count_only) {
            const size_t b = text.find_first_not_of(" \t\n\v\f\r");
            if (b != std::string::npos && text.find_first_of(" \t\n\v\f\r", b) != std::string::npos) break;
        }
    }
    const bool read_ok = !std::ferror(fp);
    std::fclose(fp);
    if (!read_ok) return UPF_ERR_IO;

    const char* const base = text.c_str();  // NUL-terminated: strtof / strtol stop at the end
    const char* const end = base + text.size();
    const char *cur = base, *tb, *te;
    int cnt = 0;
    if (!next_token(cur, end, tb, te) || !parse_int(tb, te, cnt) || cnt < 0) return UPF_ERR_ARG;
    *n_out = cnt;
    if (count_only) return UPF_OK;
    if (cnt > cap) return UPF_ERR_RANGE;
    const size_t need = (size_t)9 * cnt;

    const size_t body = (size_t)(end - cur);
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int parts = body < ((size_t)1 << 20) ? 1 : std::min(16, hw);
    std::vector<const char*> cut(parts + 1);
    cut[0] = cur;
    cut[parts] = end;
    for (int t = 1; t < parts; ++t) {
        const char* c = std::max(cut[t - 1], cur + body * t / parts);
        while (c < end && !is_space(*c)) ++c;  // a cut never splits a token
        cut[t] = c;
    }
    std::vector<size_t> first(parts + 1, 0);
    auto run = [&](auto&& fn) {
        std::vector<std::thread> pool;
        for (int t = 1; t < parts; ++t) pool.emplace_back(fn, t);
        fn(0);
        for (auto& th : pool) th.join();
    };
    run([&](int t) {
        const char *c = cut[t], *b, *e;
        size_t k = 0;
        while (next_token(c, cut[t + 1], b, e)) ++k;
        first[t + 1] = k;
    });
    for (int t = 0; t < parts; ++t) first[t + 1] += first[t];
    if (first[parts] < need) return UPF_ERR_ARG;  // fewer particles than the count says

    std::atomic<int> bad{0};
    run([&](int t) {
        const char *c = cut[t], *b, *e;
        for (size_t k = first[t]; k < need && next_token(c, cut[t + 1], b, e); ++k) {
            const size_t i = k / 9;
            const int f = (int)(k % 9);
            float v = 0.0f;
            int ti = 0;
            if (!(f == 3 ? parse_int(b, e, ti) : parse_float(b, e, v))) {
                bad.store(1, std::memory_order_relaxed);
                return;
            }
            switch (f) {
                case 0: case 1: case 2:
                    if (pos4) pos4[4 * i + f] = v;
                    break;
                case 3:
                    if (type) type[i] = ti;
                    break;
                case 4: case 5: case 6:
                    if (vel4) vel4[4 * i + f - 4] = v;
                    break;
                case 7:
                    if (corr_pressure) corr_pressure[i] = v;
                    break;
                default:
                    if (predicted_density) predicted_density[i] = v;
            }
            if (f == 8) {
                if (pos4) pos4[4 * i + 3] = 0.0f;
                if (vel4) vel4[4 * i + 3] = 0.0f;
            }
        }
    });
    return bad.load() ? UPF_ERR_ARG : UPF_OK;
}

}  // extern "C"
