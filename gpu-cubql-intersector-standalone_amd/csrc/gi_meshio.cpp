// gi_meshio.cpp -- OFF input with the reference's canonical face indexing, and
// the output formats declared in include/gi_meshio.h (host code only).
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gi_meshio.h"

struct gi_mesh_data {
    std::vector<double> xyz;
    std::vector<uint32_t> tri, rows, degenerate, duplicate;
};

namespace {

thread_local std::string g_err;

int io_fail(const std::string& msg) {
    g_err = msg;
    return GI_ERR_IO;
}

int arg_fail(const std::string& msg) {
    g_err = msg;
    return GI_ERR_INVALID_ARGUMENT;
}

bool read_file(const char* path, std::string& out) {
    FILE* f = fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out.append(buf, n);
    bool ok = !ferror(f);
    fclose(f);
    return ok;
}

// Lines of the file with '#' comments removed; blank lines skipped.
struct LineReader {
    const std::string& s;
    size_t pos = 0;
    uint64_t lineno = 0;
    explicit LineReader(const std::string& str) : s(str) {}
    bool next(std::string& line) {
        while (pos < s.size()) {
            size_t e = s.find('\n', pos);
            if (e == std::string::npos) e = s.size();
            line.assign(s, pos, e - pos);
            pos = e + 1;
            ++lineno;
            size_t h = line.find('#');
            if (h != std::string::npos) line.resize(h);
            if (line.find_first_not_of(" \t\r\f\v") != std::string::npos) return true;
        }
        return false;
    }
};

// whitespace-separated tokens of one line
std::vector<std::string> tokens(const std::string& line) {
    std::vector<std::string> t;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && isspace((unsigned char)line[i])) ++i;
        size_t j = i;
        while (j < line.size() && !isspace((unsigned char)line[j])) ++j;
        if (j > i) t.emplace_back(line, i, j - i);
        i = j;
    }
    return t;
}

bool parse_u64(const std::string& t, uint64_t& v) {
    if (t.empty() || !isdigit((unsigned char)t[0])) return false;
    errno = 0;
    char* end = nullptr;
    unsigned long long x = strtoull(t.c_str(), &end, 10);
    if (errno || *end) return false;
    v = x;
    return true;
}

bool parse_double(const std::string& t, double& v) {
    errno = 0;
    char* end = nullptr;
    v = strtod(t.c_str(), &end);
    return end != t.c_str() && *end == 0 && errno != EINVAL;
}

void put_double(std::string& o, double x) {
    char b[40];
    snprintf(b, sizeof(b), "%.17g", x);  // what a stream with precision 17 prints
    o += b;
}

void put_u64(std::string& o, uint64_t x) {
    char b[24];
    snprintf(b, sizeof(b), "%llu", (unsigned long long)x);
    o += b;
}

bool write_file(const char* path, const std::string& text) {
    FILE* f = fopen(path, "wb");
    if (!f) return false;
    bool ok = fwrite(text.data(), 1, text.size(), f) == text.size();
    ok = (fclose(f) == 0) && ok;
    return ok;
}

// Fan triangulation and the two drop rules of the reference loader
// (RobustOffLoader.h:142-178): per polygon row r, triangles (p0, pk, pk+1) for
// k = 1 .. size-2; a triangle with a repeated vertex drops (the row is listed
// once per such triangle); a polygon of fewer than 3 vertices drops; then, over
// the vertex sets sorted with the triangle's position as the last key, every
// triangle equal to its predecessor drops, so the first occurrence stays.
void triangulate(const std::vector<std::vector<uint32_t>>& polys, gi_mesh_data& d, uint64_t& extra) {
    std::vector<std::array<uint32_t, 3>> tris;
    std::vector<uint32_t> rows;
    extra = 0;
    for (size_t r = 0; r < polys.size(); ++r) {
        const auto& p = polys[r];
        if (p.size() < 3) {
            d.degenerate.push_back((uint32_t)r);
            continue;
        }
        for (size_t k = 1; k + 1 < p.size(); ++k) {
            if (p[0] == p[k] || p[k] == p[k + 1] || p[0] == p[k + 1]) {
                d.degenerate.push_back((uint32_t)r);
                continue;
            }
            if (k > 1) ++extra;
            tris.push_back({p[0], p[k], p[k + 1]});
            rows.push_back((uint32_t)r);
        }
    }
    std::vector<std::pair<std::array<uint32_t, 3>, size_t>> keys(tris.size());
    for (size_t i = 0; i < tris.size(); ++i) {
        std::array<uint32_t, 3> s = tris[i];
        std::sort(s.begin(), s.end());
        keys[i] = {s, i};
    }
    std::sort(keys.begin(), keys.end());
    std::vector<char> drop(tris.size(), 0);
    for (size_t i = 1; i < keys.size(); ++i)
        if (keys[i].first == keys[i - 1].first) drop[keys[i].second] = 1;
    d.tri.reserve(3 * tris.size());
    for (size_t i = 0; i < tris.size(); ++i) {
        if (drop[i]) {
            d.duplicate.push_back(rows[i]);
            continue;
        }
        d.tri.insert(d.tri.end(), tris[i].begin(), tris[i].end());
        d.rows.push_back(rows[i]);
    }
}

// decimal text of a little-endian two's complement integer of `limbs` words
std::string big_decimal(const uint64_t* x, int limbs) {
    std::vector<uint64_t> m(x, x + limbs);
    bool neg = limbs > 0 && (m[limbs - 1] >> 63);
    if (neg) {  // magnitude = ~x + 1
        unsigned carry = 1;
        for (int i = 0; i < limbs; ++i) {
            unsigned __int128 v = (unsigned __int128)(~m[i]) + carry;
            m[i] = (uint64_t)v;
            carry = (unsigned)(v >> 64);
        }
    }
    const uint64_t base = 10000000000000000000ull;  // 10^19
    std::vector<uint64_t> parts;
    for (;;) {
        bool zero = true;
        for (uint64_t w : m) zero = zero && w == 0;
        if (zero) break;
        unsigned __int128 rem = 0;
        for (int i = limbs - 1; i >= 0; --i) {
            unsigned __int128 cur = (rem << 64) | m[i];
            m[i] = (uint64_t)(cur / base);
            rem = cur % base;
        }
        parts.push_back((uint64_t)rem);
    }
    std::string s = neg ? "-" : "";
    if (parts.empty()) return "0";
    char b[24];
    snprintf(b, sizeof(b), "%llu", (unsigned long long)parts.back());
    s += b;
    for (size_t i = parts.size() - 1; i-- > 0;) {
        snprintf(b, sizeof(b), "%019llu", (unsigned long long)parts[i]);
        s += b;
    }
    return s;
}

struct MeshView {
    const double* xyz;
    uint32_t nv;
    const uint32_t* tri;
    uint32_t nt;
};

bool check_mesh(const MeshView& m, const char* name) {
    if ((m.nv && !m.xyz) || (m.nt && !m.tri)) {
        g_err = std::string("mesh ") + name + ": null array";
        return false;
    }
    return true;
}

const char kPairsHeader[] =
    "FaceA,FaceB,"
    "VA0_Idx,VA1_Idx,VA2_Idx,"
    "VB0_Idx,VB1_Idx,VB2_Idx,"
    "A_V0_X,A_V0_Y,A_V0_Z,A_V1_X,A_V1_Y,A_V1_Z,A_V2_X,A_V2_Y,A_V2_Z,"
    "B_V0_X,B_V0_Y,B_V0_Z,B_V1_X,B_V1_Y,B_V1_Z,B_V2_X,B_V2_Y,B_V2_Z\n";

// one record row (IntersectionHistory.h:148-158), without the line end
bool put_record(std::string& o, uint32_t fa, uint32_t fb, const MeshView& A, const MeshView& B) {
    if (fa >= A.nt || fb >= B.nt) {
        g_err = "pair (" + std::to_string(fa) + ", " + std::to_string(fb) + ") names a face that does not exist";
        return false;
    }
    const uint32_t* ta = A.tri + 3 * (size_t)fa;
    const uint32_t* tb = B.tri + 3 * (size_t)fb;
    for (int k = 0; k < 3; ++k)
        if (ta[k] >= A.nv || tb[k] >= B.nv) {
            g_err = "triangle index out of range";
            return false;
        }
    put_u64(o, fa);
    o += ',';
    put_u64(o, fb);
    for (int k = 0; k < 3; ++k) { o += ','; put_u64(o, ta[k]); }
    for (int k = 0; k < 3; ++k) { o += ','; put_u64(o, tb[k]); }
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) { o += ','; put_double(o, A.xyz[3 * (size_t)ta[k] + c]); }
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) { o += ','; put_double(o, B.xyz[3 * (size_t)tb[k] + c]); }
    return true;
}

bool host_result(const gi_result* r) {
    if (!r) {
        g_err = "null result";
        return false;
    }
    if (r->on_device) {
        g_err = "result arrays are in device memory (use output_on_device = 0)";
        return false;
    }
    if ((r->n_pairs && (!r->pairs || !r->seg_offsets)) || (r->n_segments && (!r->seg_exact || !r->seg_f64))) {
        g_err = "result without segments";
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

const char* gi_meshio_last_error(void) { return g_err.c_str(); }

int gi_off_read(const char* path, gi_off_mesh* out, gi_mesh_data** owner) {
    if (!path || !out || !owner) return arg_fail("null argument");
    *owner = nullptr;
    memset(out, 0, sizeof(*out));
    std::string text;
    if (!read_file(path, text)) return io_fail(std::string("cannot read ") + path);
    LineReader lr(text);
    std::string line;
    if (!lr.next(line)) return io_fail("empty file");
    std::vector<std::string> t = tokens(line);
    const std::string& kw = t[0];
    if (kw.size() < 3 || kw.compare(kw.size() - 3, 3, "OFF") != 0)
        return io_fail("not an OFF file (header '" + kw + "')");
    if (kw.find('4') != std::string::npos || kw.find('n') != std::string::npos)
        return io_fail("homogeneous or n-dimensional OFF is not supported");
    if (t.size() > 1 && t[1] == "BINARY") return io_fail("binary OFF is not supported");
    std::vector<std::string> counts(t.begin() + 1, t.end());
    if (counts.empty()) {
        if (!lr.next(line)) return io_fail("missing element counts");
        counts = tokens(line);
    }
    uint64_t nv = 0, nf = 0;
    if (counts.size() < 2 || !parse_u64(counts[0], nv) || !parse_u64(counts[1], nf))
        return io_fail("bad element counts line");
    if (nv >= (1ull << 32) || nf >= (1ull << 32)) return io_fail("more than 2^32 - 1 vertices or faces");
    gi_mesh_data* d = new gi_mesh_data();
    d->xyz.resize(3 * nv);
    for (uint64_t i = 0; i < nv; ++i) {
        if (!lr.next(line)) {
            delete d;
            return io_fail("file ends inside the vertex list");
        }
        t = tokens(line);
        // extra values (colours, normals, texture coordinates) are ignored
        if (t.size() < 3 || !parse_double(t[0], d->xyz[3 * i]) || !parse_double(t[1], d->xyz[3 * i + 1]) ||
            !parse_double(t[2], d->xyz[3 * i + 2])) {
            delete d;
            return io_fail("bad vertex on line " + std::to_string(lr.lineno));
        }
    }
    std::vector<std::vector<uint32_t>> polys(nf);
    for (uint64_t r = 0; r < nf; ++r) {
        if (!lr.next(line)) {
            delete d;
            return io_fail("file ends inside the face list");
        }
        t = tokens(line);
        uint64_t k = 0;
        if (t.empty() || !parse_u64(t[0], k) || k > t.size() - 1) {
            delete d;
            return io_fail("bad face on line " + std::to_string(lr.lineno));
        }
        polys[r].resize(k);
        for (uint64_t j = 0; j < k; ++j) {
            uint64_t v = 0;
            if (!parse_u64(t[1 + j], v) || v >= nv) {  // extra values after the indices (colours) are ignored
                delete d;
                return io_fail("bad or out-of-range vertex index on line " + std::to_string(lr.lineno));
            }
            polys[r][j] = (uint32_t)v;
        }
    }
    uint64_t extra = 0;
    triangulate(polys, *d, extra);
    out->soup_points = nv;
    out->soup_polygons = nf;
    out->extra_fan_triangles = extra;
    out->n_dropped_degenerate = d->degenerate.size();
    out->n_dropped_duplicate = d->duplicate.size();
    out->nv = (uint32_t)nv;
    out->nt = (uint32_t)d->rows.size();
    out->xyz = d->xyz.data();
    out->tri = d->tri.data();
    out->face_to_row = d->rows.data();
    out->dropped_degenerate = d->degenerate.data();
    out->dropped_duplicate = d->duplicate.data();
    *owner = d;
    return GI_OK;
}

int gi_off_free(gi_mesh_data* owner) {
    if (!owner) return GI_ERR_INVALID_ARGUMENT;
    delete owner;
    return GI_OK;
}

int gi_off_write(const char* path, const double* xyz, uint32_t nv, const uint32_t* tri, uint32_t nt) {
    if (!path) return arg_fail("null path");
    MeshView m{xyz, nv, tri, nt};
    if (!check_mesh(m, "")) return GI_ERR_INVALID_ARGUMENT;
    std::string o = "OFF\n";
    put_u64(o, nv);
    o += ' ';
    put_u64(o, nt);
    o += " 0\n";
    for (uint64_t i = 0; i < nv; ++i) {
        put_double(o, xyz[3 * i]);
        o += ' ';
        put_double(o, xyz[3 * i + 1]);
        o += ' ';
        put_double(o, xyz[3 * i + 2]);
        o += '\n';
    }
    for (uint64_t i = 0; i < nt; ++i) {
        o += "3";
        for (int k = 0; k < 3; ++k) {
            if (tri[3 * i + k] >= nv) return arg_fail("triangle index out of range");
            o += ' ';
            put_u64(o, tri[3 * i + k]);
        }
        o += '\n';
    }
    return write_file(path, o) ? GI_OK : io_fail(std::string("cannot write ") + path);
}

int gi_pairs_csv_write(const char* path, const uint32_t* pairs, uint64_t n, const double* xyzA, uint32_t nvA,
                       const uint32_t* triA, uint32_t ntA, const double* xyzB, uint32_t nvB, const uint32_t* triB,
                       uint32_t ntB) {
    if (!path || (n && !pairs)) return arg_fail("null argument");
    MeshView A{xyzA, nvA, triA, ntA}, B{xyzB, nvB, triB, ntB};
    if (!check_mesh(A, "A") || !check_mesh(B, "B")) return GI_ERR_INVALID_ARGUMENT;
    std::string o = kPairsHeader;
    for (uint64_t i = 0; i < n; ++i) {
        if (!put_record(o, pairs[2 * i], pairs[2 * i + 1], A, B)) return GI_ERR_INVALID_ARGUMENT;
        o += '\n';
    }
    return write_file(path, o) ? GI_OK : io_fail(std::string("cannot write ") + path);
}

int gi_pairs_diff_csv_write(const char* path, const uint32_t* pairs1, uint64_t n1, uint64_t id1,
                            const uint32_t* pairs2, uint64_t n2, uint64_t id2, const char* filter, const double* xyzA,
                            uint32_t nvA, const uint32_t* triA, uint32_t ntA, const double* xyzB, uint32_t nvB,
                            const uint32_t* triB, uint32_t ntB) {
    if (!path || !filter || (n1 && !pairs1) || (n2 && !pairs2)) return arg_fail("null argument");
    MeshView A{xyzA, nvA, triA, ntA}, B{xyzB, nvB, triB, ntB};
    if (!check_mesh(A, "A") || !check_mesh(B, "B")) return GI_ERR_INVALID_ARGUMENT;
    // records compare by (faceA, faceB), as std::set<IntersectingTriangleRecord> does
    std::set<std::pair<uint32_t, uint32_t>> s1, s2;
    for (uint64_t i = 0; i < n1; ++i) s1.insert({pairs1[2 * i], pairs1[2 * i + 1]});
    for (uint64_t i = 0; i < n2; ++i) s2.insert({pairs2[2 * i], pairs2[2 * i + 1]});
    std::vector<std::pair<uint32_t, uint32_t>> only1, only2, both;
    for (const auto& r : s1) (s2.count(r) ? both : only1).push_back(r);
    for (const auto& r : s2)
        if (!s1.count(r)) only2.push_back(r);
    std::string f = filter;
    std::transform(f.begin(), f.end(), f.begin(), [](unsigned char c) { return (char)tolower(c); });
    bool w1 = f == "all" || f == "only1" || f == "1" || f == "exclusive1";
    bool w2 = f == "all" || f == "only2" || f == "2" || f == "exclusive2";
    bool wb = f == "all" || f == "both";
    std::string o = std::string("DiffCategory,") + kPairsHeader;
    auto rows = [&](const std::string& cat, const std::vector<std::pair<uint32_t, uint32_t>>& v) {
        for (const auto& r : v) {
            o += cat;
            o += ',';
            if (!put_record(o, r.first, r.second, A, B)) return false;
            o += '\n';
        }
        return true;
    };
    if (w1 && !rows("ONLY_IN_" + std::to_string(id1), only1)) return GI_ERR_INVALID_ARGUMENT;
    if (w2 && !rows("ONLY_IN_" + std::to_string(id2), only2)) return GI_ERR_INVALID_ARGUMENT;
    if (wb && !rows("IN_BOTH", both)) return GI_ERR_INVALID_ARGUMENT;
    return write_file(path, o) ? GI_OK : io_fail(std::string("cannot write ") + path);
}

int gi_segments_csv_write(const char* path, const gi_result* res) {
    if (!path) return arg_fail("null path");
    if (!host_result(res)) return GI_ERR_INVALID_ARGUMENT;
    const int L = res->exact_limbs;
    std::string o = "pair,face_a,face_b,segment,x0,y0,z0,x1,y1,z1,X0,Y0,Z0,W0,X1,Y1,Z1,W1,exp2\n";
    for (uint64_t i = 0; i < res->n_pairs; ++i)
        for (uint32_t s = res->seg_offsets[i]; s < res->seg_offsets[i + 1]; ++s) {
            put_u64(o, i);
            o += ',';
            put_u64(o, res->pairs[2 * i]);
            o += ',';
            put_u64(o, res->pairs[2 * i + 1]);
            o += ',';
            put_u64(o, s - res->seg_offsets[i]);
            for (int k = 0; k < 6; ++k) {
                o += ',';
                put_double(o, res->seg_f64[6 * (size_t)s + k]);
            }
            for (int e = 0; e < 8; ++e) {
                o += ',';
                o += big_decimal(res->seg_exact + ((size_t)s * 8 + e) * L, L);
            }
            o += ',';
            o += std::to_string(res->seg_exp2 ? res->seg_exp2[s] : 0);
            o += '\n';
        }
    return write_file(path, o) ? GI_OK : io_fail(std::string("cannot write ") + path);
}

int gi_segments_obj_write(const char* path, const gi_result* res) {
    if (!path) return arg_fail("null path");
    if (!host_result(res)) return GI_ERR_INVALID_ARGUMENT;
    std::string o = "# intersection curve: one polyline segment per exact segment (endpoints rounded to double)\n";
    for (uint64_t s = 0; s < res->n_segments; ++s)
        for (int e = 0; e < 2; ++e) {
            o += "v";
            for (int k = 0; k < 3; ++k) {
                o += ' ';
                put_double(o, res->seg_f64[6 * s + 3 * e + k]);
            }
            o += '\n';
        }
    for (uint64_t s = 0; s < res->n_segments; ++s) {
        o += "l ";
        put_u64(o, 2 * s + 1);
        o += ' ';
        put_u64(o, 2 * s + 2);
        o += '\n';
    }
    return write_file(path, o) ? GI_OK : io_fail(std::string("cannot write ") + path);
}

}  // extern "C"
