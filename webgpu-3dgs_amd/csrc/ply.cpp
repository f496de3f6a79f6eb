// ply.cpp — native Gaussian-splat PLY ingest (host code in libgs3d.so).
//
// Restates the reference loader src/workers/SceneWorker.ts:
//   header parse and validation        :146-272
//   ascii / binary value readers       :274-333, :508-549
//   element walk, row -> splat          :336-372, :407-464
//   f_rest ordering by index            :375-388
//   SH degree from the f_rest count     :501-506
// and the upload of SceneSyncer.setSplatData (src/SceneSyncer.ts:25-43).
//
// Activations are evaluated in f64 and stored as f32, as the reference does
// with JS numbers written into Float32Arrays: colour clamp01(0.5 + C0 f_dc),
// alpha clamp01(1 / (1 + exp(-opacity))), scale exp(s), rotation divided by
// its 4-norm (1 when the norm is 0). The 4-norm follows JS Math.hypot: the
// magnitudes are divided by the largest one and summed with Kahan
// compensation, then sqrt(sum) * largest.
//
// Two readers share the row rules. The sequential reader follows the
// reference value by value (ascii bodies, several vertex elements, types the
// reference rejects). The parallel reader takes binary bodies with one vertex
// element of supported types: rows have a fixed stride, so threads take row
// ranges; the first error in stream order is the one reported, as a
// sequential reader would throw it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>
#include <vector>

#include "gs3d.h"

namespace {

constexpr double kShC0 = 0.28209479177387814;
constexpr size_t kMaxHeaderBytes = 1u << 20;

enum Kind { K_I8, K_U8, K_I16, K_U16, K_I32, K_U32, K_F32, K_F64, K_BAD };

struct Prop {
    std::string type, name;
    Kind kind = K_BAD;
    int size = 0;
};

struct Elem {
    std::string name;
    uint64_t count = 0;
    std::vector<Prop> props;
    uint64_t row_bytes = 0;  // binary, when every type is supported
    int first_bad = -1;      // index of the first unsupported type, -1 if none
};

struct Header {
    int format = -1;  // GS3D_PLY_ASCII / _BINARY_LE / _BINARY_BE
    std::vector<Elem> elems;
    int vertex = -1;
    size_t header_bytes = 0;
};

struct Failure {
    std::string msg;
};

Failure failf(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return Failure{buf};
}

bool is_ws(unsigned char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

Kind kind_of(const std::string& t, int* size) {
    struct {
        const char* a;
        const char* b;
        Kind k;
        int s;
    } static const table[] = {
        {"char", "int8", K_I8, 1},     {"uchar", "uint8", K_U8, 1},   {"short", "int16", K_I16, 2},
        {"ushort", "uint16", K_U16, 2}, {"int", "int32", K_I32, 4},   {"uint", "uint32", K_U32, 4},
        {"float", "float32", K_F32, 4}, {"double", "float64", K_F64, 8},
    };
    for (const auto& e : table)
        if (t == e.a || t == e.b) {
            *size = e.s;
            return e.k;
        }
    *size = 0;
    return K_BAD;
}

// A JS Number(token) for the forms PLY bodies and headers use: decimal
// literals with optional sign, fraction and exponent, and unsigned 0x / 0o /
// 0b integers. Anything else reads as NaN.
double js_number(const char* s, size_t n) {
    const double nan = std::nan("");
    if (n == 0) return 0.0;  // Number("") is 0; only reachable for empty header tokens
    std::string t(s, n);
    if (n > 2 && t[0] == '0' && (t[1] == 'x' || t[1] == 'X' || t[1] == 'o' || t[1] == 'O' || t[1] == 'b' || t[1] == 'B')) {
        const int base = (t[1] == 'x' || t[1] == 'X') ? 16 : (t[1] == 'o' || t[1] == 'O') ? 8 : 2;
        double v = 0.0;
        for (size_t i = 2; i < n; i++) {
            int d;
            const char c = t[i];
            if (c >= '0' && c <= '9') d = c - '0';
            else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
            else if (c >= 'A' && c <= 'F') d = c - 'A' + 10;
            else return nan;
            if (d >= base) return nan;
            v = v * base + d;
        }
        return v;
    }
    size_t i = 0;
    if (t[i] == '+' || t[i] == '-') i++;
    if (t.compare(i, std::string::npos, "Infinity") == 0) return t[0] == '-' ? -INFINITY : INFINITY;
    size_t digits = 0;
    while (i < n && t[i] >= '0' && t[i] <= '9') i++, digits++;
    if (i < n && t[i] == '.') {
        i++;
        while (i < n && t[i] >= '0' && t[i] <= '9') i++, digits++;
    }
    if (digits == 0) return nan;
    if (i < n && (t[i] == 'e' || t[i] == 'E')) {
        i++;
        if (i < n && (t[i] == '+' || t[i] == '-')) i++;
        size_t ed = 0;
        while (i < n && t[i] >= '0' && t[i] <= '9') i++, ed++;
        if (ed == 0) return nan;
    }
    if (i != n) return nan;
    return std::strtod(t.c_str(), nullptr);
}

// JS Math.min / Math.max on finite numbers (-0 below +0).
double js_min(double a, double b) {
    if (a < b) return a;
    if (b < a) return b;
    return std::signbit(a) ? a : b;
}
double js_max(double a, double b) {
    if (a > b) return a;
    if (b > a) return b;
    return std::signbit(a) ? b : a;
}

double clamp01(double x) { return std::min(1.0, std::max(0.0, x)); }

double js_hypot4(double a, double b, double c, double d) {
    const double v[4] = {std::fabs(a), std::fabs(b), std::fabs(c), std::fabs(d)};
    double mx = 0.0;
    for (double x : v)
        if (x > mx) mx = x;
    if (mx == INFINITY) return INFINITY;
    if (mx == 0.0) return 0.0;
    double sum = 0.0, comp = 0.0;
    for (double x : v) {
        const double q = x / mx;
        const double summand = q * q - comp;
        const double pre = sum + summand;
        comp = (pre - sum) - summand;
        sum = pre;
    }
    return std::sqrt(sum) * mx;
}

// ---- header ------------------------------------------------------------------

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && is_ws((unsigned char)s[a])) a++;
    while (b > a && is_ws((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::vector<std::string> split_ws(const std::string& s) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && is_ws((unsigned char)s[i])) i++;
        size_t j = i;
        while (j < s.size() && !is_ws((unsigned char)s[j])) j++;
        if (j > i) out.emplace_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

bool parse_header(const uint8_t* bytes, size_t len, Header* h, Failure* f) {
    const size_t scan = std::min(len, kMaxHeaderBytes);
    static const char kEnd[] = "end_header";
    const void* hit = memmem(bytes, scan, kEnd, sizeof kEnd - 1);
    if (!hit) {
        *f = failf("Invalid PLY: missing end_header");
        return false;
    }
    size_t hb = (size_t)((const uint8_t*)hit - bytes) + (sizeof kEnd - 1);
    while (hb < len && (bytes[hb] == 10 || bytes[hb] == 13)) hb++;
    h->header_bytes = hb;
    // lines of the header, split on \r?\n, trimmed, empty ones dropped
    std::vector<std::string> lines;
    {
        size_t i = 0;
        while (i < hb) {
            size_t j = i;
            while (j < hb && bytes[j] != '\n') j++;
            std::string line((const char*)bytes + i, j - i);
            if (!line.empty() && line.back() == '\r') line.pop_back();
            line = trim(line);
            if (!line.empty()) lines.push_back(line);
            i = j + 1;
        }
    }
    if (lines.empty() || lines[0] != "ply") {
        *f = failf("Invalid PLY: file must start with `ply`");
        return false;
    }
    std::string format;
    bool have_format = false;
    for (size_t li = 1; li < lines.size(); li++) {
        const std::string& line = lines[li];
        if (line == "end_header") break;
        if (line.rfind("comment ", 0) == 0 || line.rfind("obj_info ", 0) == 0) continue;
        const std::vector<std::string> t = split_ws(line);
        if (t[0] == "format") {
            have_format = t.size() > 1;
            if (have_format) format = t[1];
        } else if (t[0] == "element") {
            Elem e;
            e.name = t.size() > 1 ? t[1] : "undefined";
            const double c = t.size() > 2 ? js_number(t[2].data(), t[2].size()) : std::nan("");
            if (!std::isfinite(c) || c < 0) {
                *f = failf("Invalid PLY element count for '%s'", e.name.c_str());
                return false;
            }
            if (c != std::floor(c) || c > 4294967295.0) {
                *f = failf("Unsupported PLY element count for '%s': %s", e.name.c_str(), t[2].c_str());
                return false;
            }
            e.count = (uint64_t)c;
            h->elems.push_back(e);
        } else if (t[0] == "property") {
            if (h->elems.empty()) {
                *f = failf("Invalid PLY: property declared before element");
                return false;
            }
            if (t.size() > 1 && t[1] == "list") {
                *f = failf("Unsupported PLY: list properties are not expected in Gaussian Splat PLY files");
                return false;
            }
            Prop p;
            p.type = t.size() > 1 ? t[1] : "undefined";
            p.name = t.size() > 2 ? t[2] : "undefined";
            p.kind = kind_of(p.type, &p.size);
            h->elems.back().props.push_back(p);
        }
    }
    if (!have_format) {
        *f = failf("Invalid PLY: missing format line");
        return false;
    }
    if (format == "ascii") h->format = GS3D_PLY_ASCII;
    else if (format == "binary_little_endian") h->format = GS3D_PLY_BINARY_LE;
    else if (format == "binary_big_endian") h->format = GS3D_PLY_BINARY_BE;
    else {
        *f = failf("Unsupported PLY format: %s", format.c_str());
        return false;
    }
    for (auto& e : h->elems) {
        e.row_bytes = 0;
        e.first_bad = -1;
        for (size_t k = 0; k < e.props.size(); k++) {
            if (e.props[k].kind == K_BAD && e.first_bad < 0) e.first_bad = (int)k;
            e.row_bytes += (uint64_t)e.props[k].size;
        }
    }
    for (size_t i = 0; i < h->elems.size(); i++)
        if (h->elems[i].name == "vertex") {
            h->vertex = (int)i;
            break;
        }
    if (h->vertex < 0) {
        *f = failf("Invalid Gaussian Splat PLY: missing vertex element");
        return false;
    }
    static const char* kRequired[] = {"x", "y", "z", "opacity", "scale_0", "scale_1", "scale_2",
                                      "rot_0", "rot_1", "rot_2", "rot_3", "f_dc_0", "f_dc_1", "f_dc_2"};
    const Elem& v = h->elems[h->vertex];
    for (const char* r : kRequired) {
        bool found = false;
        for (const auto& p : v.props) found = found || p.name == r;
        if (!found) {
            *f = failf("Invalid Gaussian Splat PLY: missing property '%s'", r);
            return false;
        }
    }
    return true;
}

bool is_f_rest(const std::string& n, long* idx) {
    if (n.compare(0, 7, "f_rest_") != 0 || n.size() == 7) return false;
    for (size_t i = 7; i < n.size(); i++)
        if (n[i] < '0' || n[i] > '9') return false;
    *idx = std::strtol(n.c_str() + 7, nullptr, 10);
    return true;
}

// f_rest names of the (first) vertex element sorted by index, stable for ties.
std::vector<std::string> f_rest_names(const Elem& v) {
    std::vector<std::pair<double, std::string>> r;
    for (const auto& p : v.props) {
        long idx;
        if (is_f_rest(p.name, &idx)) r.emplace_back(std::strtod(p.name.c_str() + 7, nullptr), p.name);
    }
    std::stable_sort(r.begin(), r.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<std::string> out;
    for (auto& e : r) out.push_back(e.second);
    return out;
}

int sh_degree_of(uint32_t f_rest) { return f_rest >= 45 ? 3 : f_rest >= 24 ? 2 : f_rest >= 9 ? 1 : 0; }

// ---- rows --------------------------------------------------------------------

// Where each value of a vertex row lives in one element's property list (the
// last property of a name wins, as in a JS object keyed by name).
struct RowLayout {
    int x, y, z, dc[3], op, sc[3], rot[4];
    std::vector<int> rest;
};

int last_index(const Elem& e, const std::string& name) {
    int idx = -1;
    for (size_t k = 0; k < e.props.size(); k++)
        if (e.props[k].name == name) idx = (int)k;
    return idx;
}

RowLayout layout_for(const Elem& e, const std::vector<std::string>& rest_names) {
    RowLayout L;
    L.x = last_index(e, "x");
    L.y = last_index(e, "y");
    L.z = last_index(e, "z");
    for (int c = 0; c < 3; c++) L.dc[c] = last_index(e, "f_dc_" + std::to_string(c));
    L.op = last_index(e, "opacity");
    for (int c = 0; c < 3; c++) L.sc[c] = last_index(e, "scale_" + std::to_string(c));
    for (int c = 0; c < 4; c++) L.rot[c] = last_index(e, "rot_" + std::to_string(c));
    for (const auto& n : rest_names) L.rest.push_back(last_index(e, n));
    return L;
}

struct Outputs {
    float *pos, *scale, *rot, *color, *dc, *rest;
    uint64_t n;        // rows of the first vertex element (output capacity)
    uint32_t f_rest;
};

struct Bounds {
    double lo[3] = {INFINITY, INFINITY, INFINITY};
    double hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void add(double x, double y, double z) {
        lo[0] = js_min(lo[0], x), lo[1] = js_min(lo[1], y), lo[2] = js_min(lo[2], z);
        hi[0] = js_max(hi[0], x), hi[1] = js_max(hi[1], y), hi[2] = js_max(hi[2], z);
    }
    void merge(const Bounds& o) {
        for (int i = 0; i < 3; i++) lo[i] = js_min(lo[i], o.lo[i]), hi[i] = js_max(hi[i], o.hi[i]);
    }
};

// writeGaussianSplat (SceneWorker.ts:407-464) for row `vi`. Returns false with
// the name of the first required value that is absent or not finite.
bool process_row(const double* row, const RowLayout& L, const std::vector<std::string>& rest_names,
                 uint64_t vi, const Outputs& o, Bounds& b, std::string* missing) {
    auto req = [&](int idx, const char* name, double* v) {
        if (idx < 0 || !std::isfinite(row[idx])) {
            *missing = name;
            return false;
        }
        *v = row[idx];
        return true;
    };
    double x, y, z, dc[3], op, sc[3], r[4];
    if (!req(L.x, "x", &x) || !req(L.y, "y", &y) || !req(L.z, "z", &z)) return false;
    const bool store = vi < o.n;  // rows past the first vertex element's count are not stored
    b.add(x, y, z);
    static const char* kDc[3] = {"f_dc_0", "f_dc_1", "f_dc_2"};
    static const char* kSc[3] = {"scale_0", "scale_1", "scale_2"};
    static const char* kRot[4] = {"rot_0", "rot_1", "rot_2", "rot_3"};
    for (int c = 0; c < 3; c++)
        if (!req(L.dc[c], kDc[c], &dc[c])) return false;
    if (!req(L.op, "opacity", &op)) return false;
    for (int c = 0; c < 3; c++)
        if (!req(L.sc[c], kSc[c], &sc[c])) return false;
    for (int c = 0; c < 4; c++)
        if (!req(L.rot[c], kRot[c], &r[c])) return false;
    const size_t q = (size_t)vi * 4;
    if (store) {
        o.pos[q] = (float)x, o.pos[q + 1] = (float)y, o.pos[q + 2] = (float)z, o.pos[q + 3] = 1.0f;
        for (int c = 0; c < 3; c++) {
            o.color[q + c] = (float)clamp01(0.5 + kShC0 * dc[c]);
            o.dc[(size_t)vi * 3 + c] = (float)dc[c];
        }
        o.color[q + 3] = (float)clamp01(1.0 / (1.0 + std::exp(-op)));
        for (int c = 0; c < 3; c++) o.scale[q + c] = (float)std::exp(sc[c]);
        o.scale[q + 3] = 0.0f;
        const double len = js_hypot4(r[0], r[1], r[2], r[3]);
        const double inv = len > 0.0 ? 1.0 / len : 1.0;
        for (int c = 0; c < 4; c++) o.rot[q + c] = (float)(r[c] * inv);
    }
    for (size_t k = 0; k < L.rest.size(); k++) {
        const int idx = L.rest[k];
        if (idx < 0 || !std::isfinite(row[idx])) {
            *missing = rest_names[k];
            return false;
        }
        if (store && o.rest) o.rest[(size_t)vi * o.f_rest + k] = (float)row[idx];
    }
    return true;
}

double load_scalar(const uint8_t* p, Kind k, bool swap) {
    uint8_t tmp[8];
    int n = 0;
    switch (k) {
        case K_I8: return (double)(int8_t)p[0];
        case K_U8: return (double)p[0];
        case K_I16: case K_U16: n = 2; break;
        case K_I32: case K_U32: case K_F32: n = 4; break;
        case K_F64: n = 8; break;
        default: return std::nan("");
    }
    for (int i = 0; i < n; i++) tmp[i] = swap ? p[n - 1 - i] : p[i];
    switch (k) {
        case K_I16: { int16_t v; memcpy(&v, tmp, 2); return v; }
        case K_U16: { uint16_t v; memcpy(&v, tmp, 2); return v; }
        case K_I32: { int32_t v; memcpy(&v, tmp, 4); return v; }
        case K_U32: { uint32_t v; memcpy(&v, tmp, 4); return v; }
        case K_F32: { float v; memcpy(&v, tmp, 4); return v; }
        default: { double v; memcpy(&v, tmp, 8); return v; }
    }
}

std::string row_error(const std::string& name) {
    return "Invalid Gaussian Splat PLY row: missing '" + name + "'";
}

const char* kRangeError = "Offset is outside the bounds of the DataView";

// The sequential reader: every element in order, value by value.
bool parse_sequential(const uint8_t* bytes, size_t len, const Header& h, const Outputs& o,
                      const std::vector<std::string>& rest_names, Bounds* bounds, Failure* f) {
    const bool ascii = h.format == GS3D_PLY_ASCII;
    const bool swap = h.format == GS3D_PLY_BINARY_BE;
    size_t off = h.header_bytes;
    // ascii: the body is trimmed and split on whitespace
    auto next_token = [&](const char** s, size_t* n) {
        while (off < len && is_ws(bytes[off])) off++;
        if (off >= len) return false;
        size_t j = off;
        while (j < len && !is_ws(bytes[j])) j++;
        *s = (const char*)bytes + off;
        *n = j - off;
        off = j;
        return true;
    };
    uint64_t vi = 0;
    std::vector<double> row;
    for (const Elem& e : h.elems) {
        const bool vertex = e.name == "vertex";
        const RowLayout L = vertex ? layout_for(e, rest_names) : RowLayout{};
        row.assign(e.props.size(), 0.0);
        for (uint64_t r = 0; r < e.count; r++) {
            for (size_t k = 0; k < e.props.size(); k++) {
                if (ascii) {
                    const char* s;
                    size_t n;
                    if (!next_token(&s, &n)) {
                        *f = failf("Unexpected end of ASCII PLY body");
                        return false;
                    }
                    const double v = js_number(s, n);
                    if (!std::isfinite(v)) {
                        *f = failf("Invalid ASCII PLY value '%.*s'", (int)std::min<size_t>(n, 200), s);
                        return false;
                    }
                    row[k] = v;
                } else {
                    const Prop& p = e.props[k];
                    if (p.kind == K_BAD) {
                        *f = failf("Unsupported PLY scalar type: %s", p.type.c_str());
                        return false;
                    }
                    if (off + (size_t)p.size > len) {
                        *f = failf("%s", kRangeError);
                        return false;
                    }
                    row[k] = load_scalar(bytes + off, p.kind, swap);
                    off += p.size;
                }
            }
            if (vertex) {
                std::string missing;
                if (!process_row(row.data(), L, rest_names, vi, o, *bounds, &missing)) {
                    *f = failf("%s", row_error(missing).c_str());
                    return false;
                }
                vi++;
            }
        }
    }
    return true;
}

// Non-vertex element of a binary body: its bytes are skipped, with the
// errors the value reader would raise on its first row.
bool skip_binary_element(const Elem& e, size_t len, size_t* off, Failure* f) {
    if (e.count == 0) return true;
    if (e.first_bad >= 0) {
        size_t before = 0;
        for (int k = 0; k < e.first_bad; k++) before += e.props[k].size;
        if (*off + before > len) {
            *f = failf("%s", kRangeError);
            return false;
        }
        *f = failf("Unsupported PLY scalar type: %s", e.props[e.first_bad].type.c_str());
        return false;
    }
    const unsigned __int128 need = (unsigned __int128)e.count * e.row_bytes;
    if ((unsigned __int128)*off + need > len) {
        *f = failf("%s", kRangeError);
        return false;
    }
    *off += (size_t)need;
    return true;
}

int thread_count(int requested, uint64_t rows) {
    int t = requested > 0 ? requested : (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
    const uint64_t per = 16384;  // rows per thread below which more threads do not pay
    t = (int)std::min<uint64_t>((uint64_t)t, std::max<uint64_t>(1, (rows + per - 1) / per));
    return std::max(1, t);
}

// The parallel reader: binary body, one vertex element of supported types.
bool parse_binary_parallel(const uint8_t* bytes, size_t len, const Header& h, const Outputs& o,
                           const std::vector<std::string>& rest_names, int threads, Bounds* bounds,
                           Failure* f) {
    const bool swap = h.format == GS3D_PLY_BINARY_BE;
    size_t off = h.header_bytes;
    for (size_t ei = 0; ei < h.elems.size(); ei++) {
        const Elem& e = h.elems[ei];
        if ((int)ei != h.vertex) {
            if (!skip_binary_element(e, len, &off, f)) return false;
            continue;
        }
        const uint64_t stride = e.row_bytes;
        const uint64_t complete = stride ? std::min<uint64_t>(e.count, (len - std::min(len, off)) / stride) : e.count;
        const RowLayout L = layout_for(e, rest_names);
        const int T = thread_count(threads, complete);
        std::vector<uint64_t> err_row(T, UINT64_MAX);
        std::vector<std::string> err_name(T);
        std::vector<Bounds> tb(T);
        auto work = [&](int t) {
            const uint64_t r0 = complete * t / T, r1 = complete * (t + 1) / T;
            std::vector<double> row(e.props.size());
            for (uint64_t r = r0; r < r1; r++) {
                const uint8_t* p = bytes + off + r * stride;
                for (size_t k = 0; k < e.props.size(); k++) {
                    row[k] = load_scalar(p, e.props[k].kind, swap);
                    p += e.props[k].size;
                }
                std::string missing;
                if (!process_row(row.data(), L, rest_names, r, o, tb[t], &missing)) {
                    err_row[t] = r;
                    err_name[t] = missing;
                    return;
                }
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < T; t++) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
        for (int t = 0; t < T; t++) {
            if (err_row[t] != UINT64_MAX) {  // ranges are in row order: the first one is the earliest
                *f = failf("%s", row_error(err_name[t]).c_str());
                return false;
            }
            bounds->merge(tb[t]);
        }
        if (complete < e.count) {
            *f = failf("%s", kRangeError);
            return false;
        }
        off += (size_t)(e.count * stride);
    }
    return true;
}

bool parallel_eligible(const Header& h) {
    if (h.format == GS3D_PLY_ASCII) return false;
    int vertices = 0;
    for (const auto& e : h.elems) vertices += e.name == "vertex";
    return vertices == 1 && h.elems[h.vertex].first_bad < 0;
}

void set_err(char* err, size_t err_bytes, const std::string& msg) {
    if (err && err_bytes) snprintf(err, err_bytes, "%s", msg.c_str());
}

void fill_info(const Header& h, const std::vector<std::string>& rest, gs3d_ply_info* info) {
    memset(info, 0, sizeof *info);
    info->splat_count = h.elems[h.vertex].count;
    info->f_rest_count = (uint32_t)rest.size();
    info->sh_degree = sh_degree_of((uint32_t)rest.size());
    info->format = h.format;
    info->header_bytes = (uint64_t)h.header_bytes;
    for (int i = 0; i < 3; i++) info->bounds_min[i] = INFINITY, info->bounds_max[i] = -INFINITY;
}

}  // namespace

// Defined in runtime.cpp: the context's last-error slot and device selection.
__attribute__((visibility("hidden"))) int gs3d_internal_fail(gs3d_ctx* c, int code, const char* msg);
__attribute__((visibility("hidden"))) int gs3d_internal_set_device(gs3d_ctx* c);

extern "C" {

int gs3d_ply_probe(const void* data, size_t bytes, gs3d_ply_info* info, char* err, size_t err_bytes) {
    if (!data || !info) return GS3D_E_INVALID_ARG;
    Header h;
    Failure f;
    if (!parse_header((const uint8_t*)data, bytes, &h, &f)) {
        set_err(err, err_bytes, f.msg);
        return GS3D_E_FORMAT;
    }
    fill_info(h, f_rest_names(h.elems[h.vertex]), info);
    return GS3D_OK;
}

int gs3d_ply_parse(const void* data, size_t bytes, float* positions, float* scales, float* rotations,
                   float* colors, float* f_dc, float* sh_rest, int threads, gs3d_ply_info* info,
                   char* err, size_t err_bytes) {
    if (!data || !info) return GS3D_E_INVALID_ARG;
    Header h;
    Failure f;
    const uint8_t* b = (const uint8_t*)data;
    if (!parse_header(b, bytes, &h, &f)) {
        set_err(err, err_bytes, f.msg);
        return GS3D_E_FORMAT;
    }
    const std::vector<std::string> rest = f_rest_names(h.elems[h.vertex]);
    fill_info(h, rest, info);
    const uint64_t n = info->splat_count;
    if (n && (!positions || !scales || !rotations || !colors || !f_dc || (!rest.empty() && !sh_rest)))
        return GS3D_E_INVALID_ARG;
    Outputs o{positions, scales, rotations, colors, f_dc, rest.empty() ? nullptr : sh_rest, n,
              (uint32_t)rest.size()};
    Bounds bounds;
    const bool ok = parallel_eligible(h) ? parse_binary_parallel(b, bytes, h, o, rest, threads, &bounds, &f)
                                         : parse_sequential(b, bytes, h, o, rest, &bounds, &f);
    if (!ok) {
        set_err(err, err_bytes, f.msg);
        return GS3D_E_FORMAT;
    }
    for (int i = 0; i < 3; i++) info->bounds_min[i] = bounds.lo[i], info->bounds_max[i] = bounds.hi[i];
    return GS3D_OK;
}

int gs3d_load_ply(gs3d_ctx* c, const char* path, int threads, gs3d_ply_info* info) {
    if (!c || !path || !info) return GS3D_E_INVALID_ARG;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) {
        char m[600];
        snprintf(m, sizeof m, "cannot open %s: %s", path, strerror(errno));
        return gs3d_internal_fail(c, GS3D_E_IO, m);
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        char m[600];
        snprintf(m, sizeof m, "cannot stat %s: %s", path, strerror(errno));
        return gs3d_internal_fail(c, GS3D_E_IO, m);
    }
    if (st.st_size == 0) {
        close(fd);
        return gs3d_internal_fail(c, GS3D_E_FORMAT, "Invalid PLY: missing end_header");
    }
    const size_t len = (size_t)st.st_size;
    void* map = mmap(nullptr, len, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
    close(fd);
    if (map == MAP_FAILED) return gs3d_internal_fail(c, GS3D_E_IO, "mmap of the PLY file failed");
    char err[512] = {0};
    int rc = gs3d_ply_probe(map, len, info, err, sizeof err);
    if (rc != GS3D_OK) {
        munmap(map, len);
        return gs3d_internal_fail(c, rc, err);
    }
    const uint64_t n = info->splat_count;
    const uint32_t F = info->f_rest_count;
    const int deg = info->sh_degree;
    const uint32_t per_ch = (uint32_t)((deg + 1) * (deg + 1) - 1);
    if (deg > 0 && F % 3 != 0) {
        munmap(map, len);
        snprintf(err, sizeof err, "%u f_rest properties do not split into 3 colour channels", F);
        return gs3d_internal_fail(c, GS3D_E_FORMAT, err);
    }
    rc = gs3d_internal_set_device(c);
    if (rc != GS3D_OK) {
        munmap(map, len);
        return rc;
    }
    // pinned staging: one block for all arrays, so the uploads are DMA copies
    const size_t vec4 = (size_t)n * 16, dcb = (size_t)n * 12, restb = (size_t)n * F * 4;
    const size_t total = 4 * vec4 + dcb + restb + 64;
    uint8_t* pinned = nullptr;
    if (hipHostMalloc((void**)&pinned, std::max<size_t>(total, 64), hipHostMallocDefault) != hipSuccess) {
        munmap(map, len);
        return gs3d_internal_fail(c, GS3D_E_OOM, "hipHostMalloc of the PLY staging buffer failed");
    }
    float* pos = (float*)pinned;
    float* sc = (float*)(pinned + vec4);
    float* rot = (float*)(pinned + 2 * vec4);
    float* col = (float*)(pinned + 3 * vec4);
    float* dc = (float*)(pinned + 4 * vec4);
    float* rest = F ? (float*)(pinned + 4 * vec4 + dcb) : nullptr;
    rc = gs3d_ply_parse(map, len, pos, sc, rot, col, dc, rest, threads, info, err, sizeof err);
    munmap(map, len);
    if (rc != GS3D_OK) {
        hipHostFree(pinned);
        return gs3d_internal_fail(c, rc, err);
    }
    // SH coefficients for the upload: all R, then G, then B, per_ch each. A
    // file with more f_rest per channel than the degree uses keeps the first
    // per_ch of each channel.
    std::vector<float> packed;
    const float* sh = rest;
    if (deg > 0 && F != 3 * per_ch) {
        const uint32_t in_ch = F / 3;
        packed.resize((size_t)n * 3 * per_ch);
        for (uint64_t i = 0; i < n; i++)
            for (uint32_t ch = 0; ch < 3; ch++)
                memcpy(&packed[(size_t)i * 3 * per_ch + ch * per_ch], rest + (size_t)i * F + ch * in_ch,
                       per_ch * sizeof(float));
        sh = packed.data();
    }
    rc = gs3d_set_splats(c, pos, sc, rot, col, deg > 0 ? dc : nullptr, deg > 0 ? sh : nullptr, deg, (uint32_t)n);
    hipHostFree(pinned);
    return rc;
}

}  // extern "C"
