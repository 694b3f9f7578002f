// bjdata.cpp -- BJData writers for .bnii and _detp.jdb (see bjdata.h).
#include "bjdata.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "json.h"

namespace umcx {

namespace {

char int_marker(int64_t v) {
    if (v >= -128 && v <= 127) return 'i';
    if (v >= 0 && v <= 255) return 'U';
    if (v >= -32768 && v <= 32767) return 'I';
    if (v >= 0 && v <= 65535) return 'u';
    if (v >= INT32_MIN && v <= INT32_MAX) return 'l';
    if (v >= 0 && v <= (int64_t)UINT32_MAX) return 'm';
    return 'L';
}

void put_le(std::string& out, uint64_t v, int nbytes) {
    for (int i = 0; i < nbytes; i++) out += (char)((v >> (8 * i)) & 0xFF);
}

void int_payload(std::string& out, char m, int64_t v) {
    int n = (m == 'i' || m == 'U') ? 1 : (m == 'I' || m == 'u') ? 2 : (m == 'l' || m == 'm') ? 4 : 8;
    put_le(out, (uint64_t)v, n);
}

void put_int(std::string& out, int64_t v) {
    char m = int_marker(v);
    out += m;
    int_payload(out, m, v);
}

void put_key(std::string& out, const std::string& k) {
    put_int(out, (int64_t)k.size());
    out += k;
}

void put_double(std::string& out, double d) {
    out += 'D';
    uint64_t u;
    memcpy(&u, &d, 8);
    put_le(out, u, 8);
}

// integer array: typed ('$') when all elements share one marker
void put_int_array(std::string& out, const std::vector<int64_t>& v) {
    out += '[';
    bool same = !v.empty();
    char m0 = v.empty() ? 'i' : int_marker(v[0]);
    for (int64_t x : v) same = same && int_marker(x) == m0;
    if (same) {
        out += '$';
        out += m0;
    }
    out += '#';
    put_int(out, (int64_t)v.size());
    for (int64_t x : v) {
        if (same) int_payload(out, m0, x);
        else put_int(out, x);
    }
}

void put_ndarray(std::string& out, char type, const std::vector<int64_t>& dims, const void* data, size_t bytes) {
    out += '[';
    out += '$';
    out += type;
    out += '#';
    put_int_array(out, dims);
    out.append((const char*)data, bytes);
}

void save(const std::string& path, const std::string& head, const void* payload, size_t bytes) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot open output file " + path);
    bool ok = fwrite(head.data(), 1, head.size(), f) == head.size();
    if (bytes) ok = ok && fwrite(payload, 1, bytes, f) == bytes;
    ok = (fclose(f) == 0) && ok;
    if (!ok) throw std::runtime_error("failed to write " + path);
}

void write_bnii(const std::string& path, char type, const void* data, size_t elem, uint32_t nx, uint32_t ny, uint32_t nz,
                uint32_t nt) {
    std::vector<int64_t> dims = {nx, ny, nz, nt};
    std::string h;
    h += '{'; h += '#'; put_int(h, 2);
    put_key(h, "NIFTIHeader");
    h += '{'; h += '#'; put_int(h, 1);
    put_key(h, "Dim");
    put_int_array(h, dims);
    put_key(h, "NIFTIData");
    // the JData annotation object is written in its N-D array form
    size_t n = (size_t)nx * ny * nz * nt;
    std::string nd;
    nd += '['; nd += '$'; nd += type; nd += '#';
    put_int_array(nd, dims);
    h += nd;
    save(path, h, data, n * elem);
}

}  // namespace

std::string bj_int(int64_t v) {
    std::string s;
    put_int(s, v);
    return s;
}

void write_bnii_float(const std::string& path, const float* data, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nt) {
    write_bnii(path, 'd', data, 4, nx, ny, nz, nt);
}

void write_bnii_int32(const std::string& path, const int32_t* data, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nt) {
    write_bnii(path, 'l', data, 4, nx, ny, nz, nt);
}

void write_detp(const std::string& path, const DetpInfo& info, const float* rows) {
    std::string h;
    h += '{'; h += '#'; put_int(h, 1);
    put_key(h, "MCXData");
    h += '{'; h += '#'; put_int(h, 2);
    put_key(h, "Info");
    h += '{'; h += '#'; put_int(h, 8);
    put_key(h, "Version"); put_int(h, 1);
    put_key(h, "MediaNum"); put_int(h, info.medianum);
    put_key(h, "DetNum"); put_int(h, info.detnum);
    put_key(h, "ColumnNum"); put_int(h, info.columns);
    put_key(h, "TotalPhoton"); put_int(h, info.columns);
    put_key(h, "DetectedPhoton"); put_int(h, info.detected);
    put_key(h, "SavedPhoton"); put_int(h, info.saved);
    put_key(h, "LengthUnit"); put_double(h, (double)info.lengthunit);
    put_key(h, "PhotonRawData");
    size_t bytes = (size_t)info.saved * (size_t)info.columns * sizeof(float);
    std::vector<int64_t> dims = {(int64_t)info.saved, (int64_t)info.columns};
    std::string nd;
    put_ndarray(nd, 'd', dims, nullptr, 0);
    h += nd;
    save(path, h, rows, bytes);
}

// ---------------- reader (json::from_bjdata of src/umcx.cpp:477) ----------------
namespace {

class BjReader {
public:
    BjReader(const uint8_t* p, size_t n) : p_(p), n_(n) {}

    Json document() {
        Json v = value(marker(), 0);
        while (pos_ < n_ && p_[pos_] == 'N') pos_++;
        if (pos_ != n_) fail("bytes left after the top-level value");
        return v;
    }

private:
    const uint8_t* p_;
    size_t n_;
    size_t pos_ = 0;
    static constexpr int kMaxDepth = 512;

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("BJData input: " + what + " at byte " + std::to_string(pos_));
    }
    void need(size_t k) const {
        if (k > n_ - pos_) fail("unexpected end of buffer");
    }
    uint8_t byte() {
        need(1);
        return p_[pos_++];
    }
    template <typename T>
    T le() {   // BJData scalars are little-endian, as is the x86-64 host
        need(sizeof(T));
        T v;
        memcpy(&v, p_ + pos_, sizeof(T));
        pos_ += sizeof(T);
        return v;
    }
    uint8_t marker() {   // next marker, no-ops skipped
        uint8_t m;
        do m = byte();
        while (m == 'N');
        return m;
    }
    uint8_t peek() {
        while (pos_ < n_ && p_[pos_] == 'N') pos_++;
        need(1);
        return p_[pos_];
    }

    // bytes of a fixed-size scalar payload; 0 for other markers
    static size_t width(uint8_t m) {
        switch (m) {
        case 'i': case 'U': case 'C': case 'B': return 1;
        case 'I': case 'u': case 'h': return 2;
        case 'l': case 'm': case 'd': return 4;
        case 'L': case 'M': case 'D': return 8;
        default: return 0;
        }
    }

    static double half(uint16_t h) {
        const int e = (h >> 10) & 0x1F, f = h & 0x3FF;
        double v = e == 0 ? std::ldexp((double)f, -24)
                          : e == 31 ? (f ? NAN : INFINITY) : std::ldexp((double)(f | 0x400), e - 25);
        return (h & 0x8000) ? -v : v;
    }

    // an integer-marked length or count
    uint64_t length(uint8_t m) {
        switch (m) {
        case 'i': { int8_t v = le<int8_t>(); if (v < 0) fail("negative length"); return (uint64_t)v; }
        case 'U': return le<uint8_t>();
        case 'I': { int16_t v = le<int16_t>(); if (v < 0) fail("negative length"); return (uint64_t)v; }
        case 'u': return le<uint16_t>();
        case 'l': { int32_t v = le<int32_t>(); if (v < 0) fail("negative length"); return (uint64_t)v; }
        case 'm': return le<uint32_t>();
        case 'L': { int64_t v = le<int64_t>(); if (v < 0) fail("negative length"); return (uint64_t)v; }
        case 'M': return le<uint64_t>();
        default: fail(std::string("expected an integer length marker, found '") + (char)m + "'");
        }
    }

    std::string str(uint8_t len_marker) {
        const uint64_t len = length(len_marker);
        need(len);
        std::string s((const char*)p_ + pos_, (size_t)len);
        pos_ += len;
        return s;
    }

    Json scalar(uint8_t m) {
        switch (m) {
        case 'Z': return Json();
        case 'T': return Json(true);
        case 'F': return Json(false);
        case 'i': return Json((long long)le<int8_t>());
        case 'U': return Json((unsigned long long)le<uint8_t>());
        case 'I': return Json((long long)le<int16_t>());
        case 'u': return Json((unsigned long long)le<uint16_t>());
        case 'l': return Json((long long)le<int32_t>());
        case 'm': return Json((unsigned long long)le<uint32_t>());
        case 'L': return Json((long long)le<int64_t>());
        case 'M': return Json((unsigned long long)le<uint64_t>());
        case 'h': return Json(half(le<uint16_t>()));
        case 'd': return Json((double)le<float>());
        case 'D': return Json(le<double>());
        case 'B': return Json((unsigned long long)le<uint8_t>());
        case 'C': return Json(std::string(1, (char)byte()));
        case 'S': return Json(str(marker()));
        case 'H': {   // high-precision number: its decimal text
            const std::string t = str(marker());
            try {
                return Json::parse(t);
            } catch (const std::exception&) {
                fail("malformed high-precision number");
            }
        }
        default: fail(std::string("unknown marker '") + (char)m + "'");
        }
    }

    // '#' count: an integer, or (arrays only) an N-D dimension vector
    uint64_t count(bool nd_allowed, std::vector<uint64_t>& dims) {
        const uint8_t m = marker();
        if (m != '[') return length(m);
        if (!nd_allowed) fail("an N-D dimension vector needs a typed array");
        Json d = container('[', kMaxDepth - 1, true);
        if (!d.is_array()) fail("malformed N-D dimension vector");
        for (size_t i = 0; i < d.size(); i++) {
            const Json& v = d.at(i);
            if (!(v.type() == Json::Type::Int || v.type() == Json::Type::UInt) || v.as_int() < 0)
                fail("N-D dimensions must be non-negative integers");
            dims.push_back((uint64_t)v.as_int());
        }
        if (dims.empty()) return 0;
        uint64_t n = 1;
        for (uint64_t x : dims) {
            if (x == 0) return 0;
            if (n > (n_ - pos_) / x + 1) fail("N-D array larger than the buffer");
            n *= x;
        }
        return n;
    }

    static const char* nd_type_name(uint8_t t) {
        switch (t) {
        case 'B': return "byte";
        case 'C': return "char";
        case 'D': return "double";
        case 'I': return "int16";
        case 'L': return "int64";
        case 'M': return "uint64";
        case 'U': return "uint8";
        case 'd': return "single";
        case 'i': return "int8";
        case 'l': return "int32";
        case 'm': return "uint32";
        case 'u': return "uint16";
        default: return nullptr;
        }
    }

    Json container(uint8_t open, int depth, bool dims_vector = false) {
        if (depth >= kMaxDepth) fail("containers nested too deeply");
        const bool is_obj = open == '{';
        uint8_t type = 0;
        bool counted = false;
        uint64_t n = 0;
        std::vector<uint64_t> dims;
        uint8_t m = peek();
        if (m == '$') {
            pos_++;
            type = byte();
            if (type == '{') fail("structure-of-arrays containers are not supported");
            if (type == 'F' || type == 'H' || type == 'N' || type == 'S' || type == 'T' || type == 'Z' || type == '[')
                fail(std::string("'") + (char)type + "' is not a permitted optimized container type");
            if (!width(type)) fail(std::string("unknown optimized container type '") + (char)type + "'");
            if (peek() != '#') fail("expected '#' after a container type");
            pos_++;
            n = count(!is_obj && !dims_vector, dims);
            counted = true;
        } else if (m == '#') {
            pos_++;
            n = count(false, dims);
            counted = true;
        }
        if (counted && type && n > (n_ - pos_) / width(type)) fail("container larger than the buffer");
        if (counted && !type && n > n_ - pos_) fail("container larger than the buffer");

        // N-D array: dimension vectors of length 1, or 1 x n, are plain arrays
        const bool nd = dims.size() > 1 && !(dims.size() == 2 && dims[0] == 1);
        if (nd && n > 0) {
            const char* tname = nd_type_name(type);
            if (!tname) fail("N-D array of an unsupported element type");
            Json o = Json::object();
            Json size = Json::array();
            for (uint64_t x : dims) size.push_back(Json((unsigned long long)x));
            o["_ArraySize_"] = std::move(size);
            o["_ArrayType_"] = Json(tname);
            const uint8_t et = (type == 'C' || type == 'B') ? 'U' : type;
            Json data = Json::array();
            for (uint64_t i = 0; i < n; i++) data.push_back(scalar(et));
            o["_ArrayData_"] = std::move(data);
            return o;
        }

        Json out = is_obj ? Json::object() : Json::array();
        const uint8_t close = is_obj ? '}' : ']';
        for (uint64_t i = 0; counted ? i < n : true; i++) {
            if (!counted && peek() == close) {
                pos_++;
                break;
            }
            std::string key;
            if (is_obj) key = str(marker());
            Json v = type ? scalar(type) : value(marker(), depth + 1);
            if (is_obj) out[key] = std::move(v);
            else out.push_back(std::move(v));
        }
        return out;
    }

    Json value(uint8_t m, int depth) {
        if (m == '[' || m == '{') return container(m, depth);
        return scalar(m);
    }
};

}  // namespace

Json bjdata_decode(const uint8_t* buf, size_t n) {
    if (!buf || n == 0) throw std::runtime_error("BJData input: empty buffer");
    return BjReader(buf, n).document();
}

}  // namespace umcx
