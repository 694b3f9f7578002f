// bjdata.cpp -- BJData writers for .bnii and _detp.jdb (see bjdata.h).
#include "bjdata.h"

#include <cstdio>
#include <cstring>
#include <stdexcept>

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

}  // namespace umcx
