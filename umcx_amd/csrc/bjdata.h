// bjdata.h -- writers for umcx's two output files.
//
// umcx serialises with Binary JData (BJData draft 2), counted containers, typed arrays, and the
// JData N-D array form for {_ArraySize_, _ArrayType_, _ArrayData_} objects
// (src/umcx.cpp:673-705). The layout here is byte-compatible with what the reference writes:
//  * object:  '{' '#' <count> then <key length><key bytes><value> ...
//  * integer: smallest of i U I u l m L M that holds the value
//  * float:   'D' + IEEE double
//  * array:   '[' ['$' <marker>] '#' <count> elements; the '$' form only when every element
//             takes the same marker
//  * N-D:     '[' '$' <type> '#' <dims array> then the raw little-endian payload
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace umcx {

// .bnii: NIFTIHeader.Dim and NIFTIData of a [Nt][Nx][Ny][Nz] volume (:673-686)
void write_bnii_float(const std::string& path, const float* data, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nt);
void write_bnii_int32(const std::string& path, const int32_t* data, uint32_t nx, uint32_t ny, uint32_t nz, uint32_t nt);

// _detp.jdb: MCXData.Info and PhotonRawData (:687-699). TotalPhoton repeats ColumnNum, as
// the reference writes it.
struct DetpInfo {
    int32_t medianum, detnum, columns;
    uint32_t detected, saved;
    float lengthunit;
};
void write_detp(const std::string& path, const DetpInfo& info, const float* rows);

// exposed for tests: serialise one unsigned/signed integer with its smallest marker
std::string bj_int(int64_t v);

// Reader for the BJData input buffer that MCX_run_simulation takes when argn < 0
// (src/umcx.cpp:475-478: json::from_bjdata of a MATLAB-encoded config). Decodes BJData draft 2
// as the reference's JSON library does: little-endian scalars (i U I u l m L M h d D C S H, Z T F,
// N no-op), containers with optional '$' type and '#' count, and N-D arrays ('#' followed by a
// dimension vector), which become {"_ArraySize_", "_ArrayType_", "_ArrayData_"} objects (a 1-D or
// 1 x n dimension vector gives a plain array, a zero dimension an empty one). Structure-of-arrays
// containers ('$' '{') are rejected. Throws std::runtime_error on malformed or truncated input.
class Json;
Json bjdata_decode(const uint8_t* buf, size_t n);

}  // namespace umcx
