// config.h -- umcx's input contract: CLI flags, JSON sections, benchmark presets, volume
// construction and the simulation constants derived from them.
//
// Behaviour follows src/umcx.cpp:451-669 (MCX_userio) and :783-813 (MCX_run_simulation's
// parameter derivation) of fangq/umcx, including its defaults (the code's, not the README's).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "json.h"

namespace umcx {

enum SourceType { kPencil = 0, kIsotropic, kCone, kDisk, kPlanar, kFourier, kUnknownSource };   // :61-62
enum OutputType { kFluenceRate = 0, kFluence = 1, kEnergy = 2 };                                // :59-60
enum DetFlags { kDetID = 1, kPPath = 4, kExitPos = 16, kExitDir = 32 };                         // :65
constexpr uint32_t kDetMask = 0x80000000u;   // :35
constexpr uint32_t kMedMask = 0x7FFFFFFFu;   // :36

// What MCX_userio decided the command asks for (:474-532).
enum class Action { Simulate, DumpJson, DumpMask, Help };

struct ParsedInput {
    Json cfg;
    Action action = Action::Simulate;
};

// Everything one simulation needs, already derived from the JSON (MCX_param :95-99 plus the
// label volume, media, detectors, source and seed words).
struct Problem {
    Json cfg;
    uint32_t nx = 0, ny = 0, nz = 0;
    std::vector<int32_t> labels;          // [x][y][z], z fastest; bit 31 = detector mask
    float tstart = 0, tend = 0, rtstep = 0, unitinmm = 1;
    int32_t maxgate = 1, isreflect = 1, isnormalized = 1, issavevol = 1, issavedet = 1;
    int32_t savedetflag = 5, mediumnum = 0, outputtype = 0, detnum = 0, srctype = 0;
    uint32_t maxdetphotons = 1000000;
    float srcpos[4] = {0, 0, 0, 1}, srcdir[4] = {0, 0, 1, 0};
    float srcparam1[4] = {0, 0, 0, 0}, srcparam2[4] = {0, 0, 0, 0};
    std::vector<float> media;             // mediumnum x {mua, mus, g, n}, per voxel-length units
    std::vector<float> detpos;            // detnum x {x, y, z, R}
    uint32_t seeds[4] = {0, 0, 0, 0};
    int64_t rngseed = 1648335518;
    uint64_t nphoton = 0;
    std::string session_id;               // empty when Session.ID is absent
    int threadnum = 0, blocksize = 0, deviceid = 1;

    int detp_columns() const;            // :157
    int ppath_columns() const;           // :158
    bool saves_detectors() const { return issavedet != 0; }
};

ParsedInput parse_argv(int argc, char** argv);          // :474-532
ParsedInput parse_json_string(const std::string& s);    // MCX_run_json (:853-856)
Json benchmark_config(const std::string& name);         // :633-669
std::string help_text();                                // :533-536
// Label volume from Domain/Shapes, then detector masks (:537-632).
void build_domain(const Json& cfg, uint32_t& nx, uint32_t& ny, uint32_t& nz, std::vector<int32_t>& labels);
// Full problem: domain + constants + media + detectors + seed words (:725-810).
Problem make_problem(const Json& cfg);
// glibc srand(seed) followed by four rand() (:725-727)
void glibc_seed_words(int32_t seed, uint32_t out[4]);
// base64 / zlib / gzip decoding for _ArrayZipData_ (:578-597)
std::vector<uint8_t> base64_decode(const std::string& s);
std::vector<uint8_t> inflate_bytes(const std::vector<uint8_t>& in, bool gzip);

}  // namespace umcx
