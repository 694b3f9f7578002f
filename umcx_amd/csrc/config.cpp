// config.cpp -- umcx's input contract (see config.h).
#include "config.h"

#include "bjdata.h"

#include <zlib.h>

#include <array>
#include <cmath>
#include <cstring>
#include <ctime>
#include <fstream>
#include <functional>
#include <sstream>

namespace umcx {

namespace {

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("[json.exception.parse_error.101] parse error at line 1, column 1: attempting to parse an empty input; "
                 "check that your input string or stream contains the expected JSON (cannot open " + path + ")");
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// Flag table of src/umcx.cpp:454-456: short flag, long flag, JSON pointer, string-valued?
struct FlagSpec { const char* s; const char* l; const char* ptr; bool is_string; };
const FlagSpec kFlags[] = {
    {"-n", "--photon", "/Session/Photons", false},   {"-b", "--reflect", "/Session/DoMismatch", false},
    {"-s", "--session", "/Session/ID", true},        {"-H", "--maxdetphoton", "/Session/MaxDetPhoton", false},
    {"-u", "--unitinmm", "/Domain/LengthUnit", false}, {"-U", "--normalize", "/Session/DoNormalize", false},
    {"-E", "--seed", "/Session/RNGSeed", false},     {"-O", "--outputtype", "/Session/OutputType", true},
    {"-w", "--savedetflag", "/Session/SaveDetFlag", false}, {"-d", "--savedet", "/Session/DoPartialPath", false},
    {"-t", "--thread", "/Session/ThreadNum", false}, {"-T", "--blocksize", "/Session/BlockSize", false},
    {"-G", "--gpuid", "/Session/DeviceID", false},   {"-S", "--save2pt", "/Session/DoSaveVolume", false},
};

const char* kBenchNames[] = {"cube60", "cube60b", "cube60planar", "cubesph60b", "sphshells", "spherebox", "skinvessel"};

// Index of a label volume cell (MCX_volume::index, :128-130), sizes kept as in the reference
inline int32_t cell(uint32_t nx, uint32_t ny, uint32_t nz, int16_t ix, int16_t iy, int16_t iz) {
    if (ix < 0 || iy < 0 || iz < 0 || ix >= (int16_t)nx || iy >= (int16_t)ny || iz >= (int16_t)nz) return -1;
    return (int32_t)((uint64_t)ix * ((uint64_t)ny * nz) + (uint64_t)((uint32_t)iy * nz) + (uint64_t)iz);
}

float num(const Json& o, const char* k, size_t i) { return o.at(k).at(i).as_float(); }

// One rasterizable construct (:458-473), with its numbers read once.
struct Shape {
    enum Kind { Sphere, XLayer, YLayer, ZLayer, Cylinder } kind;
    float a[8] = {0};
    int tag = 0;
    int label(float x, float y, float z) const {
        switch (kind) {
            case Sphere:
                return ((x - a[0]) * (x - a[0]) + (y - a[1]) * (y - a[1]) + (z - a[2]) * (z - a[2]) < (a[3] * a[3])) ? tag : 0;
            case XLayer: return (x >= a[0] && x <= a[1]) ? tag : 0;
            case YLayer: return (y >= a[0] && y <= a[1]) ? tag : 0;
            case ZLayer: return (z >= a[0] && z <= a[1]) ? tag : 0;
            case Cylinder: {
                // a = {C0x, C0y, C0z, v0x, v0y, v0z, d0, R}; the axial test compares the projection
                // with the squared axis length, as src/umcx.cpp:470 does
                float px = x - a[0], py = y - a[1], pz = z - a[2];
                float d = a[3] * px + a[4] * py + a[5] * pz;
                return (d <= a[6] && d >= 0.f && px * px + py * py + pz * pz - d * d <= a[7] * a[7]) ? tag : 0;
            }
        }
        return 0;
    }
};

Shape make_cylinder(const Json& o) {
    Shape s;
    s.kind = Shape::Cylinder;
    float c0[3] = {num(o, "C0", 0), num(o, "C0", 1), num(o, "C0", 2)};
    float c1[3] = {num(o, "C1", 0), num(o, "C1", 1), num(o, "C1", 2)};
    float d0 = (c1[0] - c0[0]) * (c1[0] - c0[0]) + (c1[1] - c0[1]) * (c1[1] - c0[1]) + (c1[2] - c0[2]) * (c1[2] - c0[2]);
    s.a[0] = c0[0]; s.a[1] = c0[1]; s.a[2] = c0[2];
    s.a[3] = (c1[0] - c0[0]) / sqrtf(d0);
    s.a[4] = (c1[1] - c0[1]) / sqrtf(d0);
    s.a[5] = (c1[2] - c0[2]) / sqrtf(d0);
    s.a[6] = d0;
    s.a[7] = o.at("R").as_float();
    s.tag = (int)o.at("Tag").as_int();
    return s;
}

void paint(std::vector<int32_t>& vol, uint32_t nx, uint32_t ny, uint32_t nz, const std::vector<Shape>& shapes) {
    // later constructs overwrite earlier ones; a zero label leaves the voxel (MCX_volume::mask :138-140)
#pragma omp parallel for collapse(2) schedule(static)
    for (int64_t x = 0; x < (int64_t)nx; x++)
        for (int64_t y = 0; y < (int64_t)ny; y++)
            for (uint32_t z = 0; z < nz; z++) {
                int32_t idx = cell(nx, ny, nz, (int16_t)x, (int16_t)y, (int16_t)z);
                for (const Shape& s : shapes) {
                    int lab = s.label((float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f);
                    if (lab) vol[idx] = lab;
                }
            }
}

template <typename T>
void copy_cast(const std::vector<uint8_t>& raw, std::vector<int32_t>& vol) {
    size_t n = vol.size();
    if (raw.size() < n * sizeof(T)) fail("_ArrayZipData_ holds fewer elements than _ArraySize_ describes");
    for (size_t i = 0; i < n; i++) {
        T v;
        memcpy(&v, raw.data() + i * sizeof(T), sizeof(T));
        vol[i] = (int32_t)v;
    }
}

// maskdetectors (:605-632): flag voxels near each detector disk so termination can test them
void mask_detectors(const Json& dets, uint32_t nx, uint32_t ny, uint32_t nz, std::vector<int32_t>& vol) {
    static const int8_t nb[26][3] = {{-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {0, 0, -1}, {1, 0, -1},
                                     {-1, 1, -1},  {0, 1, -1},  {1, 1, -1},  {-1, -1, 0}, {0, -1, 0}, {1, -1, 0},
                                     {-1, 0, 0},   {1, 0, 0},   {-1, 1, 0},  {0, 1, 0},   {1, 1, 0},  {-1, -1, 1},
                                     {0, -1, 1},   {1, -1, 1},  {-1, 0, 1},  {0, 0, 1},   {1, 0, 1},  {-1, 1, 1},
                                     {0, 1, 1},    {1, 1, 1}};
    const int64_t total = (int64_t)vol.size();
    for (size_t d = 0; d < dets.size(); d++) {
        const Json& det = dets.at(d);
        float R = det.at("R").as_float();
        float px = num(det, "Pos", 0), py = num(det, "Pos", 1), pz = num(det, "Pos", 2);
        for (float iz = -R - 1.f + pz; iz <= R + 1.f + pz; iz += 0.5f)
            for (float iy = -R - 1.f + py; iy <= R + 1.f + py; iy += 0.5f)
                for (float ix = -R - 1.f + px; ix <= R + 1.f + px; ix += 0.5f) {
                    int16_t sx = (int16_t)ix, sy = (int16_t)iy, sz = (int16_t)iz;
                    int32_t idx = cell(nx, ny, nz, sx, sy, sz);
                    if (idx < 0 || (ix - px) * (ix - px) + (iy - py) * (iy - py) + (iz - pz) * (iz - pz) > (R + 1.f) * (R + 1.f) ||
                        ((uint32_t)vol[idx] & kMedMask) == 0)
                        continue;
                    if ((int)sx * (int)sy * (int)sz == 0 || sx == (int16_t)nx || sy == (int16_t)ny || sz == (int16_t)nz) {
                        vol[idx] = (int32_t)((uint32_t)vol[idx] | kDetMask);
                    } else {
                        // the neighbour offset goes through the bounds-checked index, so any offset
                        // with a -1 component reads the cell one z-step back (:625); offsets that
                        // leave the buffer are skipped here
                        for (int k = 0; k < 26; k++) {
                            int64_t o = idx + (int64_t)cell(nx, ny, nz, nb[k][0], nb[k][1], nb[k][2]);
                            if (o < 0 || o >= total) continue;
                            if (((uint32_t)vol[o] & kMedMask) == 0) {
                                vol[idx] = (int32_t)((uint32_t)vol[idx] | kDetMask);
                                break;
                            }
                        }
                    }
                }
    }
}

}  // namespace

void glibc_seed_words(int32_t seed, uint32_t out[4]) {
    // glibc random() TYPE_3 state: r[0] = seed, r[i] = 16807 r[i-1] mod (2^31-1) for i < 31,
    // r[31..33] = r[0..2], r[i] = r[i-31] + r[i-3]; output k is r[344+k] >> 1
    int32_t r[348];
    r[0] = seed == 0 ? 1 : seed;
    for (int i = 1; i < 31; i++) {
        int64_t v = (16807LL * (int64_t)r[i - 1]) % 2147483647LL;
        if (v < 0) v += 2147483647LL;
        r[i] = (int32_t)v;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int i = 34; i < 348; i++) r[i] = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
    for (int k = 0; k < 4; k++) out[k] = ((uint32_t)r[344 + k]) >> 1;
}

std::vector<uint8_t> base64_decode(const std::string& s) {
    // built once, thread-safely (function-local static initialisation)
    static const std::array<int8_t, 256> tab = [] {
        std::array<int8_t, 256> t;
        t.fill(-1);
        const char* a = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
        for (int i = 0; i < 64; i++) t[(uint8_t)a[i]] = (int8_t)i;
        return t;
    }();
    std::vector<uint8_t> out;
    uint32_t acc = 0;
    int bits = 0;
    for (unsigned char c : s) {
        if (c == '=') break;
        if (tab[c] < 0) continue;   // skips line breaks
        acc = (acc << 6) | (uint32_t)tab[c];
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out.push_back((uint8_t)((acc >> bits) & 0xFF));
        }
    }
    return out;
}

std::vector<uint8_t> inflate_bytes(const std::vector<uint8_t>& in, bool gzip) {
    z_stream zs;
    memset(&zs, 0, sizeof zs);
    if (inflateInit2(&zs, gzip ? (16 + MAX_WBITS) : MAX_WBITS) != Z_OK) fail("zlib: cannot initialise the decoder");
    std::vector<uint8_t> out;
    zs.next_in = const_cast<Bytef*>(in.data());
    zs.avail_in = (uInt)in.size();
    std::vector<uint8_t> buf(1 << 20);
    int rc;
    do {
        zs.next_out = buf.data();
        zs.avail_out = (uInt)buf.size();
        rc = inflate(&zs, Z_NO_FLUSH);
        if (rc != Z_OK && rc != Z_STREAM_END) {
            inflateEnd(&zs);
            fail("zlib: corrupt _ArrayZipData_ stream");
        }
        out.insert(out.end(), buf.data(), buf.data() + (buf.size() - zs.avail_out));
    } while (rc != Z_STREAM_END && (zs.avail_in > 0 || zs.avail_out == 0));
    inflateEnd(&zs);
    return out;
}

std::string help_text() {
    std::string s =
        "/uMCX/ - Portable, massively-parallel physical volumetric ray-tracer (MI355X build)\n\nFormat:\n"
        "\tumcx -flag1 value1 -flag2 value2 ...\n\t\tor\n\tumcx inputjson.json\n\tumcx benchmarkname\n\nFlags:\n"
        "\t-f/--input\tinput json file\n\t-Q/--bench\t\tbenchmark name\n\t-n/--photon\tphoton number [1e6]\n"
        "\t-s/--session\toutput name\n\t-u/--unitinmm\tvoxel size in mm [1]\n\t-E/--seed\tRNG seed [1648335518]\n"
        "\t-O/--outputtype\t[x]: fluence-rate, f: fluence, e: energy\n\t-d/--savedet\tSave detected photons [1]\n"
        "\t-S/--save2pt\tSave volumetric output [1]\n"
        "\t-w/--savedetflag\t1:detector-id, 4:partial-path, 16:exit-pos, 32:exit-dir, add to combine [5]\n"
        "\t-U/--normalize\tnormalize output [1]\n\t-j/--json\tJSON string to overwrite settings\n"
        "\t-t/--thread\tmanual total threads\n\t-T/--blocksize\tmanual thread-block size [autotuned]\n"
        "\t-G/--gpuid\tdevice ID [1]\n\t--dumpjson\tdump settings as json\n\t--dumpmask\tdump domain as binary json\n"
        "\t-h/--help\tprint help\n\nBuilt-in benchmarks: [\n";
    for (size_t i = 0; i < sizeof(kBenchNames) / sizeof(kBenchNames[0]); i++)
        s += std::string("        \"") + kBenchNames[i] + "\"" + (i + 1 < sizeof(kBenchNames) / sizeof(kBenchNames[0]) ? ",\n" : "\n");
    s += "]\n";
    return s;
}

Json benchmark_config(const std::string& name) {
    // presets of src/umcx.cpp:633-669, in the same key order
    int id = -1;
    for (int i = 0; i < 7; i++) if (name == kBenchNames[i]) id = i;
    Json cfg = Json::parse(
        "{\"Session\":{\"ID\":\"" + name + "\",\"Photons\":1000000},"
        "\"Forward\":{\"T0\":0.0,\"T1\":5e-9,\"Dt\":5e-9},"
        "\"Domain\":{\"Media\":[{\"mua\":0.0,\"mus\":0.0,\"g\":1.0,\"n\":1.0},{\"mua\":0.005,\"mus\":1.0,\"g\":0.01,\"n\":1.37},"
        "{\"mua\":0.002,\"mus\":5.0,\"g\":0.9,\"n\":1.0}],\"Dim\":[60,60,60]},"
        "\"Optode\":{\"Source\":{\"Type\":\"pencil\",\"Pos\":[29.0,29.0,0.0],\"Dir\":[0.0,0.0,1.0]},"
        "\"Detector\":[{\"Pos\":[29,19,0],\"R\":1},{\"Pos\":[29,39,0],\"R\":1},{\"Pos\":[19,29,0],\"R\":1},{\"Pos\":[39,29,0],\"R\":1}]}}");
    cfg["Shapes"] = Json::parse("[{\"Grid\":{\"Tag\":1,\"Size\":[60,60,60]}}]");
    cfg["Session"]["DoMismatch"] = Json(!(id == 0 || id == 6 || id == 5));
    if (id == 3 || id == 2) {
        cfg["Shapes"] = Json::parse("[{\"Grid\":{\"Tag\":1,\"Size\":[60,60,60]}},{\"Sphere\":{\"O\":[30,30,30],\"R\":15,\"Tag\":2}}]");
        if (id == 2)
            cfg["Optode"]["Source"] = Json::parse("{\"Type\":\"planar\",\"Pos\":[10.0,10.0,-10.0],\"Dir\":[0.0,0.0,1.0],"
                                                  "\"Param1\":[40.0,0.0,0.0,0.0],\"Param2\":[0.0,40.0,0.0,0.0]}");
    } else if (id == 6) {
        cfg["Shapes"] = Json::parse("[{\"Grid\":{\"Size\":[200,200,200],\"Tag\":1}},{\"ZLayers\":[[1,20,1],[21,32,4],[33,200,3]]},"
                                    "{\"Cylinder\":{\"Tag\":2,\"C0\":[0,100.5,100.5],\"C1\":[200,100.5,100.5],\"R\":20}}]");
        cfg["Forward"] = Json::parse("{\"T0\":0.0,\"T1\":5e-8,\"Dt\":5e-8}");
        cfg["Optode"] = Json::parse("{\"Source\":{\"Type\":\"disk\",\"Pos\":[100,100,20],\"Dir\":[0,0,1],\"Param1\":[60,0,0,0]}}");
        cfg["Domain"] = Json::parse(
            "{\"Media\":[{\"mua\":0.002,\"mus\":0.0,\"g\":1.0,\"n\":1.37},{\"mua\":3.564e-05,\"mus\":1.0,\"g\":1.0,\"n\":1.37},"
            "{\"mua\":23.05426549,\"mus\":9.398496241,\"g\":0.9,\"n\":1.37},{\"mua\":0.04584957865,\"mus\":35.65405549,\"g\":0.9,\"n\":1.37},"
            "{\"mua\":1.657237447,\"mus\":37.59398496,\"g\":0.9,\"n\":1.37}],\"LengthUnit\":0.005,\"Dim\":[200,200,200]}");
    } else if (id == 4) {
        cfg["Shapes"] = Json::parse("[{\"Grid\":{\"Size\":[60,60,60],\"Tag\":1}},{\"Sphere\":{\"O\":[30,30,30],\"R\":25,\"Tag\":2}},"
                                    "{\"Sphere\":{\"O\":[30,30,30],\"R\":23,\"Tag\":3}},{\"Sphere\":{\"O\":[30,30,30],\"R\":10,\"Tag\":4}}]");
        cfg["Domain"]["Media"] = Json::parse(
            "[{\"mua\":0.0,\"mus\":0.0,\"g\":1.0,\"n\":1.0},{\"mua\":0.02,\"mus\":7.0,\"g\":0.89,\"n\":1.37},"
            "{\"mua\":0.004,\"mus\":0.09,\"g\":0.89,\"n\":1.37},{\"mua\":0.02,\"mus\":9.0,\"g\":0.89,\"n\":1.37},"
            "{\"mua\":0.05,\"mus\":0.0,\"g\":1.0,\"n\":1.37}]");
    } else if (id == 5) {
        cfg["Forward"]["Dt"] = Json(1e-10);
        cfg["Domain"]["Media"] = Json::parse(
            "[{\"mua\":0.0,\"mus\":0.0,\"g\":1.0,\"n\":1.0},{\"mua\":0.002,\"mus\":1.0,\"g\":0.01,\"n\":1.37},"
            "{\"mua\":0.005,\"mus\":5.0,\"g\":0.9,\"n\":1.37}]");
        cfg["Shapes"] = Json::parse("[{\"Grid\":{\"Tag\":1,\"Size\":[60,60,60]}},{\"Sphere\":{\"O\":[30,30,30],\"R\":10,\"Tag\":2}}]");
    } else if (id < 0) {
        fail("benchmark " + name + " is not found");
    }
    return cfg;
}

ParsedInput parse_argv(int argc, char** argv) {
    ParsedInput in;
    if (argc < 0) {   // argv[0] is a BJData buffer of -argc bytes, as MATLAB passes it (:475-478)
        if (!argv || !argv[0]) fail("BJData input: null buffer");
        in.cfg = bjdata_decode((const uint8_t*)argv[0], (size_t)(-(int64_t)argc));
        return in;
    }
    if (argc <= 1) { in.action = Action::Help; return in; }
    std::vector<std::string> params(argv + 1, argv + argc);
    bool want_help = false;
    if (params[0].find('-') == 0) {
        for (int i = 1; i < argc;) {
            std::string arg(argv[i++]);
            if ((arg == "-f" || arg == "--input") && i < argc) {
                in.cfg = Json::parse(read_file(argv[i++]));
            } else if (arg == "-h" || arg == "--help") {
                want_help = true;
            } else if (arg == "-Q" || arg == "--bench") {
                if (i < argc) in.cfg = benchmark_config(argv[i++]);
                else want_help = true;
            } else if ((arg == "-j" || arg == "--json") && i < argc) {
                in.cfg.update(Json::parse(argv[i++]));
            } else if (arg == "-N" || arg == "--net") {
                fail("-N/--net needs network access to neurojson.io, which this build does not provide");
            } else if (arg[0] == '-' && i < argc) {
                for (const FlagSpec& f : kFlags) {
                    if (arg == f.s || arg == f.l) {
                        std::string v(argv[i++]);
                        in.cfg.set_pointer(f.ptr, f.is_string ? Json::parse("\"" + v + "\"") : Json::parse(v));
                        break;
                    }
                }
            } else if (!(arg == "--dumpjson" || arg == "--dumpmask")) {
                fail("incomplete input parameter: " + arg + "; every -flag/--flag must be followed by a valid value");
            }
        }
    } else if (argc == 2) {
        if (params[0].find('.') == std::string::npos) in.cfg = benchmark_config(params[0]);
        else in.cfg = Json::parse(read_file(params[0]));
    } else {
        fail("must use -flag or --flag to use more than 1 input");
    }
    if (want_help) { in.action = Action::Help; return in; }
    for (const std::string& p : params) {
        if (p == "--dumpjson") { in.action = Action::DumpJson; break; }
        if (p == "--dumpmask") { in.action = Action::DumpMask; break; }
    }
    return in;
}

ParsedInput parse_json_string(const std::string& s) {
    ParsedInput in;
    in.cfg.update(Json::parse(s));   // MCX_run_json passes {"", "--json", s} (:853-856)
    return in;
}

void build_domain(const Json& cfg, uint32_t& nx, uint32_t& ny, uint32_t& nz, std::vector<int32_t>& vol) {
    const Json& dom = cfg.at("Domain");
    auto reshape = [&](const Json& size, int32_t value) {
        nx = (uint32_t)size.at(0).as_uint();
        ny = (uint32_t)size.at(1).as_uint();
        nz = (uint32_t)size.at(2).as_uint();
        if (nx == 0 || ny == 0 || nz == 0 || nx > 32767 || ny > 32767 || nz > 32767)
            fail("Domain dimensions must be within 1..32767 voxels per axis");
        // voxel indices are 32-bit signed integers, as the reference's index() returns (:128-130)
        if ((uint64_t)nx * ny * nz > 0x7fffffffull) fail("Domain must hold fewer than 2^31 voxels");
        vol.assign((size_t)nx * ny * nz, value);
    };
    reshape(dom.at("Dim"), 0);
    const Json* shapes = cfg.find("Shapes");
    if (shapes && !shapes->is_null()) {
        if (!shapes->contains("_ArraySize_")) {
            const Json& first = shapes->is_array() ? shapes->at((size_t)0) : *shapes;
            if (first.contains("Grid")) reshape(first.at("Grid").at("Size"), (int32_t)first.at("Grid").at("Tag").as_int());
            std::vector<Shape> batch;
            // constructs are applied in order; consecutive ones are painted in one pass
            for (size_t k = 0; k < shapes->size(); k++) {
                const Json& obj = shapes->val(k);
                if (!obj.is_object() || obj.size() == 0) fail("shape construct must be a JSON object");
                const std::string& key = obj.key(0);
                const Json& body = obj.val(0);
                if (key == "Sphere") {
                    Shape s; s.kind = Shape::Sphere;
                    s.a[0] = num(body, "O", 0); s.a[1] = num(body, "O", 1); s.a[2] = num(body, "O", 2);
                    s.a[3] = body.at("R").as_float();
                    s.tag = (int)body.at("Tag").as_int();
                    batch.push_back(s);
                } else if (key == "XLayers" || key == "YLayers" || key == "ZLayers") {
                    for (size_t l = 0; l < body.size(); l++) {
                        const Json& layer = body.val(l);
                        Shape s;
                        s.kind = key[0] == 'X' ? Shape::XLayer : (key[0] == 'Y' ? Shape::YLayer : Shape::ZLayer);
                        s.a[0] = layer.at((size_t)0).as_float();
                        s.a[1] = layer.at((size_t)1).as_float();
                        s.tag = (int)layer.at((size_t)2).as_int();
                        batch.push_back(s);
                    }
                } else if (key == "Cylinder") {
                    batch.push_back(make_cylinder(body));
                } else if (key == "Box") {
                    // src/umcx.cpp:460 reads a misspelt "Sip.ze" key and compares x only; it aborts
                    // with a JSON type error, so the construct is rejected here as well
                    fail("shape construct Box is not supported by umcx (its Box parser is broken upstream)");
                } else if (!(key == "Name" || key == "Grid" || key == "Origin")) {
                    fail("shape construct" + key + " is not supported");
                }
            }
            if (!batch.empty()) paint(vol, nx, ny, nz, batch);
        } else {
            const Json& s = *shapes;
            reshape(s.at("_ArraySize_"), 0);
            std::string dtype = s.value_str("_ArrayType_", "int32");
            if (const Json* data = s.find("_ArrayData_")) {
                if (data->size() < vol.size()) fail("_ArrayData_ holds fewer elements than _ArraySize_ describes");
                for (size_t i = 0; i < vol.size(); i++) vol[i] = (int32_t)data->val(i).as_int();
            } else if (const Json* z = s.find("_ArrayZipData_")) {
                // a base64 string (JSON text) or an array of byte values (BJData input, :584)
                std::vector<uint8_t> comp;
                if (z->is_string()) {
                    comp = base64_decode(z->as_string());
                } else if (z->is_array()) {
                    comp.resize(z->size());
                    for (size_t i = 0; i < z->size(); i++) {
                        const uint64_t b = z->at(i).as_uint();
                        if (b > 255) fail("_ArrayZipData_ byte values must be within 0..255");
                        comp[i] = (uint8_t)b;
                    }
                } else {
                    fail("_ArrayZipData_ must be a base64 string or a byte array");
                }
                std::vector<uint8_t> raw = inflate_bytes(comp, s.value_str("_ArrayZipType_", "zlib") == "gzip");
                if (dtype == "uint8") copy_cast<uint8_t>(raw, vol);
                else if (dtype == "int8") copy_cast<int8_t>(raw, vol);
                else if (dtype == "uint16") copy_cast<uint16_t>(raw, vol);
                else if (dtype == "int16") copy_cast<int16_t>(raw, vol);
                else if (dtype == "uint32") copy_cast<uint32_t>(raw, vol);
                else copy_cast<int32_t>(raw, vol);
            }
        }
    }
    const Json& opt = cfg.at("Optode");
    if (const Json* dets = opt.find("Detector")) mask_detectors(*dets, nx, ny, nz, vol);
}

int Problem::detp_columns() const {
    return (savedetflag & kDetID ? 1 : 0) + ((savedetflag & kPPath) ? mediumnum : 0) +
           (((savedetflag & kExitPos) ? 1 : 0) + ((savedetflag & kExitDir) ? 1 : 0)) * 3;
}
int Problem::ppath_columns() const { return (issavedet && (savedetflag & kPPath)) ? mediumnum : 0; }

Problem make_problem(const Json& cfg) {
    Problem p;
    p.cfg = cfg;
    build_domain(cfg, p.nx, p.ny, p.nz, p.labels);
    const Json& S = cfg.contains("Session") ? cfg.at("Session") : Json::object();
    const Json& F = cfg.at("Forward");
    const Json& D = cfg.at("Domain");
    const Json& src = cfg.at("Optode").at("Source");
    // MCX_param (:787-796)
    float T0 = F.at("T0").as_float(), T1 = F.at("T1").as_float(), Dt = F.at("Dt").as_float();
    p.tstart = T0;
    p.tend = T1;
    p.rtstep = 1.f / Dt;
    p.unitinmm = (float)D.value("LengthUnit", (double)1.f);
    p.maxgate = (int32_t)((T1 - T0) / Dt + 0.5f);
    if (p.maxgate < 1) fail("Forward: (T1-T0)/Dt must give at least one time gate");
    p.isreflect = (int32_t)S.value_int("DoMismatch", 1);
    p.isnormalized = (int32_t)S.value_int("DoNormalize", 1);
    p.issavevol = (int32_t)S.value_int("DoSaveVolume", 1);
    p.issavedet = (int32_t)S.value_int("DoPartialPath", 1);
    p.savedetflag = (int32_t)S.value_int("SaveDetFlag", kDetID + kPPath);
    const Json& media = D.at("Media");
    p.mediumnum = (int32_t)media.size();
    std::string ot = S.value_str("OutputType", "x");
    p.outputtype = ot.empty() ? -1 : (int32_t)std::string("xfe").find(ot[0]);
    const Json* dets = cfg.at("Optode").find("Detector");
    p.detnum = dets ? (int32_t)dets->size() : 0;
    p.maxdetphotons = (uint32_t)S.value_int("MaxDetPhoton", 1000000);
    std::string st = src.value_str("Type", "pencil");
    const char* names[] = {"pencil", "isotropic", "cone", "disk", "planar", "fourier"};
    p.srctype = kUnknownSource;
    for (int i = 0; i < 6; i++) if (st == names[i]) p.srctype = i;
    for (const char* key : {"Param1", "Param2"}) {
        float* dst = key[5] == '1' ? p.srcparam1 : p.srcparam2;
        if (const Json* v = src.find(key)) {
            if (!v->is_null())
                for (size_t i = 0; i < 4 && i < v->size(); i++) dst[i] = v->at(i).as_float();
        }
    }
    // media (:803-806): objects are scaled to voxel units, arrays are taken as given
    if (p.mediumnum <= 0) fail("Domain.Media must list at least one medium");
    for (int i = 0; i < p.mediumnum; i++) {
        const Json& m = media.at((size_t)i);
        if (m.is_array()) {
            if (m.size() < 4) fail("each Domain.Media array entry needs 4 values {mua, mus, g, n}");
            for (size_t k = 0; k < 4; k++) p.media.push_back(m.at(k).as_float());
        } else {
            p.media.push_back(m.at("mua").as_float() * p.unitinmm);
            p.media.push_back(m.at("mus").as_float() * p.unitinmm);
            p.media.push_back(m.at("g").as_float());
            p.media.push_back(m.at("n").as_float());
        }
    }
    for (int i = 0; i < p.detnum; i++) {   // :808-810
        const Json& d = dets->at((size_t)i);
        p.detpos.push_back(num(d, "Pos", 0));
        p.detpos.push_back(num(d, "Pos", 1));
        p.detpos.push_back(num(d, "Pos", 2));
        p.detpos.push_back(d.at("R").as_float());
    }
    // every label must name a listed medium; the reference would index past its media table
    for (int32_t v : p.labels) {
        uint32_t m = (uint32_t)v & kMedMask;
        if (m >= (uint32_t)p.mediumnum) fail("label " + std::to_string(m) + " has no entry in Domain.Media");
    }
    // seeds and source (:725-729)
    if (!S.contains("Photons") || S.at("Photons").is_null()) fail("[json.exception.type_error.302] type must be number, but is null (Session.Photons)");
    p.nphoton = S.at("Photons").as_uint();
    int64_t seed = 1648335518;
    if (const Json* s = S.find("RNGSeed")) seed = s->as_int() > 0 ? s->as_int() : (int64_t)time(nullptr);
    p.rngseed = seed;
    glibc_seed_words((int32_t)seed, p.seeds);
    for (int i = 0; i < 3; i++) {
        p.srcpos[i] = src.at("Pos").at((size_t)i).as_float();
        p.srcdir[i] = src.at("Dir").at((size_t)i).as_float();
    }
    p.srcpos[3] = 1.f;
    p.srcdir[3] = 0.f;
    if (const Json* id = S.find("ID")) p.session_id = id->as_string();
    p.threadnum = (int)S.value_int("ThreadNum", 0);
    p.blocksize = (int)S.value_int("BlockSize", 0);
    p.deviceid = (int)S.value_int("DeviceID", 1);
    return p;
}

}  // namespace umcx
