// photon_device.h -- per-photon transport of umcx on gfx950, shared by the plain, replay and
// batched kernels.
//
// Each function restates one MCX_photon / MCX_rand method of fangq/umcx (cited per function)
// with the reference's FP32 operation order; the trajectory-affecting libm calls use the
// bit-exact replicas of glibc_libm.h. State lives in registers (Photon, Rng); geometry and
// source come in as small by-value structs so one lane's job may differ from its neighbour's.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "config.h"
#include "glibc_libm.h"

namespace umcx {
namespace dev {

constexpr float kOneOverC0 = 3.335640951981520e-12f;   // src/umcx.cpp:29
constexpr float kPi = 3.1415926535897932385f;           // :30
constexpr float kEps = 1.19209290e-07f;                 // FLT_EPSILON
constexpr uint32_t kDet = 0x80000000u, kMed = 0x7FFFFFFFu;

// label grid of one simulation (MCX_volume::size, :104-120)
struct Grid {
    uint32_t nx, ny, nz, nyz;
};

// source of one simulation (:728-729, MCX_param::srcparam1/2, srctype)
struct Source {
    float4 srcpos, srcdir, srcp1, srcp2;
    int32_t srctype;
};

// libm tables staged in LDS (per-lane gathers from constant memory would go through the vector
// memory pipeline)
struct Tabs {
    const libm::LogfEntry* log;
    const uint64_t* exp;
};

// ---------------- xorshift128+ (src/umcx.cpp:186-215) ----------------
struct Rng {
    uint64_t t0, t1;
};
__device__ __forceinline__ void rng_seed(Rng& r, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    r.t0 = ((uint64_t)a << 32) | b;
    r.t1 = ((uint64_t)c << 32) | d;
}
__device__ __forceinline__ float rng_u01(Rng& r) {
    uint64_t s0 = r.t1, s1 = r.t0;
    r.t0 = s0;
    s1 ^= s1 << 23;
    r.t1 = s1 ^ s0 ^ (s1 >> 18) ^ (s0 >> 5);
    uint32_t lo = (uint32_t)(r.t1 + s0);
    return __builtin_bit_cast(float, 0x3F800000u | (lo >> 9)) - 1.0f;
}
__device__ __forceinline__ float rng_scatlen(Rng& r, const Tabs& t) {
    return -libm::logf_tab(rng_u01(r) + kEps, t.log);
}

// ---------------- photon state (src/umcx.cpp:216-220) ----------------
struct Photon {
    float px, py, pz, pw;      // position, weight
    float vx, vy, vz, vw;      // direction, scatter count
    float rx, ry, rz;          // reciprocal direction
    float lx, ly, lz, lw;      // scattering length left, time, path since save, weight at save
    int ix, iy, iz, iw;        // voxel (short semantics) and crossing axis
    int lastvox;
    uint32_t med;
    int tvox;                  // offset of lastvox inside the LDS tally tile, -1 outside it
    float cmua, cmus;          // absorption and scattering coefficients of lastvox
    float cn;                  // refractive index of lastvox
};

// Label sources. raw() issues the load (its value is waited for where decode() uses it), decode()
// gives the kernels' form: medium in the low bits, detector flag in bit 31.
// GlobalLabels: the label volume in HBM (uint8 with the detector flag in bit 7, or uint32).
template <typename LabelT>
struct GlobalLabels {
    const LabelT* p;
    __device__ __forceinline__ uint32_t raw(int idx) const { return p[idx]; }
    __device__ __forceinline__ uint32_t decode(uint32_t v) const {
        if (sizeof(LabelT) == 1) return (v & 0x7Fu) | ((v & 0x80u) << 24);
        return v;
    }
};
// LdsLabels: media ids packed 1, 2 or 4 bits per voxel into 32-bit words staged in LDS (no
// detector flag: the kernels read that one from HBM when a photon ends). The media id is
// id0 + code * step: 2- and 4-bit codes are the id itself (id0 0, step 1); 1-bit codes, used when
// the volume holds at most two media ids, select the smaller (code 0) or the larger one.
struct LdsLabels {
    const uint32_t* w;
    uint32_t lg;      // log2(voxels per word): 5, 4 or 3 for 1-, 2- or 4-bit labels
    uint32_t bits;    // 1, 2 or 4
    uint32_t id0, step;
    __device__ __forceinline__ uint32_t raw(int idx) const {
        const uint32_t i = (uint32_t)idx;
        return (w[i >> lg] >> ((i & ((1u << lg) - 1u)) * bits)) & ((1u << bits) - 1u);
    }
    __device__ __forceinline__ uint32_t decode(uint32_t v) const { return id0 + v * step; }
};

template <typename LabelT>
__device__ __forceinline__ uint32_t load_label(const void* labels, int idx);
template <>
__device__ __forceinline__ uint32_t load_label<uint8_t>(const void* labels, int idx) {
    uint32_t v = ((const uint8_t*)labels)[idx];
    return (v & 0x7Fu) | ((v & 0x80u) << 24);
}
template <>
__device__ __forceinline__ uint32_t load_label<uint32_t>(const void* labels, int idx) {
    return ((const uint32_t*)labels)[idx];
}

__device__ __forceinline__ int to_short(float x) { return (int)(short)(int)x; }

// MCX_volume::index (:128-130)
__device__ __forceinline__ int vox_index(const Grid& a, int ix, int iy, int iz) {
    if ((unsigned)ix >= a.nx || (unsigned)iy >= a.ny || (unsigned)iz >= a.nz) return -1;
    return ix * (int)a.nyz + iy * (int)a.nz + iz;
}

// rotatevector (:418-428)
__device__ __forceinline__ void rotate(Photon& p, float st, float ct, float sp, float cp) {
    float nx, ny, nz;
    if (p.vz > -1.f + kEps && p.vz < 1.f - kEps) {
        float t0 = 1.f - p.vz * p.vz;
        float t1 = libm::div_ns(st, libm::sqrt_ns(t0));   // t0 >= 2^-23; st = sin of [0, pi], +0 or >= 2^-24
        nx = t1 * (p.vx * p.vz * cp - p.vy * sp) + p.vx * ct;
        ny = t1 * (p.vy * p.vz * cp + p.vx * sp) + p.vy * ct;
        nz = -t1 * t0 * cp + p.vz * ct;
    } else {
        nx = st * cp;
        ny = st * sp;
        nz = (p.vz > 0.f) ? ct : -ct;
    }
    float s = libm::rcp_ns(libm::sqrt_ns(nx * nx + ny * ny + nz * nz));   // of a near-unit vector's norm
    p.vx = nx * s;
    p.vy = ny * s;
    p.vz = nz * s;
}

// launch (:225-259); returns the launch weight
__device__ __forceinline__ float launch(Photon& p, Rng& r, const Source& a) {
    p.px = a.srcpos.x; p.py = a.srcpos.y; p.pz = a.srcpos.z; p.pw = 1.f;
    p.vx = a.srcdir.x; p.vy = a.srcdir.y; p.vz = a.srcdir.z; p.vw = 0.f;
    const int st = a.srctype;
    if (st == kIsotropic || st == kCone) {
        float sphi, cphi, sth, cth;
        libm::sincosf_host((kPi * 2.f) * rng_u01(r), &sphi, &cphi);
        float th;
        if (st == kIsotropic) {
            th = libm::acosf_host(2.f * rng_u01(r) - 1.f);
        } else {
            float ca = libm::cosf_any(a.srcp1.x);   // a user-set half-angle: any float
            th = libm::acosf_host(rng_u01(r) * (1.0f - ca) + ca);
        }
        libm::sincosf_host(th, &sth, &cth);
        rotate(p, sth, cth, sphi, cphi);
    } else if (st == kDisk) {
        float sphi, cphi;
        libm::sincosf_host((kPi * 2.f) * rng_u01(r), &sphi, &cphi);
        float rad = libm::sqrtf_cr(rng_u01(r) * fabsf(a.srcp1.x * a.srcp1.x - a.srcp1.y * a.srcp1.y) + a.srcp1.y * a.srcp1.y);
        if (p.vz > -1.f + kEps && p.vz < 1.f - kEps) {
            float q = 1.f - p.vz * p.vz;
            float b = rad / libm::sqrtf_cr(q);
            float nx = p.px + b * (p.vx * p.vz * cphi - p.vy * sphi);
            float ny = p.py + b * (p.vy * p.vz * cphi + p.vx * sphi);
            float nz = p.pz - b * q * cphi;
            p.px = nx; p.py = ny; p.pz = nz;
        } else {
            p.px += rad * cphi;
            p.py += rad * sphi;
        }
    } else if (st == kPlanar || st == kFourier) {
        float u = rng_u01(r), v = rng_u01(r);
        float nx = p.px + u * a.srcp1.x + v * a.srcp2.x;
        float ny = p.py + u * a.srcp1.y + v * a.srcp2.y;
        float nz = p.pz + u * a.srcp1.z + v * a.srcp2.z;
        p.px = nx; p.py = ny; p.pz = nz;
        if (st == kFourier) {
            float f1 = floorf(a.srcp1.w), f2 = floorf(a.srcp2.w);
            float arg = (f1 * u + f2 * v + a.srcp1.w - f1) * kPi * 2.f;
            p.pw *= (libm::cosf_any(arg) * (1.f - a.srcp2.w + f2) + 1.f) * 0.5f;   // user-set frequencies
        }
    }
    p.rx = 1.f / a.srcdir.x;   // from the nominal direction (:255)
    p.ry = 1.f / a.srcdir.y;
    p.rz = 1.f / a.srcdir.z;
    p.lx = __builtin_nanf("");
    p.ly = 0.f; p.lz = 0.f; p.lw = p.pw;
    p.ix = to_short(p.px); p.iy = to_short(p.py); p.iz = to_short(p.pz); p.iw = -1;
    return p.pw;
}

// skip (:339-358): bring a photon launched outside the grid to its entry point
__device__ __forceinline__ bool skip_to_box(Photon& p, const Grid& a) {
    float tx0 = -p.px * p.rx, ty0 = -p.py * p.ry, tz0 = -p.pz * p.rz;
    float tx1 = ((float)a.nx - p.px) * p.rx;
    float ty1 = ((float)a.ny - p.py) * p.ry;
    float tz1 = ((float)a.nz - p.pz) * p.rz;
    float tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    float tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    if (tmax < 0.f || tmin > tmax) {
        p.lx = tx0; p.ly = ty0; p.lz = tz0; p.lw = tx1;
        p.pw = tz1;
        return false;
    }
    float nx = p.px + tmin * p.vx, ny = p.py + tmin * p.vy, nz = p.pz + tmin * p.vz;
    p.px = nx; p.py = ny; p.pz = nz; p.pw = 1.f;
    p.lx = __builtin_nanf(""); p.ly = 0.f; p.lz = 0.f; p.lw = 1.f;
    p.ix = to_short(p.px); p.iy = to_short(p.py); p.iz = to_short(p.pz); p.iw = -1;
    p.lastvox = vox_index(a, p.ix, p.iy, p.iz);
    return true;
}

// step (:315-338); returns the voxel the photon is in afterwards (-1 outside). The label of the
// voxel across the nearest face is loaded before the weight update and handed back raw in
// `nlab`, so its memory latency overlaps the step's arithmetic and the deposit; it is valid
// whenever the photon crosses into a voxel inside the grid.
template <typename Labels>
__device__ __forceinline__ int step(Photon& p, float n, const Grid& a, const Tabs& t, const Labels& labels,
                                    uint32_t& nlab) {
    const float mua = p.cmua, mus = p.cmus;
    float hx = fabsf(((float)(p.ix + (p.vx > 0.f)) - p.px) * p.rx);
    float hy = fabsf(((float)(p.iy + (p.vy > 0.f)) - p.py) * p.ry);
    float hz = fabsf(((float)(p.iz + (p.vz > 0.f)) - p.pz) * p.rz);
    float dmin = fminf(fminf(hx, hy), hz);
    p.iw = (dmin == hx) ? 0 : ((dmin == hy) ? 1 : 2);
    // the voxel across the nearest face (value selects keep the photon in registers)
    const int ax = p.iw;
    const int cx = (int)(short)(p.ix + (ax == 0 ? (p.vx > 0.f ? 1 : -1) : 0));
    const int cy = (int)(short)(p.iy + (ax == 1 ? (p.vy > 0.f ? 1 : -1) : 0));
    const int cz = (int)(short)(p.iz + (ax == 2 ? (p.vz > 0.f ? 1 : -1) : 0));
    const int cand = vox_index(a, cx, cy, cz);
    nlab = labels.raw(cand >= 0 ? cand : 0);
    float used = fminf(dmin * mus, p.lx);
    bool cont = used != p.lx;
    float d = (mus == 0.f) ? dmin : (used / mus);
    p.px = p.px + d * p.vx;
    p.py = p.py + d * p.vy;
    p.pz = p.pz + d * p.vz;
    p.pw = p.pw * libm::expf_tab(-mua * d, t.exp);
    p.lx -= used;
    p.ly += d * n * kOneOverC0;
    p.lz += d;
    if (cont) {
        p.ix = cx;
        p.iy = cy;
        p.iz = cz;
        return cand;
    }
    return p.lastvox;
}

// scatter (:364-380)
__device__ __forceinline__ void scatter(Photon& p, Rng& r, float4 prop, const Tabs& t) {
    const float g = prop.z;
    p.lx = rng_scatlen(r, t);
    float sphi, cphi, sth, cth;
    libm::sincosf_host((2.f * kPi) * rng_u01(r), &sphi, &cphi);
    float th;
    if (fabsf(g) > kEps) {
        // |g| in (2^-23, 1]: denominators in [2^-23, 2], numerators +0 or at least 2^-46
        float t = libm::div_ns(1.f - g * g, 1.f - g + 2.f * g * rng_u01(r));
        t = libm::div_ns(1.f + g * g - t * t, 2.f * g);
        th = libm::acosf_host(fmaxf(-1.f, fminf(1.f, t)));
    } else {
        th = libm::acosf_host(2.f * rng_u01(r) - 1.f);
    }
    libm::sincosf_host(th, &sth, &cth);
    rotate(p, sth, cth, sphi, cphi);
    // one range test for the three components (a unit vector's components never exceed 2^60)
    if (fminf(fminf(fabsf(p.vx), fabsf(p.vy)), fabsf(p.vz)) >= 0x1p-60f) {
        p.rx = libm::rcp_ns(p.vx);
        p.ry = libm::rcp_ns(p.vy);
        p.rz = libm::rcp_ns(p.vz);
    } else {
        p.rx = libm::recip_dir(p.vx);
        p.ry = libm::recip_dir(p.vy);
        p.rz = libm::recip_dir(p.vz);
    }
    p.vw += 1.f;
}

// reflectcoeff (:381-396): unpolarised Fresnel reflectance, 1 under total internal reflection
__device__ __forceinline__ float fresnel(float vaxis, float n1, float n2) {
    float ci = fabsf(vaxis);
    float a2 = n1 * n1, b2 = n2 * n2;
    float ct2 = 1.f - a2 / b2 * (1.f - ci * ci);
    if (ct2 > 0.f) {
        float re = a2 * ci * ci + b2 * ct2;
        ct2 = libm::sqrtf_cr(ct2);
        float im = 2.f * n1 * n2 * ci * ct2;
        float rp = (re - im) / (re + im);
        re = b2 * ci * ci + a2 * ct2 * ct2;
        return (rp + (re - im) / (re + im)) * 0.5f;
    }
    return 1.f;
}

// reflect (:405-417) with transmit (:397-404); true when the photon is transmitted
// Register value barrier: a select between photon fields would otherwise be rewritten as a
// select of field addresses, which forces the whole photon state out to scratch memory.
__device__ __forceinline__ float in_reg(float x) {
    __asm__ volatile("" : "+v"(x));
    return x;
}
__device__ __forceinline__ uint32_t in_reg_u32(uint32_t x) {
    __asm__ volatile("" : "+v"(x));
    return x;
}
__device__ __forceinline__ float pick_axis(bool ax, bool ay, float x, float y, float z) {
    return ax ? in_reg(x) : (ay ? in_reg(y) : in_reg(z));
}

__device__ __forceinline__ bool boundary(Photon& p, float n1, float n2, Rng& r, int& nv, uint32_t& nm) {
    const bool ax = p.iw == 0, ay = p.iw == 1, az = !ax && !ay;
    const float vaxis = pick_axis(ax, ay, p.vx, p.vy, p.vz);
    float R = fresnel(vaxis, n1, n2);
    if (R < 1.f && rng_u01(r) > R) {
        float s = n1 / n2;
        p.vx *= s; p.vy *= s; p.vz *= s;
        const float c = pick_axis(ax, ay, p.vx, p.vy, p.vz);
        const float t = p.vx * p.vx + p.vy * p.vy + p.vz * p.vz - c * c;
        const float cn = (t < 1.f) ? libm::sqrtf_cr(1.f - t) * (float)((c > 0.f) - (c < 0.f)) : 0.f;
        p.vx = ax ? cn : p.vx;
        p.vy = ay ? cn : p.vy;
        p.vz = az ? cn : p.vz;
        s = 1.f / libm::sqrtf_cr(p.vx * p.vx + p.vy * p.vy + p.vz * p.vz);
        p.vx *= s; p.vy *= s; p.vz *= s;
        return true;
    }
    // mirror on the crossed axis (REFLECT_PHOTON, :31)
    p.vx = ax ? -p.vx : p.vx;  p.rx = ax ? -p.rx : p.rx;
    p.vy = ay ? -p.vy : p.vy;  p.ry = ay ? -p.ry : p.ry;
    p.vz = az ? -p.vz : p.vz;  p.rz = az ? -p.rz : p.rz;
    const float v = -vaxis;
    const float pos = pick_axis(ax, ay, p.px, p.py, p.pz);
    const float snapped = libm::nextafterf_host((float)(int)(pos + 0.5f), (float)((v > 0.f) - (v < 0.f)));
    const int cell = to_short(snapped);
    p.px = ax ? snapped : p.px;  p.ix = ax ? cell : p.ix;
    p.py = ay ? snapped : p.py;  p.iy = ay ? cell : p.iy;
    p.pz = az ? snapped : p.pz;  p.iz = az ? cell : p.iz;
    nv = p.lastvox;
    nm = p.med;
    return false;
}

// Per-lane write-back cache of the last K global deposits, most recent first. A photon that
// steps back into a voxel it left a moment ago (same time gate) adds to the cached partial sum
// instead of issuing another global deposit; an entry reaches the grid (through `emit(cell,
// sum)`) when it is evicted or flushed. Only the float summation order of a cell changes. Returns
// true on a miss (a new entry). Keys are 32-bit fluence indices (grids of fewer than 2^32 - 1
// cells; the all-ones key marks an empty entry).
constexpr uint32_t kNoKey = ~0u;
template <int K, typename Emit>
__device__ __forceinline__ bool cached_deposit(Emit&& emit, uint32_t (&key)[K], float (&sum)[K], uint32_t idx,
                                               float val) {
    int pos = K;
#pragma unroll
    for (int j = K - 1; j >= 0; j--) pos = key[j] == idx ? j : pos;
    const bool miss = pos == K;
    float cur = 0.f;
#pragma unroll
    for (int j = 0; j < K; j++) cur = pos == j ? in_reg(sum[j]) : cur;
    if (miss) {                             // the least recent entry goes to the grid
        if (key[K - 1] != kNoKey) emit(key[K - 1], sum[K - 1]);
        pos = K - 1;
    }
#pragma unroll
    for (int j = K - 1; j >= 1; j--) {      // entries above pos move down one place
        const bool sh = j <= pos;
        key[j] = sh ? in_reg_u32(key[j - 1]) : in_reg_u32(key[j]);
        sum[j] = sh ? in_reg(sum[j - 1]) : in_reg(sum[j]);
    }
    key[0] = idx;
    sum[0] = cur + val;
    return miss;
}
template <int K, typename Emit>
__device__ __forceinline__ void flush_deposits(Emit&& emit, uint32_t (&key)[K], float (&sum)[K]) {
#pragma unroll
    for (int j = 0; j < K; j++) {
        if (key[j] != kNoKey) emit(key[j], sum[j]);
        key[j] = kNoKey;
        sum[j] = 0.f;
    }
}

__device__ __forceinline__ unsigned long long lanemask_lt() {
    return (1ull << __lane_id()) - 1ull;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}


}  // namespace dev
}  // namespace umcx
