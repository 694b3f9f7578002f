/*
 * umcx_oracle.c -- CPU restatement of umcx's photon loop. TEST INFRASTRUCTURE ONLY:
 * the checker that tests/, the smoke test and bench.py's cpu_baseline leg use.
 * The product never links or calls it.
 *
 * Written from the behaviour of fangq/umcx src/umcx.cpp; each function cites the lines
 * it restates. Built like the reference CPU target (gcc -O3 -fopenmp, no -march, so no
 * FMA contraction on x86-64; see oracle/Makefile) and it calls the host libm where the
 * reference does, so on one host it follows the reference's trajectories bit for bit.
 *
 * Pinning (tests/test_oracle.py): the seed words against glibc rand() known answers;
 * detected-photon rows, fluence and energies against golden vectors written by the real
 * reference binary (tests/golden/make_golden.py); closed forms (ballistic pencil beam,
 * normal-incidence Fresnel); energy balance.
 */
#include "umcx_oracle.h"
#include <math.h>
#include <float.h>
#include <string.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_ONE_OVER_C0 3.335640951981520e-12f   /* src/umcx.cpp:29 */
#define OR_PI          3.1415926535897932385f   /* src/umcx.cpp:30 */
#define OR_DET_MASK    0x80000000u              /* src/umcx.cpp:35 */
#define OR_MED_MASK    0x7FFFFFFFu              /* src/umcx.cpp:36 */

enum { SRC_PENCIL, SRC_ISOTROPIC, SRC_CONE, SRC_DISK, SRC_PLANAR, SRC_FOURIER };  /* :62 */
enum { OUT_FLUENCE_RATE, OUT_FLUENCE, OUT_ENERGY };                               /* :60 */
enum { DP_DETID = 1, DP_PPATH = 4, DP_EXITPOS = 16, DP_EXITDIR = 32 };           /* :65 */

/* Planted mistakes for the tests that show the golden comparisons catch plausible errors.
 * 0 = faithful (default). 1 = polar angle drawn before the azimuth. 2 = fused multiply-add in the
 * position update. 3 = Fresnel draw taken even under total internal reflection. */
static int g_mutant = 0;
void oracle_set_mutant(int m) { g_mutant = m; }

float oracle_host_logf(float x) { return logf(x); }
float oracle_host_sinf(float x) { return sinf(x); }
float oracle_host_cosf(float x) { return cosf(x); }
float oracle_host_acosf(float x) { return acosf(x); }
float oracle_host_expf(float x) { return expf(x); }

void oracle_host_libm_array(int fn, const float* in, float* out, uint64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        float x = in[i];
        out[i] = fn == 0 ? logf(x) : fn == 1 ? sinf(x) : fn == 2 ? cosf(x) : fn == 3 ? acosf(x) : expf(x);
    }
}

/* glibc srandom/random TYPE_3: r[0]=seed, r[i]=16807*r[i-1] mod (2^31-1) for i<31,
 * r[31..33]=r[0..2], r[i]=r[i-31]+r[i-3]; the k-th output is r[k+344]>>1.
 * Restates std::srand(seed) followed by four std::rand() (src/umcx.cpp:725-727). */
void oracle_seed_words(int32_t seed, uint32_t out[4]) {
    int32_t r[348];
    r[0] = (seed == 0) ? 1 : seed;
    for (int i = 1; i < 31; i++) {
        long long v = (16807LL * (long long)r[i - 1]) % 2147483647LL;
        if (v < 0) v += 2147483647LL;
        r[i] = (int32_t)v;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int i = 34; i < 348; i++) r[i] = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
    for (int k = 0; k < 4; k++) out[k] = ((uint32_t)r[344 + k]) >> 1;
}

/* ---------------- xorshift128+ (src/umcx.cpp:186-215) ---------------- */
typedef struct { uint64_t t0, t1; uint32_t ndraw; } or_rng;

static inline void rng_seed(or_rng* g, uint32_t a, uint32_t b, uint32_t c, uint32_t d) { /* :193-196 */
    g->t0 = ((uint64_t)a << 32) | b;
    g->t1 = ((uint64_t)c << 32) | d;
}
static inline float rng_u01(or_rng* g) {  /* :197-211: low 32 bits of t1+s0 -> 23-bit mantissa */
    uint64_t s0 = g->t1, s1 = g->t0;
    g->t0 = s0;
    s1 ^= s1 << 23;
    g->t1 = s1 ^ s0 ^ (s1 >> 18) ^ (s0 >> 5);
    uint32_t lo = (uint32_t)(g->t1 + s0);
    uint32_t bits = 0x3F800000u | (lo >> 9);
    float f;
    memcpy(&f, &bits, 4);
    g->ndraw++;
    return f - 1.0f;
}
static inline float rng_scatlen(or_rng* g) {  /* :212-214 */
    return -logf(rng_u01(g) + FLT_EPSILON);
}

/* ---------------- photon state (src/umcx.cpp:216-220) ---------------- */
typedef struct { float x, y, z, w; } f4;
typedef struct {
    f4 pos, vec, rvec, len;   /* len = {scattering length left, time, path since save, weight at save} */
    int16_t ix, iy, iz, iw;   /* voxel and crossing axis */
    int32_t lastvox, med;
    uint32_t nstep, ncross, nscat, nfres;
} or_photon;

typedef struct {
    const oracle_problem* p;
    uint64_t dimyz, dimxyz;
    float* fluence;
    float* detp; uint32_t* detp_id; float* detp_w; uint32_t* ndet;
    int detlen, ppathlen;
} or_ctx;

/* MCX_volume::index (:128-130); the label volume has one time frame */
static inline int32_t vox_index(const or_ctx* c, int16_t ix, int16_t iy, int16_t iz) {
    const oracle_problem* p = c->p;
    if (ix < 0 || iy < 0 || iz < 0 || ix >= (int16_t)p->nx || iy >= (int16_t)p->ny || iz >= (int16_t)p->nz) return -1;
    return (int32_t)((uint64_t)ix * c->dimyz + (uint64_t)((uint32_t)iy * p->nz) + (uint64_t)iz);
}

static inline const float* medium(const or_ctx* c, int32_t med) {
    return c->p->media + 4 * (size_t)((uint32_t)med & OR_MED_MASK);
}

/* MCX_photon::rotatevector (:418-428) */
static void rotate(or_photon* ph, float st, float ct, float sp, float cp) {
    f4* v = &ph->vec;
    if (v->z > -1.f + FLT_EPSILON && v->z < 1.f - FLT_EPSILON) {
        float t0 = 1.f - v->z * v->z;
        float t1 = st / sqrtf(t0);
        float nx = t1 * (v->x * v->z * cp - v->y * sp) + v->x * ct;
        float ny = t1 * (v->y * v->z * cp + v->x * sp) + v->y * ct;
        float nz = -t1 * t0 * cp + v->z * ct;
        v->x = nx; v->y = ny; v->z = nz;
    } else {
        float nx = st * cp, ny = st * sp, nz = (v->z > 0.f) ? ct : -ct;
        v->x = nx; v->y = ny; v->z = nz;
    }
    float s = 1.f / sqrtf(v->x * v->x + v->y * v->y + v->z * v->z);
    v->x *= s; v->y *= s; v->z *= s;
}

/* MCX_photon::launch (:225-259); returns the launch weight */
static float launch(or_photon* ph, or_rng* g, const oracle_problem* p) {
    ph->pos = (f4){p->srcpos[0], p->srcpos[1], p->srcpos[2], 1.f};
    ph->vec = (f4){p->srcdir[0], p->srcdir[1], p->srcdir[2], 0.f};
    const float* s1 = p->srcparam1;
    const float* s2 = p->srcparam2;
    if (p->srctype == SRC_ISOTROPIC || p->srctype == SRC_CONE) {        /* :229-234 */
        float phi = (OR_PI * 2.f) * rng_u01(g);
        float sphi = sinf(phi), cphi = cosf(phi);
        float th;
        if (p->srctype == SRC_ISOTROPIC) {
            th = acosf(2.f * rng_u01(g) - 1.f);
        } else {
            float ca = cosf(s1[0]);
            th = acosf(rng_u01(g) * (1.0f - ca) + ca);
        }
        rotate(ph, sinf(th), cosf(th), sphi, cphi);
    } else if (p->srctype == SRC_DISK) {                                /* :235-247 */
        float phi = (OR_PI * 2.f) * rng_u01(g);
        float sphi = sinf(phi), cphi = cosf(phi);
        float r = sqrtf(rng_u01(g) * fabsf(s1[0] * s1[0] - s1[1] * s1[1]) + s1[1] * s1[1]);
        f4 v = ph->vec;
        if (v.z > -1.f + FLT_EPSILON && v.z < 1.f - FLT_EPSILON) {
            float a = 1.f - v.z * v.z;
            float b = r / sqrtf(a);
            float px = ph->pos.x + b * (v.x * v.z * cphi - v.y * sphi);
            float py = ph->pos.y + b * (v.y * v.z * cphi + v.x * sphi);
            float pz = ph->pos.z - b * a * cphi;
            ph->pos.x = px; ph->pos.y = py; ph->pos.z = pz;
        } else {
            ph->pos.x += r * cphi;
            ph->pos.y += r * sphi;
        }
    } else if (p->srctype == SRC_PLANAR || p->srctype == SRC_FOURIER) { /* :248-253 */
        float a = rng_u01(g), b = rng_u01(g);
        float px = ph->pos.x + a * s1[0] + b * s2[0];
        float py = ph->pos.y + a * s1[1] + b * s2[1];
        float pz = ph->pos.z + a * s1[2] + b * s2[2];
        ph->pos.x = px; ph->pos.y = py; ph->pos.z = pz;
        if (p->srctype == SRC_FOURIER) {
            float ph_arg = (floorf(s1[3]) * a + floorf(s2[3]) * b + s1[3] - floorf(s1[3])) * OR_PI * 2.f;
            ph->pos.w *= (cosf(ph_arg) * (1.f - s2[3] + floorf(s2[3])) + 1.f) * 0.5f;
        }
    }
    /* :255-257 -- the reciprocal direction comes from the nominal direction */
    ph->rvec = (f4){1.f / p->srcdir[0], 1.f / p->srcdir[1], 1.f / p->srcdir[2], 0.f};
    ph->len = (f4){NAN, 0.f, 0.f, ph->pos.w};
    ph->ix = (int16_t)ph->pos.x; ph->iy = (int16_t)ph->pos.y; ph->iz = (int16_t)ph->pos.z; ph->iw = -1;
    return ph->pos.w;
}

/* MCX_photon::skip (:339-358): ray/box entry for sources outside the grid */
static float skip_to_box(or_ctx* c, or_photon* ph) {
    const oracle_problem* p = c->p;
    f4 r = ph->rvec;
    float tx0 = -ph->pos.x * r.x, ty0 = -ph->pos.y * r.y, tz0 = -ph->pos.z * r.z;
    float tx1 = ((float)p->nx - ph->pos.x) * r.x;
    float ty1 = ((float)p->ny - ph->pos.y) * r.y;
    float tz1 = ((float)p->nz - ph->pos.z) * r.z;
    float tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    float tmax = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    /* the reference reuses len/rvec.w/pos.w as scratch; rvec.w and pos.w are overwritten below */
    ph->rvec.w = ty1;
    if (tmax < 0.f || tmin > tmax) {
        ph->len = (f4){tx0, ty0, tz0, tx1};
        ph->pos.w = tz1;
        return -1.f;
    }
    ph->pos = (f4){ph->pos.x + tmin * ph->vec.x, ph->pos.y + tmin * ph->vec.y, ph->pos.z + tmin * ph->vec.z, 1.f};
    ph->len = (f4){NAN, 0.f, 0.f, 1.f};
    ph->ix = (int16_t)ph->pos.x; ph->iy = (int16_t)ph->pos.y; ph->iz = (int16_t)ph->pos.z; ph->iw = -1;
    ph->lastvox = vox_index(c, ph->ix, ph->iy, ph->iz);
    return tmin;
}

/* MCX_photon::step (:315-338): move to the nearer of the voxel face or the scattering site */
static int32_t step(or_ctx* c, or_photon* ph, const float* prop) {
    float mua = prop[0], mus = prop[1], n = prop[3];
    float hx = fabsf(((float)(ph->ix + (ph->vec.x > 0.f)) - ph->pos.x) * ph->rvec.x);
    float hy = fabsf(((float)(ph->iy + (ph->vec.y > 0.f)) - ph->pos.y) * ph->rvec.y);
    float hz = fabsf(((float)(ph->iz + (ph->vec.z > 0.f)) - ph->pos.z) * ph->rvec.z);
    float dmin = fminf(fminf(hx, hy), hz);
    ph->iw = (dmin == hx) ? 0 : ((dmin == hy) ? 1 : 2);
    float used = fminf(dmin * mus, ph->len.x);
    float cont = (used != ph->len.x) ? 1.f : 0.f;
    float d = (mus == 0.f) ? dmin : (used / mus);
    ph->rvec.w = d;
    if (g_mutant == 2) {
        ph->pos.x = fmaf(d, ph->vec.x, ph->pos.x);
        ph->pos.y = fmaf(d, ph->vec.y, ph->pos.y);
        ph->pos.z = fmaf(d, ph->vec.z, ph->pos.z);
    } else {
        ph->pos.x = ph->pos.x + d * ph->vec.x;
        ph->pos.y = ph->pos.y + d * ph->vec.y;
        ph->pos.z = ph->pos.z + d * ph->vec.z;
    }
    ph->pos.w = ph->pos.w * expf(-mua * d);
    ph->len.x -= used;
    ph->len.y += d * n * OR_ONE_OVER_C0;
    ph->len.z += d;
    ph->nstep++;
    if (cont > 0.f) {
        if (ph->iw == 0) ph->ix += (ph->vec.x > 0.f ? 1 : -1);
        else if (ph->iw == 1) ph->iy += (ph->vec.y > 0.f ? 1 : -1);
        else ph->iz += (ph->vec.z > 0.f ? 1 : -1);
        return vox_index(c, ph->ix, ph->iy, ph->iz);
    }
    return ph->lastvox;
}

/* MCX_photon::save + MCX_volume::add (:359-363, :134-137) */
static void deposit(or_ctx* c, or_photon* ph, int tshift, float mua, double* edep) {
    const oracle_problem* p = c->p;
    float val = (p->outputtype == OUT_ENERGY) ? (ph->len.w - ph->pos.w)
              : ((mua < FLT_EPSILON) ? (ph->len.w * ph->len.z) : (ph->len.w - ph->pos.w) / mua);
    *edep += (double)(ph->len.w - ph->pos.w);
    if (c->fluence) {
        int64_t idx = (int64_t)ph->lastvox + (int64_t)tshift * (int64_t)c->dimxyz;
        float* cell = c->fluence + idx;
#pragma omp atomic
        *cell += val;
    }
    ph->len.w = ph->pos.w;
    ph->len.z = 0.f;
}

/* MCX_photon::scatter (:364-380) */
static void scatter(or_photon* ph, or_rng* g, const float* prop) {
    float gg = prop[2];
    ph->len.x = rng_scatlen(g);
    float uphi = (g_mutant == 1) ? 0.f : rng_u01(g);
    float phi = (2.f * OR_PI) * uphi;
    float sphi = sinf(phi), cphi = cosf(phi);
    float th;
    if (fabsf(gg) > FLT_EPSILON) {
        float t = (1.f - gg * gg) / (1.f - gg + 2.f * gg * rng_u01(g));
        t = (1.f + gg * gg - t * t) / (2.f * gg);
        th = acosf(fmaxf(-1.f, fminf(1.f, t)));
    } else {
        th = acosf(2.f * rng_u01(g) - 1.f);
    }
    if (g_mutant == 1) {
        phi = (2.f * OR_PI) * rng_u01(g);
        sphi = sinf(phi);
        cphi = cosf(phi);
    }
    float sth = sinf(th), cth = cosf(th);
    rotate(ph, sth, cth, sphi, cphi);
    ph->rvec = (f4){1.f / ph->vec.x, 1.f / ph->vec.y, 1.f / ph->vec.z, cphi};
    ph->vec.w++;
    ph->nscat++;
}

static inline float axis_comp(const f4* v, int16_t ax) { return ax == 0 ? v->x : (ax == 1 ? v->y : v->z); }

/* MCX_photon::reflectcoeff (:381-396): unpolarised Fresnel reflectance, 1 for TIR */
static float fresnel(const or_photon* ph, float n1, float n2) {
    float ci = fabsf(axis_comp(&ph->vec, ph->iw));
    float a = n1 * n1, b = n2 * n2;
    float ct2 = 1.f - a / b * (1.f - ci * ci);
    if (ct2 > 0.f) {
        float re = a * ci * ci + b * ct2;
        ct2 = sqrtf(ct2);
        float im = 2.f * n1 * n2 * ci * ct2;
        float rp = (re - im) / (re + im);
        re = b * ci * ci + a * ct2 * ct2;
        return (rp + (re - im) / (re + im)) * 0.5f;
    }
    return 1.f;
}

/* MCX_photon::transmit (:397-404) with TRANSMIT_PHOTON (:32) */
static void refract(or_photon* ph, float n1, float n2) {
    float s = n1 / n2;
    ph->vec.x *= s; ph->vec.y *= s; ph->vec.z *= s;
    float* comp = ph->iw == 0 ? &ph->vec.x : (ph->iw == 1 ? &ph->vec.y : &ph->vec.z);
    float t = ph->vec.x * ph->vec.x + ph->vec.y * ph->vec.y + ph->vec.z * ph->vec.z - *comp * *comp;
    *comp = (t < 1.f) ? sqrtf(1.f - t) * (float)((*comp > 0.f) - (*comp < 0.f)) : 0.f;
    s = 1.f / sqrtf(ph->vec.x * ph->vec.x + ph->vec.y * ph->vec.y + ph->vec.z * ph->vec.z);
    ph->vec.x *= s; ph->vec.y *= s; ph->vec.z *= s;
}

/* MCX_photon::reflect (:405-417) with REFLECT_PHOTON (:31); returns 1 when transmitted */
static int boundary(or_ctx* c, or_photon* ph, float n1, float n2, or_rng* g, int32_t* newvox, int32_t* newmed) {
    (void)c;
    float R = fresnel(ph, n1, n2);
    ph->nfres++;
    if (g_mutant == 3 && R >= 1.f) (void)rng_u01(g);
    if (R < 1.f && rng_u01(g) > R) {
        refract(ph, n1, n2);
        return 1;
    }
    float* v; float* rv; float* ps; int16_t* ip;
    if (ph->iw == 0) { v = &ph->vec.x; rv = &ph->rvec.x; ps = &ph->pos.x; ip = &ph->ix; }
    else if (ph->iw == 1) { v = &ph->vec.y; rv = &ph->rvec.y; ps = &ph->pos.y; ip = &ph->iy; }
    else { v = &ph->vec.z; rv = &ph->rvec.z; ps = &ph->pos.z; ip = &ph->iz; }
    *v = -*v;
    *rv = -*rv;
    *ps = nextafterf((float)(int)(*ps + 0.5f), (float)((*v > 0.f) - (*v < 0.f)));
    *ip = (int16_t)*ps;
    *newvox = ph->lastvox;
    *newmed = ph->med;
    return 0;
}

/* Per-voxel segment accounting of the replay (oracle_replay): at every voxel change the path
 * since the previous change (len.z, in voxels) and the scatterings taken meanwhile are credited
 * to the voxel the photon leaves. */
typedef struct {
    double coef;                  /* record coefficient */
    double* gmua; double* gmus;   /* gradient slice of the record */
    double* amua; double* amus;   /* sums of |term| per entry (may be NULL) */
    float* mpath;                 /* per-medium sums of the segments, float, in crossing order */
    const float* vprops;          /* nvox x {mua, mus} per voxel length, or NULL: the labels' media */
    float unit;                   /* mm per voxel */
    uint32_t nscat;               /* scatterings since the last voxel change */
} or_seg;

static inline void seg_add(double* a, double v) {
#pragma omp atomic
    *a += v;
}

static void seg_flush(or_seg* sg, const or_ctx* c, const or_photon* ph, int32_t nv) {
    const int32_t v = (g_mutant == 4 && nv >= 0) ? nv : ph->lastvox;   /* mutant 4: credit the new voxel */
    const float mus_vox = sg->vprops ? sg->vprops[2 * (size_t)v + 1] : medium(c, ph->med)[1];
    const double L = (double)ph->len.z * (double)sg->unit;              /* mm */
    const double tmua = -sg->coef * L;
    const double tmus = sg->coef * ((sg->nscat ? (double)sg->nscat * (double)sg->unit / (double)mus_vox : 0.0) - L);
    seg_add(sg->gmua + v, tmua);
    seg_add(sg->gmus + v, tmus);
    if (sg->amua) seg_add(sg->amua + v, fabs(tmua));
    if (sg->amus) seg_add(sg->amus + v, fabs(sg->coef) * ((sg->nscat ? (double)sg->nscat * (double)sg->unit / (double)mus_vox : 0.0) + L));
    const uint32_t m = (uint32_t)ph->med & OR_MED_MASK;
    if (sg->mpath && m > 0) sg->mpath[m - 1] += ph->len.z;
    sg->nscat = 0;
}

/* MCX_photon::sprint (:280-314): returns 1 when the photon terminates. With `sg` (replay), every
 * voxel change is also credited to the segment accounts and len.z restarts there. */
static int fly_seg(or_ctx* c, or_photon* ph, or_rng* g, float* ppath, double* edep, or_seg* sg) {
    const oracle_problem* p = c->p;
    while (ph->len.x > 0.f) {
        const float* prop = medium(c, ph->med);
        float vp[4];
        if (sg && sg->vprops) {                    /* per-voxel coefficients replace the label's */
            vp[0] = sg->vprops[2 * (size_t)ph->lastvox]; vp[1] = sg->vprops[2 * (size_t)ph->lastvox + 1];
            vp[2] = prop[2]; vp[3] = prop[3];
            prop = vp;
        }
        int32_t nv = step(c, ph, prop);
        if (nv == ph->lastvox) continue;
        ph->ncross++;
        uint32_t m = (uint32_t)ph->med & OR_MED_MASK;
        if (sg) {
            seg_flush(sg, c, ph, nv);
            ph->len.z = 0.f;
        }
        if (p->issavedet && (p->savedetflag & DP_PPATH) && m > 0) ppath[m - 1] += ph->len.z;
        if (p->issavevol) {
            int tshift = (int)fminf((float)(p->maxgate - 1), (float)(int)floorf((ph->len.y - p->tstart) * p->rtstep));
            deposit(c, ph, tshift, prop[0], edep);
        }
        if (ph->len.y > p->tend) return 1;
        int32_t nm = (nv >= 0) ? p->labels[nv] : 0;
        const float* nprop = medium(c, nm);
        if (p->isreflect && prop[3] != nprop[3]) {
            if (boundary(c, ph, prop[3], nprop[3], g, &nv, &nm) && (nv < 0 || ((uint32_t)nm & OR_MED_MASK) == 0)) return 1;
        } else if (nv < 0 || (((uint32_t)nm & OR_MED_MASK) == 0 && m != 0)) {
            return 1;
        }
        ph->lastvox = nv;
        ph->med = nm;
    }
    return 0;
}

static int fly(or_ctx* c, or_photon* ph, or_rng* g, float* ppath, double* edep) {
    return fly_seg(c, ph, g, ppath, edep, NULL);
}

/* MCX_detect::addphoton (:164-176) called from savedetector (:429-435) */
static void record(or_ctx* c, const or_photon* ph, const float* ppath, uint64_t id) {
    const oracle_problem* p = c->p;
    for (int i = 0; i < p->detnum; i++) {
        const float* d = p->detpos + 4 * i;
        float dx = d[0] - ph->pos.x, dy = d[1] - ph->pos.y, dz = d[2] - ph->pos.z;
        if (dx * dx + dy * dy + dz * dz < d[3] * d[3]) {
            uint32_t slot;
#pragma omp atomic capture
            slot = (*c->ndet)++;
            if (slot < p->maxdetphotons && c->detp) {
                float* row = c->detp + (size_t)slot * c->detlen;
                int k = 0;
                if (p->savedetflag & DP_DETID) row[k++] = (float)(i + 1);
                for (int j = 0; j < c->ppathlen; j++) row[k++] = ppath[j];
                if (p->savedetflag & DP_EXITPOS) { row[k++] = ph->pos.x; row[k++] = ph->pos.y; row[k++] = ph->pos.z; }
                if (p->savedetflag & DP_EXITDIR) { row[k++] = ph->vec.x; row[k++] = ph->vec.y; row[k++] = ph->vec.z; }
                if (c->detp_id) c->detp_id[slot] = (uint32_t)id;
                if (c->detp_w) c->detp_w[slot] = ph->pos.w;
            }
        }
    }
}

int oracle_detp_len(const oracle_problem* p) {  /* :157 */
    return (p->savedetflag & DP_DETID) + ((p->savedetflag & DP_PPATH) ? p->mediumnum : 0)
         + (((p->savedetflag & DP_EXITPOS) ? 1 : 0) + ((p->savedetflag & DP_EXITDIR) ? 1 : 0)) * 3;
}

/* MCX_kernel loop body (:752-775) plus MCX_photon::run (:260-279) */
int oracle_run(const oracle_problem* p, uint64_t first, uint64_t count, float* fluence,
               float* detp, uint32_t* detp_id, float* detp_w, uint32_t* ndetected,
               double energies[3], oracle_trace* traces, int nthreads) {
    if (!p || !p->labels || !p->media || p->mediumnum <= 0 || (p->detnum > 0 && !p->detpos) || !energies) return -1;
    or_ctx c;
    memset(&c, 0, sizeof c);
    c.p = p;
    c.dimyz = (uint64_t)p->ny * p->nz;
    c.dimxyz = c.dimyz * p->nx;
    c.fluence = p->issavevol ? fluence : NULL;
    c.detp = detp; c.detp_id = detp_id; c.detp_w = detp_w;
    uint32_t dummy = 0;
    c.ndet = ndetected ? ndetected : &dummy;
    c.detlen = oracle_detp_len(p);
    c.ppathlen = (p->issavedet && (p->savedetflag & DP_PPATH)) ? p->mediumnum : 0;
    const int savedet = p->issavedet && p->detnum > 0;
    double esc = 0.0, tot = 0.0, dep = 0.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    int64_t n = (int64_t)count;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : esc, tot, dep)
    for (int64_t k = 0; k < n; k++) {
        uint64_t i = first + (uint64_t)k;
        float ppath[64];
        float* pp = ppath;
        float* heap = NULL;
        if (c.ppathlen > 64) pp = heap = (float*)calloc((size_t)c.ppathlen, sizeof(float));
        else memset(ppath, 0, sizeof ppath);
        or_rng g;
        or_photon ph;
        memset(&ph, 0, sizeof ph);
        g.ndraw = 0;
        rng_seed(&g, p->seeds[0] ^ (uint32_t)i, p->seeds[1] | (uint32_t)i, p->seeds[2] ^ (uint32_t)i, p->seeds[3] | (uint32_t)i); /* :761 */
        float w0 = launch(&ph, &g, p);
        double edep = 0.0;
        ph.lastvox = vox_index(&c, ph.ix, ph.iy, ph.iz);                 /* :262 */
        int entered = 1;
        if (ph.lastvox < 0 && skip_to_box(&c, &ph) < 0.f) { ph.med = 0; entered = 0; }  /* :264-267 */
        if (entered && ph.lastvox < 0) { ph.med = 0; entered = 0; }      /* grid entered exactly at a far face */
        if (entered) {
            ph.med = p->labels[ph.lastvox];                              /* :269 */
            ph.len.x = rng_scatlen(&g);                                  /* :270 */
            while (fly(&c, &ph, &g, pp, &edep) == 0)                     /* :272-274 */
                scatter(&ph, &g, medium(&c, ph.med));
            if (savedet && ((uint32_t)ph.med & OR_DET_MASK) && ph.len.y <= p->tend)  /* :276-278 */
                record(&c, &ph, pp, i);
        }
        if (((uint32_t)ph.med & OR_MED_MASK) != 0) {                    /* :765-770 */
            tot += w0;
            esc += ph.pos.w;
            dep += edep;
        }
        if (traces) {
            oracle_trace* t = traces + k;
            memcpy(t->pos, &ph.pos, 16); memcpy(t->vec, &ph.vec, 16); memcpy(t->len, &ph.len, 16);
            t->mediaid = ph.med; t->lastvoxelidx = ph.lastvox; t->w0 = w0;
            t->nstep = ph.nstep; t->ncross = ph.ncross; t->nscat = ph.nscat; t->nrand = g.ndraw; t->nfresnel = ph.nfres;
        }
        free(heap);
    }
    energies[0] += esc;
    energies[1] += tot;
    energies[2] += dep;
    return 0;
}

/* Replay for the differentiable form (perturbation Monte Carlo per voxel). Re-traces photon
 * ids[k] -- photon i is a pure function of i (:761) -- launched from source src_index[k] of
 * `sources` ({pos xyz, dir xyz} per source; NULL: the problem's source), and credits every voxel
 * segment (the path between two voxel changes, as sprint :285-292 delimits it) to the voxel the
 * photon leaves:
 *   gmua[slices[k]*stride + v] += coef[k] * (-L)
 *   gmus[slices[k]*stride + v] += coef[k] * (n / mus_v - L)
 * with L the segment length in mm, n the scatterings taken in it (:364-380) and mus_v in 1/mm.
 * Summed over a photon's segments these are d w / d mua_v / w and d log p / d mus_v of its final
 * weight w and path density p. amua (may be NULL) receives the sums of the mua terms' absolute
 * values per entry, amus (may be NULL) the sums of |coef| (n / mus_v + L), the magnitudes the
 * mus terms are formed from (the scale of their rounding); final_w (may be NULL) the replayed final weights; mpath (may be NULL,
 * mediumnum floats per record) the per-medium sums of the segment lengths in voxels, accumulated
 * in float in crossing order like the reference's partial paths (:286-288). vprops (may be NULL)
 * holds per-voxel {mua, mus} in per-voxel-length units replacing the labels' (g and n stay per
 * label). Accumulation in double. Returns 0, or -1 on invalid input. */
int oracle_replay(const oracle_problem* p0, const uint32_t* ids, const int32_t* src_index, const float* sources,
                  const int32_t* slices, const double* coef, uint64_t n, const float* vprops, double* gmua,
                  double* gmus, double* amua, double* amus, uint64_t stride, float* final_w, float* mpath,
                  int nthreads) {
    if (!p0 || !p0->labels || !p0->media || p0->mediumnum <= 0 || !ids || !slices || !coef || !gmua || !gmus) return -1;
    oracle_problem pr = *p0;
    pr.issavevol = 0;           /* the replay tallies no fluence and saves no detector rows */
    pr.issavedet = 0;
    or_ctx c;
    memset(&c, 0, sizeof c);
    c.p = &pr;
    c.dimyz = (uint64_t)pr.ny * pr.nz;
    c.dimxyz = c.dimyz * pr.nx;
    uint32_t dummy = 0;
    c.ndet = &dummy;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const int nmed = pr.mediumnum;
    int64_t nn = (int64_t)n;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < nn; k++) {
        oracle_problem ps = pr;
        if (sources && src_index) {
            const float* q = sources + 6 * (size_t)src_index[k];
            ps.srcpos[0] = q[0]; ps.srcpos[1] = q[1]; ps.srcpos[2] = q[2];
            ps.srcdir[0] = q[3]; ps.srcdir[1] = q[4]; ps.srcdir[2] = q[5];
        }
        or_ctx ck = c;
        ck.p = &ps;
        const uint64_t i = ids[k];
        or_seg sg;
        memset(&sg, 0, sizeof sg);
        sg.coef = coef[k];
        sg.gmua = gmua + (uint64_t)slices[k] * stride;
        sg.gmus = gmus + (uint64_t)slices[k] * stride;
        sg.amua = amua ? amua + (uint64_t)slices[k] * stride : NULL;
        sg.amus = amus ? amus + (uint64_t)slices[k] * stride : NULL;
        sg.mpath = mpath ? mpath + (size_t)k * nmed : NULL;
        if (sg.mpath) memset(sg.mpath, 0, sizeof(float) * (size_t)nmed);
        sg.vprops = vprops;
        sg.unit = ps.unitinmm;
        or_rng g;
        or_photon ph;
        memset(&ph, 0, sizeof ph);
        g.ndraw = 0;
        rng_seed(&g, ps.seeds[0] ^ (uint32_t)i, ps.seeds[1] | (uint32_t)i, ps.seeds[2] ^ (uint32_t)i, ps.seeds[3] | (uint32_t)i);
        (void)launch(&ph, &g, &ps);
        double edep = 0.0;
        float pp[1] = {0.f};
        ph.lastvox = vox_index(&ck, ph.ix, ph.iy, ph.iz);
        int entered = 1;
        if (ph.lastvox < 0 && skip_to_box(&ck, &ph) < 0.f) entered = 0;
        if (entered && ph.lastvox < 0) entered = 0;
        if (entered) {
            ph.med = ps.labels[ph.lastvox];
            ph.len.x = rng_scatlen(&g);
            while (fly_seg(&ck, &ph, &g, pp, &edep, &sg) == 0) {
                scatter(&ph, &g, medium(&ck, ph.med));
                sg.nscat++;
            }
        }
        if (final_w) final_w[k] = ph.pos.w;
    }
    return 0;
}
