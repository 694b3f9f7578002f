// capi.cpp -- the C ABI of libumcx.so (include/umcx.h) and MCX_run_simulation.
#include "../../include/umcx.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <memory>
#include <string>

#include "batch.h"
#include "tally_box.h"
#include "bjdata.h"
#include "config.h"
#include "glibc_libm.h"
#include "runtime.h"

struct umcx_session {
    std::unique_ptr<umcx::Session> s;
};

struct umcx_batch {
    std::unique_ptr<umcx::BatchSession> b;
};

namespace {

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return 1;
    } catch (...) {
        g_last_error = "unknown error";
        return 1;
    }
}

umcx::LaunchConfig to_launch(const umcx_launch* l) {
    umcx::LaunchConfig lc;
    if (l) {
        lc.device = l->device;
        lc.block = l->block;
        lc.blocks = l->blocks;
        lc.tally_mode = l->tally_mode;
        lc.trace = l->trace != 0;
        lc.vote = l->vote;
        lc.label_cache = l->label_cache;
        lc.deposit_cache = l->deposit_cache;
        lc.deposit_log = l->deposit_log;
        lc.deposit_log_segments = l->deposit_log_segments;
    }
    return lc;
}

// handles passed in by C callers: a null or freed-looking handle is an error, not a crash
umcx::Session& session_of(const umcx_session* s, const char* fn) {
    if (!s || !s->s) throw std::runtime_error(std::string(fn) + ": null session");
    return *s->s;
}

umcx::BatchSession& batch_of(const umcx_batch* b, const char* fn) {
    if (!b || !b->b) throw std::runtime_error(std::string(fn) + ": null batch");
    return *b->b;
}

// Normaliser of src/umcx.cpp:820: double arithmetic, stored as float
float normalizer_of(const umcx::Problem& p, double energytotal) {
    if (p.outputtype == umcx::kEnergy) return (float)(1.f / energytotal);
    if (p.outputtype == umcx::kFluenceRate)
        return (float)(p.rtstep / (energytotal * p.unitinmm * p.unitinmm));
    return (float)(1.f / (energytotal * p.unitinmm * p.unitinmm));
}

std::string out_name(const umcx::Problem& p, const char* suffix, const char* dflt) {
    return p.session_id.empty() ? std::string(dflt) : p.session_id + suffix;
}

void fill_info(const umcx::Problem& p, umcx_info* info) {
    memset(info, 0, sizeof *info);
    info->dims[0] = p.nx; info->dims[1] = p.ny; info->dims[2] = p.nz; info->dims[3] = (uint32_t)p.maxgate;
    info->nphoton = p.nphoton;
    info->mediumnum = p.mediumnum; info->detnum = p.detnum;
    info->detp_columns = p.detp_columns(); info->ppath_columns = p.ppath_columns();
    info->outputtype = p.outputtype; info->isnormalized = p.isnormalized; info->issavevol = p.issavevol;
    info->issavedet = p.issavedet; info->isreflect = p.isreflect; info->srctype = p.srctype;
    info->tstart = p.tstart; info->tend = p.tend; info->rtstep = p.rtstep; info->unitinmm = p.unitinmm;
    for (int i = 0; i < 4; i++) info->seeds[i] = p.seeds[i];
    info->label_bytes = p.mediumnum <= 128 ? 1 : 4;
}

}  // namespace

int MCX_run_simulation(char* argv[], int argn, int nlhs, void* plhs[]) {
    (void)nlhs;
    (void)plhs;
    return guarded([&]() -> int {
        umcx::ParsedInput in = umcx::parse_argv(argn, argv);
        if (in.action == umcx::Action::Help) {
            std::cout << umcx::help_text() << std::flush;
            return 0;
        }
        if (in.action == umcx::Action::DumpJson) {
            uint32_t nx, ny, nz;
            std::vector<int32_t> lab;
            umcx::build_domain(in.cfg, nx, ny, nz, lab);   // the reference builds the domain first
            std::cout << in.cfg.dump(2) << std::endl;
            return 0;
        }
        if (in.action == umcx::Action::DumpMask) {
            uint32_t nx, ny, nz;
            std::vector<int32_t> lab;
            umcx::build_domain(in.cfg, nx, ny, nz, lab);
            const umcx::Json* s = in.cfg.find("Session");
            std::string name = (s && s->contains("ID")) ? s->at("ID").as_string() + "_vol.bnii" : std::string("vol.bnii");
            umcx::write_bnii_int32(name, lab.data(), nx, ny, nz, 1);
            return 0;
        }
        umcx::Problem prob = umcx::make_problem(in.cfg);
        umcx::LaunchConfig lc;
        lc.device = prob.deviceid > 0 ? prob.deviceid - 1 : 0;
        if (umcx::device_count() <= 0) throw std::runtime_error("no HIP device available: the umcx GPU path needs an MI355X (gfx950)");
        if (hipSetDevice(lc.device) != hipSuccess || hipFree(nullptr) != hipSuccess)
            throw std::runtime_error("cannot initialise HIP device " + std::to_string(lc.device));
        // host outputs are allocated before the timer, as the reference allocates its output
        // volume and detected-photon buffer (:797-799) before starting it (:812)
        std::vector<float> flu(prob.issavevol ? (size_t)prob.nx * prob.ny * prob.nz * (size_t)prob.maxgate : 0);
        std::vector<float> rows(prob.issavedet ? (size_t)prob.maxdetphotons * std::max(prob.detp_columns(), 1) : 0);
        // timing window of src/umcx.cpp:812-840: device upload, the photon loop, read-back
        auto t0 = std::chrono::steady_clock::now();
        umcx::Session sess(prob, lc);
        sess.run(0, prob.nphoton);
        if (flu.size() != sess.fluence_len()) throw std::runtime_error("fluence size mismatch");
        sess.fetch_fluence(flu.data());
        uint32_t saved = sess.saved();
        if (prob.issavedet) sess.fetch_detp(rows.data(), nullptr, nullptr);
        double e[3];
        sess.energies(e);
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        double escape = e[0], total = e[1];
        float norm = normalizer_of(prob, total);
        printf("simulated energy %.2f, speed %.2f photon/ms, duration %.6f ms, normalizer %.6f, detected %d, absorbed %.6f%%\n",
               total, (double)prob.nphoton / ms, ms, norm, (int)saved, (total - escape) / total * 100.);
        fflush(stdout);
        if (prob.issavevol) {
            if (prob.isnormalized && norm != 1.f)
                for (float& v : flu) v *= norm;
            umcx::write_bnii_float(out_name(prob, ".bnii", "output.bnii"), flu.data(), prob.nx, prob.ny, prob.nz,
                                   (uint32_t)prob.maxgate);
        }
        if (prob.issavedet) {
            umcx::DetpInfo info{prob.mediumnum, prob.detnum, prob.detp_columns(), sess.detected(), saved, prob.unitinmm};
            umcx::write_detp(out_name(prob, "_detp.jdb", "output_detp.jdb"), info, rows.data());
        }
        return 0;
    });
}

extern "C" {

int MCX_run_cmd(char* argv[], int argn) { return MCX_run_simulation(argv, argn); }

int MCX_run_json(char* jsoninput) {
    char* flags[] = {(char*)"", (char*)"--json", jsoninput};
    return MCX_run_simulation(flags, 3);
}

const char* umcx_last_error(void) { return g_last_error.c_str(); }
const char* umcx_version(void) { return "umcx-mi355x 0.1 (gfx950)"; }
int umcx_device_count(void) { return umcx::device_count(); }

int umcx_session_from_json(const char* json, const umcx_launch* launch, umcx_session** out) {
    return guarded([&]() -> int {
        if (!json || !out) throw std::runtime_error("umcx_session_from_json: null argument");
        umcx::Problem p = umcx::make_problem(umcx::Json::parse(json));
        auto* s = new umcx_session;
        try {
            s->s.reset(new umcx::Session(p, to_launch(launch)));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return 0;
    });
}

int umcx_session_from_argv(char** argv, int argn, const umcx_launch* launch, umcx_session** out) {
    return guarded([&]() -> int {
        if (!out) throw std::runtime_error("umcx_session_from_argv: null argument");
        umcx::ParsedInput in = umcx::parse_argv(argn, argv);
        umcx::Problem p = umcx::make_problem(in.cfg);
        auto* s = new umcx_session;
        try {
            s->s.reset(new umcx::Session(p, to_launch(launch)));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
        return 0;
    });
}

int umcx_session_info(const umcx_session* s, umcx_info* info) {
    return guarded([&]() -> int {
        const umcx::Session& ss = session_of(s, "umcx_session_info");
        if (!info) throw std::runtime_error("umcx_session_info: null argument");
        fill_info(ss.problem(), info);
        info->block = ss.block(); info->blocks = ss.blocks(); info->tally_mode = ss.tally_mode();
        info->lds_label_bits = ss.lds_label_bits();
        info->last_block = ss.last_block(); info->last_blocks = ss.last_blocks();
        info->deposit_cache = ss.deposit_cache() ? 1 : 0;
        info->deposit_cache_hits = ss.deposit_cache_hits();
        info->deposit_log = ss.deposit_log() ? 1 : 0;
        info->global_deposits_per_photon = ss.global_deposits_per_photon();
        info->vote = ss.vote();
        return 0;
    });
}

int umcx_session_reset(umcx_session* s) {
    return guarded([&]() -> int { session_of(s, "umcx_session_reset").reset(); return 0; });
}

int umcx_session_run(umcx_session* s, uint64_t first, uint64_t count, double* kernel_ms) {
    return guarded([&]() -> int {
        umcx::RunStats st = session_of(s, "umcx_session_run").run(first, count);
        if (kernel_ms) *kernel_ms = st.kernel_ms;
        return 0;
    });
}

int umcx_session_fluence(const umcx_session* s, float* out, uint64_t n) {
    return guarded([&]() -> int {
        const umcx::Session& ss = session_of(s, "umcx_session_fluence");
        if (!out || n < ss.fluence_len()) throw std::runtime_error("umcx_session_fluence: buffer too small");
        ss.fetch_fluence(out);
        return 0;
    });
}

int umcx_session_fluence_device(const umcx_session* s, float** dptr) {
    return guarded([&]() -> int { *dptr = session_of(s, "umcx_session_fluence_device").device_fluence(); return 0; });
}

int umcx_session_detected(const umcx_session* s, uint32_t* detected, uint32_t* saved) {
    return guarded([&]() -> int {
        const umcx::Session& ss = session_of(s, "umcx_session_detected");
        if (detected) *detected = ss.detected();
        if (saved) *saved = ss.saved();
        return 0;
    });
}

int umcx_session_detp(const umcx_session* s, float* rows, uint32_t* ids, float* weights) {
    return guarded([&]() -> int { session_of(s, "umcx_session_detp").fetch_detp(rows, ids, weights); return 0; });
}

int umcx_session_energies(const umcx_session* s, double out[3]) {
    return guarded([&]() -> int { session_of(s, "umcx_session_energies").energies(out); return 0; });
}

int umcx_session_trace(const umcx_session* s, void* out, uint64_t n) {
    return guarded([&]() -> int { session_of(s, "umcx_session_trace").fetch_trace((umcx::TraceRec*)out, n); return 0; });
}

int umcx_session_set_voxel_props(umcx_session* s, const float* mua, const float* mus, int32_t on_device) {
    return guarded([&]() -> int { session_of(s, "umcx_session_set_voxel_props").set_voxel_props(mua, mus, on_device != 0); return 0; });
}

int umcx_session_set_source(umcx_session* s, const float pos[3], const float dir[3]) {
    return guarded([&]() -> int { session_of(s, "umcx_session_set_source").set_source(pos, dir); return 0; });
}

int umcx_session_detp_device(const umcx_session* s, uint32_t** ids, float** weights, float** rows) {
    return guarded([&]() -> int {
        if (ids) *ids = session_of(s, "umcx_session_detp_device").device_detid();
        if (weights) *weights = session_of(s, "umcx_session_detp_device").device_detw();
        if (rows) *rows = session_of(s, "umcx_session_detp_device").device_detp();
        return 0;
    });
}

int umcx_session_replay(umcx_session* s, const uint32_t* ids, const int32_t* slices, const float* coef, uint64_t n,
                        float* grad_mua, float* grad_mus, uint64_t slice_stride, float* final_w, double* kernel_ms) {
    return guarded([&]() -> int {
        umcx::RunStats st = session_of(s, "umcx_session_replay").replay(ids, slices, coef, n, grad_mua, grad_mus, slice_stride, final_w);
        if (kernel_ms) *kernel_ms = st.kernel_ms;
        return 0;
    });
}

int umcx_session_set_sources(umcx_session* s, const float* sources, int32_t nsrc, uint64_t photons_per_source) {
    return guarded([&]() -> int { session_of(s, "umcx_session_set_sources").set_sources(sources, nsrc, photons_per_source); return 0; });
}

int umcx_session_replay_sources(umcx_session* s, const uint32_t* ids, const int32_t* slices, const float* coef,
                                const int32_t* src_index, const float* sources, uint64_t n, float* grad_mua,
                                float* grad_mus, uint64_t slice_stride, float* final_w, double* kernel_ms) {
    return guarded([&]() -> int {
        if (!src_index || !sources) throw std::runtime_error("umcx_session_replay_sources: null source arrays");
        umcx::RunStats st = session_of(s, "umcx_session_replay_sources").replay(ids, slices, coef, n, grad_mua, grad_mus, slice_stride, final_w, src_index,
                                         sources);
        if (kernel_ms) *kernel_ms = st.kernel_ms;
        return 0;
    });
}

int umcx_session_tally_box(const umcx_session* s, int32_t box[8]) {
    return guarded([&]() -> int {
        for (int i = 0; i < 8; i++) box[i] = session_of(s, "umcx_session_tally_box").tile()[i];
        return 0;
    });
}

float umcx_session_normalizer(const umcx_session* s, double energytotal) {
    float v = std::numeric_limits<float>::quiet_NaN();
    guarded([&]() -> int { v = normalizer_of(session_of(s, "umcx_session_normalizer").problem(), energytotal); return 0; });
    return v;
}

void umcx_session_free(umcx_session* s) { delete s; }

int umcx_batch_create(const char* const* jsons, int32_t njobs, const umcx_launch* launch, umcx_batch** out) {
    return guarded([&]() -> int {
        if (!jsons || njobs <= 0 || !out) throw std::runtime_error("umcx_batch_create: need at least one job");
        // jobs are parsed and their domains built in parallel; the first failing job's error is reported
        std::vector<umcx::Problem> probs((size_t)njobs);
        std::vector<std::string> errs((size_t)njobs);
#pragma omp parallel for schedule(dynamic, 1)
        for (int32_t j = 0; j < njobs; j++) {
            try {
                if (!jsons[j]) throw std::runtime_error("umcx_batch_create: null job config");
                probs[j] = umcx::make_problem(umcx::Json::parse(jsons[j]));
            } catch (const std::exception& e) {
                errs[j] = e.what();
            } catch (...) {
                errs[j] = "unknown error";
            }
        }
        for (int32_t j = 0; j < njobs; j++)
            if (!errs[j].empty()) throw std::runtime_error("job " + std::to_string(j) + ": " + errs[j]);
        auto* b = new umcx_batch;
        try {
            b->b.reset(new umcx::BatchSession(probs, to_launch(launch)));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
        return 0;
    });
}

int umcx_batch_run(umcx_batch* b, double* kernel_ms) {
    return guarded([&]() -> int {
        umcx::RunStats st = batch_of(b, "umcx_batch_run").run();
        if (kernel_ms) *kernel_ms = st.kernel_ms;
        return 0;
    });
}

int umcx_batch_info(const umcx_batch* b, int32_t job, umcx_info* info) {
    return guarded([&]() -> int {
        const umcx::BatchSession& bs = batch_of(b, "umcx_batch_info");
        if (!info) throw std::runtime_error("umcx_batch_info: null argument");
        fill_info(bs.problem(job), info);
        info->block = bs.block();
        info->blocks = bs.blocks();
        info->tally_mode = bs.tile_cells() > 0 ? 1 : 0;
        return 0;
    });
}

int umcx_batch_fluence(const umcx_batch* b, int32_t job, float* out, uint64_t n) {
    return guarded([&]() -> int {
        const umcx::BatchSession& bs = batch_of(b, "umcx_batch_fluence");
        if (!out || n < bs.fluence_len(job)) throw std::runtime_error("umcx_batch_fluence: buffer too small");
        bs.fetch_fluence(job, out);
        return 0;
    });
}

int umcx_batch_energies(const umcx_batch* b, int32_t job, double out[3]) {
    return guarded([&]() -> int { batch_of(b, "umcx_batch_energies").energies(job, out); return 0; });
}

int umcx_batch_detected(const umcx_batch* b, int32_t job, uint32_t* detected, uint32_t* saved) {
    return guarded([&]() -> int {
        const umcx::BatchSession& bs = batch_of(b, "umcx_batch_detected");
        if (job < 0 || job >= bs.njobs()) throw std::runtime_error("umcx_batch_detected: job index out of range");
        if (detected) *detected = bs.detected(job);
        if (saved) *saved = bs.saved(job);
        return 0;
    });
}

int umcx_batch_detp(const umcx_batch* b, int32_t job, float* rows, uint32_t* ids, float* weights) {
    return guarded([&]() -> int {
        const umcx::BatchSession& bs = batch_of(b, "umcx_batch_detp");
        if (job < 0 || job >= bs.njobs()) throw std::runtime_error("umcx_batch_detp: job index out of range");
        bs.fetch_detp(job, rows, ids, weights);
        return 0;
    });
}

int umcx_batch_trials(const umcx_batch* b, umcx_batch_trial* out, int32_t cap, int32_t* n) {
    return guarded([&]() -> int {
        const umcx::BatchSession& bs = batch_of(b, "umcx_batch_trials");
        const auto& t = bs.trials();
        if (n) *n = (int32_t)t.size();
        for (int32_t i = 0; out && i < cap && i < (int32_t)t.size(); i++) {
            out[i].block = t[i].block;
            out[i].blocks = t[i].blocks;
            out[i].tile_cells = t[i].tile_cells;
            out[i].chosen = t[i].block == bs.block() && t[i].blocks == bs.blocks() && t[i].tile_cells == bs.tile_cells();
            out[i].ms = t[i].ms;
            out[i].tile_share = t[i].tile_share;
        }
        return 0;
    });
}

int umcx_batch_tally_box(const umcx_batch* b, int32_t job, int32_t box[8], double* deposits_per_photon) {
    return guarded([&]() -> int {
        const umcx::BatchSession& bs = batch_of(b, "umcx_batch_tally_box");
        if (!box) throw std::runtime_error("umcx_batch_tally_box: null argument");
        const umcx::TallyBox& bx = bs.box(job);
        for (int i = 0; i < 8; i++) box[i] = bx[i];
        if (deposits_per_photon) *deposits_per_photon = bs.cost(job);
        return 0;
    });
}

void umcx_batch_free(umcx_batch* b) { delete b; }

int umcx_run_device_noalloc(const umcx_device_problem* p, uint64_t first, uint64_t count, void* stream) {
    return guarded([&]() -> int { umcx::run_device_noalloc(p, first, count, stream); return 0; });
}

int umcx_domain_from_json(const char* json, uint32_t dims[3], int32_t** labels) {
    return guarded([&]() -> int {
        if (!json || !dims || !labels) throw std::runtime_error("umcx_domain_from_json: null argument");
        std::vector<int32_t> lab;
        umcx::build_domain(umcx::Json::parse(json), dims[0], dims[1], dims[2], lab);
        *labels = (int32_t*)malloc(std::max<size_t>(lab.size(), 1) * 4);
        if (!*labels) throw std::runtime_error("out of host memory");
        memcpy(*labels, lab.data(), lab.size() * 4);
        return 0;
    });
}

int umcx_config_from_argv(char** argv, int argn, char** json_out) {
    return guarded([&]() -> int {
        umcx::ParsedInput in = umcx::parse_argv(argn, argv);
        std::string s = in.cfg.dump(2);
        if (!json_out) throw std::runtime_error("umcx_config_from_argv: null argument");
        *json_out = (char*)malloc(s.size() + 1);
        if (!*json_out) throw std::runtime_error("out of host memory");
        memcpy(*json_out, s.c_str(), s.size() + 1);
        return 0;
    });
}

int umcx_choose_tally_box(const float* counts, int32_t nx, int32_t ny, int32_t nz, int32_t ngates, uint64_t budget,
                          int32_t box[8]) {
    return guarded([&]() -> int {
        if (!counts || !box || nx <= 0 || ny <= 0 || nz <= 0 || ngates <= 0)
            throw std::runtime_error("umcx_choose_tally_box: empty grid or null pointer");
        const umcx::TallyBox b = umcx::choose_box(counts, nx, ny, nz, ngates, budget);
        for (int i = 0; i < 8; i++) box[i] = b[i];
        return 0;
    });
}

int umcx_choose_tally_box_device(const float* counts, int32_t nx, int32_t ny, int32_t nz, int32_t ngates,
                                 uint64_t budget, int32_t box[8], double sums[2], int32_t device) {
    return guarded([&]() -> int {
        if (!counts || !box || nx <= 0 || ny <= 0 || nz <= 0 || ngates <= 0)
            throw std::runtime_error("umcx_choose_tally_box_device: empty grid or null pointer");
        const umcx::TallyBoxSearch r = umcx::device_choose_box(counts, nx, ny, nz, ngates, budget, device);
        for (int i = 0; i < 8; i++) box[i] = r.box[i];
        if (sums) { sums[0] = r.in_box; sums[1] = r.total; }
        return 0;
    });
}

int umcx_seed_words(int32_t seed, uint32_t out[4]) {
    if (!out) {
        g_last_error = "umcx_seed_words: null argument";
        return 1;
    }
    umcx::glibc_seed_words(seed, out);
    return 0;
}

int umcx_libm_host(int32_t fn, const float* in, float* out, uint64_t n) {
    namespace L = umcx::libm;
    if (n && (!in || !out)) {
        g_last_error = "umcx_libm_host: null argument";
        return 1;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        float x = in[i];
        out[i] = fn == 0 ? L::logf_host(x) : fn == 1 ? L::sinf_any(x) : fn == 2 ? L::cosf_any(x)
                 : fn == 3 ? L::acosf_host(x) : L::expf_host(x);
    }
    return 0;
}

int umcx_libm_device(int32_t fn, const float* in, float* out, uint64_t n, int32_t device) {
    return guarded([&]() -> int { umcx::device_libm_eval(fn, in, out, n, device); return 0; });
}

int umcx_arith_check_device(int32_t which, uint64_t* tested, uint64_t* mismatches, int32_t device) {
    return guarded([&]() -> int {
        if (which < 0 || which > 4 || !mismatches) throw std::runtime_error("umcx_arith_check_device: which must be 0 to 4");
        *mismatches = umcx::device_arith_check(which, tested, device);
        return 0;
    });
}

void umcx_free(void* p) { free(p); }

}  // extern "C"
