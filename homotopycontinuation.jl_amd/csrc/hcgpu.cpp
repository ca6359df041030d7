// libhcgpu.so — C ABI of the MI355X batched path tracker (see include/hcgpu.h).
//
// hcg_program_create: generates HIP source from the homotopy's tapes (codegen.cpp),
//   compiles it with hipRTC for gfx950 (cached per process by source hash) and loads
//   the persistent tracker kernel.
// hcg_track_batch: copies a batch of start solutions to HBM, launches the persistent
//   kernel once (lanes pull paths from a device work queue) and copies the per-path
//   records back.  Counterpart of threaded_solve (src/solve.jl:628-709) for one batch.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdarg>
#include <fstream>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hcg_codegen.h"
#include "hcg_group_launch.h"
#include "hcg_mlp.h"
#include "hcg_noalloc.h"
#include "hcgpu.h"

#include "_gen/device_headers.inc"

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
#define HIPCHK(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return fail(HCG_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

struct hcg_ctx {
    int device = 0;
    int ncu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    unsigned long long last_counters[4] = {0, 0, 0, 0};
    float* d_mlp = nullptr;
    bool mlp_loaded = false;
    int32_t* d_next = nullptr;
    unsigned long long* d_counters = nullptr;
    hcg_options* d_opts = nullptr;
    hcg::MlpWork mlp_work;
};

struct hcg_program {
    hcg_ctx* ctx = nullptr;
    int n = 0, kind = 0, npar = 0;
    std::string source;
    hipModule_t module = nullptr;
    hipFunction_t kernel = nullptr;
    hcg::Flops flops;
    int waves_per_eu = 1;  // register budget the kernel was built for (waves per SIMD)
    // multi-system programs (hcg_program_create_multi): per-system size, kind, parameters
    int nsys = 0;
    std::vector<int> ns, kinds, npars;
    // certifier programs (hcg_certifier_create): the kernel is hcg_certify_kernel
    bool certifier = false;
    int cert_npar = 0;  // parameters given per hcg_certify call (0 when fixed or absent)
    // lane-group layout (hcg_build_options.layout = 1): slot programs for the precompiled kernel
    hcg::GroupPrograms* group = nullptr;
};

extern "C" const char* hcg_last_error(void) { return g_err.c_str(); }

extern "C" int hcg_ctx_create(int device, hcg_ctx** out) {
    if (!out) return fail(HCG_ERR_INVALID_ARGUMENT, "out is NULL");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(HCG_ERR_INVALID_ARGUMENT, "device %d out of range (%d devices)", device, ndev);
    HIPCHK(hipSetDevice(device));
    hcg_ctx* c = new hcg_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) { delete c; return fail(HCG_ERR_HIP, "hipGetDeviceProperties: %s", hipGetErrorString(e)); }
    c->ncu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return fail(HCG_ERR_HIP, "stream/event creation failed");
    }
    *out = c;
    return HCG_OK;
}

extern "C" void hcg_ctx_destroy(hcg_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->d_mlp) hipFree(c->d_mlp);
    if (c->d_next) hipFree(c->d_next);
    if (c->d_counters) hipFree(c->d_counters);
    if (c->d_opts) hipFree(c->d_opts);
    hcg::mlp_work_free(&c->mlp_work);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

extern "C" void hcg_default_options(hcg_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->max_steps = 10000;
    o->extended_precision = 1;
    o->max_step_size = INFINITY;
    o->max_initial_step_size = INFINITY;
    o->min_step_size = 1e-48;
    o->min_rel_step_size = 0.0;
    o->terminate_cond = 1e13;
    o->a = 0.125; o->beta_a = 1.0; o->beta_omega_p = 3.0; o->beta_tau = 0.4; o->strict_beta_tau = 0.3;
    o->min_newton_iters = 2;
    o->max_endgame_steps = 2000;
    o->endgame_start = 0.1;
    o->max_endgame_extended_steps = 400;
    o->max_winding_number = 6;
    o->min_cond = 1e6; o->min_cond_growth = 1e4; o->min_coord_growth = 100.0;
    o->zero_is_at_infinity = 0; o->at_infinity_check = 1; o->only_nonsingular = 0; o->refine_steps = 3;
    o->singular_min_accuracy = 1e-6;
    o->val_finite_tol = 0.05; o->val_at_infinity_tol = 0.01;
    o->sing_cond = 1e14; o->sing_accuracy = 1e-12; o->scaling_threshold = -30.0;
    o->scale_min = 1e-4; o->scale_abs_min = 1e-6; o->scale_max = std::ldexp(1.0, 511);
    o->learned_max_mult = 4.0; o->learned_min_t = 0.0;
    o->learned_inf_t = 0.0; o->learned_inf_logit = 4.0;
}

// ------------------------------ hipRTC ------------------------------
static std::mutex g_cache_mu;
static std::map<std::string, std::vector<char>> g_code_cache;  // source -> code object

// On-disk cache of compiled code objects: <dir>/<fnv1a64(source, options, hipRTC version)>.co,
// dir = $HCGPU_CACHE_DIR or <directory of libhcgpu.so>/rtc_cache.
static uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ULL) {
    for (unsigned char c : s) { h ^= c; h *= 1099511628211ULL; }
    return h;
}
static std::string cache_dir() {
    const char* env = std::getenv("HCGPU_CACHE_DIR");
    if (env && *env) return env;
    Dl_info info;
    if (dladdr((void*)&fnv1a, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t k = p.rfind('/');
        if (k != std::string::npos) return p.substr(0, k) + "/rtc_cache";
    }
    return "";
}
static bool cache_read(const std::string& path, std::vector<char>* code) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::vector<char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (buf.size() < 64) return false;
    *code = std::move(buf);
    return true;
}
static void cache_write(const std::string& dir, const std::string& path, const std::vector<char>& code) {
    if (dir.empty()) return;
    mkdir(dir.c_str(), 0755);
    std::string tmp = path + ".tmp" + std::to_string(getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        f.write(code.data(), (std::streamsize)code.size());
        if (!f) return;
    }
    std::rename(tmp.c_str(), path.c_str());
}

static std::vector<std::string> rtc_options() {
    std::vector<std::string> opts = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-std=c++17",
                                     "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                                     "-Wno-unused-variable", "-Wno-unused-function"};
    const char* extra = std::getenv("HCGPU_RTC_FLAGS");
    if (extra) {
        std::string s(extra), tok;
        for (char ch : s) { if (ch == ' ') { if (!tok.empty()) opts.push_back(tok); tok.clear(); } else tok += ch; }
        if (!tok.empty()) opts.push_back(tok);
    }
    return opts;
}

// Texts of the device headers the source includes, directly or through other headers.
static std::string included_headers(const std::string& src) {
    const size_t nh = sizeof(k_hdr_srcs) / sizeof(k_hdr_srcs[0]);
    std::vector<char> used(nh, 0);
    std::vector<std::string> todo = {src};
    while (!todo.empty()) {
        const std::string t = todo.back();
        todo.pop_back();
        for (size_t h = 0; h < nh; h++)
            if (!used[h] && t.find(std::string("#include \"") + k_hdr_names[h] + "\"") != std::string::npos) {
                used[h] = 1;
                todo.push_back(k_hdr_srcs[h]);
            }
    }
    std::string out;
    for (size_t h = 0; h < nh; h++)
        if (used[h]) out += k_hdr_srcs[h];
    return out;
}

static std::string cache_key(const std::string& src) {
    std::string k = src + included_headers(src);
    for (auto& o : rtc_options()) k += "\n" + o;
    int maj = 0, min = 0;
    hiprtcVersion(&maj, &min);
    k += "\nhiprtc " + std::to_string(maj) + "." + std::to_string(min);
    char buf[32];
    std::snprintf(buf, sizeof buf, "%016llx", (unsigned long long)fnv1a(k));
    return buf;
}

static int rtc_compile(const std::string& src, std::vector<char>* code) {
    const std::string key = cache_key(src);
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_code_cache.find(key);
        if (it != g_code_cache.end()) { *code = it->second; return HCG_OK; }
    }
    const std::string dir = cache_dir();
    const std::string path = dir.empty() ? "" : dir + "/" + key + ".co";
    if (!path.empty() && cache_read(path, code)) {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_code_cache[key] = *code;
        return HCG_OK;
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "hcg_program.hip", (int)(sizeof(k_hdr_names) / sizeof(k_hdr_names[0])),
                            k_hdr_srcs, k_hdr_names) != HIPRTC_SUCCESS)
        return fail(HCG_ERR_COMPILE, "hiprtcCreateProgram failed");
    std::vector<std::string> optstr = rtc_options();
    std::vector<const char*> opts;
    for (auto& o : optstr) opts.push_back(o.c_str());
    hiprtcResult r = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    size_t logsz = 0;
    hiprtcGetProgramLogSize(prog, &logsz);
    std::string log(logsz + 1, '\0');
    if (logsz) hiprtcGetProgramLog(prog, &log[0]);
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(HCG_ERR_COMPILE, "hipRTC compile failed: %.1800s", log.c_str());
    }
    size_t sz = 0;
    hiprtcGetCodeSize(prog, &sz);
    code->resize(sz);
    hiprtcGetCode(prog, code->data());
    hiprtcDestroyProgram(&prog);
    cache_write(dir, path, *code);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_code_cache[key] = *code;
    return HCG_OK;
}

// Generated source without compiling (no device needed).
static int hcg_generate_impl(const hcg_homotopy_desc* h, char* buf, size_t bufsz, size_t* needed, long* flops6) {
    std::string src, err;
    hcg::Flops fl;
    if (!hcg::generate_source(h, &src, &fl, &err)) return fail(HCG_ERR_INVALID_ARGUMENT, "%s", err.c_str());
    if (needed) *needed = src.size() + 1;
    if (buf && bufsz > 0) { std::strncpy(buf, src.c_str(), bufsz - 1); buf[bufsz - 1] = 0; }
    if (flops6) { flops6[0] = fl.eval; flops6[1] = fl.eval_jac; flops6[2] = fl.eval_dd; flops6[3] = fl.taylor[1]; flops6[4] = fl.taylor[2]; flops6[5] = fl.taylor[3]; }
    return HCG_OK;
}

// Default register budget of the tracker kernels: waves per SIMD they are built for.
static const int kDefaultWavesPerEU = 2;

static int build_source(const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t multi, const hcg_build_options* bo,
                        std::string* src, std::vector<hcg::Flops>* fl) {
    if (!hs || nsys < 1 || nsys > 8) return fail(HCG_ERR_INVALID_ARGUMENT, "between 1 and 8 homotopies per program");
    for (int s = 0; s < nsys; s++)
        if (!hs[s]) return fail(HCG_ERR_INVALID_ARGUMENT, "homotopy %d is NULL", s);
    if (nsys > 1 && !multi) return fail(HCG_ERR_INVALID_ARGUMENT, "several homotopies need a multi-system program");
    int w = bo && bo->waves_per_eu ? bo->waves_per_eu : kDefaultWavesPerEU;
    if (w != 1 && w != 2 && w != 4 && w != 8) return fail(HCG_ERR_INVALID_ARGUMENT, "waves_per_eu must be 1, 2, 4 or 8");
    std::string body, err;
    fl->assign(nsys, hcg::Flops());
    bool ok = multi ? hcg::generate_multi_source(hs, nsys, &body, fl->data(), &err)
                    : hcg::generate_source(hs[0], &body, &(*fl)[0], &err);
    if (!ok) return fail(HCG_ERR_INVALID_ARGUMENT, "%s", err.c_str());
    *src = "#define HCG_WAVES_PER_EU " + std::to_string(w) + "\n" + body;
    return HCG_OK;
}

static int hcg_compile_ex_impl(const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t multi,
                              const hcg_build_options* bo, size_t* code_size) {
    std::string src;
    std::vector<hcg::Flops> fl;
    int rc = build_source(hs, nsys, multi, bo, &src, &fl);
    if (rc) return rc;
    std::vector<char> code;
    rc = rtc_compile(src, &code);
    if (rc) return rc;
    if (code_size) *code_size = code.size();
    return HCG_OK;
}

static int group_program_create(hcg_ctx* ctx, const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t multi,
                                hcg_program** out) {
    if (!hs || nsys < 1 || nsys > HCG_MAX_SYSTEMS) return fail(HCG_ERR_INVALID_ARGUMENT, "between 1 and 8 homotopies per program");
    if (nsys > 1 && !multi) return fail(HCG_ERR_INVALID_ARGUMENT, "several homotopies need a multi-system program");
    hcg_program* p = new hcg_program();
    p->ctx = ctx;
    p->group = new hcg::GroupPrograms();
    std::string err;
    if (!hcg::group_build(hs, nsys, true, p->group, &err)) {
        delete p->group;
        delete p;
        return fail(HCG_ERR_INVALID_ARGUMENT, "%s", err.c_str());
    }
    p->waves_per_eu = 1;
    p->source = "// lane-group layout: slot programs for the precompiled hcg_group_kernel\n";
    if (multi) {
        p->nsys = nsys;
        for (int s = 0; s < nsys; s++) {
            p->ns.push_back(hs[s]->n);
            p->kinds.push_back(hs[s]->kind);
            p->npars.push_back(hs[s]->kind == HCG_PARAMETER ? hs[s]->target.eval->nparams : 0);
        }
        p->n = *std::max_element(p->ns.begin(), p->ns.end());
    } else {
        p->n = hs[0]->n;
        p->kind = hs[0]->kind;
        p->npar = hs[0]->kind == HCG_PARAMETER ? hs[0]->target.eval->nparams : 0;
    }
    *out = p;
    return HCG_OK;
}

static int hcg_program_create_ex_impl(hcg_ctx* ctx, const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t multi,
                                     const hcg_build_options* bo, hcg_program** out) {
    if (!ctx || !out) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    HIPCHK(hipSetDevice(ctx->device));
    if (bo && bo->layout == HCG_LAYOUT_GROUP) return group_program_create(ctx, hs, nsys, multi, out);
    if (bo && bo->layout != HCG_LAYOUT_LANE) return fail(HCG_ERR_INVALID_ARGUMENT, "unknown layout %d", bo->layout);
    hcg_program* p = new hcg_program();
    p->ctx = ctx;
    std::vector<hcg::Flops> fl;
    int rc = build_source(hs, nsys, multi, bo, &p->source, &fl);
    if (rc) { delete p; return rc; }
    p->waves_per_eu = bo && bo->waves_per_eu ? bo->waves_per_eu : kDefaultWavesPerEU;
    if (multi) {
        p->nsys = nsys;
        for (int s = 0; s < nsys; s++) {
            p->ns.push_back(hs[s]->n);
            p->kinds.push_back(hs[s]->kind);
            p->npars.push_back(hs[s]->kind == HCG_PARAMETER ? hs[s]->target.eval->nparams : 0);
        }
        p->n = *std::max_element(p->ns.begin(), p->ns.end());
    } else {
        p->n = hs[0]->n;
        p->kind = hs[0]->kind;
        p->npar = hs[0]->kind == HCG_PARAMETER ? hs[0]->target.eval->nparams : 0;
        p->flops = fl[0];
    }
    std::vector<char> code;
    rc = rtc_compile(p->source, &code);
    if (rc) { delete p; return rc; }
    hipError_t e = hipModuleLoadData(&p->module, code.data());
    if (e != hipSuccess) { delete p; return fail(HCG_ERR_HIP, "hipModuleLoadData: %s", hipGetErrorString(e)); }
    e = hipModuleGetFunction(&p->kernel, p->module, multi ? "hcg_track_multi_kernel" : "hcg_track_kernel");
    if (e != hipSuccess) { hipModuleUnload(p->module); delete p; return fail(HCG_ERR_HIP, "hipModuleGetFunction: %s", hipGetErrorString(e)); }
    *out = p;
    return HCG_OK;
}

// Compile-only entries (no device needed): the build check and CPU tests use them.
extern "C" int hcg_compile_only(const hcg_homotopy_desc* h, size_t* code_size) {
    return hcg_compile_ex(&h, 1, 0, nullptr, code_size);
}
extern "C" int hcg_compile_multi_only(const hcg_homotopy_desc* const* hs, int32_t nsys, size_t* code_size) {
    return hcg_compile_ex(hs, nsys, 1, nullptr, code_size);
}
extern "C" int hcg_program_create(hcg_ctx* ctx, const hcg_homotopy_desc* h, hcg_program** out) {
    if (!h) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    return hcg_program_create_ex(ctx, &h, 1, 0, nullptr, out);
}
extern "C" int hcg_program_create_multi(hcg_ctx* ctx, const hcg_homotopy_desc* const* hs, int32_t nsys,
                                        hcg_program** out) {
    return hcg_program_create_ex(ctx, hs, nsys, 1, nullptr, out);
}

extern "C" void hcg_program_destroy(hcg_program* p) {
    if (!p) return;
    if (p->group) {
        hipSetDevice(p->ctx->device);
        hcg::group_free(p->group);
        delete p->group;
        p->group = nullptr;
    }
    if (p->module) { hipSetDevice(p->ctx->device); hipModuleUnload(p->module); }
    delete p;
}

extern "C" const char* hcg_program_source(const hcg_program* p) { return p ? p->source.c_str() : ""; }

// ------------------------------ tracking ------------------------------
namespace {
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
};
}  // namespace

#include "device/hcg_kargs.h"
#include "device/hcg_cert_kargs.h"

static int validate_batch_args(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* b, const hcg_options* opts,
                               bool host) {
    if (!ctx || !prog || !b || !opts) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    if (prog->ctx != ctx) return fail(HCG_ERR_INVALID_ARGUMENT, "program belongs to another context");
    if (prog->certifier) return fail(HCG_ERR_INVALID_ARGUMENT, "certifier program: use hcg_certify");
    if (prog->nsys) return fail(HCG_ERR_INVALID_ARGUMENT, "multi-system program: use hcg_track_multi");
    const long np = b->npaths;
    if (np < 0) return fail(HCG_ERR_INVALID_ARGUMENT, "negative path count");
    if (np > (1L << 30) / prog->n) return fail(HCG_ERR_INVALID_ARGUMENT, "batch too large");
    if (np > 0 && !b->starts) return fail(HCG_ERR_INVALID_ARGUMENT, "starts is NULL");
    if (prog->kind == HCG_PARAMETER && np > 0) {
        if (!b->start_params || !b->target_params || b->ngroups < 1)
            return fail(HCG_ERR_INVALID_ARGUMENT, "parameter homotopy needs start/target parameters and ngroups >= 1");
        if (host && b->group_of_path)
            for (long k = 0; k < np; k++)
                if (b->group_of_path[k] < 0 || b->group_of_path[k] >= b->ngroups)
                    return fail(HCG_ERR_INVALID_ARGUMENT, "group_of_path[%ld] = %d out of range", k, b->group_of_path[k]);
    }
    if (opts->max_winding_number < 1 || opts->min_newton_iters < 0 || opts->min_newton_iters > 20 || opts->max_steps < 0)
        return fail(HCG_ERR_INVALID_ARGUMENT, "invalid options");
    if (opts->learned_step && !ctx->mlp_loaded) return fail(HCG_ERR_INVALID_ARGUMENT, "learned_step requested but no MLP loaded");
    if (opts->learned_step && !(opts->learned_max_mult >= 1.0))
        return fail(HCG_ERR_INVALID_ARGUMENT, "learned_max_mult must be >= 1");
    return HCG_OK;
}

// Persistent per-context device scalars: work-queue head, counters, options.
static int ctx_scalars(hcg_ctx* ctx) {
    if (ctx->d_next) return HCG_OK;
    HIPCHK(hipMalloc(&ctx->d_next, sizeof(int32_t)));
    HIPCHK(hipMalloc(&ctx->d_counters, sizeof(unsigned long long) * 4));
    HIPCHK(hipMalloc(&ctx->d_opts, sizeof(hcg_options)));
    return HCG_OK;
}

// Launch the lane-group kernel: 16 paths per workgroup of 256 lanes, at most
// blocks_per_cu workgroups per CU (default 2).
static int group_launch(hcg_ctx* ctx, const hcg_program* prog, const hcg_multi_kargs& m, long total,
                        const hcg_options* opts, hipStream_t stream) {
    const int per_block = hcg::group_block() / hcg::group_lanes_for(prog->group->nn);
    const int per_cu = opts->blocks_per_cu > 0 ? opts->blocks_per_cu : 2;
    long grid = std::min<long>((total + per_block - 1) / per_block, (long)ctx->ncu * per_cu);
    if (grid < 1) grid = 1;
    const bool mlp_lds = opts->learned_step && m.sys[0].mlp;  // the kernel's own staging condition
    hipError_t e = hcg::group_launch(prog->group->nn, (int)grid, prog->group->lds_per_path, mlp_lds, stream, m,
                                     prog->group->d_sys);
    if (e != hipSuccess) return fail(HCG_ERR_HIP, "group kernel launch: %s", hipGetErrorString(e));
    return HCG_OK;
}

// Where a launch keeps its queue head, counters and options: the context's own buffers
// (allocated once per context) or the fixed scratch of the allocation-free entry point.
struct LaunchState {
    int32_t* next;
    unsigned long long* counters;
    hcg_options* opts;
    hcg::NoallocScratch* fixed;  // non-null: state lives in the fixed scratch
};

static int ctx_launch_state(hcg_ctx* ctx, LaunchState* st) {
    int rc = ctx_scalars(ctx);
    if (rc) return rc;
    st->next = ctx->d_next;
    st->counters = ctx->d_counters;
    st->opts = ctx->d_opts;
    st->fixed = nullptr;
    return HCG_OK;
}

// Enqueue one launch on `stream`: the reset of the launch state (queue head, counters, options)
// and the persistent tracker kernel over device-resident inputs/outputs.  `timed` records the
// context's start event right before the kernel (a captured graph is timed around its replay).
static int enqueue_track(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* b, const hcg_options* opts,
                         hcg_path_result* d_res, double* d_sol, double* d_last, double* d_val, float* d_log, int logcap,
                         int32_t* d_logc, const LaunchState& st, hipStream_t stream, bool timed) {
    if (st.fixed) {
        if (hcg::noalloc_init(stream, st.fixed, opts)) return fail(HCG_ERR_HIP, "launch-state init kernel failed");
    } else {
        HIPCHK(hipMemcpyAsync(st.opts, opts, sizeof(hcg_options), hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemsetAsync(st.next, 0, sizeof(int32_t), stream));
        HIPCHK(hipMemsetAsync(st.counters, 0, sizeof(unsigned long long) * 4, stream));
    }
    hcg_kargs ka;
    std::memset(&ka, 0, sizeof ka);
    ka.npaths = b->npaths;
    ka.log_capacity = logcap;
    ka.ngroups = b->ngroups;
    ka.starts = b->starts;
    ka.t1 = b->t1;
    ka.t0 = b->t0;
    ka.start_params = b->start_params;
    ka.target_params = b->target_params;
    ka.group_of_path = b->group_of_path;
    ka.next_path = st.next;
    ka.opts = st.opts;
    ka.results = d_res;
    ka.solutions = d_sol;
    ka.last_points = d_last;
    ka.valuations = d_val;
    ka.counters = st.counters;
    ka.log_records = d_log;
    ka.log_count = d_logc;
    ka.mlp = opts->learned_step ? ctx->d_mlp : nullptr;
    if (timed) HIPCHK(hipEventRecord(ctx->ev0, stream));
    if (prog->group) {
        hcg_multi_kargs m;
        std::memset(&m, 0, sizeof m);
        m.nsys = 1;
        m.total = b->npaths;
        m.offset[0] = 0;
        m.offset[1] = b->npaths;
        m.next_path = st.next;
        m.sys[0] = ka;
        return group_launch(ctx, prog, m, b->npaths, opts, stream);
    }
    const int block = 64;
    int per_cu = opts->blocks_per_cu > 0 ? opts->blocks_per_cu : 4 * prog->waves_per_eu;  // every wave slot of the CU's 4 SIMDs
    long waves_needed = ((long)b->npaths + block - 1) / block;
    long grid = std::min<long>(waves_needed, (long)ctx->ncu * per_cu);
    if (grid < 1) grid = 1;
    void* args[] = {&ka};
    HIPCHK(hipModuleLaunchKernel(prog->kernel, (unsigned)grid, 1, 1, block, 1, 1, 0, stream, args, nullptr));
    return HCG_OK;
}

// Launch the persistent tracker kernel over device-resident inputs/outputs on the context's
// stream, timed by the context's events; the launch counters are copied back after it.
static int launch_track(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* b, const hcg_options* opts,
                        hcg_path_result* d_res, double* d_sol, double* d_last, double* d_val, float* d_log, int logcap,
                        int32_t* d_logc, const LaunchState& st) {
    int rc = enqueue_track(ctx, prog, b, opts, d_res, d_sol, d_last, d_val, d_log, logcap, d_logc, st, ctx->stream,
                           true);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->last_counters, st.counters, sizeof(unsigned long long) * 4, hipMemcpyDeviceToHost,
                          ctx->stream));
    return HCG_OK;
}

static int finish_launch(hcg_ctx* ctx) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    return HCG_OK;
}

static int hcg_track_batch_impl(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* b, const hcg_options* opts,
                               hcg_path_result* results, double* solutions, double* last_points, double* valuations,
                               hcg_step_log* log) {
    int rc = validate_batch_args(ctx, prog, b, opts, true);
    if (rc) return rc;
    if (!results || !solutions || !last_points || !valuations) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL output");
    const int n = prog->n;
    const long np = b->npaths;
    if (np == 0) return HCG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t sz_starts = sizeof(double) * 2 * n * np;
    DevBuf d_starts, d_sp, d_tp, d_grp, d_res, d_sol, d_last, d_val, d_log, d_logc;
    HIPCHK(hipMalloc(&d_starts.p, sz_starts));
    HIPCHK(hipMalloc(&d_res.p, sizeof(hcg_path_result) * np));
    HIPCHK(hipMalloc(&d_sol.p, sz_starts));
    HIPCHK(hipMalloc(&d_last.p, sz_starts));
    HIPCHK(hipMalloc(&d_val.p, sizeof(double) * n * np));
    HIPCHK(hipMemcpyAsync(d_starts.p, b->starts, sz_starts, hipMemcpyHostToDevice, ctx->stream));
    hcg_batch db = *b;
    db.starts = (const double*)d_starts.p;
    if (prog->kind == HCG_PARAMETER) {
        HIPCHK(hipMalloc(&d_sp.p, sizeof(double) * 2 * prog->npar));
        HIPCHK(hipMalloc(&d_tp.p, sizeof(double) * 2 * prog->npar * b->ngroups));
        HIPCHK(hipMemcpyAsync(d_sp.p, b->start_params, sizeof(double) * 2 * prog->npar, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_tp.p, b->target_params, sizeof(double) * 2 * prog->npar * b->ngroups, hipMemcpyHostToDevice, ctx->stream));
        db.start_params = (const double*)d_sp.p;
        db.target_params = (const double*)d_tp.p;
        if (b->group_of_path) {
            HIPCHK(hipMalloc(&d_grp.p, sizeof(int32_t) * np));
            HIPCHK(hipMemcpyAsync(d_grp.p, b->group_of_path, sizeof(int32_t) * np, hipMemcpyHostToDevice, ctx->stream));
            db.group_of_path = (const int32_t*)d_grp.p;
        }
    }
    int logcap = 0;
    if (log && opts->log_steps) {
        if (log->capacity < 0 || !log->records || !log->count) return fail(HCG_ERR_INVALID_ARGUMENT, "bad step log");
        logcap = log->capacity;
        HIPCHK(hipMalloc(&d_log.p, sizeof(float) * HCG_STEP_RECORD_FLOATS * (size_t)(logcap > 0 ? logcap : 1)));
        HIPCHK(hipMalloc(&d_logc.p, sizeof(int32_t)));
        HIPCHK(hipMemsetAsync(d_logc.p, 0, sizeof(int32_t), ctx->stream));
    }
    LaunchState st;
    rc = ctx_launch_state(ctx, &st);
    if (rc) return rc;
    rc = launch_track(ctx, prog, &db, opts, (hcg_path_result*)d_res.p, (double*)d_sol.p, (double*)d_last.p,
                      (double*)d_val.p, (float*)d_log.p, logcap, (int32_t*)d_logc.p, st);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(results, d_res.p, sizeof(hcg_path_result) * np, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(solutions, d_sol.p, sz_starts, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(last_points, d_last.p, sz_starts, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(valuations, d_val.p, sizeof(double) * n * np, hipMemcpyDeviceToHost, ctx->stream));
    rc = finish_launch(ctx);
    if (rc) return rc;
    if (d_logc.p) {
        int32_t cnt = 0;
        HIPCHK(hipMemcpy(&cnt, d_logc.p, sizeof(int32_t), hipMemcpyDeviceToHost));
        *log->count = cnt;
        int ncopy = std::min(cnt, logcap);
        if (ncopy > 0)
            HIPCHK(hipMemcpy(log->records, d_log.p, sizeof(float) * HCG_STEP_RECORD_FLOATS * ncopy, hipMemcpyDeviceToHost));
    }
    return HCG_OK;
}

static int hcg_track_batch_device_impl(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch_dev,
                                      const hcg_options* opts, hcg_path_result* results_dev, double* solutions_dev,
                                      double* last_points_dev, double* valuations_dev) {
    int rc = validate_batch_args(ctx, prog, batch_dev, opts, false);
    if (rc) return rc;
    if (!results_dev || !solutions_dev || !last_points_dev || !valuations_dev)
        return fail(HCG_ERR_INVALID_ARGUMENT, "NULL output");
    if (batch_dev->npaths == 0) return HCG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    LaunchState st;
    rc = ctx_launch_state(ctx, &st);
    if (rc) return rc;
    rc = launch_track(ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev, valuations_dev, nullptr,
                      0, nullptr, st);
    if (rc) return rc;
    return finish_launch(ctx);
}

// hcg_track_batch_device without any allocation: the launch state lives in the library's
// fixed device scratch (noalloc.hip), cleared by a kernel on the context's stream.
static int hcg_track_batch_device_noalloc_impl(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch_dev,
                                               const hcg_options* opts, hcg_path_result* results_dev,
                                               double* solutions_dev, double* last_points_dev, double* valuations_dev) {
    int rc = validate_batch_args(ctx, prog, batch_dev, opts, false);
    if (rc) return rc;
    if (!results_dev || !solutions_dev || !last_points_dev || !valuations_dev)
        return fail(HCG_ERR_INVALID_ARGUMENT, "NULL output");
    if (batch_dev->npaths == 0) return HCG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    hcg::NoallocScratch* fixed = nullptr;
    const char* err = "";
    if (hcg::noalloc_scratch(ctx->device, &fixed, &err)) return fail(HCG_ERR_HIP, "hipGetSymbolAddress(scratch): %s", err);
    LaunchState st;
    st.next = &fixed->next_path;
    st.counters = fixed->counters;
    st.opts = &fixed->opts;
    st.fixed = fixed;
    rc = launch_track(ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev, valuations_dev, nullptr,
                      0, nullptr, st);
    if (rc) return rc;
    return finish_launch(ctx);
}

// A captured device-resident launch: the launch-state reset kernel and the tracker kernel as
// one HIP graph, replayed by hcg_graph_launch.  The graph owns its launch state (queue head,
// counters, options), cleared by the reset kernel at every replay (no memset node).
struct hcg_graph {
    hcg_ctx* ctx = nullptr;
    const hcg_program* prog = nullptr;
    hcg::NoallocScratch* state = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static void graph_free(hcg_graph* g) {
    if (!g) return;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->state) hipFree(g->state);
    delete g;
}

static int hcg_graph_create_impl(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch_dev,
                                 const hcg_options* opts, hcg_path_result* results_dev, double* solutions_dev,
                                 double* last_points_dev, double* valuations_dev, hcg_graph** out) {
    if (!out) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL out");
    *out = nullptr;
    int rc = validate_batch_args(ctx, prog, batch_dev, opts, false);
    if (rc) return rc;
    if (!results_dev || !solutions_dev || !last_points_dev || !valuations_dev)
        return fail(HCG_ERR_INVALID_ARGUMENT, "NULL output");
    if (batch_dev->npaths < 1) return fail(HCG_ERR_INVALID_ARGUMENT, "a graph needs at least one path");
    HIPCHK(hipSetDevice(ctx->device));
    std::unique_ptr<hcg_graph, void (*)(hcg_graph*)> g(new hcg_graph, graph_free);
    g->ctx = ctx;
    g->prog = prog;
    HIPCHK(hipMalloc((void**)&g->state, sizeof(hcg::NoallocScratch)));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    LaunchState st;
    st.next = &g->state->next_path;
    st.counters = g->state->counters;
    st.opts = &g->state->opts;
    st.fixed = g->state;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_track(ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev, valuations_dev,
                       nullptr, 0, nullptr, st, ctx->stream, false);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    g->graph = graph;
    if (rc) return rc;
    if (e != hipSuccess) return fail(HCG_ERR_HIP, "stream capture of the launch: %s", hipGetErrorString(e));
    HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
    return HCG_OK;
}

static int hcg_graph_launch_impl(hcg_graph* g) {
    if (!g || !g->exec) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL graph");
    hcg_ctx* ctx = g->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(hipGraphLaunch(g->exec, ctx->stream));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->last_counters, g->state->counters, sizeof(unsigned long long) * 4,
                          hipMemcpyDeviceToHost, ctx->stream));
    return finish_launch(ctx);
}

static int hcg_track_multi_impl(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batches,
                               const hcg_options* opts, hcg_path_result* results, double* solutions,
                               double* last_points, double* valuations) {
    if (!ctx || !prog || !batches || !opts || !results || !solutions || !last_points || !valuations)
        return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    if (prog->ctx != ctx) return fail(HCG_ERR_INVALID_ARGUMENT, "program belongs to another context");
    if (!prog->nsys) return fail(HCG_ERR_INVALID_ARGUMENT, "single-system program: use hcg_track_batch");
    if (opts->max_winding_number < 1 || opts->min_newton_iters < 0 || opts->min_newton_iters > 20 || opts->max_steps < 0)
        return fail(HCG_ERR_INVALID_ARGUMENT, "invalid options");
    if (opts->learned_step && !ctx->mlp_loaded) return fail(HCG_ERR_INVALID_ARGUMENT, "learned_step requested but no MLP loaded");
    if (opts->learned_step && !(opts->learned_max_mult >= 1.0)) return fail(HCG_ERR_INVALID_ARGUMENT, "learned_max_mult must be >= 1");
    const int S = prog->nsys;
    long total = 0, tot_nx = 0;
    for (int s = 0; s < S; s++) {
        const hcg_batch* b = &batches[s];
        if (b->npaths < 0) return fail(HCG_ERR_INVALID_ARGUMENT, "system %d: negative path count", s);
        if (b->npaths > 0 && !b->starts) return fail(HCG_ERR_INVALID_ARGUMENT, "system %d: starts is NULL", s);
        if (prog->kinds[s] == HCG_PARAMETER && b->npaths > 0) {
            if (!b->start_params || !b->target_params || b->ngroups < 1)
                return fail(HCG_ERR_INVALID_ARGUMENT, "system %d: parameter homotopy needs parameters", s);
            if (b->group_of_path)
                for (long k = 0; k < b->npaths; k++)
                    if (b->group_of_path[k] < 0 || b->group_of_path[k] >= b->ngroups)
                        return fail(HCG_ERR_INVALID_ARGUMENT, "system %d: group_of_path[%ld] out of range", s, k);
        }
        total += b->npaths;
        tot_nx += (long)b->npaths * prog->ns[s];
    }
    if (total > (1L << 30) || tot_nx > (1L << 31)) return fail(HCG_ERR_INVALID_ARGUMENT, "batch too large");
    if (total == 0) return HCG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ctx_scalars(ctx);
    if (rc) return rc;
    // one allocation per array kind, sliced per system
    DevBuf d_starts, d_res, d_sol, d_last, d_val, d_par;
    HIPCHK(hipMalloc(&d_starts.p, sizeof(double) * 2 * tot_nx));
    HIPCHK(hipMalloc(&d_res.p, sizeof(hcg_path_result) * total));
    HIPCHK(hipMalloc(&d_sol.p, sizeof(double) * 2 * tot_nx));
    HIPCHK(hipMalloc(&d_last.p, sizeof(double) * 2 * tot_nx));
    HIPCHK(hipMalloc(&d_val.p, sizeof(double) * tot_nx));
    long par_doubles = 0, grp_ints = 0;
    for (int s = 0; s < S; s++)
        if (prog->kinds[s] == HCG_PARAMETER && batches[s].npaths > 0) {
            par_doubles += 2L * prog->npars[s] * (1 + batches[s].ngroups);
            if (batches[s].group_of_path) grp_ints += batches[s].npaths;
        }
    if (par_doubles + grp_ints > 0) HIPCHK(hipMalloc(&d_par.p, sizeof(double) * (par_doubles + grp_ints)));
    HIPCHK(hipMemcpyAsync(ctx->d_opts, opts, sizeof(hcg_options), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_next, 0, sizeof(int32_t), ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->d_counters, 0, sizeof(unsigned long long) * 4, ctx->stream));
    hcg_multi_kargs m;
    std::memset(&m, 0, sizeof m);
    m.nsys = S;
    m.total = (int32_t)total;
    m.next_path = ctx->d_next;
    long off = 0, offx = 0;
    double* parp = (double*)d_par.p;
    for (int s = 0; s < S; s++) {
        const hcg_batch* b = &batches[s];
        const int n = prog->ns[s];
        const long np = b->npaths;
        m.offset[s] = (int32_t)off;
        hcg_kargs& ka = m.sys[s];
        double* st = (double*)d_starts.p + 2 * offx;
        if (np > 0) HIPCHK(hipMemcpyAsync(st, b->starts, sizeof(double) * 2 * n * np, hipMemcpyHostToDevice, ctx->stream));
        ka.npaths = (int32_t)np;
        ka.ngroups = b->ngroups;
        ka.starts = st;
        ka.t1 = b->t1;
        ka.t0 = b->t0;
        if (prog->kinds[s] == HCG_PARAMETER && np > 0) {
            const int npar = prog->npars[s];
            HIPCHK(hipMemcpyAsync(parp, b->start_params, sizeof(double) * 2 * npar, hipMemcpyHostToDevice, ctx->stream));
            ka.start_params = parp;
            parp += 2 * npar;
            HIPCHK(hipMemcpyAsync(parp, b->target_params, sizeof(double) * 2 * npar * b->ngroups, hipMemcpyHostToDevice, ctx->stream));
            ka.target_params = parp;
            parp += 2L * npar * b->ngroups;
        }
        ka.next_path = ctx->d_next;
        ka.opts = ctx->d_opts;
        ka.results = (hcg_path_result*)d_res.p + off;
        ka.solutions = (double*)d_sol.p + 2 * offx;
        ka.last_points = (double*)d_last.p + 2 * offx;
        ka.valuations = (double*)d_val.p + offx;
        ka.counters = ctx->d_counters;
        ka.mlp = opts->learned_step ? ctx->d_mlp : nullptr;
        off += np;
        offx += np * n;
    }
    // group indices after all parameter blocks (int32, 8-byte aligned start)
    int32_t* grp = (int32_t*)parp;
    for (int s = 0; s < S; s++)
        if (prog->kinds[s] == HCG_PARAMETER && batches[s].npaths > 0 && batches[s].group_of_path) {
            HIPCHK(hipMemcpyAsync(grp, batches[s].group_of_path, sizeof(int32_t) * batches[s].npaths, hipMemcpyHostToDevice,
                                  ctx->stream));
            m.sys[s].group_of_path = grp;
            grp += batches[s].npaths;
        }
    m.offset[S] = (int32_t)total;
    const int block = 64;
    int per_cu = opts->blocks_per_cu > 0 ? opts->blocks_per_cu : 4 * prog->waves_per_eu;  // every wave slot of the CU's 4 SIMDs
    long grid = std::min<long>((total + block - 1) / block, (long)ctx->ncu * per_cu);
    if (grid < 1) grid = 1;
    void* args[] = {&m};
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    if (prog->group) {
        for (int s = 0; s < S; s++) m.sys[s].counters = ctx->d_counters, m.sys[s].opts = ctx->d_opts;
        rc = group_launch(ctx, prog, m, total, opts, ctx->stream);
        if (rc) return rc;
    } else {
        HIPCHK(hipModuleLaunchKernel(prog->kernel, (unsigned)grid, 1, 1, block, 1, 1, 0, ctx->stream, args, nullptr));
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->last_counters, ctx->d_counters, sizeof(unsigned long long) * 4, hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipMemcpyAsync(results, d_res.p, sizeof(hcg_path_result) * total, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(solutions, d_sol.p, sizeof(double) * 2 * tot_nx, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(last_points, d_last.p, sizeof(double) * 2 * tot_nx, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(valuations, d_val.p, sizeof(double) * tot_nx, hipMemcpyDeviceToHost, ctx->stream));
    return finish_launch(ctx);
}

extern "C" int hcg_last_kernel_stats(hcg_ctx* ctx, double* kernel_ms, int64_t* total_steps, int64_t* total_newton_iters) {
    if (!ctx) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL ctx");
    if (kernel_ms) *kernel_ms = ctx->last_ms;
    if (total_steps) *total_steps = (int64_t)ctx->last_counters[0];
    if (total_newton_iters) *total_newton_iters = (int64_t)ctx->last_counters[1];
    return HCG_OK;
}

// Extra counters of the last launch: [0] steps, [1] Newton iterations, [2] FP64 flops (counted), [3] learned proposals accepted.
extern "C" int hcg_last_counters(hcg_ctx* ctx, uint64_t* out4) {
    if (!ctx || !out4) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int i = 0; i < 4; i++) out4[i] = ctx->last_counters[i];
    return HCG_OK;
}

// Slot schedules of the lane-group layout without a device (checked against the tapes):
// out = [kernel size class, LDS bytes per path, then per system: eval_jac slots, registers,
// instructions, eval slots, registers, instructions].
extern "C" int hcg_group_schedule_info(const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t* out) {
    if (!hs || !out) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    try {
        hcg::GroupPrograms g;
        std::string err;
        if (!hcg::group_build(hs, nsys, false, &g, &err)) return fail(HCG_ERR_INVALID_ARGUMENT, "%s", err.c_str());
        out[0] = g.nn;
        out[1] = g.lds_per_path;
        for (int s = 0; s < nsys; s++) {
            int32_t* o = out + 2 + 6 * s;
            o[0] = g.ej_slots[s]; o[1] = g.ej_regs[s]; o[2] = g.ej_instr[s];
            o[3] = g.e_slots[s]; o[4] = g.e_regs[s]; o[5] = g.e_instr[s];
        }
        return HCG_OK;
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_group_schedule_info: %s", ex.what());
    }
}

// Resource use of a program's kernel, from its code object: [0] VGPRs per lane (as the
// loader reports them), [1] private segment bytes per lane, [2] static LDS bytes per
// workgroup, [3] waves per SIMD the kernel was built for (lane-group layout: the
// workgroups that fit on a CU, each one wave per SIMD), [4] workgroup size, [5] CUs.
extern "C" int hcg_program_kernel_info(const hcg_program* p, int32_t* out6) {
    if (!p || !out6) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    if (p->group) {
        HIPCHK(hipSetDevice(p->ctx->device));
        hipFuncAttributes fa;
        HIPCHK(hcg::group_kernel_attributes(p->group->nn, &fa));
        out6[0] = fa.numRegs;
        out6[1] = (int32_t)fa.localSizeBytes;
        // LDS of a plain launch; the learned step stages its network's weights in front of it
        // (group_mlp_lds_bytes() more per workgroup)
        out6[2] = (hcg::group_block() / hcg::group_lanes_for(p->group->nn)) * p->group->lds_per_path;
        int blocks = 0;  // workgroups per CU; each holds one wave per SIMD
        HIPCHK(hcg::group_kernel_blocks_per_cu(p->group->nn, (size_t)out6[2], &blocks));
        out6[3] = blocks;
        out6[4] = hcg::group_block();
        out6[5] = p->ctx->ncu;
        return HCG_OK;
    }
    if (!p->kernel) return fail(HCG_ERR_INVALID_ARGUMENT, "program has no kernel");
    HIPCHK(hipSetDevice(p->ctx->device));
    int regs = 0, priv = 0, lds = 0, maxthr = 0;
    HIPCHK(hipFuncGetAttribute(&regs, HIP_FUNC_ATTRIBUTE_NUM_REGS, p->kernel));
    HIPCHK(hipFuncGetAttribute(&priv, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, p->kernel));
    HIPCHK(hipFuncGetAttribute(&lds, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, p->kernel));
    HIPCHK(hipFuncGetAttribute(&maxthr, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, p->kernel));
    out6[0] = regs; out6[1] = priv; out6[2] = lds; out6[3] = p->waves_per_eu; out6[4] = 64; out6[5] = p->ctx->ncu;
    return HCG_OK;
}

// ------------------------------ learned proposer ------------------------------
extern "C" int hcg_mlp_load(hcg_ctx* ctx, const float* params) {
    if (!ctx || !params) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->d_mlp) HIPCHK(hipMalloc(&ctx->d_mlp, sizeof(float) * HCG_MLP_NPARAMS));
    HIPCHK(hipMemcpy(ctx->d_mlp, params, sizeof(float) * HCG_MLP_NPARAMS, hipMemcpyHostToDevice));
    ctx->mlp_loaded = true;
    return HCG_OK;
}

extern "C" int hcg_mlp_forward_backward(hcg_ctx* ctx, const float* params, int32_t n, const float* x, const float* y,
                                        const float* w, float* out_pred, float* loss, float* grad) {
    if (!ctx || !params || !x || !y || !w || n < 0) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    HIPCHK(hipSetDevice(ctx->device));
    std::string err;
    int rc = hcg::mlp_forward_backward(&ctx->mlp_work, ctx->stream, params, n, x, y, w, out_pred, loss, grad, &err);
    if (rc) return fail(rc, "%s", err.c_str());
    return HCG_OK;
}

extern "C" int hcg_mlp_train(hcg_ctx* ctx, float* params, int32_t n, const float* x, const float* y, const float* w,
                             int32_t epochs, int32_t batch, float lr, uint64_t seed, float* final_loss) {
    if (!ctx || !params || !x || !y || !w || n < 1 || epochs < 0 || batch < 1) return fail(HCG_ERR_INVALID_ARGUMENT, "bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    std::string err;
    int rc = hcg::mlp_train(&ctx->mlp_work, ctx->stream, params, n, x, y, w, epochs, batch, lr, seed, final_loss, &err);
    if (rc) return fail(rc, "%s", err.c_str());
    return HCG_OK;
}

// ---- C ABI: no C++ exception crosses the boundary ----
extern "C" int hcg_generate(const hcg_homotopy_desc* h, char* buf, size_t bufsz, size_t* needed, long* flops6) {
    try {
        return hcg_generate_impl(h, buf, bufsz, needed, flops6);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_generate: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_generate: unexpected error");
    }
}
extern "C" int hcg_compile_ex(const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t multi, const hcg_build_options* bo, size_t* code_size) {
    try {
        return hcg_compile_ex_impl(hs, nsys, multi, bo, code_size);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_compile_ex: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_compile_ex: unexpected error");
    }
}
extern "C" int hcg_program_create_ex(hcg_ctx* ctx, const hcg_homotopy_desc* const* hs, int32_t nsys, int32_t multi, const hcg_build_options* bo, hcg_program** out) {
    try {
        return hcg_program_create_ex_impl(ctx, hs, nsys, multi, bo, out);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_program_create_ex: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_program_create_ex: unexpected error");
    }
}
extern "C" int hcg_track_batch(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* b, const hcg_options* opts, hcg_path_result* results, double* solutions, double* last_points, double* valuations, hcg_step_log* log) {
    try {
        return hcg_track_batch_impl(ctx, prog, b, opts, results, solutions, last_points, valuations, log);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_track_batch: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_track_batch: unexpected error");
    }
}
extern "C" int hcg_track_batch_device(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch_dev, const hcg_options* opts, hcg_path_result* results_dev, double* solutions_dev, double* last_points_dev, double* valuations_dev) {
    try {
        return hcg_track_batch_device_impl(ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev, valuations_dev);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_track_batch_device: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_track_batch_device: unexpected error");
    }
}
extern "C" int hcg_track_batch_device_noalloc(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch_dev,
                                              const hcg_options* opts, hcg_path_result* results_dev,
                                              double* solutions_dev, double* last_points_dev, double* valuations_dev) {
    try {
        return hcg_track_batch_device_noalloc_impl(ctx, prog, batch_dev, opts, results_dev, solutions_dev,
                                                   last_points_dev, valuations_dev);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_track_batch_device_noalloc: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_track_batch_device_noalloc: unexpected error");
    }
}
extern "C" int hcg_graph_create(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batch_dev,
                                const hcg_options* opts, hcg_path_result* results_dev, double* solutions_dev,
                                double* last_points_dev, double* valuations_dev, hcg_graph** out) {
    try {
        return hcg_graph_create_impl(ctx, prog, batch_dev, opts, results_dev, solutions_dev, last_points_dev,
                                     valuations_dev, out);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_graph_create: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_graph_create: unexpected error");
    }
}
extern "C" int hcg_graph_launch(hcg_graph* g) {
    try {
        return hcg_graph_launch_impl(g);
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_graph_launch: unexpected error");
    }
}
extern "C" void hcg_graph_destroy(hcg_graph* g) { graph_free(g); }
extern "C" int hcg_track_multi(hcg_ctx* ctx, const hcg_program* prog, const hcg_batch* batches, const hcg_options* opts, hcg_path_result* results, double* solutions, double* last_points, double* valuations) {
    try {
        return hcg_track_multi_impl(ctx, prog, batches, opts, results, solutions, last_points, valuations);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_track_multi: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_track_multi: unexpected error");
    }
}

// ------------------------------ certification ------------------------------
static int certifier_source(const hcg_system* F, int32_t n, std::string* src) {
    if (!F) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL system");
    std::string err;
    if (!hcg::generate_certify_source(F, n, src, &err)) return fail(HCG_ERR_INVALID_ARGUMENT, "%s", err.c_str());
    return HCG_OK;
}

static int hcg_certifier_create_impl(hcg_ctx* ctx, const hcg_system* F, int32_t n, hcg_program** out) {
    if (!ctx || !out) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    HIPCHK(hipSetDevice(ctx->device));
    hcg_program* p = new hcg_program();
    p->ctx = ctx;
    int rc = certifier_source(F, n, &p->source);
    if (rc) { delete p; return rc; }
    p->n = n;
    p->certifier = true;
    p->cert_npar = F->fixed_params ? 0 : F->eval->nparams;
    std::vector<char> code;
    rc = rtc_compile(p->source, &code);
    if (rc) { delete p; return rc; }
    hipError_t e = hipModuleLoadData(&p->module, code.data());
    if (e != hipSuccess) { delete p; return fail(HCG_ERR_HIP, "hipModuleLoadData: %s", hipGetErrorString(e)); }
    e = hipModuleGetFunction(&p->kernel, p->module, "hcg_certify_kernel");
    if (e != hipSuccess) { hipModuleUnload(p->module); delete p; return fail(HCG_ERR_HIP, "hipModuleGetFunction: %s", hipGetErrorString(e)); }
    *out = p;
    return HCG_OK;
}

static int hcg_certify_impl(hcg_ctx* ctx, const hcg_program* prog, int32_t nsol, const double* cand,
                            int32_t ngroups, const double* params, const double* params_ival,
                            const double* params_ball, const int32_t* group_of, int32_t refine, int32_t extended,
                            int32_t real_system, hcg_certificate* out, double* box1, double* box0, double* refined) {
    if (!ctx || !prog) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL argument");
    if (prog->ctx != ctx) return fail(HCG_ERR_INVALID_ARGUMENT, "program belongs to another context");
    if (!prog->certifier) return fail(HCG_ERR_INVALID_ARGUMENT, "not a certifier program (use hcg_certifier_create)");
    if (nsol < 0) return fail(HCG_ERR_INVALID_ARGUMENT, "negative candidate count");
    const int n = prog->n, np = prog->cert_npar;
    if ((long)nsol > (1L << 28) / n) return fail(HCG_ERR_INVALID_ARGUMENT, "too many candidates");
    if (nsol == 0) return HCG_OK;
    if (!cand || !out || !box1 || !box0 || !refined) return fail(HCG_ERR_INVALID_ARGUMENT, "NULL input or output");
    if (np > 0) {
        if (!params) return fail(HCG_ERR_INVALID_ARGUMENT, "the system has %d parameters: params is NULL", np);
        if (ngroups < 1 || (long)ngroups * np > (1L << 24))
            return fail(HCG_ERR_INVALID_ARGUMENT, "ngroups must be in 1..%ld", (1L << 24) / np);
        if (group_of)
            for (int32_t k = 0; k < nsol; k++)
                if (group_of[k] < 0 || group_of[k] >= ngroups)
                    return fail(HCG_ERR_INVALID_ARGUMENT, "group_of_candidate[%d] = %d out of range", k, group_of[k]);
    } else {
        ngroups = 0;
        group_of = nullptr;
    }
    // per group and parameter: the FP64 box and the double-double ball (default: the point)
    std::vector<double> ival(4 * (size_t)np * ngroups), ball(6 * (size_t)np * ngroups);
    for (long q = 0; q < (long)np * ngroups; q++) {
        const double re = params[2 * q], im = params[2 * q + 1];
        double b[4] = {re, re, im, im};
        if (params_ival) for (int t = 0; t < 4; t++) b[t] = params_ival[4 * q + t];
        if (!(b[0] <= b[1]) || !(b[2] <= b[3]))
            return fail(HCG_ERR_INVALID_ARGUMENT, "parameter %ld: invalid interval", q);
        if (!(b[0] <= re && re <= b[1] && b[2] <= im && im <= b[3]))
            return fail(HCG_ERR_INVALID_ARGUMENT, "parameter %ld: the point is not in its interval", q);
        double c[6] = {re, 0.0, 0.0, im, 0.0, 0.0};
        if (params_ball) for (int t = 0; t < 6; t++) c[t] = params_ball[6 * q + t];
        if (!(c[2] >= 0.0) || !(c[5] >= 0.0)) return fail(HCG_ERR_INVALID_ARGUMENT, "parameter %ld: invalid ball radius", q);
        for (int t = 0; t < 4; t++) ival[4 * q + t] = b[t];
        for (int t = 0; t < 6; t++) ball[6 * q + t] = c[t];
    }
    HIPCHK(hipSetDevice(ctx->device));
    const size_t zc = sizeof(double) * 2 * n * (size_t)nsol, zb = 2 * zc;
    DevBuf d_c, d_p, d_ip, d_bp, d_g, d_out, d_b1, d_b0, d_r;
    HIPCHK(hipMalloc(&d_c.p, zc));
    HIPCHK(hipMalloc(&d_out.p, sizeof(hcg_certificate) * nsol));
    HIPCHK(hipMalloc(&d_b1.p, zb));
    HIPCHK(hipMalloc(&d_b0.p, zb));
    HIPCHK(hipMalloc(&d_r.p, zc));
    HIPCHK(hipMemcpyAsync(d_c.p, cand, zc, hipMemcpyHostToDevice, ctx->stream));
    if (np > 0) {
        const size_t m = (size_t)np * ngroups;
        HIPCHK(hipMalloc(&d_p.p, sizeof(double) * 2 * m));
        HIPCHK(hipMalloc(&d_ip.p, sizeof(double) * 4 * m));
        HIPCHK(hipMalloc(&d_bp.p, sizeof(double) * 6 * m));
        HIPCHK(hipMemcpyAsync(d_p.p, params, sizeof(double) * 2 * m, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_ip.p, ival.data(), sizeof(double) * 4 * m, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(d_bp.p, ball.data(), sizeof(double) * 6 * m, hipMemcpyHostToDevice, ctx->stream));
        if (group_of) {
            HIPCHK(hipMalloc(&d_g.p, sizeof(int32_t) * nsol));
            HIPCHK(hipMemcpyAsync(d_g.p, group_of, sizeof(int32_t) * nsol, hipMemcpyHostToDevice, ctx->stream));
        }
    }
    hcg_cert_kargs ka;
    std::memset(&ka, 0, sizeof ka);
    ka.nsol = nsol;
    ka.refine = refine ? 1 : 0;
    ka.real_system = real_system ? 1 : 0;
    ka.extended = extended ? 1 : 0;
    ka.candidates = (const double*)d_c.p;
    ka.params = (const double*)d_p.p;
    ka.params_ival = (const double*)d_ip.p;
    ka.params_ball = (const double*)d_bp.p;
    ka.group_of = (const int32_t*)d_g.p;
    ka.out = (hcg_certificate*)d_out.p;
    ka.box1 = (double*)d_b1.p;
    ka.box0 = (double*)d_b0.p;
    ka.refined = (double*)d_r.p;
    const unsigned grid = (unsigned)(((long)nsol + 63) / 64);  // one lane per candidate, 64 per workgroup
    void* args[] = {&ka};
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(hipModuleLaunchKernel(prog->kernel, grid, 1, 1, 64, 1, 1, 0, ctx->stream, args, nullptr));
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipMemcpyAsync(out, d_out.p, sizeof(hcg_certificate) * nsol, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(box1, d_b1.p, zb, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(box0, d_b0.p, zb, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(refined, d_r.p, zc, hipMemcpyDeviceToHost, ctx->stream));
    return finish_launch(ctx);
}

extern "C" int hcg_certifier_compile_only(const hcg_system* F, int32_t n, size_t* code_size) {
    try {
        std::string src;
        int rc = certifier_source(F, n, &src);
        if (rc) return rc;
        std::vector<char> code;
        rc = rtc_compile(src, &code);
        if (rc) return rc;
        if (code_size) *code_size = code.size();
        return HCG_OK;
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_certifier_compile_only: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_certifier_compile_only: unexpected error");
    }
}
extern "C" int hcg_certifier_create(hcg_ctx* ctx, const hcg_system* F, int32_t n, hcg_program** out) {
    try {
        return hcg_certifier_create_impl(ctx, F, n, out);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_certifier_create: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_certifier_create: unexpected error");
    }
}
extern "C" int hcg_certify(hcg_ctx* ctx, const hcg_program* cert, int32_t nsol, const double* candidates,
                           int32_t ngroups, const double* params, const double* params_ival,
                           const double* params_ball, const int32_t* group_of_candidate, int32_t refine,
                           int32_t extended, int32_t real_system, hcg_certificate* out, double* box1, double* box0,
                           double* refined) {
    try {
        return hcg_certify_impl(ctx, cert, nsol, candidates, ngroups, params, params_ival, params_ball,
                                group_of_candidate, refine, extended, real_system, out, box1, box0, refined);
    } catch (const std::exception& ex) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_certify: %s", ex.what());
    } catch (...) {
        return fail(HCG_ERR_INVALID_ARGUMENT, "hcg_certify: unexpected error");
    }
}
