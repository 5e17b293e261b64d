// The GPU engine behind include/tesseract_gpu.h: hipRTC compilation of lowered programs, device
// state in the word layout, pass launches and exhaustive enumeration.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "lower.hpp"
#include "tesseract_gpu.h"

using namespace ts;

#include "engine_internal.hpp"

struct gpu_code {
    gpu_engine            *e = nullptr;
    Lowered                lw;
    hipModule_t            mod = nullptr;
    std::vector<hipFunction_t> fns;
    std::vector<uint32_t>  slices;      // per segment
    std::vector<uint32_t>  always;      // per segment: slices every pass runs (the rest: last pass)
    gpu_code_info          info{};
    gpu_tune               tune{};
    uint32_t               ncells = 0;
    std::vector<uint8_t>   kind;        // per cell: TS_CK_*; constants as plain
    std::vector<uint8_t>   id_dim;
    // partial passes (gpu_run_steps): a copy of the program and the compiled instruction ranges
    ts_program             prog{};
    std::map<std::pair<uint64_t, uint64_t>, gpu_code *> parts;
};

struct gpu_state {
    gpu_code *c = nullptr;
    unsigned  groups = 0;
    uint32_t  nw = 0;        // words = groups * 16
    uint32_t *slots = nullptr, *init = nullptr;
    int       latest = 1;    // parity holding the newest values of written cells
    // segmented programs: `graph_passes` passes of segment launches captured once per starting
    // parity and replayed, instead of one launch per segment per pass
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    unsigned       graph_passes = 0;
    uint64_t       pc = 0;   // the next instruction of every machine (0 = between passes)
};

namespace {

void say(char *err, size_t errlen, const char *fmt, ...)
{
    if (!err || !errlen)
        return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, errlen, fmt, ap);
    va_end(ap);
}

#define HIPCHK(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            say(err, errlen, "%s failed: %s", #call, hipGetErrorString(e_));                \
            return 0;                                                                       \
        }                                                                                   \
    } while (0)

gpu_tune defaults(const gpu_engine *e, const gpu_tune *t)
{
    gpu_tune d{};
    if (t)
        d = *t;
    if (!d.block || d.block % 64 || d.block > 1024)
        d.block = 256;
    if (!d.lut3)
        d.lut3 = 1;
    if (!d.grid)  // a few words per thread: 8 per CU measured 5.7% slower on mul16 (tail waves)
        d.grid = (unsigned)std::max(1, e->cus) * 256;
    if (!d.graph_passes)
        d.graph_passes = 32;
    return d;
}

}  // namespace

// hipRTC: source -> gfx950 code object (needs no device)
const char *ts_internal::sched_strategy(unsigned sched)
{
    switch (sched) {
    case 2: return "iterative-ilp";
    case 3: return "max-ilp";
    case 4: return "iterative-minreg";
    default: return nullptr;  // the compiler's default
    }
}

int ts_internal::rtc_compile(const std::string &src, const std::string &arch_name, std::vector<char> &code,
                double *seconds, char *err, size_t errlen, unsigned sched)
{
    if (sched == 2) {
        say(err, errlen, "the iterative-ilp scheduling strategy (sched 2) is not used: the compiler "
                         "faults on some kernels under it; take 0/1, 3 or 4");
        return 0;
    }
    auto t0 = std::chrono::steady_clock::now();
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "tesseract_gen.hip", 0, nullptr, nullptr) !=
        HIPRTC_SUCCESS) {
        say(err, errlen, "hiprtcCreateProgram failed");
        return 0;
    }
    std::string arch = "--offload-arch=" + arch_name;
    const char *strategy = sched_strategy(sched);
    const std::string sched_opt = std::string("-amdgpu-sched-strategy=") + (strategy ? strategy : "");
    const char *opts[] = {arch.c_str(), "-O3", "-Xclang", "-target-feature", "-Xclang",
                          "-packed-fp32-ops", "-mllvm", sched_opt.c_str()};
    hiprtcResult r = hiprtcCompileProgram(prog, strategy ? 8 : 6, opts);
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n + 1, 0);
        hiprtcGetProgramLog(prog, &log[0]);
        say(err, errlen, "hipRTC: %s: %.900s", hiprtcGetErrorString(r), log.c_str());
        hiprtcDestroyProgram(&prog);
        return 0;
    }
    size_t size = 0;
    hiprtcGetCodeSize(prog, &size);
    code.resize(size);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 1;
}

namespace {

int build_module(gpu_engine *e, const std::string &src, hipModule_t *mod, double *seconds, char *err,
                 size_t errlen, unsigned sched = 0)
{
    std::vector<char> code;
    if (!ts_internal::rtc_compile(src, e->arch, code, seconds, err, errlen, sched))
        return 0;
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipModuleLoadData(mod, code.data()));
    return 1;
}

int launch(gpu_engine *e, hipFunction_t f, unsigned grid, unsigned block, void **args, char *err,
           size_t errlen)
{
    HIPCHK(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, e->stream, args, nullptr));
    return 1;
}

uint32_t id_pattern_word(const gpu_code *c, uint32_t cell, unsigned group, unsigned j)
{
    // word j of group g holds machines 32 j .. 32 j + 31 of that group, numbered within the cube
    unsigned cube_groups = 1u << (c->lw.dims - TS_CUBE_DIMS);
    unsigned d = c->id_dim[cell];
    uint32_t w = 0;
    for (unsigned b = 0; b < 32; b++) {
        unsigned machine = (group % cube_groups) * TS_CUBE_SIZE + 32 * j + b;
        w |= (uint32_t)(machine >> d & 1) << b;
    }
    return w;
}

}  // namespace

// ---- engine ---------------------------------------------------------------------------------

extern "C" int gpu_engine_open(int device, gpu_engine **out, char *err, size_t errlen)
{
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        say(err, errlen, "no GPU %d (%d visible)", device, n);
        return 0;
    }
    std::unique_ptr<gpu_engine> e(new gpu_engine());
    e->device = device;
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    e->cus = prop.multiProcessorCount;
    int lds = 0;
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    e->lds_per_block = (size_t)lds;
    e->arch = prop.gcnArchName;
    size_t colon = e->arch.find(':');
    if (colon != std::string::npos)
        e->arch.resize(colon);
    char d[256];
    snprintf(d, sizeof d, "%s, %d CUs, %.0f MHz, %.1f GB", e->arch.c_str(), e->cus,
             prop.clockRate / 1000.0, prop.totalGlobalMem / 1e9);
    e->desc = d;
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&e->ev0));
    HIPCHK(hipEventCreate(&e->ev1));
    if (!ts_internal::static_scratch_init(device, err, errlen))
        return 0;
    *out = e.release();
    return 1;
}

extern "C" void gpu_engine_close(gpu_engine *e)
{
    if (!e)
        return;
    (void)hipSetDevice(e->device);
    if (e->batch_scratch)
        (void)hipFree(e->batch_scratch);
    if (e->batch_stop)
        (void)hipHostFree(e->batch_stop);
    for (hipStream_t &st : e->batch_side)
        if (st)
            (void)hipStreamDestroy(st);
    if (e->ev0)
        (void)hipEventDestroy(e->ev0);
    if (e->ev1)
        (void)hipEventDestroy(e->ev1);
    if (e->stream)
        (void)hipStreamDestroy(e->stream);
    delete e;
}

extern "C" const char *gpu_engine_describe(const gpu_engine *e) { return e->desc.c_str(); }

extern "C" int ts_valu_calibrate(hipStream_t stream, int cus, double *lane_ops_per_s);

extern "C" int gpu_valu_rate(gpu_engine *e, double *lane_ops_per_s, char *err, size_t errlen)
{
    HIPCHK(hipSetDevice(e->device));
    if (!ts_valu_calibrate(e->stream, e->cus, lane_ops_per_s)) {
        say(err, errlen, "the VALU calibration launch failed");
        return 0;
    }
    return 1;
}

extern "C" int ts_hbm_calibrate(hipStream_t stream, int cus, size_t bytes, double *bytes_per_s);

extern "C" int gpu_hbm_rate(gpu_engine *e, size_t bytes, double *bytes_per_s, char *err,
                            size_t errlen)
{
    HIPCHK(hipSetDevice(e->device));
    if (!ts_hbm_calibrate(e->stream, e->cus, bytes, bytes_per_s)) {
        say(err, errlen, "the HBM copy calibration failed (bytes must be a nonzero multiple of 16 "
                         "and two buffers of that size must fit)");
        return 0;
    }
    return 1;
}
extern "C" double gpu_last_ms(const gpu_engine *e) { return e->last_ms; }

// ---- compile --------------------------------------------------------------------------------

namespace {

// program -> lowered passes -> HIP source (host only)
int generate(const ts_program *p, gpu_tune &tune, Lowered &lw, KernelSource &ks, char *err,
             size_t errlen)
{
    if (tune.sched == 2) {
        // the compiler itself faults on some pass kernels under this strategy (the reference's
        // PERM8 program, Life's segment kernel) and takes the process down with it
        say(err, errlen, "the iterative-ilp scheduling strategy (sched 2) is not used for pass "
                         "kernels: the compiler faults on some of them; take 0/1, 3 or 4");
        return 0;
    }
    LowerOptions lo;
    unsigned dims = ts_program_dims(p);
    lo.cube_words_log2 = dims - 5;
    lo.slice_nodes = tune.slice_nodes;
    CodegenOptions co;
    co.block = tune.block;
    co.lut3 = tune.lut3 == 1;
    co.waves = std::min(tune.waves, 8u);
    auto make = [&](unsigned wg_log2, Lowered &l, KernelSource &k) {
        lo.wg_words_log2 = wg_log2;
        std::string msg;
        if (!lower_program(p, lo, l, msg)) {
            say(err, errlen, "%s", msg.c_str());
            return false;
        }
        codegen_run(l, co, k);
        return true;
    };
    // A cube of 128-1,024 words may run in one workgroup (links across its waves through LDS, the
    // pass loop in-kernel) or segmented (slices of threads per segment, a launch per segment).
    // By default the form with the shorter pass for one cube is taken, from the LUTs one thread
    // evaluates in a pass (the workgroup form's whole pass; the segments' largest slices) and the
    // launches: measured on one cube, rank on 12 dims (376 LUTs in one workgroup, 0.6 us per pass,
    // against 2 segments of 1,744, 18 us) and gamma on 12 dims (6,311 LUTs in one workgroup,
    // 94 us, against 2 segments of 5,945, 76 us: the workgroup form carries ~1,850 cells in
    // registers and barriers between its levels), about 44 cycles per LUT in one workgroup, 35
    // in a segment and 12,000 per launch.
    const bool may_wg = dims >= 12 && dims <= 15 && tune.lds_links != 2;
    if (!make(may_wg ? 10 : 0, lw, ks))
        return 0;
    if (may_wg && tune.lds_links != 1 && lw.resident && lw.wg_dims > lw.wave_dims) {
        Lowered lw2;
        KernelSource ks2;
        if (!make(0, lw2, ks2))
            return 0;
        const double cyc_wg = 44.0 * (double)ks.chain_luts;
        const double cyc_seg = 12000.0 * (lw2.resident ? 0 : lw2.segs.size()) +
                               35.0 * (double)ks2.chain_luts;
        if (cyc_seg < cyc_wg) {
            lw = std::move(lw2);
            ks = std::move(ks2);
        }
    }
    if (ks.block)  // the kernels need this workgroup size (a whole cube per workgroup)
        tune.block = ks.block;
    return 1;
}

void fill_info(const Lowered &lw, const KernelSource &ks, gpu_code_info &info)
{
    uint32_t most = 0;
    for (const Segment &s : lw.segs)
        most = std::max<uint32_t>(most, (uint32_t)s.slices.size());
    info.segments = lw.resident ? 1 : (uint32_t)lw.segs.size();
    info.resident = lw.resident;
    info.slices = most;
    info.bitop3 = ks.bitop3;
    info.values = ks.values;
    info.dims = lw.dims;
    info.chain = ks.chain_luts;
}

}  // namespace

extern "C" int gpu_compile(gpu_engine *e, const ts_program *p, const gpu_tune *tune, gpu_code **out,
                           char **listing, char *err, size_t errlen)
{
    *out = nullptr;
    if (listing)
        *listing = nullptr;
    std::unique_ptr<gpu_code> c(new gpu_code());
    c->e = e;
    c->tune = defaults(e, tune);
    c->ncells = p->ncells;
    c->kind.assign(p->ncells, TS_CK_PLAIN);
    c->id_dim.assign(p->ncells, 0);
    for (uint32_t i = 0; p->special && i < p->ncells; i++) {
        c->kind[i] = p->special[i].kind;
        c->id_dim[i] = p->special[i].dim;
    }
    KernelSource ks;
    if (!generate(p, c->tune, c->lw, ks, err, errlen))
        return 0;
    if (listing)
        *listing = strdup(ks.text.c_str());
    double secs = 0;
    if (!build_module(e, ks.text, &c->mod, &secs, err, errlen, c->tune.sched))
        return 0;
    for (const std::string &name : ks.names) {
        hipFunction_t f;
        HIPCHK(hipModuleGetFunction(&f, c->mod, name.c_str()));
        c->fns.push_back(f);
    }
    for (const Segment &s : c->lw.segs) {
        c->slices.push_back((uint32_t)s.slices.size());
        c->always.push_back((uint32_t)s.always);
    }
    fill_info(c->lw, ks, c->info);
    c->info.compile_seconds = secs;
    if (!ts_program_copy(&c->prog, p)) {
        say(err, errlen, "out of memory");
        return 0;
    }
    *out = c.release();
    return 1;
}

extern "C" int gpu_source(const ts_program *p, const gpu_tune *tune, char **source,
                          gpu_code_info *info, char *err, size_t errlen)
{
    gpu_engine none;
    none.cus = 256;
    gpu_tune t = defaults(&none, tune);
    Lowered lw;
    KernelSource ks;
    if (!generate(p, t, lw, ks, err, errlen))
        return 0;
    if (info) {
        memset(info, 0, sizeof *info);
        fill_info(lw, ks, *info);
    }
    *source = strdup(ks.text.c_str());
    return 1;
}

extern "C" int gpu_rtc_check(const char *source, const char *arch, double *seconds,
                             size_t *code_bytes, char *err, size_t errlen)
{
    std::vector<char> code;
    if (!ts_internal::rtc_compile(source, arch ? arch : "gfx950", code, seconds, err, errlen))
        return 0;
    if (code_bytes)
        *code_bytes = code.size();
    return 1;
}

extern "C" void gpu_code_free(gpu_code *c)
{
    if (!c)
        return;
    if (c->mod) {
        (void)hipSetDevice(c->e->device);
        (void)hipModuleUnload(c->mod);
    }
    for (auto &kv : c->parts)
        gpu_code_free(kv.second);
    ts_program_free(&c->prog);
    delete c;
}

extern "C" int gpu_code_stats(const gpu_code *c, gpu_code_info *info)
{
    *info = c->info;
    return 1;
}

// ---- state ----------------------------------------------------------------------------------

extern "C" int gpu_state_new(gpu_code *c, unsigned groups, gpu_state **out, char *err, size_t errlen)
{
    *out = nullptr;
    unsigned cube_groups = 1u << (c->lw.dims - TS_CUBE_DIMS);
    if (!groups || groups % cube_groups) {
        say(err, errlen, "a cube of this program has %u groups of 512 machines; %u is not a multiple",
            cube_groups, groups);
        return 0;
    }
    if ((uint64_t)groups * 16 > 0xFFFFFFFFull / 2) {
        say(err, errlen, "too many groups");
        return 0;
    }
    std::unique_ptr<gpu_state> s(new gpu_state());
    s->c = c;
    s->groups = groups;
    s->nw = groups * 16;
    HIPCHK(hipSetDevice(c->e->device));
    size_t slot_bytes = (size_t)2 * c->lw.nslots * s->nw * 4, init_bytes = (size_t)c->lw.ninit * s->nw * 4;
    HIPCHK(hipMalloc(&s->slots, std::max<size_t>(slot_bytes, 4)));
    HIPCHK(hipMalloc(&s->init, std::max<size_t>(init_bytes, 4)));
    *out = s.release();
    return gpu_state_reset(*out, err, errlen);
}

extern "C" void gpu_state_free(gpu_state *s)
{
    if (!s)
        return;
    (void)hipSetDevice(s->c->e->device);
    for (int k = 0; k < 2; k++)
        if (s->graph[k])
            (void)hipGraphExecDestroy(s->graph[k]);
    (void)hipFree(s->slots);
    (void)hipFree(s->init);
    delete s;
}

namespace {

// Staging copies of the groups [g0, g0 + n): per stored cell, n * 16 words, cell-major.
struct Staging {
    std::vector<uint32_t> slot_words, init_words;
};

int put_lines(gpu_state *s, unsigned g0, unsigned n, const ts_line *mem, const ts_line *rr,
              char *err, size_t errlen)
{
    gpu_code *c = s->c;
    const Lowered &lw = c->lw;
    const size_t per = (size_t)n * 16;
    std::vector<uint32_t> slot_words((size_t)lw.nslots * per), init_words((size_t)lw.ninit * per);
    for (uint32_t cell = 0; cell <= lw.ncells; cell++) {
        int32_t fs = lw.final_slot[cell], is = lw.init_slot[cell];
        if (fs < 0 && is < 0)
            continue;
        uint32_t *dst = fs >= 0 ? &slot_words[(size_t)fs * per] : &init_words[(size_t)is * per];
        for (unsigned g = 0; g < n; g++) {
            const ts_line *line = cell == lw.ncells ? &rr[g] : &mem[(size_t)g * lw.ncells + cell];
            memcpy(dst + (size_t)g * 16, line->w, 64);
        }
    }
    hipStream_t st = c->e->stream;
    if (lw.nslots)
        HIPCHK(hipMemcpy2DAsync(s->slots + (size_t)s->latest * lw.nslots * s->nw + (size_t)g0 * 16,
                                (size_t)s->nw * 4, slot_words.data(), per * 4, per * 4, lw.nslots,
                                hipMemcpyHostToDevice, st));
    if (lw.ninit)
        HIPCHK(hipMemcpy2DAsync(s->init + (size_t)g0 * 16, (size_t)s->nw * 4, init_words.data(),
                                per * 4, per * 4, lw.ninit, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return 1;
}

}  // namespace

extern "C" int gpu_state_reset(gpu_state *s, char *err, size_t errlen)
{
    // power-on: all 0, RR 0; ONE and id cells are constants the kernels synthesise
    gpu_code *c = s->c;
    HIPCHK(hipSetDevice(c->e->device));
    s->latest = 1;
    s->pc = 0;
    size_t slot_bytes = (size_t)2 * c->lw.nslots * s->nw * 4, init_bytes = (size_t)c->lw.ninit * s->nw * 4;
    if (slot_bytes)
        HIPCHK(hipMemsetAsync(s->slots, 0, slot_bytes, c->e->stream));
    if (init_bytes)
        HIPCHK(hipMemsetAsync(s->init, 0, init_bytes, c->e->stream));
    HIPCHK(hipStreamSynchronize(c->e->stream));
    return 1;
}

extern "C" int gpu_state_put(gpu_state *s, unsigned group0, unsigned ngroups, const ts_line *mem,
                             const ts_line *rr, char *err, size_t errlen)
{
    if ((uint64_t)group0 + ngroups > s->groups) {
        say(err, errlen, "groups %u..%u outside the state's %u", group0, group0 + ngroups, s->groups);
        return 0;
    }
    HIPCHK(hipSetDevice(s->c->e->device));
    return put_lines(s, group0, ngroups, mem, rr, err, errlen);
}

extern "C" int gpu_state_get(gpu_state *s, unsigned group0, unsigned ngroups, ts_line *mem,
                             ts_line *rr, char *err, size_t errlen)
{
    if ((uint64_t)group0 + ngroups > s->groups) {
        say(err, errlen, "groups %u..%u outside the state's %u", group0, group0 + ngroups, s->groups);
        return 0;
    }
    gpu_code *c = s->c;
    const Lowered &lw = c->lw;
    HIPCHK(hipSetDevice(c->e->device));
    const size_t per = (size_t)ngroups * 16;
    std::vector<uint32_t> slot_words((size_t)lw.nslots * per), init_words((size_t)lw.ninit * per);
    hipStream_t st = c->e->stream;
    if (lw.nslots)
        HIPCHK(hipMemcpy2DAsync(slot_words.data(), per * 4,
                                s->slots + (size_t)s->latest * lw.nslots * s->nw + (size_t)group0 * 16,
                                (size_t)s->nw * 4, per * 4, lw.nslots, hipMemcpyDeviceToHost, st));
    if (lw.ninit)
        HIPCHK(hipMemcpy2DAsync(init_words.data(), per * 4, s->init + (size_t)group0 * 16,
                                (size_t)s->nw * 4, per * 4, lw.ninit, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (unsigned g = 0; g < ngroups; g++) {
        for (uint32_t cell = 0; cell <= lw.ncells; cell++) {
            ts_line *line = cell == lw.ncells ? &rr[g] : &mem[(size_t)g * lw.ncells + cell];
            int32_t fs = lw.final_slot[cell], is = lw.init_slot[cell];
            if (fs >= 0 || is >= 0) {
                const uint32_t *src = fs >= 0 ? &slot_words[(size_t)fs * per] : &init_words[(size_t)is * per];
                memcpy(line->w, src + (size_t)g * 16, 64);
            } else if (cell == TS_CELL_ONE) {
                memset(line->w, 0xFF, 64);
            } else if (cell < lw.ncells && c->kind[cell] == TS_CK_ID) {
                uint32_t words[16];
                for (unsigned j = 0; j < 16; j++)
                    words[j] = id_pattern_word(c, cell, group0 + g, j);
                memcpy(line->w, words, 64);
            } else {
                memset(line->w, 0, 64);
            }
        }
    }
    return 1;
}

// ---- run ------------------------------------------------------------------------------------

extern "C" int gpu_run(gpu_code *c, gpu_state *s, uint64_t passes, char *err, size_t errlen)
{
    if (s->c != c) {
        say(err, errlen, "the state belongs to another program");
        return 0;
    }
    if (s->pc) {
        say(err, errlen, "a pass is in progress (next instruction %llu): finish it with "
                         "gpu_run_steps", (unsigned long long)s->pc);
        return 0;
    }
    gpu_engine *e = c->e;
    HIPCHK(hipSetDevice(e->device));
    const Lowered &lw = c->lw;
    const unsigned block = c->tune.block;
    HIPCHK(hipEventRecord(e->ev0, e->stream));
    uint32_t nw = s->nw;
    if (lw.resident) {
        uint64_t left = passes;
        while (left) {
            uint32_t now = (uint32_t)std::min<uint64_t>(left, 1u << 30);
            uint32_t *prev = s->slots + (size_t)s->latest * lw.nslots * nw;
            uint32_t *cur = s->slots + (size_t)(1 - s->latest) * lw.nslots * nw;
            void *args[] = {&prev, &cur, &s->init, &nw, &now};
            if (!launch(e, c->fns[0], (nw + block - 1) / block, block, args, err, errlen))
                return 0;
            s->latest = 1 - s->latest;
            left -= now;
        }
    } else {
        // one pass: every segment's launch, parity flipping after the pass. Only the last pass of
        // the call (`full`) computes the cells no later pass reads before writing them (the
        // caller may read them afterwards); the others run each segment's `always` slices.
        auto one_pass = [&](int latest, bool full) -> int {
            uint32_t *prev = s->slots + (size_t)latest * lw.nslots * nw;
            uint32_t *cur = s->slots + (size_t)(1 - latest) * lw.nslots * nw;
            for (size_t k = 0; k < c->fns.size(); k++) {
                const uint32_t nslices = full ? c->slices[k] : c->always[k];
                if (!nslices)
                    continue;
                uint64_t threads = (uint64_t)nslices * nw;
                void *args[] = {&prev, &cur, &s->init, &nw};
                if (!launch(e, c->fns[k], (unsigned)((threads + block - 1) / block), block, args, err,
                            errlen))
                    return 0;
            }
            return 1;
        };
        // captured graphs hold passes before the last one; the last pass is launched directly
        const unsigned gp = c->tune.graph_passes;
        // the graph of gp passes starting from parity `lat`, captured on first use (kernel
        // launches only); an odd gp alternates between the two parities' graphs
        auto graph_for = [&](int lat) -> hipGraphExec_t {
            if (s->graph[lat])
                return s->graph[lat];
            hipGraph_t g;
            if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
                say(err, errlen, "graph capture could not start");
                return nullptr;
            }
            int l = lat, ok = 1;
            for (unsigned q = 0; q < gp && ok; q++, l = 1 - l)
                ok = one_pass(l, false);
            hipError_t ce = hipStreamEndCapture(e->stream, &g);
            if (!ok || ce != hipSuccess) {
                say(err, errlen, "graph capture failed: %s", hipGetErrorString(ce));
                return nullptr;
            }
            hipError_t ie = hipGraphInstantiate(&s->graph[lat], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                s->graph[lat] = nullptr;
                say(err, errlen, "graph instantiate failed: %s", hipGetErrorString(ie));
                return nullptr;
            }
            return s->graph[lat];
        };
        uint64_t p = 0;
        if (gp > 1)
            for (; p + gp < passes; p += gp) {
                hipGraphExec_t ge = graph_for(s->latest);
                if (!ge)
                    return 0;
                HIPCHK(hipGraphLaunch(ge, e->stream));
                if (gp & 1)
                    s->latest = 1 - s->latest;
            }
        for (; p < passes; p++) {
            if (!one_pass(s->latest, p + 1 == passes))
                return 0;
            s->latest = 1 - s->latest;
        }
    }
    HIPCHK(hipEventRecord(e->ev1, e->stream));
    HIPCHK(hipEventSynchronize(e->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->last_ms = ms;
    return 1;
}

// ---- one machine's cells --------------------------------------------------------------------

namespace {

// the group line holding `cell` of `machine` (after link resolution) and the bit within it
int machine_line(gpu_state *s, uint64_t machine, uint32_t cell, uint32_t *base, uint64_t *at,
                 char *err, size_t errlen)
{
    const gpu_code *c = s->c;
    if (machine >= (uint64_t)s->groups * TS_CUBE_SIZE || cell > c->ncells) {
        say(err, errlen, "machine %llu cell %u outside the state (%llu machines, %u cells + RR)",
            (unsigned long long)machine, cell, (unsigned long long)s->groups * TS_CUBE_SIZE,
            c->ncells);
        return 0;
    }
    unsigned mask = 0;
    *base = cell == c->ncells ? cell : ts_program_resolve(&c->prog, cell, &mask);
    const uint64_t cube = (uint64_t)1 << c->lw.dims;
    *at = machine / cube * cube + ((machine % cube) ^ mask);
    return 1;
}

}  // namespace

// lane_get / ref_get (src/lanes.h:60-62, src/ref.h:28): any cell, links and ids included;
// cell == ncells is RR
extern "C" int gpu_state_peek(gpu_state *s, uint64_t machine, uint32_t cell, uint8_t *v, char *err,
                              size_t errlen)
{
    uint32_t base;
    uint64_t at;
    if (!machine_line(s, machine, cell, &base, &at, err, errlen))
        return 0;
    const uint32_t n = s->c->ncells;
    std::vector<ts_line> mem(n);
    ts_line rr;
    if (!gpu_state_get(s, (unsigned)(at / TS_CUBE_SIZE), 1, mem.data(), &rr, err, errlen))
        return 0;
    const ts_line &l = base == n ? rr : mem[base];
    const unsigned k = (unsigned)(at % TS_CUBE_SIZE);
    *v = (uint8_t)(l.w[k / 64] >> (k % 64) & 1);
    return 1;
}

// lane_set / ref_set (src/lanes.h:61,63, src/ref.h:29): ignored for read-only cells
extern "C" int gpu_state_poke(gpu_state *s, uint64_t machine, uint32_t cell, uint8_t v, char *err,
                              size_t errlen)
{
    uint32_t base;
    uint64_t at;
    if (!machine_line(s, machine, cell, &base, &at, err, errlen))
        return 0;
    const uint32_t n = s->c->ncells;
    if (cell < n && ts_program_readonly(&s->c->prog, cell))
        return 1;
    const unsigned g = (unsigned)(machine / TS_CUBE_SIZE), k = (unsigned)(machine % TS_CUBE_SIZE);
    std::vector<ts_line> mem(n);
    ts_line rr;
    if (!gpu_state_get(s, g, 1, mem.data(), &rr, err, errlen))
        return 0;
    ts_line &l = cell == n ? rr : mem[cell];
    const uint64_t bit = 1ull << (k % 64);
    l.w[k / 64] = v ? l.w[k / 64] | bit : l.w[k / 64] & ~bit;
    return gpu_state_put(s, g, 1, mem.data(), &rr, err, errlen);
}

// ---- partial passes -------------------------------------------------------------------------

namespace {

// instructions [from, from + n) of c's program in every machine: the sub-program of those
// instructions, compiled once, run for one pass on a state of its own, the machines' cells and
// RR carried over in the reference's line layout (a sub-program has its own slot layout)
int run_part(gpu_code *c, gpu_state *s, uint64_t from, uint64_t n, char *err, size_t errlen)
{
    gpu_code *&sub = c->parts[{from, n}];
    if (!sub) {
        ts_program p = c->prog;
        p.code = c->prog.code + from;
        p.len = (size_t)n;
        if (!gpu_compile(c->e, &p, &c->tune, &sub, nullptr, err, errlen)) {
            c->parts.erase({from, n});
            return 0;
        }
    }
    gpu_state *ss = nullptr;
    if (!gpu_state_new(sub, s->groups, &ss, err, errlen))
        return 0;
    std::vector<ts_line> mem((size_t)s->groups * c->ncells), rr(s->groups);
    int ok = gpu_state_get(s, 0, s->groups, mem.data(), rr.data(), err, errlen) &&
             gpu_state_put(ss, 0, s->groups, mem.data(), rr.data(), err, errlen) &&
             gpu_run(sub, ss, 1, err, errlen) &&
             gpu_state_get(ss, 0, s->groups, mem.data(), rr.data(), err, errlen) &&
             gpu_state_put(s, 0, s->groups, mem.data(), rr.data(), err, errlen);
    gpu_state_free(ss);
    return ok;
}

}  // namespace

// `steps` instructions in every machine from the state's pc, whole passes through gpu_run: the
// counterpart of the reference's single steps (ref_cube_step, src/ref.c:119-130) and of finishing
// a started pass (src/jit.c:743-746, src/check.c:89-107)
extern "C" int gpu_run_steps(gpu_code *c, gpu_state *s, uint64_t steps, char *err, size_t errlen)
{
    if (s->c != c) {
        say(err, errlen, "the state belongs to another program");
        return 0;
    }
    const uint64_t len = c->prog.len;
    if (!len)
        return 1;
    while (steps) {
        if (s->pc == 0 && steps >= len) {
            const uint64_t whole = steps / len;
            if (!gpu_run(c, s, whole, err, errlen))
                return 0;
            steps -= whole * len;
            continue;
        }
        const uint64_t n = std::min(steps, len - s->pc);
        if (!run_part(c, s, s->pc, n, err, errlen))
            return 0;
        s->pc = (s->pc + n) % len;
        steps -= n;
    }
    return 1;
}

extern "C" uint64_t gpu_state_pc(const gpu_state *s) { return s->pc; }

// ---- enumeration ----------------------------------------------------------------------------

using ts_internal::ShareInfo;
using ts_internal::kListCap;

// reference src/enum.c:215-242: passes of 512 in chunks of 64; chunk c belongs to share c % parts
ShareInfo ts_internal::shares(unsigned n, int part, int part_end, int parts)
{
    ShareInfo si{};
    si.space = 1ull << n;
    si.passes = (si.space + 511) / 512;
    si.nchunks = (si.passes + 63) / 64;
    si.nshare = (unsigned)(part_end - part + 1);
    const uint64_t last = si.nchunks - 1;
    for (int r = part; r <= part_end; r++) {
        if ((uint64_t)r >= si.nchunks)
            continue;
        uint64_t count = (si.nchunks - (uint64_t)r + (uint64_t)parts - 1) / (uint64_t)parts;
        si.tried += count << 15;
        si.passes_run += count * 64;
    }
    if ((int)(last % (uint64_t)parts) >= part && (int)(last % (uint64_t)parts) <= part_end) {
        uint64_t size = si.space - (last << 15);
        si.tried -= (1ull << 15) - size;
        si.passes_run -= 64 - (si.passes - last * 64);
    }
    if (parts == 1)
        si.nact_words = (si.space + 31) / 32;
    else
        si.nact_words = ((si.nchunks + (uint64_t)parts - 1) / (uint64_t)parts) * si.nshare * 1024;
    return si;
}

namespace {

}  // namespace

// validation (the reference's messages, src/enum.c:244-268), spec parsing and code generation
namespace {

// Validation with the reference's messages (src/enum.c:244-262), name matching for equiv and spec
// parsing: everything both tiers need before they build code or a tape.
struct EnumPlan {
    ts::EnumLowering el;
    std::vector<int> in_map, out_map;
    ts_spec         *spec = nullptr;
    ~EnumPlan() { ts_spec_free(spec); }
};

int enum_plan(const gpu_enum *en, EnumPlan &plan, char *err, size_t errlen)
{
    const ts_program *a = en->a, *b = en->b;
    if (a->special || (en->mode == TS_EN_EQUIV && b && b->special)) {
        say(err, errlen, "id and link cells connect the lanes, but every input needs a lane of its "
                         "own here");
        return 0;
    }
    size_t n = a->ninputs;
    if (n > TS_EN_MAX_IN) {
        say(err, errlen, "%zu inputs is too many to try them all (at most %d)", n, TS_EN_MAX_IN);
        return 0;
    }
    if (en->parts < 1 || en->part < 0 || en->part > en->part_end || en->part_end >= en->parts ||
        en->ticks < 1) {
        say(err, errlen, "bad --part or --ticks");
        return 0;
    }
    if (en->mode == TS_EN_EQUIV && !b) {
        say(err, errlen, "equiv needs two programs");
        return 0;
    }
    plan.in_map.assign(n + 1, 0);
    plan.out_map.assign(a->noutputs + 1, 0);
    if (en->mode == TS_EN_EQUIV &&
        !ts_enum_match(a, b, plan.in_map.data(), plan.out_map.data(), err, errlen))
        return 0;
    ts::EnumLowering &el = plan.el;
    el.mode = en->mode;
    el.n = (unsigned)n;
    el.ticks = en->ticks;
    if (en->mode != TS_EN_EQUIV) {
        plan.spec = ts_spec_parse(en->spec ? en->spec : "", a, en->mode == TS_EN_VERIFY, err, errlen);
        if (!plan.spec)
            return 0;
        el.width = ts_spec_width(plan.spec);
        for (int bit = 0; bit < el.width; bit++)
            el.out_cells_a.push_back(ts_spec_out_cell(plan.spec, bit));
        if (en->mode == TS_EN_SEARCH)
            el.target = ts_spec_value(plan.spec, 0);
    }
    return 1;
}

}  // namespace

int ts_internal::enum_prepare(const gpu_engine *e, const gpu_enum *en, gpu_tune &tune,
                              KernelSource &ks, char *err, size_t errlen)
{
    EnumPlan plan;
    if (!enum_plan(en, plan, err, errlen))
        return 0;
    tune = defaults(e, en->tune);
    CodegenOptions co;
    co.block = tune.block;
    co.lut3 = tune.lut3 == 1;
    co.waves = std::min(tune.waves, 8u);
    co.group = tune.group;
    std::string msg;
    if (!codegen_enum(en->a, en->b, plan.in_map.data(), plan.out_map.data(), plan.el, co, ks, msg,
                      plan.spec)) {
        say(err, errlen, "%s", msg.c_str());
        return 0;
    }
    return 1;
}

int ts_internal::enum_tape(const gpu_engine *e, const gpu_enum *en, gpu_tune &tune,
                           unsigned max_slots, Tape &t, char *err, size_t errlen)
{
    EnumPlan plan;
    if (!enum_plan(en, plan, err, errlen))
        return 0;
    tune = defaults(e, en->tune);
    std::string msg;
    if (!build_tape(en->a, en->b, plan.in_map.data(), plan.out_map.data(), plan.el, max_slots, t,
                    msg, plan.spec)) {
        say(err, errlen, "%s", msg.c_str());
        return 0;
    }
    return 1;
}

extern "C" int gpu_enum_source(const gpu_enum *en, char **source, char *err, size_t errlen)
{
    gpu_engine none;
    none.cus = 256;
    gpu_tune tune;
    KernelSource ks;
    if (!ts_internal::enum_prepare(&none, en, tune, ks, err, errlen))
        return 0;
    *source = strdup(ks.text.c_str());
    return 1;
}

extern "C" int gpu_enum_tape(const gpu_enum *en, unsigned max_slots, gpu_tape *out, char *err,
                             size_t errlen)
{
    memset(out, 0, sizeof *out);
    gpu_engine none;
    none.cus = 256;
    gpu_tune tune;
    Tape t;
    if (!ts_internal::enum_tape(&none, en, tune, max_slots ? max_slots : 640, t, err, errlen))
        return 0;
    auto dup = [](const std::vector<uint32_t> &v) {
        uint32_t *p = (uint32_t *)malloc(std::max<size_t>(1, v.size()) * sizeof(uint32_t));
        if (p && !v.empty())
            memcpy(p, v.data(), v.size() * sizeof(uint32_t));
        return p;
    };
    out->ops = dup(t.ops);
    out->count = dup(t.count);
    out->leaves = dup(t.leaves);
    if (!out->ops || !out->count || !out->leaves) {
        gpu_tape_free(out);
        say(err, errlen, "out of memory");
        return 0;
    }
    out->ngroups = (uint32_t)t.count.size();
    out->nleaves = (uint32_t)t.leaves.size();
    out->root = t.root;
    out->slots = t.slots;
    out->and_nodes = t.and_nodes;
    return 1;
}

extern "C" void gpu_tape_free(gpu_tape *t)
{
    free(t->ops);
    free(t->count);
    free(t->leaves);
    memset(t, 0, sizeof *t);
}

struct gpu_enum_code {
    gpu_engine   *e = nullptr;
    hipModule_t   mod = nullptr;
    hipFunction_t fn = nullptr;
    bool          tape = false;  // the interpreter kernel runs `td` instead of a compiled kernel
    ts_internal::TapeDev td{};
    unsigned      tape_blocks_per_cu = 1;
    gpu_tune      tune{};
    int           part = 0, part_end = 0, parts = 1;
    unsigned      n = 0, ticks = 1;
    uint64_t      len_total = 0;  // instructions per pass of a (+ b for equiv)
    uint64_t     *d_out = nullptr, *d_list = nullptr;
    double        compile_seconds = 0;
    unsigned      group = 1;          // words a thread of the kernel takes per step
    unsigned      sched = 1;          // the scheduling strategy of the kept kernel
    double        sched_ms[5] = {0, 0, 0, 0, 0};
    uint64_t      sample_words = 0;
};

namespace {

// the hit list of an enumeration code, allocated once and kept with it
int enum_buffers(gpu_enum_code *c, char *err, size_t errlen)
{
    if (c->d_out)
        return 1;
    if (hipMalloc(&c->d_out, 2 * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&c->d_list, 2 * kListCap * sizeof(uint64_t)) != hipSuccess) {
        if (c->d_out)
            (void)hipFree(c->d_out);
        c->d_out = nullptr;
        say(err, errlen, "hipMalloc failed for the hit list");
        return 0;
    }
    return 1;
}

// Compiles the enumeration kernel with scheduling strategies 1 and 3 (in parallel host threads),
// times each on the first `sample` words of the share (a warm-up launch, then the best of two)
// and keeps the fastest module in c.
int enum_autotune(gpu_enum_code *c, const std::string &src, uint64_t sample, char *err,
                  size_t errlen)
{
    gpu_engine *e = c->e;
    constexpr int kCand = 2;
    const unsigned cand[kCand] = {1, 3};
    std::vector<char> code[kCand];
    std::string msg[kCand];
    double secs[kCand] = {0, 0};
    int ok[kCand] = {0, 0};
    {
        std::vector<std::thread> th;
        for (int k = 0; k < kCand; k++)
            th.emplace_back([&, k] {
                char m[1024] = "";
                ok[k] = ts_internal::rtc_compile(src, e->arch, code[k], &secs[k], m, sizeof m, cand[k]);
                msg[k] = m;
            });
        for (std::thread &t : th)
            t.join();
    }
    HIPCHK(hipSetDevice(e->device));
    if (!enum_buffers(c, err, errlen))
        return 0;
    hipModule_t mods[kCand] = {nullptr, nullptr};
    double best = 0;
    int keep = -1;
    ShareInfo si = ts_internal::shares(c->n, c->part, c->part_end, c->parts);
    for (int k = 0; k < kCand; k++) {
        if (!ok[k] || hipModuleLoadData(&mods[k], code[k].data()) != hipSuccess)
            continue;
        hipFunction_t fn;
        if (hipModuleGetFunction(&fn, mods[k], "ts_enum") != hipSuccess)
            continue;
        uint64_t *d_out = c->d_out, *d_list = c->d_list, capv = kListCap, i0 = 0, i1 = sample;
        uint64_t space = si.space;
        uint32_t parts = (uint32_t)c->parts, part = (uint32_t)c->part, nshare = si.nshare;
        void *args[] = {&d_out, &d_list, &capv, &i0, &i1, &space, &parts, &part, &nshare};
        const gpu_tune &tune = c->tune;
        const uint64_t units = (sample + c->group - 1) / c->group;
        unsigned grid = (unsigned)std::max<uint64_t>(
            1, std::min<uint64_t>((units + tune.block - 1) / tune.block, tune.grid));
        double ms_best = 0;
        for (int rep = 0; rep < 3; rep++) {
            HIPCHK(hipEventRecord(e->ev0, e->stream));
            HIPCHK(hipModuleLaunchKernel(fn, grid, 1, 1, tune.block, 1, 1, 0, e->stream, args,
                                         nullptr));
            HIPCHK(hipEventRecord(e->ev1, e->stream));
            HIPCHK(hipEventSynchronize(e->ev1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
            if (rep && (ms_best == 0 || ms < ms_best))
                ms_best = ms;
        }
        c->sched_ms[cand[k]] = ms_best;
        if (keep < 0 || ms_best < best) {
            best = ms_best;
            keep = k;
        }
    }
    for (int k = 0; k < kCand; k++)
        if (mods[k] && k != keep)
            (void)hipModuleUnload(mods[k]);
    if (keep < 0) {
        say(err, errlen, "%s", msg[0].empty() ? "no scheduling strategy gave a kernel" : msg[0].c_str());
        return 0;
    }
    c->mod = mods[keep];
    c->sched = cand[keep];
    c->sample_words = sample;
    c->compile_seconds = secs[0] + secs[1];
    HIPCHK(hipModuleGetFunction(&c->fn, c->mod, "ts_enum"));
    return 1;
}

}  // namespace


extern "C" int gpu_enum_compile(gpu_engine *e, const gpu_enum *en, gpu_enum_code **out, char *err,
                                size_t errlen)
{
    *out = nullptr;
    std::unique_ptr<gpu_enum_code> c(new gpu_enum_code());
    c->e = e;
    auto fail = [&](void) {
        if (c->mod)
            (void)hipModuleUnload(c->mod);
        c->mod = nullptr;
        ts_internal::tape_free(c->td);
    };
    if (en->tune && en->tune->tape) {
        auto t0 = std::chrono::steady_clock::now();
        Tape t;
        if (!ts_internal::enum_tape(e, en, c->tune, ts_internal::tape_max_slots(e), t, err,
                                    errlen))
            return 0;
        HIPCHK(hipSetDevice(e->device));
        if (!ts_internal::tape_upload(t, c->td) ||
            !ts_internal::tape_prepare(e->device, c->td, &c->tape_blocks_per_cu)) {
            fail();
            say(err, errlen, "could not place the tape on the device");
            return 0;
        }
        c->tape = true;
        c->compile_seconds =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    } else {
        KernelSource ks;
        if (!ts_internal::enum_prepare(e, en, c->tune, ks, err, errlen))
            return 0;
        c->group = ks.group;
        c->part = en->part;
        c->part_end = en->part_end;
        c->parts = en->parts;
        c->n = (unsigned)en->a->ninputs;
        c->ticks = en->ticks;
        c->len_total = en->a->len + (en->mode == TS_EN_EQUIV ? en->b->len : 0);
        // long enumerations choose their scheduling strategy by timing a sample of their words
        const ShareInfo si = ts_internal::shares(c->n, c->part, c->part_end, c->parts);
        const double work = (double)si.tried * (double)c->len_total * c->ticks;
        if (c->tune.sched == 0 && work >= 1099511627776.0) {  // 2^40 1-bit ops
            const uint64_t sample =
                std::min<uint64_t>(si.nact_words, std::max<uint64_t>(1u << 20, si.nact_words / 16));
            if (!enum_autotune(c.get(), ks.text, sample, err, errlen)) {
                fail();
                return 0;
            }
        } else {
            c->sched = c->tune.sched ? c->tune.sched : ts_internal::kEnumSched;
            if (!build_module(e, ks.text, &c->mod, &c->compile_seconds, err, errlen, c->sched))
                return 0;
            if (hipModuleGetFunction(&c->fn, c->mod, "ts_enum") != hipSuccess) {
                fail();
                say(err, errlen, "hipModuleGetFunction failed");
                return 0;
            }
        }
    }
    c->part = en->part;
    c->part_end = en->part_end;
    c->parts = en->parts;
    c->n = (unsigned)en->a->ninputs;
    c->ticks = en->ticks;
    c->len_total = en->a->len + (en->mode == TS_EN_EQUIV ? en->b->len : 0);
    *out = c.release();
    return 1;
}

extern "C" void gpu_enum_free(gpu_enum_code *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->e->device);
    if (c->d_out)
        (void)hipFree(c->d_out);
    if (c->d_list)
        (void)hipFree(c->d_list);
    if (c->mod)
        (void)hipModuleUnload(c->mod);
    ts_internal::tape_free(c->td);
    delete c;
}

namespace {

// The whole enumeration of c into r with the hits counted and listed in b. Host side it allocates
// nothing: the listed hit words come back in fixed chunks through a stack buffer.
struct EnumExec {
    gpu_enum_code  *c;
    ts_internal::HitBuffers b;
    ts_enum_result *r;
    char           *err;
    size_t          errlen;
    ShareInfo       si;

    // one launch over words [i0, i1): the hit count and the number of hit words listed
    int launch_range(uint64_t i0, uint64_t i1, uint64_t *hits, uint64_t *nlist, bool timed)
    {
        gpu_engine *e = c->e;
        const gpu_tune &tune = c->tune;
        // cleared by a kernel: no memset node when a caller captures the stream
        if (!ts_internal::clear_device_words(b.d_out, 2, e->stream)) {
            say(err, errlen, "the hit counters could not be cleared");
            return 0;
        }
        uint64_t space = si.space, capv = b.cap;
        uint32_t parts = (uint32_t)c->parts, part = (uint32_t)c->part, nshare = si.nshare;
        uint64_t *d_out = b.d_out, *d_list = b.d_list;
        if (timed)
            HIPCHK(hipEventRecord(e->ev0, e->stream));
        if (c->tape) {
            // one wave per workgroup, four rounds of the resident workgroups
            uint64_t blocks = (i1 - i0 + 63) / 64;
            uint64_t most = (uint64_t)e->cus * c->tape_blocks_per_cu * 4;
            unsigned grid = (unsigned)std::max<uint64_t>(1, std::min(blocks, most));
            if (!ts_internal::tape_launch(c->td, e->stream, grid, d_out, d_list, capv, i0, i1,
                                          space, parts, part, nshare)) {
                say(err, errlen, "the tape launch failed");
                return 0;
            }
        } else {
            // a thread takes c->group consecutive words (a group starts at a multiple of it)
            const uint64_t units = (i1 + c->group - 1) / c->group - i0 / c->group;
            uint64_t blocks = (units + tune.block - 1) / tune.block;
            unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, tune.grid));
            void *args[] = {&d_out, &d_list, &capv, &i0, &i1, &space, &parts, &part, &nshare};
            // a cap on resident waves: LDS the kernel does not use, reserved per workgroup so
            // that only max_waves waves per SIMD fit a CU's LDS
            unsigned lds = 0;
            if (tune.max_waves) {
                const unsigned per_cu = std::max(1u, tune.max_waves * 4 / (tune.block / 64));
                lds = (unsigned)std::min<size_t>(e->lds_per_block, e->lds_per_block / per_cu + 1024);
            }
            HIPCHK(hipModuleLaunchKernel(c->fn, grid, 1, 1, tune.block, 1, 1, lds, e->stream, args,
                                         nullptr));
        }
        if (timed)
            HIPCHK(hipEventRecord(e->ev1, e->stream));
        uint64_t host[2];
        HIPCHK(hipMemcpyAsync(host, d_out, sizeof host, hipMemcpyDeviceToHost, e->stream));
        HIPCHK(hipStreamSynchronize(e->stream));
        if (timed) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
            e->last_ms = ms;
        }
        *hits = host[0];
        *nlist = host[1];
        return 1;
    }

    // the listed hit words into r (the 16 smallest hits are kept)
    int collect(uint64_t nlist)
    {
        constexpr uint64_t kChunk = 256;
        uint64_t buf[2 * kChunk];
        const uint64_t keep = std::min(nlist, b.cap);
        for (uint64_t k = 0; k < keep; k += kChunk) {
            const uint64_t n = std::min(kChunk, keep - k);
            HIPCHK(hipMemcpy(buf, b.d_list + 2 * k, 2 * n * sizeof(uint64_t),
                             hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < n; i++)
                for (uint32_t m = (uint32_t)buf[2 * i + 1]; m; m &= m - 1)
                    ts_enum_add_hit(r, buf[2 * i] * 32 + (uint64_t)__builtin_ctz(m));
        }
        return 1;
    }

    // Words are visited in increasing x within the share, so a range's hits are all smaller than
    // the next range's: when a range has more hit words than the list holds, its left half is
    // listed first and the right half only while fewer than 16 hits are known.
    int collect_range(uint64_t i0, uint64_t i1)
    {
        if (i0 >= i1 || r->nhit >= TS_EN_SHOW)
            return 1;
        uint64_t h = 0, nl = 0;
        if (!launch_range(i0, i1, &h, &nl, false))
            return 0;
        if (nl <= b.cap)
            return collect(nl);
        const uint64_t mid = i0 + (i1 - i0) / 2;
        return collect_range(i0, mid) && collect_range(mid, i1);
    }

    int run()
    {
        memset(r, 0, sizeof *r);
        gpu_engine *e = c->e;
        HIPCHK(hipSetDevice(e->device));
        si = ts_internal::shares(c->n, c->part, c->part_end, c->parts);
        r->space = si.space;
        r->tried = si.tried;
        r->ops = si.passes_run * c->ticks * c->len_total * 512;
        r->threads = 1;
        uint64_t hits = 0, nlist = 0;
        auto t0 = std::chrono::steady_clock::now();
        if (!launch_range(0, si.nact_words, &hits, &nlist, true))
            return 0;
        r->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        r->hits = hits;
        if (nlist <= b.cap)
            return collect(nlist);
        return collect_range(0, si.nact_words);
    }
};

}  // namespace

extern "C" int gpu_enum_stats(const gpu_enum_code *c, gpu_enum_info *info)
{
    memset(info, 0, sizeof *info);
    info->compile_seconds = c->compile_seconds;
    info->tape = c->tape;
    info->sched = c->tape ? 0 : c->sched;
    for (int k = 0; k < 5; k++)
        info->sched_ms[k] = c->sched_ms[k];
    info->sample_words = c->sample_words;
    return 1;
}

extern "C" int gpu_enum_exec(gpu_enum_code *c, ts_enum_result *r, char *err, size_t errlen)
{
    memset(r, 0, sizeof *r);
    HIPCHK(hipSetDevice(c->e->device));
    // the hit list, allocated on the first execution (or by the strategy timing) and kept
    if (!enum_buffers(c, err, errlen))
        return 0;
    return ts_internal::enum_exec(c, {c->d_out, c->d_list, kListCap}, r, err, errlen);
}

int ts_internal::enum_device(const gpu_enum_code *c)
{
    return c->e->device;
}

int ts_internal::enum_exec(gpu_enum_code *c, const HitBuffers &b, ts_enum_result *r, char *err,
                           size_t errlen)
{
    EnumExec x{c, b, r, err, errlen, {}};
    return x.run();
}

extern "C" int gpu_enum_run(gpu_engine *e, const gpu_enum *en, ts_enum_result *r, char *err,
                            size_t errlen)
{
    memset(r, 0, sizeof *r);
    gpu_enum_code *c = nullptr;
    if (!gpu_enum_compile(e, en, &c, err, errlen))
        return 0;
    int ok = gpu_enum_exec(c, r, err, errlen);
    gpu_enum_free(c);
    return ok;
}
