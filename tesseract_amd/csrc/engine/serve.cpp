// Serving: verify / equiv / search requests arriving over time, with continuous batching.
//
// A submitted request is copied (programs and spec), code-generated and compiled through hipRTC on a
// pool of host threads, then admitted to the device when its footprint (code object, launch
// buffers) fits the memory budget and an admission slot is free. Every admitted enumeration is cut
// into chunks of words; a scheduler thread keeps `streams` device slots busy, each running one
// chunk at a time, picking the next chunk from the admitted request with the least work left — so
// a request that arrives while a large one is running starts within a chunk's time, and chunks of
// different requests run side by side on the device. A chunk's grid takes a 1/streams share of the
// device's resident waves. Results are exact and independent of the mix: words are enumerated in
// increasing x, a chunk's hits all precede the next chunk's, so hit counts add and the 16 smallest
// come from the earliest chunks (a chunk whose hit list overflowed is listed again by halves).
// Compiled modules stay resident, keyed by their source, for requests that repeat a circuit.
//
// Two tiers: hipRTC takes 0.3-3 s for a kernel of 500-35,000 AND nodes, while the tape
// interpreter (tape_kernel.hip) starts at once but runs about 100 times slower than a compiled
// kernel. An equiv/search request whose estimated tape time is at most tape_max_ms runs on the
// tape; one of up to 0.3 s of tape starts there too, in chunks of about 2 ms, while its kernel
// compiles, and its remaining chunks run compiled once the module is in (both tiers list the same
// hits for the same words, so which tier ran a chunk changes nothing in the result). A verify's
// spec is bit-blasted into the tape's graph (a spec with a right shift has no tape form). Larger
// requests wait for their compile. Chunks are cut as
// they are launched, each to the size its tier wants. Tape buffers are recycled through a pool,
// never freed while the server runs (hipFree would wait for the whole device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine_internal.hpp"

using namespace ts_internal;

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void put_err(char *err, size_t errlen, const char *fmt, ...)
{
    if (!err || !errlen)
        return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, errlen, fmt, ap);
    va_end(ap);
}

enum { ST_QUEUED = 0, ST_COMPILING = 1, ST_WAITING = 2, ST_RUNNING = 3, ST_DONE = 4, ST_FAILED = 5 };

struct Chunk {
    uint64_t i0, i1;
    uint64_t hits = 0;
    std::vector<uint64_t> list;  // (word, mask) pairs
    bool overflow = false, done = false;
};

struct Module {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    size_t bytes = 0;
    int users = 0;
    double last_use = 0;
    unsigned blocks_per_cu = 1;
    unsigned group = 1;  // words a thread of the kernel takes per step
    unsigned vgprs = 0, lds_bytes = 0;  // per thread, per workgroup (function attributes)
    uint64_t scratch_per_thread = 0;    // private memory of a kernel that spills
};

// AND nodes per second of the tape interpreter with a file of about 230 slots (two workgroups of
// four waves per CU): measured 5.0e13 on MI355X for a pair of 3,366 nodes; the estimate errs on
// the slow side.
constexpr double kTapeAndPerSecond = 4.0e13;
// default estimated tape time below which a request worth a compile also starts on the tape
// meanwhile: about a compile's time, so that the tape may finish first; a larger tape share is
// mostly spent work (a compiled kernel does the same words about 80 times faster) that other
// requests wait for
constexpr double kTapeRaceMs = 300;

struct TapeBuffer {
    uint32_t *ptr = nullptr;
    size_t bytes = 0;
};

struct Request {
    uint64_t id = 0;
    bool tape = false;      // tier: tape interpreter instead of a compiled module
    bool upgrade = false;   // on the tape while its kernel compiles
    bool upgrade_done = false, upgrade_ok = false;  // the compile finished (under the lock)
    uint64_t tape_nodes = 0;
    ts::Tape host_tape;     // until admission
    TapeDev td{};
    TapeBuffer tbuf{};
    unsigned tape_bpc = 1;  // resident tape workgroups per CU
    gpu_enum en{};
    ts_program a{}, b{};
    bool has_b = false;
    std::string spec;
    std::string source, error;
    std::vector<char> code;
    gpu_tune tune{};
    gpu_tune upgrade_tune{};
    unsigned group = 1, upgrade_group = 1;  // words per thread step of the compiled kernel
    ShareInfo si{};
    Module *mod = nullptr;
    std::vector<Chunk> chunks;  // cut as launched, in increasing words
    size_t chunks_done = 0;
    uint64_t pos = 0;           // the next word to hand out
    uint64_t words_left = 0;
    ts_enum_result result{};
    int state = ST_QUEUED;
    gpu_request_times times{};
    uint64_t len_total = 0;
    bool in_compile = false;       // a compile thread reads the request's programs
    bool release_wanted = false;   // released while in_compile: dropped when the compile ends
    uint64_t held_bytes = 0;       // device bytes counted for it while admitted
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint64_t *d_out = nullptr, *d_list = nullptr;
    uint64_t *h_out = nullptr;  // pinned
    Request *req = nullptr;
    size_t chunk = 0;
};

}  // namespace

struct gpu_server {
    gpu_engine *e = nullptr;
    gpu_serve_tune tune{};
    std::mutex mu;
    std::condition_variable cv_plan, cv_compile, cv_sched, cv_done;
    std::deque<Request *> to_plan, to_compile, ready;
    unsigned preparing = 0;  // requests taken by a planner and not yet ready or failed
    std::vector<Request *> active;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> all;
    std::unordered_map<std::string, Module> modules;
    std::vector<Slot> slots;
    Slot relist_slot;  // synchronous re-listing of overflowed chunks, off the dispatch slots
    std::vector<std::thread> planners, compilers;
    std::thread scheduler;
    std::atomic<bool> stop{false};
    uint64_t next_id = 1;
    size_t module_bytes = 0;
    uint64_t held_bytes = 0;  // tapes and scratch of the admitted requests (codes: module_bytes)
    std::string fatal;
    std::vector<TapeBuffer> tape_pool;  // free device buffers for tapes
    std::vector<TapeBuffer> tape_all;   // every tape buffer, freed at close
};

namespace {

// the end of a request's preparation, under the lock: on to admission, or failed
void prepared(gpu_server *s, Request *r, bool ok, const char *err)
{
    s->preparing--;
    s->cv_compile.notify_all();  // compile threads leave once nothing is left to prepare
    r->times.compiled_s = now_s();
    r->times.tier = r->tape ? 1 : 0;
    if (!ok) {
        r->error = err;
        r->state = ST_FAILED;
        r->times.done_s = r->times.compiled_s;
        s->cv_done.notify_all();
    } else {
        r->state = ST_WAITING;
        s->ready.push_back(r);
        s->cv_sched.notify_all();
    }
}

// Frees a finished request (under the lock): its copied programs and its entry.
void drop_request(gpu_server *s, Request *r)
{
    auto it = s->all.find(r->id);
    if (it == s->all.end())
        return;
    ts_program_free(&r->a);
    if (r->has_b)
        ts_program_free(&r->b);
    s->all.erase(it);
}

// Planner threads pick each request's tier as it arrives: the tape when its estimated device
// time is small (building it takes milliseconds, so these requests never wait behind a compile),
// otherwise the compile queue. A refusal (e.g. too many slots) leaves the request to the compiler,
// which reports validation errors with the same messages.
void plan_loop(gpu_server *s)
{
    for (;;) {
        Request *r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_plan.wait(lk, [&] { return s->stop || !s->to_plan.empty(); });
            if (s->stop && s->to_plan.empty())
                return;
            r = s->to_plan.front();
            s->to_plan.pop_front();
            r->state = ST_COMPILING;
            s->preparing++;
        }
        if (s->tune.tape_max_ms >= 0 && r->en.ticks == 1) {
            char err[1024] = "";
            ts::Tape tp;
            gpu_tune tn{};
            if (enum_tape(s->e, &r->en, tn, tape_max_slots(s->e), tp, err, sizeof err)) {
                ShareInfo si = shares((unsigned)r->en.a->ninputs, r->en.part, r->en.part_end,
                                      r->en.parts);
                double ms = (double)si.tried * (double)tp.and_nodes / kTapeAndPerSecond * 1e3;
                // up to tape_max_ms: the tape alone; up to tape_race_ms: the tape while the kernel
                // compiles (it may finish first); beyond, the compile alone
                if (ms <= std::max(s->tune.tape_max_ms, s->tune.tape_race_ms)) {
                    r->tape_nodes = tp.and_nodes;
                    r->host_tape = std::move(tp);
                    r->tune = tn;
                    r->tape = true;
                    r->upgrade = ms > s->tune.tape_max_ms;
                }
            }
        }
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->tape)
            prepared(s, r, true, "");
        if (!r->tape || r->upgrade) {
            s->to_compile.push_back(r);
            s->cv_compile.notify_one();
        }
    }
}

void compile_loop(gpu_server *s)
{
    for (;;) {
        Request *r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_compile.wait(lk, [&] {
                return !s->to_compile.empty() ||
                       (s->stop && s->to_plan.empty() && s->preparing == 0);
            });
            if (s->to_compile.empty())
                return;
            r = s->to_compile.front();
            s->to_compile.pop_front();
            r->in_compile = true;
        }
        char err[1024] = "";
        ts::KernelSource ks;
        gpu_tune tn = r->upgrade ? gpu_tune{} : r->tune;
        bool ok = enum_prepare(s->e, &r->en, tn, ks, err, sizeof err) != 0;
        std::string source;
        std::vector<char> code;
        if (ok) {
            source = ks.text;
            bool cached;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                cached = s->modules.count(source) != 0;
            }
            if (!cached)
                ok = rtc_compile(source, s->e->arch, code, nullptr, err, sizeof err, kEnumSched) != 0;
        }
        std::lock_guard<std::mutex> lk(s->mu);
        r->in_compile = false;
        if (r->release_wanted) {  // finished on the tape and released meanwhile
            drop_request(s, r);
            continue;
        }
        if (r->upgrade) {  // running on the tape: the scheduler switches it over if still running
            if (r->state == ST_RUNNING || r->state == ST_WAITING) {
                r->source = std::move(source);
                r->code = std::move(code);
                r->upgrade_tune = tn;
                r->upgrade_group = ks.group;
                r->upgrade_ok = ok;
                r->upgrade_done = true;
                s->cv_sched.notify_all();
            }
            continue;
        }
        r->source = std::move(source);
        r->code = std::move(code);
        r->tune = tn;
        r->group = ks.group;
        prepared(s, r, ok, err);
    }
}

// a pooled device buffer of at least `bytes` (sizes are powers of two from 64 KiB)
bool take_tape_buffer(gpu_server *s, size_t bytes, TapeBuffer &out)
{
    size_t best = SIZE_MAX;
    for (size_t k = 0; k < s->tape_pool.size(); k++)
        if (s->tape_pool[k].bytes >= bytes &&
            (best == SIZE_MAX || s->tape_pool[k].bytes < s->tape_pool[best].bytes))
            best = k;
    if (best != SIZE_MAX) {
        out = s->tape_pool[best];
        s->tape_pool.erase(s->tape_pool.begin() + (long)best);
        return true;
    }
    size_t cap = 64 * 1024;
    while (cap < bytes)
        cap *= 2;
    if (hipMalloc(&out.ptr, cap) != hipSuccess)
        return false;
    out.bytes = cap;
    s->tape_all.push_back(out);
    return true;
}

// device side of a tape request: its tape in a pooled buffer
bool admit_tape(gpu_server *s, Request *r)
{
    if (!take_tape_buffer(s, tape_bytes(r->host_tape), r->tbuf)) {
        r->error = "no device memory for the tape";
        return false;
    }
    if (!tape_upload(r->host_tape, r->td, r->tbuf.ptr) ||
        !tape_prepare(s->e->device, r->td, &r->tape_bpc)) {
        s->tape_pool.push_back(r->tbuf);
        r->tbuf = TapeBuffer();
        r->error = "could not place the tape on the device";
        return false;
    }
    r->host_tape = ts::Tape();
    return true;
}

// device side of a compiled request: load (or reuse) its module
bool admit_module(gpu_server *s, Request *r)
{
    Module &m = s->modules[r->source];
    if (!m.mod) {
        if (r->code.empty()) {  // its module was evicted after the compile found it resident
            char err[512];
            if (!rtc_compile(r->source, s->e->arch, r->code, nullptr, err, sizeof err, kEnumSched)) {
                r->error = err;
                s->modules.erase(r->source);
                return false;
            }
        }
        if (hipModuleLoadData(&m.mod, r->code.data()) != hipSuccess ||
            hipModuleGetFunction(&m.fn, m.mod, "ts_enum") != hipSuccess) {
            r->error = "module load failed";
            s->modules.erase(r->source);
            return false;
        }
        m.bytes = r->code.size();
        s->module_bytes += m.bytes;
        int blocks = 1;
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, m.fn, r->tune.block, 0) !=
                hipSuccess ||
            blocks < 1)
            blocks = 1;
        m.blocks_per_cu = (unsigned)blocks;
        m.group = r->group;
        int v = 0;
        if (hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, m.fn) == hipSuccess)
            m.vgprs = (unsigned)v;
        if (hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, m.fn) == hipSuccess)
            m.lds_bytes = (unsigned)v;
        if (hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, m.fn) == hipSuccess)
            m.scratch_per_thread = (uint64_t)v;
    }
    m.users++;
    m.last_use = now_s();
    r->mod = &m;
    r->code.clear();
    r->code.shrink_to_fit();
    return true;
}

// the device bytes a compiled request's chunks hold beyond the code: the scratch of a kernel
// that spills, for the resident threads of its grid share (a 1/streams share of the device)
uint64_t scratch_bytes(const gpu_server *s, const Module &m, unsigned block)
{
    return m.scratch_per_thread * s->e->cus * (uint64_t)m.blocks_per_cu * block /
           std::max<size_t>(1, s->slots.size());
}

// the request's footprint, recorded in its times and counted while it is admitted
void hold(gpu_server *s, Request *r, uint64_t code_bytes)
{
    uint64_t bytes = code_bytes;
    if (r->tape) {
        bytes += r->tbuf.bytes;
        r->times.lds_bytes = (uint32_t)tape_lds_bytes(r->td);
        r->times.waves_per_cu = r->tape_bpc * 4;
    } else if (r->mod) {
        const uint64_t scratch = scratch_bytes(s, *r->mod, r->tune.block);
        bytes += scratch;
        r->held_bytes += scratch;
        s->held_bytes += scratch;
        r->times.vgprs = r->mod->vgprs;
        r->times.lds_bytes = r->mod->lds_bytes;
        r->times.waves_per_cu = r->mod->blocks_per_cu * (r->tune.block / 64);
    }
    r->times.hbm_bytes = bytes;
}

// returns what a finished request held
void unhold(gpu_server *s, Request *r)
{
    s->held_bytes -= std::min(s->held_bytes, r->held_bytes);
    r->held_bytes = 0;
}

// device side of one request, then its share cut into chunks
bool admit(gpu_server *s, Request *r)
{
    const bool loads = !r->tape && !s->modules.count(r->source);
    const uint64_t code_bytes = loads ? r->code.size() : 0;
    if (!(r->tape ? admit_tape(s, r) : admit_module(s, r)))
        return false;
    if (r->tape) {
        r->held_bytes += r->tbuf.bytes;
        s->held_bytes += r->tbuf.bytes;
    }
    hold(s, r, code_bytes);
    r->si = shares((unsigned)r->en.a->ninputs, r->en.part, r->en.part_end, r->en.parts);
    r->result.space = r->si.space;
    r->result.tried = r->si.tried;
    r->result.ops = r->si.passes_run * r->en.ticks * r->len_total * 512;
    r->result.threads = 1;
    r->words_left = r->si.nact_words;
    r->state = ST_RUNNING;
    r->times.admitted_s = now_s();
    return true;
}

// drop modules nobody uses, oldest first, until the budget holds `need` more bytes
void evict(gpu_server *s, size_t need)
{
    while (s->module_bytes + need > s->tune.memory_budget) {
        auto victim = s->modules.end();
        for (auto it = s->modules.begin(); it != s->modules.end(); ++it)
            if (it->second.users == 0 && it->second.mod &&
                (victim == s->modules.end() || it->second.last_use < victim->second.last_use))
                victim = it;
        if (victim == s->modules.end())
            return;
        (void)hipModuleUnload(victim->second.mod);
        s->module_bytes -= victim->second.bytes;
        s->modules.erase(victim);
    }
}

// the next chunk of a request, cut for the tier that will run it: the configured size compiled
// (and on the tape for a request that stays there), about 2 ms of tape while a compile is pending,
// so that a request arriving meanwhile finds a slot within 2 ms (a chunk is not preempted)
size_t cut_chunk(gpu_server *s, Request *r)
{
    uint64_t step = std::max<uint64_t>(1024, s->tune.chunk_words);
    if (r->tape && r->upgrade && r->tape_nodes) {
        const double words = 2e-3 * kTapeAndPerSecond / (32.0 * (double)r->tape_nodes);
        step = std::min<uint64_t>(step, std::max<uint64_t>(1024, (uint64_t)words));
    }
    Chunk c;
    c.i0 = r->pos;
    c.i1 = std::min(r->si.nact_words, r->pos + step);
    r->pos = c.i1;
    r->chunks.push_back(c);
    return r->chunks.size() - 1;
}

// a request on the tape whose kernel has compiled: later chunks run compiled. A failed compile,
// or no room for the module, leaves it on the tape (its result is the same).
void upgrade(gpu_server *s, Request *r)
{
    r->upgrade_done = false;
    if (!r->upgrade_ok || r->pos >= r->si.nact_words)
        return;
    size_t need = s->modules.count(r->source) ? 0 : r->code.size();
    evict(s, need);
    if (need && s->module_bytes + s->held_bytes + need > s->tune.memory_budget)
        return;
    const uint64_t code_bytes = need;
    gpu_tune keep = r->tune;
    r->tune = r->upgrade_tune;
    r->group = r->upgrade_group;
    std::string err = r->error;
    if (!admit_module(s, r)) {
        r->tune = keep;
        r->error = err;
        return;
    }
    r->tape = false;
    r->upgrade = false;
    r->times.tier = 2;
    hold(s, r, code_bytes);
    r->times.hbm_bytes += r->tbuf.bytes;  // the tape's buffer stays with it until it finishes
}

bool launch_chunk(gpu_server *s, Slot &sl, Request *r, size_t ci)
{
    Chunk &c = r->chunks[ci];
    const unsigned block = r->tape ? 64 : r->tune.block;
    const unsigned per_cu = r->tape ? r->tape_bpc : r->mod->blocks_per_cu;
    const uint64_t group = r->tape ? 1 : r->mod->group;
    uint64_t blocks = ((c.i1 + group - 1) / group - c.i0 / group + block - 1) / block;
    uint64_t share = std::max<uint64_t>(1, (uint64_t)s->e->cus * per_cu / s->slots.size());
    unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(blocks, share));
    uint64_t space = r->si.space, cap = kListCap, i0 = c.i0, i1 = c.i1;
    uint32_t parts = (uint32_t)r->en.parts, part = (uint32_t)r->en.part, nshare = r->si.nshare;
    void *args[] = {&sl.d_out, &sl.d_list, &cap, &i0, &i1, &space, &parts, &part, &nshare};
    if (hipMemsetAsync(sl.d_out, 0, 16, sl.stream) != hipSuccess)
        return false;
    if (r->tape) {
        if (!tape_launch(r->td, sl.stream, grid, sl.d_out, sl.d_list, cap, i0, i1, space, parts,
                         part, nshare))
            return false;
    } else if (hipModuleLaunchKernel(r->mod->fn, grid, 1, 1, block, 1, 1, 0, sl.stream, args,
                                     nullptr) != hipSuccess) {
        return false;
    }
    if (hipMemcpyAsync(sl.h_out, sl.d_out, 16, hipMemcpyDeviceToHost, sl.stream) != hipSuccess ||
        hipEventRecord(sl.done, sl.stream) != hipSuccess)
        return false;
    sl.req = r;
    sl.chunk = ci;
    if (r->times.started_s == 0)
        r->times.started_s = now_s();
    r->times.launches++;
    return true;
}

// a finished chunk: counts from the pinned copy, then its hit list (a short synchronous copy)
bool finish_chunk(gpu_server *s, Slot &sl)
{
    (void)s;
    Request *r = sl.req;
    Chunk &c = r->chunks[sl.chunk];
    c.hits = sl.h_out[0];
    uint64_t nlist = sl.h_out[1];
    if (nlist > kListCap) {
        c.overflow = true;
    } else if (nlist) {
        c.list.resize(2 * nlist);
        if (hipMemcpyAsync(c.list.data(), sl.d_list, 16 * nlist, hipMemcpyDeviceToHost, sl.stream) !=
                hipSuccess ||
            hipStreamSynchronize(sl.stream) != hipSuccess)
            return false;
    }
    c.done = true;
    r->chunks_done++;
    r->words_left -= c.i1 - c.i0;
    sl.req = nullptr;
    return true;
}

// the 16 smallest hits of an overflowed range, by halves, on one slot's stream (synchronous)
bool relist(gpu_server *s, Slot &sl, Request *r, uint64_t i0, uint64_t i1, ts_enum_result &res)
{
    if (i0 >= i1 || res.nhit >= TS_EN_SHOW)
        return true;
    Chunk tmp;
    tmp.i0 = i0;
    tmp.i1 = i1;
    r->chunks.push_back(tmp);
    size_t ci = r->chunks.size() - 1;
    if (!launch_chunk(s, sl, r, ci) || hipEventSynchronize(sl.done) != hipSuccess)
        return false;
    Chunk &c = r->chunks[ci];
    c.hits = sl.h_out[0];
    uint64_t nlist = sl.h_out[1];
    sl.req = nullptr;
    if (nlist > kListCap) {
        uint64_t mid = i0 + (i1 - i0) / 2;
        return relist(s, sl, r, i0, mid, res) && relist(s, sl, r, mid, i1, res);
    }
    std::vector<uint64_t> list(2 * nlist);
    if (nlist && (hipMemcpy(list.data(), sl.d_list, 16 * nlist, hipMemcpyDeviceToHost) != hipSuccess))
        return false;
    for (size_t k = 0; k + 1 < list.size(); k += 2)
        for (uint32_t m = (uint32_t)list[k + 1]; m; m &= m - 1)
            ts_enum_add_hit(&res, list[k] * 32 + (uint64_t)__builtin_ctz(m));
    return true;
}

bool finalize(gpu_server *s, Request *r)
{
    ts_enum_result &res = r->result;
    const size_t n = r->chunks.size();
    for (size_t k = 0; k < n; k++)
        res.hits += r->chunks[k].hits;
    for (size_t k = 0; k < n && res.nhit < TS_EN_SHOW; k++) {
        Chunk &c = r->chunks[k];
        if (c.overflow) {
            if (!relist(s, s->relist_slot, r, c.i0, c.i1, res))
                return false;
            continue;
        }
        for (size_t q = 0; q + 1 < c.list.size(); q += 2)
            for (uint32_t m = (uint32_t)c.list[q + 1]; m; m &= m - 1)
                ts_enum_add_hit(&res, c.list[q] * 32 + (uint64_t)__builtin_ctz(m));
    }
    r->chunks.clear();
    r->chunks.shrink_to_fit();
    return true;
}

void schedule_loop(gpu_server *s)
{
    if (hipSetDevice(s->e->device) != hipSuccess) {
        s->fatal = "hipSetDevice failed";
        return;
    }
    for (;;) {
        std::unique_lock<std::mutex> lk(s->mu);
        bool busy = false;
        // admission: budget and active-request limit
        while (!s->ready.empty() && s->active.size() < s->tune.max_active) {
            Request *r = s->ready.front();
            // its footprint: a code object not yet resident, a tape, a spilling kernel's scratch
            // (known once its module was loaded for an earlier request)
            size_t need = r->tape ? tape_bytes(r->host_tape)
                                  : s->modules.count(r->source) ? 0 : r->code.size();
            auto m = s->modules.find(r->source);
            if (!r->tape && m != s->modules.end())
                need += scratch_bytes(s, m->second, r->tune.block);
            evict(s, need);
            if (s->module_bytes + s->held_bytes + need > s->tune.memory_budget && !s->active.empty())
                break;  // wait for memory to free up
            s->ready.pop_front();
            if (!admit(s, r)) {
                r->state = ST_FAILED;
                r->times.done_s = now_s();
                s->cv_done.notify_all();
                continue;
            }
            s->active.push_back(r);
            busy = true;
        }
        // completions
        for (Slot &sl : s->slots) {
            if (!sl.req)
                continue;
            hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady)
                continue;
            if (q != hipSuccess || !finish_chunk(s, sl)) {
                s->fatal = "a chunk failed on the device";
                s->stop = true;
                s->cv_done.notify_all();
                return;
            }
            busy = true;
        }
        // finished requests
        for (size_t k = 0; k < s->active.size();) {
            Request *r = s->active[k];
            if (r->chunks_done == r->chunks.size() && r->pos == r->si.nact_words) {
                bool ok = finalize(s, r);
                unhold(s, r);
                if (r->mod)
                    r->mod->users--;
                if (r->tbuf.ptr) {  // every chunk and re-listing of this tape has completed
                    s->tape_pool.push_back(r->tbuf);
                    r->tbuf = TapeBuffer();
                    r->td = TapeDev();
                }
                r->state = ok ? ST_DONE : ST_FAILED;
                if (!ok)
                    r->error = "listing hits failed";
                r->times.done_s = now_s();
                s->active.erase(s->active.begin() + (long)k);
                s->cv_done.notify_all();
                busy = true;
            } else {
                k++;
            }
        }
        // compiles that finished for requests running on the tape
        for (Request *r : s->active)
            if (r->upgrade_done) {
                upgrade(s, r);
                busy = true;
            }
        // dispatch: each idle slot takes the next chunk of the request with the least work left;
        // requests on the tape while their kernel compiles hold at most half the slots, the rest
        // staying for arrivals
        unsigned speculative = 0;
        for (Slot &sl : s->slots)
            speculative += sl.req && sl.req->tape && sl.req->upgrade;
        for (Slot &sl : s->slots) {
            if (sl.req)
                continue;
            Request *best = nullptr;
            const bool room = 2 * (speculative + 1) <= s->slots.size();
            for (Request *r : s->active)
                if (r->pos < r->si.nact_words && (room || !(r->tape && r->upgrade)) &&
                    (!best || r->words_left < best->words_left))
                    best = r;
            if (!best)
                break;
            speculative += best->tape && best->upgrade;
            if (!launch_chunk(s, sl, best, cut_chunk(s, best))) {
                s->fatal = "launch failed";
                s->stop = true;
                s->cv_done.notify_all();
                return;
            }
            busy = true;
        }
        bool in_flight = false;
        for (Slot &sl : s->slots)
            in_flight |= sl.req != nullptr;
        if (s->stop && !in_flight && s->ready.empty() && s->to_plan.empty() && s->preparing == 0 &&
            s->active.empty())
            return;
        if (!busy) {
            if (in_flight)
                s->cv_sched.wait_for(lk, std::chrono::microseconds(50));
            else
                s->cv_sched.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

}  // namespace

extern "C" int gpu_server_open(gpu_engine *e, const gpu_serve_tune *tune, gpu_server **out,
                               char *err, size_t errlen)
{
    *out = nullptr;
    std::unique_ptr<gpu_server> s(new gpu_server());
    s->e = e;
    if (tune)
        s->tune = *tune;
    if (!s->tune.streams)
        s->tune.streams = 4;
    if (!s->tune.compile_threads)
        s->tune.compile_threads = 8;
    if (!s->tune.memory_budget)
        s->tune.memory_budget = 1ull << 30;
    if (!s->tune.chunk_words)
        s->tune.chunk_words = 1ull << 24;
    if (!s->tune.max_active)
        s->tune.max_active = 64;
    if (s->tune.tape_max_ms == 0)
        s->tune.tape_max_ms = 50;
    if (s->tune.tape_race_ms == 0)
        s->tune.tape_race_ms = kTapeRaceMs;
    if (hipSetDevice(e->device) != hipSuccess) {
        put_err(err, errlen, "hipSetDevice failed");
        return 0;
    }
    s->slots.resize(s->tune.streams);
    std::vector<Slot *> all_slots;
    for (Slot &sl : s->slots)
        all_slots.push_back(&sl);
    all_slots.push_back(&s->relist_slot);
    for (Slot *slp : all_slots) {
        Slot &sl = *slp;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc(&sl.d_out, 16) != hipSuccess ||
            hipMalloc(&sl.d_list, 16 * kListCap) != hipSuccess ||
            hipHostMalloc(&sl.h_out, 16, hipHostMallocDefault) != hipSuccess) {
            put_err(err, errlen, "could not allocate the server's device slots");
            gpu_server_close(s.release());
            return 0;
        }
    }
    gpu_server *raw = s.get();
    for (unsigned k = 0; k < 2; k++)
        raw->planners.emplace_back(plan_loop, raw);
    for (unsigned k = 0; k < raw->tune.compile_threads; k++)
        raw->compilers.emplace_back(compile_loop, raw);
    raw->scheduler = std::thread(schedule_loop, raw);
    *out = s.release();
    return 1;
}

extern "C" int gpu_server_submit(gpu_server *s, const gpu_enum *en, uint64_t *ticket, char *err,
                                 size_t errlen)
{
    std::unique_ptr<Request> r(new Request());
    if (!ts_program_copy(&r->a, en->a) || (en->b && !ts_program_copy(&r->b, en->b))) {
        put_err(err, errlen, "out of memory copying the request");
        return 0;
    }
    r->has_b = en->b != nullptr;
    r->en = *en;
    r->en.a = &r->a;
    r->en.b = r->has_b ? &r->b : nullptr;
    if (en->spec) {
        r->spec = en->spec;
        r->en.spec = r->spec.c_str();
    }
    r->en.tune = nullptr;
    r->len_total = r->a.len + (r->has_b && en->mode == TS_EN_EQUIV ? r->b.len : 0);
    r->times.submitted_s = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        put_err(err, errlen, "the server is closed: %s", s->fatal.c_str());
        return 0;
    }
    r->id = s->next_id++;
    *ticket = r->id;
    s->to_plan.push_back(r.get());
    s->all[r->id] = std::move(r);
    s->cv_plan.notify_one();
    return 1;
}

extern "C" int gpu_server_poll(gpu_server *s, uint64_t ticket, ts_enum_result *res, int *state,
                               char *err, size_t errlen)
{
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) {
        put_err(err, errlen, "no request %llu", (unsigned long long)ticket);
        return 0;
    }
    Request *r = it->second.get();
    *state = r->state;
    if (r->state == ST_DONE && res)
        *res = r->result;
    if (r->state == ST_FAILED)
        put_err(err, errlen, "%s", r->error.c_str());
    return 1;
}

extern "C" int gpu_server_wait(gpu_server *s, uint64_t ticket, ts_enum_result *res, char *err,
                               size_t errlen)
{
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) {
        put_err(err, errlen, "no request %llu", (unsigned long long)ticket);
        return 0;
    }
    Request *r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->state == ST_DONE || r->state == ST_FAILED || s->stop; });
    if (r->state != ST_DONE) {
        put_err(err, errlen, "%s", r->state == ST_FAILED ? r->error.c_str() : s->fatal.c_str());
        return 0;
    }
    *res = r->result;
    return 1;
}

extern "C" int gpu_server_times(gpu_server *s, uint64_t ticket, gpu_request_times *t)
{
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end())
        return 0;
    *t = it->second->times;
    return 1;
}

extern "C" int gpu_server_release(gpu_server *s, uint64_t ticket, char *err, size_t errlen)
{
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) {
        put_err(err, errlen, "no request %llu", (unsigned long long)ticket);
        return 0;
    }
    Request *r = it->second.get();
    if (r->state != ST_DONE && r->state != ST_FAILED) {
        put_err(err, errlen, "request %llu has not finished", (unsigned long long)ticket);
        return 0;
    }
    // a request that finished on the tape may still wait for, or be in, its kernel's compile
    for (size_t k = 0; k < s->to_compile.size(); k++)
        if (s->to_compile[k] == r) {
            s->to_compile.erase(s->to_compile.begin() + (long)k);
            break;
        }
    if (r->in_compile)
        r->release_wanted = true;
    else
        drop_request(s, r);
    return 1;
}

extern "C" uint64_t gpu_server_live(gpu_server *s)
{
    std::lock_guard<std::mutex> lk(s->mu);
    uint64_t n = 0;
    for (auto &kv : s->all)
        n += !kv.second->release_wanted;
    return n;
}

extern "C" void gpu_server_close(gpu_server *s)
{
    if (!s)
        return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_plan.notify_all();
    s->cv_compile.notify_all();
    s->cv_sched.notify_all();
    for (std::thread &t : s->planners)
        t.join();
    for (std::thread &t : s->compilers)
        t.join();
    if (s->scheduler.joinable())
        s->scheduler.join();
    (void)hipSetDevice(s->e->device);
    s->slots.push_back(s->relist_slot);
    for (Slot &sl : s->slots) {
        if (sl.stream)
            (void)hipStreamSynchronize(sl.stream);
        if (sl.d_out)
            (void)hipFree(sl.d_out);
        if (sl.d_list)
            (void)hipFree(sl.d_list);
        if (sl.h_out)
            (void)hipHostFree(sl.h_out);
        if (sl.done)
            (void)hipEventDestroy(sl.done);
        if (sl.stream)
            (void)hipStreamDestroy(sl.stream);
    }
    for (auto &kv : s->modules)
        if (kv.second.mod)
            (void)hipModuleUnload(kv.second.mod);
    for (TapeBuffer &b : s->tape_all)
        (void)hipFree(b.ptr);
    for (auto &kv : s->all) {
        ts_program_free(&kv.second->a);
        if (kv.second->has_b)
            ts_program_free(&kv.second->b);
    }
    delete s;
}
