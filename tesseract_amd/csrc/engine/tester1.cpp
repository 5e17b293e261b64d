// Tester 1 on the device (reference src/tester.c:166-279): the random programs of each machine
// size, run as one batch of the program-as-data interpreter (one 512-machine group per program),
// summed per size in the reference's program order.
#include <algorithm>
#include <chrono>
#include <thread>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "engine_internal.hpp"

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

// the programs of one round, freed with it
struct Round {
    std::vector<ts_program> progs;
    std::vector<uint64_t> seeds;
    std::vector<int> size_of;
    ~Round()
    {
        for (ts_program &p : progs)
            ts_program_free(&p);
    }
};

constexpr uint64_t kMostPerSize = 4096;   // programs per size in one batch (programs > 0)
constexpr uint64_t kSampleMost = 1024;    // programs of a size in its residency sample
constexpr uint32_t kSamplePasses = 32;    // passes at most per program of the sample
constexpr uint32_t kItemRow = 16 * 4;     // bytes of one cell of a one-group item (16 words)

// the 1-bit ops per second of the first `most` programs of size index si in r, run as one batch
// with every machine in LDS (residency 1) or HBM (2), each at most kSamplePasses passes
int sample_rate(gpu_engine *e, const Round &r, int si, unsigned residency, uint64_t most,
                uint32_t max_passes, double *rate, uint64_t *programs, char *err, size_t errlen)
{
    std::vector<gpu_batch_item> items;
    for (size_t j = 0; j < r.progs.size() && items.size() < most; j++)
        if (r.size_of[j] == si) {
            gpu_batch_item it{};
            it.prog = &r.progs[j];
            it.seed = r.seeds[j];
            it.residency = residency;
            items.push_back(it);
        }
    gpu_batch_tune tune{};
    tune.groups = 1;
    tune.max_passes = std::min(max_passes, kSamplePasses);
    std::vector<gpu_batch_stats> st(items.size());
    gpu_batch_info info{};
    if (!gpu_batch_run(e, items.data(), items.size(), &tune, st.data(), &info, err, errlen))
        return 0;
    *rate = info.ms > 0 ? (double)info.ops / (info.ms * 1e-3) : 0;
    *programs = items.size();
    return 1;
}

}  // namespace

extern "C" int gpu_t1_choices(const gpu_engine *e, gpu_t1_choice *out, int max)
{
    const int n = (int)std::min<size_t>(e->t1_choices.size(), (size_t)std::max(0, max));
    for (int k = 0; k < n; k++)
        out[k] = e->t1_choices[k];
    return n;
}

extern "C" int gpu_t1_run(gpu_engine *e, const ts_t1_config *cfg, ts_t1_size *out, double *seconds,
                          char *err, size_t errlen)
{
    if (!cfg || !out) {
        say(err, errlen, "%s", "no configuration or no output");
        return 0;
    }
    if (cfg->nsizes < 1 || cfg->nsizes > TS_T1_SIZES) {
        say(err, errlen, "between 1 and %d machine sizes, not %d", TS_T1_SIZES, cfg->nsizes);
        return 0;
    }
    for (int si = 0; si < cfg->nsizes; si++)
        if (cfg->cells[si] < 1 || cfg->cells[si] > 100000) {
            say(err, errlen, "%s", "machine sizes are 1 to 100000 cells");
            return 0;
        }
    if (!(cfg->ratio > 0) || cfg->max_passes == 0 || (cfg->programs == 0 && !(cfg->seconds > 0))) {
        say(err, errlen, "%s", "--ratio, --max-passes and --programs or --seconds must be positive");
        return 0;
    }
    if (cfg->max_passes > 0xFFFFFFFFull) {
        say(err, errlen, "%s", "the device cycle search runs at most 2^32 - 1 passes");
        return 0;
    }
    std::memset(out, 0, (size_t)cfg->nsizes * sizeof *out);
    for (int si = 0; si < cfg->nsizes; si++) {
        out[si].cells = cfg->cells[si];
        out[si].len = ts_t1_program_len(cfg, cfg->cells[si]);
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    gpu_batch_tune tune{};
    tune.groups = 1;
    tune.max_passes = (unsigned)cfg->max_passes;
    // a round in seconds mode: two programs for every wave slot (16 per CU), sizes in turn
    const uint64_t slots = (uint64_t)std::max(1, e->cus) * 16;
    const uint64_t fill = std::max<uint64_t>(64, (2 * slots + cfg->nsizes - 1) / cfg->nsizes);
    // the largest machine whose one-group items fit two to a CU's LDS (the sink cell included)
    const uint64_t lds_cells = e->lds_per_block / (2 * kItemRow) - 1;
    std::vector<unsigned> residency((size_t)cfg->nsizes, 2);
    e->t1_choices.clear();
    bool stopped = false;
    // rounds of programs k0 .. k0 + per_size - 1 of every size: the reference's order, a round at
    // a time (the summaries add up in that order whatever the rounds)
    for (uint64_t k0 = 0; !stopped;) {
        const uint64_t per_size = cfg->programs ? std::min(kMostPerSize, cfg->programs - k0) : fill;
        Round r;
        const size_t n = (size_t)(per_size * (uint64_t)cfg->nsizes);
        r.progs.resize(n);
        r.seeds.resize(n);
        r.size_of.resize(n);
        // sizes take turns (src/tester.c:190-192); the programs are generated on host threads
        const size_t threads = std::min<size_t>(
            {16, std::max(1u, std::thread::hardware_concurrency()), (n + 63) / 64});
        auto make = [&](size_t j) {
            const int si = (int)(j % (size_t)cfg->nsizes);
            const uint64_t k = k0 + j / (size_t)cfg->nsizes;
            const uint32_t cells = cfg->cells[si];
            r.seeds[j] = ts_t1_program_seed(cfg->seed, cells, k);
            ts_random_program(&r.progs[j], cells, out[si].len, r.seeds[j]);
            r.size_of[j] = si;
        };
        std::vector<std::thread> pool;
        for (size_t t = 1; t < threads; t++)
            pool.emplace_back([&, t] {
                for (size_t j = t; j < n; j += threads)
                    make(j);
            });
        for (size_t j = 0; j < n; j += std::max<size_t>(threads, 1))
            make(j);
        for (std::thread &th : pool)
            th.join();
        if (k0 == 0)  // each size's residency, by timing a sample of this round's programs
            for (int si = 0; si < cfg->nsizes; si++) {
                gpu_t1_choice ch{};
                ch.cells = cfg->cells[si];
                ch.residency = 2;
                if (cfg->cells[si] + 2 <= lds_cells) {
                    if (!sample_rate(e, r, si, 1, kSampleMost, tune.max_passes, &ch.rate[0],
                                     &ch.sample_programs, err, errlen) ||
                        !sample_rate(e, r, si, 2, kSampleMost, tune.max_passes, &ch.rate[1],
                                     &ch.sample_programs, err, errlen))
                        return 0;
                    ch.residency = ch.rate[0] >= ch.rate[1] ? 1 : 2;
                }
                residency[(size_t)si] = ch.residency;
                e->t1_choices.push_back(ch);
            }
        std::vector<gpu_batch_item> items(n);
        for (size_t j = 0; j < n; j++) {
            items[j] = gpu_batch_item{};
            items[j].prog = &r.progs[j];
            items[j].seed = r.seeds[j];
            items[j].residency = residency[(size_t)r.size_of[j]];
        }
        tune.seconds = cfg->programs ? 0.0 : std::max(1e-3, cfg->seconds - elapsed());
        std::vector<gpu_batch_stats> st(n);
        if (!gpu_batch_run(e, items.data(), n, &tune, st.data(), nullptr, err, errlen))
            return 0;
        for (size_t j = 0; j < n; j++) {
            ts_t1_size x;
            std::memset(&x, 0, sizeof x);
            if (st[j].unfinished) {  // stopped: its work counts, its cycle lengths do not
                x.passes = st[j].passes;
                x.unfinished = 1;
                stopped = true;
            } else if (st[j].passes == 0) {  // never started before the stop
                stopped = true;
                continue;
            } else {
                x.programs = 1;
                x.starts = 512;
                x.passes = st[j].passes;
                std::memcpy(x.hist, st[j].hist, sizeof x.hist);
                x.longest = st[j].longest;
                x.longest_seed = r.seeds[j];
            }
            ts_t1_add_size(&out[r.size_of[j]], &x);
        }
        k0 += per_size;
        if (cfg->programs ? k0 >= cfg->programs : elapsed() >= cfg->seconds)
            break;
    }
    if (seconds)
        *seconds = elapsed();
    return 1;
}
