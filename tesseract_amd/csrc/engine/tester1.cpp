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

constexpr uint64_t kRoundPerSize = 64;     // programs per size and round when running for a time
constexpr uint64_t kMostPerSize = 4096;    // programs per size in one batch otherwise

}  // namespace

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
    // rounds of programs k0 .. k0 + per_size - 1 of every size: the reference's order, a round at
    // a time (the summaries add up in that order whatever the rounds)
    for (uint64_t k0 = 0;;) {
        const uint64_t per_size = cfg->programs ? std::min(kMostPerSize, cfg->programs - k0)
                                                : kRoundPerSize;
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
        std::vector<gpu_batch_item> items(n);
        for (size_t j = 0; j < n; j++)
            items[j] = gpu_batch_item{&r.progs[j], r.seeds[j], nullptr};
        std::vector<gpu_batch_stats> st(n);
        if (!gpu_batch_run(e, items.data(), n, &tune, st.data(), nullptr, err, errlen))
            return 0;
        for (size_t j = 0; j < n; j++) {
            ts_t1_size x;
            std::memset(&x, 0, sizeof x);
            x.programs = 1;
            x.starts = 512;
            x.passes = st[j].passes;
            std::memcpy(x.hist, st[j].hist, sizeof x.hist);
            x.longest = st[j].longest;
            x.longest_seed = r.seeds[j];
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
