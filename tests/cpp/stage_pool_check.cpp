// Exercises gss::StagePool (gpusharesat_amd/csrc/stage_pool.h) on the host:
// every part of every job runs exactly once, jobs never overlap, wait()
// returns only after every part ended, an early return through StageJoin
// waits, and destruction joins sleeping and spinning workers.  Built with
// -fsanitize=thread by tests/test_stage_pool.py.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <thread>
#include <vector>

#include "stage_pool.h"

static int fail(const char *what, long a, long b) {
    std::fprintf(stderr, "FAIL %s: %ld vs %ld\n", what, a, b);
    return 1;
}

static int early_return(gss::StagePool &pool, std::vector<int> &out, int k) {
    gss::StageJoin join;
    pool.start([&](int p) {
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        out[p] = k;
    });
    join.pool = &pool;
    if (k % 2) return 1;  // leaves while the parts still run
    pool.wait();
    return 0;
}

int main() {
    for (int workers : {1, 2, 3, 4, 7}) {
        gss::StagePool pool(workers);
        if (pool.size() != workers) return fail("size", pool.size(), workers);
        std::vector<long> count(workers, 0);
        std::atomic<int> inside{0};
        int overlap = 0;
        for (int job = 0; job < 2000; job++) {
            std::vector<int> seen(workers, -1);
            pool.run([&](int p) {
                if (inside.fetch_add(1) >= workers) overlap = 1;
                seen[p] = job;
                count[p]++;
                inside.fetch_sub(1);
            });
            for (int p = 0; p < workers; p++)
                if (seen[p] != job) return fail("part not run", p, job);
            if (job % 500 == 499)  // let the workers fall asleep
                std::this_thread::sleep_for(std::chrono::milliseconds(2));
        }
        if (overlap) return fail("jobs overlapped", overlap, 0);
        for (int p = 0; p < workers; p++)
            if (count[p] != 2000) return fail("runs per part", count[p], 2000);
        for (int k = 0; k < 200; k++) {
            std::vector<int> out(workers, -1);
            early_return(pool, out, k);
            for (int p = 0; p < workers; p++)
                if (out[p] != k) return fail("early return did not wait", out[p], k);
        }
        // a started job with main-thread work in between
        std::vector<long> sums(workers, 0);
        pool.start([&](int p) {
            long s = 0;
            for (long i = p; i < 1000000; i += workers) s += i;
            sums[p] = s;
        });
        long local = 0;
        for (long i = 0; i < 1000; i++) local += i;
        pool.wait();
        long tot = 0;
        for (long s : sums) tot += s;
        if (tot != 1000000L * 999999L / 2) return fail("sum", tot, 1000000L * 999999L / 2);
        (void)local;
    }
    {  // destroyed while its workers sleep
        gss::StagePool pool(3);
        std::this_thread::sleep_for(std::chrono::milliseconds(5));
    }
    {  // destroyed right after a job, workers still spinning
        gss::StagePool pool(3);
        pool.run([](int) {});
    }
    std::puts("ok");
    return 0;
}
