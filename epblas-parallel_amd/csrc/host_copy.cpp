// host_copy.cpp -- parallel strided host memcpy for the host-pointer staging pipeline.
//
// The drop-in entry points receive pageable host arrays.  Transfers run at PCIe speed only
// from pinned memory, so operands are copied column-block by column-block into pinned
// chunks (and C back out of them).  Those host copies are split over a small pool of
// worker threads that lives for the whole process.
#include "host_copy.hpp"

#include <algorithm>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#if defined(__linux__)
#include <pthread.h>
#include <sched.h>
#include <strings.h>
#endif

namespace mw {
namespace {

#if defined(__linux__)
// Whether an OpenMP runtime in the process was asked to bind its threads.  The binding
// applies to the process's initial thread too, so a thread that thread creates inherits a
// mask of one place (one core, say), while the CPUs the process may use are many more.
bool omp_binding_requested()
{
    const char *b = std::getenv("OMP_PROC_BIND");
    if (b && *b && strcasecmp(b, "false") != 0) return true;
    return std::getenv("OMP_PLACES") || std::getenv("GOMP_CPU_AFFINITY") ||
           std::getenv("KMP_AFFINITY");
}

// Parses a CPU list such as "0-15,32-47" into set; false if it names no CPU.
bool parse_cpu_list(const char *text, cpu_set_t *set)
{
    CPU_ZERO(set);
    const char *p = text;
    while (*p) {
        char *end = nullptr;
        long a = std::strtol(p, &end, 10);
        if (end == p) break;
        long b = a;
        p = end;
        if (*p == '-') {
            b = std::strtol(p + 1, &end, 10);
            if (end == p + 1) break;
            p = end;
        }
        for (long c = a; c <= b && c >= 0 && c < CPU_SETSIZE; ++c) CPU_SET((int)c, set);
        if (*p == ',') ++p;
        else break;
    }
    return CPU_COUNT(set) > 0;
}

// The CPUs of the process's cgroup cpuset (the allotment of a container or a batch job),
// or every CPU when no cpuset file is readable (the kernel then keeps a thread to its
// cpuset anyway).
void allotted_cpus(cpu_set_t *set)
{
    for (const char *path : {"/sys/fs/cgroup/cpuset.cpus.effective",
                             "/sys/fs/cgroup/cpuset/cpuset.effective_cpus"}) {
        if (FILE *f = std::fopen(path, "r")) {
            char buf[4096] = {0};
            const size_t got = std::fread(buf, 1, sizeof buf - 1, f);
            std::fclose(f);
            buf[got] = 0;
            if (parse_cpu_list(buf, set)) return;
        }
    }
    CPU_ZERO(set);
    for (int c = 0; c < CPU_SETSIZE; ++c) CPU_SET(c, set);
}
#endif

class Pool {
public:
    explicit Pool(int n) : n_(n)
    {
#if defined(__linux__)
        // When an OpenMP binding has narrowed the creating thread to fewer CPUs than the pool
        // has threads, the workers would all share those; they run on the allotted CPUs
        // instead.  An unbound caller's mask (taskset, a launcher) is inherited as is.
        cpu_set_t mine;
        if (omp_binding_requested() &&
            sched_getaffinity(0, sizeof mine, &mine) == 0 && CPU_COUNT(&mine) < n_) {
            widen_ = true;
            allotted_cpus(&wide_);
        }
#endif
        for (int i = 1; i < n_; ++i) th_.emplace_back([this, i] { worker(i); });
    }
    ~Pool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    int size() const { return n_; }
    // Runs f(i) for every i in [0, n); the calling thread takes i = 0.  One job at a time.
    void run(const std::function<void(int)> &f)
    {
        std::lock_guard<std::mutex> job_lk(job_mu_);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            pending_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return pending_ == 0; });
        job_ = nullptr;
    }

private:
    void worker(int i)
    {
#if defined(__linux__)
        if (widen_) (void)pthread_setaffinity_np(pthread_self(), sizeof wide_, &wide_);
#endif
        unsigned seen = 0;
        for (;;) {
            const std::function<void(int)> *f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                f = job_;
            }
            (*f)(i);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--pending_ == 0) done_cv_.notify_one();
            }
        }
    }
    int n_;
#if defined(__linux__)
    bool widen_ = false;
    cpu_set_t wide_;
#endif
    std::vector<std::thread> th_;
    std::mutex mu_, job_mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)> *job_ = nullptr;
    int pending_ = 0;
    unsigned gen_ = 0;
    bool stop_ = false;
};

int pool_threads()
{
    if (const char *e = std::getenv("MWGEMM_HOST_THREADS")) {
        const int v = std::atoi(e);
        if (v >= 1 && v <= 256) return v;
    }
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
}

Pool &pool()
{
    static Pool p(pool_threads());
    return p;
}

}  // namespace

void host_copy_columns(void *dst, size_t dpitch, const void *src, size_t spitch, size_t seg,
                       int64_t cols)
{
    if (cols <= 0 || seg == 0) return;
    char *d = static_cast<char *>(dst);
    const char *s = static_cast<const char *>(src);
    const size_t total = seg * (size_t)cols;
    // Below a few MiB the pool's wake-up costs more than it saves.
    if (total < (size_t(4) << 20) || pool().size() == 1) {
        for (int64_t c = 0; c < cols; ++c) std::memcpy(d + c * dpitch, s + c * spitch, seg);
        return;
    }
    // Split the linearised byte range [0, total) into equal contiguous pieces; a piece may
    // start and end inside a column.
    const int n = pool().size();
    pool().run([&](int i) {
        const size_t lo = total * (size_t)i / (size_t)n, hi = total * (size_t)(i + 1) / (size_t)n;
        size_t pos = lo;
        while (pos < hi) {
            const size_t c = pos / seg, off = pos % seg;
            const size_t len = std::min(seg - off, hi - pos);
            std::memcpy(d + c * dpitch + off, s + c * spitch + off, len);
            pos += len;
        }
    });
}

bool host_has_avx2_fma()
{
#if defined(__x86_64__)
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
#else
    return false;
#endif
}

}  // namespace mw
