// csx_serve -- serving-loop benchmark of the continuous-batching server.
//
// Requests arrive as a Poisson process (default 20000/s) for a fixed time.
// Each has an element type drawn uniformly from {uint32, float32, uint64}, a
// length log-uniform in [2^a, 2^b] (default [2^12, 2^26]) and a Bernoulli
// mask of selectivity uniform in [0, 1].
//
// Two host threads.  The client thread plays the clients: it draws the
// arrivals, admits them in arrival order when input, mask and output fit in
// the HBM budget (a csx_arena), produces their inputs on the device on its
// own stream and hands each admission round (with one ready event) to the
// server thread.  The server thread drives the csx_server: it submits the
// handed-over requests, steps (batches launch on a ring of streams) and
// polls, and recomputes every k-th finished request alone with csx_compact,
// comparing element for element on the device, asynchronously.  Reported:
// request latency from arrival to the poll that saw it finish (p50/p99/max)
// and the aggregate effective GB/s (input + mask + selected output bytes) of
// completed requests.
//
//   csx_serve [-r rate] [-t seconds] [-B budget_GiB] [-a log2min] [-b log2max]
//             [-S streams] [-m max_batch] [-s seed] [-v verify_every]
#include <csx.h>
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <random>
#include <unordered_map>
#include <vector>

namespace {

[[noreturn]] void die(const char* what, const char* detail)
{
    fprintf(stderr, "csx_serve: %s: %s\n", what, detail);
    exit(EXIT_FAILURE);
}

void ck(csx_status s, const char* what)
{
    if (s != CSX_OK) die(what, csx_last_error());
}

void hk(hipError_t e, const char* what)
{
    if (e != hipSuccess) die(what, hipGetErrorString(e));
}

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Spec {
    uint64_t id;
    double arrival_ms;
    uint64_t n;
    uint32_t w;
    int kind;  // CSX_VALUES_INTEGER / CSX_VALUES_FLOAT
    double p;
};

struct ReadyEvent {
    hipEvent_t ev;
    int refs;
};

struct Live {
    Spec spec;
    void *in, *mask, *out;
    ReadyEvent* ready;
};

// One asynchronous check of a finished request: csx_compact of the same
// request alone into v_out, then csx_validate against the served output.
// The request's buffers stay allocated until the check's event has fired.
struct Check {
    void* v_out = nullptr;
    void* ws = nullptr;
    uint64_t* d_res = nullptr;  // [0] count alone, [1] differing elements
    uint64_t* h_res = nullptr;  // pinned copy of d_res
    hipEvent_t done = nullptr;
    bool busy = false;
    Live req{};
    uint64_t served_count = 0;
};

} // namespace

int main(int argc, char** argv)
{
    double rate = 20000, seconds = 60, budget_gib = 64;
    int lmin = 12, lmax = 26, streams = 4, max_batch = 512, verify_every = 64;
    uint64_t seed = 7;
    int opt;
    while ((opt = getopt(argc, argv, "r:t:B:a:b:S:m:s:v:")) != -1) {
        switch (opt) {
            case 'r': rate = atof(optarg); break;
            case 't': seconds = atof(optarg); break;
            case 'B': budget_gib = atof(optarg); break;
            case 'a': lmin = atoi(optarg); break;
            case 'b': lmax = atoi(optarg); break;
            case 'S': streams = atoi(optarg); break;
            case 'm': max_batch = atoi(optarg); break;
            case 's': seed = strtoull(optarg, nullptr, 10); break;
            case 'v': verify_every = atoi(optarg); break;
            default: die("usage", "csx_serve [-r rate] [-t seconds] [-B GiB] [-a log2min] [-b log2max] [-S streams] [-m batch] [-s seed] [-v k]");
        }
    }
    if (rate <= 0 || seconds <= 0 || lmin < 1 || lmax < lmin || lmax > 30 || streams < 1 || max_batch < 1) die("arguments", "out of range");

    csx_arena* arena = nullptr;
    ck(csx_arena_create((uint64_t)(budget_gib * (1ull << 30)), &arena), "arena");
    csx_server_config cfg{};
    cfg.streams = (uint32_t)streams;
    cfg.max_batch = (uint32_t)max_batch;
    cfg.max_batch_elements = 1ull << 32;
    csx_server* server = nullptr;
    ck(csx_server_create(&cfg, &server), "server");
    hipStream_t gen, ver;
    hk(hipStreamCreate(&gen), "stream");
    hk(hipStreamCreate(&ver), "stream");
    // verification contexts, each sized for the largest request alone
    const uint64_t nmax = 1ull << lmax;
    size_t v_ws_bytes = 0;
    ck(csx_workspace_size(nmax, 8, 1024, &v_ws_bytes), "workspace size");
    std::vector<Check> checks(4);
    for (Check& c : checks) {
        hk(hipMalloc(&c.v_out, nmax * 8), "hipMalloc");
        hk(hipMalloc(&c.ws, v_ws_bytes), "hipMalloc");
        hk(hipMalloc((void**)&c.d_res, 2 * sizeof(uint64_t)), "hipMalloc");
        hk(hipHostMalloc((void**)&c.h_res, 2 * sizeof(uint64_t), hipHostMallocDefault), "hipHostMalloc");
        hk(hipEventCreateWithFlags(&c.done, hipEventDisableTiming), "event");
    }

    // ---- hand-off from the client thread to the server thread ----
    struct Handoff {
        std::mutex mu;
        std::vector<Live> ready;  // admitted, inputs being produced; arrival order
        bool done = false;        // no more arrivals
    } hand;
    std::atomic<uint64_t> produced{0}, waits{0}, peak_used{0};
    std::atomic<double> host_admit{0};
    const double t0 = now_ms();
    const double t_end = t0 + seconds * 1000.0;

    std::thread client([&] {
        std::mt19937_64 rng(seed);
        std::exponential_distribution<double> gap(rate / 1000.0);  // per ms
        std::uniform_real_distribution<double> unif(0.0, 1.0);
        std::deque<Spec> arrivals;  // arrived, not yet admitted
        double next_arrival = t0 + gap(rng);
        uint64_t next_id = 0;
        double admit_ms = 0;
        while (true) {
            const double t = now_ms();
            while (next_arrival <= t && next_arrival < t_end) {
                Spec sp{};
                sp.id = next_id++;
                sp.arrival_ms = next_arrival;
                sp.n = (uint64_t)std::exp(std::log((double)(1ull << lmin)) +
                                          unif(rng) * (std::log((double)(1ull << lmax)) - std::log((double)(1ull << lmin))));
                const int type = (int)(unif(rng) * 3.0);  // 0 uint32, 1 float32, 2 uint64
                sp.w = type == 2 ? 8 : 4;
                sp.kind = type == 1 ? CSX_VALUES_FLOAT : CSX_VALUES_INTEGER;
                sp.p = unif(rng);
                arrivals.push_back(sp);
                next_arrival += gap(rng);
            }
            // admission in arrival order while the budget allows
            ReadyEvent* round = nullptr;
            std::vector<Live> fresh;
            while (!arrivals.empty()) {
                const Spec& sp = arrivals.front();
                void *in = nullptr, *mask = nullptr, *out = nullptr;
                bool fits = csx_arena_alloc(arena, sp.n * sp.w, &in) == CSX_OK;
                if (fits && csx_arena_alloc(arena, (sp.n + 7) / 8, &mask) != CSX_OK) {
                    csx_arena_free(arena, in);
                    fits = false;
                }
                if (fits && csx_arena_alloc(arena, sp.n * sp.w, &out) != CSX_OK) {
                    csx_arena_free(arena, in);
                    csx_arena_free(arena, mask);
                    fits = false;
                }
                if (!fits) {
                    if (csx_arena_used(arena) == 0) die("admission", "a request is larger than the whole budget");
                    break;
                }
                ck(csx_gen_values(in, sp.n, sp.w, sp.kind, seed * 1000003ull + sp.id, gen), "generate values");
                ck(csx_gen_mask_bernoulli(static_cast<uint8_t*>(mask), sp.n, sp.p, seed * 7919ull + sp.id, CSX_MSB_FIRST, gen), "generate mask");
                if (!round) {
                    round = new ReadyEvent{nullptr, 0};
                    hk(hipEventCreateWithFlags(&round->ev, hipEventDisableTiming), "event");
                }
                round->refs++;
                fresh.push_back(Live{sp, in, mask, out, round});
                arrivals.pop_front();
            }
            if (round) {
                hk(hipEventRecord(round->ev, gen), "event record");
                std::lock_guard<std::mutex> lk(hand.mu);
                for (Live& L : fresh) hand.ready.push_back(L);
            }
            if (!arrivals.empty()) waits++;
            uint64_t used = csx_arena_used(arena), prev = peak_used.load();
            while (used > prev && !peak_used.compare_exchange_weak(prev, used)) {
            }
            admit_ms += now_ms() - t;
            if (t >= t_end && arrivals.empty()) break;
            if (fresh.empty()) {
                // idle until the next arrival (at most 20 us), or 10 us while waiting for budget
                const double wait_us = arrivals.empty() ? std::min(20.0, (next_arrival - now_ms()) * 1000.0) : 10.0;
                if (wait_us >= 2.0) usleep((useconds_t)wait_us);
            }
        }
        produced = next_id;
        host_admit = admit_ms;
        std::lock_guard<std::mutex> lk(hand.mu);
        hand.done = true;
    });

    // ---- the server thread (this one) ----
    std::unordered_map<uint64_t, Live> live;
    std::vector<double> latencies;
    std::vector<csx_completion> done(4096);
    std::vector<Live> incoming;
    uint64_t completed = 0, verified = 0, mismatches = 0;
    double bytes = 0.0, last_done_ms = 0.0;
    double host_step = 0, host_poll = 0, host_check = 0;
    auto release = [&](Live& L) {
        csx_arena_free(arena, L.in);
        csx_arena_free(arena, L.mask);
        csx_arena_free(arena, L.out);
        if (--L.ready->refs == 0) {
            hk(hipEventDestroy(L.ready->ev), "event destroy");
            delete L.ready;
        }
    };
    auto retire = [&](Check& c) {
        verified++;
        if (c.h_res[0] != c.served_count || c.h_res[1] != 0) mismatches++;
        release(c.req);
        c.busy = false;
    };
    // retire finished checks; with wait, block for every busy one
    auto sweep_checks = [&](bool wait) {
        for (Check& c : checks) {
            if (!c.busy) continue;
            if (wait) hk(hipEventSynchronize(c.done), "check wait");
            const hipError_t q = hipEventQuery(c.done);
            if (q == hipSuccess) retire(c);
            else if (q != hipErrorNotReady) hk(q, "check query");
        }
    };
    auto launch_check = [&](Live& L, uint64_t served) {
        Check* c = nullptr;
        for (Check& k : checks)
            if (!k.busy) c = &k;
        if (!c) {
            sweep_checks(true);
            for (Check& k : checks)
                if (!k.busy) c = &k;
        }
        c->busy = true;
        c->req = L;
        c->served_count = served;
        hk(hipMemsetAsync(c->d_res, 0, 2 * sizeof(uint64_t), ver), "memset");
        ck(csx_compact(L.in, static_cast<const uint8_t*>(L.mask), c->v_out, L.spec.n, L.spec.w, 1024, CSX_MSB_FIRST, CSX_PIPELINE_THREE_PHASE,
                       c->d_res, nullptr, c->ws, v_ws_bytes, ver),
           "check compact");
        ck(csx_validate(c->v_out, L.out, served, L.spec.w, c->d_res + 1, ver), "check validate");
        hk(hipMemcpyAsync(c->h_res, c->d_res, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, ver), "check copy");
        hk(hipEventRecord(c->done, ver), "check record");
    };

    while (true) {
        bool client_done;
        {
            std::lock_guard<std::mutex> lk(hand.mu);
            incoming.swap(hand.ready);
            client_done = hand.done;
        }
        for (Live& L : incoming) {
            live[L.spec.id] = L;
            csx_request r{};
            r.id = L.spec.id;
            r.in = L.in;
            r.mask = static_cast<const uint8_t*>(L.mask);
            r.out = L.out;
            r.n = L.spec.n;
            r.elem_size = L.spec.w;
            r.ready = L.ready->ev;
            ck(csx_server_submit(server, &r), "submit");
        }
        const bool got = !incoming.empty();
        incoming.clear();
        const double ts = now_ms();
        uint32_t launched = 0, nd = 0;
        ck(csx_server_step(server, &launched), "step");
        const double tq = now_ms();
        host_step += tq - ts;
        ck(csx_server_poll(server, done.data(), (uint32_t)done.size(), &nd), "poll");
        const double tp = now_ms();
        host_poll += tp - tq;
        for (uint32_t i = 0; i < nd; i++) {
            auto it = live.find(done[i].id);
            if (it == live.end()) die("poll", "unknown request id");
            Live& L = it->second;
            latencies.push_back(tp - L.spec.arrival_ms);
            bytes += (double)L.spec.n * L.spec.w + (double)((L.spec.n + 7) / 8) + (double)done[i].count * L.spec.w;
            completed++;
            last_done_ms = tp;
            if (verify_every > 0 && L.spec.id % (uint64_t)verify_every == 0) launch_check(L, done[i].count);
            else release(L);
            live.erase(it);
        }
        sweep_checks(false);
        host_check += now_ms() - tp;
        if (client_done && live.empty()) {
            std::lock_guard<std::mutex> lk(hand.mu);
            if (hand.ready.empty()) break;
        }
        if (!got && nd == 0 && launched == 0) usleep(10);
    }
    sweep_checks(true);
    client.join();
    const double wall_ms = last_done_ms - t0;
    std::sort(latencies.begin(), latencies.end());
    auto pct = [&](double q) { return latencies.empty() ? 0.0 : latencies[std::min(latencies.size() - 1, (size_t)(q * latencies.size()))]; };
    double mean = 0;
    for (double l : latencies) mean += l;
    mean = latencies.empty() ? 0 : mean / latencies.size();
    printf("{\"requests\": %llu, \"completed\": %llu, \"seconds\": %.3f, \"offered_rate\": %.0f, \"achieved_rate\": %.1f, \"GB/s\": %.1f, "
           "\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"max_ms\": %.3f, \"mean_ms\": %.3f, \"verified\": %llu, \"mismatches\": %llu, "
           "\"peak_budget_GiB\": %.2f, \"budget_GiB\": %.1f, \"admission_waits\": %llu, \"streams\": %d, \"max_batch\": %d, "
           "\"host_ms\": {\"client_admit\": %.0f, \"step\": %.0f, \"poll\": %.0f, \"check\": %.0f}}\n",
           (unsigned long long)produced.load(), (unsigned long long)completed, wall_ms / 1000.0, rate, completed / (wall_ms / 1000.0),
           bytes / (wall_ms * 1e-3) / 1e9, pct(0.50), pct(0.99), latencies.empty() ? 0.0 : latencies.back(), mean, (unsigned long long)verified,
           (unsigned long long)mismatches, peak_used.load() / double(1ull << 30), budget_gib, (unsigned long long)waits.load(), streams, max_batch,
           host_admit.load(), host_step, host_poll, host_check);
    ck(csx_server_destroy(server), "server destroy");
    ck(csx_arena_destroy(arena), "arena destroy");
    for (Check& c : checks) {
        (void)hipFree(c.v_out);
        (void)hipFree(c.ws);
        (void)hipFree(c.d_res);
        (void)hipHostFree(c.h_res);
        (void)hipEventDestroy(c.done);
    }
    (void)hipStreamDestroy(gen);
    (void)hipStreamDestroy(ver);
    return mismatches ? 1 : 0;
}
