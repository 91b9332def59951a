// csx_serve -- serving-loop benchmark of the continuous-batching server.
//
// Requests arrive as a Poisson process (default 20000/s) for a fixed time.
// Each has an element type drawn uniformly from {uint32, float32, uint64}, a
// length log-uniform in [2^a, 2^b] (default [2^12, 2^26]) and a Bernoulli
// mask of selectivity uniform in [0, 1].  A request is admitted when its
// input, mask and output fit in the HBM budget (a csx_arena, FIFO order);
// its inputs are then produced on a generator stream and it is submitted to
// the server, whose batches overlap on a ring of streams.  Reported: request
// latency from arrival to the poll that saw it finish (p50/p99/max) and the
// aggregate effective GB/s (input + mask + selected output bytes) of
// completed requests.  Every k-th request is recomputed alone with
// csx_compact and compared element for element.
//
//   csx_serve [-r rate] [-t seconds] [-B budget_GiB] [-a log2min] [-b log2max]
//             [-S streams] [-m max_batch] [-s seed] [-v verify_every]
#include <csx.h>
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
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

    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(rate / 1000.0);  // per ms
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    std::deque<Spec> arrivals;  // arrived, not yet admitted
    std::unordered_map<uint64_t, Live> live;
    std::vector<double> latencies;
    std::vector<csx_completion> done(4096);
    uint64_t next_id = 0, admitted = 0, completed = 0, verified = 0, mismatches = 0, waits = 0;
    double bytes = 0.0, last_done_ms = 0.0;
    uint64_t peak_used = 0;
    double host_admit = 0, host_step = 0, host_poll = 0, host_check = 0;
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
    // retire finished checks; with wait, block for the oldest busy one
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
    const double t0 = now_ms();
    double next_arrival = t0 + gap(rng);
    const double t_end = t0 + seconds * 1000.0;

    while (true) {
        const double t = now_ms();
        // arrivals up to now
        while (next_arrival <= t && next_arrival < t_end) {
            Spec s{};
            s.id = next_id++;
            s.arrival_ms = next_arrival;
            s.n = (uint64_t)std::exp(std::log((double)(1ull << lmin)) + unif(rng) * (std::log((double)(1ull << lmax)) - std::log((double)(1ull << lmin))));
            const int type = (int)(unif(rng) * 3.0);  // 0 uint32, 1 float32, 2 uint64
            s.w = type == 2 ? 8 : 4;
            s.kind = type == 1 ? CSX_VALUES_FLOAT : CSX_VALUES_INTEGER;
            s.p = unif(rng);
            arrivals.push_back(s);
            next_arrival += gap(rng);
        }
        // admission in arrival order while the budget allows
        const double ta = now_ms();
        ReadyEvent* round = nullptr;
        std::vector<uint64_t> fresh;
        while (!arrivals.empty()) {
            const Spec& s = arrivals.front();
            void *in = nullptr, *mask = nullptr, *out = nullptr;
            bool fits = csx_arena_alloc(arena, s.n * s.w, &in) == CSX_OK;
            if (fits && csx_arena_alloc(arena, (s.n + 7) / 8, &mask) != CSX_OK) {
                csx_arena_free(arena, in);
                fits = false;
            }
            if (fits && csx_arena_alloc(arena, s.n * s.w, &out) != CSX_OK) {
                csx_arena_free(arena, in);
                csx_arena_free(arena, mask);
                fits = false;
            }
            if (!fits) {
                if (live.empty()) die("admission", "a request is larger than the whole budget");
                break;
            }
            ck(csx_gen_values(in, s.n, s.w, s.kind, seed * 1000003ull + s.id, gen), "generate values");
            ck(csx_gen_mask_bernoulli(static_cast<uint8_t*>(mask), s.n, s.p, seed * 7919ull + s.id, CSX_MSB_FIRST, gen), "generate mask");
            if (!round) {
                round = new ReadyEvent{nullptr, 0};
                hk(hipEventCreateWithFlags(&round->ev, hipEventDisableTiming), "event");
            }
            round->refs++;
            live[s.id] = Live{s, in, mask, out, round};
            fresh.push_back(s.id);
            arrivals.pop_front();
            admitted++;
        }
        if (round) {
            hk(hipEventRecord(round->ev, gen), "event record");
            for (uint64_t id : fresh) {
                const Live& L = live[id];
                csx_request r{};
                r.id = L.spec.id;
                r.in = L.in;
                r.mask = static_cast<const uint8_t*>(L.mask);
                r.out = L.out;
                r.n = L.spec.n;
                r.elem_size = L.spec.w;
                r.ready = round->ev;
                ck(csx_server_submit(server, &r), "submit");
            }
        }
        if (!arrivals.empty()) waits++;
        peak_used = std::max(peak_used, csx_arena_used(arena));
        const double ts = now_ms();
        host_admit += ts - ta;
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
        const double tc = now_ms();
        host_check += tc - tp;
        if (tp >= t_end && arrivals.empty() && live.empty()) {
            sweep_checks(true);
            break;
        }
        if (nd == 0 && launched == 0 && arrivals.empty() && next_arrival > tp + 0.05 && tp < t_end) usleep(20);
    }
    const double wall_ms = last_done_ms - t0;
    std::sort(latencies.begin(), latencies.end());
    auto pct = [&](double q) { return latencies.empty() ? 0.0 : latencies[std::min(latencies.size() - 1, (size_t)(q * latencies.size()))]; };
    double mean = 0;
    for (double l : latencies) mean += l;
    mean = latencies.empty() ? 0 : mean / latencies.size();
    printf("{\"requests\": %llu, \"completed\": %llu, \"seconds\": %.3f, \"offered_rate\": %.0f, \"achieved_rate\": %.1f, \"GB/s\": %.1f, "
           "\"p50_ms\": %.3f, \"p99_ms\": %.3f, \"max_ms\": %.3f, \"mean_ms\": %.3f, \"verified\": %llu, \"mismatches\": %llu, "
           "\"peak_budget_GiB\": %.2f, \"budget_GiB\": %.1f, \"admission_waits\": %llu, \"streams\": %d, \"max_batch\": %d, "
           "\"host_ms\": {\"admit\": %.0f, \"step\": %.0f, \"poll\": %.0f, \"check\": %.0f}}\n",
           (unsigned long long)next_id, (unsigned long long)completed, wall_ms / 1000.0, rate, completed / (wall_ms / 1000.0),
           bytes / (wall_ms * 1e-3) / 1e9, pct(0.50), pct(0.99), latencies.empty() ? 0.0 : latencies.back(), mean, (unsigned long long)verified,
           (unsigned long long)mismatches, peak_used / double(1ull << 30), budget_gib, (unsigned long long)waits, streams, max_batch, host_admit,
           host_step, host_poll, host_check);
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
