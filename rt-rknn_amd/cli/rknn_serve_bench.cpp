// rknn-serve-bench -- open-loop load generator for the serving loop
// (rk_serve_*): Poisson arrivals at each offered rate, requests (dataset, q,
// k) with the dataset uniform over the resident ones, q from that dataset's
// seeded query pool and k uniform in 1..kmax (or fixed), one JSON line per
// rate with achieved queries/s and p50/p99 latency.  Latency is the server's
// own submission-to-completion time.
//
//   rknn-serve-bench [-c config] [-u users] [-f facilities] [-s seconds]
//                    [-k kmax] [--fixed-k] [--qps r1,r2,...] [--max-batch n]
//                    [--wait seconds] [--pool n]
//                    [--datasets users:facilities,users:facilities,...]
//
// --datasets keeps several datasets of the config's generator resident at
// once (e.g. 1000000:10000,10000000:100000,50000000:200000); without it the
// one dataset is -u users and -f facilities.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "rknn.h"
#include "rknn_io.h"

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int die(const char* what, rk_ctx* c, rk_status s)
{
    std::fprintf(stderr, "%s: %s (%s)\n", what, rk_status_string(s), c ? rk_last_error(c) : "");
    return 1;
}

} // namespace

int main(int argc, char** argv)
{
    int config = 4, kmax = 50, pool = 4096;
    uint64_t nu = 10000000, nf = 100000;
    double seconds = 1.0, wait = 0.0005;
    bool fixed_k = false;
    uint32_t max_batch = 1024;
    std::vector<double> rates = {50000, 100000, 200000};
    std::vector<std::pair<uint64_t, uint64_t>> sizes;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value after %s\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-c") config = std::atoi(next());
        else if (a == "-u") nu = std::strtoull(next(), nullptr, 10);
        else if (a == "-f") nf = std::strtoull(next(), nullptr, 10);
        else if (a == "-s") seconds = std::atof(next());
        else if (a == "-k") kmax = std::atoi(next());
        else if (a == "--fixed-k") fixed_k = true;
        else if (a == "--max-batch") max_batch = (uint32_t)std::atoi(next());
        else if (a == "--wait") wait = std::atof(next());
        else if (a == "--pool") pool = std::atoi(next());
        else if (a == "--qps") {
            rates.clear();
            std::string v = next();
            for (size_t p = 0; p < v.size();) {
                size_t e = v.find(',', p);
                rates.push_back(std::atof(v.substr(p, e - p).c_str()));
                p = e == std::string::npos ? v.size() : e + 1;
            }
        } else if (a == "--datasets") {
            std::string v = next();
            for (size_t p = 0; p < v.size();) {
                size_t e = v.find(',', p);
                const std::string item = v.substr(p, e - p);
                const size_t c = item.find(':');
                if (c == std::string::npos) {
                    std::fprintf(stderr, "--datasets takes users:facilities pairs\n");
                    return 2;
                }
                sizes.emplace_back(std::strtoull(item.substr(0, c).c_str(), nullptr, 10),
                                   std::strtoull(item.substr(c + 1).c_str(), nullptr, 10));
                p = e == std::string::npos ? v.size() : e + 1;
            }
        } else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    if (sizes.empty()) sizes.emplace_back(nu, nf);
    for (const auto& z : sizes)
        if (z.second < 2 || z.first < 1) {
            std::fprintf(stderr, "need >= 2 facilities and >= 1 user per dataset\n");
            return 2;
        }
    if (kmax < 1 || pool < 1) {
        std::fprintf(stderr, "need k >= 1, pool >= 1\n");
        return 2;
    }
    const uint64_t seed = rk_default_seed(config);
    rk_ctx* ctx = nullptr;
    rk_status st = rk_init(0, &ctx);
    if (st != RK_OK) return die("rk_init", ctx, st);
    struct Resident {
        rk_dataset* ds;
        std::vector<uint32_t> qs;
    };
    std::vector<Resident> res;
    uint64_t max_nu = 0;
    for (const auto& z : sizes) {
        const uint64_t n_u = z.first, n_f = z.second;
        const int n_pool = (int)std::min<uint64_t>((uint64_t)pool, n_f);
        std::vector<float> ux(n_u), uy(n_u), fx(n_f), fy(n_f);
        std::vector<int32_t> uid(n_u), fid(n_f);
        Resident r{nullptr, std::vector<uint32_t>(n_pool)};
        st = rk_generate(config, n_u, n_f, seed, n_pool, seed, ux.data(), uy.data(), uid.data(), fx.data(), fy.data(),
                         fid.data());
        if (st != RK_OK) return die("rk_generate", nullptr, st);
        st = rk_generate_queries(n_f, n_pool, seed, r.qs.data());
        if (st != RK_OK) return die("rk_generate_queries", nullptr, st);
        st = rk_dataset_create(ctx, ux.data(), uy.data(), uid.data(), n_u, fx.data(), fy.data(), fid.data(), n_f,
                               &r.ds);
        if (st != RK_OK) return die("rk_dataset_create", ctx, st);
        res.push_back(std::move(r));
        max_nu = std::max(max_nu, n_u);
        std::fprintf(stderr, "serve bench: config %d, %llu users, %llu facilities resident, k %s %d\n", config,
                     (unsigned long long)n_u, (unsigned long long)n_f, fixed_k ? "=" : "in 1..", kmax);
    }

    std::vector<int32_t> out(max_nu + 64);
    // one server for the whole sweep, as a serving process runs: started and
    // warmed once, then every rate in turn (each rate's requests all complete
    // before the next rate starts)
    rk_serve_config cfg{max_batch, 0, wait};
    rk_server* srv = nullptr;
    if ((st = rk_serve_start(ctx, &cfg, &srv)) != RK_OK) return die("rk_serve_start", ctx, st);
    {
        // untimed warm-up: four rounds of a full step's worth per dataset, all
        // in flight together, grow the server's buffers to their working size
        for (int round = 0; round < 4; ++round) {
            std::vector<uint64_t> wt;
            for (const Resident& r : res)
                for (size_t i = 0; i < std::min<size_t>(max_batch, r.qs.size()); ++i) {
                    uint64_t t = 0;
                    const uint32_t k = fixed_k ? (uint32_t)kmax : 1u + (uint32_t)((i + round) % kmax);
                    if (rk_serve_submit(srv, r.ds, r.qs[(i * 7 + round) % r.qs.size()], k, &t) == RK_OK)
                        wt.push_back(t);
                }
            for (uint64_t t : wt) {
                size_t n = 0;
                while (rk_serve_poll(srv, t, out.data(), out.size(), &n, nullptr) == RK_ERR_NOT_READY)
                    std::this_thread::yield();
            }
        }
    }
    for (double rate : rates) {
        std::mt19937_64 rng(seed ^ (uint64_t)rate);
        std::exponential_distribution<double> gap(rate);
        std::uniform_int_distribution<int> kd(1, kmax);
        std::uniform_int_distribution<size_t> which(0, res.size() - 1);
        const int n_req = std::max(1, (int)(rate * seconds));
        std::vector<double> t_sub(n_req);
        std::vector<uint64_t> tickets(n_req);
        std::atomic<int> submitted{0};
        bool failed = false;
        const double t0 = now_s() + 0.01;
        std::thread submitter([&] {
            double t = t0;
            for (int i = 0; i < n_req; ++i) {
                t += gap(rng);
                while (now_s() < t) {
                }
                const Resident& r = res[which(rng)];
                const uint32_t q = r.qs[std::uniform_int_distribution<size_t>(0, r.qs.size() - 1)(rng)];
                const uint32_t k = fixed_k ? (uint32_t)kmax : (uint32_t)kd(rng);
                t_sub[i] = now_s();
                if (rk_serve_submit(srv, r.ds, q, k, &tickets[i]) != RK_OK) {
                    failed = true;
                    tickets[i] = 0;
                }
                submitted.store(i + 1, std::memory_order_release);
            }
        });
        std::vector<double> lat;
        lat.reserve(n_req);
        double last_done = t0;
        for (int i = 0; i < n_req; ++i) {
            while (submitted.load(std::memory_order_acquire) <= i) std::this_thread::yield();
            if (!tickets[i]) continue;
            while (true) {
                size_t n = 0;
                double l = 0.0;
                const rk_status ps = rk_serve_poll(srv, tickets[i], out.data(), out.size(), &n, &l);
                if (ps == RK_ERR_NOT_READY) {
                    std::this_thread::yield();
                    continue;
                }
                if (ps != RK_OK) failed = true;
                lat.push_back(l);
                last_done = std::max(last_done, t_sub[i] + l);
                break;
            }
        }
        submitter.join();
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double p) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, (size_t)(p * lat.size()))]; };
        const double span = last_done - t0;
        std::printf("{\"offered_qps\": %.1f, \"achieved_qps\": %.1f, \"completed\": %zu, \"p50_ms\": %.4f, "
                    "\"p99_ms\": %.4f, \"max_ms\": %.4f, \"datasets\": %zu, \"errors\": %s}\n",
                    rate, span > 0 ? lat.size() / span : 0.0, lat.size(), pct(0.50) * 1e3, pct(0.99) * 1e3,
                    lat.empty() ? 0.0 : lat.back() * 1e3, res.size(), failed ? "true" : "false");
        std::fflush(stdout);
    }
    rk_serve_stop(srv); // before the datasets it served go
    for (Resident& r : res) rk_dataset_destroy(r.ds);
    rk_shutdown(ctx);
    return 0;
}
