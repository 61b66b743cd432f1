// host_check.cpp -- drives the host-side C/C++ code under the sanitizers
// (AddressSanitizer + UndefinedBehaviorSanitizer, or ThreadSanitizer): the
// seeded generators and the text format (rk_io.cpp, including malformed
// files), the host influence-zone filter on one and on several threads
// (rk_filter.cpp), the thread gang of the device-to-host copies
// (rk_gang.hpp), and the oracle (oracle.c, oracle_tree.c) on its worker
// threads.  Checks results as well: filter_many equals filter, the
// candidate lists are sufficient (the oracle over them equals the oracle
// over every facility), the tree restatement equals the brute force.
// Exit status 0 = every check held and the sanitizer reported nothing.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"
#include "rk_filter.hpp"
#include "rk_gang.hpp"
#include "rknn_io.h"

namespace {

int failures = 0;

void check(bool ok, const char* what, long a = 0, long b = 0)
{
    if (!ok) {
        std::fprintf(stderr, "FAIL: %s (%ld, %ld)\n", what, a, b);
        ++failures;
    }
}

struct Data {
    std::vector<float> ux, uy, fx, fy;
    std::vector<int32_t> uid, fid;
};

Data generate(int config, uint64_t nu, uint64_t nf, uint64_t ties)
{
    Data d;
    d.ux.resize(nu);
    d.uy.resize(nu);
    d.uid.resize(nu);
    d.fx.resize(nf);
    d.fy.resize(nf);
    d.fid.resize(nf);
    const uint64_t seed = rk_default_seed(config);
    rk_status s = rk_generate(config, nu, nf, seed, ties, seed, d.ux.data(), d.uy.data(), d.uid.data(),
                              d.fx.data(), d.fy.data(), d.fid.data());
    check(s == RK_OK, "rk_generate", config, s);
    return d;
}

void io_round_trip(const Data& d, const std::string& dir)
{
    const std::string path = dir + "/host_check_ds.txt";
    check(rk_write_text(path.c_str(), d.fx.data(), d.fy.data(), d.fid.data(), d.fx.size(), d.ux.data(),
                        d.uy.data(), d.uid.data(), d.ux.size()) == RK_OK,
          "rk_write_text");
    float *fx, *fy, *ux, *uy;
    int32_t *fid, *uid;
    uint64_t nf, nu;
    check(rk_load_text(path.c_str(), &fx, &fy, &fid, &nf, &ux, &uy, &uid, &nu) == RK_OK, "rk_load_text");
    check(nf == d.fx.size() && nu == d.ux.size(), "loaded sizes", (long)nf, (long)nu);
    check(!std::memcmp(fx, d.fx.data(), nf * 4) && !std::memcmp(fy, d.fy.data(), nf * 4) &&
              !std::memcmp(fid, d.fid.data(), nf * 4) && !std::memcmp(ux, d.ux.data(), nu * 4) &&
              !std::memcmp(uy, d.uy.data(), nu * 4) && !std::memcmp(uid, d.uid.data(), nu * 4),
          "loaded values");
    for (void* p : {(void*)fx, (void*)fy, (void*)fid, (void*)ux, (void*)uy, (void*)uid}) rk_free(p);
    std::remove(path.c_str());

    // malformed files: every one is an error, none reads out of bounds or
    // allocates what its header claims
    const char* bad[] = {"", "3\n1 0 0\n2 1 1\n", "1\n1 0\n0\n", "1\n1 0 x\n0\n", "2147483648\n1 0 0\n",
                         "1\n1 0 0\n99999999999999999999\n", "-1\n", "1\n1 nan 0\n0\n", "1\n1 0 0\n1\n5 1e39 0\n"};
    for (const char* text : bad) {
        const std::string p = dir + "/host_check_bad.txt";
        FILE* f = std::fopen(p.c_str(), "wb");
        std::fwrite(text, 1, std::strlen(text), f);
        std::fclose(f);
        check(rk_load_text(p.c_str(), &fx, &fy, &fid, &nf, &ux, &uy, &uid, &nu) == RK_ERR_IO, "malformed file");
        std::remove(p.c_str());
    }
    check(rk_load_text((dir + "/no_such_file").c_str(), &fx, &fy, &fid, &nf, &ux, &uy, &uid, &nu) == RK_ERR_IO,
          "missing file");
}

void filter_and_oracle(const Data& d, unsigned threads)
{
    const uint64_t nu = d.ux.size(), nf = d.fx.size();
    double box[4] = {1e300, 1e300, -1e300, -1e300};
    for (uint64_t i = 0; i < nu; ++i) {
        box[0] = std::min(box[0], (double)d.ux[i]);
        box[1] = std::min(box[1], (double)d.uy[i]);
        box[2] = std::max(box[2], (double)d.ux[i]);
        box[3] = std::max(box[3], (double)d.uy[i]);
    }
    rk::FacilityIndex idx(d.fx.data(), d.fy.data(), nf, box);
    std::vector<uint32_t> qs(std::min<uint64_t>(24, nf)), ks(qs.size());
    check(rk_generate_queries(nf, qs.size(), 7, qs.data()) == RK_OK, "rk_generate_queries");
    const uint32_t kk[] = {1, 2, 5, 10, 25, 50};
    for (size_t i = 0; i < ks.size(); ++i) ks[i] = kk[i % 6];
    std::vector<std::vector<uint32_t>> many;
    rk::filter_many(idx, qs.data(), ks.data(), qs.size(), many, threads);
    oracle_set_threads((int)threads);
    oracle_tree* tree = oracle_tree_build(d.ux.data(), d.uy.data(), d.uid.data(), nu);
    std::vector<int32_t> a(nu), b(nu), c(nu);
    std::vector<uint64_t> off(qs.size() + 1, 0);
    std::vector<uint32_t> flat;
    for (size_t i = 0; i < qs.size(); ++i) {
        std::vector<uint32_t> one;
        rk::FilterStats st;
        idx.filter(qs[i], ks[i], one, &st);
        check(one == many[i], "filter_many == filter", (long)i);
        const int64_t na = oracle_rknn(d.ux.data(), d.uy.data(), d.uid.data(), nu, d.fx.data(), d.fy.data(), nf,
                                       qs[i], ks[i], one.data(), one.size(), a.data());
        const int64_t nb = oracle_rknn(d.ux.data(), d.uy.data(), d.uid.data(), nu, d.fx.data(), d.fy.data(), nf,
                                       qs[i], ks[i], nullptr, 0, b.data());
        const int64_t nc = oracle_tree_rknn(tree, d.fx.data(), d.fy.data(), nf, qs[i], ks[i], one.data(),
                                            one.size(), c.data());
        check(na == nb && !std::memcmp(a.data(), b.data(), na * 4), "candidates sufficient", (long)na, (long)nb);
        check(nc == na && !std::memcmp(c.data(), a.data(), na * 4), "tree == brute force", (long)nc, (long)na);
        flat.insert(flat.end(), one.begin(), one.end());
        off[i + 1] = flat.size();
    }
    std::vector<int64_t> counts(qs.size());
    check(oracle_tree_rknn_many(tree, d.fx.data(), d.fy.data(), nf, qs.size(), qs.data(), ks.data(), off.data(),
                                flat.data(), counts.data()) == 0,
          "oracle_tree_rknn_many");
    for (size_t i = 0; i < qs.size(); ++i) {
        const int64_t n = oracle_tree_rknn(tree, d.fx.data(), d.fy.data(), nf, qs[i], ks[i], nullptr, 0, nullptr);
        check(counts[i] == n, "tree many == tree one", (long)counts[i], (long)n);
    }
    oracle_tree_free(tree);
    // argument checks
    check(oracle_rknn(d.ux.data(), d.uy.data(), d.uid.data(), nu, d.fx.data(), d.fy.data(), nf, (uint32_t)nf, 1,
                      nullptr, 0, a.data()) == -1,
          "q out of range");
    check(oracle_rknn(d.ux.data(), d.uy.data(), d.uid.data(), nu, d.fx.data(), d.fy.data(), nf, 0, 0, nullptr, 0,
                      a.data()) == -1,
          "k = 0");
}

void gang(unsigned threads)
{
    rk::Gang g(threads);
    std::vector<unsigned char> src(3'000'001), dst(src.size());
    for (size_t i = 0; i < src.size(); ++i) src[i] = (unsigned char)(i * 131 + 7);
    for (int round = 0; round < 50; ++round) {
        const size_t len = src.size() - (size_t)round * 997;
        std::fill(dst.begin(), dst.end(), 0);
        const unsigned t = g.size();
        g.run([&](unsigned k) {
            const size_t lo = len * k / t, hi = len * (k + 1) / t;
            std::memcpy(dst.data() + lo, src.data() + lo, hi - lo);
        });
        check(!std::memcmp(dst.data(), src.data(), len), "gang copy", round);
        check(len == dst.size() || dst[len] == 0, "gang copy bound", round);
    }
}

} // namespace

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : "/tmp";
    const unsigned threads = argc > 2 ? (unsigned)std::atoi(argv[2]) : 4;
    for (int config : {1, 2, 3, 4, 5}) {
        Data d = generate(config, config == 1 ? 2000 : 12000, config == 1 ? 200 : 600, config >= 3 ? 24 : 0);
        if (config == 1) io_round_trip(d, dir);
        filter_and_oracle(d, threads);
    }
    // edge sizes: one facility, two facilities at one position, no users
    {
        Data d = generate(2, 500, 2, 0);
        d.fx[1] = d.fx[0];
        d.fy[1] = d.fy[0];
        filter_and_oracle(d, threads);
    }
    gang(threads);
    gang(1);
    std::printf("host_check: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
