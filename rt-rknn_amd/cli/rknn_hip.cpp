// rknn-hip -- command-line front end of the gfx950 RkNN engine.
//
// Keeps the reference's interface: -if/--infile, -k (default 4), -q (facility
// index, default 0), -al/--algorithm (parsed, unused) as in
// rknn-inf/inf.cpp:38-119, and prints the stdout layout of rknn-inf-gpu
// (rknn-inf-gpu/inf.cpp:164-658).  Additional flags only append output:
//   --bruteforce     candidate set = every facility but q (rknn-naive-gpu form)
//   --ids <path>     write the sorted result ids, one per line
//   --device <n>     HIP device index (default 0)
// Errors print a message and exit 1, as the reference does; there is no CPU
// fallback (rknn-inf-gpu/inf.cpp:659-670 has one; this engine reports "no
// device" instead).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>
#include <chrono>

#include "rknn.h"
#include "rknn_io.h"

namespace {

void usage_and_exit(const char* argv0)
{
    std::cerr << "Usage  : " << argv0 << " [options]\n";
    std::cerr << "Options: --infile | -if <filename>   Specify file for txt intput\n";
    std::cerr << "         -k <k>                      k of the RkNN query (default 4)\n";
    std::cerr << "         -q <index>                  query facility index (default 0)\n";
    std::cerr << "         --bruteforce                use every facility as a candidate\n";
    std::cerr << "         --ids <filename>            write the sorted result ids\n";
    std::cerr << "         --device <n>                HIP device index\n";
    std::cerr << "         --help | -h                 Print this usage message\n";
    std::exit(1);
}

double wall()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

long parse_int(const char* s, const char* argv0)
{
    char* e = nullptr;
    long v = std::strtol(s, &e, 10);
    if (!e || *e != '\0') usage_and_exit(argv0);
    return v;
}

} // namespace

int main(int argc, char* argv[])
{
    std::string infile, ids_path, algorithm = "all";
    long k = 4, q = 0;
    int device = 0;
    bool brute = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a(argv[i]);
        if (a == "--help" || a == "-h") {
            usage_and_exit(argv[0]);
        } else if (a == "--infile" || a == "-if") {
            if (i < argc - 1) infile = argv[++i];
            else usage_and_exit(argv[0]);
        } else if (a == "--algorithm" || a == "-al") {
            if (i < argc - 1) algorithm = argv[++i];
        } else if (a == "-k") {
            if (i < argc - 1) {
                k = parse_int(argv[++i], argv[0]);
                if (k <= 0) {
                    std::cerr << "Invalid value for k: " << k << ". It must be a positive integer." << std::endl;
                    return 1;
                }
            } else {
                std::cerr << "Using default value k = 4." << std::endl;
            }
        } else if (a == "-q") {
            if (i < argc - 1) {
                q = parse_int(argv[++i], argv[0]);
                if (q < 0) {
                    std::cerr << "Invalid value for q: " << q << ". It must be a >= 0 number." << std::endl;
                    usage_and_exit(argv[0]);
                }
            } else {
                std::cerr << "Using default value q = 0." << std::endl;
            }
        } else if (a == "--bruteforce") {
            brute = true;
        } else if (a == "--ids") {
            if (i < argc - 1) ids_path = argv[++i];
            else usage_and_exit(argv[0]);
        } else if (a == "--device") {
            if (i < argc - 1) device = (int)parse_int(argv[++i], argv[0]);
            else usage_and_exit(argv[0]);
        } else {
            std::cerr << "Unknown option '" << a << "'\n";
            usage_and_exit(argv[0]);
        }
    }

    std::cout << "Loading data..." << std::endl;
    float *fx = nullptr, *fy = nullptr, *ux = nullptr, *uy = nullptr;
    int32_t *fid = nullptr, *uid = nullptr;
    uint64_t nf = 0, nu = 0;
    if (rk_load_text(infile.c_str(), &fx, &fy, &fid, &nf, &ux, &uy, &uid, &nu) != RK_OK) {
        std::cerr << "Error opening file: " << infile << std::endl;
        return 1;
    }
    std::cout << "Loaded " << nf << " facilities and " << nu << " users." << std::endl;
    if ((uint64_t)q >= nf) {
        std::cerr << "Invalid value for q: " << q << ". It must be < " << nf << "." << std::endl;
        return 1;
    }

    rk_ctx* ctx = nullptr;
    rk_status s = rk_init(device, &ctx);
    if (s != RK_OK) {
        std::cerr << "Caught exception: " << rk_status_string(s) << "\n";
        return 1;
    }
    std::cout << "building R*-trees..." << std::endl;
    double t0 = wall();
    rk_dataset* ds = nullptr;
    s = rk_dataset_create(ctx, ux, uy, uid, nu, fx, fy, fid, nf, &ds);
    double t1 = wall();
    if (s != RK_OK) {
        std::cerr << "Caught exception: " << rk_status_string(s) << ": " << rk_last_error(ctx) << "\n";
        return 1;
    }
    std::cout << "R*-tree is built in " << t1 - t0 << "[s]." << std::endl << std::endl;

    std::cout << "Filtering..." << std::endl;
    t0 = wall();
    std::vector<uint32_t> cand;
    size_t nc = 0;
    if (brute) {
        for (uint64_t f = 0; f < nf; ++f)
            if (f != (uint64_t)q) cand.push_back((uint32_t)f);
        nc = cand.size();
    } else {
        s = rk_filter(ds, (uint32_t)q, (uint32_t)k, nullptr, 0, &nc);
        if (s == RK_ERR_CAPACITY || s == RK_OK) {
            cand.resize(nc);
            s = rk_filter(ds, (uint32_t)q, (uint32_t)k, cand.data(), cand.size(), &nc);
        }
        if (s != RK_OK) {
            std::cerr << "Caught exception: " << rk_status_string(s) << ": " << rk_last_error(ctx) << "\n";
            return 1;
        }
    }
    t1 = wall();
    const double filtering_time = t1 - t0;
    std::cout << "Filtering time: " << filtering_time << "[s]." << std::endl;
    std::cout << "Total bisectors created: " << nc << std::endl << std::endl;

    std::cout << "Verifying..." << std::endl;
    std::cout << "Using GPU acceleration for verification" << std::endl;
    t0 = wall();
    std::vector<int32_t> ids(nu ? nu : 1);
    size_t n_out = 0;
    rk_timing tm{};
    s = rk_query_plain(ds, (uint32_t)q, (uint32_t)k, cand.data(), nc, ids.data(), ids.size(), &n_out, &tm);
    t1 = wall();
    if (s != RK_OK) {
        std::cerr << "Caught exception: " << rk_status_string(s) << ": " << rk_last_error(ctx) << "\n";
        return 1;
    }
    const double verification_time = t1 - t0;
    std::cout << "Found " << n_out << " RkNN results." << std::endl;
    std::cout << "Verification time (GPU): " << verification_time << "[s]." << std::endl << std::endl;
    std::cout << "Total time: " << filtering_time + verification_time << "[s]." << std::endl << std::endl;
    // appended lines
    std::cout << "Kernel time: " << tm.kernel_s << "[s]. FP64 re-checks: " << tm.n_fp64_rechecks << std::endl;
    if (!ids_path.empty()) {
        FILE* f = std::fopen(ids_path.c_str(), "w");
        if (!f) {
            std::cerr << "Error opening file: " << ids_path << std::endl;
            return 1;
        }
        for (size_t i = 0; i < n_out; ++i) std::fprintf(f, "%d\n", ids[i]);
        std::fclose(f);
    }
    rk_dataset_destroy(ds);
    rk_shutdown(ctx);
    rk_free(fx);
    rk_free(fy);
    rk_free(fid);
    rk_free(ux);
    rk_free(uy);
    rk_free(uid);
    return 0;
}
