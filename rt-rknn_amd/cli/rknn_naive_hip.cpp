// rknn-naive-hip -- the rknn-naive-gpu command line on the gfx950 engine.
//
// Keeps the reference's interface (rknn-naive-gpu/rknn.cpp:262-352): -if/--infile,
// -k (default 4), -q (query facility, an index into the file's facilities,
// default 0), --gpu/-gpu, -al/--algorithm (parsed, unused), and its stdout
// lines.  The query is looked up as the reference's kernel does
// (rknn-naive-gpu/rknn_kernel.cu:135-141): by the id of facility q, the first
// facility carrying that id.  The candidate set is every other facility
// (rk_query_bruteforce), exact against the binary64 oracle.  The search
// always runs on the GPU: --gpu is accepted and changes nothing, and a
// missing device is an error, not the reference's CPU fallback
// (rknn-naive-gpu/rknn.cpp:404-414).  Additional flags only append output:
//   --ids <path>     write the sorted result ids, one per line
//   --device <n>     HIP device index (default 0)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "rknn.h"
#include "rknn_io.h"

namespace {

void usage_and_exit(const char* argv0)
{
    std::cerr << "Usage  : " << argv0 << " [options]\n";
    std::cerr << "Options: --infile | -if <filename>   Specify file for txt intput\n";
    std::cerr << "         --gpu                       Use GPU acceleration for RkNN computation\n";
    std::cerr << "         -k <number>                 Set k value for k-nearest neighbors (default: 4)\n";
    std::cerr << "         -q <number>                 Set query facility index (default: 0)\n";
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
    std::string infile, ids_path;
    long k = 4, q = 0;
    int device = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a(argv[i]);
        if (a == "--help" || a == "-h") {
            usage_and_exit(argv[0]);
        } else if (a == "--infile" || a == "-if") {
            if (i < argc - 1) infile = argv[++i];
            else usage_and_exit(argv[0]);
        } else if (a == "--gpu" || a == "-gpu") {
            std::cout << "GPU acceleration enabled" << std::endl;
        } else if (a == "--algorithm" || a == "-al") {
            if (i < argc - 1) ++i;
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
    // the reference kernel's lookup: the first facility with the query's id
    uint64_t qi = 0;
    while (fid[qi] != fid[q]) ++qi;

    rk_ctx* ctx = nullptr;
    rk_status s = rk_init(device, &ctx);
    if (s != RK_OK) {
        std::cerr << "Caught exception: " << rk_status_string(s) << "\n";
        return 1;
    }
    std::cout << "Using GPU: HIP device " << device << std::endl;
    rk_dataset* ds = nullptr;
    s = rk_dataset_create(ctx, ux, uy, uid, nu, fx, fy, fid, nf, &ds);
    if (s != RK_OK) {
        std::cerr << "Caught exception: " << rk_status_string(s) << ": " << rk_last_error(ctx) << "\n";
        return 1;
    }
    std::vector<int32_t> ids(nu ? nu : 1);
    size_t n_out = 0;
    rk_timing tm{};
    const double t0 = wall();
    s = rk_query_bruteforce(ds, (uint32_t)qi, (uint32_t)k, ids.data(), ids.size(), &n_out, &tm);
    const double t1 = wall();
    if (s != RK_OK) {
        std::cerr << "Caught exception: " << rk_status_string(s) << ": " << rk_last_error(ctx) << "\n";
        return 1;
    }
    std::cout << "RkNN search (GPU) completed in " << std::fixed << std::setprecision(6) << (t1 - t0) << "[s].\n";
    std::cout << "Found " << n_out << " RkNN results\n";
    // appended line
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
