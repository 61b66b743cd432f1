// rk_tune.cpp -- the autotuner's memory: which kernel variant won for a
// workload class.
//
// Choices are made by timing the variants on the caller's own data the first
// time a class is seen (rk_batch.hip: traversal block shape; rk_api.hip:
// count-all kernel).  They are kept for the process and in a JSON file
// (RK_TUNE_FILE, default $HOME/.cache/rknn_amd/autotune.json), one line per
// class, so later processes skip the timing.  Writes are best effort: a
// read-only home only loses the persistence.
#include <hip/hip_runtime.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "rk_internal.hpp"

namespace rk {

namespace {

struct TuneCache {
    std::mutex mu;
    bool loaded = false;
    std::map<std::string, std::pair<int, std::string>> choice; // class -> (variant, label)
};

TuneCache& cache()
{
    static TuneCache c;
    return c;
}

std::string tune_path()
{
    if (const char* p = std::getenv("RK_TUNE_FILE")) return p;
    const char* h = std::getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/rknn_amd/autotune.json";
}

// lines of the form   "class": {"choice": N, "variant": "label"},
void load(TuneCache& tc)
{
    tc.loaded = true;
    FILE* f = std::fopen(tune_path().c_str(), "r");
    if (!f) return;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        const char* q0 = std::strchr(line, '"');
        const char* q1 = q0 ? std::strchr(q0 + 1, '"') : nullptr;
        const char* ch = q1 ? std::strstr(q1, "\"choice\":") : nullptr;
        if (!ch) continue;
        const char* lb = std::strstr(ch, "\"variant\": \"");
        std::string label;
        if (lb) {
            lb += 12;
            const char* le = std::strchr(lb, '"');
            if (le) label.assign(lb, le);
        }
        tc.choice[std::string(q0 + 1, q1)] = {std::atoi(ch + 9), label};
    }
    std::fclose(f);
}

void save(const TuneCache& tc)
{
    const std::string path = tune_path();
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos) {
        const std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); ++i)
            if (i == dir.size() || dir[i] == '/') mkdir(dir.substr(0, i).c_str(), 0755);
    }
    const std::string tmp = path + ".tmp" + std::to_string(getpid());
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return;
    std::fprintf(f, "{\n");
    size_t i = 0;
    for (const auto& kv : tc.choice)
        std::fprintf(f, "  \"%s\": {\"choice\": %d, \"variant\": \"%s\"}%s\n", kv.first.c_str(), kv.second.first,
                     kv.second.second.c_str(), ++i < tc.choice.size() ? "," : "");
    std::fprintf(f, "}\n");
    std::fclose(f);
    std::rename(tmp.c_str(), path.c_str());
}

} // namespace

std::string device_arch(rk_ctx* c)
{
    if (!c->arch.empty()) return c->arch;
    hipDeviceProp_t prop{};
    std::string arch = hipGetDeviceProperties(&prop, c->device) == hipSuccess ? prop.gcnArchName : "unknown";
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);
    c->arch = arch;
    return arch;
}

int tuned_choice(const std::string& key)
{
    TuneCache& tc = cache();
    std::lock_guard<std::mutex> lk(tc.mu);
    if (!tc.loaded) load(tc);
    auto it = tc.choice.find(key);
    return it == tc.choice.end() ? -1 : it->second.first;
}

void tuned_store(const std::string& key, int choice, const char* label)
{
    TuneCache& tc = cache();
    std::lock_guard<std::mutex> lk(tc.mu);
    if (!tc.loaded) load(tc);
    tc.choice[key] = {choice, label ? label : ""};
    save(tc);
}

} // namespace rk
