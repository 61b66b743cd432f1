// rk_io.cpp -- reference-format text I/O and seeded synthetic datasets
// (see include/rknn_io.h).
#include "rknn_io.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

namespace {

// SplitMix64 (Steele, Lea, Flood 2014): state += golden gamma, then the
// variant-13 finaliser.
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u53() { return (double)(next() >> 11) * 0x1p-53; }
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
    // Box-Muller, one value per call (no cached pair: keeps streams simple)
    double normal()
    {
        double a = u53();
        const double b = u53();
        if (a < 0x1p-60) a = 0x1p-60;
        return std::sqrt(-2.0 * std::log(a)) * std::cos(6.283185307179586476925286766559 * b);
    }
};

constexpr double kSide = 1000.0;

inline float snap7(double v)
{
    v = std::min(kSide, std::max(0.0, v));
    return (float)(std::nearbyint(v * 128.0) / 128.0);
}

// Clustered, road-like model of config 3.
struct Clustered {
    struct Comp {
        double cx, cy, sx, sy, c, s;
    };
    struct Line {
        std::vector<double> x, y, cum; // vertices and cumulative length
    };
    std::vector<Comp> comp;
    std::vector<Line> line;

    explicit Clustered(uint64_t seed)
    {
        SplitMix r(seed ^ 0xC1A55EEDull);
        comp.resize(64);
        for (Comp& c : comp) {
            c.cx = 50.0 + 900.0 * r.u53();
            c.cy = 50.0 + 900.0 * r.u53();
            c.sx = 0.5 + 19.5 * r.u53();
            c.sy = 0.5 + 19.5 * r.u53();
            const double th = 3.141592653589793238462643383279 * r.u53();
            c.c = std::cos(th);
            c.s = std::sin(th);
        }
        line.resize(2000);
        for (Line& l : line) {
            const int nv = 2 + (int)r.below(5);
            double x = kSide * r.u53(), y = kSide * r.u53();
            l.x.push_back(x);
            l.y.push_back(y);
            l.cum.push_back(0.0);
            for (int v = 1; v < nv; ++v) {
                const double len = 10.0 + 90.0 * r.u53();
                const double a = 6.283185307179586476925286766559 * r.u53();
                const double nx = std::min(kSide, std::max(0.0, x + len * std::cos(a)));
                const double ny = std::min(kSide, std::max(0.0, y + len * std::sin(a)));
                l.cum.push_back(l.cum.back() + std::hypot(nx - x, ny - y));
                l.x.push_back(nx);
                l.y.push_back(ny);
                x = nx;
                y = ny;
            }
        }
    }

    void sample(SplitMix& r, float* ox, float* oy) const
    {
        double x, y;
        if (r.u53() < 0.6) {
            const Comp& c = comp[r.below(comp.size())];
            const double z1 = r.normal() * c.sx, z2 = r.normal() * c.sy;
            x = c.cx + c.c * z1 - c.s * z2;
            y = c.cy + c.s * z1 + c.c * z2;
        } else {
            const Line& l = line[r.below(line.size())];
            const double pos = l.cum.back() * r.u53();
            size_t i = 1;
            while (i + 1 < l.cum.size() && l.cum[i] < pos) ++i;
            const double seg = l.cum[i] - l.cum[i - 1];
            const double t = seg > 0.0 ? (pos - l.cum[i - 1]) / seg : 0.0;
            x = l.x[i - 1] + t * (l.x[i] - l.x[i - 1]) + 0.05 * r.normal();
            y = l.y[i - 1] + t * (l.y[i] - l.y[i - 1]) + 0.05 * r.normal();
        }
        *ox = snap7(x);
        *oy = snap7(y);
    }
};

void point(int config, const Clustered* cl, SplitMix& r, float* x, float* y)
{
    if (config <= 2) {
        *x = (float)(kSide * r.u53());
        *y = (float)(kSide * r.u53());
    } else if (config == 3) {
        cl->sample(r, x, y);
    } else if (r.u53() < 0.5) {
        *x = (float)(kSide * r.u53());
        *y = (float)(kSide * r.u53());
    } else {
        cl->sample(r, x, y);
    }
}

bool parse_u64(const char*& p, const char* end, uint64_t* v)
{
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    if (p >= end) return false;
    char* e = nullptr;
    errno = 0;
    unsigned long long x = std::strtoull(p, &e, 10);
    if (e == p || errno) return false;
    p = e;
    *v = x;
    return true;
}

bool parse_i32(const char*& p, const char* end, int32_t* v)
{
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    if (p >= end) return false;
    char* e = nullptr;
    errno = 0;
    long x = std::strtol(p, &e, 10);
    if (e == p || errno || x < INT32_MIN || x > INT32_MAX) return false;
    p = e;
    *v = (int32_t)x;
    return true;
}

bool parse_f32(const char*& p, const char* end, float* v)
{
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    if (p >= end) return false;
    char* e = nullptr;
    float x = std::strtof(p, &e);
    if (e == p || !std::isfinite(x)) return false;
    p = e;
    *v = x;
    return true;
}

} // namespace

extern "C" {

uint64_t rk_default_seed(int config) { return 0x524B4E4Eull ^ (uint64_t)config; }

void rk_free(void* p) { std::free(p); }

rk_status rk_load_text(const char* path, float** fx, float** fy, int32_t** fid, uint64_t* nf, float** ux,
                       float** uy, int32_t** uid, uint64_t* nu)
{
    if (!path || !fx || !fy || !fid || !nf || !ux || !uy || !uid || !nu) return RK_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return RK_ERR_IO;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) {
        std::fclose(f);
        return RK_ERR_IO;
    }
    std::vector<char> buf((size_t)sz + 1);
    size_t got = std::fread(buf.data(), 1, (size_t)sz, f);
    std::fclose(f);
    buf[got] = '\0';
    const char* p = buf.data();
    const char* end = p + got;
    uint64_t counts[2];
    float* xs[2] = {nullptr, nullptr};
    float* ys[2] = {nullptr, nullptr};
    int32_t* ids[2] = {nullptr, nullptr};
    for (int part = 0; part < 2; ++part) {
        // a record takes at least 5 bytes ("i x y" and a separator but the
        // last), so a count the rest of the file cannot hold fails before
        // anything is allocated for it
        if (!parse_u64(p, end, &counts[part]) || counts[part] > (1ull << 31) ||
            counts[part] > (uint64_t)(end - p) / 5)
            goto fail;
        xs[part] = (float*)std::malloc(sizeof(float) * (counts[part] ? counts[part] : 1));
        ys[part] = (float*)std::malloc(sizeof(float) * (counts[part] ? counts[part] : 1));
        ids[part] = (int32_t*)std::malloc(sizeof(int32_t) * (counts[part] ? counts[part] : 1));
        if (!xs[part] || !ys[part] || !ids[part]) goto fail;
        for (uint64_t i = 0; i < counts[part]; ++i) {
            if (!parse_i32(p, end, &ids[part][i]) || !parse_f32(p, end, &xs[part][i]) ||
                !parse_f32(p, end, &ys[part][i]))
                goto fail;
        }
    }
    *fx = xs[0];
    *fy = ys[0];
    *fid = ids[0];
    *nf = counts[0];
    *ux = xs[1];
    *uy = ys[1];
    *uid = ids[1];
    *nu = counts[1];
    return RK_OK;
fail:
    for (int part = 0; part < 2; ++part) {
        std::free(xs[part]);
        std::free(ys[part]);
        std::free(ids[part]);
    }
    return RK_ERR_IO;
}

rk_status rk_write_text(const char* path, const float* fx, const float* fy, const int32_t* fid, uint64_t nf,
                        const float* ux, const float* uy, const int32_t* uid, uint64_t nu)
{
    if (!path || (nf && (!fx || !fy || !fid)) || (nu && (!ux || !uy || !uid))) return RK_ERR_INVALID_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return RK_ERR_IO;
    std::vector<char> line(128);
    const float* xs[2] = {fx, ux};
    const float* ys[2] = {fy, uy};
    const int32_t* ids[2] = {fid, uid};
    const uint64_t cnt[2] = {nf, nu};
    bool ok = true;
    for (int part = 0; part < 2 && ok; ++part) {
        ok = std::fprintf(f, "%llu\n", (unsigned long long)cnt[part]) > 0;
        for (uint64_t i = 0; i < cnt[part] && ok; ++i)
            ok = std::fprintf(f, "%d %.17g %.17g\n", ids[part][i], (double)xs[part][i], (double)ys[part][i]) > 0;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? RK_OK : RK_ERR_IO;
}

rk_status rk_generate_queries(uint64_t nf, uint64_t nq, uint64_t seed, uint32_t* q_out)
{
    if (!q_out || nq > nf || nf > 0xffffffffull) return RK_ERR_INVALID_ARG;
    SplitMix r(seed ^ 0x5155455259ull);
    std::vector<uint32_t> perm(nf);
    for (uint64_t i = 0; i < nf; ++i) perm[i] = (uint32_t)i;
    for (uint64_t i = 0; i < nq; ++i) {
        const uint64_t j = i + r.below(nf - i);
        std::swap(perm[i], perm[j]);
        q_out[i] = perm[i];
    }
    return RK_OK;
}

rk_status rk_generate(int config, uint64_t nu, uint64_t nf, uint64_t seed, uint64_t nq_ties, uint64_t seed_ties,
                      float* ux, float* uy, int32_t* uid, float* fx, float* fy, int32_t* fid)
{
    if (config < 1 || config > 5) return RK_ERR_INVALID_ARG;
    if ((nu && (!ux || !uy || !uid)) || (nf && (!fx || !fy || !fid))) return RK_ERR_INVALID_ARG;
    if (nf + nu > 0x7fffffffull) return RK_ERR_INVALID_ARG;
    Clustered* cl = config >= 3 ? new Clustered(seed) : nullptr;
    SplitMix r(seed);
    for (uint64_t i = 0; i < nf; ++i) {
        point(config, cl, r, &fx[i], &fy[i]);
        fid[i] = (int32_t)i;
    }
    for (uint64_t i = 0; i < nu; ++i) {
        point(config, cl, r, &ux[i], &uy[i]);
        uid[i] = (int32_t)(nf + i);
    }
    delete cl;
    if (config < 3 || nu == 0 || nf < 2) return RK_OK;

    // 1% of users: exact bisector points of (query, near facility) pairs, or
    // duplicates of another user.
    std::vector<uint32_t> qs;
    std::vector<std::vector<uint32_t>> near;
    if (nq_ties > 0 && nq_ties <= nf) {
        qs.resize(nq_ties);
        rk_generate_queries(nf, nq_ties, seed_ties, qs.data());
        near.resize(nq_ties);
        for (uint64_t qi = 0; qi < nq_ties; ++qi) {
            const uint32_t q = qs[qi];
            std::vector<std::pair<double, uint32_t>> best;
            for (uint64_t f = 0; f < nf; ++f) {
                if (f == q || (fx[f] == fx[q] && fy[f] == fy[q])) continue;
                const double dx = (double)fx[f] - fx[q], dy = (double)fy[f] - fy[q];
                const double d = dx * dx + dy * dy;
                if (best.size() < 8) {
                    best.push_back({d, (uint32_t)f});
                    std::push_heap(best.begin(), best.end());
                } else if (d < best.front().first) {
                    std::pop_heap(best.begin(), best.end());
                    best.back() = {d, (uint32_t)f};
                    std::push_heap(best.begin(), best.end());
                }
            }
            for (auto& b : best) near[qi].push_back(b.second);
        }
    }
    SplitMix t(seed ^ 0x71E5ull);
    const uint64_t n_tie = nu / 100;
    for (uint64_t i = 0; i < n_tie; ++i) {
        const uint64_t u = t.below(nu);
        if ((i & 1) == 0 && !qs.empty()) {
            const uint64_t qi = t.below(qs.size());
            if (near[qi].empty()) continue;
            const uint32_t q = qs[qi], f = near[qi][t.below(near[qi].size())];
            int tt = (int)t.below(3) - 1;
            const double mx = 0.5 * ((double)fx[q] + fx[f]), my = 0.5 * ((double)fy[q] + fy[f]);
            const double px = -((double)fy[f] - fy[q]), py = (double)fx[f] - fx[q];
            double x = mx + tt * px, y = my + tt * py;
            if (x < 0.0 || x > kSide || y < 0.0 || y > kSide) {
                x = mx;
                y = my;
            }
            ux[u] = (float)x;
            uy[u] = (float)y;
        } else {
            const uint64_t src = t.below(nu);
            ux[u] = ux[src];
            uy[u] = uy[src];
        }
    }
    return RK_OK;
}

} // extern "C"
