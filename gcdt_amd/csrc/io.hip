// Input and output formats around the solve: the reference's text point/
// constraint file (kernel.cu:5159-5291 readFile), its double->grid map
// (kernel.cu:171-188 transPoints) and a canonical triangle list.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <random>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gcdt.h"
#include "gcdt_common.hpp"

namespace gcdt {
int fail(int status, const std::string& msg);
}

using gcdt::fail;

namespace {

// Whitespace-separated tokens as the reference's `infile >>` reads them.
struct Reader {
    FILE* f;
    bool num(long long& v) { return fscanf(f, "%lld", &v) == 1; }
    bool num(double& v) { return fscanf(f, "%lf", &v) == 1; }
};

// ((v - c) / l) * 0.99 + 0.5 as three rounded double operations, in the
// reference's order (kernel.cu:5227-5234); contraction into an FMA would
// change the last bit.
#pragma clang fp contract(off)
double normalise(double v, double c, double l)
{
    double t = v - c;
    t = t / l * 0.99;
    return t + 0.5;
}

__global__ void k_trans_points(const double* __restrict__ f, int n, long long res, int2* __restrict__ out)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double fx = f[2 * i], fy = f[2 * i + 1];
        // truncation toward zero, then clamp into [0, res-1]; v*res is exact
        // for a power-of-two res
        long long x = (long long)(fx * (double)res);
        long long y = (long long)(fy * (double)res);
        x = x >= res ? res - 1 : (x < 0 ? 0 : x);
        y = y >= res ? res - 1 : (y < 0 ? 0 : y);
        out[i] = make_int2((int)x, (int)y);
    }
}

}  // namespace

#define IO_BEGIN try {
#define IO_END                                                   \
    }                                                            \
    catch (const std::bad_alloc&) { return fail(GCDT_ERR_OOM, "host allocation failed"); } \
    catch (const gcdt::Error& e) { return fail(e.status, e.what()); }

extern "C" {

int gcdt_read_file(const char* path, int shuffle, double** points, int* numPoints, gcdt_int2** cons, int* numCons)
{
    IO_BEGIN
    if (!path || !points || !numPoints || !cons || !numCons) return fail(GCDT_ERR_INVALID, "null argument");
    *points = nullptr;
    *cons = nullptr;
    *numPoints = *numCons = 0;
    FILE* f = fopen(path, "r");
    if (!f) return fail(GCDT_ERR_INVALID, std::string("cannot open ") + path + ": " + strerror(errno));
    struct Closer {
        FILE* f;
        ~Closer() { fclose(f); }
    } closer{f};
    Reader rd{f};
    long long n_file = 0;
    if (!rd.num(n_file) || n_file < 0 || n_file > (1LL << 28)) return fail(GCDT_ERR_INVALID, "bad point count");
    // merge points with identical coordinates; the first occurrence keeps its
    // slot and later file ids map onto it (kernel.cu:5176-5203)
    std::vector<double> pts;
    pts.reserve(2 * (size_t)n_file);
    std::vector<int> idxmap((size_t)n_file, -1);
    std::map<std::pair<double, double>, int> seen;
    double minx = 0, miny = 0, maxx = 0, maxy = 0;
    for (long long i = 0; i < n_file; i++) {
        long long idx;
        double x, y;
        if (!rd.num(idx) || !rd.num(x) || !rd.num(y)) return fail(GCDT_ERR_INVALID, "truncated point list");
        if (idx < 0 || idx >= n_file) return fail(GCDT_ERR_INVALID, "point index out of range");
        if (!std::isfinite(x) || !std::isfinite(y)) return fail(GCDT_ERR_INVALID, "non-finite coordinate");
        auto key = std::make_pair(x, y);
        auto it = seen.find(key);
        if (it == seen.end()) {
            int r = (int)(pts.size() / 2);
            seen.emplace(key, r);
            pts.push_back(x);
            pts.push_back(y);
            idxmap[idx] = r;
            if (r == 0) {
                minx = maxx = x;
                miny = maxy = y;
            } else {
                minx = std::min(minx, x);
                maxx = std::max(maxx, x);
                miny = std::min(miny, y);
                maxy = std::max(maxy, y);
            }
        } else {
            idxmap[idx] = it->second;
        }
    }
    const int n = (int)(pts.size() / 2);
    const double cx = (minx + maxx) / 2, cy = (miny + maxy) / 2;
    double lx = maxx - minx;
    if (maxy - miny > lx) lx = maxy - miny;
    if (lx == 0) lx = 1;  // zero or one distinct point: keep them at the centre
    // optional shuffle: point i moves to slot random[i] (kernel.cu:5212-5227)
    std::vector<int> random;
    if (shuffle) {
        random.resize(n);
        std::iota(random.begin(), random.end(), 0);
        std::mt19937 gen(19999);
        std::shuffle(random.begin(), random.end(), gen);
        std::vector<double> moved(pts.size());
        for (int i = 0; i < n; i++) {
            moved[2 * (size_t)random[i]] = pts[2 * (size_t)i];
            moved[2 * (size_t)random[i] + 1] = pts[2 * (size_t)i + 1];
        }
        pts.swap(moved);
    }
    for (int i = 0; i < n; i++) {
        pts[2 * (size_t)i] = normalise(pts[2 * (size_t)i], cx, lx);
        pts[2 * (size_t)i + 1] = normalise(pts[2 * (size_t)i + 1], cy, lx);
    }
    // constraints: map through the merge (and shuffle); drop self-loops and
    // repeats in either orientation (kernel.cu:5238-5272)
    long long c_file = 0;
    std::vector<gcdt_int2> cs;
    if (rd.num(c_file)) {
        if (c_file < 0 || c_file > (1LL << 28)) return fail(GCDT_ERR_INVALID, "bad constraint count");
        std::set<std::pair<int, int>> have;
        for (long long i = 0; i < c_file; i++) {
            long long k, a, b;
            if (!rd.num(k) || !rd.num(a) || !rd.num(b)) return fail(GCDT_ERR_INVALID, "truncated constraint list");
            if (a < 0 || a >= n_file || b < 0 || b >= n_file || idxmap[a] < 0 || idxmap[b] < 0)
                return fail(GCDT_ERR_INVALID, "constraint endpoint is not a listed point");
            int x = idxmap[a], y = idxmap[b];
            if (shuffle) {
                x = random[x];
                y = random[y];
            }
            if (x == y || have.count({x, y}) || have.count({y, x})) continue;
            have.insert({x, y});
            cs.push_back(gcdt_int2{x, y});
        }
    }
    double* P = (double*)malloc(sizeof(double) * (pts.size() ? pts.size() : 1));
    gcdt_int2* C = (gcdt_int2*)malloc(sizeof(gcdt_int2) * (cs.size() ? cs.size() : 1));
    if (!P || !C) {
        free(P);
        free(C);
        return fail(GCDT_ERR_OOM, "host allocation failed");
    }
    if (!pts.empty()) memcpy(P, pts.data(), sizeof(double) * pts.size());
    if (!cs.empty()) memcpy(C, cs.data(), sizeof(gcdt_int2) * cs.size());
    *points = P;
    *numPoints = n;
    *cons = C;
    *numCons = (int)cs.size();
    return GCDT_OK;
    IO_END
}

void gcdt_free_host(void* p) { free(p); }

int gcdt_trans_points(const double* fpoints, int n, long long res, gcdt_int2* points, void* stream)
{
    IO_BEGIN
    if (n < 0 || (n > 0 && (!fpoints || !points)) || res <= 0 || res > (1LL << 31))
        return fail(GCDT_ERR_INVALID, "bad arguments");
    if (n == 0) return GCDT_OK;
    hipStream_t st = (hipStream_t)stream;
    k_trans_points<<<gcdt::grid_for(n), gcdt::kBlock, 0, st>>>(fpoints, n, res, (int2*)points);
    GCDT_LAUNCH_CHECK();
    GCDT_HIP(hipStreamSynchronize(st));
    return GCDT_OK;
    IO_END
}

int gcdt_canonicalize_host(gcdt_int3* tris, int64_t T)
{
    IO_BEGIN
    if (T < 0 || (T > 0 && !tris)) return fail(GCDT_ERR_INVALID, "bad arguments");
    for (int64_t i = 0; i < T; i++) {
        gcdt_int3 t = tris[i];
        // rotate to the first position holding the smallest id
        if (t.y < t.x && t.y <= t.z)
            tris[i] = gcdt_int3{t.y, t.z, t.x};
        else if (t.z < t.x && t.z < t.y)
            tris[i] = gcdt_int3{t.z, t.x, t.y};
    }
    std::sort(tris, tris + T, [](const gcdt_int3& a, const gcdt_int3& b) {
        if (a.x != b.x) return a.x < b.x;
        if (a.y != b.y) return a.y < b.y;
        return a.z < b.z;
    });
    return GCDT_OK;
    IO_END
}

int gcdt_write_mesh(const char* path, const gcdt_int2* points, int numPoints, const gcdt_int3* tris, int64_t T)
{
    IO_BEGIN
    if (!path || numPoints < 0 || T < 0 || (numPoints > 0 && !points) || (T > 0 && !tris))
        return fail(GCDT_ERR_INVALID, "bad arguments");
    FILE* f = fopen(path, "w");
    if (!f) return fail(GCDT_ERR_INVALID, std::string("cannot create ") + path + ": " + strerror(errno));
    bool ok = fprintf(f, "%d\n", numPoints) > 0;
    for (int i = 0; ok && i < numPoints; i++) ok = fprintf(f, "%d %d %d\n", i, points[i].x, points[i].y) > 0;
    ok = ok && fprintf(f, "%lld\n", (long long)T) > 0;
    for (int64_t i = 0; ok && i < T; i++)
        ok = fprintf(f, "%lld %d %d %d\n", (long long)i, tris[i].x, tris[i].y, tris[i].z) > 0;
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(GCDT_ERR_INVALID, std::string("write failed: ") + path);
    return GCDT_OK;
    IO_END
}

}  // extern "C"
