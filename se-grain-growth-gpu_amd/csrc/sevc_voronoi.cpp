// sevc_voronoi.cpp -- periodic Voronoi tessellation of the unit torus.
//
// Plays the role of voronoi/voronoi.py:12-109 (3x3 periodic images, merge
// of equivalent vertices, trivalency check) for meshes of millions of
// grains.  Each Voronoi cell is built independently by clipping a square
// with the bisectors of nearby seeds (periodic images included) found on a
// bucket grid, until no farther seed can cut it.  A cell corner is the
// circumcentre of its seed and the two seeds whose bisectors meet there, so
// vertices are identified by that seed triple (with image shifts,
// normalised to the lowest seed) instead of by rounded coordinates.  Every
// vertex must be reported by exactly three cells and every edge by exactly
// two; otherwise the seed set is rejected (SEVC_E_TOPOLOGY), as voronoi.py
// retries when its trivalency check fails.
//
// Seeds are sorted along a Morton curve before numbering, so that
// consecutive grains (and, through the loader's face walk, consecutive
// vertices) are spatial neighbours.
#include "../../include/sevc.h"
#include "../../include/sevc_math.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

struct Nb {            // a neighbour seed image
    int id;
    int sx, sy;        // image shift
};

struct PVert {
    double x, y;
    Nb lab;            // generator of the edge leaving this vertex (id -1: box side)
};

struct Key {
    int32_t v[9];
    bool operator==(const Key& o) const { return std::memcmp(v, o.v, sizeof v) == 0; }
};

struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t h = 1469598103934665603ull;
        for (int i = 0; i < 9; ++i) {
            h ^= static_cast<uint32_t>(k.v[i]);
            h *= 1099511628211ull;
        }
        return static_cast<size_t>(h ^ (h >> 29));
    }
};

inline uint32_t spread16(uint32_t x) {
    x &= 0xFFFF;
    x = (x | (x << 8)) & 0x00FF00FF;
    x = (x | (x << 4)) & 0x0F0F0F0F;
    x = (x | (x << 2)) & 0x33333333;
    x = (x | (x << 1)) & 0x55555555;
    return x;
}

// canonical key of the corner defined by seed i (no shift) and neighbours a, b
Key corner_key(int i, const Nb& a, const Nb& b) {
    Nb t[3] = {Nb{i, 0, 0}, a, b};
    int lo = 0;
    for (int k = 1; k < 3; ++k)
        if (t[k].id < t[lo].id || (t[k].id == t[lo].id && (t[k].sx < t[lo].sx ||
                                                           (t[k].sx == t[lo].sx && t[k].sy < t[lo].sy))))
            lo = k;
    const int bx = t[lo].sx, by = t[lo].sy;
    for (int k = 0; k < 3; ++k) { t[k].sx -= bx; t[k].sy -= by; }
    std::sort(t, t + 3, [](const Nb& p, const Nb& q) {
        if (p.id != q.id) return p.id < q.id;
        if (p.sx != q.sx) return p.sx < q.sx;
        return p.sy < q.sy;
    });
    Key k;
    for (int q = 0; q < 3; ++q) { k.v[3 * q] = t[q].id; k.v[3 * q + 1] = t[q].sx; k.v[3 * q + 2] = t[q].sy; }
    return k;
}

}  // namespace

extern "C" int sevc_voronoi_generate(int32_t n, uint64_t seed, int32_t* n_vertices, double** unit_pos,
                                     int32_t* n_ridges, int32_t** ridges) {
    if (n < 2 || !n_vertices || !unit_pos || !n_ridges || !ridges) return SEVC_E_INVALID_ARG;
    *unit_pos = nullptr;
    *ridges = nullptr;
    // seeds: Philox(seed; k, 0, 4, 0) -> two uniforms in [0,1)
    std::vector<double> sx(static_cast<size_t>(n)), sy(static_cast<size_t>(n));
    {
        std::vector<std::pair<uint32_t, int>> order(static_cast<size_t>(n));
        std::vector<double> tx(static_cast<size_t>(n)), ty(static_cast<size_t>(n));
        for (int k = 0; k < n; ++k) {
            uint32_t ctr[4] = {static_cast<uint32_t>(k), 0u, 4u, 0u};
            uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
            uint32_t o[4];
            sevc_philox4x32_10(ctr, key, o);
            tx[k] = static_cast<double>(((static_cast<uint64_t>(o[0]) << 32) | o[1]) >> 11) * (1.0 / 9007199254740992.0);
            ty[k] = static_cast<double>(((static_cast<uint64_t>(o[2]) << 32) | o[3]) >> 11) * (1.0 / 9007199254740992.0);
            uint32_t qx = static_cast<uint32_t>(tx[k] * 65536.0), qy = static_cast<uint32_t>(ty[k] * 65536.0);
            order[k] = {spread16(qx) | (spread16(qy) << 1), k};
        }
        std::sort(order.begin(), order.end());
        for (int k = 0; k < n; ++k) { sx[k] = tx[order[k].second]; sy[k] = ty[order[k].second]; }
    }
    // bucket grid, about 2 seeds per bucket
    const int M = std::max(1, static_cast<int>(std::sqrt(n / 2.0)));
    const double w = 1.0 / M;
    std::vector<int> head(static_cast<size_t>(M) * M + 1, 0), items(static_cast<size_t>(n));
    auto bucket = [&](double x, double y) {
        int cx = std::min(M - 1, static_cast<int>(x * M)), cy = std::min(M - 1, static_cast<int>(y * M));
        return cy * M + cx;
    };
    for (int k = 0; k < n; ++k) head[bucket(sx[k], sy[k]) + 1]++;
    for (int c = 0; c < M * M; ++c) head[c + 1] += head[c];
    {
        std::vector<int> fill(head.begin(), head.end() - 1);
        for (int k = 0; k < n; ++k) items[fill[bucket(sx[k], sy[k])]++] = k;
    }

    // per-cell clipping
    std::vector<std::vector<PVert>> cells(static_cast<size_t>(n));
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bad)
    for (int i = 0; i < n; ++i) {
        const double px = sx[i], py = sy[i];
        std::vector<PVert> poly, next;
        const double H = 1.0;
        poly.push_back({px - H, py - H, {-1, 0, 0}});
        poly.push_back({px + H, py - H, {-1, 0, 0}});
        poly.push_back({px + H, py + H, {-1, 0, 0}});
        poly.push_back({px - H, py + H, {-1, 0, 0}});
        const int cx = std::min(M - 1, static_cast<int>(px * M)), cy = std::min(M - 1, static_cast<int>(py * M));
        bool done = false;
        for (int r = 0; !done && r <= M + 2; ++r) {
            for (int dy = -r; dy <= r; ++dy)
                for (int dx = -r; dx <= r; ++dx) {
                    if (std::max(std::abs(dx), std::abs(dy)) != r) continue;
                    const int ux = cx + dx, uy = cy + dy;
                    const int bx = ((ux % M) + M) % M, by = ((uy % M) + M) % M;
                    const int shx = (ux - bx) / M, shy = (uy - by) / M;
                    const int c = by * M + bx;
                    for (int t = head[c]; t < head[c + 1]; ++t) {
                        const int j = items[t];
                        if (j == i && shx == 0 && shy == 0) continue;
                        const double qx = sx[j] + shx, qy = sy[j] + shy;
                        const double nx = qx - px, ny = qy - py;
                        const double mx = 0.5 * (px + qx), my = 0.5 * (py + qy);
                        next.clear();
                        const size_t L = poly.size();
                        for (size_t k = 0; k < L; ++k) {
                            const PVert& a = poly[k];
                            const PVert& b = poly[(k + 1) % L];
                            const double ha = (a.x - mx) * nx + (a.y - my) * ny;
                            const double hb = (b.x - mx) * nx + (b.y - my) * ny;
                            const bool ia = ha <= 0, ib = hb <= 0;
                            if (ia) next.push_back(a);
                            if (ia != ib) {
                                const double tt = ha / (ha - hb);
                                PVert q{a.x + tt * (b.x - a.x), a.y + tt * (b.y - a.y), a.lab};
                                if (ia) q.lab = Nb{j, shx, shy};
                                next.push_back(q);
                            }
                        }
                        poly.swap(next);
                    }
                }
            // security radius: seeds outside ring r are at least r*w away
            double md = 0;
            bool box = false;
            for (const PVert& v : poly) {
                md = std::max(md, (v.x - px) * (v.x - px) + (v.y - py) * (v.y - py));
                if (v.lab.id < 0) box = true;
            }
            if (!box && 4.0 * md <= (r * w) * (r * w)) done = true;
        }
        if (!done || poly.size() < 3) bad += 1;
        cells[static_cast<size_t>(i)] = poly;
    }
    if (bad) return SEVC_E_TOPOLOGY;
    // a cell bounded by its own periodic image cannot be stored (a boundary
    // must separate two different grains, geometry.h:837-848)
    for (int i = 0; i < n; ++i)
        for (const PVert& v : cells[static_cast<size_t>(i)])
            if (v.lab.id == i) return SEVC_E_TOPOLOGY;

    // vertices by seed triple, numbered in order of first appearance
    std::unordered_map<Key, int, KeyHash> vid;
    vid.reserve(static_cast<size_t>(2.2 * n));
    std::vector<Key> vkeys;
    std::vector<int> vcount;
    std::vector<std::vector<int>> cellv(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const auto& poly = cells[static_cast<size_t>(i)];
        const size_t L = poly.size();
        auto& cv = cellv[static_cast<size_t>(i)];
        cv.resize(L);
        for (size_t k = 0; k < L; ++k) {
            const Nb& before = poly[(k + L - 1) % L].lab;
            const Nb& after = poly[k].lab;
            Key key = corner_key(i, before, after);
            auto it = vid.find(key);
            int id;
            if (it == vid.end()) {
                id = static_cast<int>(vkeys.size());
                vid.emplace(key, id);
                vkeys.push_back(key);
                vcount.push_back(0);
            } else {
                id = it->second;
            }
            vcount[static_cast<size_t>(id)]++;
            cv[k] = id;
        }
    }
    const int nv = static_cast<int>(vkeys.size());
    for (int v = 0; v < nv; ++v) if (vcount[static_cast<size_t>(v)] != 3) return SEVC_E_TOPOLOGY;
    // edges: consecutive corners of a cell, each must be shared by two cells
    std::unordered_map<uint64_t, int> eid;
    eid.reserve(static_cast<size_t>(3.3 * n));
    std::vector<int32_t> rid;
    std::vector<int> ecount;
    for (int i = 0; i < n; ++i) {
        const auto& cv = cellv[static_cast<size_t>(i)];
        const size_t L = cv.size();
        for (size_t k = 0; k < L; ++k) {
            int a = cv[k], b = cv[(k + 1) % L];
            if (a == b) return SEVC_E_TOPOLOGY;
            const int lo = std::min(a, b), hi = std::max(a, b);
            const uint64_t key = (static_cast<uint64_t>(lo) << 32) | static_cast<uint32_t>(hi);
            auto it = eid.find(key);
            if (it == eid.end()) {
                eid.emplace(key, static_cast<int>(ecount.size()));
                ecount.push_back(1);
                rid.push_back(lo);
                rid.push_back(hi);
            } else {
                ecount[static_cast<size_t>(it->second)]++;
            }
        }
    }
    for (int c : ecount) if (c != 2) return SEVC_E_TOPOLOGY;
    // positions: circumcentre of the canonical triple, wrapped into [0,1)
    double* P = static_cast<double*>(std::malloc(sizeof(double) * 2 * static_cast<size_t>(nv)));
    int32_t* R = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * rid.size()));
    if (!P || !R) { std::free(P); std::free(R); return SEVC_E_NOMEM; }
#pragma omp parallel for schedule(static)
    for (int v = 0; v < nv; ++v) {
        const Key& k = vkeys[static_cast<size_t>(v)];
        double ax = sx[k.v[0]] + k.v[1], ay = sy[k.v[0]] + k.v[2];
        double bx = sx[k.v[3]] + k.v[4], by = sy[k.v[3]] + k.v[5];
        double cx = sx[k.v[6]] + k.v[7], cy = sy[k.v[6]] + k.v[8];
        double bxr = bx - ax, byr = by - ay, cxr = cx - ax, cyr = cy - ay;
        double dd = 2.0 * (bxr * cyr - byr * cxr);
        double b2 = bxr * bxr + byr * byr, c2 = cxr * cxr + cyr * cyr;
        double ux = ax + (cyr * b2 - byr * c2) / dd;
        double uy = ay + (bxr * c2 - cxr * b2) / dd;
        ux -= std::floor(ux);
        uy -= std::floor(uy);
        if (ux >= 1.0) ux = 0.0;
        if (uy >= 1.0) uy = 0.0;
        P[2 * v] = ux;
        P[2 * v + 1] = uy;
    }
    std::memcpy(R, rid.data(), sizeof(int32_t) * rid.size());
    *n_vertices = nv;
    *unit_pos = P;
    *n_ridges = static_cast<int32_t>(rid.size() / 2);
    *ridges = R;
    return SEVC_OK;
}
