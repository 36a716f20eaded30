// Host-only check of the forward projector's depth-strip orderings (designs
// 2, 3 and 4), built from geometry.cpp and run on the CPU:
//  * design 2: for every angle and contributing column, the strips holding
//    the column are exactly those whose window [s*W - 2, (s+1)*W) contains
//    its first tap, each once; strip offsets are monotone and end at the
//    angle's entry count;
//  * design 3: every (angle, strip, column) contribution of design 2 appears
//    in exactly one angle-pair entry (the entry's flag for that angle set),
//    and no entry carries a contribution design 2 does not have;
//  * design 4: the same contributions appear exactly once as staged-strip
//    entries (chunk slot -> column through the stage list, angle of the
//    group), every chunk slot referenced holds a column, every staged column
//    is referenced, offsets within a (chunk, angle) are non-decreasing with
//    one offset per batch of kStripBatch (zero-weight padding entries fill
//    the runs), at most kSgEnt entries per (chunk, angle), the entry ranges
//    tile the table and chunk counts match the stage lists.
// Exit status 0 = all checks passed.  argv: method nx ny na.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "arec_internal.hpp"

using namespace arec;

int main(int argc, char **argv) {
    const int method = argc > 1 ? atoi(argv[1]) : kSQ;
    const int nx = argc > 2 ? atoi(argv[2]) : 48, ny = argc > 3 ? atoi(argv[3]) : 4;
    const int na = argc > 4 ? atoi(argv[4]) : 13;
    std::vector<float> ang(na);
    for (int a = 0; a < na; a++) ang[a] = (float)(a * 3.14159265358979 / na);
    Geometry g;
    std::string err;
    if (geometry_build_host(g, method, nx, ny, 0.5 * (nx - 1), ang.data(), na, &err)) {
        fprintf(stderr, "geometry: %s\n", err.c_str());
        return 2;
    }
    int fails = 0;
    // design 2
    geometry_strip_count(g);
    const int nb = g.nstrips * g.nbins;
    std::vector<uint32_t> order;
    geometry_strip_angles(g, 0, na, order);
    std::map<std::tuple<int, int, int>, int> d2;  // (angle, strip, column) -> count
    for (int a = 0; a < na; a++) {
        const int64_t e0 = g.strip_base[a], n = g.strip_base[a + 1] - e0;
        const int *off = &g.strip_off[(size_t)a * (nb + 1)];
        for (int i = 0; i < nb; i++)
            if (off[i] > off[i + 1]) fails++;
        if (off[0] != 0 || off[nb] != n) fails++;
        for (int64_t e = 0; e < n; e++) {
            const uint32_t o = order[e0 + e];
            const int j = (int)(o & (kMaxCols - 1)), s = (int)(o >> kColBits);
            d2[{a, s, j}]++;
            const int p0 = host_first_tap(g, a, j);
            if (p0 < s * kStripW - 2 || p0 >= (s + 1) * kStripW) fails++;
        }
        for (int j = 0; j < g.nvalid; j++) {
            const int p0 = host_first_tap(g, a, j);
            for (int s = 0; s < g.nstrips; s++) {
                const bool in = p0 >= s * kStripW - 2 && p0 < (s + 1) * kStripW;
                auto it = d2.find({a, s, j});
                const int c = it == d2.end() ? 0 : it->second;
                if (c != (in ? 1 : 0)) fails++;
            }
        }
    }
    // design 3
    Geometry h = g;
    geometry_pair_count(h);
    const int nu = (na + 1) / 2;
    geometry_pair_units(h, 0, nu, order);
    std::map<std::tuple<int, int, int>, int> d3;
    for (int u = 0; u < nu; u++) {
        const int64_t e0 = h.strip_base[u], n = h.strip_base[u + 1] - e0;
        const int *off = &h.strip_off[(size_t)u * (nb + 1)];
        if (off[0] != 0 || off[nb] != n) fails++;
        for (int64_t e = 0; e < n; e++) {
            const uint32_t o = order[e0 + e];
            const int j = (int)(o & (kMaxCols - 1)), s = (int)((o >> kColBits) & 255);
            const unsigned m = o >> 30;
            if (m == 0) fails++;
            for (int k = 0; k < 2; k++)
                if (m & (1u << k)) {
                    if (2 * u + k >= na) fails++;
                    d3[{2 * u + k, s, j}]++;
                }
        }
    }
    if (d2 != d3) fails++;
    // design 4
    Geometry q = g;
    geometry_staged_count(q);
    const int ngroups = (na + kSgAngles - 1) / kSgAngles, ns = q.nstrips;
    std::vector<int> stage;
    std::vector<int64_t> off;
    std::vector<uint32_t> col, angi, meta;
    if (!geometry_staged_groups(q, 0, ngroups, stage, off, col, angi, meta)) fails++;
    std::map<std::tuple<int, int, int>, int> d4;
    const int64_t nchunk = q.sg_chunk_base.back(), total = q.sg_group_ent.back();
    if ((int64_t)stage.size() != nchunk * kSgChunk || (int64_t)off.size() != nchunk * kSgAngles ||
        (int64_t)col.size() != total)
        fails++;
    off.push_back(total);
    for (int G = 0; G < ngroups; G++)
        for (int s = 0; s < ns; s++)
            for (int64_t K = q.sg_chunk_base[(size_t)G * ns + s]; K < q.sg_chunk_base[(size_t)G * ns + s + 1]; K++) {
                std::vector<int> used(kSgChunk, 0);
                for (int i = 0; i < kSgAngles; i++) {
                    const int64_t e0 = off[(size_t)K * kSgAngles + i], e1 = off[(size_t)K * kSgAngles + i + 1];
                    if (e1 < e0) fails++;
                    int dprev = -1;
                    if ((e1 - e0) % kStripBatch || e1 - e0 > kSgEnt) fails++;  // whole batches, staged size
                    for (int64_t e = e0; e < e1; e++) {
                        const int slot = (int)(meta[e] & 255u), d = (int)((meta[e] >> 8) & 31u);
                        const int ms = (int)(meta[e] >> 16);
                        if ((int)angi[e] != G * kSgAngles + i || ms != s || d < dprev) fails++;
                        if ((e - e0) % kStripBatch && d != dprev) fails++;  // one offset per batch
                        dprev = d;
                        if (col[e] == 0xffffffffu) continue;  // zero-weight padding
                        const int j = stage[(size_t)K * kSgChunk + slot];
                        if (j < 0 || j != (int)col[e]) fails++;
                        if (j >= 0) used[slot]++;
                        if (host_first_tap(q, (int)angi[e], (int)col[e]) - (s * kStripW - 2) != d) fails++;
                        d4[{(int)angi[e], s, (int)col[e]}]++;
                    }
                }
                for (int slot = 0; slot < kSgChunk; slot++)
                    if ((stage[(size_t)K * kSgChunk + slot] >= 0) != (used[slot] > 0)) fails++;
            }
    if (d2 != d4) fails++;
    printf(fails ? "strip order check: %d failures\n" : "strip order check: all checks passed%.0d\n", fails);
    return fails ? 1 : 0;
}
