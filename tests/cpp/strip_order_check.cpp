// Host-only check of the forward projector's depth-strip orderings (designs
// 2 and 3), built from geometry.cpp and run on the CPU:
//  * design 2: for every angle and contributing column, the strips holding
//    the column are exactly those whose window [s*W - 2, (s+1)*W) contains
//    its first tap, each once; strip offsets are monotone and end at the
//    angle's entry count;
//  * design 3: every (angle, strip, column) contribution of design 2 appears
//    in exactly one angle-pair entry (the entry's flag for that angle set),
//    and no entry carries a contribution design 2 does not have.
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
    printf(fails ? "strip order check: %d failures\n" : "strip order check: all checks passed%.0d\n", fails);
    return fails ? 1 : 0;
}
