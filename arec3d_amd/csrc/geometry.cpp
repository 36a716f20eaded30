// Host-side geometry: cylinder mask, per-angle constants, back-projection
// tiles and the ordering of the forward table.  Compiled with
// -ffp-contract=off so that every FP64 product and sum here is a separate
// rounding, exactly like the device weight function that uses the same
// formulas.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>

#include <unistd.h>

#include "arec_internal.hpp"

namespace arec {

// arecAllocateCylImage, AREC3D/AREC3D/arecImage.cpp:29-86.  The mask keeps
// (ix,iz) with (int)(ix-r)^2 + (int)(iz-r)^2 <= (int)(r^2), every conversion
// truncating toward zero, rays listed iz-major.
int cyl_mask(double radius, std::vector<int> *cord) {
    int n = (int)std::lround(2 * radius + 1);
    int r2 = (int)(radius * radius);
    int nrays = 0;
    if (cord) cord->clear();
    for (int iz = 0; iz < n; iz++) {
        int zs = (int)(iz - radius);
        for (int ix = 0; ix < n; ix++) {
            int xs = (int)(ix - radius);
            if (zs * zs + xs * xs <= r2) {
                if (cord) {
                    cord->push_back(ix);
                    cord->push_back(iz);
                }
                nrays++;
            }
        }
    }
    return nrays;
}

static AngleConst angle_const(float angle) {
    AngleConst a;
    a.c = (double)cosf(angle);
    a.s = (double)sinf(angle);
    double cd = std::cos((double)angle), sd = std::sin((double)angle);
    double cms = std::fabs(0.5 * (cd - sd));
    double cps = std::fabs(0.5 * (cd + sd));
    a.a = std::min(cms, cps);
    a.b = std::max(cms, cps) + 1e-6;
    a.lmax = 1 / ((a.b - a.a) + 2 * a.a);
    return a;
}

int geometry_build_host(Geometry &g, int method, int nx, int ny, double radius,
                        const float *angles, int na, std::string *err) {
    if (method < 1 || method > 3) {
        *err = "method must be 1 (linear), 2 (KB) or 3 (SQ)";
        return AREC_EINVAL;
    }
    if (nx <= 0 || ny <= 0 || na <= 0 || !(radius > 0)) {
        *err = "sizes and radius must be positive";
        return AREC_EINVAL;
    }
    g.method = method;
    g.nx = nx;
    g.ny = ny;
    g.na = na;
    g.radius = radius;
    g.ncyl = (int)std::lround(2 * radius + 1);
    g.nrays = cyl_mask(radius, &g.cord);
    g.nnz = (int64_t)g.nrays * ny;
    g.nproj = (int64_t)nx * ny * na;
    g.xcent = (int)radius + 1;
    if (g.nrays >= kMaxCols) {
        *err = "more voxel columns than the forward table encodes (2^22)";
        return AREC_ESIZE;
    }
    g.angles.assign(angles, angles + na);
    g.ang.resize(na);
    for (int a = 0; a < na; a++) g.ang[a] = angle_const(angles[a]);

    // Coordinates fed to the weight function.  SQ and KB use the column's
    // own (ix, iz).  The linear pair indexes the 0-based cord table with a
    // 1-based loop (arecproject.cpp:15,75-77), so column j is placed at
    // X = iz of column j+1, Z = ix of column j+2; the last two columns would
    // read past the table and contribute nothing here.
    g.coord.assign(2 * (size_t)g.nrays, 0);
    g.nvalid = g.nrays;
    if (method == kLinear) {
        g.nvalid = std::max(0, g.nrays - 2);
        for (int j = 0; j < g.nvalid; j++) {
            g.coord[2 * j] = g.cord[2 * j + 3];
            g.coord[2 * j + 1] = g.cord[2 * j + 4];
        }
    } else {
        g.coord = g.cord;
    }

    // Back-projection tiles: columns bucketed by kTileBox x kTileBox boxes of
    // their coordinates (in column order), buckets split into tiles of at
    // most kTileCols columns.  Coordinates need not be distinct (the linear
    // pair's shifted coordinates repeat), so slots are filled in order, not
    // by position.
    std::map<std::pair<int, int>, std::vector<int>> boxes;
    for (int j = 0; j < g.nvalid; j++)
        boxes[{g.coord[2 * j + 1] / kTileBox, g.coord[2 * j] / kTileBox}].push_back(j);
    g.tile_cols.clear();
    g.tile_org.clear();
    for (auto &kv : boxes) {
        const std::vector<int> &cols = kv.second;
        for (size_t s = 0; s < cols.size(); s += kTileCols) {
            for (int c = 0; c < kTileCols; c++)
                g.tile_cols.push_back(s + c < cols.size() ? cols[s + c] : -1);
            g.tile_org.push_back(kv.first.second * kTileBox);
            g.tile_org.push_back(kv.first.first * kTileBox);
        }
    }
    g.ntiles = (int)(g.tile_org.size() / 2);
    return 0;
}

// Same arithmetic as tap_weights<> in projector.hip (first tap only).
int host_first_tap(const Geometry &g, int a, int j) {
    const AngleConst &A = g.ang[a];
    int X = g.coord[2 * j], Z = g.coord[2 * j + 1];
    if (g.method == kLinear) {
        int xc = X - g.xcent, zc = Z - g.xcent;
        double t1 = (double)xc * A.c;
        double t2 = (double)zc * A.s;
        float xb = (float)((t1 + t2) + (double)g.xcent);
        return (int)std::floor(xb) - 1;
    }
    double xc = (double)X - g.radius, zc = (double)Z - g.radius;
    double t1 = xc * A.c;
    double t2 = zc * A.s;
    double xb = (t1 + t2) + g.radius;
    return (int)std::nearbyint(xb) - 1;
}

bool fwd_rpl_allowed(int ny, int rpl) {
    if (rpl == 1) return true;
    if (rpl == 2) return ny % 2 == 0;
    if (rpl == 4) return ny % 4 == 0;
    return false;
}

// Launch shape of the forward projector: rows per lane (rpl; 0 picks 4 with
// 16-byte loads when the row count allows), row chunks of 64 lanes, and
// enough detector segments per angle that the grid has >= min_groups
// workgroups of kFwdAngles waves (0 picks 1024).  Returns false for an rpl
// the row count does not allow.
bool geometry_plan_forward(Geometry &g, int rpl, int min_groups) {
    if (rpl == 0) rpl = (g.ny >= 256 && g.ny % 4 == 0) ? 4 : (g.ny >= 128 && g.ny % 2 == 0) ? 2 : 1;
    if (!fwd_rpl_allowed(g.ny, rpl)) return false;
    if (min_groups <= 0) min_groups = 1024;
    g.fwd_rpl = rpl;
    g.fwd_min_groups = min_groups;
    g.fwd_chunks = (g.ny + 64 * g.fwd_rpl - 1) / (64 * g.fwd_rpl);
    int64_t groups = (int64_t)((g.na + kFwdAngles - 1) / kFwdAngles) * g.fwd_chunks;
    int nseg = (int)std::max<int64_t>(1, (min_groups + groups - 1) / groups);
    nseg = std::min(nseg, std::max(1, g.nx / (4 * kFwdOut)));
    int segw = (g.nx + nseg - 1) / nseg;
    segw = (segw + kFwdOut - 1) / kFwdOut * kFwdOut;
    g.segw = segw;
    g.nseg = (g.nx + segw - 1) / segw;
    return true;
}

// Segment table of the current plan from pixstart: segment s (pixels
// [s*segw, (s+1)*segw)) needs the entries whose first tap lies in
// [s*segw - 2, (s+1)*segw), and the first tap of the entry before them.
void geometry_segments(Geometry &g) { segment_table(g, g.segw, g.nseg, g.seg); }

void segment_table(const Geometry &g, int segw, int nseg, std::vector<int> &seg) {
    const int nv = g.nvalid, npix = g.npix;
    seg.assign((size_t)g.na * nseg * 4, 0);
    for (int a = 0; a < g.na; a++) {
        const int *ps = &g.pixstart[(size_t)a * npix];
        // first entry whose first tap is >= k
        auto first_at = [&](int k) {
            int q = k - kTapBase;
            return q <= 0 ? 0 : q >= npix ? nv : ps[q];
        };
        // The first tap of entry e is the largest k with first_at(k) <= e;
        // qk tracks it for the (non-decreasing) entries start - 1.
        int qk = 0;
        for (int s = 0; s < nseg; s++) {
            int start = first_at(s * segw - 2), end = first_at((s + 1) * segw);
            int *e = &seg[((size_t)a * nseg + s) * 4];
            e[0] = start;
            e[1] = end;
            if (start == 0) {
                e[2] = kTapBase;
            } else {
                while (qk + 1 < npix && ps[qk + 1] <= start - 1) qk++;
                e[2] = qk + kTapBase;
            }
            e[3] = 0;
        }
    }
}

// Depth along the ray of column j at angle a (pixels, centred), the same
// coordinates as host_first_tap.
static double host_depth(const Geometry &g, int a, int j) {
    const AngleConst &A = g.ang[a];
    double c = g.method == kLinear ? (double)g.xcent : g.radius;
    double xc = (double)g.coord[2 * j] - c, zc = (double)g.coord[2 * j + 1] - c;
    return zc * A.c - xc * A.s;
}

static int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// The strips whose entry ranges hold a column with first tap p0: s0 = p0
// div kStripW and, when p0 is within 2 pixels of the next strip, s0 + 1.
static int strips_of(int p0, int nstrips, int *s) {
    int n = 0, s0 = floor_div(p0, kStripW);
    if (s0 >= 0 && s0 < nstrips) s[n++] = s0;
    if (p0 - s0 * kStripW >= kStripW - 2 && s0 + 1 >= 0 && s0 + 1 < nstrips) s[n++] = s0 + 1;
    return n;
}

void geometry_strip_count(Geometry &g) {
    g.nstrips = (g.nx + kStripW - 1) / kStripW;
    double c = g.method == kLinear ? (double)g.xcent : g.radius;
    g.nbins = (int)std::ceil((2 * c + 4) / kDepthBin) + 1;
    g.strip_base.assign((size_t)g.na + 1, 0);
    std::vector<int64_t> cnt(g.na, 0);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nth = std::min<unsigned>(nth, (unsigned)g.na);
    auto work = [&](unsigned t) {
        int s[2];
        for (int a = (int)t; a < g.na; a += (int)nth) {
            int64_t n = 0;
            for (int j = 0; j < g.nvalid; j++) n += strips_of(host_first_tap(g, a, j), g.nstrips, s);
            cnt[a] = n;
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++) pool.emplace_back(work, t);
    for (auto &th : pool) th.join();
    for (int a = 0; a < g.na; a++) g.strip_base[a + 1] = g.strip_base[a] + cnt[a];
    g.strip_off.assign((size_t)g.na * ((size_t)g.nstrips * g.nbins + 1), 0);
}

void geometry_strip_angles(Geometry &g, int a0, int a1, std::vector<uint32_t> &order) {
    const int nb = g.nstrips * g.nbins, nk = nb * kStripGroups;
    const int64_t e0 = g.strip_base[a0];
    order.resize((size_t)(g.strip_base[a1] - e0));
    double c = g.method == kLinear ? (double)g.xcent : g.radius;
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nth = std::min<unsigned>(nth, (unsigned)(a1 - a0));
    auto work = [&](unsigned t) {
        std::vector<int> cnt(nk + 1);
        std::vector<int> nstr(g.nvalid);
        std::vector<int> str(2 * (size_t)g.nvalid), key(2 * (size_t)g.nvalid);
        for (int a = a0 + (int)t; a < a1; a += (int)nth) {
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int j = 0; j < g.nvalid; j++) {
                int b = (int)std::floor((host_depth(g, a, j) + c + 1.0) / kDepthBin);
                b = std::min(std::max(b, 0), g.nbins - 1);
                const int p0 = host_first_tap(g, a, j);
                nstr[j] = strips_of(p0, g.nstrips, &str[2 * (size_t)j]);
                for (int k = 0; k < nstr[j]; k++) {
                    const int sj = str[2 * (size_t)j + k], d = p0 - (sj * kStripW - 2);
                    key[2 * (size_t)j + k] = (sj * g.nbins + b) * kStripGroups + d;
                    cnt[key[2 * (size_t)j + k] + 1]++;
                }
            }
            for (int i = 1; i <= nk; i++) cnt[i] += cnt[i - 1];
            int *off = &g.strip_off[(size_t)a * (nb + 1)];
            for (int i = 0; i <= nb; i++) off[i] = cnt[(size_t)i * kStripGroups];
            uint32_t *out = order.data() + (g.strip_base[a] - e0);
            for (int j = 0; j < g.nvalid; j++)
                for (int k = 0; k < nstr[j]; k++)
                    out[cnt[key[2 * (size_t)j + k]]++] =
                        (uint32_t)j | ((uint32_t)str[2 * (size_t)j + k] << kColBits);
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++) pool.emplace_back(work, t);
    for (auto &th : pool) th.join();
}

// Angle pairs (design 3): unit u holds angles 2u and 2u + 1 (the second
// absent when 2u + 1 == na).  A column is an entry of (u, strip s) when s
// is one of its strips at either angle; the entry carries a present flag
// per angle, so every (angle, strip, column) contribution of design 2 is in
// exactly one entry.  Entries are ordered by the depth bin at angle 2u and,
// inside a bin, by (d0, d1), the first-tap offsets in the strip (absent =
// kStripGroups).
static int pair_strips(const Geometry &g, int a, int j, int *s, unsigned *mask) {
    int sa[2], sb[2];
    const int n0 = strips_of(host_first_tap(g, a, j), g.nstrips, sa);
    const int n1 = a + 1 < g.na ? strips_of(host_first_tap(g, a + 1, j), g.nstrips, sb) : 0;
    int n = 0;
    auto add = [&](int st, unsigned bit) {
        for (int k = 0; k < n; k++)
            if (s[k] == st) {
                mask[k] |= bit;
                return;
            }
        s[n] = st;
        mask[n++] = bit;
    };
    for (int k = 0; k < n0; k++) add(sa[k], 1u);
    for (int k = 0; k < n1; k++) add(sb[k], 2u);
    return n;
}

void geometry_pair_count(Geometry &g) {
    g.nstrips = (g.nx + kStripW - 1) / kStripW;
    double c = g.method == kLinear ? (double)g.xcent : g.radius;
    g.nbins = (int)std::ceil((2 * c + 4) / kDepthBin) + 1;
    const int nu = (g.na + 1) / 2;
    g.strip_base.assign((size_t)nu + 1, 0);
    std::vector<int64_t> cnt(nu, 0);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nth = std::min<unsigned>(nth, (unsigned)nu);
    auto work = [&](unsigned t) {
        int s[4];
        unsigned m[4];
        for (int u = (int)t; u < nu; u += (int)nth) {
            int64_t n = 0;
            for (int j = 0; j < g.nvalid; j++) n += pair_strips(g, 2 * u, j, s, m);
            cnt[u] = n;
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++) pool.emplace_back(work, t);
    for (auto &th : pool) th.join();
    for (int u = 0; u < nu; u++) g.strip_base[u + 1] = g.strip_base[u] + cnt[u];
    g.strip_off.assign((size_t)nu * ((size_t)g.nstrips * g.nbins + 1), 0);
}

void geometry_pair_units(Geometry &g, int u0, int u1, std::vector<uint32_t> &order) {
    constexpr int G = kStripGroups + 1, G2 = G * G;
    const int nb = g.nstrips * g.nbins, nk = nb * G2;
    const int64_t e0 = g.strip_base[u0];
    order.resize((size_t)(g.strip_base[u1] - e0));
    double c = g.method == kLinear ? (double)g.xcent : g.radius;
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nth = std::min<unsigned>(nth, (unsigned)(u1 - u0));
    auto work = [&](unsigned t) {
        std::vector<int> cnt(nk + 1);
        std::vector<int> nstr(g.nvalid);
        std::vector<int> str(4 * (size_t)g.nvalid), key(4 * (size_t)g.nvalid);
        std::vector<unsigned> msk(4 * (size_t)g.nvalid);
        for (int u = u0 + (int)t; u < u1; u += (int)nth) {
            const int a = 2 * u;
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int j = 0; j < g.nvalid; j++) {
                int b = (int)std::floor((host_depth(g, a, j) + c + 1.0) / kDepthBin);
                b = std::min(std::max(b, 0), g.nbins - 1);
                const int p0 = host_first_tap(g, a, j);
                const int p1 = a + 1 < g.na ? host_first_tap(g, a + 1, j) : 0;
                nstr[j] = pair_strips(g, a, j, &str[4 * (size_t)j], &msk[4 * (size_t)j]);
                for (int k = 0; k < nstr[j]; k++) {
                    const size_t q = 4 * (size_t)j + k;
                    const int sj = str[q];
                    const int d0 = (msk[q] & 1u) ? p0 - (sj * kStripW - 2) : kStripGroups;
                    const int d1 = (msk[q] & 2u) ? p1 - (sj * kStripW - 2) : kStripGroups;
                    key[q] = (sj * g.nbins + b) * G2 + d0 * G + d1;
                    cnt[key[q] + 1]++;
                }
            }
            for (int i = 1; i <= nk; i++) cnt[i] += cnt[i - 1];
            int *off = &g.strip_off[(size_t)u * (nb + 1)];
            for (int i = 0; i <= nb; i++) off[i] = cnt[(size_t)i * G2];
            uint32_t *out = order.data() + (g.strip_base[u] - e0);
            for (int j = 0; j < g.nvalid; j++)
                for (int k = 0; k < nstr[j]; k++) {
                    const size_t q = 4 * (size_t)j + k;
                    out[cnt[key[q]]++] = (uint32_t)j | ((uint32_t)str[q] << kColBits) | (msk[q] << 30);
                }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++) pool.emplace_back(work, t);
    for (auto &th : pool) th.join();
}

// Staged strips (design 4).  Angle group G holds angles [G*kSgAngles,
// min(na, (G+1)*kSgAngles)); its central angle is the middle one.  A column
// belongs to the stage list of (G, s) when strip s is one of its strips at
// any of the group's angles, and to one entry of (G, s) per such angle, so
// every (angle, strip, column) contribution of design 2 is in exactly one
// entry.  The stage list is ordered by the central angle's first tap (bands
// one pixel wide along its rays) and inside a band by depth along the ray:
// a chunk of kSgChunk consecutive columns then spans a short stretch of one
// or two bands, where each angle of the group sees one or two first-tap
// offsets.  A (chunk, angle)'s entries are ordered by offset and every
// offset's run is padded with zero-weight entries (column -1) to a multiple
// of kStripBatch, so each batch of the walk has one offset; a chunk ends
// early when one more column would give an angle more than kSgEnt entries
// (groups of widely spaced angles see many offsets in a chunk).
namespace {

struct StagedItem {
    int s, band;
    double t;
    int j;
    int8_t d[kSgAngles];  // first-tap offset in strip s per angle of the group, -1 outside [0, kStripGroups)
};

// One angle group's chunks and entry counts, in table order: stage holds
// kSgChunk slots per chunk of strips 0, 1, ...; nchunk[s] the chunks of
// strip s; cnt kSgAngles entry counts per chunk (angle i's entries follow
// angle i - 1's in each chunk); nent their sum.  With a sink the entries
// themselves go straight to the upload arrays: column (-1: padding), angle,
// and slot | d << 8 | strip << 16.
struct StagedKey {
    double t;
    int j, q;
};

struct StagedGroup {
    std::vector<StagedItem> items, sorted, gather, band;
    std::vector<StagedKey> keys;
    std::vector<int64_t> sbase, bucket, pos;  // counting-sort offsets per strip and per band of a strip
    std::vector<int> stage, nchunk, cnt;
    int64_t nent = 0;
};

struct StagedSink {
    uint32_t *col, *ang, *meta;  // the group's first entry in each array
    int64_t cap;                 // the group's entries (from the counting pass)
};

// With keep, the group's ordered columns and chunk sizes are also stored
// there; with replay (from such a keep), they are taken from it instead of
// being ordered again.
void staged_group(const Geometry &g, int G, StagedGroup &o, const StagedSink *sink, StagedOrder *keep,
                  const StagedOrder *replay) {
    const int ns = g.nstrips, a0 = G * kSgAngles, na = std::min(kSgAngles, g.na - a0), ac = a0 + (na - 1) / 2;
    o.items.clear();
    if (replay) {
        o.items.resize(replay->col.size());
        for (size_t q = 0; q < o.items.size(); q++) {
            StagedItem &it = o.items[q];
            it.j = replay->col[q];
            std::memcpy(it.d, &replay->offs[q], sizeof(it.d));
        }
    } else {
    int sa[2], un[2 * kSgAngles], p0[kSgAngles];
    for (int j = 0; j < g.nvalid; j++) {
        int nu = 0;
        for (int i = 0; i < na; i++) {
            const int q = host_first_tap(g, a0 + i, j);
            p0[i] = q;
            const int n = strips_of(q, ns, sa);
            for (int k = 0; k < n; k++) {
                bool seen = false;
                for (int u = 0; u < nu; u++) seen = seen || un[u] == sa[k];
                if (!seen) un[nu++] = sa[k];
            }
        }
        if (!nu) continue;
        const int band = host_first_tap(g, ac, j);
        const double t = host_depth(g, ac, j);
        for (int u = 0; u < nu; u++) {
            StagedItem it{un[u], band, t, j, {}};
            for (int i = 0; i < kSgAngles; i++) {
                const int d = i < na ? p0[i] - (un[u] * kStripW - 2) : -1;
                it.d[i] = (int8_t)(d >= 0 && d < kStripGroups ? d : -1);
            }
            o.items.push_back(it);
        }
    }
    }  // ordered below unless replayed
    // order by (strip, band, depth, column): counting sorts by strip (one
    // write stream per strip) and, inside each strip, by band; then each
    // band (the columns along one ray of the central angle) by depth -- the
    // order of one full sort, in linear time but for the small buckets
    if (!replay && !o.items.empty()) {
        const size_t n = o.items.size();
        std::vector<int64_t> &pos = o.pos;
        o.sbase.assign((size_t)ns + 1, 0);
        for (const StagedItem &it : o.items) o.sbase[(size_t)it.s + 1]++;
        for (int s = 0; s < ns; s++) o.sbase[(size_t)s + 1] += o.sbase[(size_t)s];
        o.gather.resize(n);
        pos.assign(o.sbase.begin(), o.sbase.end() - 1);
        for (const StagedItem &it : o.items) o.gather[(size_t)pos[(size_t)it.s]++] = it;
        o.sorted.resize(n);
        for (int s = 0; s < ns; s++) {
            const int64_t s0 = o.sbase[(size_t)s], s1 = o.sbase[(size_t)s + 1];
            if (s0 == s1) continue;
            int bmin = o.gather[(size_t)s0].band, bmax = bmin;
            for (int64_t q = s0; q < s1; q++) {
                bmin = std::min(bmin, o.gather[(size_t)q].band);
                bmax = std::max(bmax, o.gather[(size_t)q].band);
            }
            const int nb = bmax - bmin + 1;
            o.bucket.assign((size_t)nb + 1, 0);
            o.bucket[0] = s0;
            for (int64_t q = s0; q < s1; q++) o.bucket[(size_t)(o.gather[(size_t)q].band - bmin) + 1]++;
            for (int k = 0; k < nb; k++) o.bucket[(size_t)k + 1] += o.bucket[(size_t)k];
            pos.assign(o.bucket.begin(), o.bucket.end() - 1);
            for (int64_t q = s0; q < s1; q++) {
                const StagedItem &it = o.gather[(size_t)q];
                o.sorted[(size_t)pos[(size_t)(it.band - bmin)]++] = it;
            }
            // each band by (depth, column) through 16-byte keys (the column
            // is unique in a band of one strip), the items gathered back in
            // that order
            for (int k = 0; k < nb; k++) {
                const int64_t b0 = o.bucket[(size_t)k], nbk = o.bucket[(size_t)k + 1] - b0;
                if (nbk < 2) continue;
                o.keys.resize((size_t)nbk);
                for (int64_t q = 0; q < nbk; q++) {
                    const StagedItem &it = o.sorted[(size_t)(b0 + q)];
                    o.keys[(size_t)q] = StagedKey{it.t, it.j, (int)q};
                }
                std::sort(o.keys.begin(), o.keys.end(), [](const StagedKey &x, const StagedKey &y) {
                    if (x.t != y.t) return x.t < y.t;
                    return x.j < y.j;
                });
                o.band.resize((size_t)nbk);
                for (int64_t q = 0; q < nbk; q++) o.band[(size_t)q] = o.sorted[(size_t)(b0 + o.keys[(size_t)q].q)];
                std::copy(o.band.begin(), o.band.end(), o.sorted.begin() + b0);
            }
        }
        o.items.swap(o.sorted);
    }
    o.stage.clear();
    o.nchunk.assign(ns, 0);
    o.cnt.clear();
    o.nent = 0;
    int runs[kStripGroups + 1];
    int fill[kSgAngles][kStripGroups];  // per angle and offset: entries of the chunk being filled
    if (keep) {
        keep->col.resize(o.items.size());
        keep->offs.resize(o.items.size());
        for (size_t q = 0; q < o.items.size(); q++) {
            keep->col[q] = o.items[q].j;
            std::memcpy(&keep->offs[q], o.items[q].d, sizeof(o.items[q].d));
        }
        keep->nq.clear();
        keep->nchunk.assign(ns, 0);
    }
    auto up = [](int c) { return (c + kStripBatch - 1) / kStripBatch * kStripBatch; };
    size_t idx = 0, kc = 0;
    for (int s = 0; s < ns; s++) {
        size_t n = 0;
        if (replay) {
            for (int c = 0; c < replay->nchunk[(size_t)s]; c++) n += replay->nq[kc + (size_t)c];
        } else {
            while (idx + n < o.items.size() && o.items[idx + n].s == s) n++;
        }
        for (size_t q0 = idx; q0 < idx + n;) {
            // the chunk takes up to kSgChunk columns, fewer when one more
            // column would give an angle more than kSgEnt entries once its
            // offset runs are padded to whole batches (widely spaced angles)
            int nq = 0;
            if (replay) nq = replay->nq[kc++];
            std::fill(&fill[0][0], &fill[0][0] + kSgAngles * kStripGroups, 0);
            int padded[kSgAngles] = {};  // per angle: its entries with every offset's run padded
            while (!replay && nq < kSgChunk && q0 + nq < idx + n) {
                const StagedItem &it = o.items[q0 + nq];
                bool fits = true;
                for (int i = 0; i < na && fits; i++) {
                    const int d = it.d[i];
                    if (d < 0) continue;
                    fits = padded[i] - up(fill[i][d]) + up(fill[i][d] + 1) <= kSgEnt;
                }
                if (!fits) break;
                for (int i = 0; i < na; i++) {
                    const int d = it.d[i];
                    if (d >= 0) {
                        padded[i] += up(fill[i][d] + 1) - up(fill[i][d]);
                        fill[i][d]++;
                    }
                }
                nq++;
            }
            o.nchunk[s]++;
            if (keep) {
                keep->nq.push_back((uint8_t)nq);
                keep->nchunk[(size_t)s]++;
            }
            for (int q = 0; q < kSgChunk; q++) o.stage.push_back(q < nq ? o.items[q0 + q].j : -1);
            // (chunk, angle)'s entries by offset, each offset's columns in
            // stage order (a stable counting sort), each run padded to whole
            // batches with zero-weight entries
            for (int i = 0; i < kSgAngles; i++) {
                int tot = 0;
                if (i < na) {
                    std::fill(runs, runs + kStripGroups + 1, 0);
                    for (int q = 0; q < nq; q++) {
                        const int d = o.items[q0 + q].d[i];
                        if (d >= 0) runs[d]++;
                    }
                    int at[kStripGroups];
                    for (int d = 0; d < kStripGroups; d++) {
                        at[d] = tot;
                        tot += up(runs[d]);
                    }
                    if (sink && o.nent + tot <= sink->cap) {
                        uint32_t *col = sink->col + o.nent, *ang = sink->ang + o.nent, *meta = sink->meta + o.nent;
                        const uint32_t sd = (uint32_t)s << 16;
                        for (int q = 0; q < nq; q++) {
                            const StagedItem &it = o.items[q0 + q];
                            const int d = it.d[i];
                            if (d < 0) continue;
                            const int k = at[d]++;
                            col[k] = (uint32_t)it.j;
                            ang[k] = (uint32_t)(a0 + i);
                            meta[k] = (uint32_t)q | ((uint32_t)d << 8) | sd;
                        }
                        for (int d = 0, end = 0; d < kStripGroups; d++) {
                            end += up(runs[d]);
                            for (int k = at[d]; k < end; k++) {
                                col[k] = 0xffffffffu;
                                ang[k] = (uint32_t)(a0 + i);
                                meta[k] = ((uint32_t)d << 8) | sd;
                            }
                        }
                    }
                }
                o.cnt.push_back(tot);
                o.nent += tot;
            }
            q0 += nq;
        }
        idx += n;
    }
}

unsigned staged_threads(int ngroups) {
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    return std::min<unsigned>(nth, (unsigned)std::max(1, ngroups));
}

}  // namespace

int geometry_staged_count(Geometry &g) {
    g.nstrips = (g.nx + kStripW - 1) / kStripW;
    const int ns = g.nstrips, ngroups = (g.na + kSgAngles - 1) / kSgAngles;
    std::vector<int64_t> chunks((size_t)ngroups * ns, 0), ent(ngroups, 0);
    std::vector<int> most(ngroups, 0);
    const unsigned nth = staged_threads(ngroups);
    g.sg_order.assign((size_t)ngroups, StagedOrder());
    std::atomic<int64_t> kept_bytes{0};
    const long pages = sysconf(_SC_PHYS_PAGES), page = sysconf(_SC_PAGE_SIZE);
    const int64_t keep_budget = pages > 0 && page > 0 ? (int64_t)pages * page / 8 : 0;
    auto work = [&](unsigned t) {
        StagedGroup o;
        for (int G = (int)t; G < ngroups; G += (int)nth) {
            // keep the group's order for the filling pass while the kept
            // orders stay within an eighth of the host's memory
            StagedOrder &k = g.sg_order[(size_t)G];
            staged_group(g, G, o, nullptr, &k, nullptr);
            const int64_t kb = (int64_t)(k.col.size() * 12 + k.nq.size() + k.nchunk.size() * 4);
            if (kept_bytes.fetch_add(kb) + kb <= keep_budget) {
                k.kept = true;
            } else {
                kept_bytes.fetch_sub(kb);
                k = StagedOrder();
            }
            for (int s = 0; s < ns; s++) chunks[(size_t)G * ns + s] = o.nchunk[s];
            ent[G] = o.nent;
            for (int c : o.cnt) most[G] = std::max(most[G], c);
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++) pool.emplace_back(work, t);
    for (auto &th : pool) th.join();
    g.sg_chunk_base.assign((size_t)ngroups * ns + 1, 0);
    for (size_t k = 0; k < chunks.size(); k++) g.sg_chunk_base[k + 1] = g.sg_chunk_base[k] + chunks[k];
    g.sg_group_ent.assign((size_t)ngroups + 1, 0);
    for (int G = 0; G < ngroups; G++) g.sg_group_ent[G + 1] = g.sg_group_ent[G] + ent[G];
    return *std::max_element(most.begin(), most.end());
}

bool geometry_staged_groups(Geometry &g, int G0, int G1, int *stage, int64_t *off, uint32_t *col, uint32_t *ang,
                            uint32_t *meta) {
    const int ns = g.nstrips;
    const int64_t K0 = g.sg_chunk_base[(size_t)G0 * ns];
    const int64_t E0 = g.sg_group_ent[G0];
    const unsigned nth = staged_threads(G1 - G0);
    std::atomic<bool> bad{false};
    auto work = [&](unsigned t) {
        StagedGroup o;
        for (int G = G0 + (int)t; G < G1; G += (int)nth) {
            const int64_t o0 = g.sg_group_ent[G] - E0;
            const int64_t cap = g.sg_group_ent[G + 1] - g.sg_group_ent[G];
            const StagedSink sink{col + o0, ang + o0, meta + o0, cap};
            StagedOrder *kept = (size_t)G < g.sg_order.size() && g.sg_order[(size_t)G].kept ? &g.sg_order[(size_t)G] : nullptr;
            staged_group(g, G, o, &sink, nullptr, kept);
            if (kept) *kept = StagedOrder();  // used once
            if (o.nent != cap) bad = true;
            const int64_t Kg = g.sg_chunk_base[(size_t)G * ns] - K0;
            const int64_t nk = g.sg_chunk_base[(size_t)(G + 1) * ns] - g.sg_chunk_base[(size_t)G * ns];
            if ((int64_t)o.stage.size() != nk * kSgChunk || (int64_t)o.cnt.size() != nk * kSgAngles) {
                bad = true;
                continue;
            }
            std::copy(o.stage.begin(), o.stage.end(), stage + Kg * kSgChunk);
            int64_t e = g.sg_group_ent[G];
            for (size_t k = 0; k < o.cnt.size(); k++) {
                off[(size_t)Kg * kSgAngles + k] = e;
                e += o.cnt[k];
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++) pool.emplace_back(work, t);
    for (auto &th : pool) th.join();
    return !bad;
}

bool geometry_staged_groups(Geometry &g, int G0, int G1, std::vector<int> &stage, std::vector<int64_t> &off,
                            std::vector<uint32_t> &col, std::vector<uint32_t> &ang, std::vector<uint32_t> &meta) {
    const int ns = g.nstrips;
    const int64_t K0 = g.sg_chunk_base[(size_t)G0 * ns], K1 = g.sg_chunk_base[(size_t)G1 * ns];
    const int64_t n = g.sg_group_ent[G1] - g.sg_group_ent[G0];
    stage.assign((size_t)(K1 - K0) * kSgChunk, -1);
    off.assign((size_t)(K1 - K0) * kSgAngles, 0);
    col.resize((size_t)n);
    ang.resize((size_t)n);
    meta.resize((size_t)n);
    return geometry_staged_groups(g, G0, G1, stage.data(), off.data(), col.data(), ang.data(), meta.data());
}

// Stable counting sort of the contributing columns of angles [a0, a1) by
// their first tap; order receives (a1-a0)*nvalid column indices.  Also
// fills pixstart: for each q, the first entry whose first tap is >=
// q + kTapBase.
void geometry_sort_angles(Geometry &g, int a0, int a1, std::vector<int> &order) {
    const int nv = g.nvalid;
    order.resize((size_t)(a1 - a0) * nv);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    nth = std::min<unsigned>(nth, (unsigned)(a1 - a0));
    auto work = [&](unsigned t) {
        std::vector<int> key(nv);
        std::vector<int> count;
        for (int a = a0 + (int)t; a < a1; a += (int)nth) {
            int kmin = 1 << 30, kmax = -(1 << 30);
            for (int j = 0; j < nv; j++) {
                key[j] = host_first_tap(g, a, j);
                kmin = std::min(kmin, key[j]);
                kmax = std::max(kmax, key[j]);
            }
            count.assign((size_t)(kmax - kmin + 2), 0);
            for (int j = 0; j < nv; j++) count[key[j] - kmin + 1]++;
            for (size_t k = 1; k < count.size(); k++) count[k] += count[k - 1];
            int *out = order.data() + (size_t)(a - a0) * nv;
            for (int j = 0; j < nv; j++) out[count[key[j] - kmin]++] = j;
            int *ps = &g.pixstart[(size_t)a * g.npix];
            int ii = 0;
            for (int q = 0; q < g.npix; q++) {
                while (ii < nv && key[out[ii]] < q + kTapBase) ii++;
                ps[q] = ii;
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++) pool.emplace_back(work, t);
    for (auto &th : pool) th.join();
}

}  // namespace arec
