// Internal declarations shared by the translation units of libarec3d_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "arec3d_amd.h"

namespace arec {

// Projector methods, numbered like the reference's "pmethod" input key
// (AREC3D/AREC3D/AREC3D.cpp:145-169).
enum Method { kLinear = 1, kKB = 2, kSQ = 3 };

// Per-angle constants.  c and s are the FP32 cos/sin of the FP32 angle,
// widened to FP64: the reference evaluates cos(angles[ia]) on a float
// argument, which selects the float overload (arecproject.cpp:71-72,
// 393-394, 599-600).  a, b and lmax are sqpoints (arecproject.h:13-25),
// computed from FP64 cos/sin because sqpoints takes a double.
struct AngleConst {
    double c, s;
    double a, b, lmax;
};

// Columns of one back-projection tile: up to 256 voxel columns whose
// coordinates lie in one 16x16 box (one column per thread).
constexpr int kTileCols = 256;
constexpr int kTileBox = 16;

// Entry of the forward table: voxel columns of one angle in order of their
// first detector tap.  x = column index (low 22 bits) | increase of the first
// tap over the previous entry (high 10 bits); y,z,w = tap weights.
constexpr int kColBits = 22;
constexpr int kMaxCols = 1 << kColBits;
constexpr int kMaxTapStep = 1023;
constexpr int kTapBase = -4;  // first tap of the virtual entry before entry 0
constexpr int kFwdAngles = 8;  // angles (wavefronts) per forward workgroup
constexpr int kFwdOut = 8;     // pixels per staged output segment (32 bytes)
constexpr int kFwdPhase = 2;   // detector pixels per lockstep phase of a forward workgroup
// Depth-strip forward design: detector strips of kStripW pixels, entries of
// a strip walked in depth bins of kDepthBin pixels along the ray, and inside
// a bin ordered by their first tap relative to the strip (kStripGroups
// values: a column's three taps start at most 2 pixels before the strip).
constexpr int kStripW = 16;  // 23.8 ms at 512^3/720 (8: 24.8, 12: 24.3, 20: 27.8)
constexpr int kDepthBin = 16;
constexpr int kStripGroups = kStripW + 2;
constexpr int kStripAngles = 4;  // angles (wavefronts) per depth-strip workgroup
constexpr int kStripBatch = 4;   // depth-strip table entries whose loads are in flight together (8: 3% slower)
// Staged-strip forward design (4): workgroups of kSgAngles neighbouring
// angles walk one detector strip (kStripW pixels) through chunks of
// kSgChunk voxel columns staged in LDS.  A group's columns that reach the
// strip at any of its angles are listed once (the stage list), ordered by
// the central angle's first tap and then by depth along its ray, and cut
// into chunks; each angle's entries reference a chunk slot.
constexpr int kSgAngles = 8;
constexpr int kSgChunk = 32;
constexpr int kSgEnt = 48;  // most entries of one (chunk, angle), padding included
// Entry word of the staged table: the slot's byte offset in an LDS column
// buffer (slot * 64 * RPL * 4 < 1 << kSgOffShift) plus the first tap's offset
// d << kSgOffShift.
constexpr int kSgOffShift = 16;
// The back projector's tap table is built when it takes at most this much.
// Back-projector tile shapes (rows per workgroup x angles staged per LDS
// pass) a geometry may select; shape 0 is the default (12.5 ms at 512^3/720
// against 14.3 ms for the other two, which suit other row counts).
constexpr int kBpShapes = 3;
constexpr int kBpShapeRows[kBpShapes] = {32, 16, 64};
constexpr int kBpShapeAngles[kBpShapes] = {2, 4, 2};
constexpr int64_t kMaxBackTableBytes = 16ll << 30;
// The depth-strip forward table is built (on request) up to this size.
constexpr int64_t kMaxStripTableBytes = 32ll << 30;

// One angle group's design-4 host ordering, kept by the counting pass for the
// filling pass (which then skips ordering the group again): its staged
// columns in stage order with their per-angle offsets in the strip (int8
// each, -1 outside, packed), the columns of each chunk and the chunks of each
// strip.
struct StagedOrder {
    bool kept = false;
    std::vector<int32_t> col;
    std::vector<uint64_t> offs;
    std::vector<uint8_t> nq;
    std::vector<int> nchunk;
};

struct Geometry {
    int method = kSQ;
    int nx = 0, ny = 0, na = 0;  // detector width, rows (height), angles
    double radius = 0;
    int ncyl = 0;     // cylinder width round(2r+1)
    int nrays = 0;    // voxel columns
    int nvalid = 0;   // columns that contribute (linear: nrays-2)
    int xcent = 0;    // linear pair's integer centre (int)r + 1
    int64_t nnz = 0, nproj = 0;
    std::vector<int> cord;        // (ix, iz) per column, iz-major
    std::vector<int> coord;       // (X, Z) used by the weight function
    std::vector<float> angles;
    std::vector<AngleConst> ang;
    int ntiles = 0;
    std::vector<int> tile_cols;   // ntiles * 64, -1 = empty
    std::vector<int> tile_org;    // ntiles * 2 box origin
    // device copies
    AngleConst *d_ang = nullptr;
    int *d_coord = nullptr;
    int *d_tile_cols = nullptr;
    int *d_tile_org = nullptr;
    uint4 *d_fwd = nullptr;       // na * nvalid entries
    // back projector's taps in tile order, ((tile * na) + a) * kTileCols + slot:
    // {first tap, w0, w1, w2}; nullptr when too large (taps computed on the fly)
    uint4 *d_btab = nullptr;
    // forward sweep split into detector segments: per (angle, segment)
    // {first entry, end entry, first tap of the entry before the first}
    int nseg = 1, segw = 0, fwd_rpl = 1, fwd_chunks = 1, fwd_min_groups = 1024;
    std::vector<int> seg;         // na * nseg * 4
    int *d_seg = nullptr;
    // first table entry whose first tap is >= q, for q = kTapBase .. nx+3
    // (index q - kTapBase); the forward kernel's lockstep phases use it
    int npix = 0;
    std::vector<int> pixstart;    // na * npix
    int *d_pixstart = nullptr;
    int *d_err = nullptr;         // geometry self-check flag
    // Forward design: 1 = detector sweep (the table above, pixstart, seg);
    // 2 = depth strips (built on request).  Per angle a, the strip table
    // holds, for every strip s (pixels [s*kStripW, (s+1)*kStripW)), the
    // entries whose first tap lies in [s*kStripW - 2, (s+1)*kStripW),
    // ordered by depth bin b along the ray and inside a bin by d = first
    // tap - (s*kStripW - 2) in [0, kStripGroups); entry = {column | d << 22,
    // w0, w1, w2}.  strip_off[a][s*nbins + b] is the first entry of (s, b)
    // relative to strip_base[a].
    // Design 3 (depth strips, angle pairs) uses the same fields per pair of
    // angles (strip_unit = 2): an entry is two uint4, {column | d0 << 22 |
    // d1 << 27, w0[3]} and {w1[3], 0}, d = kStripGroups for an angle absent
    // from the entry, and entries are ordered by (bin, d0, d1).
    int fwd_design = 1;
    int nstrips = 0, nbins = 0, strip_rpl = 1, strip_unit = 1;
    std::vector<int64_t> strip_base;  // na + 1
    std::vector<int> strip_off;       // na * (nstrips * nbins + 1)
    uint4 *d_strip = nullptr;
    int64_t *d_strip_base = nullptr;
    int *d_strip_off = nullptr;
    // Design 4 (staged strips; rows per lane in strip_rpl).  Angle group G
    // holds angles [G*kSgAngles, ...); chunk K of (G, s) is global chunk
    // sg_chunk_base[G*nstrips + s] + k.  sg_stage[K*kSgChunk + q] is the
    // column in slot q of chunk K (-1: empty).  Entries are stored chunk
    // after chunk, inside a chunk angle after angle (i = a - G*kSgAngles),
    // inside an angle by d: sg_off[K*kSgAngles + i] is the first entry of
    // (K, i), sg_off[nchunks*kSgAngles] the total.  Entry = {slot | d << 8,
    // w0, w1, w2}, d = first tap - (s*kStripW - 2).
    std::vector<int64_t> sg_chunk_base;  // ngroups * nstrips + 1
    std::vector<int64_t> sg_group_ent;   // ngroups + 1: first entry of each group
    std::vector<StagedOrder> sg_order;   // per group, while the filling pass has not used it
    int *d_sg_stage = nullptr;
    int64_t *d_sg_chunk_base = nullptr;
    int64_t *d_sg_off = nullptr;
    uint4 *d_sg_tab = nullptr;
    int sg_rows4 = 0;  // column stride (bytes) the staged entry words are encoded for
    int bp_shape = 0;  // back-projector tile shape (index into kBpShapeRows)
    bool device_ready = false;
};

// Host geometry (geometry.cpp).
int cyl_mask(double radius, std::vector<int> *cord);
int geometry_build_host(Geometry &g, int method, int nx, int ny, double radius,
                        const float *angles, int na, std::string *err);
// First detector tap of column j at angle a, the same arithmetic as the
// device weight function (used to order the forward table).
int host_first_tap(const Geometry &g, int a, int j);
void geometry_sort_angles(Geometry &g, int a0, int a1, std::vector<int> &order);
bool fwd_rpl_allowed(int ny, int rpl);
bool geometry_plan_forward(Geometry &g, int rpl, int min_groups);
void geometry_segments(Geometry &g);
// Depth-strip ordering (design 2): strip_count fills nstrips, nbins and
// strip_base; strip_angles fills strip_off for angles [a0, a1) and order
// with (column | strip << 22) per entry, angle after angle.
void geometry_strip_count(Geometry &g);
void geometry_strip_angles(Geometry &g, int a0, int a1, std::vector<uint32_t> &order);
// Angle-pair ordering (design 3): the same fields per pair of angles
// (strip_base has (na + 1) / 2 + 1 entries); order gets (column | strip << 22
// | present flags << 30) per entry, pair after pair.
void geometry_pair_count(Geometry &g);
void geometry_pair_units(Geometry &g, int u0, int u1, std::vector<uint32_t> &order);
// Staged-strip ordering (design 4).  staged_count fills nstrips,
// sg_chunk_base and sg_group_ent (host counting pass) and returns the most
// entries of one (chunk, angle), padding included.  staged_groups fills,
// for angle groups [G0, G1): stage (all chunks of those groups, kSgChunk
// slots each, starting at chunk sg_chunk_base[G0*nstrips]), off (the groups'
// sg_off values, kSgAngles per chunk) and, per entry of the groups in table
// order, col (column; 0xffffffff for a zero-weight padding entry), ang
// (angle) and meta (slot | d << 8 | strip << 16); false if a group's
// entries do not come out as the counting pass sized them (nothing is
// written past a group's share then).
int geometry_staged_count(Geometry &g);
bool geometry_staged_groups(Geometry &g, int G0, int G1, std::vector<int> &stage, std::vector<int64_t> &off,
                            std::vector<uint32_t> &col, std::vector<uint32_t> &ang, std::vector<uint32_t> &meta);
// The same into caller buffers of at least the groups' chunk and entry counts
// (the stage slots of chunks without a column are written too).
bool geometry_staged_groups(Geometry &g, int G0, int G1, int *stage, int64_t *off, uint32_t *col, uint32_t *ang,
                            uint32_t *meta);
// Segment table of segments of width segw (nseg of them) for g's forward table.
void segment_table(const Geometry &g, int segw, int nseg, std::vector<int> &seg);

// Arguments of one forward / back application (device code reads them).
struct FwdArgs {
    const float *x;
    float *p;
    const uint4 *table;
    const int4 *seg;
    const int *pixstart;
    double *partials;
    int npix, nvalid, nseg, segw, nx, ny, na, nchunks;
};
struct BpArgs {
    const float *p;
    float *out;
    const float *bvec;
    const uint4 *btab;
    const int *tile_cols, *tile_org, *coord;
    const AngleConst *ang;
    double *partials;
    int *err;
    double radius;
    int xcent, nx, ny, na, nrowblk;
};
// One problem of a persistent batched back projection: its arguments, a
// skip flag (nullptr: never skipped) and the index of its first work item.
struct BatchBp {
    BpArgs a;
    const int *skip;
    int item0, pad;
};

// One problem of a batched staged-strip forward projection (design 4): its
// tables and sizes, a skip flag (nullptr: never skipped), and its first
// workgroup in the launch divided by 8 (every problem takes a multiple of 8
// workgroups, so that the 8 XCDs get equal shares of it).
struct BatchFwd {
    const float *x;
    float *p;
    const uint4 *tab;
    const int *stage;
    const int64_t *chunk_base, *off;
    const int *skip;
    int nstrips, nx, ny, na, nchunks, nblk;  // nblk: the problem's workgroups before padding
    int64_t blk8;
};

// Device setup and operators (projector.hip).
int geometry_upload(Geometry &g, hipStream_t st, std::string *err);
void geometry_free_device(Geometry &g);
// Re-plan the forward launch (rows per lane, minimum workgroup count) of an
// uploaded geometry: recomputes and uploads the segment table.
int geometry_replan_forward(Geometry &g, int rpl, int min_groups, hipStream_t st);
// Select the forward design (1 sweep, 2 depth strips; 2 builds its table on
// first use) and rows per lane.
int geometry_set_forward(Geometry &g, int design, int rpl, int min_groups, hipStream_t st);
// Modes of the back projector's epilogue.
enum BpMode { kBpStore = 0, kBpSirtGrad = 1, kBpStoreNorm = 2 };
int launch_forward(const Geometry &g, const float *x, float *p, double *partials, hipStream_t st,
                   const int *skip);
int launch_backward(const Geometry &g, const float *p, float *x, int mode, const float *bvec,
                    double *partials, hipStream_t st, const int *skip);
int64_t forward_partial_count(const Geometry &g);
int64_t backward_partial_count(const Geometry &g);
// Work items (workgroup jobs) of one application and their arguments.
int64_t forward_items(const Geometry &g);
int64_t backward_items(const Geometry &g);
FwdArgs fwd_args(const Geometry &g, const float *x, float *p, double *partials);
BpArgs bp_args(const Geometry &g, const float *p, float *x, const float *bvec, double *partials);
// Persistent back projection over nprob problems (device array of
// descriptors, item0 ascending, `total` items, all with the same method and
// tile shape); d_counter is an int of scratch.
int launch_backward_batch(int method, int mode, int shape, const BatchBp *d_probs, int nprob, int64_t total,
                          int *d_counter, hipStream_t st);
// Select the back projector's tile shape (0 .. kBpShapes - 1).
int geometry_set_back(Geometry &g, int shape);
// Descriptor of g's staged-strip forward (g on design 4) at workgroup
// blk8 * 8 of a batched launch; returns the workgroups it takes (a
// multiple of 8).
int64_t forward_batch_item(const Geometry &g, const float *x, float *p, const int *skip, int64_t blk8,
                           BatchFwd &d);
// One launch of the staged-strip forward over several problems, all at rpl
// rows per lane: nblk workgroups (a multiple of 8), owner[i] = the problem
// of workgroups 8i .. 8i + 7 (device arrays).
int launch_forward_batch(int method, int rpl, const BatchFwd *d_probs, const int *d_owner, int64_t nblk,
                         hipStream_t st);

// One problem of a (batched) SIRT solve (solver.hip).
struct SirtProblem {
    const Geometry *g;
    const float *b;
    float *x;
    float lam;
};
int sirt_batch_run(const std::vector<SirtProblem> &P, hipStream_t st, int maxit, float tol, int warm,
                   double reject_above, int check_pass, arec_hist *hist, int hist_cap, int *nhist,
                   float *lam_out, bool use_graph);
int sirt_run(const Geometry &g, hipStream_t st, const float *b, float *x, float lam, int maxit, float tol,
             int warm, double reject_above, int check_pass, arec_hist *hist, int hist_cap, int *nhist,
             float *lam_out, bool use_graph);

int cgls_run(const Geometry &g, hipStream_t st, const float *b, const float *w, float *x, int warm,
             int maxit, float tol, arec_hist *hist, int hist_cap, int *nhist, bool use_graph);
int init_weights(const float *b, float *w, int64_t nx, int64_t ny, int64_t nz, hipStream_t st);
int soft_noneg(float *x, int64_t n, hipStream_t st);
int dot(const float *a, const float *b, int64_t n, double *out, hipStream_t st);
int reduce_max(const float *x, int64_t n, float *out, hipStream_t st);
int copy(float *dst, const float *src, int64_t n, hipStream_t st);
int scale(float *x, float a, int64_t n, hipStream_t st);
// hipGraph replay of solver passes unless AREC_NO_GRAPH is set (capi.cpp)
bool use_graphs();

// Alignment (align.hip): per-view cross-correlation maps and/or sub-pixel
// peaks; Fourier shift of every view by sgn*(sx, sy).
// Device address of the no-allocation scratch array for `device` (which must
// be current), looked up once per device; cc_views_noalloc runs cc_views's
// kernels on it (sx, sy required).
int noalloc_setup(int device);
int cc_views_noalloc(int device, const float *proj, const float *meas, int nx, int ny, int nv, float *sx, float *sy,
                     hipStream_t st);
int cc_views(const float *proj, const float *meas, int nx, int ny, int nv, float *cmap_out, float *sx,
             float *sy, hipStream_t st);
int shift_views(float *data, int nx, int ny, int nv, const float *sx, const float *sy, float sgn,
                hipStream_t st);
// Preprocessing: -log of positive values (else 0); Moisan's periodic
// decomposition of every view in place.
int take_log(float *x, int64_t n, hipStream_t st);
int periodic_views(float *data, int nx, int ny, int nv, hipStream_t st);
// Per-view moments {sum p, sum u p, sum y p} (FP64, deterministic) into
// mom_dev[3 * nv].
int view_moments(const float *views, int nx, int ny, int nv, double *mom_dev, hipStream_t st);
// Rotational alignment (rotate.hip); angles are host arrays, in degrees.
int rot_check(int nx, int ny, int nv, const char *what);
int rot_cc_views(const float *img1, const float *img2, int nx, int ny, int nv, int r2, float *angles,
                 hipStream_t st);
int rotate_views(float *views, int nx, int ny, int nv, float *angles, int rule, hipStream_t st);
int rotate_views_safe(float *views, const float *ref, int nx, int ny, int nv, float *angles, int *kept,
                      double *mean_cc, hipStream_t st);
// Projection matching (matching.cpp).  flags bit 0: x shifts from the views'
// centroids (centroid_x_shifts) instead of cross-correlation; bit 1: the
// rotational step every third outer loop (the driver's align = 2).
constexpr int kMatchXCentroid = AREC_MATCH_X_CENTROID;
constexpr int kMatchRotate = AREC_MATCH_ROTATE;
int centroid_x_shifts(const Geometry &g, hipStream_t st, const float *views, float *sx_host);
int match_run(const Geometry &g, hipStream_t st, const float *views0, float *x, float lam, int outer,
              int maxit, float tol, int cc_passes, int method, int flags, float *sx_host, float *sy_host,
              float *rot_host, arec_hist *loop_hist);

// Fill with a constant by a kernel (graph-capture safe); skip as above.
int fill(float *x, float v, int64_t n, hipStream_t st, const int *skip);

// Error plumbing (capi.cpp).
void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);
#define AREC_HIP(call)                                   \
    do {                                                 \
        hipError_t e__ = (call);                         \
        if (e__ != hipSuccess) return arec::hip_fail(e__, #call); \
    } while (0)

}  // namespace arec

struct arec_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
};

struct arec_geom {
    arec::Geometry g;
};
