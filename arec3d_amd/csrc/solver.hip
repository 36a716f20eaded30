// Device-resident SIRT (Landweber) and CGLS, gfx950.
//
// The reference's solver loops (arecsirt.cpp:11-191, areccgls.cpp:12-243,
// areccgls_stat.cpp:49-275) make every control decision on the host from
// FP64 norms.  Here those decisions run in one-workgroup control kernels that
// read the norms on the device, so one pass of a loop is a fixed sequence of
// kernels: it is captured once as a hipGraph and replayed, and the host only
// polls the "done" flag between batches of replays.  Once the solver has
// stopped, further replays are no-ops (every kernel checks a skip flag).
//
// Norms follow the reference's arithmetic: squares of FP32 values are formed
// in FP32 or FP64 exactly where the reference forms them, and always summed
// in FP64 (per-thread FP64 partials, wavefront shuffles, then one workgroup
// summing the per-block partials in a fixed order, so results are
// deterministic from run to run).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cmath>
#include <vector>

#include "arec_internal.hpp"

namespace arec {

constexpr int kVecThreads = 256;
constexpr int kVecBlocks = 1024;

struct SirtState {
    int iter, restarts, done, first, zero_x, update, skip_all, skip_proj;
    int nhist, hist_cap, maxit, warm;
    int npass, check_pass;  // passes decided so far; pass at which a warm start is checked
    float lam, tol;
    double bnorm2, bnorm, relnrm0, reject_above;
};

struct CglsState {
    int iter, done, skip, best, iterbest, nhist, hist_cap, maxit;
    double gamma, gamma0, alpha, beta, rnorm0, rnorm_min, tol;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum of n FP64 partials by one workgroup of kVecThreads, fixed order.  The
// result is valid in every thread.
__device__ double block_sum(const double *__restrict__ part, int64_t n) {
    __shared__ double s[kVecThreads / 64];
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kVecThreads) v += part[i];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < kVecThreads / 64; w++) t += s[w];
    return t;
}

// Per-block partial of a grid-stride reduction (written to out[blockIdx.x]).
__device__ __forceinline__ void block_partial(double v, double *__restrict__ out) {
    __shared__ double s[kVecThreads / 64];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kVecThreads / 64; w++) t += s[w];
        out[blockIdx.x] = t;
    }
}

#define GRID_STRIDE(i, n) \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

__global__ void fill_kernel(float *__restrict__ x, float v, int64_t n, const int *skip) {
    if (skip && *skip) return;
    GRID_STRIDE(i, n) x[i] = v;
}

static dim3 vec_grid(int64_t n) {
    int64_t b = (n + kVecThreads - 1) / kVecThreads;
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(b, kVecBlocks)));
}

int fill(float *x, float v, int64_t n, hipStream_t st, const int *skip) {
    if (n <= 0) return 0;
    fill_kernel<<<vec_grid(n), kVecThreads, 0, st>>>(x, v, n, skip);
    AREC_HIP(hipGetLastError());
    return 0;
}

__device__ __forceinline__ void push_hist(arec_hist *h, int cap, int *n, int iter, int event,
                                          double relnrm, double rnorm) {
    if (*n < cap) {
        h[*n].iter = iter;
        h[*n].event = event;
        h[*n].relnrm = relnrm;
        h[*n].rnorm = rnorm;
    }
    (*n)++;
}

// ------------------------------------------------------------------ SIRT --
// The SIRT kernels run one or many independent problems: problem k is
// blockIdx.x (control) or blockIdx.y (vector passes), with its own state,
// vectors and history.  A vector pass of problem k uses v[k].grid
// workgroups (its single-problem grid), so every problem's partial sums have
// the same order in a batch as alone.

struct SirtVec {
    float *x;
    const float *bdata;  // A^T b
    float *grad;
    const double *bp_part;  // back projector partials, n_bp of them
    arec_hist *hist;
    int64_t nnz, n_bp;
    int grid, pad;
};

__global__ void sirt_begin(SirtState *S) {
    SirtState *st = S + blockIdx.x;
    st->skip_all = st->done;
    st->skip_proj = st->done || (st->first && !st->warm);
}

// First pass after the start or a restart (arecsirt.cpp:96-108):
// grad = bdata; partials hold sum (double)(float)(b*b) in [0, G) and
// sum grad*(double)grad in [G, 2G) (G = the problem's grid).
__global__ __launch_bounds__(kVecThreads) void sirt_first_pass(const SirtVec *__restrict__ V,
                                                               double *__restrict__ fp_part,
                                                               const SirtState *__restrict__ S) {
    const int k = blockIdx.y;
    const SirtState *st = S + k;
    const SirtVec v = V[k];
    if (st->skip_all || !st->first || (int)blockIdx.x >= v.grid) return;
    const int warm = st->warm;  // warm start: grad comes from the projection pass
    double sb = 0.0, sg = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kVecThreads + threadIdx.x; i < v.nnz; i += (int64_t)v.grid * kVecThreads) {
        float b = v.bdata[i];
        float s = b * b;
        sb += (double)s;
        sg += (double)b * (double)b;
        if (!warm) v.grad[i] = b;
    }
    double *part = fp_part + (size_t)k * 2 * kVecBlocks;
    block_partial(sb, part);
    block_partial(sg, part + v.grid);
}

// Decisions of one pass (arecsirt.cpp:121-177).  A warm start with
// reject_above > 0 is checked by SIRT's own residual: it is rejected (event
// 3, solve ends) if the residual printed at pass check_pass is not below
// reject_above, or if a restart would come first.
__global__ __launch_bounds__(kVecThreads) void sirt_control(SirtState *S, const SirtVec *__restrict__ V,
                                                            const double *__restrict__ fp_part) {
    const int k = blockIdx.x;
    SirtState *st = S + k;
    const SirtVec v = V[k];
    if (st->skip_all) return;
    const int first = st->first, warm = st->warm;
    const double *fp = fp_part + (size_t)k * 2 * kVecBlocks;
    double s_b = 0.0, s_g;
    if (first) {
        s_b = block_sum(fp, v.grid);
        s_g = warm ? block_sum(v.bp_part, v.n_bp) : block_sum(fp + v.grid, v.grid);
    } else {
        s_g = block_sum(v.bp_part, v.n_bp);
    }
    if (threadIdx.x != 0) return;
    arec_hist *hist = v.hist;
    if (first) {
        st->bnorm2 += s_b;  // accumulates across restarts (:91,104)
        st->bnorm = sqrt(st->bnorm2);
    }
    (void)warm;
    st->warm = 0;
    double rnorm = sqrt(s_g);
    double relnrm = rnorm / st->bnorm;
    st->zero_x = 0;
    st->update = 0;
    st->npass++;
    if (st->reject_above > 0.0 && st->npass <= st->check_pass &&
        ((st->npass == st->check_pass && relnrm >= st->reject_above) || relnrm > st->relnrm0)) {
        push_hist(hist, st->hist_cap, &st->nhist, st->iter, 3, relnrm, rnorm);
        st->done = 1;
        return;
    }
    if (relnrm > st->relnrm0) {
        if (st->restarts > 20) {
            push_hist(hist, st->hist_cap, &st->nhist, st->iter, 2, relnrm, rnorm);
            st->done = 1;
            return;
        }
        push_hist(hist, st->hist_cap, &st->nhist, st->iter, 1, relnrm, rnorm);
        st->restarts++;
        st->iter = 1;
        st->lam = (float)(st->lam / 2.0);
        st->relnrm0 = 1.0001;
        st->zero_x = 1;
        st->first = 1;
        return;
    }
    if (relnrm < st->tol) {
        push_hist(hist, st->hist_cap, &st->nhist, st->iter, 2, relnrm, rnorm);
        st->done = 1;
        return;
    }
    push_hist(hist, st->hist_cap, &st->nhist, st->iter, 0, relnrm, rnorm);
    st->relnrm0 = relnrm;
    st->update = 1;
    st->first = 0;
    st->iter++;
    if (st->iter > st->maxit) st->done = 1;
}

// x += lam*grad (FP32), or x = 0 after a restart (arecsirt.cpp:145-177).
__global__ __launch_bounds__(kVecThreads) void sirt_update(const SirtVec *__restrict__ V,
                                                           const SirtState *__restrict__ S) {
    const int k = blockIdx.y;
    const SirtState *st = S + k;
    const SirtVec v = V[k];
    if (st->skip_all || (int)blockIdx.x >= v.grid) return;
    const int64_t stride = (int64_t)v.grid * kVecThreads;
    if (st->zero_x) {
        for (int64_t i = blockIdx.x * (int64_t)kVecThreads + threadIdx.x; i < v.nnz; i += stride) v.x[i] = 0.0f;
    } else if (st->update) {
        const float lam = st->lam;
        for (int64_t i = blockIdx.x * (int64_t)kVecThreads + threadIdx.x; i < v.nnz; i += stride) {
            float t = lam * v.grad[i];
            v.x[i] = v.x[i] + t;
        }
    }
}

// *all = 1 when every problem is done.
__global__ void sirt_all_done(const SirtState *__restrict__ S, int nprob, int *all) {
    __shared__ int s_open;
    if (threadIdx.x == 0) s_open = 0;
    __syncthreads();
    for (int k = threadIdx.x; k < nprob; k += blockDim.x)
        if (!S[k].done) atomicOr(&s_open, 1);
    __syncthreads();
    if (threadIdx.x == 0) *all = !s_open;
}

// ------------------------------------------------------------------ CGLS --

__global__ void cgls_begin(CglsState *st) { st->skip = st->done; }

// s = b - prj, then s *= w for the weighted solver (areccgls.cpp:111-112,
// areccgls_stat.cpp:138-149).
__global__ __launch_bounds__(kVecThreads) void cgls_residual(const float *__restrict__ b,
                                                             const float *__restrict__ prj,
                                                             const float *__restrict__ w,
                                                             float *__restrict__ s, int64_t n) {
    GRID_STRIDE(i, n) {
        float v = b[i] - prj[i];
        if (w) v = v * w[i];
        s[i] = v;
    }
}

// gamma = ||g||^2, rnorm0 (areccgls.cpp:127-139); p = g.
__global__ __launch_bounds__(kVecThreads) void cgls_init_control(CglsState *st,
                                                                 const double *__restrict__ part,
                                                                 int64_t n, arec_hist *hist) {
    double s = block_sum(part, n);
    if (threadIdx.x != 0) return;
    st->gamma = s;
    st->rnorm0 = sqrt(s);
    st->rnorm_min = st->rnorm0;
    push_hist(hist, st->hist_cap, &st->nhist, 0, 0, 1.0, st->rnorm0);
}

// Weighted ||q||^2: per-block partials of (double)(float)((float)(q*q) * w).
__global__ __launch_bounds__(kVecThreads) void weighted_sq(const float *__restrict__ q,
                                                           const float *__restrict__ w, int64_t n,
                                                           double *__restrict__ part,
                                                           const CglsState *st) {
    if (st->skip) return;
    double acc = 0.0;
    GRID_STRIDE(i, n) {
        float t = q[i] * q[i];
        float u = t * w[i];
        acc += (double)u;
    }
    block_partial(acc, part);
}

// alpha = gamma / ||q||^2 (areccgls.cpp:155-163).
__global__ __launch_bounds__(kVecThreads) void cgls_alpha(CglsState *st,
                                                          const double *__restrict__ part, int64_t n) {
    if (st->skip) return;
    double s = block_sum(part, n);
    if (threadIdx.x == 0) st->alpha = st->gamma / s;
}

// s -= alpha q (or alpha q w), evaluated in FP64 and rounded (:170-171).
__global__ __launch_bounds__(kVecThreads) void cgls_update_s(float *__restrict__ s,
                                                             const float *__restrict__ q,
                                                             const float *__restrict__ w, int64_t n,
                                                             const CglsState *st) {
    if (st->skip) return;
    const double alpha = st->alpha;
    GRID_STRIDE(i, n) {
        double t = alpha * (double)q[i];
        if (w) t = t * (double)w[i];
        s[i] = (float)((double)s[i] - t);
    }
}

// gamma, relnrm, best iterate, termination, beta (areccgls.cpp:173-220).
__global__ __launch_bounds__(kVecThreads) void cgls_gamma(CglsState *st,
                                                          const double *__restrict__ part, int64_t n,
                                                          arec_hist *hist) {
    if (st->skip) return;
    double s = block_sum(part, n);
    if (threadIdx.x != 0) return;
    st->gamma0 = st->gamma;
    st->gamma = s;
    double rnorm = sqrt(s);
    double relnrm = rnorm / st->rnorm0;
    st->best = 0;
    if (rnorm < st->rnorm_min) {
        st->rnorm_min = rnorm;
        st->iterbest = st->iter;
        st->best = 1;
    }
    int stop = relnrm < st->tol;
    push_hist(hist, st->hist_cap, &st->nhist, st->iter, stop ? 2 : 0, relnrm, rnorm);
    st->beta = st->gamma / st->gamma0;
    st->iter++;
    if (stop || st->iter > st->maxit) st->done = 1;
}

// x += alpha p; xsave = x when this iterate is the best so far; p = g + beta p
// for the next pass (areccgls.cpp:143-151,165-168,190-195).
__global__ __launch_bounds__(kVecThreads) void cgls_update_x(float *__restrict__ x,
                                                             float *__restrict__ p,
                                                             const float *__restrict__ g,
                                                             float *__restrict__ xsave, int64_t n,
                                                             const CglsState *st) {
    if (st->skip) return;
    const double alpha = st->alpha, beta = st->beta;
    const int best = st->best;
    GRID_STRIDE(i, n) {
        float pi = p[i];
        float xi = (float)((double)x[i] + alpha * (double)pi);
        x[i] = xi;
        if (best) xsave[i] = xi;
        p[i] = (float)((double)g[i] + beta * (double)pi);
    }
}

__global__ void copy_kernel(float *__restrict__ dst, const float *__restrict__ src, int64_t n) {
    GRID_STRIDE(i, n) dst[i] = src[i];
}

// ------------------------------------------------------------ host drivers --

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) hipFree(p);
    }
    template <typename T>
    T *as() {
        return (T *)p;
    }
};

int alloc(DevBuf &b, size_t bytes) {
    AREC_HIP(hipMalloc(&b.p, std::max<size_t>(bytes, 16)));
    return 0;
}

// Replays a captured pass until the state's done flag is set, polling every
// `batch` replays through pinned memory.
struct PassRunner {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~PassRunner() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
    }
};

}  // namespace

// SIRT over nprob independent problems (nprob = 1: the reference's solver).
// One pass = forward of every live problem (persistent launches, one per
// forward rows-per-lane class; a single problem uses the plain launch), back
// projection with the gradient epilogue, then the control and update kernels
// of all problems.  The pass is captured once as a hipGraph and replayed
// until every problem is done.
int sirt_batch_run(const std::vector<SirtProblem> &P, hipStream_t st, int maxit, float tol, int warm,
                   double reject_above, int check_pass, arec_hist *hist, int hist_cap, int *nhist,
                   float *lam_out, bool use_graph) {
    const int nprob = (int)P.size();
    if (nprob == 0) return 0;
    const int method = P[0].g->method;
    for (const SirtProblem &q : P)
        if (q.g->method != method) {
            set_error("batched SIRT: all problems must use the same projector method");
            return AREC_EINVAL;
        }
    // workspace offsets
    std::vector<int64_t> o_nnz(nprob + 1, 0), o_proj(nprob + 1, 0), o_bp(nprob + 1, 0);
    for (int k = 0; k < nprob; k++) {
        o_nnz[k + 1] = o_nnz[k] + P[k].g->nnz;
        o_proj[k + 1] = o_proj[k] + P[k].g->nproj;
        o_bp[k + 1] = o_bp[k] + std::max<int64_t>(1, backward_partial_count(*P[k].g));
    }
    const int cap = std::max(1, hist_cap);
    DevBuf bdata, grad, proj, bp_part, fp_part, dstate, dhist, dvec, dall, dbp, dbp0, dcnt;
    int rc;
    if ((rc = alloc(bdata, sizeof(float) * o_nnz[nprob])) || (rc = alloc(grad, sizeof(float) * o_nnz[nprob])) ||
        (rc = alloc(proj, sizeof(float) * o_proj[nprob])) || (rc = alloc(bp_part, sizeof(double) * o_bp[nprob])) ||
        (rc = alloc(fp_part, sizeof(double) * 2 * kVecBlocks * nprob)) ||
        (rc = alloc(dstate, sizeof(SirtState) * nprob)) || (rc = alloc(dhist, sizeof(arec_hist) * cap * nprob)) ||
        (rc = alloc(dvec, sizeof(SirtVec) * nprob)) || (rc = alloc(dall, sizeof(int))) ||
        (rc = alloc(dcnt, sizeof(int))))
        return rc;
    SirtState *S = dstate.as<SirtState>();
    std::vector<SirtState> s0(nprob);
    std::vector<SirtVec> vec(nprob);
    for (int k = 0; k < nprob; k++) {
        SirtState &z = s0[k];
        z = SirtState{};
        z.iter = 1;
        z.first = 1;
        z.lam = P[k].lam;
        z.tol = tol;
        z.maxit = maxit;
        z.hist_cap = hist_cap;
        z.relnrm0 = 1.00001;
        z.done = maxit < 1;
        z.warm = warm != 0;
        z.reject_above = warm ? reject_above : 0.0;
        z.check_pass = std::max(1, check_pass);
        SirtVec &v = vec[k];
        v.x = P[k].x;
        v.bdata = bdata.as<float>() + o_nnz[k];
        v.grad = grad.as<float>() + o_nnz[k];
        v.bp_part = bp_part.as<double>() + o_bp[k];
        v.hist = dhist.as<arec_hist>() + (size_t)cap * k;
        v.nnz = P[k].g->nnz;
        v.n_bp = backward_partial_count(*P[k].g);
        v.grid = (int)vec_grid(v.nnz).x;
        v.pad = 0;
    }
    AREC_HIP(hipMemcpyAsync(S, s0.data(), sizeof(SirtState) * nprob, hipMemcpyHostToDevice, st));
    AREC_HIP(hipMemcpyAsync(dvec.p, vec.data(), sizeof(SirtVec) * nprob, hipMemcpyHostToDevice, st));

    // Projector launches of a pass.  Batched: the problems on the staged-strip
    // forward design go into one forward launch per rows-per-lane class that
    // holds at least two of them (every problem keeps its own tables and
    // rows per lane; AREC_NO_FWD_BATCH set: one launch per problem); the
    // others get a launch each with their own plans.  The back projections of
    // all problems with one tile shape share one persistent launch, problems
    // in decreasing order of angle count so that the long items start first.  (A persistent
    // forward over all problems, a loop over work items, measured 0.66-0.85x
    // of the problems solved one after the other: the loop kept the inlined
    // body's invariants in registers, 86 instead of 44 VGPRs at 2 rows per
    // lane, halving occupancy.  The batched forward launch is one grid with
    // no loop.)
    // Back classes: problems of one tile shape, in decreasing order of angle
    // count, item0 counted within the class; one persistent launch each.
    struct BpClass {
        int shape, first, count;
        int64_t total;
    };
    std::vector<BpClass> bclass;
    std::vector<BatchBp> hb0;  // A^T b (store)
    std::vector<BatchBp> hb;   // gradient pass
    if (nprob > 1) {
        std::vector<int> border(nprob);
        for (int k = 0; k < nprob; k++) border[k] = k;
        std::stable_sort(border.begin(), border.end(), [&](int a, int b) { return P[a].g->na > P[b].g->na; });
        for (int shape = 0; shape < kBpShapes; shape++) {
            BpClass c{shape, (int)hb0.size(), 0, 0};
            for (int k : border) {
                const Geometry &g = *P[k].g;
                if (g.bp_shape != shape) continue;
                BatchBp e;
                e.a = bp_args(g, P[k].b, bdata.as<float>() + o_nnz[k], nullptr, nullptr);
                e.skip = nullptr;
                e.item0 = (int)c.total;
                e.pad = 0;
                hb0.push_back(e);
                e.a = bp_args(g, proj.as<float>() + o_proj[k], vec[k].grad, vec[k].bdata,
                              bp_part.as<double>() + o_bp[k]);
                e.skip = &S[k].skip_proj;
                hb.push_back(e);
                c.total += backward_items(g);
                c.count++;
            }
            if (c.count) bclass.push_back(c);
        }
        if ((rc = alloc(dbp0, sizeof(BatchBp) * nprob)) || (rc = alloc(dbp, sizeof(BatchBp) * nprob))) return rc;
        AREC_HIP(hipMemcpyAsync(dbp0.p, hb0.data(), sizeof(BatchBp) * nprob, hipMemcpyHostToDevice, st));
        AREC_HIP(hipMemcpyAsync(dbp.p, hb.data(), sizeof(BatchBp) * nprob, hipMemcpyHostToDevice, st));
    }
    auto back_batch = [&](int mode, DevBuf &d) -> int {
        for (const BpClass &c : bclass) {
            int r = launch_backward_batch(method, mode, c.shape, d.as<BatchBp>() + c.first, c.count, c.total,
                                          dcnt.as<int>(), st);
            if (r) return r;
        }
        return 0;
    };

    struct FwdClass {
        int rpl, prob0, own0;
        int64_t nblk;
    };
    std::vector<FwdClass> fclass;
    std::vector<int> fsolo;  // problems whose forward is launched on its own
    DevBuf dfwd, down;
    if (nprob > 1) {
        const bool batch_fwd = getenv("AREC_NO_FWD_BATCH") == nullptr;
        std::vector<BatchFwd> hf;
        std::vector<int> own;
        for (int rpl : {4, 2, 1}) {
            std::vector<int> ks;
            for (int k = 0; k < nprob; k++)
                if (P[k].g->fwd_design == 4 && P[k].g->strip_rpl == rpl) ks.push_back(k);
            if (!batch_fwd || ks.size() < 2) continue;
            FwdClass c{rpl, (int)hf.size(), (int)own.size(), 0};
            for (int i = 0; i < (int)ks.size(); i++) {
                const int k = ks[i];
                BatchFwd d;
                const int64_t n = forward_batch_item(*P[k].g, P[k].x, proj.as<float>() + o_proj[k], &S[k].skip_proj,
                                                     c.nblk / 8, d);
                hf.push_back(d);
                own.insert(own.end(), (size_t)(n / 8), i);
                c.nblk += n;
            }
            if (c.nblk > INT32_MAX) {
                set_error("batched SIRT: too many forward workgroups in one launch");
                return AREC_ESIZE;
            }
            fclass.push_back(c);
        }
        for (int k = 0; k < nprob; k++) {
            bool in = false;
            for (const FwdClass &c : fclass) in |= P[k].g->fwd_design == 4 && P[k].g->strip_rpl == c.rpl;
            if (!in) fsolo.push_back(k);
        }
        if (!hf.empty()) {
            if ((rc = alloc(dfwd, sizeof(BatchFwd) * hf.size())) || (rc = alloc(down, sizeof(int) * own.size())))
                return rc;
            AREC_HIP(hipMemcpyAsync(dfwd.p, hf.data(), sizeof(BatchFwd) * hf.size(), hipMemcpyHostToDevice, st));
            AREC_HIP(hipMemcpyAsync(down.p, own.data(), sizeof(int) * own.size(), hipMemcpyHostToDevice, st));
        }
    }

    // Columns that never contribute (linear pair's last two) stay zero in
    // every vector the back projector writes.
    for (int k = 0; k < nprob; k++) {
        const Geometry &g = *P[k].g;
        if (g.nvalid < g.nrays) {
            int64_t n0 = (int64_t)g.nvalid * g.ny, n1 = (int64_t)(g.nrays - g.nvalid) * g.ny;
            if ((rc = fill(vec[k].grad + n0, 0.0f, n1, st, nullptr)) ||
                (rc = fill(bdata.as<float>() + o_nnz[k] + n0, 0.0f, n1, st, nullptr)))
                return rc;
        }
        // x = 0 (arecsirt.cpp:53-56) unless warm-started
        if (!warm && (rc = fill(P[k].x, 0.0f, P[k].g->nnz, st, nullptr))) return rc;
    }
    // bdata = A^T b (:98)
    if (nprob == 1) {
        if ((rc = launch_backward(*P[0].g, P[0].b, bdata.as<float>(), kBpStore, nullptr, nullptr, st, nullptr)))
            return rc;
    } else if ((rc = back_batch(kBpStore, dbp0))) {
        return rc;
    }

    const SirtVec *V = dvec.as<SirtVec>();
    int gmax = 1;
    for (const SirtVec &v : vec) gmax = std::max(gmax, v.grid);
    auto pass = [&]() -> int {
        sirt_begin<<<nprob, 1, 0, st>>>(S);
        int r;
        if (nprob == 1) {
            const Geometry &g = *P[0].g;
            if ((r = launch_forward(g, P[0].x, proj.as<float>(), nullptr, st, &S->skip_proj))) return r;
            if ((r = launch_backward(g, proj.as<float>(), vec[0].grad, kBpSirtGrad, vec[0].bdata,
                                     bp_part.as<double>(), st, &S->skip_proj)))
                return r;
        } else {
            for (const FwdClass &c : fclass)
                if ((r = launch_forward_batch(method, c.rpl, dfwd.as<BatchFwd>() + c.prob0, down.as<int>() + c.own0,
                                              c.nblk, st)))
                    return r;
            for (int k : fsolo)
                if ((r = launch_forward(*P[k].g, P[k].x, proj.as<float>() + o_proj[k], nullptr, st, &S[k].skip_proj)))
                    return r;
            if ((r = back_batch(kBpSirtGrad, dbp)))
                return r;
        }
        sirt_first_pass<<<dim3(gmax, nprob), kVecThreads, 0, st>>>(V, fp_part.as<double>(), S);
        sirt_control<<<nprob, kVecThreads, 0, st>>>(S, V, fp_part.as<double>());
        sirt_update<<<dim3(gmax, nprob), kVecThreads, 0, st>>>(V, S);
        sirt_all_done<<<1, 256, 0, st>>>(S, nprob, dall.as<int>());
        AREC_HIP(hipGetLastError());
        return 0;
    };

    PassRunner runner;
    if (use_graph) {
        AREC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = pass();
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc) return rc;
        AREC_HIP(e);
        runner.graph = graph;
        AREC_HIP(hipGraphInstantiate(&runner.exec, graph, nullptr, nullptr, 0));
    }
    int *done_h = nullptr;
    AREC_HIP(hipHostMalloc((void **)&done_h, sizeof(int), 0));
    *done_h = 0;
    // At most 22 runs of maxit passes (21 restarts), plus slack.
    const int64_t max_passes = (int64_t)(maxit + 1) * 22 + 2;
    int64_t passes = 0;
    int batch = 4;
    while (passes < max_passes) {
        for (int k = 0; k < batch; k++) {
            if (use_graph) {
                hipError_t e = hipGraphLaunch(runner.exec, st);
                if (e != hipSuccess) {
                    hipHostFree(done_h);
                    return hip_fail(e, "hipGraphLaunch");
                }
            } else if ((rc = pass())) {
                hipHostFree(done_h);
                return rc;
            }
        }
        passes += batch;
        hipMemcpyAsync(done_h, dall.p, sizeof(int), hipMemcpyDeviceToHost, st);
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            hipHostFree(done_h);
            return hip_fail(e, "hipStreamSynchronize");
        }
        if (*done_h) break;
        batch = std::min(batch * 2, 32);
    }
    hipHostFree(done_h);
    std::vector<SirtState> sf(nprob);
    AREC_HIP(hipMemcpy(sf.data(), S, sizeof(SirtState) * nprob, hipMemcpyDeviceToHost));
    for (int k = 0; k < nprob; k++) {
        int n = std::min(sf[k].nhist, hist_cap);
        if (hist && n > 0)
            AREC_HIP(hipMemcpy(hist + (size_t)hist_cap * k, vec[k].hist, sizeof(arec_hist) * n, hipMemcpyDeviceToHost));
        if (nhist) nhist[k] = sf[k].nhist;
        if (lam_out) lam_out[k] = sf[k].lam;
    }
    return 0;
}

int sirt_run(const Geometry &g, hipStream_t st, const float *b, float *x, float lam, int maxit,
             float tol, int warm, double reject_above, int check_pass, arec_hist *hist, int hist_cap, int *nhist,
             float *lam_out, bool use_graph) {
    std::vector<SirtProblem> P{SirtProblem{&g, b, x, lam}};
    return sirt_batch_run(P, st, maxit, tol, warm, reject_above, check_pass, hist, hist_cap, nhist, lam_out,
                          use_graph);
}

int cgls_run(const Geometry &g, hipStream_t st, const float *b, const float *w, float *x, int warm,
             int maxit, float tol, arec_hist *hist, int hist_cap, int *nhist, bool use_graph) {
    const int64_t nnz = g.nnz, nproj = g.nproj;
    DevBuf prj, s, gv, pv, xsave, bp_part, fw_part, dstate, dhist;
    int rc;
    const int64_t nfp = std::max<int64_t>(forward_partial_count(g), kVecBlocks);
    if ((rc = alloc(prj, sizeof(float) * nproj)) || (rc = alloc(s, sizeof(float) * nproj)) ||
        (rc = alloc(gv, sizeof(float) * nnz)) || (rc = alloc(pv, sizeof(float) * nnz)) ||
        (rc = alloc(xsave, sizeof(float) * nnz)) ||
        (rc = alloc(bp_part, sizeof(double) * std::max<int64_t>(1, backward_partial_count(g)))) ||
        (rc = alloc(fw_part, sizeof(double) * nfp)) || (rc = alloc(dstate, sizeof(CglsState))) ||
        (rc = alloc(dhist, sizeof(arec_hist) * std::max(1, hist_cap))))
        return rc;
    CglsState c0 = {};
    c0.iter = 1;
    c0.iterbest = maxit - 1;
    c0.maxit = maxit;
    c0.hist_cap = hist_cap;
    c0.tol = tol;
    c0.done = maxit < 1;
    CglsState *S = dstate.as<CglsState>();
    AREC_HIP(hipMemcpyAsync(S, &c0, sizeof c0, hipMemcpyHostToDevice, st));
    if (warm) {
        if ((rc = launch_forward(g, x, prj.as<float>(), nullptr, st, nullptr))) return rc;
    } else {
        if ((rc = fill(x, 0.0f, nnz, st, nullptr)) || (rc = fill(prj.as<float>(), 0.0f, nproj, st, nullptr)))
            return rc;
    }
    cgls_residual<<<vec_grid(nproj), kVecThreads, 0, st>>>(b, prj.as<float>(), w, s.as<float>(), nproj);
    if ((rc = launch_backward(g, s.as<float>(), gv.as<float>(), kBpStoreNorm, nullptr, bp_part.as<double>(), st, nullptr)))
        return rc;
    cgls_init_control<<<1, kVecThreads, 0, st>>>(S, bp_part.as<double>(), backward_partial_count(g), dhist.as<arec_hist>());
    copy_kernel<<<vec_grid(nnz), kVecThreads, 0, st>>>(pv.as<float>(), gv.as<float>(), nnz);
    AREC_HIP(hipGetLastError());

    const int64_t nfwd = w ? (int64_t)vec_grid(nproj).x : forward_partial_count(g);
    auto pass = [&]() -> int {
        cgls_begin<<<1, 1, 0, st>>>(S);
        int r;
        if ((r = launch_forward(g, pv.as<float>(), prj.as<float>(), w ? nullptr : fw_part.as<double>(), st, &S->skip)))
            return r;
        if (w)
            weighted_sq<<<vec_grid(nproj), kVecThreads, 0, st>>>(prj.as<float>(), w, nproj, fw_part.as<double>(), S);
        cgls_alpha<<<1, kVecThreads, 0, st>>>(S, fw_part.as<double>(), nfwd);
        cgls_update_s<<<vec_grid(nproj), kVecThreads, 0, st>>>(s.as<float>(), prj.as<float>(), w, nproj, S);
        if ((r = launch_backward(g, s.as<float>(), gv.as<float>(), kBpStoreNorm, nullptr, bp_part.as<double>(), st, &S->skip)))
            return r;
        cgls_gamma<<<1, kVecThreads, 0, st>>>(S, bp_part.as<double>(), backward_partial_count(g), dhist.as<arec_hist>());
        cgls_update_x<<<vec_grid(nnz), kVecThreads, 0, st>>>(x, pv.as<float>(), gv.as<float>(), xsave.as<float>(), nnz, S);
        AREC_HIP(hipGetLastError());
        return 0;
    };
    PassRunner runner;
    if (use_graph) {
        AREC_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        rc = pass();
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc) return rc;
        AREC_HIP(e);
        runner.graph = graph;
        AREC_HIP(hipGraphInstantiate(&runner.exec, graph, nullptr, nullptr, 0));
    }
    int *done_h = nullptr;
    AREC_HIP(hipHostMalloc((void **)&done_h, sizeof(int), 0));
    *done_h = 0;
    int64_t passes = 0;
    int batch = 4;
    while (passes < (int64_t)maxit + 1) {
        int nb = (int)std::min<int64_t>(batch, maxit + 1 - passes);
        for (int k = 0; k < nb; k++) {
            if (use_graph) {
                hipError_t e = hipGraphLaunch(runner.exec, st);
                if (e != hipSuccess) {
                    hipHostFree(done_h);
                    return hip_fail(e, "hipGraphLaunch");
                }
            } else if ((rc = pass())) {
                hipHostFree(done_h);
                return rc;
            }
        }
        passes += nb;
        hipMemcpyAsync(done_h, &S->done, sizeof(int), hipMemcpyDeviceToHost, st);
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) {
            hipHostFree(done_h);
            return hip_fail(e, "hipStreamSynchronize");
        }
        if (*done_h) break;
        batch = std::min(batch * 2, 32);
    }
    hipHostFree(done_h);
    CglsState cf;
    AREC_HIP(hipMemcpy(&cf, S, sizeof cf, hipMemcpyDeviceToHost));
    // retrieve the best approximation (areccgls.cpp:226-230)
    if (cf.iterbest < maxit - 1) {
        copy_kernel<<<vec_grid(nnz), kVecThreads, 0, st>>>(x, xsave.as<float>(), nnz);
        AREC_HIP(hipGetLastError());
    }
    int n = std::min(cf.nhist, hist_cap);
    if (hist && n > 0) AREC_HIP(hipMemcpy(hist, dhist.p, sizeof(arec_hist) * n, hipMemcpyDeviceToHost));
    if (nhist) *nhist = cf.nhist;
    AREC_HIP(hipStreamSynchronize(st));
    return 0;
}

// ------------------------------------------------------------- utilities --

// init_weights (areccgls_stat.cpp:12-47) with tukey_filter_x_inplace
// (imagetools.cpp:369-391): w = (float)exp(-b) then w = (float)(w * hx).
__global__ void init_weights_kernel(const float *__restrict__ b, float *__restrict__ w, int64_t nx,
                                    int64_t n) {
    const double alpha = 0.5, pi = 3.141592653589793;
    const double xa = 0.5 * alpha * (double)(nx - 1);
    const double xbnd = (1.0 - 0.5 * alpha) * (double)(nx - 1);
    const double A = 2.0 * pi / alpha;
    GRID_STRIDE(i, n) {
        int64_t x = i % nx;
        float e = (float)exp(-(double)b[i]);
        double hx = 1.0;
        if ((double)x < xa)
            hx = 0.5 * (1.0 + cos(A * (1.0 * (double)x / (double)(nx - 1) - alpha / 2)));
        else if ((double)x > xbnd)
            hx = 0.5 * (1.0 + cos(A * (1.0 * (double)x / (double)(nx - 1) - 1.0 + alpha / 2)));
        w[i] = (float)((double)e * hx);
    }
}

int copy(float *dst, const float *src, int64_t n, hipStream_t st) {
    copy_kernel<<<vec_grid(n), kVecThreads, 0, st>>>(dst, src, n);
    AREC_HIP(hipGetLastError());
    return 0;
}

int init_weights(const float *b, float *w, int64_t nx, int64_t ny, int64_t nz, hipStream_t st) {
    int64_t n = nx * ny * nz;
    init_weights_kernel<<<vec_grid(n), kVecThreads, 0, st>>>(b, w, nx, n);
    AREC_HIP(hipGetLastError());
    return 0;
}

__global__ void soft_noneg_kernel(float *__restrict__ x, int64_t n) {
    GRID_STRIDE(i, n) {
        float v = x[i];
        if (v < 0.0f) x[i] = v * 0.5f;
    }
}

__global__ void scale_kernel(float *__restrict__ x, float a, int64_t n) { GRID_STRIDE(i, n) x[i] = a * x[i]; }

int scale(float *x, float a, int64_t n, hipStream_t st) {
    if (n <= 0) return 0;
    scale_kernel<<<vec_grid(n), kVecThreads, 0, st>>>(x, a, n);
    AREC_HIP(hipGetLastError());
    return 0;
}

int soft_noneg(float *x, int64_t n, hipStream_t st) {
    soft_noneg_kernel<<<vec_grid(n), kVecThreads, 0, st>>>(x, n);
    AREC_HIP(hipGetLastError());
    return 0;
}

// Dot product: FP64 sum of exact FP64 products of FP32 values.
__global__ __launch_bounds__(kVecThreads) void dot_kernel(const float *__restrict__ a,
                                                          const float *__restrict__ b, int64_t n,
                                                          double *__restrict__ part) {
    double acc = 0.0;
    GRID_STRIDE(i, n) acc += (double)a[i] * (double)b[i];
    block_partial(acc, part);
}

__global__ __launch_bounds__(kVecThreads) void sum_kernel(const double *__restrict__ part, int64_t n,
                                                          double *__restrict__ out) {
    double s = block_sum(part, n);
    if (threadIdx.x == 0) *out = s;
}

int dot(const float *a, const float *b, int64_t n, double *out, hipStream_t st) {
    dim3 grid = vec_grid(n);
    DevBuf part;
    int rc;
    if ((rc = alloc(part, sizeof(double) * (grid.x + 1)))) return rc;
    dot_kernel<<<grid, kVecThreads, 0, st>>>(a, b, n, part.as<double>());
    sum_kernel<<<1, kVecThreads, 0, st>>>(part.as<double>(), grid.x, part.as<double>() + grid.x);
    AREC_HIP(hipGetLastError());
    AREC_HIP(hipMemcpyAsync(out, part.as<double>() + grid.x, sizeof(double), hipMemcpyDeviceToHost, st));
    AREC_HIP(hipStreamSynchronize(st));
    return 0;
}

__global__ __launch_bounds__(kVecThreads) void max_kernel(const float *__restrict__ x, int64_t n,
                                                          float *__restrict__ part) {
    __shared__ float s[kVecThreads / 64];
    float m = -INFINITY;
    GRID_STRIDE(i, n) m = fmaxf(m, x[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = s[0];
        for (int w = 1; w < kVecThreads / 64; w++) t = fmaxf(t, s[w]);
        part[blockIdx.x] = t;
    }
}

int reduce_max(const float *x, int64_t n, float *out, hipStream_t st) {
    dim3 grid = vec_grid(n);
    DevBuf part, fin;
    int rc;
    if ((rc = alloc(part, sizeof(float) * grid.x)) || (rc = alloc(fin, sizeof(float)))) return rc;
    max_kernel<<<grid, kVecThreads, 0, st>>>(x, n, part.as<float>());
    max_kernel<<<1, kVecThreads, 0, st>>>(part.as<float>(), grid.x, fin.as<float>());
    AREC_HIP(hipGetLastError());
    AREC_HIP(hipMemcpyAsync(out, fin.p, sizeof(float), hipMemcpyDeviceToHost, st));
    AREC_HIP(hipStreamSynchronize(st));
    return 0;
}

}  // namespace arec
