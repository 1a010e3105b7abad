// fp64.hip -- FP64 synchronous solvers behind bin/centrality_reference and
// bin/excent_params (tools/centrality_reference.cu:32-59, 184-214 and
// tools/excent_params.cu:27-60, 132-163, paths relative to the reference root).
//
// The reference pulls over CSR lists and so assumes a symmetric graph.  These solvers
// pull over in-neighbour lists (the transpose, built here on the device) and divide by
// the source's out-degree, which equals the reference on symmetric graphs and matches
// the push semantics (push to out-neighbours) on directed ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "excent_hip.h"

void exc_internal_set_error(const char* msg);   // kernels.hip: exc_last_error() text

namespace {

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    exc_internal_set_error(buf);
    return code;
}
#define HIPCHK(call)                                                                                        \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) return fail(EXC_EHIP, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

constexpr int BLOCK = 256;
constexpr int GL = 16;   // lanes per vertex in the pull

__global__ void k_count_in(uint64_t m, const int32_t* __restrict__ adj, unsigned long long* cnt) {
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < (long long)m; e += (long long)gridDim.x * blockDim.x)
        atomicAdd(&cnt[adj[e] + 1], 1ull);
}
__global__ void k_fill_in(int n, const uint64_t* __restrict__ idx, const int32_t* __restrict__ adj,
                          unsigned long long* pos, int32_t* tadj) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x)
        for (uint64_t e = idx[v]; e < idx[v + 1]; e++) tadj[atomicAdd(&pos[adj[e]], 1ull)] = (int32_t)v;
}
// xs[u] = x[u] / outdeg(u) (PageRank) or x[u]
__global__ void k_prescale(int n, const uint64_t* __restrict__ idx, const double* __restrict__ x, double* xs, int divdeg) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) {
        const uint64_t d = idx[v + 1] - idx[v];
        xs[v] = divdeg ? (d ? x[v] / (double)d : 0.0) : x[v];
    }
}
// y[w] = c + s * sum_{u -> w} xs[u]; mx[0] = max |y - x|, mx[1] = max |y| (as ordered bits).
// GL lanes per vertex, vertex groups strided over the grid (the whole group takes the same
// trip count, so the shuffles see every lane).
__global__ __launch_bounds__(BLOCK) void k_pull(int n, const uint64_t* __restrict__ tidx, const int32_t* __restrict__ tadj,
                                                 const double* __restrict__ xs, const double* __restrict__ x, double* y,
                                                 double c, const double* __restrict__ cvec, double s,
                                                 unsigned long long* mx) {
    const int gl = threadIdx.x % GL;
    const long long groups = (long long)gridDim.x * (BLOCK / GL);
    for (long long w = (blockIdx.x * (long long)BLOCK + threadIdx.x) / GL; w < n; w += groups) {
        double acc = 0;
        for (uint64_t e = tidx[w] + gl; e < tidx[w + 1]; e += GL) acc += xs[tadj[e]];
#pragma unroll
        for (int o = GL / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, GL);
        if (gl == 0) {
            const double nv = (cvec ? cvec[w] : c) + s * acc;
            y[w] = nv;
            atomicMax(&mx[0], (unsigned long long)__double_as_longlong(fabs(nv - x[w])));
            atomicMax(&mx[1], (unsigned long long)__double_as_longlong(fabs(nv)));
        }
    }
}
__global__ void k_sums(int n, const double* __restrict__ x, double* out) {   // out[0] += x^2, out[1] += x
    double a = 0, b = 0;
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) {
        a += x[v] * x[v];
        b += x[v];
    }
    for (int o = 32; o > 0; o >>= 1) { a += __shfl_xor(a, o, 64); b += __shfl_xor(b, o, 64); }
    if ((threadIdx.x & 63) == 0) { atomicAdd(&out[0], a); atomicAdd(&out[1], b); }
}
__global__ void k_scale_diff(int n, double* y, const double* __restrict__ x, double f, unsigned long long* mx) {
    double md = 0;
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) {
        y[v] *= f;
        md = fmax(md, fabs(x[v] - y[v]));
    }
    for (int o = 32; o > 0; o >>= 1) md = fmax(md, __shfl_xor(md, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(mx, (unsigned long long)__double_as_longlong(md));
}
__global__ void k_absdiff_sum(int n, const double* __restrict__ y, const double* __restrict__ x, double* out) {
    double a = 0;
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x)
        a += fabs(y[v] - x[v]);
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, a);
}
// Transpose refill that carries each edge's coefficient with its in-list entry.
__global__ void k_fill_in_coef(int n, const uint64_t* __restrict__ idx, const int32_t* __restrict__ adj,
                               const double* __restrict__ coef, unsigned long long* pos, int32_t* tadj, double* tcoef) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x)
        for (uint64_t e = idx[v]; e < idx[v + 1]; e++) {
            const unsigned long long q = atomicAdd(&pos[adj[e]], 1ull);
            tadj[q] = (int32_t)v;
            tcoef[q] = coef[e];
        }
}
// y[w] = cvec[w] + s * sum_{e in [lidx[w], lidx[w+1])} lcoef[e] x[ladj[e]] over lists given
// as offsets / neighbours / coefficients: the transpose's in-lists (x = c + s W x) or the
// graph's out-lists (x = c + s W^T x); mx as in k_pull.
__global__ __launch_bounds__(BLOCK) void k_pull_coef(int n, const uint64_t* __restrict__ lidx,
                                                      const int32_t* __restrict__ ladj,
                                                      const double* __restrict__ lcoef, const double* __restrict__ x,
                                                      double* y, const double* __restrict__ cvec, double s,
                                                      unsigned long long* mx) {
    const int gl = threadIdx.x % GL;
    const long long groups = (long long)gridDim.x * (BLOCK / GL);
    for (long long w = (blockIdx.x * (long long)BLOCK + threadIdx.x) / GL; w < n; w += groups) {
        double acc = 0;
        for (uint64_t e = lidx[w] + gl; e < lidx[w + 1]; e += GL) acc += lcoef[e] * x[ladj[e]];
#pragma unroll
        for (int o = GL / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, GL);
        if (gl == 0) {
            const double nv = cvec[w] + s * acc;
            y[w] = nv;
            atomicMax(&mx[0], (unsigned long long)__double_as_longlong(fabs(nv - x[w])));
            atomicMax(&mx[1], (unsigned long long)__double_as_longlong(fabs(nv)));
        }
    }
}
__global__ void k_fill(int n, double* x, double f) {
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) x[v] = f;
}

struct Solver {
    int dev = 0, n = 0, cus = 256;
    uint64_t m = 0;
    uint64_t* d_idx = nullptr;
    uint64_t* d_tidx = nullptr;
    int32_t* d_tadj = nullptr;
    double *x = nullptr, *y = nullptr, *xs = nullptr, *red = nullptr;
    double* cvec = nullptr;   // per-vertex constant of the last vector solve / residual
    unsigned long long* mx = nullptr;
    // per-edge coefficients (exc_fp64_set_coef): out-lists in the graph's order, and the
    // transpose (refilled so that tcoef is aligned with tadj)
    int32_t* d_adj = nullptr;
    double* d_coef = nullptr;
    double* d_tcoef = nullptr;
};

double bits_to_d(unsigned long long b) { double d; memcpy(&d, &b, 8); return d; }

int pull(Solver* S, const double* x, double* y, double c, double s, int divdeg, double* mxd, double* mxy,
         const double* cvec = nullptr) {
    HIPCHK(hipMemset(S->mx, 0, 2 * sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_prescale, dim3(S->cus * 8), dim3(BLOCK), 0, 0, S->n, S->d_idx, x, S->xs, divdeg);
    HIPCHK(hipGetLastError());
    // at most 64 blocks per CU (the kernel strides over vertex groups): a grid of one group
    // per vertex would pass 2^32 threads at n = 2^28
    const long long want = ((long long)S->n * GL + BLOCK - 1) / BLOCK;
    const unsigned grid = (unsigned)std::min<long long>(want, (long long)S->cus * 64);
    hipLaunchKernelGGL(k_pull, dim3(grid), dim3(BLOCK), 0, 0, S->n, S->d_tidx, S->d_tadj, S->xs, x, y, c, cvec, s, S->mx);
    HIPCHK(hipGetLastError());
    unsigned long long h[2];
    HIPCHK(hipMemcpy(h, S->mx, sizeof h, hipMemcpyDeviceToHost));
    *mxd = bits_to_d(h[0]);
    *mxy = bits_to_d(h[1]);
    return EXC_OK;
}

int sums(Solver* S, const double* x, double* sq, double* sm) {
    HIPCHK(hipMemset(S->red, 0, 2 * sizeof(double)));
    hipLaunchKernelGGL(k_sums, dim3(S->cus * 4), dim3(BLOCK), 0, 0, S->n, x, S->red);
    HIPCHK(hipGetLastError());
    double h[2];
    HIPCHK(hipMemcpy(h, S->red, sizeof h, hipMemcpyDeviceToHost));
    *sq = h[0];
    *sm = h[1];
    return EXC_OK;
}

int scale_diff(Solver* S, double* y, const double* x, double f, double* md) {
    HIPCHK(hipMemset(S->mx, 0, sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_scale_diff, dim3(S->cus * 4), dim3(BLOCK), 0, 0, S->n, y, x, f, S->mx);
    HIPCHK(hipGetLastError());
    unsigned long long h;
    HIPCHK(hipMemcpy(&h, S->mx, sizeof h, hipMemcpyDeviceToHost));
    *md = bits_to_d(h);
    return EXC_OK;
}

}  // namespace

extern "C" {

int exc_fp64_create(int device, int n, uint64_t m, const uint64_t* h_idx, const int32_t* h_adj, void** out) {
    if (!out || n < 1 || !h_idx || (m && !h_adj)) return fail(EXC_EINVAL, "exc_fp64_create: bad arguments");
    Solver* S = new Solver();
    S->dev = device; S->n = n; S->m = m;
    *out = nullptr;
    auto bail = [&](int rc) { exc_fp64_destroy(S); return rc; };
#define CK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) \
    return bail(fail(EXC_EHIP, "HIP error %s at %s:%d", hipGetErrorString(e_), __FILE__, __LINE__)); } while (0)
    CK(hipSetDevice(device));
    hipDeviceProp_t p;
    CK(hipGetDeviceProperties(&p, device));
    S->cus = p.multiProcessorCount;
    int32_t* d_adj = nullptr;
    unsigned long long* cnt = nullptr;
    CK(hipMalloc(&S->d_idx, ((size_t)n + 1) * 8));
    CK(hipMemcpy(S->d_idx, h_idx, ((size_t)n + 1) * 8, hipMemcpyHostToDevice));
    CK(hipMalloc(&d_adj, (m ? m : 1) * 4));
    if (m) CK(hipMemcpy(d_adj, h_adj, m * 4, hipMemcpyHostToDevice));
    CK(hipMalloc(&cnt, ((size_t)n + 1) * 8));
    CK(hipMemset(cnt, 0, ((size_t)n + 1) * 8));
    hipLaunchKernelGGL(k_count_in, dim3(S->cus * 16), dim3(BLOCK), 0, 0, m, d_adj, cnt);
    CK(hipGetLastError());
    std::vector<unsigned long long> h((size_t)n + 1);
    CK(hipMemcpy(h.data(), cnt, h.size() * 8, hipMemcpyDeviceToHost));
    for (size_t v = 0; v < (size_t)n; v++) h[v + 1] += h[v];
    CK(hipMalloc(&S->d_tidx, ((size_t)n + 1) * 8));
    CK(hipMemcpy(S->d_tidx, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    CK(hipMemcpy(cnt, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    CK(hipMalloc(&S->d_tadj, (m ? m : 1) * 4));
    hipLaunchKernelGGL(k_fill_in, dim3(S->cus * 8), dim3(BLOCK), 0, 0, n, S->d_idx, d_adj, cnt, S->d_tadj);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    hipFree(d_adj);
    hipFree(cnt);
    CK(hipMalloc(&S->x, (size_t)n * 8));
    CK(hipMalloc(&S->y, (size_t)n * 8));
    CK(hipMalloc(&S->xs, (size_t)n * 8));
    CK(hipMalloc(&S->red, 2 * sizeof(double)));
    CK(hipMalloc(&S->mx, 2 * sizeof(unsigned long long)));
#undef CK
    *out = S;
    return EXC_OK;
}

int exc_fp64_destroy(void* p) {
    Solver* S = static_cast<Solver*>(p);
    if (!S) return EXC_OK;
    void* q[] = {S->d_idx, S->d_tidx, S->d_tadj, S->x, S->y, S->xs, S->red, S->mx, S->cvec, S->d_adj, S->d_coef, S->d_tcoef};
    for (void* z : q) if (z) (void)hipFree(z);
    delete S;
    return EXC_OK;
}

// tools/centrality_reference.cu:184-214
int exc_fp64_solve(void* p, int kind, double alpha, double beta, int max_iter, double tol,
                   const double* h_x0, double* h_out, double* rel_out, double* lambda_out, int* iters) {
    Solver* S = static_cast<Solver*>(p);
    if (!S || !h_out) return fail(EXC_EINVAL, "exc_fp64_solve: bad arguments");
    HIPCHK(hipSetDevice(S->dev));
    const int n = S->n;
    if (h_x0) HIPCHK(hipMemcpy(S->x, h_x0, (size_t)n * 8, hipMemcpyHostToDevice));
    else hipLaunchKernelGGL(k_fill, dim3(S->cus * 4), dim3(BLOCK), 0, 0, n, S->x, kind == EXC_PR ? 0.15 : 1.0);
    HIPCHK(hipGetLastError());
    double c = 0, s = 1;
    int divdeg = 0;
    if (kind == EXC_PR) { c = 0.15; s = 0.85; divdeg = 1; }
    else if (kind == EXC_KATZ) { c = beta; s = alpha; }
    int it = 0;
    double rel = 1, lam = 0;
    double* x = S->x;
    double* y = S->y;
    for (; it < max_iter && rel > tol; it++) {
        double mxd, mxy;
        int rc = pull(S, x, y, c, s, divdeg, &mxd, &mxy);
        if (rc) return rc;
        if (kind == EXC_EV) {
            double sq, sm;
            rc = sums(S, y, &sq, &sm);
            if (rc) return rc;
            lam = std::sqrt(sq);
            double md;
            rc = scale_diff(S, y, x, (sm < 0 ? -1.0 : 1.0) / lam, &md);   // unit L2, sign by sum
            if (rc) return rc;
            rel = it == 0 ? 1.0 : md;
        } else {
            rel = mxd / mxy;
        }
        std::swap(x, y);
    }
    HIPCHK(hipMemcpy(h_out, x, (size_t)n * 8, hipMemcpyDeviceToHost));
    if (rel_out) *rel_out = rel;
    if (lambda_out) *lambda_out = lam;
    if (iters) *iters = it;
    return EXC_OK;
}

int upload_cvec(Solver* S, const double* h_c) {
    if (!h_c) return EXC_OK;
    if (!S->cvec) HIPCHK(hipMalloc(&S->cvec, (size_t)S->n * 8));
    HIPCHK(hipMemcpy(S->cvec, h_c, (size_t)S->n * 8, hipMemcpyHostToDevice));
    return EXC_OK;
}

// ||c + s W x - x||_1 in FP64 for a float32 state x (W = A^T, scaled by 1/outdeg when
// divdeg; c a constant or, when h_c is given, a per-vertex vector): the global residual
// the push's termination bound is stated in (src/excent.cu:100-124 keeps the invariant
// x + r = c + s W x).
int exc_fp64_residual(void* p, double c, const double* h_c, double s, int divdeg, const float* h_x, double* l1) {
    Solver* S = static_cast<Solver*>(p);
    if (!S || !h_x || !l1) return fail(EXC_EINVAL, "exc_fp64_residual: bad arguments");
    HIPCHK(hipSetDevice(S->dev));
    const int n = S->n;
    int rc = upload_cvec(S, h_c);
    if (rc) return rc;
    std::vector<double> hx(h_x, h_x + n);
    HIPCHK(hipMemcpy(S->x, hx.data(), (size_t)n * 8, hipMemcpyHostToDevice));
    double mxd, mxy;
    rc = pull(S, S->x, S->y, c, s, divdeg, &mxd, &mxy, h_c ? S->cvec : nullptr);
    if (rc) return rc;
    HIPCHK(hipMemset(S->red, 0, sizeof(double)));
    hipLaunchKernelGGL(k_absdiff_sum, dim3(S->cus * 4), dim3(BLOCK), 0, 0, n, S->y, S->x, S->red);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(l1, S->red, sizeof(double), hipMemcpyDeviceToHost));
    return EXC_OK;
}

// Jacobi x <- c + s W x from x = c for a per-vertex constant c (personalised PageRank:
// c = (1 - d) * teleport, s = d, divdeg = 1), until max|x' - x| / max|x'| < tol.
int exc_fp64_solve_vec(void* p, const double* h_c, double s, int divdeg, int max_iter, double tol, double* h_out,
                       double* rel_out, int* iters) {
    Solver* S = static_cast<Solver*>(p);
    if (!S || !h_c || !h_out) return fail(EXC_EINVAL, "exc_fp64_solve_vec: bad arguments");
    HIPCHK(hipSetDevice(S->dev));
    const int n = S->n;
    int rc = upload_cvec(S, h_c);
    if (rc) return rc;
    HIPCHK(hipMemcpy(S->x, S->cvec, (size_t)n * 8, hipMemcpyDeviceToDevice));
    double* x = S->x;
    double* y = S->y;
    double rel = 1;
    int it = 0;
    for (; it < max_iter && rel > tol; it++) {
        double mxd, mxy;
        rc = pull(S, x, y, 0.0, s, divdeg, &mxd, &mxy, S->cvec);
        if (rc) return rc;
        rel = mxy > 0 ? mxd / mxy : 0.0;
        std::swap(x, y);
    }
    HIPCHK(hipMemcpy(h_out, x, (size_t)n * 8, hipMemcpyDeviceToHost));
    if (rel_out) *rel_out = rel;
    if (iters) *iters = it;
    return EXC_OK;
}

// tools/excent_params.cu:132-147
int exc_fp64_lambda(void* p, const double* h_start, int max_iter, double tol, double* lambda, double* mean_ev, int* iters) {
    Solver* S = static_cast<Solver*>(p);
    if (!S || !lambda) return fail(EXC_EINVAL, "exc_fp64_lambda: bad arguments");
    HIPCHK(hipSetDevice(S->dev));
    const int n = S->n;
    if (h_start) HIPCHK(hipMemcpy(S->x, h_start, (size_t)n * 8, hipMemcpyHostToDevice));
    else hipLaunchKernelGGL(k_fill, dim3(S->cus * 4), dim3(BLOCK), 0, 0, n, S->x, 1.0);
    HIPCHK(hipGetLastError());
    double sq, sm, md;
    int rc = sums(S, S->x, &sq, &sm);
    if (rc) return rc;
    rc = scale_diff(S, S->x, S->x, 1.0 / std::sqrt(sq), &md);
    if (rc) return rc;
    double lam = 0, lamPrev = 0, a, b;
    double* x = S->x;
    double* y = S->y;
    int pit = 0;
    for (; pit < max_iter; pit++) {
        rc = pull(S, x, y, 0.0, 1.0, 0, &a, &b);
        if (rc) return rc;
        rc = sums(S, y, &sq, &sm);
        if (rc) return rc;
        lam = std::sqrt(sq);
        rc = scale_diff(S, y, x, 1.0 / lam, &md);
        if (rc) return rc;
        std::swap(x, y);
        if (pit > 10 && std::fabs(lam - lamPrev) / lam < tol) break;
        lamPrev = lam;
    }
    rc = sums(S, x, &sq, &sm);
    if (rc) return rc;
    *lambda = lam;
    if (mean_ev) *mean_ev = sm * sm / n;
    if (iters) *iters = pit;
    if (x != S->x) std::swap(S->x, S->y);
    return EXC_OK;
}

// Per-edge coefficients for exc_fp64_solve_coef: h_adj and h_coef are the graph's out-lists
// (the CSR given to exc_fp64_create) and one coefficient per edge.
int exc_fp64_set_coef(void* p, const int32_t* h_adj, const double* h_coef) {
    Solver* S = static_cast<Solver*>(p);
    if (!S || (S->m && (!h_adj || !h_coef))) return fail(EXC_EINVAL, "exc_fp64_set_coef: bad arguments");
    HIPCHK(hipSetDevice(S->dev));
    const size_t m = S->m ? S->m : 1;
    if (!S->d_adj) HIPCHK(hipMalloc(&S->d_adj, m * 4));
    if (!S->d_coef) HIPCHK(hipMalloc(&S->d_coef, m * 8));
    if (!S->d_tcoef) HIPCHK(hipMalloc(&S->d_tcoef, m * 8));
    if (S->m) {
        HIPCHK(hipMemcpy(S->d_adj, h_adj, S->m * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(S->d_coef, h_coef, S->m * 8, hipMemcpyHostToDevice));
    }
    unsigned long long* pos = nullptr;
    HIPCHK(hipMalloc(&pos, ((size_t)S->n + 1) * 8));
    hipError_t e = hipMemcpy(pos, S->d_tidx, ((size_t)S->n + 1) * 8, hipMemcpyDeviceToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_in_coef, dim3(S->cus * 8), dim3(BLOCK), 0, 0, S->n, S->d_idx, S->d_adj, S->d_coef, pos,
                           S->d_tadj, S->d_tcoef);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(pos);
    if (e != hipSuccess) return fail(EXC_EHIP, "exc_fp64_set_coef: %s", hipGetErrorString(e));
    return EXC_OK;
}

// Jacobi x = c + s W x (transpose = 0: W[u][v] = coef of edge v -> u, the pull over in-lists)
// or x = c + s W^T x (transpose = 1: the pull over out-lists, the adjoint system) from x = c
// until max|x' - x| / max|x'| < tol or max_iter.
int exc_fp64_solve_coef(void* p, int transpose, const double* h_c, double s, int max_iter, double tol, double* h_out,
                        double* rel_out, int* iters) {
    Solver* S = static_cast<Solver*>(p);
    if (!S || !h_c || !h_out) return fail(EXC_EINVAL, "exc_fp64_solve_coef: bad arguments");
    if (!S->d_coef) return fail(EXC_ESTATE, "exc_fp64_solve_coef: no coefficients (exc_fp64_set_coef)");
    HIPCHK(hipSetDevice(S->dev));
    const int n = S->n;
    int rc = upload_cvec(S, h_c);
    if (rc) return rc;
    HIPCHK(hipMemcpy(S->x, S->cvec, (size_t)n * 8, hipMemcpyDeviceToDevice));
    const uint64_t* lidx = transpose ? S->d_idx : S->d_tidx;
    const int32_t* ladj = transpose ? S->d_adj : S->d_tadj;
    const double* lcoef = transpose ? S->d_coef : S->d_tcoef;
    const long long want = ((long long)n * GL + BLOCK - 1) / BLOCK;
    const unsigned grid = (unsigned)std::min<long long>(want, (long long)S->cus * 64);
    double* x = S->x;
    double* y = S->y;
    double rel = 1;
    int it = 0;
    for (; it < max_iter && rel > tol; it++) {
        HIPCHK(hipMemset(S->mx, 0, 2 * sizeof(unsigned long long)));
        hipLaunchKernelGGL(k_pull_coef, dim3(grid), dim3(BLOCK), 0, 0, n, lidx, ladj, lcoef, x, y, S->cvec, s, S->mx);
        HIPCHK(hipGetLastError());
        unsigned long long h[2];
        HIPCHK(hipMemcpy(h, S->mx, sizeof h, hipMemcpyDeviceToHost));
        const double mxd = bits_to_d(h[0]), mxy = bits_to_d(h[1]);
        rel = mxy > 0 ? mxd / mxy : 0.0;
        std::swap(x, y);
    }
    HIPCHK(hipMemcpy(h_out, x, (size_t)n * 8, hipMemcpyDeviceToHost));
    if (rel_out) *rel_out = rel;
    if (iters) *iters = it;
    return EXC_OK;
}

// tools/excent_params.cu:149-163
int exc_fp64_katz_mean(void* p, double alpha, double beta, double* mean, int* iters) {
    Solver* S = static_cast<Solver*>(p);
    if (!S || !mean) return fail(EXC_EINVAL, "exc_fp64_katz_mean: bad arguments");
    HIPCHK(hipSetDevice(S->dev));
    const int n = S->n;
    hipLaunchKernelGGL(k_fill, dim3(S->cus * 4), dim3(BLOCK), 0, 0, n, S->x, beta);
    HIPCHK(hipGetLastError());
    double* x = S->x;
    double* y = S->y;
    double meanK = 0, a, b, sq, sm;
    int it = 0;
    for (; it < 5000; it++) {
        int rc = pull(S, x, y, beta, alpha, 0, &a, &b);
        if (rc) return rc;
        std::swap(x, y);
        rc = sums(S, x, &sq, &sm);
        if (rc) return rc;
        const double nm = sm / n;
        if (std::fabs(nm - meanK) / nm < 1e-12) { meanK = nm; break; }
        meanK = nm;
    }
    *mean = meanK;
    if (iters) *iters = it;
    return EXC_OK;
}

}  // extern "C"
