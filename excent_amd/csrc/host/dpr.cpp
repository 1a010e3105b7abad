// dpr.cpp -- differentiable personalised PageRank (no reference counterpart).
//
// Model: x = (1 - d) s + d A x with A[u][v] = w_vu / S_v for every edge v -> u, where w are
// raw positive edge weights and S_v = sum of v's out-weights.  Loss L = sum_v g_v x_v.
//   forward : the delta push over the CSR with per-edge coefficients w_vu / S_v;
//   adjoint : lambda = (I - d A^T)^-1 g, the same delta push over the transpose (the
//             in-neighbour lists) with the coefficient of edge v -> u moved to u's list;
//   grads   : dL/ds = (1 - d) lambda;  dL/d(w_vu/S_v) = d lambda_u x_v;
//             dL/dw_vu = (d x_v / S_v) (lambda_u - sum_{u'} (w_vu'/S_v) lambda_u').
// Both solves run through the device C ABI on fully cached graphs, as pushes and, while the
// demand is high, as dense passes (fold + pull weighted by the edge coefficients); the
// transpose and the gradient assembly run on the host with OpenMP.
#include <omp.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "excent.h"
#include "excent_hip.h"

namespace {

thread_local std::string g_err;
int err(const std::string& s) { g_err = s; return -1; }

struct Dpr {
    int dev = 0, n = 0;
    uint64_t m = 0;
    float damp = 0.85f, eps = 1e-7f;
    std::vector<uint64_t> idx, tidx;
    std::vector<int32_t> adj, tadj;
    std::vector<float> S, wn, x;
    void* fctx = nullptr;
    void* bctx = nullptr;
    int passes_f = 0, passes_b = 0;
    ~Dpr() {
        if (fctx) exc_destroy(fctx);
        if (bctx) exc_destroy(bctx);
    }
};

int make_ctx(Dpr* D, const std::vector<uint64_t>& off, const std::vector<int32_t>& lists, const std::vector<float>& w,
             void** out) {
    exc_measures M{};
    M.nm = 1;
    M.kind[0] = EXC_PR;
    M.eps[0] = D->eps;
    M.scale[0] = D->damp;
    M.divdeg[0] = 0;
    int cv0 = 0;
    exc_config cfg{};
    cfg.device = D->dev; cfg.n = D->n; cfg.m = D->m; cfg.nc = 1; cfg.cv0 = &cv0;
    cfg.max_chunk_edges = 1; cfg.max_chunk_verts = 1; cfg.push_kernel = EXC_KERNEL_LB;
    int rc = exc_create(&cfg, &M, out);
    if (rc) return rc;
    if ((rc = exc_upload_offsets(*out, off.data(), 0, off.size()))) return rc;
    if ((rc = exc_cache_chunks(*out, D->m))) return rc;
    if (D->m && (rc = exc_cache_upload(*out, lists.data(), 0, D->m))) return rc;
    if ((rc = exc_set_edge_weights(*out, w.data(), D->m))) return rc;
    // dense passes (fold + weighted pull over the device transpose) while the demand is high;
    // without the memory for the transpose the push alone solves it
    rc = exc_enable_pull(*out, 0.05);
    return rc == EXC_ENOMEM ? EXC_OK : rc;
}

// Runs passes (snapshot, push over all vertices) from x = 0, r = r0 until nothing is pending.
int solve(Dpr* D, void* ctx, const float* r0, float* x_out, int* passes) {
    std::vector<float> zero(D->n, 0.f);
    int rc = exc_set_state(ctx, 0, zero.data(), r0);
    if (rc) return rc;
    uint64_t cnt[2], hp = 0;
    exc_chunk all{0, D->n, 0, D->m};
    int p = 0;
    for (; p < 1000000; p++) {
        if ((rc = exc_snapshot(ctx, 0, cnt, &hp))) return rc;
        if (cnt[0] == 0) break;
        if ((rc = exc_push_resident(ctx, &all, (double)cnt[0] < 0.5 * D->n, 0, 0))) return rc;
        if ((rc = exc_pass_sync(ctx, nullptr))) return rc;
    }
    *passes = p;
    return exc_read_state(ctx, 0, 0, x_out, 0, D->n);
}

}  // namespace

extern "C" {

const char* excent_dpr_error(void) { return g_err.empty() ? exc_last_error() : g_err.c_str(); }

int excent_dpr_create(int device, int n, uint64_t m, const uint64_t* idx, const int32_t* adj, const float* w_raw,
                      float damp, float eps, void** out) {
    g_err.clear();
    if (!out || n < 1 || !idx || (m && (!adj || !w_raw)) || idx[n] != m) return err("excent_dpr_create: bad arguments");
    Dpr* D = new Dpr();
    D->dev = device; D->n = n; D->m = m; D->damp = damp; D->eps = eps;
    D->idx.assign(idx, idx + n + 1);
    D->adj.assign(adj, adj + m);
    for (uint64_t e = 0; e < m; e++)
        if (adj[e] < 0 || adj[e] >= n || !(w_raw[e] > 0.f)) { delete D; return err("excent_dpr_create: bad edge or weight"); }
    D->S.assign(n, 0.f);
    D->wn.resize(m);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int v = 0; v < n; v++) {
        double s = 0;
        for (uint64_t e = idx[v]; e < idx[v + 1]; e++) s += w_raw[e];
        D->S[v] = (float)s;
        for (uint64_t e = idx[v]; e < idx[v + 1]; e++) D->wn[e] = (float)(w_raw[e] / s);
    }
    // transpose with the coefficients carried along (counting sort, sources ascending)
    D->tidx.assign((size_t)n + 1, 0);
    for (uint64_t e = 0; e < m; e++) D->tidx[adj[e] + 1]++;
    for (int v = 0; v < n; v++) D->tidx[v + 1] += D->tidx[v];
    D->tadj.resize(m);
    std::vector<float> wt(m);
    std::vector<uint64_t> pos(D->tidx.begin(), D->tidx.end() - 1);
    for (int v = 0; v < n; v++)
        for (uint64_t e = idx[v]; e < idx[v + 1]; e++) {
            const uint64_t q = pos[adj[e]]++;
            D->tadj[q] = v;
            wt[q] = D->wn[e];
        }
    int rc = make_ctx(D, D->idx, D->adj, D->wn, &D->fctx);
    if (!rc) rc = make_ctx(D, D->tidx, D->tadj, wt, &D->bctx);
    if (rc) { delete D; return rc; }
    *out = D;
    return 0;
}

int excent_dpr_forward(void* h, const float* s, float* x_out) {
    Dpr* D = static_cast<Dpr*>(h);
    if (!D || !s) return err("excent_dpr_forward: bad arguments");
    g_err.clear();
    std::vector<float> r0(D->n);
    for (int v = 0; v < D->n; v++) r0[v] = (float)((1.0 - (double)D->damp) * s[v]);
    D->x.resize(D->n);
    int rc = solve(D, D->fctx, r0.data(), D->x.data(), &D->passes_f);
    if (rc) return rc;
    if (x_out) memcpy(x_out, D->x.data(), D->n * sizeof(float));
    return 0;
}

int excent_dpr_backward(void* h, const float* g, float* grad_s, float* grad_w, float* grad_wn) {
    Dpr* D = static_cast<Dpr*>(h);
    if (!D || !g || D->x.size() != (size_t)D->n) return err("excent_dpr_backward: run the forward first");
    g_err.clear();
    std::vector<float> lam(D->n);
    int rc = solve(D, D->bctx, g, lam.data(), &D->passes_b);
    if (rc) return rc;
    const double d = D->damp;
    if (grad_s)
        for (int v = 0; v < D->n; v++) grad_s[v] = (float)((1.0 - d) * lam[v]);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int v = 0; v < D->n; v++) {
        const uint64_t b = D->idx[v], e1 = D->idx[v + 1];
        if (b == e1) continue;
        double t = 0;
        for (uint64_t e = b; e < e1; e++) t += (double)D->wn[e] * lam[D->adj[e]];
        const double xv = D->x[v], f = d * xv / D->S[v];
        for (uint64_t e = b; e < e1; e++) {
            const double lu = lam[D->adj[e]];
            if (grad_w) grad_w[e] = (float)(f * (lu - t));
            if (grad_wn) grad_wn[e] = (float)(d * lu * xv);
        }
    }
    return 0;
}

int excent_dpr_passes(void* h, int* forward, int* backward) {
    Dpr* D = static_cast<Dpr*>(h);
    if (!D) return err("excent_dpr_passes: null handle");
    *forward = D->passes_f;
    *backward = D->passes_b;
    return 0;
}

int excent_dpr_destroy(void* h) {
    delete static_cast<Dpr*>(h);
    return 0;
}

}  // extern "C"
