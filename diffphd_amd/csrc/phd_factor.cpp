// Host-side factor construction: ordering, symbolic analysis, multifrontal
// FP64 Cholesky and the index-free layout of S = L^-1 (see phd_factor.h).
#include "phd_factor.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace phd {

namespace {

// Runs body(i) for i in [0, n) on `threads` threads in dynamic chunks.
void parallel_for(int threads, int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& body) {
    if (threads <= 1 || n <= grain) { body(0, n); return; }
    std::atomic<int64_t> next(0);
    auto work = [&]() {
        for (;;) {
            const int64_t b = next.fetch_add(grain);
            if (b >= n) break;
            body(b, std::min(n, b + grain));
        }
    };
    std::vector<std::thread> pool;
    const int t = (int)std::min<int64_t>(threads, (n + grain - 1) / grain);
    for (int i = 1; i < t; ++i) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

// Nested dissection by recursive coordinate bisection: the median plane of the
// widest extent splits the node set, the lower-side nodes adjacent to the
// upper side form the separator, ordered last.
void nd_order(std::vector<int>& nodes, const std::vector<double>& X, const std::vector<int>& Ap,
              const std::vector<int>& Ai, std::vector<char>& mark, std::vector<int>& out) {
    if (nodes.size() <= 48) { out.insert(out.end(), nodes.begin(), nodes.end()); return; }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v : nodes)
        for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], X[3 * v + d]); hi[d] = std::max(hi[d], X[3 * v + d]); }
    int ax = 0;
    for (int d = 1; d < 3; ++d) if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    std::vector<double> key(nodes.size());
    for (size_t i = 0; i < nodes.size(); ++i) key[i] = X[3 * nodes[i] + ax];
    std::vector<double> sk = key;
    std::nth_element(sk.begin(), sk.begin() + sk.size() / 2, sk.end());
    const double med = sk[sk.size() / 2];
    for (size_t i = 0; i < nodes.size(); ++i) mark[nodes[i]] = (key[i] < med) ? 1 : 2;
    std::vector<int> L, U, S;
    for (int v : nodes) {
        if (mark[v] == 2) { U.push_back(v); continue; }
        bool touch = false;
        for (int p = Ap[v]; p < Ap[v + 1] && !touch; ++p) touch = (mark[Ai[p]] == 2);
        (touch ? S : L).push_back(v);
    }
    for (int v : nodes) mark[v] = 0;
    if (L.empty() || U.empty()) { out.insert(out.end(), nodes.begin(), nodes.end()); return; }
    std::vector<int>().swap(nodes);
    nd_order(L, X, Ap, Ai, mark, out);
    nd_order(U, X, Ap, Ai, mark, out);
    out.insert(out.end(), S.begin(), S.end());
}

// Elimination tree of P A P^T (Liu's algorithm with path compression).
std::vector<int> etree(int n, const std::vector<int>& Ap, const std::vector<int>& Ai, const std::vector<int>& perm,
                       const std::vector<int>& iperm) {
    std::vector<int> parent(n, -1), anc(n, -1);
    for (int k = 0; k < n; ++k) {
        const int v = perm[k];
        for (int p = Ap[v]; p < Ap[v + 1]; ++p) {
            int i = iperm[Ai[p]];
            while (i != -1 && i < k) {
                const int nxt = anc[i];
                anc[i] = k;
                if (nxt == -1) { parent[i] = k; break; }
                i = nxt;
            }
        }
    }
    return parent;
}

// Partial Cholesky of a dense front (row-major, lower triangle, leading dim f):
// factors the leading ns columns and leaves the Schur complement in F[ns:, ns:].
void partial_cholesky(double* F, int f, int ns, int threads) {
    const int B = 48;
    const bool par = (int64_t)f * f > 256 * 256;
    for (int p0 = 0; p0 < ns; p0 += B) {
        const int p1 = std::min(p0 + B, ns);
        // diagonal block
        for (int j = p0; j < p1; ++j) {
            double* rj = F + (int64_t)j * f;
            double d = rj[j];
            for (int k = p0; k < j; ++k) d -= rj[k] * rj[k];
            if (!(d > 0)) throw std::runtime_error("Cholesky factorization failed: matrix is not positive definite.");
            d = std::sqrt(d);
            rj[j] = d;
            for (int i = j + 1; i < p1; ++i) {
                double* ri = F + (int64_t)i * f;
                double s = ri[j];
                for (int k = p0; k < j; ++k) s -= ri[k] * rj[k];
                ri[j] = s / d;
            }
        }
        // panel rows below the block: F[i, p0:p1] <- F[i, p0:p1] L_bb^-T
        auto trsm = [&](int64_t a, int64_t b) {
            for (int64_t i = p1 + a; i < p1 + b; ++i) {
                double* ri = F + i * f;
                for (int j = p0; j < p1; ++j) {
                    const double* rj = F + (int64_t)j * f;
                    double s = ri[j];
                    for (int k = p0; k < j; ++k) s -= ri[k] * rj[k];
                    ri[j] = s / rj[j];
                }
            }
        };
        if (par) parallel_for(threads, f - p1, 16, trsm); else trsm(0, f - p1);
        // trailing update of the lower triangle
        const int w = p1 - p0;
        auto syrk = [&](int64_t a, int64_t b) {
            for (int64_t i = p1 + a; i < p1 + b; ++i) {
                double* ri = F + i * f;
                const double* li = ri + p0;
                for (int64_t j = p1; j <= i; ++j) {
                    const double* lj = F + j * f + p0;
                    double s = 0;
                    for (int k = 0; k < w; ++k) s += li[k] * lj[k];
                    ri[j] -= s;
                }
            }
        };
        if (par) parallel_for(threads, f - p1, 8, syrk); else syrk(0, f - p1);
    }
}

}  // namespace

void Factor::build(int n_, const std::vector<int>& Ap, const std::vector<int>& Ai, const std::vector<double>& Ax,
                   const std::vector<double>& X, int threads) {
    auto t0 = std::chrono::steady_clock::now();
    n = n_;
    // 1. nested dissection, then postorder of the elimination tree.
    std::vector<int> nodes(n), perm0;
    std::iota(nodes.begin(), nodes.end(), 0);
    std::vector<char> mark(n, 0);
    perm0.reserve(n);
    nd_order(nodes, X, Ap, Ai, mark, perm0);
    std::vector<int> ip0(n);
    for (int i = 0; i < n; ++i) ip0[perm0[i]] = i;
    std::vector<int> par0 = etree(n, Ap, Ai, perm0, ip0);
    // postorder (children visited in increasing index order)
    std::vector<int> head(n, -1), next(n, -1), stack, post;
    post.reserve(n);
    for (int j = n - 1; j >= 0; --j)
        if (par0[j] != -1) { next[j] = head[par0[j]]; head[par0[j]] = j; }
    for (int r = 0; r < n; ++r) {
        if (par0[r] != -1) continue;
        stack.push_back(r);
        while (!stack.empty()) {
            const int p = stack.back();
            const int c = head[p];
            if (c == -1) { stack.pop_back(); post.push_back(p); }
            else { head[p] = next[c]; stack.push_back(c); }
        }
    }
    perm.resize(n);
    for (int k = 0; k < n; ++k) perm[k] = perm0[post[k]];
    iperm.resize(n);
    for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
    parent = etree(n, Ap, Ai, perm, iperm);
    auto t1 = std::chrono::steady_clock::now();
    order_seconds = std::chrono::duration<double>(t1 - t0).count();

    // 2. depth, first descendants.
    depth.assign(n, 0);
    for (int k = n - 1; k >= 0; --k) depth[k] = parent[k] == -1 ? 0 : depth[parent[k]] + 1;
    fd.resize(n);
    for (int k = 0; k < n; ++k) fd[k] = k;
    for (int k = 0; k < n; ++k)
        if (parent[k] != -1) fd[parent[k]] = std::min(fd[parent[k]], fd[k]);
    max_depth = 0;
    for (int k = 0; k < n; ++k) max_depth = std::max(max_depth, depth[k]);
    for (int k = 0; k < n; ++k)
        if (parent[k] != -1 && parent[k] <= k) throw std::runtime_error("elimination tree is not postordered");

    // 3. column patterns of L via row subtrees (row k: ereach of its entries).
    std::vector<int> cc(n, 1), flag(n, -1);
    auto row_walk = [&](int k, const std::function<void(int)>& visit) {
        flag[k] = k;
        const int v = perm[k];
        for (int p = Ap[v]; p < Ap[v + 1]; ++p) {
            int i = iperm[Ai[p]];
            if (i >= k) continue;
            for (; flag[i] != k; i = parent[i]) { flag[i] = k; visit(i); }
        }
    };
    for (int k = 0; k < n; ++k) row_walk(k, [&](int j) { cc[j]++; });
    Lp.assign(n + 1, 0);
    for (int j = 0; j < n; ++j) Lp[j + 1] = Lp[j] + cc[j];
    Li.assign(Lp[n], 0);
    std::vector<int64_t> fill(n);
    for (int j = 0; j < n; ++j) { fill[j] = Lp[j] + 1; Li[Lp[j]] = j; }
    std::fill(flag.begin(), flag.end(), -1);
    for (int k = 0; k < n; ++k) row_walk(k, [&](int j) { Li[fill[j]++] = k; });
    Lx.assign(Lp[n], 0.0);

    // 4. fundamental supernodes.
    std::vector<int> nchild(n, 0);
    for (int k = 0; k < n; ++k) if (parent[k] != -1) nchild[parent[k]]++;
    std::vector<int> sn_start;
    for (int j = 0; j < n; ++j) {
        const bool cont = j > 0 && parent[j - 1] == j && cc[j - 1] == cc[j] + 1 && nchild[j] == 1;
        if (!cont) sn_start.push_back(j);
    }
    const int ns_total = (int)sn_start.size();
    sn_start.push_back(n);
    std::vector<int> sn_of(n);
    for (int s = 0; s < ns_total; ++s)
        for (int j = sn_start[s]; j < sn_start[s + 1]; ++j) sn_of[j] = s;

    // 5. multifrontal numeric factorization (supernodes in postorder).
    struct Update { std::vector<int> rows; std::vector<double> U; };
    std::vector<Update> stack_u;
    std::vector<int> map(n, -1);
    std::vector<int> sn_nchild(ns_total, 0);
    for (int s = 0; s < ns_total; ++s) {
        const int last = sn_start[s + 1] - 1;
        if (parent[last] != -1) sn_nchild[sn_of[parent[last]]]++;
    }
    std::vector<double> F;
    for (int s = 0; s < ns_total; ++s) {
        const int c0 = sn_start[s], c1 = sn_start[s + 1];
        const int ns = c1 - c0;
        const int f = cc[c0];
        const int* R = &Li[Lp[c0]];
        for (int t = 0; t < f; ++t) map[R[t]] = t;
        F.assign((size_t)f * f, 0.0);
        for (int j = c0; j < c1; ++j) {
            const int v = perm[j];
            for (int p = Ap[v]; p < Ap[v + 1]; ++p) {
                const int i = iperm[Ai[p]];
                if (i < j) continue;
                F[(size_t)map[i] * f + map[j]] += Ax[p];
            }
        }
        for (int c = 0; c < sn_nchild[s]; ++c) {
            Update& u = stack_u.back();
            const int m = (int)u.rows.size();
            std::vector<int> loc(m);
            for (int a = 0; a < m; ++a) loc[a] = map[u.rows[a]];
            for (int a = 0; a < m; ++a) {
                double* Fa = &F[(size_t)loc[a] * f];
                const double* Ua = &u.U[(size_t)a * m];
                for (int b = 0; b <= a; ++b) Fa[loc[b]] += Ua[b];
            }
            stack_u.pop_back();
        }
        partial_cholesky(F.data(), f, ns, threads);
        for (int t = 0; t < ns; ++t) {
            const int64_t p = Lp[c0 + t];
            for (int r = t; r < f; ++r) Lx[p + (r - t)] = F[(size_t)r * f + t];
        }
        const int m = f - ns;
        if (m > 0) {
            Update u;
            u.rows.assign(R + ns, R + f);
            u.U.resize((size_t)m * m);
            for (int a = 0; a < m; ++a)
                std::memcpy(&u.U[(size_t)a * m], &F[(size_t)(ns + a) * f + ns], sizeof(double) * (a + 1));
            stack_u.push_back(std::move(u));
        }
        for (int t = 0; t < f; ++t) map[R[t]] = -1;
    }
    if (!stack_u.empty()) throw std::runtime_error("multifrontal update stack not empty");

    // 6. index-free layout of S: base[k] + j, base[k] multiple of 64, rows packed.
    base.resize(n);
    int64_t end = 0;
    nnz = 0;
    for (int k = 0; k < n; ++k) {
        int64_t b = end - fd[k];
        b = (b + 63) / 64 * 64;
        base[k] = b;
        end = b + k + 1;
        nnz += k - fd[k] + 1;
    }
    v_len = (end + 63) / 64 * 64;
    factor_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
}

void SpmvPlan::build(const Factor& f, int unit_items, int group_chunk_w) {
    chunk_w = group_chunk_w;
    const int n = f.n;
    const int ntiles = (n + 63) / 64;
    items.clear(); cu_tile.clear(); cu_i0.clear(); cu_i1.clear();
    tile_unit_ptr.assign(ntiles + 1, 0);
    for (int t = 0; t < ntiles; ++t) {
        const int t0 = 64 * t, t1 = std::min(n - 1, 64 * t + 63);
        const int i0 = (int)items.size();
        for (int k = t0; k <= t1; ++k) items.push_back(k);
        for (int k = f.parent[t1]; k != -1; k = f.parent[k])
            if (k > t1) items.push_back(k);
        const int i1 = (int)items.size();
        for (int a = i0; a < i1; a += unit_items) {
            cu_tile.push_back(t);
            cu_i0.push_back(a);
            cu_i1.push_back(std::min(i1, a + unit_items));
        }
        tile_unit_ptr[t + 1] = (int)cu_tile.size();
    }
    // S w by row groups x column chunks; per-(row, chunk) partial slots, contiguous per row
    gu_group.clear(); gu_chunk.clear();
    const int ngroups = (n + kRowGroup - 1) / kRowGroup;
    for (int g = 0; g < ngroups; ++g) {
        int lo = n, hi = -1;
        for (int k = g * kRowGroup; k < std::min(n, (g + 1) * kRowGroup); ++k) { lo = std::min(lo, f.fd[k]); hi = std::max(hi, k); }
        for (int c = lo / chunk_w; c <= hi / chunk_w; ++c) { gu_group.push_back(g); gu_chunk.push_back(c); }
    }
    rbase.assign(n + 1, 0);
    for (int k = 0; k < n; ++k) rbase[k + 1] = rbase[k] + (k / chunk_w - f.fd[k] / chunk_w + 1);
    // Largest units first: the separators near the root cross every chunk and
    // would otherwise be dispatched last, leaving a tail. Each slot belongs to
    // one unit, so the order does not change any result.
    const size_t nu = gu_group.size();
    std::vector<int64_t> work(nu, 0);
    for (size_t u = 0; u < nu; ++u) {
        const int g = gu_group[u], c0 = gu_chunk[u] * chunk_w, c1 = c0 + chunk_w - 1;
        for (int k = g * kRowGroup; k < std::min(n, (g + 1) * kRowGroup); ++k)
            work[u] += std::max(0, std::min(k, c1) - std::max(f.fd[k], c0) + 1);
    }
    std::vector<size_t> ord(nu);
    for (size_t u = 0; u < nu; ++u) ord[u] = u;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return work[a] > work[b]; });
    std::vector<int> gg(nu), gc(nu);
    for (size_t u = 0; u < nu; ++u) { gg[u] = gu_group[ord[u]]; gc[u] = gu_chunk[ord[u]]; }
    gu_group.swap(gg);
    gu_chunk.swap(gc);
}


// Host emulation of the device data path for small systems (tests): expands
// S = L^-1 column by column, checks S^T S A = I, the subtree-range structure,
// and replays the two SpMV plans against dense products.
// out: [0] max |S^T S A - I|, [1] max |plan(S w) - S w|, [2] max |plan(S^T y) - S^T y|,
//      [3] nnz(S), [4] max depth, [5] nnz(L), [6] structural violations.
void host_factor_check(int n, const std::vector<int>& Ap, const std::vector<int>& Ai, const std::vector<double>& Ax,
                       const std::vector<double>& X, double* out) {
    Factor f;
    f.build(n, Ap, Ai, Ax, X, 2);
    SpmvPlan plan;
    plan.build(f, 5, 16);  // small chunks so that rows split and chains split
    std::vector<double> V(f.v_len, 0.0);
    std::vector<double> x(n);
    double viol = 0;
    for (int j = 0; j < n; ++j) {
        std::fill(x.begin(), x.end(), 0.0);
        x[j] = 1.0;
        for (int i = j; i < n; ++i) {
            if (x[i] == 0.0) continue;
            x[i] /= f.Lx[f.Lp[i]];
            for (int64_t p = f.Lp[i] + 1; p < f.Lp[i + 1]; ++p) x[f.Li[p]] -= f.Lx[p] * x[i];
        }
        for (int k = 0; k < n; ++k) {
            if (x[k] == 0.0) continue;
            if (!(f.fd[k] <= j && j <= k)) { viol += 1; continue; }
            V[f.base[k] + j] = x[k];
        }
    }
    auto S = [&](int k, int j) { return (f.fd[k] <= j && j <= k) ? V[f.base[k] + j] : 0.0; };
    // S^T S A = I on the permuted matrix: (P A P^T)^{-1} = S^T S
    double err = 0;
    std::vector<double> col(n), y(n), z(n);
    for (int c = 0; c < n; ++c) {
        std::fill(col.begin(), col.end(), 0.0);
        const int vc = f.perm[c];
        for (int p = Ap[vc]; p < Ap[vc + 1]; ++p) col[f.iperm[Ai[p]]] = Ax[p];
        for (int k = 0; k < n; ++k) { double s = 0; for (int j = f.fd[k]; j <= k; ++j) s += S(k, j) * col[j]; y[k] = s; }
        for (int j = 0; j < n; ++j) { double s = 0; for (int k = j; k < n; ++k) s += S(k, j) * y[k]; z[j] = s; }
        for (int j = 0; j < n; ++j) err = std::max(err, std::fabs(z[j] - (j == c ? 1.0 : 0.0)));
    }
    // plan replay with a pseudo-random vector
    std::vector<double> w(n), yref(n), ypl(n, 0.0), xref(n), xpl(n, 0.0);
    for (int j = 0; j < n; ++j) w[j] = std::sin(0.37 * j + 1.0);
    for (int k = 0; k < n; ++k) { double s = 0; for (int j = f.fd[k]; j <= k; ++j) s += S(k, j) * w[j]; yref[k] = s; }
    // S w through the row-group units: per-(row, chunk) slots, then per-row sums
    std::vector<double> pslot(plan.rbase[n], 0.0);
    std::vector<char> seen(plan.rbase[n], 0);
    for (size_t u = 0; u < plan.gu_group.size(); ++u) {
        const int g = plan.gu_group[u], ch = plan.gu_chunk[u];
        const int c0 = ch * plan.chunk_w, c1 = c0 + plan.chunk_w - 1;
        for (int k = g * SpmvPlan::kRowGroup; k < std::min(n, (g + 1) * SpmvPlan::kRowGroup); ++k) {
            const int lo = std::max(f.fd[k], c0), hi = std::min(k, c1);
            if (lo > hi) continue;
            const int slot = plan.rbase[k] + ch - f.fd[k] / plan.chunk_w;
            if (slot < plan.rbase[k] || slot >= plan.rbase[k + 1] || seen[slot]) viol += 1;
            seen[slot] = 1;
            for (int c = lo; c <= hi; ++c) pslot[slot] += V[f.base[k] + c] * w[c];
        }
    }
    for (int k = 0; k < n; ++k)
        for (int s = plan.rbase[k]; s < plan.rbase[k + 1]; ++s) {
            if (!seen[s]) viol += 1;
            ypl[k] += pslot[s];
        }
    for (int j = 0; j < n; ++j) { double s = 0; for (int k = j; k < n; ++k) s += S(k, j) * yref[k]; xref[j] = s; }
    for (size_t u = 0; u < plan.cu_tile.size(); ++u) {
        const int t = plan.cu_tile[u];
        for (int it = plan.cu_i0[u]; it < plan.cu_i1[u]; ++it) {
            const int k = plan.items[it];
            for (int l = 0; l < 64; ++l) {
                const int j = 64 * t + l;
                if (j < n && j >= f.fd[k] && j <= k) xpl[j] += V[f.base[k] + j] * yref[k];
            }
        }
    }
    double e1 = 0, e2 = 0;
    for (int k = 0; k < n; ++k) { e1 = std::max(e1, std::fabs(ypl[k] - yref[k])); e2 = std::max(e2, std::fabs(xpl[k] - xref[k])); }
    out[0] = err; out[1] = e1; out[2] = e2; out[3] = (double)f.nnz; out[4] = f.max_depth; out[5] = (double)f.Lx.size();
    out[6] = viol;
}

}  // namespace phd
