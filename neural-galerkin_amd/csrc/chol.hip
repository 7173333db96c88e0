// Sparse direct solve of a level system: supernodal multifrontal Cholesky in FP64.
//
// Takes the role of the reference's "cholmod" solver (jspsr/core/solver.py:147-157 ->
// _sparse_solve.solve_cusparse, csrc/sparse_solve/solve_cusparse.cu:15-42, a cuSOLVER sparse
// Cholesky with reordering). Three phases:
//  - analysis (host): a fill-reducing nested-dissection order from BFS level structures of the
//    matrix graph (every BFS level separates the levels before it from those after it), the
//    supernodes (the dissection tree's separators and leaf blocks, numbered in postorder), the row
//    structure of every front, the map of every lower-triangle entry of A into its front, and the
//    map of every update row into the parent's front;
//  - factorization (GPU): fronts processed by height in the tree, each height one batch: A scattered
//    into the fronts, children's update matrices extend-added, then a blocked partial Cholesky per
//    front (diagonal block in LDS, panel triangular solve, lower-tiled Schur update), FP64 throughout;
//  - solve (GPU): forward substitution up the tree, backward substitution down it, FP64.
// The factor keeps every front (f x f, column-major, lower part used) in one device pool.
#include "common.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <memory>
#include <numeric>
#include <vector>

namespace ngs {

namespace {

constexpr int CHOL_NB = 32;       // block width of the partial factorization and the solves
constexpr int CHOL_TILE = 64;     // Schur-update tile
constexpr int CHOL_LEAF = 64;     // dissection stops at this many vertices
constexpr int CHOL_THREADS = 256;

// Per-node description read by the kernels.
struct NodeDesc {
    int64_t f_off;     // front offset in the pool (doubles)
    int64_t idx_off;   // offset of the front's index list (f entries: pivot columns, then update rows)
    int64_t cmap_off;  // offset of the update rows' positions in the parent front (-1: root)
    int64_t child_off; // offset of the node's children (with update rows) in the children list
    int32_t f;         // front size
    int32_t ns;        // pivot columns
    int32_t parent;    // parent node (-1: root)
    int32_t start;     // first pivot column (new numbering)
    int32_t nchild;    // children with update rows
    int32_t pad;
};

struct HostNode {
    int32_t start = 0, end = 0, parent = -1, height = 0;
    std::vector<int32_t> rows;  // update rows, new numbering, ascending (all >= end)
    std::vector<int32_t> children;
};

}  // namespace

}  // namespace ngs

// Opaque handle of the C ABI.
struct ngs_chol_s {
    int64_t M = 0, nnz = 0;
    std::vector<int32_t> perm, iperm;        // perm[new] = old
    std::vector<ngs::HostNode> nodes;        // postorder
    std::vector<ngs::NodeDesc> desc;
    std::vector<int32_t> front_idx;          // concatenated front index lists (new numbering)
    std::vector<int32_t> cmap;               // concatenated child -> parent front positions
    std::vector<int64_t> amap;               // CSR entry -> pool position (-1: upper triangle)
    std::vector<std::vector<int32_t>> level_nodes;     // node ids by height
    std::vector<int32_t> level_max_ns, level_max_f;
    // Extend-add batches: for height h and slot k, the k-th child (with update rows) of every node of
    // height h; one launch per slot, so each parent front receives its children in child order.
    std::vector<std::vector<int64_t>> slot_off, slot_n, slot_maxu;
    std::vector<int32_t> children;  // per node (child_off, nchild), the children with update rows
    std::vector<int32_t> lvl;       // uploaded: node lists by height, slot lists, perm, iperm
    int64_t lvl_perm_off = 0;
    int64_t pool_doubles = 0;
    int64_t fill = 0;  // stored entries of L (lower fronts' pivot columns)
    // workspace layout (bytes)
    size_t off_pool = 0, off_amap = 0, off_desc = 0, off_idx = 0, off_cmap = 0, off_kids = 0, off_lvl = 0,
           off_diag = 0, off_sing = 0, off_y = 0, off_yv = 0, total = 0;
    std::vector<int64_t> lvl_node_off;  // offsets (int32 elements) of the node lists in lvl
    const void* uploaded = nullptr;                    // workspace holding the index arrays
};

namespace ngs {
namespace {

// ------------------------------------------------------------------------------- analysis
struct Graph {
    int32_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int32_t> adj;  // symmetric pattern without the diagonal
};

Graph build_graph(const std::vector<int32_t>& rowptr, const std::vector<int32_t>& cols, int32_t n) {
    Graph g;
    g.n = n;
    std::vector<std::vector<int32_t>> nb(n);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            const int32_t j = cols[k];
            if (j == i || j < 0 || j >= n) continue;
            nb[i].push_back(j);
            nb[j].push_back(i);  // symmetrise the pattern
        }
    g.ptr.assign(n + 1, 0);
    for (int32_t i = 0; i < n; ++i) {
        auto& v = nb[i];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        g.ptr[i + 1] = g.ptr[i] + (int64_t)v.size();
    }
    g.adj.resize(g.ptr[n]);
    for (int32_t i = 0; i < n; ++i) std::copy(nb[i].begin(), nb[i].end(), g.adj.begin() + g.ptr[i]);
    return g;
}

class Dissection {
  public:
    Dissection(const Graph& g, ngs_chol_s* h) : g_(g), h_(h), stamp_(g.n, 0), mark2_(g.n, 0) {}

    // Orders `verts` (all in the current region, marked with `region`): returns the roots of the
    // subtrees created for its connected components.
    std::vector<int32_t> build(std::vector<int32_t> verts) {
        std::vector<int32_t> roots;
        // connected components of the region
        const int32_t reg = ++epoch_;
        for (int32_t v : verts) stamp_[v] = reg;
        std::vector<std::vector<int32_t>> comps;
        const int32_t seen = ++epoch_;
        for (int32_t v : verts) {
            if (stamp_[v] != reg) continue;
            std::vector<int32_t> comp{v};
            stamp_[v] = seen;
            for (size_t q = 0; q < comp.size(); ++q) {
                const int32_t u = comp[q];
                for (int64_t k = g_.ptr[u]; k < g_.ptr[u + 1]; ++k) {
                    const int32_t w = g_.adj[k];
                    if (stamp_[w] == reg) {
                        stamp_[w] = seen;
                        comp.push_back(w);
                    }
                }
            }
            comps.push_back(std::move(comp));
        }
        for (auto& comp : comps) roots.push_back(build_connected(std::move(comp)));
        return roots;
    }

  private:
    // BFS level structure of a connected region rooted at r; returns the levels.
    std::vector<std::vector<int32_t>> bfs_levels(int32_t r, int32_t reg) {
        std::vector<std::vector<int32_t>> levels;
        const int32_t vis = ++epoch_;
        std::vector<int32_t> cur{r};
        mark2_[r] = vis;
        while (!cur.empty()) {
            std::vector<int32_t> nxt;
            for (int32_t u : cur)
                for (int64_t k = g_.ptr[u]; k < g_.ptr[u + 1]; ++k) {
                    const int32_t w = g_.adj[k];
                    if (stamp_[w] == reg && mark2_[w] != vis) {
                        mark2_[w] = vis;
                        nxt.push_back(w);
                    }
                }
            levels.push_back(std::move(cur));
            cur = std::move(nxt);
        }
        return levels;
    }

    int32_t degree_in(int32_t u, int32_t reg) const {
        int32_t d = 0;
        for (int64_t k = g_.ptr[u]; k < g_.ptr[u + 1]; ++k) d += stamp_[g_.adj[k]] == reg;
        return d;
    }

    int32_t make_node(const std::vector<int32_t>& cols, const std::vector<int32_t>& children) {
        HostNode nd;
        nd.start = next_;
        for (int32_t v : cols) {
            h_->perm[next_] = v;
            h_->iperm[v] = next_;
            ++next_;
        }
        nd.end = next_;
        nd.children = children;
        int32_t hgt = 0;
        for (int32_t c : children) hgt = std::max(hgt, h_->nodes[c].height + 1);
        nd.height = hgt;
        const int32_t id = (int32_t)h_->nodes.size();
        for (int32_t c : children) h_->nodes[c].parent = id;
        h_->nodes.push_back(std::move(nd));
        return id;
    }

    int32_t build_connected(std::vector<int32_t> verts) {
        if ((int64_t)verts.size() <= CHOL_LEAF) {
            std::sort(verts.begin(), verts.end());
            return make_node(verts, {});
        }
        const int32_t reg = ++epoch_;
        for (int32_t v : verts) stamp_[v] = reg;
        // pseudo-peripheral root: repeated BFS from a minimum-degree vertex of the last level
        int32_t r = verts[0];
        std::vector<std::vector<int32_t>> lv = bfs_levels(r, reg);
        for (int it = 0; it < 4; ++it) {
            const auto& last = lv.back();
            int32_t best = last[0], bd = INT_MAX;
            for (int32_t u : last) {
                const int32_t d = degree_in(u, reg);
                if (d < bd) { bd = d; best = u; }
            }
            auto lv2 = bfs_levels(best, reg);
            if (lv2.size() <= lv.size()) break;
            lv = std::move(lv2);
        }
        if (lv.size() < 3) {  // no separating level: one dense block
            std::sort(verts.begin(), verts.end());
            return make_node(verts, {});
        }
        // the middle level by vertex count (both sides non-empty)
        const int64_t half = (int64_t)verts.size() / 2;
        int64_t acc = 0;
        size_t m = 1;
        for (size_t i = 0; i < lv.size(); ++i) {
            acc += (int64_t)lv[i].size();
            if (acc >= half) { m = i; break; }
        }
        m = std::min(std::max(m, (size_t)1), lv.size() - 2);
        std::vector<int32_t> a, b;
        for (size_t i = 0; i < m; ++i) a.insert(a.end(), lv[i].begin(), lv[i].end());
        for (size_t i = m + 1; i < lv.size(); ++i) b.insert(b.end(), lv[i].begin(), lv[i].end());
        std::vector<int32_t> sep = lv[m];
        std::sort(sep.begin(), sep.end());
        std::vector<int32_t> children = build(std::move(a));
        std::vector<int32_t> cb = build(std::move(b));
        children.insert(children.end(), cb.begin(), cb.end());
        return make_node(sep, children);
    }

    const Graph& g_;
    ngs_chol_s* h_;
    std::vector<int32_t> stamp_, mark2_;
    int32_t epoch_ = 0;
    int32_t next_ = 0;
};

// Row structure of every front (postorder): the rows of A below the node's columns among its
// columns' neighbours, merged with the children's update rows, keeping those beyond the node.
void symbolic(const Graph& g, ngs_chol_s* h) {
    const int32_t n = g.n;
    std::vector<int32_t> mark(n, -1);
    for (int32_t id = 0; id < (int32_t)h->nodes.size(); ++id) {
        HostNode& nd = h->nodes[id];
        std::vector<int32_t> rows;
        for (int32_t j = nd.start; j < nd.end; ++j) {
            const int32_t v = h->perm[j];
            for (int64_t k = g.ptr[v]; k < g.ptr[v + 1]; ++k) {
                const int32_t r = h->iperm[g.adj[k]];
                if (r >= nd.end && mark[r] != id) {
                    mark[r] = id;
                    rows.push_back(r);
                }
            }
        }
        for (int32_t c : nd.children)
            for (int32_t r : h->nodes[c].rows)
                if (r >= nd.end && mark[r] != id) {
                    mark[r] = id;
                    rows.push_back(r);
                }
        std::sort(rows.begin(), rows.end());
        nd.rows = std::move(rows);
    }
}

}  // namespace

// -------------------------------------------------------------------------------- kernels
__global__ void chol_fill_kernel(double* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void chol_init_sing_kernel(int32_t* sing) { *sing = INT32_MAX; }

__global__ void chol_scatter_kernel(const float* __restrict__ vals, const int64_t* __restrict__ amap, int64_t nnz,
                                    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cols,
                                    const int32_t* __restrict__ iperm, int32_t M, double* __restrict__ pool,
                                    double* __restrict__ diag) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const int64_t pos = amap[k];
    if (pos < 0) return;
    const double v = (double)vals[k];
    atomicAdd(&pool[pos], v);  // padded CSR rows repeat their diagonal entry with value 0
    // original diagonal (new numbering) for the relative pivot test: binary search for the row
    int32_t lo = 0, hi = M;
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) / 2;
        if (rowptr[mid] <= k) lo = mid;
        else hi = mid;
    }
    if (cols[k] == lo) atomicAdd(&diag[iperm[lo]], v);
}

// Extend-add of one child slot of one height into the parents' fronts (a child's update matrix is
// the trailing (f - ns)^2 part of its front; its rows map to parent positions through cmap). A launch
// holds at most one child per parent and the slots run in order: plain adds, in child order.
__global__ void chol_extend_add_kernel(const NodeDesc* __restrict__ desc, const int32_t* __restrict__ children,
                                       const int32_t* __restrict__ cmap, double* __restrict__ pool) {
    const NodeDesc C = desc[children[blockIdx.y]];
    const NodeDesc P = desc[C.parent];
    const int64_t nu = C.f - C.ns;
    const double* __restrict__ F = pool + C.f_off;
    double* __restrict__ G = pool + P.f_off;
    const int32_t* __restrict__ map = cmap + C.cmap_off;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nu * nu; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t / nu, a = t % nu;  // column b, row a (a >= b: lower)
        if (a < b) continue;
        const double u = F[(C.ns + b) * (int64_t)C.f + (C.ns + a)];
        G[(int64_t)map[b] * P.f + map[a]] += u;
    }
}

// Cholesky of the nb x nb diagonal block at column k of each front of the batch (LDS). A pivot
// below tol times the column's original diagonal marks the matrix singular (the first such column
// is kept) and is replaced by that diagonal, so the factorization continues.
__global__ void __launch_bounds__(CHOL_THREADS) chol_potrf_kernel(const NodeDesc* __restrict__ desc,
                                                                  const int32_t* __restrict__ nodes, int k,
                                                                  double* __restrict__ pool,
                                                                  const double* __restrict__ diag, double tol,
                                                                  int32_t* __restrict__ sing) {
    const NodeDesc D = desc[nodes[blockIdx.x]];
    if (k >= D.ns) return;
    const int nb = min(CHOL_NB, D.ns - k);
    const int64_t f = D.f;
    double* __restrict__ F = pool + D.f_off;
    __shared__ double a[CHOL_NB][CHOL_NB + 1];
    for (int t = threadIdx.x; t < nb * nb; t += blockDim.x) {
        const int c = t / nb, r = t % nb;
        a[r][c] = (r >= c) ? F[(k + c) * f + (k + r)] : 0.0;
    }
    __syncthreads();
    for (int c = 0; c < nb; ++c) {
        if (threadIdx.x == 0) {
            double d = a[c][c];
            const double d0 = diag[D.start + k + c];
            if (!(d > tol * fabs(d0))) {
                atomicMin(sing, D.start + k + c);
                d = fabs(d0) > 0.0 ? fabs(d0) : 1.0;
            }
            a[c][c] = sqrt(d);
        }
        __syncthreads();
        const double inv = 1.0 / a[c][c];
        for (int r = c + 1 + threadIdx.x; r < nb; r += blockDim.x) a[r][c] *= inv;
        __syncthreads();
        const int m = nb - c - 1;
        for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
            const int r = c + 1 + t / m, cc = c + 1 + t % m;
            if (cc <= r) a[r][cc] -= a[r][c] * a[cc][c];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < nb * nb; t += blockDim.x) {
        const int c = t / nb, r = t % nb;
        if (r >= c) F[(k + c) * f + (k + r)] = a[r][c];
    }
}

// Panel: rows [k + nb, f) of columns [k, k + nb) become L_rk = F_rk L_kk^-T (one row per thread).
__global__ void __launch_bounds__(CHOL_THREADS) chol_trsm_kernel(const NodeDesc* __restrict__ desc,
                                                                 const int32_t* __restrict__ nodes, int k,
                                                                 double* __restrict__ pool) {
    const NodeDesc D = desc[nodes[blockIdx.y]];
    if (k >= D.ns) return;
    const int nb = min(CHOL_NB, D.ns - k);
    const int64_t f = D.f;
    const int64_t r0 = k + nb;
    if (r0 + (int64_t)blockIdx.x * blockDim.x >= f) return;
    double* __restrict__ F = pool + D.f_off;
    __shared__ double l[CHOL_NB][CHOL_NB + 1];
    for (int t = threadIdx.x; t < nb * nb; t += blockDim.x) {
        const int c = t / nb, r = t % nb;
        l[r][c] = (r >= c) ? F[(k + c) * f + (k + r)] : 0.0;
    }
    __syncthreads();
    const int64_t r = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= f) return;
    double x[CHOL_NB];
#pragma unroll
    for (int c = 0; c < CHOL_NB; ++c) {
        if (c < nb) {
            double v = F[(k + c) * f + r];
            for (int cc = 0; cc < c; ++cc) v -= x[cc] * l[c][cc];
            x[c] = v / l[c][c];
            F[(k + c) * f + r] = x[c];
        }
    }
}

// Schur update of the trailing lower part: F[i][j] -= sum_c L[i][k+c] L[j][k+c], i >= j >= k + nb,
// one 64 x 64 lower tile per workgroup (4 x 4 entries per thread), panels staged in LDS.
__global__ void __launch_bounds__(CHOL_THREADS) chol_syrk_kernel(const NodeDesc* __restrict__ desc,
                                                                 const int32_t* __restrict__ nodes, int k,
                                                                 double* __restrict__ pool) {
    const NodeDesc D = desc[nodes[blockIdx.y]];
    if (k >= D.ns) return;
    const int nb = min(CHOL_NB, D.ns - k);
    const int64_t f = D.f;
    const int64_t base = k + nb;
    const int64_t m = f - base;
    if (m <= 0) return;
    const int64_t T = (m + CHOL_TILE - 1) / CHOL_TILE;
    const int64_t x = blockIdx.x;
    if (x >= T * (T + 1) / 2) return;
    int64_t ti = (int64_t)((sqrt(8.0 * (double)x + 1.0) - 1.0) * 0.5);
    while (ti * (ti + 1) / 2 > x) --ti;
    while ((ti + 1) * (ti + 2) / 2 <= x) ++ti;
    const int64_t tj = x - ti * (ti + 1) / 2;
    double* __restrict__ F = pool + D.f_off;
    __shared__ double pa[CHOL_NB][CHOL_TILE];
    __shared__ double pb[CHOL_NB][CHOL_TILE];
    const int64_t i0 = base + ti * CHOL_TILE, j0 = base + tj * CHOL_TILE;
    for (int t = threadIdx.x; t < CHOL_NB * CHOL_TILE; t += blockDim.x) {
        const int c = t / CHOL_TILE, rr = t % CHOL_TILE;
        const bool okc = c < nb;
        pa[c][rr] = (okc && i0 + rr < f) ? F[(k + c) * f + i0 + rr] : 0.0;
        pb[c][rr] = (okc && j0 + rr < f) ? F[(k + c) * f + j0 + rr] : 0.0;
    }
    __syncthreads();
    const int tr = threadIdx.x % 16, tc = threadIdx.x / 16;
    double acc[4][4] = {};
    for (int c = 0; c < nb; ++c) {
        double av[4], bv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            av[q] = pa[c][tr + 16 * q];
            bv[q] = pb[c][tc + 16 * q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int p = 0; p < 4; ++p) acc[q][p] = fma(av[q], bv[p], acc[q][p]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int64_t i = i0 + tr + 16 * q, j = j0 + tc + 16 * p;
            if (i < f && j < f && i >= j) F[j * f + i] -= acc[q][p];
        }
}

__global__ void chol_permute_in_kernel(const float* __restrict__ b, const double* __restrict__ b64,
                                       const int32_t* __restrict__ perm, int64_t M, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    y[i] = b64 ? b64[perm[i]] : (double)b[perm[i]];
}

__global__ void chol_permute_out_kernel(const double* __restrict__ y, const int32_t* __restrict__ perm, int64_t M,
                                        float* __restrict__ x, double* __restrict__ x64) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    if (x) x[perm[i]] = (float)y[i];
    if (x64) x64[perm[i]] = y[i];
}

// Forward substitution of the fronts of one height (one workgroup per front), multifrontal: the
// front vector (f entries, kept in yv at the front's index-list offset) starts as the right-hand
// side on the pivot rows and zero elsewhere, receives the children's update vectors in child order,
// then y_S = L11^-1 v_S and the update vector v_U -= L21 y_S stay for the parent. No atomics.
__global__ void __launch_bounds__(CHOL_THREADS) chol_forward_kernel(const NodeDesc* __restrict__ desc,
                                                                    const int32_t* __restrict__ nodes,
                                                                    const int32_t* __restrict__ children,
                                                                    const int32_t* __restrict__ cmap,
                                                                    const double* __restrict__ pool,
                                                                    double* __restrict__ y, double* __restrict__ yv) {
    const NodeDesc D = desc[nodes[blockIdx.x]];
    const int64_t f = D.f;
    const double* __restrict__ F = pool + D.f_off;
    double* __restrict__ v = yv + D.idx_off;
    for (int64_t r = threadIdx.x; r < f; r += blockDim.x) v[r] = r < D.ns ? y[D.start + r] : 0.0;
    __syncthreads();
    for (int q = 0; q < D.nchild; ++q) {
        const NodeDesc C = desc[children[D.child_off + q]];
        const int32_t nu = C.f - C.ns;
        const double* __restrict__ u = yv + C.idx_off + C.ns;
        const int32_t* __restrict__ map = cmap + C.cmap_off;
        for (int32_t a = threadIdx.x; a < nu; a += blockDim.x) v[map[a]] += u[a];  // map is one-to-one
        __syncthreads();
    }
    __shared__ double z[CHOL_NB];
    for (int k = 0; k < D.ns; k += CHOL_NB) {
        const int nb = min(CHOL_NB, D.ns - k);
        if (threadIdx.x == 0) {
            for (int c = 0; c < nb; ++c) {
                double t = v[k + c];
                for (int cc = 0; cc < c; ++cc) t -= F[(k + cc) * f + (k + c)] * z[cc];
                z[c] = t / F[(k + c) * f + (k + c)];
                v[k + c] = z[c];
            }
        }
        __syncthreads();
        for (int64_t r = k + nb + threadIdx.x; r < f; r += blockDim.x) {
            double acc = 0.0;
            for (int c = 0; c < nb; ++c) acc += F[(k + c) * f + r] * z[c];
            v[r] -= acc;
        }
        __syncthreads();
    }
    for (int r = threadIdx.x; r < D.ns; r += blockDim.x) y[D.start + r] = v[r];
}

// Backward substitution of the fronts of one height: x_S = L11^-T (y_S - L21^T x[update rows]).
__global__ void __launch_bounds__(CHOL_THREADS) chol_backward_kernel(const NodeDesc* __restrict__ desc,
                                                                     const int32_t* __restrict__ nodes,
                                                                     const int32_t* __restrict__ front_idx,
                                                                     const double* __restrict__ pool,
                                                                     double* __restrict__ y) {
    const NodeDesc D = desc[nodes[blockIdx.x]];
    const int64_t f = D.f;
    const double* __restrict__ F = pool + D.f_off;
    const int32_t* __restrict__ idx = front_idx + D.idx_off;
    // y_S -= L21^T x[rows]: one thread per pivot column
    for (int c = threadIdx.x; c < D.ns; c += blockDim.x) {
        double acc = 0.0;
        for (int64_t r = D.ns; r < f; ++r) acc += F[(int64_t)c * f + r] * y[idx[r]];
        y[D.start + c] -= acc;
    }
    __syncthreads();
    __shared__ double z[CHOL_NB];
    const int nblk = (D.ns + CHOL_NB - 1) / CHOL_NB;
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int k = bk * CHOL_NB;
        const int nb = min(CHOL_NB, D.ns - k);
        if (threadIdx.x == 0) {
            for (int c = nb - 1; c >= 0; --c) {
                double v = y[D.start + k + c];
                for (int cc = c + 1; cc < nb; ++cc) v -= F[(k + c) * f + (k + cc)] * z[cc];
                z[c] = v / F[(k + c) * f + (k + c)];
                y[D.start + k + c] = z[c];
            }
        }
        __syncthreads();
        // earlier pivot columns: y_c -= sum_{r in block} L[r][c] z_r
        for (int c = threadIdx.x; c < k; c += blockDim.x) {
            double acc = 0.0;
            for (int q = 0; q < nb; ++q) acc += F[(int64_t)c * f + (k + q)] * z[q];
            y[D.start + c] -= acc;
        }
        __syncthreads();
    }
}

}  // namespace ngs

using namespace ngs;

extern "C" {

static int chol_analyze(std::vector<int32_t> rp, std::vector<int32_t> cl, int64_t M, ngs_chol** out);

int ngs_chol_analyze(const int32_t* rowptr, const int32_t* cols, int64_t M, ngs_stream_t s, ngs_chol** out) {
    NGS_REQUIRE(out && M >= 0 && M < INT32_MAX, "ngs_chol_analyze: bad arguments");
    NGS_REQUIRE(M == 0 || (rowptr && cols), "ngs_chol_analyze: null CSR arrays");
    std::vector<int32_t> rp(M + 1, 0), cl;
    if (M) {
        hipStream_t st = S(s);
        NGS_CHECK_HIP(hipMemcpyAsync(rp.data(), rowptr, sizeof(int32_t) * (M + 1), hipMemcpyDeviceToHost, st));
        NGS_CHECK_HIP(hipStreamSynchronize(st));
        NGS_REQUIRE(rp[0] == 0 && rp[M] >= 0, "ngs_chol_analyze: malformed row pointer");
        cl.resize(rp[M]);
        if (rp[M]) {
            NGS_CHECK_HIP(hipMemcpyAsync(cl.data(), cols, sizeof(int32_t) * rp[M], hipMemcpyDeviceToHost, st));
            NGS_CHECK_HIP(hipStreamSynchronize(st));
        }
    }
    return chol_analyze(std::move(rp), std::move(cl), M, out);
}

int ngs_chol_analyze_host(const int32_t* rowptr, const int32_t* cols, int64_t M, ngs_chol** out) {
    NGS_REQUIRE(out && M >= 0 && M < INT32_MAX, "ngs_chol_analyze_host: bad arguments");
    NGS_REQUIRE(M == 0 || (rowptr && cols), "ngs_chol_analyze_host: null CSR arrays");
    std::vector<int32_t> rp(M + 1, 0), cl;
    if (M) {
        std::copy(rowptr, rowptr + M + 1, rp.begin());
        NGS_REQUIRE(rp[0] == 0 && rp[M] >= 0, "ngs_chol_analyze_host: malformed row pointer");
        cl.assign(cols, cols + rp[M]);
    }
    return chol_analyze(std::move(rp), std::move(cl), M, out);
}

int ngs_chol_structure(const ngs_chol* h, int64_t* nodes, int32_t* front_idx, int32_t* cmap, int64_t* amap) {
    NGS_REQUIRE(h, "ngs_chol_structure: null handle");
    for (size_t i = 0; i < h->desc.size(); ++i) {
        const NodeDesc& D = h->desc[i];
        int64_t* o = nodes + 8 * i;
        o[0] = D.start; o[1] = D.ns; o[2] = D.f; o[3] = D.parent;
        o[4] = D.f_off; o[5] = D.idx_off; o[6] = D.cmap_off; o[7] = h->nodes[i].height;
    }
    std::copy(h->front_idx.begin(), h->front_idx.end(), front_idx);
    std::copy(h->cmap.begin(), h->cmap.end(), cmap);
    std::copy(h->amap.begin(), h->amap.end(), amap);
    return NGS_OK;
}

static int chol_analyze(std::vector<int32_t> rp, std::vector<int32_t> cl, int64_t M, ngs_chol** out) {
    std::unique_ptr<ngs_chol_s> h(new ngs_chol_s());
    h->M = M;
    const int32_t n = (int32_t)M;
    h->nnz = rp[M];
    h->perm.assign(n, -1);
    h->iperm.assign(n, -1);
    Graph g = build_graph(rp, cl, n);
    {
        Dissection nd(g, h.get());
        std::vector<int32_t> all(n);
        std::iota(all.begin(), all.end(), 0);
        nd.build(std::move(all));
    }
    symbolic(g, h.get());
    // fronts, index lists and the pool layout
    const int32_t nn = (int32_t)h->nodes.size();
    h->desc.resize(nn);
    int64_t pool = 0, idx = 0, fill = 0;
    int32_t H = 0;
    for (int32_t id = 0; id < nn; ++id) {
        const HostNode& nd = h->nodes[id];
        NodeDesc& D = h->desc[id];
        D.ns = nd.end - nd.start;
        D.f = D.ns + (int32_t)nd.rows.size();
        D.start = nd.start;
        D.parent = nd.parent;
        D.f_off = pool;
        D.idx_off = idx;
        D.cmap_off = -1;
        pool += (int64_t)D.f * D.f;
        idx += D.f;
        fill += (int64_t)D.ns * D.f - (int64_t)D.ns * (D.ns - 1) / 2;
        for (int32_t j = nd.start; j < nd.end; ++j) h->front_idx.push_back(j);
        h->front_idx.insert(h->front_idx.end(), nd.rows.begin(), nd.rows.end());
        H = std::max(H, nd.height);
    }
    h->pool_doubles = pool;
    h->fill = fill;
    // child -> parent front positions
    for (int32_t id = 0; id < nn; ++id) {
        const HostNode& nd = h->nodes[id];
        if (nd.parent < 0) continue;
        const HostNode& p = h->nodes[nd.parent];
        h->desc[id].cmap_off = (int64_t)h->cmap.size();
        for (int32_t r : nd.rows) {
            int32_t pos;
            if (r < p.end) {
                NGS_REQUIRE(r >= p.start, "ngs_chol_analyze: update row outside the parent front");
                pos = r - p.start;
            } else {
                auto it = std::lower_bound(p.rows.begin(), p.rows.end(), r);
                NGS_REQUIRE(it != p.rows.end() && *it == r, "ngs_chol_analyze: update row outside the parent front");
                pos = (p.end - p.start) + (int32_t)(it - p.rows.begin());
            }
            h->cmap.push_back(pos);
        }
    }
    // A's lower entries in the caller's numbering (the triangle cuSOLVER reads; the other one may
    // differ from it by rounding) -> pool positions of the permuted lower triangle
    std::vector<int32_t> owner(n);
    for (int32_t id = 0; id < nn; ++id)
        for (int32_t j = h->nodes[id].start; j < h->nodes[id].end; ++j) owner[j] = id;
    h->amap.assign(h->nnz, -1);
    for (int32_t io = 0; io < n; ++io)
        for (int32_t k = rp[io]; k < rp[io + 1]; ++k) {
            const int32_t jo = cl[k];
            if (jo < 0 || jo >= n || jo > io) continue;
            const int32_t i = std::max(h->iperm[io], h->iperm[jo]), j = std::min(h->iperm[io], h->iperm[jo]);
            const int32_t id = owner[j];
            const HostNode& nd = h->nodes[id];
            const NodeDesc& D = h->desc[id];
            int32_t lr;
            if (i < nd.end) lr = i - nd.start;
            else {
                auto it = std::lower_bound(nd.rows.begin(), nd.rows.end(), i);
                NGS_REQUIRE(it != nd.rows.end() && *it == i, "ngs_chol_analyze: entry outside its front");
                lr = D.ns + (int32_t)(it - nd.rows.begin());
            }
            h->amap[k] = D.f_off + (int64_t)(j - nd.start) * D.f + lr;
        }
    // batches by height, children lists and the extend-add slots
    h->level_nodes.assign(H + 1, {});
    h->level_max_ns.assign(H + 1, 0);
    h->level_max_f.assign(H + 1, 0);
    std::vector<std::vector<int32_t>> kids(nn);
    for (int32_t id = 0; id < nn; ++id) {
        const int32_t ht = h->nodes[id].height;
        h->level_nodes[ht].push_back(id);
        h->level_max_ns[ht] = std::max(h->level_max_ns[ht], h->desc[id].ns);
        h->level_max_f[ht] = std::max(h->level_max_f[ht], h->desc[id].f);
        for (int32_t c : h->nodes[id].children)
            if (h->desc[c].f > h->desc[c].ns) kids[id].push_back(c);
        h->desc[id].child_off = (int64_t)h->children.size();
        h->desc[id].nchild = (int32_t)kids[id].size();
        h->desc[id].pad = 0;
        h->children.insert(h->children.end(), kids[id].begin(), kids[id].end());
    }
    for (const auto& l : h->level_nodes)
        NGS_REQUIRE(l.size() <= 65535, "ngs_chol_analyze: more than 65535 supernodes of one height");
    h->slot_off.assign(H + 1, {});
    h->slot_n.assign(H + 1, {});
    h->slot_maxu.assign(H + 1, {});
    for (int32_t ht = 0; ht <= H; ++ht) {
        h->lvl_node_off.push_back((int64_t)h->lvl.size());
        h->lvl.insert(h->lvl.end(), h->level_nodes[ht].begin(), h->level_nodes[ht].end());
        size_t maxk = 0;
        for (int32_t id : h->level_nodes[ht]) maxk = std::max(maxk, kids[id].size());
        for (size_t k = 0; k < maxk; ++k) {
            h->slot_off[ht].push_back((int64_t)h->lvl.size());
            int64_t cnt = 0, maxu = 0;
            for (int32_t id : h->level_nodes[ht])
                if (kids[id].size() > k) {
                    const int32_t c = kids[id][k];
                    h->lvl.push_back(c);
                    ++cnt;
                    maxu = std::max<int64_t>(maxu, h->desc[c].f - h->desc[c].ns);
                }
            h->slot_n[ht].push_back(cnt);
            h->slot_maxu[ht].push_back(maxu);
        }
    }
    h->lvl_perm_off = (int64_t)h->lvl.size();
    h->lvl.insert(h->lvl.end(), h->perm.begin(), h->perm.end());
    h->lvl.insert(h->lvl.end(), h->iperm.begin(), h->iperm.end());
    // workspace layout
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t off = 0;
    h->off_pool = off; off += al(sizeof(double) * (size_t)pool);
    h->off_amap = off; off += al(sizeof(int64_t) * (size_t)h->nnz);
    h->off_desc = off; off += al(sizeof(NodeDesc) * (size_t)nn);
    h->off_idx = off; off += al(sizeof(int32_t) * (size_t)h->front_idx.size());
    h->off_cmap = off; off += al(sizeof(int32_t) * (size_t)h->cmap.size());
    h->off_kids = off; off += al(sizeof(int32_t) * (size_t)h->children.size());
    h->off_lvl = off; off += al(sizeof(int32_t) * h->lvl.size());
    h->off_diag = off; off += al(sizeof(double) * (size_t)n);
    h->off_sing = off; off += 256;
    h->off_y = off; off += al(sizeof(double) * (size_t)n);
    h->off_yv = off; off += al(sizeof(double) * (size_t)h->front_idx.size());
    h->total = off;
    *out = h.release();
    return NGS_OK;
}

int64_t ngs_chol_workspace_bytes(const ngs_chol* h) { return h ? (int64_t)h->total : -1; }

int ngs_chol_info(const ngs_chol* h, int64_t* stats) {
    NGS_REQUIRE(h && stats, "ngs_chol_info: null argument");
    int32_t maxf = 0;
    for (auto f : h->level_max_f) maxf = std::max(maxf, f);
    stats[0] = (int64_t)h->nodes.size();
    stats[1] = (int64_t)h->level_nodes.size();
    stats[2] = maxf;
    stats[3] = h->fill;
    stats[4] = h->pool_doubles;
    stats[5] = (int64_t)h->front_idx.size();
    stats[6] = (int64_t)h->cmap.size();
    stats[7] = h->nnz;
    return NGS_OK;
}

int ngs_chol_permutation(const ngs_chol* h, int32_t* perm) {
    NGS_REQUIRE(h && (perm || h->M == 0), "ngs_chol_permutation: null argument");
    std::copy(h->perm.begin(), h->perm.end(), perm);
    return NGS_OK;
}

void ngs_chol_free(ngs_chol* h) { delete h; }

static int chol_upload(ngs_chol* h, char* ws, hipStream_t st) {
    if (h->uploaded == ws) return NGS_OK;
    auto up = [&](size_t off, const void* src, size_t bytes) -> int {
        if (bytes) NGS_CHECK_HIP(hipMemcpyAsync(ws + off, src, bytes, hipMemcpyHostToDevice, st));
        return NGS_OK;
    };
    if (int r = up(h->off_amap, h->amap.data(), sizeof(int64_t) * h->amap.size())) return r;
    if (int r = up(h->off_desc, h->desc.data(), sizeof(NodeDesc) * h->desc.size())) return r;
    if (int r = up(h->off_idx, h->front_idx.data(), sizeof(int32_t) * h->front_idx.size())) return r;
    if (int r = up(h->off_cmap, h->cmap.data(), sizeof(int32_t) * h->cmap.size())) return r;
    if (int r = up(h->off_kids, h->children.data(), sizeof(int32_t) * h->children.size())) return r;
    if (int r = up(h->off_lvl, h->lvl.data(), sizeof(int32_t) * h->lvl.size())) return r;
    // the pageable sources must stay valid until the copies land
    NGS_CHECK_HIP(hipStreamSynchronize(st));
    h->uploaded = ws;
    return NGS_OK;
}

int ngs_chol_factor(ngs_chol* h, const int32_t* rowptr, const int32_t* cols, const float* vals, double tol,
                    void* workspace, ngs_stream_t s, int64_t* singularity) {
    NGS_REQUIRE(h && workspace, "ngs_chol_factor: bad arguments");
    if (singularity) *singularity = -1;
    if (h->M == 0) return NGS_OK;
    NGS_REQUIRE(rowptr && cols && vals, "ngs_chol_factor: null CSR arrays");
    hipStream_t st = S(s);
    char* ws = reinterpret_cast<char*>(workspace);
    if (int r = chol_upload(h, ws, st)) return r;
    double* pool = reinterpret_cast<double*>(ws + h->off_pool);
    const int64_t* amap = reinterpret_cast<const int64_t*>(ws + h->off_amap);
    const NodeDesc* desc = reinterpret_cast<const NodeDesc*>(ws + h->off_desc);
    const int32_t* cmap = reinterpret_cast<const int32_t*>(ws + h->off_cmap);
    const int32_t* lvl = reinterpret_cast<const int32_t*>(ws + h->off_lvl);
    const int32_t* iperm = lvl + h->lvl_perm_off + h->M;
    double* diag = reinterpret_cast<double*>(ws + h->off_diag);
    int32_t* sing = reinterpret_cast<int32_t*>(ws + h->off_sing);
    chol_fill_kernel<<<2048, 256, 0, st>>>(pool, h->pool_doubles);
    chol_fill_kernel<<<grid1d(h->M, 256), 256, 0, st>>>(diag, h->M);
    chol_init_sing_kernel<<<1, 1, 0, st>>>(sing);
    if (h->nnz)
        chol_scatter_kernel<<<grid1d(h->nnz, 256), 256, 0, st>>>(vals, amap, h->nnz, rowptr, cols, iperm,
                                                                 (int32_t)h->M, pool, diag);
    NGS_CHECK_LAUNCH();
    for (size_t ht = 0; ht < h->level_nodes.size(); ++ht) {
        const int32_t* nodes = lvl + h->lvl_node_off[ht];
        const unsigned nnode = (unsigned)h->level_nodes[ht].size();
        for (size_t k = 0; k < h->slot_off[ht].size(); ++k) {
            const int64_t maxu = h->slot_maxu[ht][k];
            const unsigned gx = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (maxu * maxu + 255) / 256), 1024);
            chol_extend_add_kernel<<<dim3(gx, (unsigned)h->slot_n[ht][k]), 256, 0, st>>>(desc, lvl + h->slot_off[ht][k],
                                                                                      cmap, pool);
            NGS_CHECK_LAUNCH();
        }
        const int maxns = h->level_max_ns[ht];
        const int64_t maxf = h->level_max_f[ht];
        for (int k = 0; k < maxns; k += CHOL_NB) {
            chol_potrf_kernel<<<nnode, CHOL_THREADS, 0, st>>>(desc, nodes, k, pool, diag, tol, sing);
            const int64_t rows = maxf - k;
            if (rows > 0) {
                const unsigned gx = (unsigned)((rows + CHOL_THREADS - 1) / CHOL_THREADS);
                chol_trsm_kernel<<<dim3(gx, nnode), CHOL_THREADS, 0, st>>>(desc, nodes, k, pool);
                const int64_t T = (rows + CHOL_TILE - 1) / CHOL_TILE;
                chol_syrk_kernel<<<dim3((unsigned)(T * (T + 1) / 2), nnode), CHOL_THREADS, 0, st>>>(desc, nodes, k,
                                                                                                     pool);
            }
            NGS_CHECK_LAUNCH();
        }
    }
    if (singularity) {
        int32_t sv = 0;
        NGS_CHECK_HIP(hipMemcpyAsync(&sv, sing, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        NGS_CHECK_HIP(hipStreamSynchronize(st));
        *singularity = (sv == INT32_MAX) ? -1 : (int64_t)h->perm[sv];
    }
    return NGS_OK;
}

int ngs_chol_solve(const ngs_chol* h, const void* workspace, const float* b, const double* b64, float* x,
                   double* x64, ngs_stream_t s) {
    NGS_REQUIRE(h && workspace, "ngs_chol_solve: bad arguments");
    if (h->M == 0) return NGS_OK;
    NGS_REQUIRE((b || b64) && (x || x64), "ngs_chol_solve: null vector");
    NGS_REQUIRE(h->uploaded == workspace, "ngs_chol_solve: the workspace holds no factorization of this handle");
    hipStream_t st = S(s);
    const char* ws = reinterpret_cast<const char*>(workspace);
    const double* pool = reinterpret_cast<const double*>(ws + h->off_pool);
    const NodeDesc* desc = reinterpret_cast<const NodeDesc*>(ws + h->off_desc);
    const int32_t* fidx = reinterpret_cast<const int32_t*>(ws + h->off_idx);
    const int32_t* lvl = reinterpret_cast<const int32_t*>(ws + h->off_lvl);
    const int32_t* perm = lvl + h->lvl_perm_off;
    const int32_t* kids = reinterpret_cast<const int32_t*>(ws + h->off_kids);
    const int32_t* cmap = reinterpret_cast<const int32_t*>(ws + h->off_cmap);
    double* y = reinterpret_cast<double*>(const_cast<char*>(ws) + h->off_y);
    double* yv = reinterpret_cast<double*>(const_cast<char*>(ws) + h->off_yv);
    chol_permute_in_kernel<<<grid1d(h->M, 256), 256, 0, st>>>(b, b64, perm, h->M, y);
    for (size_t ht = 0; ht < h->level_nodes.size(); ++ht)
        chol_forward_kernel<<<(unsigned)h->level_nodes[ht].size(), CHOL_THREADS, 0, st>>>(
            desc, lvl + h->lvl_node_off[ht], kids, cmap, pool, y, yv);
    for (size_t ht = h->level_nodes.size(); ht-- > 0;)
        chol_backward_kernel<<<(unsigned)h->level_nodes[ht].size(), CHOL_THREADS, 0, st>>>(
            desc, lvl + h->lvl_node_off[ht], fidx, pool, y);
    chol_permute_out_kernel<<<grid1d(h->M, 256), 256, 0, st>>>(y, perm, h->M, x, x64);
    NGS_CHECK_LAUNCH();
    return NGS_OK;
}

}  // extern "C"
