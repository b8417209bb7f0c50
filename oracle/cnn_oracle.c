/*
 * C restatement of the CommonNN fit hot path (cnnclustering 0.5.0).
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for large inputs and the timed
 * CPU baseline ("port").  Only tests/, the smoke entry point and bench.py's
 * cpu_baseline leg load it.  The product library never links or calls it.
 *
 * Follows the reference files:
 *   metric   _types.pyx:2253-2267  total = total + (a_k - b_k)*(a_k - b_k), k ascending,
 *                                  every operation separately rounded (no FMA:
 *                                  built with -ffp-contract=off)
 *   radius   _types.pyx:2286-2287  r2 = r*r (the caller passes r2)
 *   getter   _types.pyx:1692-1710  N(i) = { j : s_ij <= r2 }, ascending j, i included
 *   enough   _types.pyx:1561-1564  |N(i)| > n_member_cutoff
 *   checker  _types.pyx:2654-2706  two-pointer merge on sorted lists, early exit at c
 *   fitter   _fit.pyx:408-491      BFS (oracle_labels_bfs, literal) and its
 *                                  connected-components equivalent (oracle_labels_cc)
 *   sort     _types.pyx:212-286    sort_by_size (size desc, first occurrence asc)
 *
 * The optional uniform-grid prefilter (d <= 3) only skips pairs whose
 * per-dimension separation already exceeds the cell size, which is at least r
 * with a margin for rounding; every pair that can satisfy s_ij <= r2 is still
 * evaluated with the reference arithmetic, so the sets are unchanged.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef int64_t i64;
typedef int32_t i32;

static inline double reduced_distance(const double *a, const double *b, int d) {
    double total = 0.0;
    for (int k = 0; k < d; ++k) {
        double t = a[k] - b[k];
        double sq = t * t;
        total = total + sq;
    }
    return total;
}

/* ------------------------------------------------------------ grid prefilter */

typedef struct {
    int d;
    i64 n;
    double lo[3];
    double cs;
    i64 ncell[3];
    i64 *order;   /* points sorted by cell key */
    i64 *keys;    /* sorted keys */
} grid_t;

static i64 cell_of(const grid_t *g, const double *x, int k) {
    i64 c = (i64)floor((x[k] - g->lo[k]) / g->cs);
    if (c < 0) c = 0;
    if (c >= g->ncell[k]) c = g->ncell[k] - 1;
    return c;
}

static i64 key_of_cells(const grid_t *g, const i64 *c) {
    i64 key = 0;
    for (int k = 0; k < g->d; ++k) key = key * g->ncell[k] + c[k];
    return key;
}

static const i64 *g_sort_keys;
static int cmp_by_key(const void *a, const void *b) {
    i64 ka = g_sort_keys[*(const i64 *)a], kb = g_sort_keys[*(const i64 *)b];
    if (ka < kb) return -1;
    if (ka > kb) return 1;
    i64 ia = *(const i64 *)a, ib = *(const i64 *)b;
    return (ia > ib) - (ia < ib);
}

static int grid_build(grid_t *g, const double *X, i64 n, int d, double r2) {
    g->d = d;
    g->n = n;
    double hi[3];
    for (int k = 0; k < d; ++k) { g->lo[k] = INFINITY; hi[k] = -INFINITY; }
    for (i64 i = 0; i < n; ++i)
        for (int k = 0; k < d; ++k) {
            double v = X[i * d + k];
            if (!(v == v) || isinf(v)) return -1;  /* non-finite: no grid */
            if (v < g->lo[k]) g->lo[k] = v;
            if (v > hi[k]) hi[k] = v;
        }
    double r = sqrt(r2);
    double extent = 0.0;
    for (int k = 0; k < d; ++k) extent = fmax(extent, hi[k] - g->lo[k]);
    g->cs = r * (1.0 + 1e-6) + extent * 1e-12 + 1e-300;
    /* keep the number of cells bounded; a larger cell only weakens the filter */
    for (;;) {
        double total = 1.0;
        for (int k = 0; k < d; ++k) {
            g->ncell[k] = (i64)floor((hi[k] - g->lo[k]) / g->cs) + 1;
            total *= (double)g->ncell[k];
        }
        if (total <= 4.0 * (double)n + 64.0) break;
        g->cs *= 2.0;
    }
    i64 *raw = (i64 *)malloc(sizeof(i64) * n);
    g->order = (i64 *)malloc(sizeof(i64) * n);
    g->keys = (i64 *)malloc(sizeof(i64) * n);
    if (!raw || !g->order || !g->keys) return -2;
    #pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
        i64 c[3];
        for (int k = 0; k < d; ++k) c[k] = cell_of(g, X + i * d, k);
        raw[i] = key_of_cells(g, c);
        g->order[i] = i;
    }
    g_sort_keys = raw;
    qsort(g->order, n, sizeof(i64), cmp_by_key);
    for (i64 i = 0; i < n; ++i) g->keys[i] = raw[g->order[i]];
    free(raw);
    return 0;
}

static void grid_free(grid_t *g) {
    free(g->order);
    free(g->keys);
}

static i64 lower_key(const grid_t *g, i64 key) {
    i64 lo = 0, hi = g->n;
    while (lo < hi) {
        i64 mid = (lo + hi) >> 1;
        if (g->keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static int cmp_i32(const void *a, const void *b) {
    i32 x = *(const i32 *)a, y = *(const i32 *)b;
    return (x > y) - (x < y);
}

/* Visit every candidate j of point i (grid neighbours), calling emit for s_ij <= r2.
 * Returns the number of neighbours.  When out != NULL the neighbours are written
 * there and sorted ascending. */
static i64 grid_row(const grid_t *g, const double *X, int d, double r2, i64 i, i32 *out) {
    const double *xi = X + i * d;
    i64 c0[3], c[3];
    for (int k = 0; k < d; ++k) c0[k] = cell_of(g, xi, k);
    i64 cnt = 0;
    int ncomb = 1;
    for (int k = 0; k < d; ++k) ncomb *= 3;
    for (int comb = 0; comb < ncomb; ++comb) {
        int t = comb, ok = 1;
        for (int k = d - 1; k >= 0; --k) {
            c[k] = c0[k] + (t % 3) - 1;
            t /= 3;
            if (c[k] < 0 || c[k] >= g->ncell[k]) ok = 0;
        }
        if (!ok) continue;
        i64 key = key_of_cells(g, c);
        for (i64 p = lower_key(g, key); p < g->n && g->keys[p] == key; ++p) {
            i64 j = g->order[p];
            double s = reduced_distance(xi, X + j * d, d);
            if (s <= r2) {
                if (out) out[cnt] = (i32)j;
                ++cnt;
            }
        }
    }
    if (out) qsort(out, cnt, sizeof(i32), cmp_i32);
    return cnt;
}

/* ------------------------------------------------------------ neighbourhoods */

/* deg[i] = |N(i)|.  use_grid: 0 brute force, 1 grid prefilter (d <= 3). */
int oracle_neighbours_count(const double *X, i64 n, int d, double r2, int use_grid, i64 *deg) {
    if (use_grid && d >= 1 && d <= 3 && n > 0) {
        grid_t g;
        int rc = grid_build(&g, X, n, d, r2);
        if (rc == 0) {
            #pragma omp parallel for schedule(dynamic, 256)
            for (i64 i = 0; i < n; ++i) deg[i] = grid_row(&g, X, d, r2, i, NULL);
            grid_free(&g);
            return 0;
        }
        if (rc == -2) return -2;
    }
    #pragma omp parallel for schedule(dynamic, 64)
    for (i64 i = 0; i < n; ++i) {
        i64 cnt = 0;
        const double *xi = X + i * d;
        for (i64 j = 0; j < n; ++j)
            if (reduced_distance(xi, X + j * d, d) <= r2) ++cnt;
        deg[i] = cnt;
    }
    return 0;
}

/* col[row_ptr[i] .. row_ptr[i+1]) = N(i), ascending. */
int oracle_neighbours_fill(const double *X, i64 n, int d, double r2, int use_grid,
                           const i64 *row_ptr, i32 *col) {
    if (use_grid && d >= 1 && d <= 3 && n > 0) {
        grid_t g;
        int rc = grid_build(&g, X, n, d, r2);
        if (rc == 0) {
            #pragma omp parallel for schedule(dynamic, 256)
            for (i64 i = 0; i < n; ++i) grid_row(&g, X, d, r2, i, col + row_ptr[i]);
            grid_free(&g);
            return 0;
        }
        if (rc == -2) return -2;
    }
    #pragma omp parallel for schedule(dynamic, 64)
    for (i64 i = 0; i < n; ++i) {
        i64 p = row_ptr[i];
        const double *xi = X + i * d;
        for (i64 j = 0; j < n; ++j)
            if (reduced_distance(xi, X + j * d, d) <= r2) col[p++] = (i32)j;
    }
    return 0;
}

/* ------------------------------------------------------------ brute force, exact early exit */

/* Brute-force neighbour rows for large inputs (the same sets as
 * oracle_neighbours_fill without the grid).  Two exact shortcuts, neither of
 * which changes a single decision:
 *   - two columns per SSE2 vector: the vector subtract, multiply and add are
 *     the same correctly rounded IEEE operations as the scalar ones (built
 *     with -ffp-contract=off, so no FMA), applied to each lane in ascending k;
 *   - early exit: every term fl(t*t) is >= +0 and round-to-nearest addition of
 *     a non-negative term never yields less than the running sum (t + x >= t
 *     and rounding is monotone), so partial sums never decrease.  Once the
 *     partial sums of all columns of a block exceed r2, their full sums do
 *     too and none of them is a neighbour.  A NaN partial sum never compares
 *     greater and keeps its block going; its full sum is NaN, which fails
 *     s <= r2 as in the reference.
 * rows: the row indices to compute (NULL: all n, in order); row_ptr receives
 * n_rows + 1 offsets and *col_out a malloc'ed array of the entries, every row
 * ascending (free it with oracle_free). */
typedef double v2d __attribute__((vector_size(16)));

#define BF_COLS 8      /* columns per block: 4 vectors of 2 */
#define BF_CHUNK 8     /* dimensions between early-exit checks */

static i64 bf_row(const double *XT, i64 n, i64 n_pad, int d, double r2, const double *xi, i32 *out) {
    i64 cnt = 0;
    for (i64 j0 = 0; j0 < n; j0 += BF_COLS) {
        v2d s0 = {0.0, 0.0}, s1 = s0, s2 = s0, s3 = s0;
        int k = 0;
        for (;;) {
            const int kend = k + BF_CHUNK < d ? k + BF_CHUNK : d;
            for (; k < kend; ++k) {
                const double *col = XT + (i64)k * n_pad + j0;
                const v2d a = {xi[k], xi[k]};
                v2d t0 = a - *(const v2d *)(col + 0);
                v2d t1 = a - *(const v2d *)(col + 2);
                v2d t2 = a - *(const v2d *)(col + 4);
                v2d t3 = a - *(const v2d *)(col + 6);
                t0 = t0 * t0; t1 = t1 * t1; t2 = t2 * t2; t3 = t3 * t3;
                s0 = s0 + t0; s1 = s1 + t1; s2 = s2 + t2; s3 = s3 + t3;
            }
            if (k >= d) break;
            if (s0[0] > r2 && s0[1] > r2 && s1[0] > r2 && s1[1] > r2 &&
                s2[0] > r2 && s2[1] > r2 && s3[0] > r2 && s3[1] > r2)
                break;
        }
        if (k < d) continue;  /* every column of the block is beyond r2 */
        const double s[BF_COLS] = {s0[0], s0[1], s1[0], s1[1], s2[0], s2[1], s3[0], s3[1]};
        for (int c = 0; c < BF_COLS; ++c) {
            const i64 j = j0 + c;
            if (j < n && s[c] <= r2) {
                if (out) out[cnt] = (i32)j;
                ++cnt;
            }
        }
    }
    return cnt;
}

int oracle_neighbours_bf(const double *X, i64 n, int d, double r2, const i64 *rows, i64 n_rows,
                         i64 *row_ptr, i32 **col_out) {
    *col_out = NULL;
    const i64 n_pad = (n + BF_COLS - 1) / BF_COLS * BF_COLS;
    double *XT = (double *)aligned_alloc(64, sizeof(double) * (size_t)(n_pad > 0 ? n_pad : 1) * (d > 0 ? d : 1));
    i32 **lists = (i32 **)calloc((size_t)(n_rows > 0 ? n_rows : 1), sizeof(i32 *));
    i64 *deg = (i64 *)calloc((size_t)(n_rows > 0 ? n_rows : 1), sizeof(i64));
    if (!XT || !lists || !deg) { free(XT); free(lists); free(deg); return -2; }
    /* columns padded with +inf: (x - inf)^2 = inf, never a neighbour */
    #pragma omp parallel for schedule(static)
    for (i64 j = 0; j < n_pad; ++j)
        for (int k = 0; k < d; ++k) XT[(i64)k * n_pad + j] = j < n ? X[j * d + k] : INFINITY;
    int failed = 0;
    #pragma omp parallel
    {
        i32 *buf = (i32 *)malloc(sizeof(i32) * (size_t)(n > 0 ? n : 1));
        if (!buf) {
            #pragma omp atomic write
            failed = 1;
        }
        #pragma omp for schedule(dynamic, 16)
        for (i64 t = 0; t < n_rows; ++t) {
            if (!buf) continue;
            const i64 i = rows ? rows[t] : t;
            const i64 cnt = bf_row(XT, n, n_pad, d, r2, X + i * d, buf);
            lists[t] = (i32 *)malloc(sizeof(i32) * (size_t)(cnt > 0 ? cnt : 1));
            if (!lists[t]) {
                #pragma omp atomic write
                failed = 1;
                continue;
            }
            memcpy(lists[t], buf, sizeof(i32) * (size_t)cnt);
            deg[t] = cnt;
        }
        free(buf);
    }
    free(XT);
    i32 *col = NULL;
    if (!failed) {
        row_ptr[0] = 0;
        for (i64 t = 0; t < n_rows; ++t) row_ptr[t + 1] = row_ptr[t] + deg[t];
        col = (i32 *)malloc(sizeof(i32) * (size_t)(row_ptr[n_rows] > 0 ? row_ptr[n_rows] : 1));
        if (!col) failed = 1;
    }
    if (!failed) {
        #pragma omp parallel for schedule(dynamic, 64)
        for (i64 t = 0; t < n_rows; ++t) memcpy(col + row_ptr[t], lists[t], sizeof(i32) * (size_t)deg[t]);
    }
    for (i64 t = 0; t < n_rows; ++t) free(lists[t]);
    free(lists);
    free(deg);
    if (failed) { free(col); return -2; }
    *col_out = col;
    return 0;
}

void oracle_free(void *p) { free(p); }

/* ------------------------------------------------------------ similarity */

/* Two-pointer merge with early exit at c (_types.pyx:2654-2706).  *la / *lb
 * receive the number of elements consumed from each list. */
static int screen_sorted(const i32 *a, i64 na, const i32 *b, i64 nb, i64 c, i64 *la, i64 *lb) {
    *la = 0; *lb = 0;
    if (c == 0) return 1;
    if (na == 0 || nb == 0) return 0;
    i64 ia = 0, ib = 0, common = 0;
    int pass = 0;
    for (;;) {
        i32 x = a[ia], y = b[ib];
        if (x == y) {
            if (++common == c) { pass = 1; break; }
            ++ia; ++ib;
            if (ia == na || ib == nb) break;
        } else if (x < y) {
            if (++ia == na) break;
        } else {
            if (++ib == nb) break;
        }
    }
    *la = ia < na ? ia + 1 : na;
    *lb = ib < nb ? ib + 1 : nb;
    return pass;
}

int oracle_check_similarity(const i32 *a, i64 na, const i32 *b, i64 nb, i64 c) {
    i64 la, lb;
    return screen_sorted(a, na, b, nb, c, &la, &lb);
}

/* ------------------------------------------------------------ labelling */

/* Literal BFS (_fit.pyx:408-491) over precomputed sorted neighbourhoods.
 * labels / consider are updated in place.  Sequential by construction. */
int oracle_labels_bfs(const i64 *row_ptr, const i32 *col, i64 n, i64 c_sim, i64 m_core,
                      uint8_t *consider, i64 current_start, i64 *labels) {
    i64 *queue = (i64 *)malloc(sizeof(i64) * (n > 0 ? n : 1));
    if (!queue) return -2;
    i64 current = current_start;
    for (i64 init = 0; init < n; ++init) {
        if (consider[init] == 0) continue;
        consider[init] = 0;
        i64 point = init;
        if (!(row_ptr[point + 1] - row_ptr[point] > m_core)) continue;
        labels[init] = current;
        i64 qh = 0, qt = 0;
        for (;;) {
            const i32 *na = col + row_ptr[point];
            i64 nna = row_ptr[point + 1] - row_ptr[point];
            for (i64 mi = 0; mi < nna; ++mi) {
                i64 member = na[mi];
                if (consider[member] == 0) continue;
                const i32 *nb = col + row_ptr[member];
                i64 nnb = row_ptr[member + 1] - row_ptr[member];
                if (!(nnb > m_core)) { consider[member] = 0; continue; }
                i64 la, lb;
                if (screen_sorted(na, nna, nb, nnb, c_sim, &la, &lb)) {
                    consider[member] = 0;
                    labels[member] = current;
                    queue[qt++] = member;
                }
            }
            if (qh == qt) break;
            point = queue[qh++];
        }
        ++current;
    }
    free(queue);
    return 0;
}

static i64 uf_find(i64 *parent, i64 x) {
    for (;;) {
        i64 p = __atomic_load_n(&parent[x], __ATOMIC_RELAXED);
        if (p == x) return x;
        i64 gp = __atomic_load_n(&parent[p], __ATOMIC_RELAXED);
        if (gp != p) __atomic_compare_exchange_n(&parent[x], &p, gp, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
        x = gp;
    }
}

static void uf_union(i64 *parent, i64 a, i64 b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a > b) { i64 t = a; a = b; b = t; }
        i64 expected = b;
        if (__atomic_compare_exchange_n(&parent[b], &expected, a, 0, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED))
            return;
    }
}

/* Connected-components form of the same fit, parallel over rows.  Core points
 * are considered points with more than m_core neighbours; an edge (i, j),
 * j in N(i), passes when both are core and the merge count reaches c_sim.
 * Components are numbered from current_start by ascending minimum index.
 * Every consider flag ends at 0, as after the BFS. */
int oracle_labels_cc(const i64 *row_ptr, const i32 *col, i64 n, i64 c_sim, i64 m_core,
                     uint8_t *consider, i64 current_start, i64 *labels) {
    i64 *parent = (i64 *)malloc(sizeof(i64) * (n > 0 ? n : 1));
    uint8_t *core = (uint8_t *)malloc(n > 0 ? n : 1);
    if (!parent || !core) return -2;
    #pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
        parent[i] = i;
        core[i] = consider[i] && (row_ptr[i + 1] - row_ptr[i] > m_core);
    }
    #pragma omp parallel for schedule(dynamic, 64)
    for (i64 i = 0; i < n; ++i) {
        if (!core[i]) continue;
        const i32 *na = col + row_ptr[i];
        i64 nna = row_ptr[i + 1] - row_ptr[i];
        for (i64 mi = 0; mi < nna; ++mi) {
            i64 j = na[mi];
            if (j <= i || !core[j]) continue;
            if (uf_find(parent, i) == uf_find(parent, j)) continue;  /* already joined */
            i64 la, lb;
            if (screen_sorted(na, nna, col + row_ptr[j], row_ptr[j + 1] - row_ptr[j], c_sim, &la, &lb))
                uf_union(parent, i, j);
        }
    }
    i64 *label_of_root = (i64 *)malloc(sizeof(i64) * (n > 0 ? n : 1));
    if (!label_of_root) return -2;
    i64 current = current_start;
    for (i64 i = 0; i < n; ++i) {
        if (!core[i]) continue;
        i64 r = uf_find(parent, i);
        if (r == i) label_of_root[i] = current++;
        labels[i] = label_of_root[r];
    }
    memset(consider, 0, n);
    free(label_of_root);
    free(parent);
    free(core);
    return 0;
}

/* Canonical merge work of the similarity phase without any skipping:
 * sum over candidate edges (i < j, both core, j in N(i)) of the elements the
 * early-exit two-pointer merge consumes from both lists.  Also returns the
 * number of candidate edges and of passing edges. */
int oracle_merge_work(const i64 *row_ptr, const i32 *col, i64 n, i64 c_sim, i64 m_core,
                      const uint8_t *consider, i64 *elements, i64 *candidates, i64 *passing) {
    i64 el = 0, cand = 0, pass = 0;
    #pragma omp parallel for schedule(dynamic, 64) reduction(+:el,cand,pass)
    for (i64 i = 0; i < n; ++i) {
        i64 nna = row_ptr[i + 1] - row_ptr[i];
        if (!(consider[i] && nna > m_core)) continue;
        const i32 *na = col + row_ptr[i];
        for (i64 mi = 0; mi < nna; ++mi) {
            i64 j = na[mi];
            if (j <= i) continue;
            i64 nnb = row_ptr[j + 1] - row_ptr[j];
            if (!(consider[j] && nnb > m_core)) continue;
            i64 la, lb;
            pass += screen_sorted(na, nna, col + row_ptr[j], nnb, c_sim, &la, &lb);
            el += la + lb;
            ++cand;
        }
    }
    *elements = el; *candidates = cand; *passing = pass;
    return 0;
}

/* ------------------------------------------------------------ sort_by_size */

typedef struct { i64 label, size, first; } cl_t;

static int cmp_label(const void *a, const void *b) {
    const i64 *x = (const i64 *)a, *y = (const i64 *)b;
    if (x[0] != y[0]) return (x[0] > y[0]) - (x[0] < y[0]);
    return (x[1] > y[1]) - (x[1] < y[1]);
}

static int cmp_rank(const void *a, const void *b) {
    const cl_t *x = (const cl_t *)a, *y = (const cl_t *)b;
    if (x->size != y->size) return (x->size < y->size) - (x->size > y->size);
    return (x->first > y->first) - (x->first < y->first);
}

/* Labels.sort_by_size (_types.pyx:212-286): clusters ranked by size descending,
 * ties by first occurrence; ranks beyond max_clusters (< 0: no limit) and
 * clusters smaller than member_cutoff become 0; survivors renumbered 1..K. */
int oracle_sort_by_size(i64 *labels, i64 n, i64 member_cutoff, i64 max_clusters, i64 *n_clusters) {
    i64 *pairs = (i64 *)malloc(sizeof(i64) * 2 * (n > 0 ? n : 1));
    if (!pairs) return -2;
    i64 m = 0;
    for (i64 i = 0; i < n; ++i)
        if (labels[i] != 0) { pairs[2 * m] = labels[i]; pairs[2 * m + 1] = i; ++m; }
    qsort(pairs, m, 2 * sizeof(i64), cmp_label);
    cl_t *cl = (cl_t *)malloc(sizeof(cl_t) * (m > 0 ? m : 1));
    i64 k = 0;
    for (i64 p = 0; p < m; ) {
        i64 q = p;
        while (q < m && pairs[2 * q] == pairs[2 * p]) ++q;
        cl[k].label = pairs[2 * p];
        cl[k].size = q - p;
        cl[k].first = pairs[2 * p + 1];
        ++k;
        p = q;
    }
    qsort(cl, k, sizeof(cl_t), cmp_rank);
    if (max_clusters < 0) max_clusters = k;
    /* new label per old label, looked up by binary search over labels sorted */
    i64 *olds = (i64 *)malloc(sizeof(i64) * 2 * (k > 0 ? k : 1));
    i64 next = 1;
    for (i64 c = 0; c < k; ++c) {
        olds[2 * c] = cl[c].label;
        olds[2 * c + 1] = (c + 1 <= max_clusters && cl[c].size >= member_cutoff) ? next++ : 0;
    }
    qsort(olds, k, 2 * sizeof(i64), cmp_label);
    for (i64 i = 0; i < n; ++i) {
        if (labels[i] == 0) continue;
        i64 lo = 0, hi = k;
        while (lo < hi) {
            i64 mid = (lo + hi) >> 1;
            if (olds[2 * mid] < labels[i]) lo = mid + 1; else hi = mid;
        }
        labels[i] = olds[2 * lo + 1];
    }
    if (n_clusters) *n_clusters = next - 1;
    free(olds);
    free(cl);
    free(pairs);
    return 0;
}

int oracle_num_threads(void) { return omp_get_max_threads(); }

void oracle_set_num_threads(int k) { if (k > 0) omp_set_num_threads(k); }
