/* CPU oracle (test infrastructure only): plain-C restatement of the neighbour sets and the
 * 3DRCM permutation.  It is an independent second statement of oracle/geometry.py and
 * oracle/rcm.py, used to cross-check them and the library's own host RCM.
 *
 * Neighbours (InchingLiteInteger/Fuel/Coordinate/T1.py:219,239,276-277; Fuel/T1.py:291-299):
 *   j is a neighbour of i iff d2 <= rc*rc with d = x_j - x_i and
 *   d2 = (dx*dx + dy*dy) + dz*dz in float64, no contraction (build with -ffp-contract=off);
 *   self included; columns ascending.  Candidates come from a uniform grid of cell size rc.
 *
 * 3DRCM (InchingLiteInteger/Fuel/Coordinate/T1.py:171-323, Tex/main.tex:63-146):
 *   degree = row length (self included); seeds in stable ascending-degree order (or a given
 *   seed order); BFS where a dequeued node's neighbours are scanned in descending index
 *   order, its new children stably sorted by degree; the final order reversed.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef struct { int64_t key; int64_t idx; } kv_t;

static int cmp_kv(const void *a, const void *b)
{
    const kv_t *x = (const kv_t *)a, *y = (const kv_t *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static int cmp_i32(const void *a, const void *b)
{
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Pass 1 (indices == NULL): counts per row into indptr[1..n] and prefix sums; returns total.
 * Pass 2: fills indices (ascending per row).  Returns -1 on allocation failure. */
int64_t oracle_neighbours(int64_t n, const double *xyz, double rc, int64_t *indptr, int32_t *indices)
{
    if (n <= 0) { indptr[0] = 0; return 0; }
    double lo[3] = {xyz[0], xyz[1], xyz[2]}, hi[3] = {xyz[0], xyz[1], xyz[2]};
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            double v = xyz[3 * i + a];
            if (v < lo[a]) lo[a] = v;
            if (v > hi[a]) hi[a] = v;
        }
    int64_t dim[3];
    for (int a = 0; a < 3; ++a) {
        dim[a] = (int64_t)floor((hi[a] - lo[a]) / rc) + 1;
        if (dim[a] < 1) dim[a] = 1;
    }
    int64_t ncell = dim[0] * dim[1] * dim[2];
    int64_t *cstart = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *sorted = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *fillp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncell + 1));
    if (!cstart || !cell || !sorted || !fillp) { free(cstart); free(cell); free(sorted); free(fillp); return -1; }
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3];
        for (int a = 0; a < 3; ++a) {
            c[a] = (int64_t)floor((xyz[3 * i + a] - lo[a]) / rc);
            if (c[a] >= dim[a]) c[a] = dim[a] - 1;
            if (c[a] < 0) c[a] = 0;
        }
        cell[i] = (c[0] * dim[1] + c[1]) * dim[2] + c[2];
        cstart[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    memcpy(fillp, cstart, sizeof(int64_t) * (size_t)(ncell + 1));
    for (int64_t i = 0; i < n; ++i) sorted[fillp[cell[i]]++] = i;
    const double rc2 = rc * rc;
    int64_t total = 0;
    if (!indices) indptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t c0 = cell[i] / (dim[1] * dim[2]), c1 = (cell[i] / dim[2]) % dim[1], c2 = cell[i] % dim[2];
        int64_t cnt = 0;
        int32_t *row = indices ? indices + indptr[i] : NULL;
        for (int64_t a = c0 - 1; a <= c0 + 1; ++a) {
            if (a < 0 || a >= dim[0]) continue;
            for (int64_t b = c1 - 1; b <= c1 + 1; ++b) {
                if (b < 0 || b >= dim[1]) continue;
                for (int64_t e = c2 - 1; e <= c2 + 1; ++e) {
                    if (e < 0 || e >= dim[2]) continue;
                    int64_t cc = (a * dim[1] + b) * dim[2] + e;
                    for (int64_t p = cstart[cc]; p < cstart[cc + 1]; ++p) {
                        int64_t j = sorted[p];
                        double dx = xyz[3 * j] - xyz[3 * i];
                        double dy = xyz[3 * j + 1] - xyz[3 * i + 1];
                        double dz = xyz[3 * j + 2] - xyz[3 * i + 2];
                        double d2 = (dx * dx + dy * dy) + dz * dz;
                        if (d2 <= rc2) {
                            if (row) row[cnt] = (int32_t)j;
                            cnt++;
                        }
                    }
                }
            }
        }
        if (row) qsort(row, (size_t)cnt, sizeof(int32_t), cmp_i32);
        else indptr[i + 1] = indptr[i] + cnt;
        total += cnt;
    }
    free(cstart); free(cell); free(sorted); free(fillp);
    return total;
}

/* order (int32, n) and undo (int64, n).  seed_order may be NULL (stable argsort of degree).
 * Returns 0, or -1 on allocation failure. */
int oracle_rcm(int64_t n, const int64_t *indptr, const int32_t *indices, const int64_t *seed_order,
               int reverse, int32_t *order_out, int64_t *undo_out)
{
    if (n <= 0) return 0;
    int64_t *inds = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    char *visited = (char *)calloc((size_t)n, 1);
    kv_t *tmp = (kv_t *)malloc(sizeof(kv_t) * (size_t)n);
    if (!inds || !order || !visited || !tmp) { free(inds); free(order); free(visited); free(tmp); return -1; }
    if (seed_order) {
        memcpy(inds, seed_order, sizeof(int64_t) * (size_t)n);
    } else {
        for (int64_t i = 0; i < n; ++i) { tmp[i].key = indptr[i + 1] - indptr[i]; tmp[i].idx = i; }
        qsort(tmp, (size_t)n, sizeof(kv_t), cmp_kv);     /* (degree, index): a stable argsort */
        for (int64_t i = 0; i < n; ++i) inds[i] = tmp[i].idx;
    }
    int64_t N = 0;
    for (int64_t z = 0; z < n && N < n; ++z) {
        int64_t seed = inds[z];
        if (visited[seed]) continue;
        visited[seed] = 1;
        order[N++] = seed;
        int64_t head = N - 1;
        while (head < N) {                     /* queue order == level order */
            int64_t i = order[head++];
            int64_t n_old = N;
            for (int64_t p = indptr[i + 1] - 1; p >= indptr[i]; --p) {
                int64_t j = indices[p];
                if (!visited[j]) { visited[j] = 1; order[N++] = j; }
            }
            int64_t m = N - n_old;
            if (m > 1) {
                /* stable by degree: key degree, tie by position of appearance */
                for (int64_t t = 0; t < m; ++t) {
                    int64_t v = order[n_old + t];
                    tmp[t].key = indptr[v + 1] - indptr[v];
                    tmp[t].idx = t;
                }
                qsort(tmp, (size_t)m, sizeof(kv_t), cmp_kv);
                int64_t *kids = (int64_t *)malloc(sizeof(int64_t) * (size_t)m);
                if (!kids) { free(inds); free(order); free(visited); free(tmp); return -1; }
                for (int64_t t = 0; t < m; ++t) kids[t] = order[n_old + tmp[t].idx];
                memcpy(order + n_old, kids, sizeof(int64_t) * (size_t)m);
                free(kids);
            }
        }
    }
    for (int64_t p = 0; p < n; ++p) {
        int64_t v = reverse ? order[n - 1 - p] : order[p];
        order_out[p] = (int32_t)v;
        undo_out[v] = p;
    }
    free(inds); free(order); free(visited); free(tmp);
    return 0;
}
