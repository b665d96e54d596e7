/*
 * poy_oracle.c -- CPU restatement of POY's direct-optimization hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see poy_oracle.h).  Single-threaded, built with
 * the reference's default flags (-O2, no -march), so that it also serves as
 * the timed CPU baseline.  Every function restates the reference function
 * named in its comment; the reference is GPL OCaml+C that needs the OCaml
 * runtime headers, so it is not compiled here.
 *
 * Direction codes (src/matrices.h:22-42): ALIGN=1, INSERT=2, DELETE=4 for
 * the linear DP; P1..SS = 1..64 for the cube.  Affine flags:
 * src/algn.c:1695-1708.
 */
#include "poy_oracle.h"

#include <stdlib.h>
#include <string.h>

#define D_ALIGN 1
#define D_INSERT 2
#define D_DELETE 4

#define C_P1 (1 << 0)
#define C_P2 (1 << 1)
#define C_P3 (1 << 2)
#define C_S1 (1 << 3)
#define C_S2 (1 << 4)
#define C_S3 (1 << 5)
#define C_SS (1 << 6)

#define A_ALIGN_TO_ALIGN 1
#define A_ALIGN_TO_VERTICAL 2
#define A_ALIGN_TO_HORIZONTAL 4
#define A_ALIGN_TO_DIAGONAL 8
#define A_BEGIN_BLOCK 16
#define A_END_BLOCK 32
#define A_BEGIN_VERTICAL 64
#define A_END_VERTICAL 128
#define A_BEGIN_HORIZONTAL 256
#define A_END_HORIZONTAL 512
#define A_DO_ALIGN 1024
#define A_DO_VERTICAL 2048
#define A_DO_HORIZONTAL 4096
#define A_DO_DIAGONAL 8192

#define TMPGAP 16
#define NTMPGAP 15

/* ------------------------------------------------------------------------ */
/* Grow-only scratch, the analogue of the reference's Matrix.default        */
/* (src/matrix.ml:28, src/matrices.c:77-119): reused by every call.         */
/* ------------------------------------------------------------------------ */
static void *scratch_grow(void **buf, size_t *cap, size_t need) {
    if (*cap < need) {
        free(*buf);
        *buf = malloc(need);
        *cap = *buf ? need : 0;
    }
    return *buf;
}

static void *g_dm;     static size_t g_dm_cap;
static void *g_rows;   static size_t g_rows_cap;
static void *g_prec;   static size_t g_prec_cap;
static void *g_cube;   static size_t g_cube_cap;
static void *g_cubed;  static size_t g_cubed_cap;

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int popcount(int v) { int n = 0; while (v) { n += v & 1; v >>= 1; } return n; }

/* ======================================================================== */
/* R1: cost matrices                                                        */
/* ======================================================================== */

static int bit_length(int v) { int i = 0; while (v) { i++; v >>= 1; } return i; }

/* cm_set_val, src/cm.c:318-361; ceil_log_2 is src/cm.c:36-43. */
po_cm2 *po_cm2_create(int a_sz, int combinations, int cost_model, int gap_open, int all_elements) {
    po_cm2 *c = (po_cm2 *)calloc(1, sizeof(po_cm2));
    if (!c) return NULL;
    if (combinations) {
        c->gap = 1 << (a_sz - 1);
        c->a_sz = (1 << a_sz) - 1;
        c->lcm = a_sz;
        c->combinations = 1;
    } else {
        c->gap = a_sz;
        c->a_sz = a_sz;
        c->lcm = bit_length(a_sz + 1) + 1;
        c->combinations = 0;
    }
    c->all_elements = all_elements;
    c->cost_model = cost_model;
    c->gap_open = gap_open;
    c->is_metric = 0;
    if (c->lcm > 5) { free(c); return NULL; }   /* tables sized for 5-bit masks */
    return c;
}

void po_cm2_free(po_cm2 *c) { free(c); }

static int cpos(const po_cm2 *c, int a, int b) { return (a << c->lcm) + b; }   /* src/cm.c:484-489 */
int po_cm2_cost(const po_cm2 *c, int a, int b) { return c->cost[cpos(c, a, b)]; }
int po_cm2_worst(const po_cm2 *c, int a, int b) { return c->worst[cpos(c, a, b)]; }
int po_cm2_median(const po_cm2 *c, int a, int b) { return c->median[cpos(c, a, b)]; }
static void set_cost(po_cm2 *c, int a, int b, int v) { c->cost[cpos(c, a, b)] = v; }
static void set_worst(po_cm2 *c, int a, int b, int v) { c->worst[cpos(c, a, b)] = v; }
static void set_median(po_cm2 *c, int a, int b, int v) { c->median[cpos(c, a, b)] = (uint8_t)v; }

/* Cost_matrix.Two_D.cleanup, src/cost_matrix.ml:255-265. */
static int cleanup(const po_cm2 *c, int item) {
    if (c->combinations && c->cost_model == 1 && item != c->gap && (item & c->gap)) return c->gap;
    return item;
}

/* fill_best_cost_and_median_for_all_combinations, src/cost_matrix.ml:323-354,
 * with test_combinations / test_all_combinations, :223-249.  The median is
 * the union of every single state v minimising cost(a,v)+cost(v,b) over the
 * member pairs (a,b); ties OR in, so the visiting order does not matter. */
static void fill_metric(po_cm2 *m, int a_sz) {
    int ncomb = (1 << a_sz) - 1;
    for (int i = 1; i <= ncomb; i++) {
        for (int j = 1; j <= ncomb; j++) {
            int best = 0, cost = PO_OCAML_MAX_INT, worst = 0;
            for (int ab = 0; ab < a_sz; ab++) {
                int a = 1 << ab;
                if (!(i & a)) continue;
                for (int bb = 0; bb < a_sz; bb++) {
                    int b = 1 << bb;
                    if (!(j & b)) continue;
                    for (int v = 0; v < a_sz; v++) {
                        int vb = 1 << v;
                        int goa = (m->cost_model == 1 && vb == m->gap && (a & m->gap) && (b & m->gap))
                                      ? m->gap_open : 0;
                        int tc = po_cm2_cost(m, a, vb) + po_cm2_cost(m, vb, b) + goa;
                        if (tc < cost) { cost = tc; best = vb; }
                        else if (tc == cost) best |= vb;
                    }
                    worst = imax(worst, po_cm2_cost(m, a, b));
                }
            }
            if (popcount(i) == 1 && popcount(j) == 1) {
                set_median(m, i, j, i | j);
            } else {
                set_cost(m, i, j, cost);
                set_median(m, i, j, cleanup(m, best));
            }
            set_worst(m, i, j, worst);
        }
    }
}

/* find_best_combination of the bitwise fill, src/cost_matrix.ml:270-286:
 * folds over (l2 x l1) first, then (l1 x l2). */
static void best_combination(const po_cm2 *m, int i, int j, int a_sz, int *best, int *med, int *worst) {
    int bst = PO_OCAML_MAX_INT, md = 0, w = 0;
    for (int pass = 0; pass < 2; pass++) {
        int l1 = pass == 0 ? j : i, l2 = pass == 0 ? i : j;
        for (int xb = 0; xb < a_sz; xb++) {
            int x = 1 << xb;
            if (!(l1 & x)) continue;
            for (int yb = 0; yb < a_sz; yb++) {
                int y = 1 << yb;
                if (!(l2 & y)) continue;
                int mb1 = po_cm2_cost(m, x, y);
                if (mb1 < bst) { bst = mb1; md = x | y; }
                else if (mb1 == bst) md = x | y | md;
                w = imax(w, mb1);
            }
        }
    }
    *best = bst; *med = md; *worst = w;
}

/* fill_best_cost_and_median_for_all_combinations_bitwise,
 * src/cost_matrix.ml:270-321. */
static void fill_bitwise(po_cm2 *m, int a_sz) {
    int ncomb = (1 << a_sz) - 1;
    for (int i = 1; i <= ncomb; i++) {
        for (int j = 1; j <= ncomb; j++) {
            int best, med, worst;
            if (popcount(i) == 1 && popcount(j) == 1) {
                set_median(m, i, j, i | j);
                set_worst(m, i, j, po_cm2_cost(m, i, j));
            } else if (i & j) {
                best_combination(m, i, j, a_sz, &best, &med, &worst);
                set_median(m, i, j, i & j);
                set_cost(m, i, j, 0);
                set_worst(m, i, j, worst);
            } else {
                best_combination(m, i, j, a_sz, &best, &med, &worst);
                set_median(m, i, j, med);
                set_cost(m, i, j, best);
                set_worst(m, i, j, worst);
            }
        }
    }
}

/* fill_default_prepend_tail, src/cost_matrix.ml:397-403. */
static void fill_prepend_tail(po_cm2 *m) {
    for (int i = 1; i <= m->a_sz; i++) {
        m->tail[i] = po_cm2_cost(m, i, m->gap);
        m->prepend[i] = po_cm2_cost(m, m->gap, i);
    }
}

/* input_is_metric, src/cost_matrix.ml:488-518. */
static int input_is_metric(const int *l, int w) {
    for (int i = 0; i < w; i++)
        for (int j = 0; j < w; j++) {
            if (l[i * w + j] < 0) return 0;
            if (l[i * w + j] != l[j * w + i]) return 0;
            for (int k = 0; k < w; k++)
                if (l[i * w + j] > l[i * w + k] + l[k * w + j]) return 0;
        }
    for (int i = 0; i < w; i++) if (l[i * w + i] != 0) return 0;
    return 1;
}

/* fill_cost_matrix, src/cost_matrix.ml:520-538, with
 * store_input_list_in_cost_matrix_all_combinations, :156-174.  Only the
 * combination alphabets (5-bit DNA masks) are supported: fill_medians of the
 * non-combination case draws from Random (:384-391). */
po_cm2 *po_cm2_of_list(const int *mat, int w, int use_comb, int all_elements) {
    if (!use_comb) return NULL;
    po_cm2 *m = po_cm2_create(w, 1, 0, 0, all_elements);
    if (!m) return NULL;
    for (int e1 = 1; e1 <= w; e1++)
        for (int e2 = 1; e2 <= w; e2++)
            set_cost(m, 1 << (e1 - 1), 1 << (e2 - 1), mat[(e1 - 1) * w + (e2 - 1)]);
    if (input_is_metric(mat, w)) {
        m->is_metric = 1;
        fill_metric(m, w);
    } else {
        fill_bitwise(m, w);
    }
    fill_prepend_tail(m);
    return m;
}

/* of_transformations_and_gaps true 5 sub indel 31, src/cost_matrix.ml:645-656. */
po_cm2 *po_cm2_of_sub_indel(int sub, int indel) {
    int mat[25];
    for (int x = 0; x < 5; x++)
        for (int pos = 0; pos < 5; pos++) {
            int v;
            if (x == pos) v = 0;
            else if (x == 4 || pos == 4) v = indel;
            else v = sub;
            mat[pos * 5 + x] = v;
        }
    return po_cm2_of_list(mat, 5, 1, 31);
}

/* set_affine, src/cost_matrix.ml:405-416. */
int po_cm2_gap_open(const po_cm2 *c) { return c->gap_open; }

void po_cm2_set_affine(po_cm2 *m, int model, int gap_open) {
    m->cost_model = model;
    m->gap_open = model == 1 ? gap_open : 0;
    if (m->combinations) fill_bitwise(m, m->lcm);
    fill_prepend_tail(m);
}

static int c3pos(const po_cm3 *c, int a, int b, int d) {   /* src/cm.c:499-503 */
    return (((a << c->lcm) + b) << c->lcm) + d;
}
int po_cm3_cost(const po_cm3 *c, int a, int b, int d) { return c->cost[c3pos(c, a, b, d)]; }
int po_cm3_median(const po_cm3 *c, int a, int b, int d) { return c->median[c3pos(c, a, b, d)]; }

/* Three_D.of_two_dim (src/cost_matrix.ml:853-861) through cm_CAML_clone_to_3d
 * (src/cm.c:1435-1446) and of_two_dim_comb (:803-851). */
po_cm3 *po_cm3_of_two_dim(const po_cm2 *m) {
    if (!m->combinations) return NULL;
    po_cm3 *n = (po_cm3 *)calloc(1, sizeof(po_cm3));
    if (!n) return NULL;
    n->gap = 1 << (m->lcm - 1);
    n->a_sz = (1 << m->lcm) - 1;
    n->lcm = m->lcm;
    n->combinations = 1;
    n->cost_model = m->cost_model;
    n->gap_open = m->gap_open;
    n->all_elements = m->all_elements;
    int alph = m->a_sz, gap = m->gap, lcm = m->lcm;
    for (int i = 1; i <= alph; i++)
        for (int j = 1; j <= alph; j++)
            for (int k = 1; k <= alph; k++) {
                int p = c3pos(n, i, j, k);
                n->cost[p] = PO_OCAML_MAX_INT;
                for (int l = 0; l < lcm; l++) {
                    int inter = 1 << l;
                    int shared = ((i & inter) != 0) + ((j & inter) != 0) + ((k & inter) != 0);
                    int cost = (m->is_metric || shared >= 2 || inter != gap)
                                   ? po_cm2_cost(m, inter, i) + po_cm2_cost(m, inter, j) +
                                         po_cm2_cost(m, inter, k)
                                   : PO_OCAML_MAX_INT;
                    if (cost < n->cost[p]) { n->cost[p] = cost; n->median[p] = (uint8_t)inter; }
                    else if (cost == n->cost[p]) n->median[p] |= (uint8_t)inter;
                }
                /* pick_bit 1: keep the lowest set bit */
                int v = n->median[p], b = 1;
                while (b < 256 && !(v & b)) b <<= 1;
                n->median[p] = (uint8_t)(b < 256 ? b : 0);
            }
    return n;
}

void po_cm3_free(po_cm3 *c) { free(c); }

/* ======================================================================== */
/* R3: band width                                                           */
/* ======================================================================== */

/* seq_CAML_count, src/seq.c:571-581: positions that carry the gap bit. */
int po_count_gaps(const uint8_t *s, int len, int gap) {
    int n = 0;
    for (int i = 0; i < len; i++) if (s[i] & gap) n++;
    return n;
}

/* deltaw_calc of Sequence.Align.cost_2, src/sequence.ml:691-702. */
static int deltaw_calc(int s1len, int s2len, int has_deltaw, int v) {
    int dif = s1len - s2len;
    int lower_limit = (int)((double)s1len * 0.10);
    if (!has_deltaw) return dif < lower_limit ? lower_limit / 2 : 2;
    return dif < lower_limit ? lower_limit : v;
}

int po_deltaw(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const po_cm2 *c,
              int has_deltaw, int deltaw) {
    int gaps = imax(po_count_gaps(s1, l1, c->gap), po_count_gaps(s2, l2, c->gap));
    if (l1 >= l2) return gaps + deltaw_calc(l1, l2, has_deltaw, deltaw);
    return gaps + deltaw_calc(l2, l1, has_deltaw, deltaw);
}

/* ======================================================================== */
/* R4/R5: banded linear NW                                                  */
/* ======================================================================== */

/* Row layout of algn_fill_plane_2 (src/algn.c:872-968) and algn_fill_plane
 * (:812-855).  Every row i >= 1 fills columns [lo, hi]; the first cell is
 * either column 0 (DELETE only; its DELETE cost is tail[s1[i]] in the banded
 * phases, :610-620, and cost(s1[i],gap) in the full plane, :567-581) or a
 * left band cell (ALIGN/DELETE only, :507-541); the last cell is either a
 * right band cell (ALIGN/INSERT only, :461-500) or a regular cell followed by
 * the last-column DELETE re-check with tail (:548-560).  Row 0 fills
 * [0, hi0] with prepend costs (:584-607). */
enum { LK_FIRST_TAIL = 0, LK_FIRST_COST = 1, LK_LEFT = 2 };
enum { RK_RIGHT = 0, RK_LAST = 1 };

typedef struct { int lo, hi, lk, rk; } row_desc;

/* Returns 1 for a full plane, 0 for a band; fills desc[0..l1-1]. */
static int linear_rows(int l1, int l2, int deltawh, row_desc *desc) {
    int width = 50 + deltawh, height = (l1 - l2) + 50 + deltawh;
    if (width > l2) width = l2;
    if (height > l1) height = l1;
    int full = 0;
    if ((float)l1 >= (float)((float)((float)3 / (float)2) * (float)l2)) full = 1;
    else if (!(2 * height < l1) && 8 >= (l1 - height)) full = 1;
    if (full) {
        desc[0].lo = 0; desc[0].hi = l2 - 1; desc[0].lk = LK_FIRST_COST; desc[0].rk = RK_LAST;
        for (int i = 1; i < l1; i++) {
            desc[i].lo = 0; desc[i].hi = l2 - 1; desc[i].lk = LK_FIRST_COST; desc[i].rk = RK_LAST;
        }
        return 1;
    }
    desc[0].lo = 0; desc[0].hi = width - 1; desc[0].lk = LK_FIRST_TAIL; desc[0].rk = RK_RIGHT;
    if (2 * height < l1) {
        /* Case 2: extending right, left-right band, extending left. */
        int i = 1;
        for (; i < height; i++) { desc[i].lo = 0; desc[i].hi = width + i - 1; desc[i].lk = LK_FIRST_TAIL; desc[i].rk = RK_RIGHT; }
        for (; i < l1 - (height - 1); i++) {
            int sc = 1 + (i - height);
            desc[i].lo = sc; desc[i].hi = sc + width + height - 2; desc[i].lk = LK_LEFT; desc[i].rk = RK_RIGHT;
        }
        int sc0 = l2 - (width + height - 2), r0 = i;
        for (; i < l1; i++) { desc[i].lo = sc0 + (i - r0); desc[i].hi = l2 - 1; desc[i].lk = LK_LEFT; desc[i].rk = RK_LAST; }
    } else {
        /* Case 3: extending right, full rows, extending left. */
        int i = 1;
        for (; i < (l2 - width) + 1; i++) { desc[i].lo = 0; desc[i].hi = width + i - 1; desc[i].lk = LK_FIRST_TAIL; desc[i].rk = RK_RIGHT; }
        for (; i < l1 - (l2 - width) + 1; i++) { desc[i].lo = 0; desc[i].hi = l2 - 1; desc[i].lk = LK_FIRST_TAIL; desc[i].rk = RK_LAST; }
        int r0 = i;
        for (; i < l1; i++) { desc[i].lo = 1 + (i - r0); desc[i].hi = l2 - 1; desc[i].lk = LK_LEFT; desc[i].rk = RK_LAST; }
    }
    return 0;
}

int64_t po_algn_nw_cells(int l1, int l2, int deltawh) {
    row_desc *d = (row_desc *)malloc(sizeof(row_desc) * (size_t)(l1 > 0 ? l1 : 1));
    if (!d) return -1;
    linear_rows(l1, l2, deltawh, d);
    int64_t n = 0;
    for (int i = 0; i < l1; i++) n += d[i].hi - d[i].lo + 1;
    free(d);
    return n;
}

/* algn_nw (src/algn.c:3276-3283) -> algn_nw_limit (:3246-3268) ->
 * cm_precalc_4algn (src/cm.c:691-723) + algn_fill_plane_2. */
int po_algn_nw(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const po_cm2 *c,
               int deltawh, uint16_t *dm, int *band_lo, int *band_hi) {
    if (l1 < l2 || l2 < 1) return -1;
    row_desc *desc = (row_desc *)malloc(sizeof(row_desc) * (size_t)l1);
    /* prec rows: 0 = prepend, 1..a_sz = cost rows (column 0 = tail). */
    size_t nprec = (size_t)(1 << c->lcm) * (size_t)l2;
    int *prec = (int *)scratch_grow(&g_prec, &g_prec_cap, nprec * sizeof(int));
    int *rows = (int *)scratch_grow(&g_rows, &g_rows_cap, (size_t)2 * l2 * sizeof(int));
    if (!desc || !prec || !rows) { free(desc); return -1; }
    for (int k = 0; k < l2; k++) prec[k] = c->prepend[s2[k]];
    for (int r = 1; r <= c->a_sz; r++) {
        int *row = prec + (size_t)r * l2;
        row[0] = c->tail[r];
        for (int k = 1; k < l2; k++) row[k] = po_cm2_cost(c, r, s2[k]);
    }
    linear_rows(l1, l2, deltawh, desc);
    int *pm = rows, *mm = rows + l2;
    const int *gap_row = prec + (size_t)c->gap * l2;
    /* row 0 */
    pm[0] = 0;
    if (dm) dm[0] = D_ALIGN;
    for (int k = 1; k <= desc[0].hi; k++) {
        pm[k] = pm[k - 1] + prec[k];
        if (dm) dm[k] = D_INSERT;
    }
    if (band_lo) { band_lo[0] = desc[0].lo; band_hi[0] = desc[0].hi; }
    for (int i = 1; i < l1; i++) {
        const row_desc *d = desc + i;
        int si = s1[i];
        int c_del = po_cm2_cost(c, si, c->gap);
        int tail = c->tail[si];
        const int *alg = prec + (size_t)si * l2;
        uint16_t *drow = dm ? dm + (size_t)i * l2 : NULL;
        int j = d->lo;
        if (d->lk != LK_LEFT) {
            mm[0] = pm[0] + (d->lk == LK_FIRST_TAIL ? alg[0] : c_del);
            if (drow) drow[0] = D_DELETE;
            j = 1;
        }
        for (; j <= d->hi; j++) {
            int has_del = !(j == d->hi && d->rk == RK_RIGHT);
            int has_ins = !(j == d->lo && d->lk == LK_LEFT);
            int t3 = pm[j - 1] + alg[j];
            int best = t3, mask = D_ALIGN;
            if (has_del) {
                int t1 = pm[j] + c_del;
                if (t1 < best) { best = t1; mask = D_DELETE; } else if (t1 == best) mask |= D_DELETE;
            }
            if (has_ins) {
                int t2 = mm[j - 1] + gap_row[j];
                if (t2 < best) { best = t2; mask = D_INSERT; } else if (t2 == best) mask |= D_INSERT;
            }
            if (j == d->hi && d->rk == RK_LAST && j > 0) {
                int cst = tail + pm[j];
                if (cst < best) { best = cst; mask = D_DELETE; } else if (cst == best) mask |= D_DELETE;
            }
            mm[j] = best;
            if (drow) drow[j] = (uint16_t)mask;
        }
        if (band_lo) { band_lo[i] = d->lo; band_hi[i] = d->hi; }
        int *t = pm; pm = mm; mm = t;
    }
    int res = pm[l2 - 1];
    free(desc);
    return res;
}

/* ======================================================================== */
/* R6: linear backtrace                                                     */
/* ======================================================================== */

/* backtrack_2d, linear branch (src/algn.c:3551-3666).  ALIGN wins first;
 * then INSERT before DELETE when swaped, else DELETE before INSERT.  The
 * walk consumes cell (0,0), so the leading gap column is emitted. */
int po_backtrack_2d(const uint16_t *dm, const int *band_lo, const int *band_hi,
                    const uint8_t *s1, int l1, const uint8_t *s2, int l2,
                    const po_cm2 *c, int swaped, uint8_t *r1, uint8_t *r2, int cap) {
    int i = l1 - 1, j = l2 - 1, n = 0;
    while (i >= 0 && j >= 0) {
        if (band_lo && (j < band_lo[i] || j > band_hi[i])) return -1;
        int d = dm[(size_t)i * l2 + j];
        int mv;
        if (d & D_ALIGN) mv = D_ALIGN;
        else if (swaped) mv = (d & D_INSERT) ? D_INSERT : D_DELETE;
        else mv = (d & D_DELETE) ? D_DELETE : D_INSERT;
        if (n >= cap) return -1;
        n++;
        if (mv == D_ALIGN) { r1[cap - n] = s1[i]; r2[cap - n] = s2[j]; i--; j--; }
        else if (mv == D_INSERT) { r1[cap - n] = (uint8_t)c->gap; r2[cap - n] = s2[j]; j--; }
        else { r1[cap - n] = s1[i]; r2[cap - n] = (uint8_t)c->gap; i--; }
        /* the reference walks a flat pointer; a move left of column 0 wraps
         * into the previous row, which the band geometry never produces */
        if (j < 0 && i >= 0) return -1;
    }
    return n;
}

/* ======================================================================== */
/* Sequence.Align wrappers                                                  */
/* ======================================================================== */

/* Sequence.Align.cost_2 (src/sequence.ml:691-723): the longer sequence goes
 * first; affine matrices dispatch to cost_2_affine. */
int po_cost_2(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const po_cm2 *c,
              int has_deltaw, int deltaw) {
    if (c->cost_model == 1) return po_cost_affine_3(s1, l1, s2, l2, c);
    int dw = po_deltaw(s1, l1, s2, l2, c, has_deltaw, deltaw);
    if (l1 >= l2) return po_algn_nw(s1, l1, s2, l2, c, dw, NULL, NULL, NULL);
    return po_algn_nw(s2, l2, s1, l1, c, dw, NULL, NULL, NULL);
}

/* Sequence.Align.align_2 ~first_gap:true (src/sequence.ml:849-869) with
 * create_edited_2 (:813-823). */
int po_align_2(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const po_cm2 *c,
               uint8_t *o1, uint8_t *o2, int cap, int *olen) {
    if (c->cost_model == 1) {
        int lens[4];
        uint8_t *med = (uint8_t *)malloc((size_t)cap + 2), *mwg = (uint8_t *)malloc((size_t)cap + 2);
        int cost = po_align_affine_3(s1, l1, s2, l2, c, med, mwg, o1, o2, cap, lens);
        free(med); free(mwg);
        *olen = lens[2];
        return cost;
    }
    const uint8_t *a = s1, *b = s2;
    int la = l1, lb = l2, swaped = l1 >= l2;
    uint8_t *oa = o1, *ob = o2;
    if (!swaped) { a = s2; b = s1; la = l2; lb = l1; oa = o2; ob = o1; }
    int dw = po_deltaw(s1, l1, s2, l2, c, 0, 0);
    uint16_t *dm = (uint16_t *)scratch_grow(&g_dm, &g_dm_cap, (size_t)la * lb * sizeof(uint16_t));
    int *lo = (int *)malloc(sizeof(int) * (size_t)la), *hi = (int *)malloc(sizeof(int) * (size_t)la);
    int cost = po_algn_nw(a, la, b, lb, c, dw, dm, lo, hi);
    int n = po_backtrack_2d(dm, lo, hi, a, la, b, lb, c, swaped, oa, ob, cap);
    free(lo); free(hi);
    *olen = n;
    return n < 0 ? -1 : cost;
}

/* ======================================================================== */
/* R7: medians                                                              */
/* ======================================================================== */

/* algn_get_median_2d_no_gaps, src/algn.c:4042-4055. */
int po_median_2_no_gaps(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c,
                        uint8_t *out, int cap) {
    int n = 0;
    for (int i = len - 1; i >= 0; i--) {
        int m = po_cm2_median(c, a[i], b[i]);
        if (m != c->gap) out[cap - ++n] = (uint8_t)m;
    }
    out[cap - ++n] = (uint8_t)c->gap;
    return n;
}

/* algn_get_median_2d_with_gaps, src/algn.c:4024-4035. */
int po_median_2_with_gaps(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c,
                          uint8_t *out, int cap) {
    for (int i = len - 1, n = 1; i >= 0; i--, n++) out[cap - n] = (uint8_t)po_cm2_median(c, a[i], b[i]);
    return len;
}

/* algn_remove_gaps (src/algn.c:4058-4077) on a right-aligned buffer. */
static int remove_gaps(int gap, uint8_t *out, int cap, int len) {
    int src = cap - 1, dst = cap - 1, n = 0;
    for (int i = len - 1; i >= 0; i--, src--)
        if (out[src] != gap) { out[dst--] = out[src]; n++; }
    out[cap - ++n] = (uint8_t)gap;
    return n;
}

/* algn_correct_blocks_affine, src/algn.c:4080-4119; s has `len` symbols at
 * out[cap-len..], aligned with a and b. */
static int correct_blocks_affine(int gap, uint8_t *out, int cap, int len, const uint8_t *a, const uint8_t *b) {
    int extending_gap = 0, inside_block = 0, prev_block = 0;
    uint8_t *s = out + cap - len;
    for (int i = 0; i < len; i++) {
        int ab = a[i], bb = b[i], sb = s[i];
        if (!inside_block && (!(ab & gap) || !(bb & gap))) inside_block = 0;
        else if (inside_block && ((!(ab & gap)) || (!(bb & gap)))) inside_block = 0;
        else if (((ab & gap) || (bb & gap)) && ((ab != gap) || (bb != gap))) inside_block = 1;
        else inside_block = 0;
        if (((gap & ab) || (gap & bb)) && (!(sb & gap)) && !extending_gap) {
            prev_block = inside_block;
            extending_gap = 1;
        } else if ((gap & ab) && (gap & bb) && (sb & gap) && (sb != gap) && extending_gap && inside_block &&
                   !prev_block) {
            sb = (~gap) & sb;
            prev_block = 0;
        } else if ((gap & ab) && (gap & bb) && (1 == extending_gap)) {
            prev_block = inside_block;
            extending_gap = 0;
        }
        s[i] = (uint8_t)sb;
    }
    return remove_gaps(gap, out, cap, len);
}

/* algn_ancestor_2, src/algn.c:4126-4147. */
int po_ancestor_2(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c, uint8_t *out, int cap) {
    int n = 0, gap = c->gap;
    int keep_all = c->combinations && c->cost_model == 1;
    for (int i = len - 1; i >= 0; i--) {
        int m = po_cm2_median(c, a[i], b[i]);
        if (!keep_all) { if (m != gap) out[cap - ++n] = (uint8_t)m; }
        else out[cap - ++n] = (uint8_t)m;
    }
    if (!c->combinations || (c->cost_model != 1 && (n == 0 || gap != out[cap - n]))) {
        out[cap - ++n] = (uint8_t)gap;
    } else if (c->combinations && c->cost_model == 1) {
        n = correct_blocks_affine(gap, out, cap, n, a, b);
    }
    return n;
}

/* algn_union, src/algn.c:4176-4184. */
int po_union(const uint8_t *a, const uint8_t *b, int len, uint8_t *out, int cap) {
    for (int i = len - 1, n = 1; i >= 0; i--, n++) out[cap - n] = (uint8_t)(a[i] | b[i]);
    return len;
}

/* Sequence unions: union_merge, src/union.c:102-164, with the helpers
 * union_prepend_pair :29-38, union_prepend_item :40-46, union_copy_single
 * :48-59, union_copy_non_homologous :61-69, union_prepend_and_fill_items
 * :71-91 and union_move_left :93-100.  The unions are the structure of
 * src/union.h (a sequence plus the arrays offset / union_c1 / union_c2 of
 * Sequence.Unions.u, src/sequence.ml:1833-1837), restated here with
 * pointers that walk right-aligned buffers exactly as the reference's do:
 * `begin` is the first item, `end` the current entry of the offset array,
 * `position` the current item.  The children are given left-aligned (n
 * items); the result is written right-aligned into capacity-`cap` buffers.
 * Returns its length, or -1 when a child runs out of items or a buffer is
 * too small (the reference would read or write out of bounds there). */
typedef struct {
    const uint8_t *s;          /* items */
    const int32_t *off;        /* offset array, same indexing as s */
    int position;              /* current item */
    int end;                   /* current entry of off */
    int begin;
} po_uin;

typedef struct {
    uint8_t *s;
    int32_t *off, *ca, *cb;
    int sbegin;                /* first item of s (sequences grow leftwards) */
    int end, ca_at, cb_at;     /* next entry of each offset array */
    int cap;
    int fail;
} po_uout;

static void uo_prepend(po_uout *c, int v) {
    if (c->sbegin == 0) { c->fail = 1; return; }
    c->s[--c->sbegin] = (uint8_t)v;
}

static int ui_get(po_uin *u, po_uout *c) {
    if (u->position < 0) { c->fail = 1; return 0; }
    return u->s[u->position];
}

static void u_prepend_pair(po_uin *a, po_uin *b, po_uout *c, int or_) {
    uo_prepend(c, or_ | ui_get(a, c) | ui_get(b, c));
    a->position--; b->position--; a->end--; b->end--;
}

static void u_prepend_item(po_uin *src, po_uout *c, int or_) {
    uo_prepend(c, or_ | ui_get(src, c));
    src->position--; src->end--;
}

static int u_copy_single(po_uin *a, po_uout *c, int it, int or_) {
    int i = a->off[a->end];
    it += i;
    while (i > 0 && !c->fail) { u_prepend_item(a, c, or_); i--; }
    return it;
}

static int u_copy_non_homologous(po_uin *au, po_uin *bu, po_uout *c, int items, int gap) {
    if (au->position != -1) items = u_copy_single(au, c, items, gap);
    if (bu->position != -1 && !c->fail) items = u_copy_single(bu, c, items, gap);
    return items;
}

static void u_slot(po_uout *c, int *at) {
    if (*at < 0) c->fail = 1;
}

static void u_prepend_and_fill(int interm, po_uout *c, int prep, int apos, int bpos) {
    for (int i = prep; i > 0 && !c->fail; i--) {
        u_slot(c, &c->end); u_slot(c, &c->ca_at); u_slot(c, &c->cb_at);
        if (c->fail) return;
        c->off[c->end--] = i;
        c->ca[c->ca_at--] = 0;
        c->cb[c->cb_at--] = 0;
    }
    u_slot(c, &c->end); u_slot(c, &c->ca_at); u_slot(c, &c->cb_at);
    if (c->fail) return;
    c->ca[c->ca_at--] = apos;
    c->cb[c->cb_at--] = bpos;
    uo_prepend(c, interm);
    c->off[c->end--] = 0;
}

static int u_move_left(po_uin *a, po_uout *c) {
    int r = ui_get(a, c);
    a->position--; a->end--;
    return r;
}

int po_union_make(const uint8_t *a, const uint8_t *b, const uint8_t *median, int lena, int gap,
                  const uint8_t *ua, const int32_t *ua_off, int na,
                  const uint8_t *ub, const int32_t *ub_off, int nb,
                  uint8_t *out, int32_t *out_off, int32_t *out_ca, int32_t *out_cb, int cap) {
    po_uin au = {ua, ua_off, na - 1, na - 1, 0}, bu = {ub, ub_off, nb - 1, nb - 1, 0};
    po_uout c = {out, out_off, out_ca, out_cb, cap, cap - 1, cap - 1, cap - 1, cap, 0};
    int items_prepended = 0, i = lena - 1, interm, apos, bpos;
    long guard = 0;
    while ((au.begin <= au.end) || (bu.begin <= bu.end)) {
        if (++guard > 4L * (na + nb + lena + 4)) return -1;   /* the reference would not end */
        items_prepended = u_copy_non_homologous(&au, &bu, &c, items_prepended, gap);
        if (c.fail) return -1;
        if (i >= 0) {
            interm = median[i];
            if (gap == interm) {
                if (i != 0) {
                    if (gap == a[i]) u_prepend_item(&bu, &c, gap);
                    else if (gap == b[i]) u_prepend_item(&au, &c, gap);
                    else u_prepend_pair(&au, &bu, &c, gap);
                } else {
                    u_prepend_and_fill(gap, &c, items_prepended, 0, 0);
                    if (c.fail || c.end + 1 != c.sbegin) return -1;
                    return cap - c.sbegin;
                }
                items_prepended += 1;
            } else {
                if (gap == a[i]) {
                    apos = 0;
                    bpos = bu.position;
                    interm = interm | u_move_left(&bu, &c);
                } else if (gap == b[i]) {
                    bpos = 0;
                    apos = au.position;
                    interm = interm | u_move_left(&au, &c);
                } else {
                    bpos = bu.position;
                    apos = au.position;
                    interm = interm | u_move_left(&au, &c);
                    interm = interm | u_move_left(&bu, &c);
                }
                u_prepend_and_fill(interm, &c, items_prepended, apos, bpos);
                items_prepended = 0;
            }
            if (c.fail) return -1;
            i--;
        }
    }
    if (c.sbegin == cap || gap != c.s[c.sbegin]) {
        if (c.sbegin == 0 || c.end < 0) return -1;
        uo_prepend(&c, gap);
        c.ca[c.ca_at] = 0;
        c.cb[c.cb_at] = 0;
        c.off[c.end] = 0;
        c.end--;
    }
    if (c.end + 1 != c.sbegin) return -1;
    return cap - c.sbegin;
}

/* algn_calculate_from_2_aligned, src/algn.c:3316-3370. */
static int calc_from_2_aligned(const uint8_t *s1, const uint8_t *s2, int len, const po_cm2 *c, const int *tab) {
    int res = 0, gap_row = 0, gap = c->gap, go = c->gap_open, comb = c->combinations, i;
    if (len == 0) return 0;
    int s1b = s1[0], s2b = s2[0];
    i = ((comb && (gap & s1b) && (gap & s2b)) || (!comb && gap == s1b && gap == s2b)) ? 1 : 0;
    for (; i < len; i++) {
        s1b = s1[i]; s2b = s2[i];
        if (gap_row == 0) {
            if ((comb && (s1b & gap) && !(s2b & gap)) || (!comb && s1b == gap)) { res += go; gap_row = 1; }
            else if ((comb && (s2b & gap) && !(s1b & gap)) || (!comb && s2b == gap)) { res += go; gap_row = 2; }
        } else if (gap_row == 1) {
            if ((comb && !(s1b & gap)) || (!comb && s1b != gap)) {
                if ((comb && (s2b & gap) && !(s1b & gap)) || (!comb && s2b == gap)) { res += go; gap_row = 2; }
                else gap_row = 0;
            }
        } else {
            if ((comb && !(s2b & gap)) || (!comb && s2b != gap)) {
                if ((comb && (s1b & gap)) || (!comb && s1b == gap)) { res += go; gap_row = 1; }
                else gap_row = 0;
            }
        }
        res += tab[(s1[i] << c->lcm) + s2[i]];
    }
    return res;
}

int po_verify_2(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c) {
    return calc_from_2_aligned(a, b, len, c, c->cost);
}
int po_worst_2(const uint8_t *a, const uint8_t *b, int len, const po_cm2 *c) {
    return calc_from_2_aligned(a, b, len, c, c->worst);
}

/* ======================================================================== */
/* R8/R9: affine band                                                       */
/* ======================================================================== */

/* HAS_GAP_OPENING, src/algn.c:1729-1733. */
static int gap_opening(int prev, int curr, int gap, int go) {
    return ((!(gap & prev)) && (gap & curr)) ? 0 : go;
}

typedef struct {
    int *eh, *ev, *eb, *cb;   /* two rolling rows each, stride lb+1 */
    int *fc;                  /* one row (the reference keeps a single row) */
    int *gap_row, *gap_open_prec, *hx;
} aff_bufs;

static int aff_alloc(aff_bufs *b, int lb) {
    size_t w = (size_t)lb + 1;
    int *base = (int *)scratch_grow(&g_rows, &g_rows_cap, sizeof(int) * w * 12);
    if (!base) return -1;
    b->eh = base; b->ev = base + 2 * w; b->eb = base + 4 * w; b->cb = base + 6 * w;
    b->fc = base + 8 * w; b->gap_row = base + 9 * w; b->gap_open_prec = base + 10 * w; b->hx = base + 11 * w;
    return 0;
}

/* The four-state band fill.  `bt` selects algn_fill_plane_3_aff
 * (src/algn.c:2411-2548, with initialize_matrices_affine :2183-2248) versus
 * algn_fill_plane_3_aff_nobt (:2287-2403, init :2123-2181).  The two differ
 * in FILL_EXTEND_BLOCK_DIAGONAL: the backtrace variant (:1861-1882) adds
 * `diag` to both candidates, the cost-only variant (:1837-1857) adds
 * `open_diag` to the closing candidate.  A is the shorter sequence (rows). */
static int aff_fill(const uint8_t *A, int la, const uint8_t *B, int lb, const po_cm2 *c, int bt,
                    uint16_t *dm, aff_bufs *bf) {
    const int gap = c->gap, go = c->gap_open, W = lb + 1, H = PO_HIGH_NUM;
    int *EH[2] = {bf->eh, bf->eh + W}, *EV[2] = {bf->ev, bf->ev + W};
    int *EB[2] = {bf->eb, bf->eb + W}, *CB[2] = {bf->cb, bf->cb + W};
    int *FC = bf->fc, *gap_row = bf->gap_row, *gop = bf->gap_open_prec, *hx = bf->hx;
    for (int j = 0; j <= lb; j++) gap_row[j] = c->prepend[B[j]];   /* precalc row 0 */
    /* row 0 */
    EH[0][0] = go; EV[0][0] = go; CB[0][0] = 0; EB[0][0] = 0; FC[0] = 0;
    if (dm) dm[0] = 0xFFFF;
    for (int j = 1; j <= lb; j++) {
        int r = EH[0][j - 1] + gap_row[j];
        EH[0][j] = r; CB[0][j] = r; FC[j] = r; EB[0][j] = H; EV[0][j] = H;
        if (dm) dm[j] = A_DO_HORIZONTAL | A_END_HORIZONTAL;
    }
    if (la == 0) {
        if (bt) return FC[lb];
        int r = EH[0][lb];
        r = imin(r, EV[0][lb]); r = imin(r, EB[0][lb]); r = imin(r, CB[0][lb]);
        return r;
    }
    int end_pos = (lb - la) + 8;
    if (end_pos < 40) end_pos = 40;
    if (end_pos > lb) end_pos = lb;
    int start_pos = 1;
    for (int j = 1; j <= lb; j++) {
        gop[j] = gap_opening(B[j - 1], B[j], gap, go);
        hx[j] = ((B[j - 1] & gap) && !(B[j] & gap)) ? gop[j] + gap_row[j] : gap_row[j];
    }
    if (lb >= 1) hx[1] = gap_row[1];
    int ic = A[0];
    for (int i = 1; i <= la; i++) {
        int *peh = EH[(i - 1) & 1], *pev = EV[(i - 1) & 1], *peb = EB[(i - 1) & 1], *pcb = CB[(i - 1) & 1];
        int *eh = EH[i & 1], *ev = EV[i & 1], *eb = EB[i & 1], *cb = CB[i & 1];
        uint16_t *drow = dm ? dm + (size_t)i * W : NULL;
        if (i > 40) start_pos++;
        int ip = ic;
        ic = A[i];
        int si_ge = po_cm2_cost(c, ic, gap);
        int si_go = gap_opening(ip, ic, gap, go);
        int si_ng = ic & NTMPGAP;
        int sive = (i > 1 && (ip & gap) && !(ic & gap)) ? si_go + si_ge : si_ge;
        int r = pev[start_pos - 1] + sive;
        eh[start_pos - 1] = H; cb[start_pos - 1] = H; eb[start_pos - 1] = H; ev[start_pos - 1] = r;
        FC[start_pos - 1] = r;
        if (drow) drow[start_pos - 1] = A_DO_VERTICAL | A_END_VERTICAL;
        int jc = B[start_pos - 1];
        const int *crow = c->cost + (si_ng << c->lcm);
        for (int j = start_pos; j <= end_pos; j++) {
            int jp = jc;
            jc = B[j];
            int sj_ng = jc & NTMPGAP, sj_ge = gap_row[j], sj_go = gop[j];
            int mask = 0, ext, open;
            /* FILL_EXTEND_HORIZONTAL, :1761-1787 */
            ext = eh[j - 1] + hx[j];
            open = cb[j - 1] + sj_go + sj_ge;
            if (ext < open) { eh[j] = ext; mask |= A_BEGIN_HORIZONTAL; } else { eh[j] = open; mask |= A_END_HORIZONTAL; }
            /* FILL_EXTEND_VERTICAL, :1812-1830 */
            ext = pev[j] + sive;
            open = pcb[j] + si_go + si_ge;
            if (ext < open) { ev[j] = ext; mask |= A_BEGIN_VERTICAL; } else { ev[j] = open; mask |= A_END_VERTICAL; }
            /* FILL_EXTEND_BLOCK_DIAGONAL(_NOBT) */
            int both_gap = (TMPGAP & ic) && (TMPGAP & jc);
            int diag = both_gap ? 0 : H;
            if (bt) {
                ext = peb[j - 1] + diag;
                open = pcb[j - 1] + diag;
            } else {
                int flag2 = !(TMPGAP & ip) && !(TMPGAP & jc);
                int open_diag = both_gap ? (flag2 ? 0 : 2 * go) : H;
                ext = peb[j - 1] + diag;
                open = pcb[j - 1] + open_diag;
            }
            (void)jp;
            if (ext < open) { eb[j] = ext; mask |= A_BEGIN_BLOCK; } else { eb[j] = open; mask |= A_END_BLOCK; }
            /* FILL_CLOSE_BLOCK_DIAGONAL, :1929-1977 */
            int dg = crow[sj_ng];
            int xgo = (sj_go < si_go) ? si_go : sj_go;
            int algn = pcb[j - 1] + dg;
            int fv = (ic == si_ng) ? pev[j - 1] + dg : pev[j - 1] + dg + sj_go;
            int fh = (jc == sj_ng) ? peh[j - 1] + dg : peh[j - 1] + dg + si_go;
            int fd = peb[j - 1] + dg + xgo;
            int v = algn, am = A_ALIGN_TO_ALIGN;
            if (v >= fv) { if (v > fv) { v = fv; am = A_ALIGN_TO_VERTICAL; } else am |= A_ALIGN_TO_VERTICAL; }
            if (v >= fh) { if (v > fh) { v = fh; am = A_ALIGN_TO_HORIZONTAL; } else am |= A_ALIGN_TO_HORIZONTAL; }
            if (v >= fd) { if (v > fd) { v = fd; am = A_ALIGN_TO_DIAGONAL; } else am |= A_ALIGN_TO_DIAGONAL; }
            cb[j] = v;
            mask |= am;
            /* ASSIGN_MINIMUM, :2250-2280 */
            int f = eh[j], dmk = A_DO_HORIZONTAL;
            if (f >= ev[j]) { if (f > ev[j]) { f = ev[j]; dmk = A_DO_VERTICAL; } else dmk |= A_DO_VERTICAL; }
            if (f >= eb[j]) { if (f > eb[j]) { f = eb[j]; dmk = A_DO_DIAGONAL; } else dmk |= A_DO_DIAGONAL; }
            if (f >= cb[j]) { if (f > cb[j]) { f = cb[j]; dmk = A_DO_ALIGN; } else dmk |= A_DO_ALIGN; }
            FC[j] = f;
            if (drow) drow[j] = (uint16_t)(mask | dmk);
        }
        if (end_pos < lb) {
            end_pos++;
            if (drow) drow[end_pos] = A_DO_HORIZONTAL | A_END_HORIZONTAL;
            ev[end_pos] = H; cb[end_pos] = H; eh[end_pos] = H; eb[end_pos] = H;
        }
    }
    if (bt) return FC[lb];
    int k = la & 1;
    int r = EH[k][lb];
    if (r > EV[k][lb]) r = EV[k][lb];
    if (r > EB[k][lb]) r = EB[k][lb];
    if (r > CB[k][lb]) r = CB[k][lb];
    return r;
}

/* Filled column range of affine row i >= 1 including both boundary cells. */
static void aff_row_range(int la, int lb, int i, int *lo, int *hi) {
    int e0 = (lb - la) + 8;
    if (e0 < 40) e0 = 40;
    if (e0 > lb) e0 = lb;
    int sp = 1 + (i > 40 ? i - 40 : 0);
    int ep = e0 + (i - 1);
    if (ep > lb) ep = lb;
    *lo = sp - 1;
    *hi = ep < lb ? ep + 1 : ep;
}

int64_t po_affine_cells(int li, int lj) {
    int la = (li <= lj ? li : lj) - 1, lb = (li <= lj ? lj : li) - 1;
    int e0 = imin(lb, imax(40, (lb - la) + 8));
    int64_t n = 0;
    for (int i = 1; i <= la; i++) {
        int sp = 1 + (i > 40 ? i - 40 : 0);
        int ep = imin(lb, e0 + i - 1);
        n += ep - sp + 1;
    }
    return n;
}

/* backtrace_affine, src/algn.c:1982-2097.  Writes four right-aligned
 * sequences; returns -1 when a read falls outside the written band. */
static int aff_backtrace(const uint16_t *dm, const uint8_t *A, int la, const uint8_t *B, int lb,
                         const po_cm2 *c, uint8_t *median, uint8_t *mwg, uint8_t *ra, uint8_t *rb,
                         int cap, int *lens) {
    enum { M_TODO, M_V, M_H, M_D, M_A } mode = M_TODO;
    int i = la, j = lb, nm = 0, nw = 0, na = 0, nb = 0;
    const int W = lb + 1;
    int ic = A[i], jc = B[j];
#define PUT(buf, n, v) do { if ((n) >= cap) return -1; (buf)[cap - ++(n)] = (uint8_t)(v); } while (0)
    while (i != 0 && j != 0) {
        int lo, hi;
        aff_row_range(la, lb, i, &lo, &hi);
        if (j < lo || j > hi) return -1;
        int f = dm[(size_t)i * W + j];
        if (mode == M_TODO) {
            if (f & A_DO_HORIZONTAL) mode = M_H;
            else if (f & A_DO_ALIGN) mode = M_A;
            else if (f & A_DO_VERTICAL) mode = M_V;
            else mode = M_D;
        } else if (mode == M_V) {
            if (f & A_END_VERTICAL) mode = M_TODO;
            if (!(ic & TMPGAP)) { PUT(median, nm, ic | TMPGAP); PUT(mwg, nw, ic | TMPGAP); }
            else PUT(mwg, nw, TMPGAP);
            PUT(ra, na, ic); PUT(rb, nb, TMPGAP);
            i--; ic = A[i];
        } else if (mode == M_H) {
            if (f & A_END_HORIZONTAL) mode = M_TODO;
            if (!(jc & TMPGAP)) { PUT(median, nm, jc | TMPGAP); PUT(mwg, nw, jc | TMPGAP); }
            else PUT(mwg, nw, TMPGAP);
            PUT(ra, na, TMPGAP); PUT(rb, nb, jc);
            j--; jc = B[j];
        } else if (mode == M_D) {
            if (f & A_END_BLOCK) mode = M_TODO;
            PUT(ra, na, ic); PUT(rb, nb, jc); PUT(mwg, nw, TMPGAP);
            i--; j--; jc = B[j]; ic = A[i];
        } else {
            if (f & A_ALIGN_TO_HORIZONTAL) mode = M_H;
            else if (f & A_ALIGN_TO_DIAGONAL) mode = M_D;
            else if (f & A_ALIGN_TO_VERTICAL) mode = M_V;
            int prep = po_cm2_median(c, ic & NTMPGAP, jc & NTMPGAP);
            PUT(median, nm, prep); PUT(mwg, nw, prep); PUT(ra, na, ic); PUT(rb, nb, jc);
            i--; j--; jc = B[j]; ic = A[i];
        }
    }
    while (i != 0) {
        if (!(ic & TMPGAP)) { PUT(median, nm, ic | TMPGAP); PUT(mwg, nw, ic | TMPGAP); }
        else PUT(mwg, nw, TMPGAP);
        PUT(ra, na, ic); PUT(rb, nb, TMPGAP);
        i--; ic = A[i];
    }
    while (j != 0) {
        if (!(jc & TMPGAP)) { PUT(median, nm, jc | TMPGAP); PUT(mwg, nw, jc | TMPGAP); }
        else PUT(mwg, nw, TMPGAP);
        PUT(ra, na, TMPGAP); PUT(rb, nb, jc);
        j--; jc = B[j];
    }
    PUT(ra, na, TMPGAP); PUT(rb, nb, TMPGAP); PUT(mwg, nw, TMPGAP);
    /* an empty median reads past its buffer in the reference; a gap is
     * prepended in that case here */
    if (nm == 0 || median[cap - nm] != TMPGAP) PUT(median, nm, TMPGAP);
#undef PUT
    lens[0] = nm; lens[1] = nw; lens[2] = na; lens[3] = nb;
    return 0;
}

/* algn_CAML_align_affine_3, src/algn.c:2550-2614. */
int po_align_affine_3(const uint8_t *si, int li, const uint8_t *sj, int lj, const po_cm2 *c,
                      uint8_t *median, uint8_t *medianwg, uint8_t *resi, uint8_t *resj,
                      int cap, int *lens) {
    int swap = li > lj;
    const uint8_t *A = swap ? sj : si, *B = swap ? si : sj;
    int la = (swap ? lj : li) - 1, lb = (swap ? li : lj) - 1;
    aff_bufs bf;
    if (aff_alloc(&bf, lb) != 0) return -1;
    uint16_t *dm = (uint16_t *)scratch_grow(&g_dm, &g_dm_cap, sizeof(uint16_t) * (size_t)(la + 1) * (lb + 1));
    if (!dm) return -1;
    int cost = aff_fill(A, la, B, lb, c, 1, dm, &bf);
    int l4[4];
    int rc = aff_backtrace(dm, A, la, B, lb, c, median, medianwg, swap ? resj : resi, swap ? resi : resj, cap, l4);
    if (rc != 0) return -1;
    lens[0] = l4[0]; lens[1] = l4[1];
    lens[2] = swap ? l4[3] : l4[2];
    lens[3] = swap ? l4[2] : l4[3];
    return cost;
}

/* algn_CAML_cost_affine_3, src/algn.c:2627-2680. */
int po_cost_affine_3(const uint8_t *si, int li, const uint8_t *sj, int lj, const po_cm2 *c) {
    int swap = li > lj;
    const uint8_t *A = swap ? sj : si, *B = swap ? si : sj;
    int la = (swap ? lj : li) - 1, lb = (swap ? li : lj) - 1;
    aff_bufs bf;
    if (aff_alloc(&bf, lb) != 0) return -1;
    return aff_fill(A, la, B, lb, c, 0, NULL, &bf);
}

/* ======================================================================== */
/* R10/R11: three-dimensional alignment                                     */
/* ======================================================================== */

/* algn_fill_cube, src/algn.c:2885-3055, as the reference executes it.  Row r
 * of the flat (plane x line) row space holds n3 cells.  Planes p >= 1 read
 * their "diagonal", "upper" and "previous" rows through three pointers that
 * advance n2-1 rows per plane while the output advances n2 rows (the first
 * row of a plane advances only the output, :2988-3017); the rows read are
 * therefore D(p)+r-1, U(p)+r-1 and P(p)+r-1 for line r >= 1, with
 * D(p) = (p-1)(n2-1), U(p) = D(p)+1, P(p) = n2-1+(p-1)(n2-1).  For p = 1 and
 * lines >= 2 the "previous" row lags the textbook one by a row, and the lag
 * grows by one row per plane. */
static int cube_fill(const uint8_t *s1, int n1, const uint8_t *s2, int n2, const int *prec, int n3,
                     int *mm, uint16_t *dm, int gap, int a_sz) {
    const int stride = (a_sz + 1) * n3;   /* cm_get_pos_in_precalc, src/cm.c:735-741 */
#define PREC_ROW(x, y) (prec + (size_t)(x) * stride + (size_t)n3 * (y))
    const int *ggs3 = PREC_ROW(gap, gap);
    int *m0 = mm;
    uint16_t *d0 = dm;
    /* plane 0 */
    m0[0] = 0; d0[0] = C_S2;
    for (int k = 1; k < n3; k++) { m0[k] = m0[k - 1] + ggs3[k]; d0[k] = C_SS; }
    for (int r = 1; r < n2; r++) {
        const int *gs2s3 = PREC_ROW(gap, s2[r]);
        int *cur = mm + (size_t)r * n3, *prv = cur - n3;
        uint16_t *dc = dm + (size_t)r * n3;
        cur[0] = prv[0] + gs2s3[0]; dc[0] = C_P1;
        for (int k = 1; k < n3; k++) {
            int v = prv[k] + gs2s3[0], code = C_P1, t;
            t = prv[k - 1] + gs2s3[k]; if (t < v) { v = t; code = C_S1; }
            t = cur[k - 1] + ggs3[k]; if (t < v) { v = t; code = C_SS; }
            cur[k] = v; dc[k] = (uint16_t)code;
        }
    }
    size_t out = (size_t)n2;   /* next output row */
    size_t D = 0, U = 1, P = out - 1;
    for (int p = 1; p < n1; p++) {
        int s1_it = s1[p];
        const int *s1gs3 = PREC_ROW(s1_it, gap);
        {
            int *cur = mm + out * n3;
            const int *dg = mm + D * n3;
            uint16_t *dc = dm + out * n3;
            cur[0] = dg[0] + s1gs3[0]; dc[0] = C_P3;
            for (int k = 1; k < n3; k++) {
                int v = dg[k] + s1gs3[0], code = C_P3, t;
                t = dg[k - 1] + s1gs3[k]; if (t < v) { v = t; code = C_S3; }
                t = ggs3[k] + cur[k - 1]; if (t < v) { v = t; code = C_SS; }
                cur[k] = v; dc[k] = (uint16_t)code;
            }
            out++;
        }
        for (int r = 1; r < n2; r++, D++, U++, P++, out++) {
            int s2_it = s2[r];
            const int *gs2s3 = PREC_ROW(gap, s2_it), *s1s2s3 = PREC_ROW(s1_it, s2_it);
            int *cur = mm + out * n3;
            const int *dg = mm + D * n3, *up = mm + U * n3, *pv = mm + P * n3;
            uint16_t *dc = dm + out * n3;
            int s1gg = s1gs3[0], gs2g = gs2s3[0], s1s2g = s1s2s3[0];
            for (int k = 0; k < n3; k++) {   /* fill_parallel, :2840-2858 */
                int v = up[k] + s1gg, code = C_P3, t;
                t = pv[k] + gs2g; if (t < v) { v = t; code = C_P1; }
                t = dg[k] + s1s2g; if (t < v) { v = t; code = C_P2; }
                cur[k] = v; dc[k] = (uint16_t)code;
            }
            for (int k = 1; k < n3; k++) {   /* fill_moved, :2812-2833 */
                int t;
                t = up[k - 1] + s1gs3[k]; if (t < cur[k]) { cur[k] = t; dc[k] = C_S3; }
                t = pv[k - 1] + gs2s3[k]; if (t < cur[k]) { cur[k] = t; dc[k] = C_S1; }
                t = dg[k - 1] + s1s2s3[k]; if (t < cur[k]) { cur[k] = t; dc[k] = C_S2; }
            }
            for (int k = 1; k < n3; k++) {   /* self pass along s3, :3039-3047 */
                int t = cur[k - 1] + ggs3[k];
                if (t < cur[k]) { cur[k] = t; dc[k] = C_SS; }
            }
        }
    }
#undef PREC_ROW
    return mm[out * n3 - 1];
}

/* algn_nw_3d (src/algn.c:3290-3313) with cm_precalc_4algn_3d
 * (src/cm.c:748-764).  dm, when given, receives the n1*n2*n3 codes. */
static int nw_3d(const uint8_t *s1, int n1, const uint8_t *s2, int n2, const uint8_t *s3, int n3,
                 const po_cm3 *c, uint16_t **dm_out) {
    size_t cells = (size_t)n1 * n2 * n3;
    int *mm = (int *)scratch_grow(&g_cube, &g_cube_cap, cells * sizeof(int));
    uint16_t *dm = (uint16_t *)scratch_grow(&g_cubed, &g_cubed_cap, cells * sizeof(uint16_t));
    size_t nprec = (size_t)(c->a_sz + 1) * (c->a_sz + 1) * n3;
    int *prec = (int *)scratch_grow(&g_prec, &g_prec_cap, nprec * sizeof(int));
    if (!mm || !dm || !prec) return -1;
    for (int j = 1; j <= c->a_sz; j++)
        for (int k = 1; k <= c->a_sz; k++) {
            const int *row = c->cost + ((((size_t)j << c->lcm) + k) << c->lcm);
            int *to = prec + (size_t)j * (c->a_sz + 1) * n3 + (size_t)n3 * k;
            for (int i = 0; i < n3; i++) to[i] = row[s3[i]];
        }
    int res = cube_fill(s1, n1, s2, n2, prec, n3, mm, dm, c->gap, c->a_sz);
    if (dm_out) *dm_out = dm;
    return res;
}

int po_cost_3(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const uint8_t *s3, int l3,
              const po_cm3 *c) {
    return nw_3d(s1, l1, s2, l2, s3, l3, c, NULL);
}

/* Sequence.Align.align_3 ~first_gap:true (src/sequence.ml:871-895) ->
 * algn_CAML_align_3d (src/algn.c:4004-4012) -> backtrack_3d (:3829-3905).
 * The walk stops on reaching cell (0,0,0) without emitting it. */
int po_align_3(const uint8_t *s1, int l1, const uint8_t *s2, int l2, const uint8_t *s3, int l3,
               const po_cm3 *c, uint8_t *o1, uint8_t *o2, uint8_t *o3, int cap, int *olen) {
    uint16_t *dm;
    int cost = nw_3d(s1, l1, s2, l2, s3, l3, c, &dm);
    if (cost < 0 && !dm) return -1;
    size_t a_plane = (size_t)l2 * l3, a_line = (size_t)l3;
    size_t e = (size_t)l1 * l2 * l3 - 1;
    int i1 = l1 - 1, i2 = l2 - 1, i3 = l3 - 1, n = 0;
    const int g = c->gap;
    while (e > 0) {
        int d = dm[e];
        if (n >= cap) return -1;
        n++;
        if (d & C_S2) { o1[cap - n] = s1[i1--]; o2[cap - n] = s2[i2--]; o3[cap - n] = s3[i3--]; e -= a_plane + a_line + 1; }
        else if (d & C_S3) { o1[cap - n] = s1[i1--]; o2[cap - n] = (uint8_t)g; o3[cap - n] = s3[i3--]; e -= a_plane + 1; }
        else if (d & C_S1) { o1[cap - n] = (uint8_t)g; o2[cap - n] = s2[i2--]; o3[cap - n] = s3[i3--]; e -= a_line + 1; }
        else if (d & C_P3) { o1[cap - n] = s1[i1--]; o2[cap - n] = (uint8_t)g; o3[cap - n] = (uint8_t)g; e -= a_plane; }
        else if (d & C_SS) { o1[cap - n] = (uint8_t)g; o2[cap - n] = (uint8_t)g; o3[cap - n] = s3[i3--]; e -= 1; }
        else if (d & C_P1) { o1[cap - n] = (uint8_t)g; o2[cap - n] = s2[i2--]; o3[cap - n] = (uint8_t)g; e -= a_line; }
        else if (d & C_P2) { o1[cap - n] = s1[i1--]; o2[cap - n] = s2[i2--]; o3[cap - n] = (uint8_t)g; e -= a_plane + a_line; }
        else return -1;
        if (i1 < -1 || i2 < -1 || i3 < -1) return -1;
    }
    *olen = n;
    return cost;
}

/* algn_get_median_3d (src/algn.c:4160-4173): the end pointers are never
 * moved, so every position receives the median of the three last symbols. */
int po_median_3(const uint8_t *a, const uint8_t *b, const uint8_t *d, int len, const po_cm3 *c,
                uint8_t *out, int cap) {
    if (len <= 0) return 0;
    int m = po_cm3_median(c, a[len - 1], b[len - 1], d[len - 1]);
    for (int n = 1; n <= len; n++) out[cap - n] = (uint8_t)m;
    return len;
}

/* ======================================================================== */
/* R12/R13: Fitch                                                           */
/* ======================================================================== */

/* nonadd_median (src/nonaddCSc.c:422-462): res = a&b if non-empty else a|b;
 * the added cost counts the characters with an empty intersection. */
int64_t po_nonadd_median(const uint8_t *a, const uint8_t *b, uint8_t *res, int len) {
    int64_t added = 0;
    for (int i = 0; i < len; i++) {
        uint8_t in = a[i] & b[i], un = a[i] | b[i];
        uint8_t emp = in == 0 ? 0xFF : 0;
        if (emp) added++;
        res[i] = (uint8_t)((emp & un) | in);
    }
    return added;
}

/* nonadd_median_3 (src/nonaddCSc.c:472-544), the Fitch final pass. */
void po_nonadd_median_3(const uint8_t *A, const uint8_t *N, const uint8_t *D1, const uint8_t *D2,
                        uint8_t *F, int len) {
    for (int i = 0; i < len; i++) {
        uint8_t inter = (uint8_t)(N[i] | (A[i] & (D1[i] | D2[i])));
        uint8_t temp = inter;
        uint8_t was_union = (uint8_t)((D1[i] & D2[i]) == 0 ? 0xFF : 0);
        inter = (uint8_t)((inter | A[i]) & was_union);
        temp |= inter;
        uint8_t contained = (uint8_t)((A[i] & N[i]) == A[i] ? 0xFF : 0);
        uint8_t f = contained & A[i];
        uint8_t not_contained = (uint8_t)(contained == 0 ? 0xFF : 0);
        F[i] = (uint8_t)(f | (not_contained & temp));
    }
}

/* The same two passes for characters of `width` bytes (1, 2 or 4): the
 * reference compiles nonaddCSc.c with CHARSIZE 8, 16 and 32
 * (src/myocamlbuild.ml.in:255-269) and picks the width from the number of
 * states (src/node.ml:1179-1234). */
static uint32_t po_get_w(const void *p, int width, int i) {
    if (width == 1) return ((const uint8_t *)p)[i];
    if (width == 2) return ((const uint16_t *)p)[i];
    return ((const uint32_t *)p)[i];
}
static void po_put_w(void *p, int width, int i, uint32_t v) {
    if (width == 1) ((uint8_t *)p)[i] = (uint8_t)v;
    else if (width == 2) ((uint16_t *)p)[i] = (uint16_t)v;
    else ((uint32_t *)p)[i] = v;
}

int64_t po_nonadd_median_w(int width, const void *a, const void *b, void *res, int len) {
    int64_t added = 0;
    for (int i = 0; i < len; i++) {
        uint32_t x = po_get_w(a, width, i), y = po_get_w(b, width, i);
        uint32_t in = x & y;
        if (in == 0) added++;
        po_put_w(res, width, i, in ? in : (x | y));
    }
    return added;
}

void po_nonadd_median_3_w(int width, const void *A, const void *N, const void *D1, const void *D2, void *F,
                          int len) {
    for (int i = 0; i < len; i++) {
        uint32_t a = po_get_w(A, width, i), n = po_get_w(N, width, i);
        uint32_t d1 = po_get_w(D1, width, i), d2 = po_get_w(D2, width, i);
        uint32_t inter = n | (a & (d1 | d2));
        uint32_t temp = inter;
        uint32_t was_union = (d1 & d2) == 0 ? 0xFFFFFFFFu : 0;
        inter = (inter | a) & was_union;
        temp |= inter;
        uint32_t f = (a & n) == a ? a : temp;
        po_put_w(F, width, i, f);
    }
}

/* char_nonadd_CAML_distance, src/nonaddCSc.c:1076-1105. */
int64_t po_nonadd_distance(const uint8_t *a, const uint8_t *b, int len) {
    int64_t r = 0;
    for (int i = 0; i < len; i++) if ((a[i] & b[i]) == 0) r++;
    return r;
}

/* char_nonadd_CAML_dist_2, src/nonaddCSc.c:1349-1376. */
int64_t po_nonadd_dist_2(const uint8_t *a, const uint8_t *b, const uint8_t *c, int len) {
    int64_t r = 0;
    for (int i = 0; i < len; i++) if (!(a[i] & b[i]) && !(a[i] & c[i])) r++;
    return r;
}

/* nonadd_reroot_median, src/nonaddCSc.c:556-575. */
void po_nonadd_reroot_median(const uint8_t *a, const uint8_t *b, uint8_t *res, int len) {
    for (int i = 0; i < len; i++) res[i] = a[i] | b[i];
}

/* ---- R15: powell_3D_align (src/ukkCommon.c:110-151, 236-330, 380-445;
 * src/ukk.checkp.c:113-212, 446-500) -------------------------------------
 * Powell's three-way alignment with affine gaps.  The reference runs an
 * Ukkonen-style search (cost-bounded furthest reaching with check points),
 * which returns the least cost of any alignment under its model; this
 * restatement finds the same least cost with an exhaustive DP over the model:
 *   states: per sequence one of match / del / ins (enumerated as
 *     s = t0 + 3 t1 + 9 t2, match 0, del 1, ins 2), kept when at least one
 *     sequence matches and at most one inserts (setup, :259-269);
 *   a state consumes the characters of its matching sequences when it has no
 *     insert, else the inserting sequence's only (neighbours, :274-285);
 *   entering state t from s costs go for every sequence whose state in t is
 *     del or ins and differs from s (transCost, :310-320), then ge for an
 *     insert column, else ge per del plus mm per mismatch among the matched
 *     characters (0 / 1 / 2 for three: whichCharCost, :157-184; 0 / 1 for
 *     two), as alignmentCost (:380-445) recomputes a finished alignment;
 *   the alignment starts in the all-match state at cost 0.
 * Characters are the first set bit of each 5-bit mask after the leading
 * gap (copySequence, :88-108); outputs get the leading gap back
 * (printTraceBack, :456-500).  The least cost equals the reference's; among
 * alignments of that cost this one takes, at every cell, the first least
 * previous state and, at the end, the first least state. */
#define PW_NSTATES 16
typedef struct {
  int st[3], adv[3], n_ins, n_del;
} pw_state;

static int pw_setup(pw_state *S, int trans[PW_NSTATES][PW_NSTATES], int go) {
  int ns = 0;
  for (int s = 0; s < 27; ++s) {
    int st[3] = {s % 3, (s / 3) % 3, (s / 9) % 3};
    int nm = 0, nd = 0, ni = 0;
    for (int q = 0; q < 3; ++q) {
      nm += st[q] == 0;
      nd += st[q] == 1;
      ni += st[q] == 2;
    }
    if (nm == 0 || ni > 1) continue;
    for (int q = 0; q < 3; ++q) {
      S[ns].st[q] = st[q];
      S[ns].adv[q] = ni == 0 ? st[q] == 0 : st[q] == 2;
    }
    S[ns].n_ins = ni;
    S[ns].n_del = nd;
    ++ns;
  }
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b < ns; ++b) {
      int c = 0;
      for (int q = 0; q < 3; ++q)
        if (S[b].st[q] != 0 && S[b].st[q] != S[a].st[q]) c += go;
      trans[a][b] = c;
    }
  return ns;
}

static int pw_base(uint8_t m) {
  return (m & 1) ? 'A' : (m & 2) ? 'C' : (m & 4) ? 'G' : (m & 8) ? 'T' : 0;
}

/* column cost of state t whose consumed characters end at (i, j, k) */
static int pw_step(const pw_state *t, const char *x[3], const int pos[3], int mm, int ge) {
  if (t->n_ins) return ge;
  int c = ge * t->n_del, ch[3], n = 0;
  for (int q = 0; q < 3; ++q)
    if (t->st[q] == 0) ch[n++] = x[q][pos[q] - 1];
  if (n == 3) c += mm * ((ch[0] == ch[1] && ch[0] == ch[2]) ? 0 : (ch[0] == ch[1] || ch[0] == ch[2] || ch[1] == ch[2]) ? 1 : 2);
  else if (n == 2) c += mm * (ch[0] != ch[1]);
  return c;
}

/* Inputs include POY's leading gap (la, lb, lc >= 1).  Returns the cost, or
 * -1 when a character has no base bit (the reference fails there) or memory
 * runs out.  When oa is given, writes the three aligned sequences (masks,
 * gap 16, leading gap first) and their length to *olen; cap >= la+lb+lc. */
int po_powell_3d(const uint8_t *A, int la, const uint8_t *B, int lb, const uint8_t *Cs, int lc, int mm,
                 int go, int ge, uint8_t *oa, uint8_t *ob, uint8_t *oc, int cap, int *olen) {
  pw_state S[PW_NSTATES];
  int trans[PW_NSTATES][PW_NSTATES];
  const int ns = pw_setup(S, trans, go);
  const int n1 = la - 1, n2 = lb - 1, n3 = lc - 1;
  if (n1 < 0 || n2 < 0 || n3 < 0) return -1;
  char *x[3];
  const uint8_t *src[3] = {A, B, Cs};
  const int len[3] = {n1, n2, n3};
  for (int q = 0; q < 3; ++q) {
    x[q] = (char *)malloc((size_t)len[q] + 1);
    for (int i = 0; i < len[q]; ++i) {
      x[q][i] = (char)pw_base(src[q][i + 1]);
      if (!x[q][i]) {
        for (int r = 0; r <= q; ++r) free(x[r]);
        return -1;
      }
    }
  }
  const int64_t d1 = n1 + 1, d2 = n2 + 1, d3 = n3 + 1, cells = d1 * d2 * d3;
  int *V = (int *)malloc(sizeof(int) * cells * ns);
  uint8_t *from = (uint8_t *)malloc((size_t)cells * ns);
  if (!V || !from) {
    free(V); free(from);
    for (int q = 0; q < 3; ++q) free(x[q]);
    return -1;
  }
  const int INF = 0x3fffffff;
  const char *xc[3] = {x[0], x[1], x[2]};
  for (int i = 0; i <= n1; ++i)
    for (int j = 0; j <= n2; ++j)
      for (int k = 0; k <= n3; ++k) {
        const int64_t c = (i * d2 + j) * d3 + k;
        for (int t = 0; t < ns; ++t) {
          int best = INF, arg = 0;
          if (i == 0 && j == 0 && k == 0) {
            best = t == 0 ? 0 : INF;
          } else {
            const int pi = i - S[t].adv[0], pj = j - S[t].adv[1], pk = k - S[t].adv[2];
            if (pi >= 0 && pj >= 0 && pk >= 0) {
              const int64_t pc = (pi * d2 + pj) * d3 + pk;
              for (int s = 0; s < ns; ++s) {
                const int v = V[pc * ns + s];
                if (v >= INF) continue;
                const int w = v + trans[s][t];
                if (w < best) { best = w; arg = s; }
              }
              if (best < INF) {
                const int pos[3] = {i, j, k};
                best += pw_step(&S[t], xc, pos, mm, ge);
              }
            }
          }
          V[c * ns + t] = best;
          from[c * ns + t] = (uint8_t)arg;
        }
      }
  const int64_t end = cells - 1;
  int cost = INF, st = 0;
  for (int t = 0; t < ns; ++t)
    if (V[end * ns + t] < cost) { cost = V[end * ns + t]; st = t; }
  if (oa) {
    int i = n1, j = n2, k = n3, n = 0;
    uint8_t *o[3] = {oa, ob, oc};
    /* collect backwards, then reverse */
    while (i > 0 || j > 0 || k > 0) {
      const int pos[3] = {i, j, k};
      for (int q = 0; q < 3; ++q) {
        const int ch = S[st].adv[q] ? xc[q][pos[q] - 1] : '-';
        o[q][n] = ch == 'A' ? 1 : ch == 'C' ? 2 : ch == 'G' ? 4 : ch == 'T' ? 8 : 16;
      }
      ++n;
      const int64_t c = (i * d2 + j) * d3 + k;
      const int prev = from[c * ns + st];
      i -= S[st].adv[0]; j -= S[st].adv[1]; k -= S[st].adv[2];
      st = prev;
      if (n >= cap - 1) break;
    }
    for (int q = 0; q < 3; ++q) {
      for (int a = 0, b = n - 1; a < b; ++a, --b) {
        const uint8_t t = o[q][a]; o[q][a] = o[q][b]; o[q][b] = t;
      }
      memmove(o[q] + 1, o[q], (size_t)n);
      o[q][0] = 16;
    }
    *olen = n + 1;
  }
  free(V); free(from);
  for (int q = 0; q < 3; ++q) free(x[q]);
  return cost;
}

/* ---- R15: the check-point Ukkonen search itself (src/ukk.checkp.c) -------
 * Restates the search powell_3D_align runs, so that the alignment it picks
 * among those of least cost is the reference's:
 *   store: a cell per (ab = a - b, ac = a - c, cost d, state s) holding the
 *     furthest a reached (dist), a tag (d + the pass's offset: a cell is
 *     valid when its tag matches) and its "from" record; the reference keeps
 *     2 maxSingleStep cost planes and reuses plane d mod that count
 *     (src/ukkCommon.h:101-260 with FIXED_NUM_PLANES, src/ukk.checkp.c:133);
 *     `planes_mode` 1 keeps every plane instead (no reuse);
 *   check-point cells: (dist, d + 1) for d in [CPcost, CPcost + CPwidth),
 *     CPwidth = maxSingleStep planes;
 *   ukk (Ukk, :578-606): outside withinMatrix (:545-576) -INFINITY; a valid
 *     cell's dist; else calc, store the check point, raise furthestReached;
 *   calc (calcUkk, :608-781): the state's one neighbour diagonal, every
 *     previous state in order (first the cost without the second mismatch
 *     when whichCharCost says 1, else, for states with a second cost, the
 *     cost one mismatch lower), then the same state one cost lower, then
 *     for the all-match state the furthest state of this cost extended
 *     along equal characters; a candidate replaces the best only when
 *     strictly further; the from record is the candidate's coordinates in a
 *     base case, the candidate's own record above the check-point band,
 *     this cell inside the band, cost -1 below it;
 *   the driver (doUkk, :113-214): the common prefix, costs 1, 2, ... at the
 *     end diagonal until the best state reaches |A|, the check-point cost
 *     set to d + 1 after the first cost whose furthest reach is |A| / 2;
 *     then the end cell's from record splits the alignment at the check
 *     point (getSplitRecurse, :324-368; the later half first), or, when it
 *     has none, the whole range is searched again (doUkkInLimits,
 *     :216-322: a fresh offset, the start cell preset, base case when the
 *     range is at most CPwidth, else CPcost = (fCost + sCost - CPwidth + 1)
 *     / 2);
 *   traceBack (:370-454): from the end cell back along from records, a run
 *     of match columns down to the previous cell's dist, then one column of
 *     the cell's state; printTraceBack (:456-500): the common prefix, the
 *     columns reversed, bases as masks (char_to_base), the leading gap.
 * Model tables as setup (src/ukkCommon.c:247-330); okIndex and
 * whichCharCost as src/ukkCommon.c:155-195. */
#define UK_INF 5000   /* INFINITY = MAXINT / 2 */
#define UK_BLK 8

typedef struct {
  int dist;
  int64_t tag;
  int fab, fac, fcost, fstate;
} uk_cell;

typedef struct {
  int dist, cost;
} uk_cp;

typedef struct {
  void ***plane;   /* plane -> block table -> block */
  int nplanes, cap, aliasing;
  size_t elem;
} uk_store;

typedef struct {
  /* model */
  int ns, nb[PW_NSTATES], cont[PW_NSTATES], second[PW_NSTATES], trans[PW_NSTATES][PW_NSTATES];
  int mis, go, ge, max_step;
  /* strings, each followed by a 0 as copySequence leaves them */
  char *A, *B, *C;
  int la, lb, lc;
  /* stores */
  uk_store U, CP;
  int ab_off, ac_off, ab_blocks, ac_blocks;
  /* search state */
  int64_t offset, final_cost;
  int furthest, cp_on_dist, cp_cost, cp_width, complete_from;
  int sab, sac, scost, sstate;
  int endA, endB, endC;
  /* output columns, in reverse */
  char *ra, *rb, *rc;
  int n, out_cap, oom;
} uk_ctx;

static void *uk_slot(uk_ctx *k, uk_store *st, int ab, int ac, int d, int s) {
  int p = st->aliasing ? d % st->nplanes : d;
  if (p >= st->cap) {
    int ncap = st->cap ? st->cap : 16;
    while (ncap <= p) ncap *= 2;
    void ***np = (void ***)realloc(st->plane, sizeof(void **) * (size_t)ncap);
    if (!np) { k->oom = 1; return NULL; }
    memset(np + st->cap, 0, sizeof(void **) * (size_t)(ncap - st->cap));
    st->plane = np;
    st->cap = ncap;
  }
  if (!st->plane[p]) {
    st->plane[p] = (void **)calloc((size_t)k->ab_blocks * k->ac_blocks, sizeof(void *));
    if (!st->plane[p]) { k->oom = 1; return NULL; }
  }
  const int i = (ab + k->ab_off) / UK_BLK, j = (ac + k->ac_off) / UK_BLK;
  void **blk = &st->plane[p][(size_t)i * k->ac_blocks + j];
  if (!*blk) {
    *blk = calloc((size_t)UK_BLK * UK_BLK * k->ns, st->elem);
    if (!*blk) { k->oom = 1; return NULL; }
  }
  const int ai = (ab + k->ab_off) - i * UK_BLK, ci = (ac + k->ac_off) - j * UK_BLK;
  return (char *)*blk + (((size_t)ai * UK_BLK + ci) * k->ns + s) * st->elem;
}

static uk_cell uk_dummy;
static uk_cell *uk_U(uk_ctx *k, int ab, int ac, int d, int s) {
  uk_cell *c = (uk_cell *)uk_slot(k, &k->U, ab, ac, d, s);
  return c ? c : &uk_dummy;
}
static uk_cp uk_cp_dummy;
static uk_cp *uk_CP(uk_ctx *k, int ab, int ac, int d, int s) {
  uk_cp *c = (uk_cp *)uk_slot(k, &k->CP, ab, ac, d, s);
  return c ? c : &uk_cp_dummy;
}

static void uk_free_store(uk_ctx *k, uk_store *st) {
  for (int p = 0; p < st->cap; ++p) {
    if (!st->plane[p]) continue;
    for (size_t b = 0; b < (size_t)k->ab_blocks * k->ac_blocks; ++b) free(st->plane[p][b]);
    free(st->plane[p]);
  }
  free(st->plane);
}

static int uk_ok(int a, int da, int end) {
  if (a < 0) return 0;
  return da ? a < end : a <= end;
}

static int uk_chars(char a, char b, char c) {
  if (a == b && a == c) return 0;
  return (a == b || a == c || b == c) ? 1 : 2;
}

static int uk_within(uk_ctx *k, int ab, int ac, int d) {
  if (d < 0) return 0;
  int v[3] = {abs(k->sab - ab), abs(k->sac - ac), abs((k->sac - k->sab) - (ac - ab))};
  /* the two smallest */
  for (int x = 0; x < 2; ++x)
    for (int y = x + 1; y < 3; ++y)
      if (v[y] < v[x]) { int t = v[x]; v[x] = v[y]; v[y] = t; }
  const int g = v[0], h = v[1];
  int cheapest;
  if (k->sstate == 0)
    cheapest = (g ? k->go + g * k->ge : 0) + (h ? k->go + h * k->ge : 0);
  else
    cheapest = g * k->ge + h * k->ge;
  return cheapest + k->scost <= d;
}

static int uk_calc(uk_ctx *k, int ab, int ac, int d, int to);

static int uk_ukk(uk_ctx *k, int ab, int ac, int d, int s) {
  if (!uk_within(k, ab, ac, d)) return -UK_INF;
  uk_cell *c = uk_U(k, ab, ac, d, s);
  if (c->tag == d + k->offset) return c->dist;
  uk_calc(k, ab, ac, d, s);
  c = uk_U(k, ab, ac, d, s);
  if (d >= k->cp_cost && d < k->cp_cost + k->cp_width) {
    uk_cp *p = uk_CP(k, ab, ac, d, s);
    p->dist = c->dist;
    p->cost = d + 1;
  }
  if (c->dist > k->furthest) k->furthest = c->dist;
  return c->dist;
}

static int uk_calc(uk_ctx *k, int ab, int ac, int d, int to) {
  int fab = 0, fac = 0, fcost = -1, fstate = 0;
  int best = -UK_INF;
  if (d >= k->cp_cost && d < k->cp_cost + k->cp_width) {
    fab = ab; fac = ac; fcost = d; fstate = to;
  }
  const int da = k->nb[to] & 1, db = (k->nb[to] >> 1) & 1, dc = (k->nb[to] >> 2) & 1;
  const int ab1 = ab - da + db, ac1 = ac - da + dc;
#define UK_TAKE(cab, cac, ccost, cstate)                                   \
  do {                                                                     \
    if (k->complete_from) {                                                \
      fab = (cab); fac = (cac); fcost = (ccost); fstate = (cstate);        \
    } else if (d >= k->cp_cost + k->cp_width) {                            \
      const uk_cell *q = uk_U(k, (cab), (cac), (ccost), (cstate));         \
      fab = q->fab; fac = q->fac; fcost = q->fcost; fstate = q->fstate;    \
    }                                                                      \
  } while (0)
  if (ab1 >= -k->endB && ab1 <= k->endA && ac1 >= -k->endC && ac1 <= k->endA) {
    for (int fs = 0; fs < k->ns; ++fs) {
      const int cost = d - k->trans[fs][to] - k->cont[to];
      int dist = -UK_INF, from_cost = -UK_INF;
      const int a1 = uk_ukk(k, ab1, ac1, cost, fs);
      if (uk_ok(a1, da, k->endA) && uk_ok(a1 - ab1, db, k->endB) && uk_ok(a1 - ac1, dc, k->endC) &&
          uk_chars(da ? k->A[a1] : '-', db ? k->B[a1 - ab1] : '-', dc ? k->C[a1 - ac1] : '-') == 1) {
        from_cost = cost;
        dist = a1 + da;
      } else {
        if (!k->second[to]) continue;
        const int a2 = uk_ukk(k, ab1, ac1, cost - k->mis, fs);
        if (uk_ok(a2, da, k->endA) && uk_ok(a2 - ab1, db, k->endB) && uk_ok(a2 - ac1, dc, k->endC)) {
          from_cost = cost - k->mis;
          dist = a2 + da;
        }
      }
      if (best < dist) {
        best = dist;
        UK_TAKE(ab1, ac1, from_cost, fs);
      }
    }
  }
  {
    const int dist = uk_ukk(k, ab, ac, d - 1, to);
    if (uk_ok(dist, 0, k->endA) && uk_ok(dist - ab, 0, k->endB) && uk_ok(dist - ac, 0, k->endC) && best < dist) {
      best = dist;
      UK_TAKE(ab, ac, d - 1, to);
    }
  }
  if (to == 0) {
    int dist = -UK_INF, from_state = -1;
    for (int s = 0; s < k->ns; ++s) {
      const int x = s == 0 ? best : uk_ukk(k, ab, ac, d, s);
      if (x > dist) { dist = x; from_state = s; }
    }
    while (uk_ok(dist, 1, k->endA) && uk_ok(dist - ab, 1, k->endB) && uk_ok(dist - ac, 1, k->endC) &&
           k->A[dist] == k->B[dist - ab] && k->A[dist] == k->C[dist - ac])
      ++dist;
    if (dist > best) {
      best = dist;
      if (from_state != 0) UK_TAKE(ab, ac, d, from_state);
    }
  }
#undef UK_TAKE
  uk_cell *c = uk_U(k, ab, ac, d, to);
  c->dist = best;
  c->tag = d + k->offset;
  c->fab = fab; c->fac = fac; c->fcost = fcost; c->fstate = fstate;
  return best;
}

static int uk_best(uk_ctx *k, int ab, int ac, int d, int want_state) {
  int best = -UK_INF, st = -1;
  for (int s = 0; s < k->ns; ++s) {
    const uk_cell *c = uk_U(k, ab, ac, d, s);
    if (c->tag == d + k->offset && c->dist > best) { best = c->dist; st = s; }
  }
  return want_state ? st : best;
}

static void uk_emit(uk_ctx *k, char a, char b, char c) {
  if (k->n >= k->out_cap) { k->oom = 1; return; }
  k->ra[k->n] = a; k->rb[k->n] = b; k->rc[k->n] = c;
  ++k->n;
}

static void uk_trace(uk_ctx *k, int sab, int sac, int scost, int sstate, int fab, int fac, int fcost, int fstate) {
  int ab = fab, ac = fac, d = fcost, s = fstate;
  while ((ab != sab || ac != sac || d != scost || s != sstate) && !k->oom) {
    const uk_cell *c = uk_U(k, ab, ac, d, s);
    int a = c->dist;
    const int nab = c->fab, nac = c->fac, nd = c->fcost, nst = c->fstate;
    int b = a - ab, cc = a - ac;
    const int a1 = uk_U(k, nab, nac, nd, nst)->dist;
    const int b1 = a1 - nab, c1 = a1 - nac;
    while (a > a1 && b > b1 && cc > c1) {
      --a; --b; --cc;
      uk_emit(k, k->A[a], k->B[b], k->C[cc]);
    }
    if (a != a1 || b != b1 || cc != c1) {
      const char x = a > a1 ? k->A[--a] : '-';
      const char y = b > b1 ? k->B[--b] : '-';
      const char z = cc > c1 ? k->C[--cc] : '-';
      uk_emit(k, x, y, z);
    }
    if (nd < 0 || nd > d) { k->oom = 1; return; }   /* a broken chain: give up rather than loop */
    ab = nab; ac = nac; d = nd; s = nst;
  }
}

static int uk_limits(uk_ctx *k, int sab, int sac, int scost, int sstate, int sdist, int fab, int fac, int fcost,
                     int fstate, int fdist);

static int uk_split(uk_ctx *k, int sab, int sac, int scost, int sstate, int sdist, int fab, int fac, int fcost,
                    int fstate, int fdist) {
  const uk_cell *c = uk_U(k, fab, fac, fcost, fstate);
  const int f_ab = c->fab, f_ac = c->fac, f_cost = c->fcost, f_state = c->fstate;
  if (f_cost < 0) { k->oom = 1; return -1; }
  uk_cp *p = uk_CP(k, f_ab, f_ac, f_cost, f_state);
  if (p->cost == 0) p->cost = 1;
  const int cpdist = p->dist;
  const int len = uk_limits(k, f_ab, f_ac, f_cost, f_state, cpdist, fab, fac, fcost, fstate, fdist);
  if (k->oom) return -1;
  uk_limits(k, sab, sac, scost, sstate, sdist, f_ab, f_ac, f_cost, f_state, cpdist);
  return len;
}

static int uk_limits(uk_ctx *k, int sab, int sac, int scost, int sstate, int sdist, int fab, int fac, int fcost,
                     int fstate, int fdist) {
  if (k->oom) return -1;
  k->sab = sab; k->sac = sac; k->scost = scost; k->sstate = sstate;
  k->endA = fdist; k->endB = fdist - fab; k->endC = fdist - fac;
  k->complete_from = 0;
  k->offset += k->final_cost + 1;
  uk_cell *s0 = uk_U(k, sab, sac, scost, sstate);
  s0->dist = sdist;
  s0->tag = scost + k->offset;
  const int guard = fcost + k->ns * k->max_step + 64;   /* the search reaches fdist at fcost */
  if (fcost - scost <= k->cp_width) {
    k->complete_from = 1;
    int i = scost - 1, dist;
    do {
      ++i;
      dist = uk_ukk(k, fab, fac, i, fstate);
    } while (dist < fdist && i < guard && !k->oom);
    if (dist < fdist) { k->oom = 1; return -1; }
    fcost = i;
    uk_trace(k, sab, sac, scost, sstate, fab, fac, fcost, fstate);
    k->complete_from = 0;
    return uk_best(k, fab, fac, fcost, 0);
  }
  k->cp_cost = (fcost + scost - k->cp_width + 1) / 2;
  int i = scost - 1, dist;
  do {
    ++i;
    uk_ukk(k, fab, fac, i, 0);
    dist = uk_ukk(k, fab, fac, i, fstate);
  } while (dist < fdist && i < guard && !k->oom);
  if (dist < fdist) { k->oom = 1; return -1; }
  return uk_split(k, sab, sac, scost, sstate, sdist, fab, fac, i, fstate, fdist);
}

static void uk_model(uk_ctx *k, int mm, int go, int ge) {
  pw_state S[PW_NSTATES];
  int trans[PW_NSTATES][PW_NSTATES];
  k->ns = pw_setup(S, trans, go);
  k->mis = mm; k->go = go; k->ge = ge;
  int max_cost = 0;
  for (int t = 0; t < k->ns; ++t) {
    k->nb[t] = S[t].adv[0] | (S[t].adv[1] << 1) | (S[t].adv[2] << 2);
    const int nm = 3 - S[t].n_del - S[t].n_ins;
    if (S[t].n_ins) { k->cont[t] = ge; k->second[t] = 0; }
    else if (nm == 3) { k->cont[t] = mm; k->second[t] = 1; }
    else if (S[t].n_del == 1) { k->cont[t] = ge; k->second[t] = 1; }
    else { k->cont[t] = 2 * ge; k->second[t] = 0; }
  }
  for (int a = 0; a < k->ns; ++a)
    for (int b = 0; b < k->ns; ++b) {
      k->trans[a][b] = trans[a][b];
      const int nm = 3 - S[b].n_del - S[b].n_ins;
      const int c = trans[a][b] + k->cont[b] + mm * (nm - 1);
      if (c > max_cost) max_cost = c;
    }
  k->max_step = max_cost;
}

/* The reference's own search: returns the cost (or -1: a character without
 * a base bit, memory, or a search that did not close) and, when oa is given,
 * the alignment it prints (masks, gap 16, leading gap first; cap >=
 * la + lb + lc).  planes_mode 0 reuses cost planes as the reference does, 1
 * keeps them all. */
int po_powell_ukk(const uint8_t *A, int la, const uint8_t *B, int lb, const uint8_t *Cs, int lc, int mm, int go,
                  int ge, int planes_mode, uint8_t *oa, uint8_t *ob, uint8_t *oc, int cap, int *olen) {
  if (la < 1 || lb < 1 || lc < 1) return -1;
  uk_ctx k;
  memset(&k, 0, sizeof k);
  uk_model(&k, mm, go, ge);
  const uint8_t *src[3] = {A, B, Cs};
  const int n[3] = {la - 1, lb - 1, lc - 1};
  char *str[3];
  for (int q = 0; q < 3; ++q) {
    str[q] = (char *)calloc((size_t)n[q] + 2, 1);
    if (!str[q]) { for (int r = 0; r < q; ++r) free(str[r]); return -1; }
    for (int i = 0; i < n[q]; ++i) {
      str[q][i] = (char)pw_base(src[q][i + 1]);
      if (!str[q][i]) { for (int r = 0; r <= q; ++r) free(str[r]); return -1; }
    }
  }
  k.A = str[0]; k.B = str[1]; k.C = str[2];
  k.la = n[0]; k.lb = n[1]; k.lc = n[2];
  k.ab_off = k.lb; k.ac_off = k.lc;
  k.ab_blocks = (k.la + k.lb + 1) / UK_BLK + 1;
  k.ac_blocks = (k.la + k.lc + 1) / UK_BLK + 1;
  k.U.elem = sizeof(uk_cell);
  k.U.aliasing = planes_mode == 0;
  k.U.nplanes = 2 * k.max_step;
  k.cp_width = k.max_step;
  k.CP.elem = sizeof(uk_cp);
  k.CP.aliasing = 1;
  k.CP.nplanes = k.cp_width;
  k.out_cap = k.la + k.lb + k.lc + 2;
  k.ra = (char *)malloc((size_t)k.out_cap);
  k.rb = (char *)malloc((size_t)k.out_cap);
  k.rc = (char *)malloc((size_t)k.out_cap);
  k.offset = 1;
  k.furthest = -1;
  int cost = -1;
  if (k.ra && k.rb && k.rc) {
    int start = 0;
    while (start < k.la && k.A[start] == k.B[start] && k.A[start] == k.C[start]) ++start;
    uk_cell *c0 = uk_U(&k, 0, 0, 0, 0);
    c0->dist = start;
    c0->tag = 0 + k.offset;
    const int fab = k.la - k.lb, fac = k.la - k.lc;
    k.endA = k.la; k.endB = k.lb; k.endC = k.lc;
    k.cp_on_dist = 1;
    k.cp_cost = UK_INF;
    int d = 0;
    const int guard = (k.la + k.lb + k.lc + 1) * (k.max_step + 2 * (mm > 0 ? mm : 1)) + 8;
    do {
      ++d;
      uk_ukk(&k, fab, fac, d, 0);
      if (k.cp_on_dist && k.furthest >= k.la / 2) {
        k.cp_cost = d + 1;
        k.cp_on_dist = 0;
      }
    } while (uk_best(&k, fab, fac, d, 0) < k.la && d < guard && !k.oom);
    if (!k.oom && uk_best(&k, fab, fac, d, 0) == k.la) {
      k.cp_on_dist = 0;
      k.final_cost = d;
      const int fstate = uk_best(&k, fab, fac, d, 1);
      if (uk_U(&k, fab, fac, d, fstate)->fcost <= 0)
        uk_limits(&k, 0, 0, 0, 0, start, fab, fac, d, fstate, k.la);
      else
        uk_split(&k, 0, 0, 0, 0, start, fab, fac, d, fstate, k.la);
      if (!k.oom) cost = d;
    }
  }
  if (cost >= 0 && oa) {
    /* the common prefix, then every column in order */
    int pre = 0;
    while (pre < k.la && k.A[pre] == k.B[pre] && k.A[pre] == k.C[pre]) ++pre;
    const int total = pre + k.n;
    if (total + 1 > cap) {
      cost = -1;
    } else {
      uint8_t *o[3] = {oa, ob, oc};
      const char *col[3] = {k.ra, k.rb, k.rc};
      for (int q = 0; q < 3; ++q) {
        o[q][0] = 16;
        for (int i = 0; i < pre; ++i) o[q][1 + i] = (uint8_t)(str[0][i] == 'A' ? 1 : str[0][i] == 'C' ? 2 : str[0][i] == 'G' ? 4 : 8);
        for (int i = 0; i < k.n; ++i) {
          const char ch = col[q][k.n - 1 - i];
          o[q][1 + pre + i] = (uint8_t)(ch == 'A' ? 1 : ch == 'C' ? 2 : ch == 'G' ? 4 : ch == 'T' ? 8 : 16);
        }
      }
      *olen = total + 1;
    }
  }
  uk_free_store(&k, &k.U);
  uk_free_store(&k, &k.CP);
  free(k.ra); free(k.rb); free(k.rc);
  for (int q = 0; q < 3; ++q) free(str[q]);
  return cost;
}
