// Cost-matrix construction of the engine (host side).
//
// POY builds its transformation cost matrices in OCaml
// (Cost_matrix.Two_D, src/cost_matrix.ml) through the cm_CAML_* setters of
// src/cm.c; this file offers the same setters plus the OCaml fills, so that
// a host without the OCaml layer can build the tables the kernels read.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "poyg_internal.h"

namespace poyg {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

static int pos2(const poyg_cm *c, int a, int b) { return (a << c->lcm) + b; }

static std::vector<int> bits_of(int v, int nbits) {
  std::vector<int> out;
  for (int b = 0; b < nbits; ++b)
    if (v & (1 << b)) out.push_back(1 << b);
  return out;
}

// Cost_matrix.Two_D.cleanup, src/cost_matrix.ml:255-265.
static int cleanup(const poyg_cm *c, int item) {
  if (c->combinations && c->cost_model == 1 && item != c->gap && (item & c->gap)) return c->gap;
  return item;
}

// fill_best_cost_and_median_for_all_combinations (src/cost_matrix.ml:323-354):
// the cost of two masks is the best cost(a,v)+cost(v,b) over members a, b
// and single states v; the median is the union of the optimal v.
static void fill_metric(poyg_cm *m, int a_sz) {
  const int ncomb = (1 << a_sz) - 1;
  for (int i = 1; i <= ncomb; ++i) {
    auto li = bits_of(i, a_sz);
    for (int j = 1; j <= ncomb; ++j) {
      auto lj = bits_of(j, a_sz);
      int best = 0, cost = kOcamlMaxInt, worst = 0;
      for (int a : li)
        for (int b : lj) {
          for (int v = 0; v < a_sz; ++v) {
            const int vb = 1 << v;
            const int goa =
                (m->cost_model == 1 && vb == m->gap && (a & m->gap) && (b & m->gap)) ? m->gap_open : 0;
            const int tc = m->cost[pos2(m, a, vb)] + m->cost[pos2(m, vb, b)] + goa;
            if (tc < cost) {
              cost = tc;
              best = vb;
            } else if (tc == cost) {
              best |= vb;
            }
          }
          worst = std::max(worst, m->cost[pos2(m, a, b)]);
        }
      if (li.size() == 1 && lj.size() == 1) {
        m->median[pos2(m, i, j)] = (uint8_t)(i | j);
      } else {
        m->cost[pos2(m, i, j)] = cost;
        m->median[pos2(m, i, j)] = (uint8_t)cleanup(m, best);
      }
      m->worst[pos2(m, i, j)] = worst;
    }
  }
}

// fill_best_cost_and_median_for_all_combinations_bitwise
// (src/cost_matrix.ml:270-321): shared states cost nothing; otherwise the
// cheapest member pair, with the union of the optimal pairs as median.
static void fill_bitwise(poyg_cm *m, int a_sz) {
  const int ncomb = (1 << a_sz) - 1;
  for (int i = 1; i <= ncomb; ++i) {
    auto li = bits_of(i, a_sz);
    for (int j = 1; j <= ncomb; ++j) {
      auto lj = bits_of(j, a_sz);
      if (li.size() == 1 && lj.size() == 1) {
        m->median[pos2(m, i, j)] = (uint8_t)(i | j);
        m->worst[pos2(m, i, j)] = m->cost[pos2(m, i, j)];
        continue;
      }
      // both orders are folded, (lj x li) first, then (li x lj)
      int best = kOcamlMaxInt, med = 0, worst = 0;
      for (int pass = 0; pass < 2; ++pass) {
        const auto &l1 = pass == 0 ? lj : li;
        const auto &l2 = pass == 0 ? li : lj;
        for (int x : l1)
          for (int y : l2) {
            const int v = m->cost[pos2(m, x, y)];
            if (v < best) {
              best = v;
              med = x | y;
            } else if (v == best) {
              med |= x | y;
            }
            worst = std::max(worst, v);
          }
      }
      if (i & j) {
        m->median[pos2(m, i, j)] = (uint8_t)(i & j);
        m->cost[pos2(m, i, j)] = 0;
      } else {
        m->median[pos2(m, i, j)] = (uint8_t)med;
        m->cost[pos2(m, i, j)] = best;
      }
      m->worst[pos2(m, i, j)] = worst;
    }
  }
}

// fill_default_prepend_tail, src/cost_matrix.ml:397-403.
static void fill_prepend_tail(poyg_cm *m) {
  for (int i = 1; i <= m->a_sz; ++i) {
    m->tail[i] = m->cost[pos2(m, i, m->gap)];
    m->prepend[i] = m->cost[pos2(m, m->gap, i)];
  }
}

static bool input_is_metric(const int *l, int w) {
  for (int i = 0; i < w; ++i) {
    if (l[i * w + i] != 0) return false;
    for (int j = 0; j < w; ++j) {
      if (l[i * w + j] < 0 || l[i * w + j] != l[j * w + i]) return false;
      for (int k = 0; k < w; ++k)
        if (l[i * w + j] > l[i * w + k] + l[k * w + j]) return false;
    }
  }
  return true;
}

void to_dev_tcm(const poyg_cm *cm, DevTcm *out) {
  memset(out, 0, sizeof *out);
  for (int a = 0; a < 32; ++a) {
    for (int b = 0; b < 32; ++b) {
      out->cost[a * 32 + b] = cm->cost[(a << cm->lcm) + b];
      out->median[a * 32 + b] = cm->median[(a << cm->lcm) + b];
    }
    out->prepend[a] = cm->prepend[a];
    out->tail[a] = cm->tail[a];
  }
  out->gap = cm->gap;
  out->lcm = cm->lcm;
  out->gap_open = cm->gap_open;
  out->cost_model = cm->cost_model;
  out->combinations = cm->combinations;
}

void to_dev_tcm3(const poyg_cm3 *cm, DevTcm3 *out) {
  memset(out, 0, sizeof *out);
  for (int a = 0; a < 32; ++a)
    for (int b = 0; b < 32; ++b)
      for (int c = 0; c < 32; ++c) {
        const int p = (((a << cm->lcm) + b) << cm->lcm) + c;
        out->cost[(a * 32 + b) * 32 + c] = cm->cost[p];
        out->median[(a * 32 + b) * 32 + c] = cm->median[p];
      }
  out->gap = cm->gap;
  out->lcm = cm->lcm;
  out->a_sz = cm->a_sz;
}

}  // namespace poyg

using namespace poyg;

extern "C" {

const char *poyg_last_error(void) { return poyg::g_err; }
const char *poyg_version(void) { return "poyg 0.1 (gfx950)"; }

// cm_set_val, src/cm.c:318-361.
int poyg_cm_create(int a_sz, int combinations, int cost_model, int gap_open, int all_elements,
                   poyg_cm **out) {
  if (!out || a_sz < 1) {
    set_error("poyg_cm_create: bad arguments");
    return POYG_EINVAL;
  }
  int lcm = 0;
  if (combinations) {
    lcm = a_sz;
  } else {
    int v = a_sz + 1, bl = 0;
    while (v) { ++bl; v >>= 1; }
    lcm = bl + 1;
  }
  if (lcm > 5) {
    set_error("poyg_cm_create: alphabets wider than 5 bits are not supported (lcm=%d)", lcm);
    return POYG_EUNSUPPORTED;
  }
  poyg_cm *c = (poyg_cm *)calloc(1, sizeof(poyg_cm));
  if (!c) {
    set_error("poyg_cm_create: out of memory");
    return POYG_ENOMEM;
  }
  c->lcm = lcm;
  if (combinations) {
    c->gap = 1 << (a_sz - 1);
    c->a_sz = (1 << a_sz) - 1;
    c->combinations = 1;
  } else {
    c->gap = a_sz;
    c->a_sz = a_sz;
  }
  c->all_elements = all_elements;
  c->cost_model = cost_model;
  c->gap_open = gap_open;
  *out = c;
  return POYG_OK;
}

void poyg_cm_destroy(poyg_cm *cm) { free(cm); }

#define CHECK_AB(c, a, b)                                              \
  if (!(c) || (a) < 0 || (b) < 0 || (a) >= (1 << (c)->lcm) * 2 || (b) >= (1 << (c)->lcm)) { \
    set_error("cost matrix index out of range");                       \
    return POYG_EINVAL;                                                \
  }

int poyg_cm_set_cost(poyg_cm *c, int a, int b, int v) { CHECK_AB(c, a, b); c->cost[pos2(c, a, b)] = v; return 0; }
int poyg_cm_set_worst(poyg_cm *c, int a, int b, int v) { CHECK_AB(c, a, b); c->worst[pos2(c, a, b)] = v; return 0; }
int poyg_cm_set_median(poyg_cm *c, int a, int b, int v) { CHECK_AB(c, a, b); c->median[pos2(c, a, b)] = (uint8_t)v; return 0; }
int poyg_cm_set_prepend(poyg_cm *c, int a, int v) { CHECK_AB(c, a, 0); c->prepend[a] = v; return 0; }
int poyg_cm_set_tail(poyg_cm *c, int a, int v) { CHECK_AB(c, a, 0); c->tail[a] = v; return 0; }
int poyg_cm_set_affine_flag(poyg_cm *c, int model, int go) {
  if (!c) return POYG_EINVAL;
  c->cost_model = model;
  c->gap_open = go;
  return 0;
}
int poyg_cm_get_cost(const poyg_cm *c, int a, int b) { return c->cost[pos2(c, a, b)]; }
int poyg_cm_get_median(const poyg_cm *c, int a, int b) { return c->median[pos2(c, a, b)]; }
int poyg_cm_get_worst(const poyg_cm *c, int a, int b) { return c->worst[pos2(c, a, b)]; }
int poyg_cm_get_prepend(const poyg_cm *c, int a) { return c->prepend[a]; }
int poyg_cm_get_tail(const poyg_cm *c, int a) { return c->tail[a]; }
int poyg_cm_get_gap(const poyg_cm *c) { return c->gap; }
int poyg_cm_get_affine(const poyg_cm *c) { return c->cost_model; }
int poyg_cm_get_gap_opening(const poyg_cm *c) { return c->gap_open; }

// Cost_matrix.Two_D.fill_cost_matrix, src/cost_matrix.ml:520-538.
int poyg_cm_of_list(const int *mat, int w, poyg_cm **out) {
  if (!mat || w < 1 || w > 5) {
    set_error("poyg_cm_of_list: base matrix must be 1..5 wide");
    return POYG_EINVAL;
  }
  poyg_cm *m = nullptr;
  int rc = poyg_cm_create(w, 1, 0, 0, (1 << w) - 1, &m);
  if (rc) return rc;
  for (int e1 = 1; e1 <= w; ++e1)
    for (int e2 = 1; e2 <= w; ++e2) m->cost[pos2(m, 1 << (e1 - 1), 1 << (e2 - 1))] = mat[(e1 - 1) * w + e2 - 1];
  if (input_is_metric(mat, w)) {
    m->is_metric = 1;
    fill_metric(m, w);
  } else {
    fill_bitwise(m, w);
  }
  fill_prepend_tail(m);
  *out = m;
  return POYG_OK;
}

// Cost_matrix.Two_D.set_affine, src/cost_matrix.ml:405-416.
int poyg_cm_set_affine(poyg_cm *m, int model, int gap_open) {
  if (!m || model < 0 || model > 2) return POYG_EINVAL;
  m->cost_model = model;
  m->gap_open = model == 1 ? gap_open : 0;
  if (!m->combinations) {
    set_error("poyg_cm_set_affine: non-combination alphabets are not supported");
    return POYG_EUNSUPPORTED;
  }
  fill_bitwise(m, m->lcm);
  fill_prepend_tail(m);
  return POYG_OK;
}

// of_transformations_and_gaps true 5 sub indel 31 (src/cost_matrix.ml:645-656).
int poyg_cm_of_sub_indel(int sub, int indel, int gap_open, poyg_cm **out) {
  int mat[25];
  for (int r = 0; r < 5; ++r)
    for (int c = 0; c < 5; ++c) mat[r * 5 + c] = r == c ? 0 : ((r == 4 || c == 4) ? indel : sub);
  int rc = poyg_cm_of_list(mat, 5, out);
  if (rc == POYG_OK && gap_open >= 0) rc = poyg_cm_set_affine(*out, 1, gap_open);
  return rc;
}

// Three_D.of_two_dim_comb, src/cost_matrix.ml:803-851.
int poyg_cm3_of_two_dim(const poyg_cm *m, poyg_cm3 **out) {
  if (!m || !out || !m->combinations) {
    set_error("poyg_cm3_of_two_dim: combination alphabets only");
    return POYG_EUNSUPPORTED;
  }
  poyg_cm3 *n = (poyg_cm3 *)calloc(1, sizeof(poyg_cm3));
  if (!n) return POYG_ENOMEM;
  n->lcm = m->lcm;
  n->gap = 1 << (m->lcm - 1);
  n->a_sz = (1 << m->lcm) - 1;
  n->combinations = 1;
  n->cost_model = m->cost_model;
  n->gap_open = m->gap_open;
  n->all_elements = m->all_elements;
  const int alph = m->a_sz, lcm = m->lcm;
  for (int i = 1; i <= alph; ++i)
    for (int j = 1; j <= alph; ++j)
      for (int k = 1; k <= alph; ++k) {
        const int p = (((i << lcm) + j) << lcm) + k;
        int cost = kOcamlMaxInt, med = 0;
        for (int l = 0; l < lcm; ++l) {
          const int inter = 1 << l;
          const int shared = !!(i & inter) + !!(j & inter) + !!(k & inter);
          const int c = (m->is_metric || shared >= 2 || inter != m->gap)
                            ? m->cost[pos2(m, inter, i)] + m->cost[pos2(m, inter, j)] + m->cost[pos2(m, inter, k)]
                            : kOcamlMaxInt;
          if (c < cost) {
            cost = c;
            med = inter;
          } else if (c == cost) {
            med |= inter;
          }
        }
        n->cost[p] = cost;
        n->median[p] = (uint8_t)(med & -med);   // pick_bit 1: lowest set bit
      }
  *out = n;
  return POYG_OK;
}

void poyg_cm3_destroy(poyg_cm3 *c) { free(c); }

// ---- custom-block wire format ---------------------------------------------
// The byte stream the OCaml runtime produces for cm_CAML_serialize /
// cm_CAML_serialize_3d (src/cm.c:909-952): caml_serialize_int_4 and
// caml_serialize_block_4 store 4-byte big-endian values, block_1 raw bytes
// (SEQT is unsigned char, src/seq.h:43-45).  A stub writes this buffer with
// caml_serialize_block_1 and reads it back with caml_deserialize_block_1,
// which gives the same bytes on the wire as the reference.
static uint8_t *put_be32(uint8_t *p, int32_t v) {
  const uint32_t u = (uint32_t)v;
  p[0] = (uint8_t)(u >> 24); p[1] = (uint8_t)(u >> 16); p[2] = (uint8_t)(u >> 8); p[3] = (uint8_t)u;
  return p + 4;
}
static const uint8_t *get_be32(const uint8_t *p, int32_t *v) {
  *v = (int32_t)(((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3]);
  return p + 4;
}
static int64_t cm2_len(int lcm) { return 2 * (int64_t(1) << lcm) * (int64_t(1) << lcm); }
static int64_t cm3_len(int lcm) { return (int64_t(1) << (lcm + 1)) * (int64_t(1) << (lcm + 1)) * (int64_t(1) << (lcm + 1)); }

int64_t poyg_cm_serialized_size(const poyg_cm *c) {
  return c ? 8 * 4 + cm2_len(c->lcm) * (4 + 1 + 4 + 4 + 4) : -1;
}

int poyg_cm_serialize(const poyg_cm *c, uint8_t *buf, int64_t cap) {
  if (!c || !buf || cap < poyg_cm_serialized_size(c)) {
    set_error("poyg_cm_serialize: buffer too small");
    return POYG_EINVAL;
  }
  const int64_t len = cm2_len(c->lcm);
  uint8_t *p = buf;
  for (int v : {c->a_sz, c->lcm, c->gap, c->cost_model, c->combinations, c->gap_open, c->is_metric,
                c->all_elements})
    p = put_be32(p, v);
  for (int64_t k = 0; k < len; ++k) p = put_be32(p, c->cost[k]);
  for (int64_t k = 0; k < len; ++k) *p++ = c->median[k];
  for (int64_t k = 0; k < len; ++k) p = put_be32(p, c->worst[k]);
  for (int64_t k = 0; k < len; ++k) p = put_be32(p, c->prepend[k]);
  for (int64_t k = 0; k < len; ++k) p = put_be32(p, c->tail[k]);
  return POYG_OK;
}

// cm_CAML_deserialize (src/cm.c:860-888).
int poyg_cm_deserialize(const uint8_t *buf, int64_t size, poyg_cm **out) {
  if (!buf || !out || size < 8 * 4) {
    set_error("poyg_cm_deserialize: truncated header");
    return POYG_EINVAL;
  }
  int32_t h[8];
  const uint8_t *p = buf;
  for (int k = 0; k < 8; ++k) p = get_be32(p, &h[k]);
  if (h[1] < 1 || h[1] > 5) {
    set_error("poyg_cm_deserialize: lcm %d is not a 5-bit alphabet", h[1]);
    return POYG_EUNSUPPORTED;
  }
  const int64_t len = cm2_len(h[1]);
  if (size < 8 * 4 + len * 17) {
    set_error("poyg_cm_deserialize: truncated tables");
    return POYG_EINVAL;
  }
  poyg_cm *c = (poyg_cm *)calloc(1, sizeof(poyg_cm));
  if (!c) return POYG_ENOMEM;
  c->a_sz = h[0]; c->lcm = h[1]; c->gap = h[2]; c->cost_model = h[3];
  c->combinations = h[4]; c->gap_open = h[5]; c->is_metric = h[6]; c->all_elements = h[7];
  for (int64_t k = 0; k < len; ++k) p = get_be32(p, &c->cost[k]);
  for (int64_t k = 0; k < len; ++k) c->median[k] = *p++;
  for (int64_t k = 0; k < len; ++k) p = get_be32(p, &c->worst[k]);
  for (int64_t k = 0; k < len; ++k) p = get_be32(p, &c->prepend[k]);
  for (int64_t k = 0; k < len; ++k) p = get_be32(p, &c->tail[k]);
  *out = c;
  return POYG_OK;
}

int64_t poyg_cm3_serialized_size(const poyg_cm3 *c) { return c ? 7 * 4 + cm3_len(c->lcm) * (4 + 1) : -1; }

int poyg_cm3_serialize(const poyg_cm3 *c, uint8_t *buf, int64_t cap) {
  if (!c || !buf || cap < poyg_cm3_serialized_size(c)) {
    set_error("poyg_cm3_serialize: buffer too small");
    return POYG_EINVAL;
  }
  const int64_t len = cm3_len(c->lcm);
  uint8_t *p = buf;
  for (int v : {c->a_sz, c->lcm, c->gap, c->cost_model, c->combinations, c->gap_open, c->all_elements})
    p = put_be32(p, v);
  for (int64_t k = 0; k < len; ++k) p = put_be32(p, c->cost[k]);
  for (int64_t k = 0; k < len; ++k) *p++ = c->median[k];
  return POYG_OK;
}

// cm_CAML_deserialize_3d (src/cm.c:890-907).
int poyg_cm3_deserialize(const uint8_t *buf, int64_t size, poyg_cm3 **out) {
  if (!buf || !out || size < 7 * 4) {
    set_error("poyg_cm3_deserialize: truncated header");
    return POYG_EINVAL;
  }
  int32_t h[7];
  const uint8_t *p = buf;
  for (int k = 0; k < 7; ++k) p = get_be32(p, &h[k]);
  if (h[1] < 1 || h[1] > 5) {
    set_error("poyg_cm3_deserialize: lcm %d is not a 5-bit alphabet", h[1]);
    return POYG_EUNSUPPORTED;
  }
  const int64_t len = cm3_len(h[1]);
  if (size < 7 * 4 + len * 5) {
    set_error("poyg_cm3_deserialize: truncated tables");
    return POYG_EINVAL;
  }
  poyg_cm3 *c = (poyg_cm3 *)calloc(1, sizeof(poyg_cm3));
  if (!c) return POYG_ENOMEM;
  c->a_sz = h[0]; c->lcm = h[1]; c->gap = h[2]; c->cost_model = h[3];
  c->combinations = h[4]; c->gap_open = h[5]; c->all_elements = h[6];
  for (int64_t k = 0; k < len; ++k) p = get_be32(p, &c->cost[k]);
  for (int64_t k = 0; k < len; ++k) c->median[k] = *p++;
  *out = c;
  return POYG_OK;
}

// cm_compare (src/cm.c:980-1010): the order of the custom block's compare,
// and the content equality under which device copies may be shared.
int poyg_cm_compare(const poyg_cm *a, const poyg_cm *b) {
  if (a->a_sz != b->a_sz) return a->a_sz - b->a_sz;
  if (a->combinations != b->combinations) return a->combinations - b->combinations;
  if (a->cost_model != b->cost_model) return a->cost_model - b->cost_model;
  if (a->gap_open != b->gap_open) return a->gap_open - b->gap_open;
  if (a->is_metric != b->is_metric) return a->is_metric - b->is_metric;
  const size_t len = (size_t)cm2_len(a->lcm);
  int cmp = memcmp(a->cost, b->cost, len * sizeof(int));
  if (cmp) return cmp;
  cmp = memcmp(a->median, b->median, len);
  if (cmp) return cmp;
  return memcmp(a->worst, b->worst, len * sizeof(int));
}
int poyg_cm3_get_cost(const poyg_cm3 *c, int a, int b, int d) { return c->cost[(((a << c->lcm) + b) << c->lcm) + d]; }
int poyg_cm3_get_median(const poyg_cm3 *c, int a, int b, int d) { return c->median[(((a << c->lcm) + b) << c->lcm) + d]; }

}  // extern "C"
