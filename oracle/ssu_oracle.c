/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C, FP64, CPU restatement of Strided State UniFrac and Faith's PD as
 * biocore/unifrac-binaries computes them.  It exists to check the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it, and only as the checker (or as the timed CPU baseline).  The
 * product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning (what
 * checks each function independently of this file) is listed per function;
 * the tests in tests/test_oracle_*.py carry those checks.
 *
 * Build: oracle/Makefile -> oracle/liboracle.so (gcc -O3 -fopenmp).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Balanced-parentheses tree                                           */
/* ------------------------------------------------------------------ */

typedef struct {
    int n;            /* number of parentheses */
    uint8_t *s;       /* 1 = open, 0 = close */
    double *len;      /* per open position */
    char **name;      /* per open position (malloc'd strings, "" if none) */
    int *oc;          /* open<->close partner */
} orc_tree;

static void tree_free(orc_tree *t) {
    if (!t) return;
    if (t->name) { for (int i = 0; i < t->n; i++) free(t->name[i]); free(t->name); }
    free(t->s); free(t->len); free(t->oc);
    memset(t, 0, sizeof(*t));
}

/* open/close partner table; src/tree.cpp:365-380 (structure_to_openclose) */
static int tree_index(orc_tree *t) {
    t->oc = (int *)malloc(sizeof(int) * (t->n ? t->n : 1));
    int *stk = (int *)malloc(sizeof(int) * (t->n ? t->n : 1));
    int sp = 0;
    for (int i = 0; i < t->n; i++) {
        if (t->s[i]) stk[sp++] = i;
        else {
            if (sp == 0) { free(stk); return -1; }
            int o = stk[--sp];
            t->oc[i] = o; t->oc[o] = i;
        }
    }
    free(stk);
    return sp == 0 ? 0 : -1;
}

static inline int t_open(const orc_tree *t, int i) { return t->s[i] ? i : t->oc[i]; }
static inline int t_close(const orc_tree *t, int i) { return t->s[i] ? t->oc[i] : i; }
/* src/tree.cpp:264-266 */
static inline int t_isleaf(const orc_tree *t, int i) { return t->s[i] && (i + 1 < t->n) && !t->s[i + 1]; }

static char *xstrdup(const char *a) { size_t l = strlen(a); char *r = (char *)malloc(l + 1); memcpy(r, a, l + 1); return r; }

static int is_struct_char(char c) { return c == '(' || c == ')' || c == ',' || c == ';'; }

/*
 * Newick -> parentheses, src/tree.cpp:316-363 (newick_to_bp).  Quoted text is
 * skipped; "(x)" and ",)" produce a tip "10"; a ")" after ")" closes a node.
 */
static int newick_structure(const char *nw, uint8_t **out, int *n_out) {
    size_t cap = 1024, n = 0;
    uint8_t *s = (uint8_t *)malloc(cap);
#define PUSH(v) do { if (n == cap) { cap *= 2; s = (uint8_t *)realloc(s, cap); } s[n++] = (v); } while (0)
    char last = '\0';
    int psd = 0, inq = 0;
    for (const char *p = nw; *p; p++) {
        char c = *p;
        if (c == '\'') inq = !inq;
        if (inq) continue;
        if (c == '(') { PUSH(1); last = c; psd = 1; }
        else if (c == ')') {
            if (psd || last == ',') { PUSH(1); PUSH(0); PUSH(0); psd = 0; }
            else PUSH(0);
            last = c;
        } else if (c == ',') {
            if (last != ')') { PUSH(1); PUSH(0); }
            psd = 0; last = c;
        }
    }
#undef PUSH
    *out = s; *n_out = (int)n;
    return 0;
}

/*
 * One newick token, src/tree.cpp:455-490 (tokenize).  Newlines are dropped,
 * quotes are stripped, and a structure character that ends a non-empty token
 * is consumed with it (the metadata pass below compensates for that).
 */
static char *next_token(const char **start, const char *end) {
    size_t cap = 64, n = 0;
    char *tok = (char *)malloc(cap);
    int inq = 0;
    do {
        char c = **start; (*start)++;
        if (c == '\n') continue;
        int isq = (c == '\'');
        if (inq && isq) { inq = 0; continue; }
        else if (!inq && isq) { inq = 1; continue; }
        if (is_struct_char(c) && !inq) {
            if (n == 0) tok[n++] = c;
            break;
        }
        if (n + 1 >= cap) { cap *= 2; tok = (char *)realloc(tok, cap); }
        tok[n++] = c;
    } while (*start != end);
    tok[n] = 0;
    return tok;
}

/* src/tree.cpp:432-448 (set_node_metadata): name before the last ':', length after it, parsed as float */
static void set_meta(orc_tree *t, int open_idx, const char *tok) {
    const char *colon = strrchr(tok, ':');
    double length = 0.0;
    free(t->name[open_idx]);
    if (colon == tok) {
        length = (double)strtof(tok + 1, NULL);
        t->name[open_idx] = xstrdup("");
    } else if (colon) {
        size_t l = (size_t)(colon - tok);
        char *nm = (char *)malloc(l + 1); memcpy(nm, tok, l); nm[l] = 0;
        t->name[open_idx] = nm;
        length = (double)strtof(colon + 1, NULL);
    } else {
        t->name[open_idx] = xstrdup(tok);
    }
    t->len[open_idx] = length;
}

/* src/tree.cpp:384-425 (newick_to_metadata) */
static void newick_metadata(orc_tree *t, const char *nw) {
    const char *start = nw, *end = nw + strlen(nw);
    while (end != start && (end[-1] == ' ' || end[-1] == '\t' || end[-1] == '\n' ||
                            end[-1] == '\r' || end[-1] == '\f' || end[-1] == '\v')) end--;
    char last = '\0';
    int idx = 0, lag = 0;
    while (start != end) {
        char *tok = next_token(&start, end);
        if (strlen(tok) == 1 && is_struct_char(tok[0])) {
            if (tok[0] == '(') idx++;
            else if (tok[0] == ')' || tok[0] == ',') { idx++; if (last == ')') lag++; }
        } else {
            idx += lag; lag = 0;
            if (idx < t->n) {
                int o = t_open(t, idx);
                set_meta(t, o, tok);
                idx += t_isleaf(t, o) ? 2 : 1;
            }
        }
        last = tok[0];
        free(tok);
    }
}

static int tree_from_newick(const char *nw, orc_tree *t) {
    memset(t, 0, sizeof(*t));
    newick_structure(nw, &t->s, &t->n);
    t->len = (double *)calloc(t->n ? t->n : 1, sizeof(double));
    t->name = (char **)malloc(sizeof(char *) * (t->n ? t->n : 1));
    for (int i = 0; i < t->n; i++) t->name[i] = xstrdup("");
    if (tree_index(t)) return -1;
    newick_metadata(t, nw);
    return 0;
}

static int tree_from_arrays(int n, const uint8_t *s, const double *len, const char *const *names, orc_tree *t) {
    memset(t, 0, sizeof(*t));
    t->n = n;
    t->s = (uint8_t *)malloc(n ? n : 1); memcpy(t->s, s, n);
    t->len = (double *)malloc(sizeof(double) * (n ? n : 1)); memcpy(t->len, len, sizeof(double) * n);
    t->name = (char **)malloc(sizeof(char *) * (n ? n : 1));
    for (int i = 0; i < n; i++) t->name[i] = xstrdup(names[i] ? names[i] : "");
    return tree_index(t);
}

/* src/tree.cpp:126-143 (mask): keep flagged parentheses, lengths from new_len */
static int tree_mask(const orc_tree *t, const uint8_t *keep, const double *new_len, orc_tree *o) {
    memset(o, 0, sizeof(*o));
    int c = 0;
    for (int i = 0; i < t->n; i++) c += keep[i] ? 1 : 0;
    o->n = c;
    o->s = (uint8_t *)malloc(c ? c : 1);
    o->len = (double *)malloc(sizeof(double) * (c ? c : 1));
    o->name = (char **)malloc(sizeof(char *) * (c ? c : 1));
    int j = 0;
    for (int i = 0; i < t->n; i++) if (keep[i]) {
        o->s[j] = t->s[i]; o->len[j] = new_len[i]; o->name[j] = xstrdup(t->name[i]); j++;
    }
    return tree_index(o);
}

static int cmp_str(const void *a, const void *b) { return strcmp(*(const char *const *)a, *(const char *const *)b); }

/* src/tree.cpp:145-163 (shear): keep named tips in the set and every ancestor */
static int tree_shear(const orc_tree *t, char **keep_sorted, int n_keep, orc_tree *o) {
    uint8_t *mask = (uint8_t *)calloc(t->n ? t->n : 1, 1);
    int *parent = (int *)malloc(sizeof(int) * (t->n ? t->n : 1));
    int *stk = (int *)malloc(sizeof(int) * (t->n ? t->n : 1));
    int sp = 0;
    for (int i = 0; i < t->n; i++) {
        if (t->s[i]) { parent[i] = sp ? stk[sp - 1] : -1; stk[sp++] = i; }
        else sp--;
    }
    for (int i = 0; i < t->n; i++) {
        if (t_isleaf(t, i)) {
            const char *nm = t->name[i];
            if (bsearch(&nm, keep_sorted, n_keep, sizeof(char *), cmp_str)) {
                mask[i] = 1; mask[i + 1] = 1;
                int p = parent[i];
                while (p != -1 && !mask[p]) { mask[p] = 1; mask[t_close(t, p)] = 1; p = parent[p]; }
            }
        }
    }
    int r = tree_mask(t, mask, t->len, o);
    free(mask); free(parent); free(stk);
    return r;
}

/* src/tree.cpp:165-191 (collapse): fold single-child nodes into their child, preorder */
static int tree_collapse(const orc_tree *t, orc_tree *o) {
    uint8_t *mask = (uint8_t *)calloc(t->n ? t->n : 1, 1);
    double *nl = (double *)malloc(sizeof(double) * (t->n ? t->n : 1));
    memcpy(nl, t->len, sizeof(double) * t->n);
    for (int cur = 0; cur < t->n; cur++) {
        if (!t->s[cur]) continue;          /* preorder = open positions in order */
        if (t_isleaf(t, cur) || cur == 0) { mask[cur] = 1; mask[t_close(t, cur)] = 1; }
        else {
            int first = cur + 1;
            int last = t_open(t, t_close(t, cur) - 1);
            if (first == last) nl[first] = nl[first] + nl[cur];
            else { mask[cur] = 1; mask[t_close(t, cur)] = 1; }
        }
    }
    int r = tree_mask(t, mask, nl, o);
    free(mask); free(nl);
    return r;
}

/* ------------------------------------------------------------------ */
/* Postorder node list                                                 */
/* ------------------------------------------------------------------ */

typedef struct {
    int m;            /* non-root nodes in postorder (src/unifrac_cmp.cpp:100: nparens/2 - 1) */
    int *parent;      /* postorder index of parent, m for the root */
    double *len;
    int *is_leaf;
    int *obs;         /* table row of a leaf, -1 otherwise */
} orc_post;

static void post_free(orc_post *p) { free(p->parent); free(p->len); free(p->is_leaf); free(p->obs); memset(p, 0, sizeof(*p)); }

/*
 * Postorder = order of closing parentheses; src/tree.cpp:248-250
 * (postorderselect = open(select_0(k))).  Children of a node precede it and
 * appear left to right, which is the order src/unifrac_internal.cpp:240-251
 * sums them in.
 */
static int tree_postorder(const orc_tree *t, char **obs_sorted, const int *obs_row_sorted, int n_obs, orc_post *p) {
    int total = t->n / 2;
    p->m = total > 0 ? total - 1 : 0;
    int m = p->m;
    p->parent = (int *)malloc(sizeof(int) * (total ? total : 1));
    p->len = (double *)malloc(sizeof(double) * (total ? total : 1));
    p->is_leaf = (int *)malloc(sizeof(int) * (total ? total : 1));
    p->obs = (int *)malloc(sizeof(int) * (total ? total : 1));
    int *post_of_open = (int *)malloc(sizeof(int) * (t->n ? t->n : 1));
    int k = 0;
    for (int i = 0; i < t->n; i++) if (!t->s[i]) post_of_open[t->oc[i]] = k++;
    int *stk = (int *)malloc(sizeof(int) * (t->n ? t->n : 1));
    int sp = 0;
    for (int i = 0; i < t->n; i++) {
        if (t->s[i]) { stk[sp++] = i; continue; }
        int o = stk[--sp];
        int kk = post_of_open[o];
        if (kk >= total) continue;
        p->parent[kk] = sp ? post_of_open[stk[sp - 1]] : m;
        p->len[kk] = t->len[o];
        p->is_leaf[kk] = t_isleaf(t, o);
        p->obs[kk] = -1;
        if (p->is_leaf[kk]) {
            const char *nm = t->name[o];
            char **hit = (char **)bsearch(&nm, obs_sorted, n_obs, sizeof(char *), cmp_str);
            if (hit) p->obs[kk] = obs_row_sorted[hit - obs_sorted];
        }
    }
    free(stk); free(post_of_open);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Table                                                               */
/* ------------------------------------------------------------------ */

/* src/biom_inmem.cpp:513-526 (compute_sample_counts): per-sample totals in observation order */
static void sample_counts(int n_obs, int n_samples, const uint32_t *indptr, const uint32_t *indices,
                          const double *data, double *out) {
    for (int s = 0; s < n_samples; s++) out[s] = 0.0;
    for (int i = 0; i < n_obs; i++)
        for (uint32_t j = indptr[i]; j < indptr[i + 1]; j++) out[indices[j]] += data[j];
}

/*
 * Node embedding in FP64; src/unifrac_internal.cpp:230-258 (set_proportions_range)
 * with src/biom_inmem.cpp:473-497 (get_obs_data_range): a leaf holds
 * count/total (or the raw count when normalize is off); an internal node
 * holds the running sum of its children, left to right, starting from 0.
 * Output is node-major: emb[k * n + s].
 */
static void embed(const orc_post *p, int n_samples, const uint32_t *indptr, const uint32_t *indices,
                  const double *data, const double *totals, int normalize, double *emb) {
    const size_t n = (size_t)n_samples;
    for (size_t i = 0; i < (size_t)p->m * n; i++) emb[i] = 0.0;
    for (int k = 0; k < p->m; k++) {
        double *row = emb + (size_t)k * n;
        if (p->is_leaf[k]) {
            int r = p->obs[k];
            if (r >= 0)
                for (uint32_t j = indptr[r]; j < indptr[r + 1]; j++)
                    row[indices[j]] = normalize ? data[j] / totals[indices[j]] : data[j];
        }
        /* children were completed before k (postorder); add k into its parent */
        int par = p->parent[k];
        if (par < p->m) {
            double *prow = emb + (size_t)par * n;
            for (size_t s = 0; s < n; s++) prow[s] += row[s];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Per-pair accumulation                                               */
/* ------------------------------------------------------------------ */

enum { M_UNW = 0, M_UNW_UNNORM = 1, M_W_NORM = 2, M_W_UNNORM = 3, M_GEN = 4 };

/*
 * Distance of one pair (k, l) over all processed nodes, FP64.
 *  unweighted:              src/unifrac_task_impl.hpp:1193-1265 (xor / or of presence, times length)
 *  unweighted_unnormalized: src/unifrac_task_impl.hpp:1592-1753 (xor only)
 *  weighted_normalized:     src/unifrac_task_impl.hpp:506-551   (|u-v| l, total (u+v) l)
 *  weighted_unnormalized:   src/unifrac_task_impl.hpp:151-194   (|u-v| l)
 *  generalized:             src/unifrac_task_impl.hpp:925-1040  ((u+v)^a l |u-v|/(u+v), total (u+v)^a l)
 *  VAW variants:            src/unifrac_task_impl.hpp:436-502, 849-922, 1043-1119, 1756-1932
 *                           (m = c_k + c_l, mi = node counts, vaw = sqrt(mi (m-mi)), skip vaw == 0)
 *  final ratio:             src/unifrac_task_impl.hpp:35-46 (dm / tot)
 */
static double pair_distance(int method, int vaw, double alpha, int bypass, const orc_post *p,
                            const double *u, const double *c, const double *totals, size_t n,
                            size_t kk, size_t ll) {
    double dm = 0.0, tot = 0.0;
    const double m = vaw ? totals[kk] + totals[ll] : 0.0;
    for (int e = 0; e < p->m; e++) {
        if (bypass && p->is_leaf[e]) continue;                /* src/unifrac_cmp.cpp:123-124 */
        const double L = p->len[e];
        const double a = u[(size_t)e * n + kk], b = u[(size_t)e * n + ll];
        double v = 1.0;
        if (vaw) {
            const double mi = c[(size_t)e * n + kk] + c[(size_t)e * n + ll];
            v = sqrt(mi * (m - mi));
            if (!(v > 0)) continue;
        }
        switch (method) {
        case M_UNW: {
            const int pa = a > 0, pb = b > 0;
            const double lv = vaw ? L / v : L;
            dm += (pa ^ pb) * lv; tot += (pa | pb) * lv;
            break; }
        case M_UNW_UNNORM: {
            const int pa = a > 0, pb = b > 0;
            dm += (pa ^ pb) * (vaw ? L / v : L);
            break; }
        case M_W_NORM:
            if (vaw) { dm += (fabs(a - b) * L) / v; tot += ((a + b) * L) / v; }
            else { dm += fabs(a - b) * L; tot += (a + b) * L; }
            break;
        case M_W_UNNORM:
            dm += vaw ? (fabs(a - b) * L) / v : fabs(a - b) * L;
            break;
        case M_GEN:
            if (vaw) {
                const double s1 = (a + b) / v, d1 = fabs(a - b) / v;
                const double sp = pow(s1, alpha) * L;
                dm += sp * (d1 / s1); tot += sp;
            } else {
                const double s1 = a + b;
                if (s1 != 0.0) {
                    const double sp = pow(s1, alpha) * L;
                    dm += sp * (fabs(a - b) / s1); tot += sp;
                }
            }
            break;
        }
    }
    if (method == M_UNW || method == M_W_NORM || method == M_GEN) return dm / tot;
    return dm;
}

/* ------------------------------------------------------------------ */
/* Exported entry points (ctypes)                                      */
/* ------------------------------------------------------------------ */

typedef struct {
    orc_tree tree;    /* sheared + collapsed */
    orc_post post;
} orc_prepared;

static int prepare(int n_parens, const uint8_t *structure, const double *lengths, const char *const *names,
                   int n_obs, const char *const *obs_ids, orc_prepared *out) {
    orc_tree t, sh;
    if (tree_from_arrays(n_parens, structure, lengths, names, &t)) return -1;
    char **keep = (char **)malloc(sizeof(char *) * (n_obs ? n_obs : 1));
    int *rows = (int *)malloc(sizeof(int) * (n_obs ? n_obs : 1));
    for (int i = 0; i < n_obs; i++) keep[i] = (char *)obs_ids[i];
    qsort(keep, n_obs, sizeof(char *), cmp_str);
    /* row of each sorted id */
    for (int i = 0; i < n_obs; i++)
        for (int r = 0; r < n_obs; r++) if (obs_ids[r] == keep[i]) { rows[i] = r; break; }
    /* src/api.cpp:90-92 (SYNC_TREE_TABLE): shear to the table, then collapse */
    tree_shear(&t, keep, n_obs, &sh);
    tree_collapse(&sh, &out->tree);
    tree_postorder(&out->tree, keep, rows, n_obs, &out->post);
    tree_free(&t); tree_free(&sh);
    free(keep); free(rows);
    return 0;
}

/* newick -> parentheses arrays (for comparing tree handling) */
int orc_newick_to_bp(const char *newick, int *n_parens, uint8_t **structure, double **lengths, char ***names) {
    orc_tree t;
    if (tree_from_newick(newick, &t)) { tree_free(&t); return -1; }
    *n_parens = t.n; *structure = t.s; *lengths = t.len; *names = t.name;
    free(t.oc);
    return 0;
}

void orc_free_bp(int n_parens, uint8_t *structure, double *lengths, char **names) {
    if (names) { for (int i = 0; i < n_parens; i++) free(names[i]); free(names); }
    free(structure); free(lengths);
}

/*
 * Sheared + collapsed postorder arrays, for comparing host tree handling.
 * Arrays must hold n_parens/2 entries.  Returns the number of non-root nodes.
 */
int orc_postorder(int n_parens, const uint8_t *structure, const double *lengths, const char *const *names,
                  int n_obs, const char *const *obs_ids,
                  int *parent, double *len, int *is_leaf, int *obs_row) {
    orc_prepared pr;
    if (prepare(n_parens, structure, lengths, names, n_obs, obs_ids, &pr)) return -1;
    int m = pr.post.m;
    for (int k = 0; k < m; k++) {
        parent[k] = pr.post.parent[k]; len[k] = pr.post.len[k];
        is_leaf[k] = pr.post.is_leaf[k]; obs_row[k] = pr.post.obs[k];
    }
    post_free(&pr.post); tree_free(&pr.tree);
    return m;
}

/*
 * UniFrac stripes [stripe_start, stripe_stop) in FP64; stripe s, element k
 * holds the pair (k, (k+s+1) mod n) (src/unifrac_task_impl.hpp:367).
 * out has (stripe_stop - stripe_start) * n_samples doubles.
 */
int orc_unifrac_stripes(int n_parens, const uint8_t *structure, const double *lengths, const char *const *names,
                        int n_obs, int n_samples, const char *const *obs_ids,
                        const uint32_t *indptr, const uint32_t *indices, const double *data,
                        int method, int variance_adjust, double alpha, int bypass_tips, int normalize,
                        int n_threads, unsigned int stripe_start, unsigned int stripe_stop, double *out) {
    orc_prepared pr;
    if (prepare(n_parens, structure, lengths, names, n_obs, obs_ids, &pr)) return -1;
    const size_t n = (size_t)n_samples;
    double *totals = (double *)malloc(sizeof(double) * n);
    sample_counts(n_obs, n_samples, indptr, indices, data, totals);
    double *u = (double *)malloc(sizeof(double) * (size_t)(pr.post.m ? pr.post.m : 1) * n);
    embed(&pr.post, n_samples, indptr, indices, data, totals, normalize, u);
    double *c = NULL;
    if (variance_adjust) {
        /* counts are propagated without normalisation: src/unifrac_cmp.cpp:250-251 */
        c = (double *)malloc(sizeof(double) * (size_t)(pr.post.m ? pr.post.m : 1) * n);
        embed(&pr.post, n_samples, indptr, indices, data, totals, 0, c);
    }
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    const long n_str = (long)stripe_stop - (long)stripe_start;
#pragma omp parallel for collapse(2) schedule(dynamic, 64)
    for (long si = 0; si < n_str; si++)
        for (long k = 0; k < (long)n; k++) {
            const size_t s = (size_t)stripe_start + (size_t)si;
            const size_t l = ((size_t)k + s + 1) % n;
            out[(size_t)si * n + (size_t)k] =
                pair_distance(method, variance_adjust, alpha, bypass_tips, &pr.post, u, c, totals, n, (size_t)k, l);
        }
    free(u); free(c); free(totals);
    post_free(&pr.post); tree_free(&pr.tree);
    return 0;
}

/*
 * Faith's PD, src/unifrac.cpp:362-387: for every non-root node in postorder,
 * result[s] += (prop > 0) * length, in FP64 and in that order.  Samples are
 * expected pre-filtered (src/api.cpp:578 drops samples with total < 1).
 */
int orc_faith_pd(int n_parens, const uint8_t *structure, const double *lengths, const char *const *names,
                 int n_obs, int n_samples, const char *const *obs_ids,
                 const uint32_t *indptr, const uint32_t *indices, const double *data, double *out) {
    orc_prepared pr;
    if (prepare(n_parens, structure, lengths, names, n_obs, obs_ids, &pr)) return -1;
    const size_t n = (size_t)n_samples;
    double *totals = (double *)malloc(sizeof(double) * n);
    sample_counts(n_obs, n_samples, indptr, indices, data, totals);
    double *u = (double *)malloc(sizeof(double) * (size_t)(pr.post.m ? pr.post.m : 1) * n);
    embed(&pr.post, n_samples, indptr, indices, data, totals, 1, u);
    for (size_t s = 0; s < n; s++) out[s] = 0.0;
    for (int k = 0; k < pr.post.m; k++)
        for (size_t s = 0; s < n; s++) out[s] += (u[(size_t)k * n + s] > 0) * pr.post.len[k];
    free(u); free(totals);
    post_free(&pr.post); tree_free(&pr.tree);
    return 0;
}

/*
 * Stripes -> full symmetric n x n matrix, zero diagonal.  Follows the pair
 * map of src/unifrac.cpp:228-340 (stripes_to_matrix_T): stripe s holds
 * (k, (k+s+1) mod n), so every unordered pair is covered by the first
 * (n+1)/2 stripes.
 */
void orc_stripes_to_matrix(const double *stripes, unsigned int n_samples, unsigned int n_stripes, double *mat) {
    const size_t n = n_samples;
    for (size_t i = 0; i < n; i++) mat[i * n + i] = 0.0;
    for (size_t s = 0; s < n_stripes; s++)
        for (size_t k = 0; k < n; k++) {
            const size_t l = (k + s + 1) % n;
            const double v = stripes[s * n + k];
            mat[k * n + l] = v; mat[l * n + k] = v;
        }
}

/* src/unifrac.cpp:99-120 (stripes_to_condensed_form): scipy squareform upper-triangle order */
void orc_stripes_to_condensed(const double *stripes, unsigned int n_samples, unsigned int start, unsigned int stop, double *cf) {
    const uint64_t n = n_samples;
    const uint64_t combN = n * (n - 1) / 2;
    for (uint64_t s = start; s < stop; s++) {
        uint64_t i = 0, j = s + 1;
        for (uint64_t k = 0; k < n; k++, i++, j++) {
            if (j == n) { i = 0; j = n - (s + 1); }
            const uint64_t ni = n - i;
            cf[combN - ni * (ni - 1) / 2 + (j - i - 1)] = stripes[(s - start) * n + k];
        }
    }
}
