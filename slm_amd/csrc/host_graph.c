/*
 * host_graph.c -- host-side graph helpers of libslm_host.so.
 *
 * slm_greedy_colour colours the subsegments of the FMM hillslope lengths
 * (slm_fmm_hsl): subsegments joined by a pair (their dilated domains can
 * share a pixel) get different colours, so the domains of one colour are
 * disjoint and can be solved in one pass.  Greedy in the order of
 * decreasing degree (ties by label), smallest free colour.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { int64_t deg; uint32_t label; } Order;

static int cmp_order(const void* a, const void* b) {
    const Order* x = (const Order*)a;
    const Order* y = (const Order*)b;
    if (x->deg != y->deg) return x->deg > y->deg ? -1 : 1;
    return x->label < y->label ? -1 : (x->label > y->label);
}

/* colour: uint8 [n_labels], 0xff on entry for labels that are not
 * subsegments (they stay 0xff), anything else for subsegments.  pairs:
 * n_pairs unique (lo << 32 | hi) with lo, hi < n_labels.  Returns the number
 * of colours, or -1 (allocation failure) / -2 (more than 254 colours). */
int slm_greedy_colour(uint32_t n_labels, uint8_t* colour, const uint64_t* pairs, int64_t n_pairs) {
    int64_t* start = (int64_t*)calloc((size_t)n_labels + 1, sizeof(int64_t));
    uint32_t* nbr = (uint32_t*)malloc((size_t)(2 * n_pairs + 1) * sizeof(uint32_t));
    if (!start || !nbr) {
        free(start); free(nbr);
        return -1;
    }
    for (int64_t k = 0; k < n_pairs; k++) {
        start[(uint32_t)(pairs[k] >> 32) + 1]++;
        start[(uint32_t)pairs[k] + 1]++;
    }
    for (uint32_t a = 0; a < n_labels; a++) start[a + 1] += start[a];
    int64_t* fill = (int64_t*)malloc((size_t)(n_labels + 1) * sizeof(int64_t));
    int64_t n_prob = 0;
    for (uint32_t a = 0; a < n_labels; a++) n_prob += colour[a] != 0xff;
    Order* ord = (Order*)malloc((size_t)(n_prob + 1) * sizeof(Order));
    if (!fill || !ord) {
        free(start); free(nbr); free(fill); free(ord);
        return -1;
    }
    memcpy(fill, start, (size_t)(n_labels + 1) * sizeof(int64_t));
    for (int64_t k = 0; k < n_pairs; k++) {
        const uint32_t lo = (uint32_t)(pairs[k] >> 32), hi = (uint32_t)pairs[k];
        nbr[fill[lo]++] = hi;
        nbr[fill[hi]++] = lo;
    }
    int64_t m = 0;
    for (uint32_t a = 0; a < n_labels; a++)
        if (colour[a] != 0xff) {
            ord[m].deg = start[a + 1] - start[a];
            ord[m].label = a;
            m++;
        }
    qsort(ord, (size_t)m, sizeof(Order), cmp_order);
    for (int64_t k = 0; k < m; k++) colour[ord[k].label] = 0xfe;   /* uncoloured subsegment */
    int n_col = 0;
    for (int64_t k = 0; k < m; k++) {
        const uint32_t a = ord[k].label;
        uint8_t used[256];
        memset(used, 0, sizeof(used));
        for (int64_t e = start[a]; e < start[a + 1]; e++) used[colour[nbr[e]]] = 1;
        int c = 0;
        while (used[c]) c++;
        if (c >= 0xfe) {
            free(start); free(nbr); free(fill); free(ord);
            return -2;
        }
        colour[a] = (uint8_t)c;
        if (c + 1 > n_col) n_col = c + 1;
    }
    free(start); free(nbr); free(fill); free(ord);
    return n_col;
}
