/*
 * dtm_fill.c -- priority-flood sink filling with an epsilon gradient, used by
 * the synthetic DTM generator (slm_amd/synth.py).  Host code, no GPU.
 *
 * Priority-flood (Barnes, Lehman & Mulla 2014, "Priority-flood: an optimal
 * depression-filling and watershed-labeling algorithm", algorithm 2 with the
 * epsilon variant): the grid edge and every cell next to a no-data cell seed a
 * min-heap; cells are popped lowest first and each unvisited 8-neighbour is
 * raised to at least (popped elevation + eps) before being pushed.  Every
 * filled cell therefore drains monotonically to the edge.
 */
#include <stdint.h>
#include <stdlib.h>
#include <math.h>

typedef struct { double h; int64_t i; } Node;

static void heap_push(Node* hp, int64_t* n, Node v) {
    int64_t k = (*n)++;
    while (k > 0) {
        int64_t p = (k - 1) / 2;
        if (hp[p].h < v.h || (hp[p].h == v.h && hp[p].i < v.i)) break;
        hp[k] = hp[p];
        k = p;
    }
    hp[k] = v;
}

static Node heap_pop(Node* hp, int64_t* n) {
    Node top = hp[0], last = hp[--(*n)];
    int64_t k = 0;
    for (;;) {
        int64_t c = 2 * k + 1;
        if (c >= *n) break;
        if (c + 1 < *n && (hp[c + 1].h < hp[c].h || (hp[c + 1].h == hp[c].h && hp[c + 1].i < hp[c].i))) c++;
        if (last.h < hp[c].h || (last.h == hp[c].h && last.i < hp[c].i)) break;
        hp[k] = hp[c];
        k = c;
    }
    hp[k] = last;
    return top;
}

/* h: float64 [nx*ny] C-order (x major), modified in place; NaN = no data.
 * Returns 0 on success, -1 on allocation failure. */
int slm_fill_sinks(double* h, int64_t nx, int64_t ny, double eps) {
    const int64_t n = nx * ny;
    uint8_t* seen = (uint8_t*)calloc((size_t)n, 1);
    Node* hp = (Node*)malloc(sizeof(Node) * (size_t)(n > 0 ? n : 1));
    if (!seen || !hp) { free(seen); free(hp); return -1; }
    int64_t nh = 0;
    for (int64_t x = 0; x < nx; x++)
        for (int64_t y = 0; y < ny; y++) {
            const int64_t i = x * ny + y;
            if (isnan(h[i])) { seen[i] = 1; continue; }
            int edge = (x == 0 || y == 0 || x == nx - 1 || y == ny - 1);
            for (int dx = -1; dx <= 1 && !edge; dx++)
                for (int dy = -1; dy <= 1 && !edge; dy++) {
                    int64_t xx = x + dx, yy = y + dy;
                    if (xx >= 0 && yy >= 0 && xx < nx && yy < ny && isnan(h[xx * ny + yy])) edge = 1;
                }
            if (edge) { seen[i] = 1; Node v = {h[i], i}; heap_push(hp, &nh, v); }
        }
    while (nh > 0) {
        Node c = heap_pop(hp, &nh);
        const int64_t x = c.i / ny, y = c.i % ny;
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++) {
                if (!dx && !dy) continue;
                int64_t xx = x + dx, yy = y + dy;
                if (xx < 0 || yy < 0 || xx >= nx || yy >= ny) continue;
                int64_t j = xx * ny + yy;
                if (seen[j]) continue;
                seen[j] = 1;
                if (h[j] < c.h + eps) h[j] = c.h + eps;
                Node v = {h[j], j};
                heap_push(hp, &nh, v);
            }
    }
    free(seen);
    free(hp);
    return 0;
}
