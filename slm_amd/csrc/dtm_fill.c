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

int slm_v_valley_segments(double* out, int64_t nx, int64_t ny, int n, const double* cx, const double* cy,
                          const double* s, const double* c, double depth, double hw, double half_len);

/* V-shaped incisions for the synthetic DTM generator: for each valley k,
 * out[x][y] -= depth * clip(1 - |(x - cx)*s - (y - cy)*c| / hw, 0, 1), valleys
 * applied in order.  Only the band of columns where the distance can be
 * below hw is visited (outside it the term is exactly zero), so the result
 * has the bits of the same expression evaluated over the whole grid.
 * out: float64 [nx*ny] C-order (x major), accumulated in place. */
int slm_v_valleys(double* out, int64_t nx, int64_t ny, int n, const double* cx, const double* cy,
                  const double* s, const double* c, double depth, double hw) {
    return slm_v_valley_segments(out, nx, ny, n, cx, cy, s, c, depth, hw, 0.0);
}

/* As slm_v_valleys, for segments: with half_len > 0 each valley also fades
 * out along its axis, its depth times clip((half_len - |a|) / hw, 0, 1)
 * where a = (x - cx)*c + (y - cy)*s is the position along the valley. */
int slm_v_valley_segments(double* out, int64_t nx, int64_t ny, int n, const double* cx, const double* cy,
                          const double* s, const double* c, double depth, double hw, double half_len) {
    if (!out || nx <= 0 || ny <= 0 || hw <= 0.0) return -1;
    for (int k = 0; k < n; k++) {
        for (int64_t x = 0; x < nx; x++) {
            const double a = ((double)x - cx[k]) * s[k];
            int64_t y0 = 0, y1 = ny - 1;
            if (fabs(c[k]) > 1e-9) {
                const double ya = cy[k] + (a - hw) / c[k], yb = cy[k] + (a + hw) / c[k];
                const double lo = floor(fmin(ya, yb)) - 2.0, hi = ceil(fmax(ya, yb)) + 2.0;
                if (hi < 0.0 || lo > (double)(ny - 1)) continue;
                y0 = lo > 0.0 ? (int64_t)lo : 0;
                y1 = hi < (double)(ny - 1) ? (int64_t)hi : ny - 1;
            }
            double* row = out + x * ny;
            for (int64_t y = y0; y <= y1; y++) {
                const double d = fabs(a - ((double)y - cy[k]) * c[k]);
                double t = 1.0 - d / hw;
                t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
                if (half_len > 0.0) {
                    const double along = fabs(((double)x - cx[k]) * c[k] + ((double)y - cy[k]) * s[k]);
                    double f = (half_len - along) / hw;
                    f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
                    t *= f;
                }
                row[y] -= depth * t;
            }
        }
    }
    return 0;
}
