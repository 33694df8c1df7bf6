/*
 * slm_oracle_fmm.c -- CPU oracle (test infrastructure only) of the eikonal
 * distance behind the reference's FMM hillslope lengths
 * (python/streamlines/mapping.py:841-900, Mapping.fmm_lengths, which calls
 * scikit-fmm's skfmm.distance(phi, dx) once per subsegment).
 *
 * scikit-fmm is not installed here.  This is a restatement of the published
 * fast marching method (Sethian) for a binary phi (+1 / -1) on a masked grid:
 *   - initialisation as scikit-fmm does it: a point with a sign change to an
 *     unmasked 4-neighbour lies next to the zero contour; along each axis
 *     the contour sits at the linear-interpolation fraction
 *     phi_p / (phi_p - phi_q) of a cell (1/2 for a binary phi); the point's
 *     distance is 1 / sqrt(sum over crossing axes of 1 / d_axis^2);
 *   - first-order upwind marching: u solves sum over axes of
 *     max(u - m_axis, 0)^2 = dx^2 with m_axis the smaller frozen neighbour
 *     along that axis, i.e. a + dx when b - a >= dx (a <= b), otherwise
 *     (a + b + sqrt(2 dx^2 - (a - b)^2)) / 2;
 *   - the result carries the sign of phi; masked points are NaN.
 * scikit-fmm's default is second order, so parity with the reference is
 * unpinned; this oracle is pinned by closed forms in tests/test_fmm.py.
 * Double precision throughout, as scikit-fmm computes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double v; int32_t i; } HeapItem;

static void heap_push(HeapItem* h, int* n, HeapItem it) {
    int k = (*n)++;
    while (k > 0) {
        int p = (k - 1) / 2;
        if (h[p].v < it.v || (h[p].v == it.v && h[p].i <= it.i)) break;
        h[k] = h[p];
        k = p;
    }
    h[k] = it;
}

static HeapItem heap_pop(HeapItem* h, int* n) {
    HeapItem top = h[0], last = h[--(*n)];
    int k = 0;
    for (;;) {
        int c = 2 * k + 1;
        if (c >= *n) break;
        if (c + 1 < *n && (h[c + 1].v < h[c].v || (h[c + 1].v == h[c].v && h[c + 1].i < h[c].i))) c++;
        if (last.v < h[c].v || (last.v == h[c].v && last.i <= h[c].i)) break;
        h[k] = h[c];
        k = c;
    }
    h[k] = last;
    return top;
}

/* the first-order update from the smaller frozen neighbour of each axis */
double slmo_fmm_update(double a, double b, double dx) {
    if (a > b) { double t = a; a = b; b = t; }
    if (isinf(b) || b - a >= dx) return a + dx;
    const double d = a - b;
    return (a + b + sqrt(2.0 * dx * dx - d * d)) * 0.5;
}

/* the initial distance of a point next to the zero contour, or -1 */
double slmo_fmm_init(int nx, int ny, const int8_t* phi, const uint8_t* dom, int x, int y, double dx) {
    const int i = x * ny + y;
    double inv = 0.0;
    int any = 0;
    for (int axis = 0; axis < 2; axis++) {
        double best = INFINITY;
        for (int s = -1; s <= 1; s += 2) {
            const int xx = axis == 0 ? x + s : x, yy = axis == 1 ? y + s : y;
            if (xx < 0 || yy < 0 || xx >= nx || yy >= ny) continue;
            const int j = xx * ny + yy;
            if (!dom[j] || phi[j] == phi[i]) continue;
            const double f = (double)phi[i] / ((double)phi[i] - (double)phi[j]) * dx;
            if (f < best) best = f;
        }
        if (best < INFINITY) {
            inv += 1.0 / (best * best);
            any = 1;
        }
    }
    return any ? 1.0 / sqrt(inv) : -1.0;
}

/* phi: +1 / -1 int8 [nx*ny]; dom: 1 where unmasked; out: signed distance,
 * NaN outside dom.  Returns 0, or -1 on allocation failure. */
int slmo_fmm_distance(int nx, int ny, const int8_t* phi, const uint8_t* dom, double dx, double* out) {
    const int n = nx * ny;
    uint8_t* state = (uint8_t*)calloc((size_t)n, 1);   /* 0 far, 1 trial, 2 initial, 3 accepted */
    double* u = (double*)malloc((size_t)n * sizeof(double));
    HeapItem* heap = (HeapItem*)malloc((size_t)(5 * n + 4) * sizeof(HeapItem));   /* <= 4 pushes a point + seeds */
    if (!state || !u || !heap) {
        free(state); free(u); free(heap);
        return -1;
    }
    int hn = 0;
    for (int i = 0; i < n; i++) u[i] = INFINITY;
    for (int x = 0; x < nx; x++)
        for (int y = 0; y < ny; y++) {
            const int i = x * ny + y;
            if (!dom[i]) continue;
            const double d = slmo_fmm_init(nx, ny, phi, dom, x, y, dx);
            if (d >= 0.0) {
                u[i] = d;
                state[i] = 2;
            }
        }
    /* the initial band seeds the march */
    for (int i = 0; i < n; i++) {
        if (state[i] != 2) continue;
        HeapItem it = {u[i], i};
        heap_push(heap, &hn, it);
    }
    /* initial points (state 2) are known; each popped point is accepted
       (state 3) and updates its unknown neighbours from the known ones */
    while (hn > 0) {
        const HeapItem it = heap_pop(heap, &hn);
        const int i = it.i;
        if (state[i] == 3 || it.v != u[i]) continue;               /* done, or a stale entry */
        state[i] = 3;
        const int x = i / ny, y = i % ny;
        const int nb[4][2] = {{x - 1, y}, {x + 1, y}, {x, y - 1}, {x, y + 1}};
        for (int k = 0; k < 4; k++) {
            const int xx = nb[k][0], yy = nb[k][1];
            if (xx < 0 || yy < 0 || xx >= nx || yy >= ny) continue;
            const int j = xx * ny + yy;
            if (!dom[j] || state[j] >= 2) continue;
            double m[2] = {INFINITY, INFINITY};
            for (int axis = 0; axis < 2; axis++)
                for (int s = -1; s <= 1; s += 2) {
                    const int ax = axis == 0 ? xx + s : xx, ay = axis == 1 ? yy + s : yy;
                    if (ax < 0 || ay < 0 || ax >= nx || ay >= ny) continue;
                    const int q = ax * ny + ay;
                    if (dom[q] && state[q] >= 2 && u[q] < m[axis]) m[axis] = u[q];
                }
            const double v = slmo_fmm_update(m[0], m[1], dx);
            if (v < u[j]) {
                u[j] = v;
                state[j] = 1;
                HeapItem nt = {v, j};
                heap_push(heap, &hn, nt);
            }
        }
    }
    for (int i = 0; i < n; i++) out[i] = dom[i] ? (phi[i] < 0 ? -u[i] : u[i]) : NAN;
    free(state); free(u); free(heap);
    return 0;
}
