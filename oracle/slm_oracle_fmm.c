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
 * Second order (scikit-fmm's default, distanceMarcher::updatePointOrderTwo
 * as published): along each axis the smaller frozen neighbour value v1, and,
 * when the next point beyond it in the same direction is frozen with a value
 * not above v1, that value v2 (negated when it lies across the zero contour
 * from the neighbour or from the point), give the term (9/4)(u - t)^2 / dx^2
 * with t = (4 v1 - v2) / 3; otherwise the first-order term (u - v1)^2 / dx^2;
 * u is the larger root of sum of terms = 1.  As published, v2 is not reset
 * when the other direction of the same axis gives a smaller v1 without a
 * usable second point.  A negative discriminant makes scikit-fmm raise; the
 * reference catches the exception and leaves that subsegment's length unset
 * (mapping.py:871-879): here the call reports it.
 *   slmo_fmm_distance2(order 2): the fast march with that update, a point's
 *     value replaced at each update, points accepted in (value, index)
 *     order (heap).
 * Parity with scikit-fmm itself is unpinned (absent here); these are pinned
 * by closed forms in tests/test_fmm.py.  Double precision throughout, as
 * scikit-fmm computes.
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

/* ------------------------------------------------------------ second order */

static int in_grid(int nx, int ny, int x, int y) { return x >= 0 && y >= 0 && x < nx && y < ny; }

/* The second-order update of point (x, y).  frozen(q) says whether point q
 * counts as frozen for the direct neighbours; frozen2(q) for the second
 * points.  Returns the root; *bad = 1 on a negative discriminant. */
typedef int (*FrozenFn)(const void* ctx, int q);

static double update2(int nx, int ny, const int8_t* phi, const uint8_t* dom, const double* u, int x, int y,
                      double dx, FrozenFn frozen, FrozenFn frozen2, const void* ctx, int* bad) {
    const double idx2 = 1.0 / (dx * dx), aa = 9.0 / 4.0, third = 1.0 / 3.0;
    const int i = x * ny + y;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int axis = 0; axis < 2; axis++) {
        double v1 = INFINITY, v2 = INFINITY;
        for (int s = -1; s <= 1; s += 2) {
            const int xx = axis == 0 ? x + s : x, yy = axis == 1 ? y + s : y;
            if (!in_grid(nx, ny, xx, yy)) continue;
            const int n = xx * ny + yy;
            if (!dom[n] || !frozen(ctx, n) || !(u[n] < v1)) continue;
            v1 = u[n];
            const int x2 = axis == 0 ? x + 2 * s : x, y2 = axis == 1 ? y + 2 * s : y;
            if (!in_grid(nx, ny, x2, y2)) continue;
            const int n2 = x2 * ny + y2;
            if (!dom[n2] || !frozen2(ctx, n2) || !(u[n2] <= v1)) continue;
            v2 = u[n2];
            if (phi[n2] * phi[n] < 0 || phi[n2] * phi[i] < 0) v2 = -v2;
        }
        if (v2 < INFINITY) {
            const double tp = third * (4.0 * v1 - v2);
            a += idx2 * aa;
            b -= idx2 * 2.0 * aa * tp;
            c += idx2 * aa * tp * tp;
        } else if (v1 < INFINITY) {
            a += idx2;
            b -= idx2 * 2.0 * v1;
            c += idx2 * v1 * v1;
        }
    }
    c -= 1.0;
    const double det = b * b - 4.0 * a * c;
    if (det < 0.0) {
        *bad = 1;
        return INFINITY;
    }
    return (-b + sqrt(det)) / 2.0 / a;
}

static int frozen_state(const void* ctx, int q) { return ((const uint8_t*)ctx)[q] >= 2; }

/* order 1: slmo_fmm_distance; order 2: scikit-fmm's march with the update
 * above.  Returns 0, -1 on allocation failure, -2 on a negative
 * discriminant (out is then incomplete). */
int slmo_fmm_distance2(int nx, int ny, const int8_t* phi, const uint8_t* dom, double dx, int order, double* out) {
    if (order != 2) return slmo_fmm_distance(nx, ny, phi, dom, dx, out);
    const int n = nx * ny;
    uint8_t* state = (uint8_t*)calloc((size_t)n, 1);
    double* u = (double*)malloc((size_t)n * sizeof(double));
    HeapItem* heap = (HeapItem*)malloc((size_t)(5 * n + 4) * sizeof(HeapItem));
    if (!state || !u || !heap) {
        free(state); free(u); free(heap);
        return -1;
    }
    int hn = 0, bad = 0;
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
    for (int i = 0; i < n; i++) {
        if (state[i] != 2) continue;
        HeapItem it = {u[i], i};
        heap_push(heap, &hn, it);
    }
    while (hn > 0 && !bad) {
        const HeapItem it = heap_pop(heap, &hn);
        const int i = it.i;
        if (state[i] == 3 || it.v != u[i]) continue;
        state[i] = 3;
        const int x = i / ny, y = i % ny;
        const int nb[4][2] = {{x - 1, y}, {x + 1, y}, {x, y - 1}, {x, y + 1}};
        for (int k = 0; k < 4; k++) {
            const int xx = nb[k][0], yy = nb[k][1];
            if (!in_grid(nx, ny, xx, yy)) continue;
            const int j = xx * ny + yy;
            if (!dom[j] || state[j] >= 2) continue;
            const double v = update2(nx, ny, phi, dom, u, xx, yy, dx, frozen_state, frozen_state, state, &bad);
            if (bad) break;
            if (v != 0.0) {                    /* scikit-fmm skips a zero update */
                u[j] = v;                      /* replaced, not min-ed */
                state[j] = 1;
                HeapItem nt = {v, j};
                heap_push(heap, &hn, nt);
            }
        }
    }
    for (int i = 0; i < n; i++) out[i] = dom[i] ? (phi[i] < 0 ? -u[i] : u[i]) : NAN;
    free(state); free(u); free(heap);
    return bad ? -2 : 0;
}
