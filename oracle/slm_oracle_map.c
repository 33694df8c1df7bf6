/*
 * slm_oracle_map.c -- CPU restatement of slm's channel-mapping and
 * segmentation kernels (FP32 trace arithmetic of oracle_math.h).
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle for the HIP mapping kernels.
 * tests/ call it as the checker; nothing in the product loads it.
 *
 * Seeds are the float2 pixel lists of pick_seeds (useful.py:519-538), seed k
 * for k in [seeds_chunk_offset, n_seed_points).  Grids are [nx_padded][ny_padded].
 *
 * Kernels whose reference result does not depend on the work-item schedule
 * are restated as a plain loop over the seeds:
 *   connect_channels       connect.cl:26-108 (+ rungekutta.cl:280-309)
 *   map_channel_heads      channelheads.cl:32-93 (+ rungekutta.cl:335-353)
 *   label_confluences      label.cl:21-134
 *   link_hillslopes        hillslopes.cl:31-91 (+ rungekutta.cl:382-406)
 *   segment_hillslopes     segment.cl:133-188 (+ rungekutta.cl:435-459)
 *   subsegment_channel_edges segment.cl:195-334
 *   subsegment_flanks      segment.cl:360-414
 * The others race in the reference (values read while other work items write
 * them); each is restated with a fixed, schedule-free semantics, described at
 * the function, that equals the reference's result whenever that result does
 * not depend on the schedule:
 *   count_downchannels     countlink.cl:34-114
 *   flag_downchannels      countlink.cl:134-210
 *   segment_downchannels   segment.cl:38-107
 *   fix_right_flanks       segment.cl:441-504
 *   fix_left_flanks        segment.cl:530-594
 * thin_channels is a stand-in for scikit-image's skeletonize(medial_axis(.))
 * (mapping.py:707-719): Zhang & Suen's two-subiteration parallel thinning
 * (Comm. ACM 27(3), 1984).
 */
#include <stdlib.h>
#include "oracle_math.h"

#define F_CHANNEL SLM_IS_CHANNEL
#define F_THIN SLM_IS_THINCHANNEL
#define F_INTER SLM_IS_INTERCHANNEL
#define F_HEAD SLM_IS_CHANNELHEAD
#define F_TAIL SLM_IS_CHANNELTAIL
#define F_MAJCONF SLM_IS_MAJORCONFLUENCE
#define F_MAJIN SLM_IS_MAJORINFLOW
#define F_MININ SLM_IS_MINORINFLOW
#define F_LEFT SLM_IS_LEFTFLANK
#define F_SSH SLM_IS_SUBSEGMENTHEAD
#define LFA SLM_LEFT_FLANK_ADDITION

/* link[i], with a link outside the grid read as a self link (end of the
 * chain); the reference would read outside its arrays there. */
static inline uint32_t lk(const SlmParams* p, const uint32_t* link, uint32_t i) {
    const uint32_t j = link[i];
    return j < p->nxy_padded ? j : i;
}

/*
 * Cycle fast-forward.  A trace (or link walk) whose state repeats is periodic
 * and would run on to its step cap.  Every shared-grid update in the loops
 * below is a max of a counter that grows by the same amount each lap, an or,
 * or an and, so all laps but the last two are dominated: on detecting a cycle
 * (Brent's algorithm on the whole state) the loop skips them, advancing the
 * step and counter by whole laps, and runs the rest normally.  The result is
 * the same bits as running every step; slmo_map_set_fast_forward(0) runs every
 * step (tests compare the two).
 */
static int g_fast_forward = 1;
void slmo_map_set_fast_forward(int on) { g_fast_forward = on; }

static inline uint32_t seed_pixel(const SlmParams* p, const float* seeds, uint32_t k) {
    return array_idx(p, (real)seeds[2 * (size_t)k], (real)seeds[2 * (size_t)k + 1]);
}

#define KEY(hi, lo) (((uint64_t)(hi) << 32) | (uint64_t)(uint32_t)(lo))
/* Shared-grid updates.  Every update here is a max, or or and: the result
 * does not depend on the order of the seeds, which is what lets the loops
 * over seeds run in parallel (OpenMP) and still give one answer. */
static inline void kmax(uint64_t* a, uint64_t v) {
    uint64_t cur = __atomic_load_n(a, __ATOMIC_RELAXED);
    while (v > cur && !__atomic_compare_exchange_n(a, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
}
static inline void umax(uint32_t* a, uint32_t v) {
    uint32_t cur = __atomic_load_n(a, __ATOMIC_RELAXED);
    while (v > cur && !__atomic_compare_exchange_n(a, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
}
static inline void aor(uint32_t* a, uint32_t v) { __atomic_fetch_or(a, v, __ATOMIC_RELAXED); }
static inline void aand(uint32_t* a, uint32_t v) { __atomic_fetch_and(a, v, __ATOMIC_RELAXED); }

/*
 * The common body of the mapping RK2 steps (rungekutta.cl:280-459): given the
 * compute_step_vec result st at position vec, apply the dxy1 retry for a
 * short step, move vec, and return the step length dl.  The caller decides
 * what a stuck step (dl < INTEGRATION_HALT_THRESHOLD) means and whether dt
 * is updated.
 */
static inline real map_move(const SlmParams* p, const Step* st, real vec[2], uint32_t* idx) {
    const int c = (int)p->fp_contract;
    real nx = st->next[0], ny = st->next[1];
    real dl = fast_length(c, st->dxy2[0], st->dxy2[1]);
    if (dl < (real)p->integration_halt_threshold) {
        nx = vec[0] + st->dxy1[0];
        ny = vec[1] + st->dxy1[1];
        dl = fast_length(c, st->dxy1[0], st->dxy1[1]);
    }
    vec[0] = nx;
    vec[1] = ny;
    *idx = array_idx(p, nx, ny);
    return dl;
}

/* ------------------------------------------------------------ connect_channels */

/*
 * connect_channels (connect.cl:26-108).  Trace from a channel pixel until the
 * trace leaves the pixel, then until it lands on a channel pixel; record each
 * step compress(vec-prev_vec) (prev_vec is not advanced by a stuck step,
 * rungekutta.cl:291-303); if 2 < n_steps < INTERCHANNEL_MAX_N_STEPS, replay
 * the record from the seed and flag every visited pixel whose mapping is 0
 * as IS_INTERCHANNEL.  The reference's private record holds
 * INTERCHANNEL_MAX_N_STEPS entries but its first loop may write more (no
 * bound); such a trace ends with n_steps >= INTERCHANNEL_MAX_N_STEPS and flags
 * nothing, so here the first loop also stops there.
 */
void slmo_connect_channels(const SlmParams* p, const float* seeds, const uint8_t* mask,
                           const float* uv, uint32_t* mapping) {
    const uint32_t cap = p->interchannel_max_n_steps;
    int8_t* rec = (int8_t*)malloc(2 * (size_t)(cap ? cap : 1));
    for (uint32_t k = p->seeds_chunk_offset; k < p->n_seed_points; k++) {
        real vec[2] = {(real)seeds[2 * (size_t)k], (real)seeds[2 * (size_t)k + 1]};
        real prev[2] = {vec[0], vec[1]};
        real dt = (real)p->dt_max;
        uint32_t n = 0, idx = array_idx(p, vec[0], vec[1]);
        const uint32_t seed_idx = idx;
        Step st;
        for (int loop = 0; loop < 2; loop++) {
            for (;;) {
                if (loop == 0) {
                    if (!(seed_idx == idx && !mask[idx] && n != p->max_n_steps && n < cap)) break;
                } else {
                    if (!((mapping[idx] & F_CHANNEL) == 0 && !mask[idx] && n < cap)) break;
                }
                compute_step(p, uv, dt, vec, NULL, 0, &st);
                idx = st.idx;
                if (mask[idx]) continue;
                const real err = step_error(p, &st);
                const real dl = map_move(p, &st, vec, &idx);
                const int stuck = dl < (real)p->integration_halt_threshold;
                if (!stuck) dt = next_dt(p, dt, err);
                const real res = (real)p->trajectory_resolution;
                rec[2 * n] = cvt_i8((vec[0] - prev[0]) * res);
                rec[2 * n + 1] = cvt_i8((vec[1] - prev[1]) * res);
                n++;
                if (!stuck) { prev[0] = vec[0]; prev[1] = vec[1]; }
            }
        }
        if (n > 2 && n < cap) {
            real v[2] = {(real)seeds[2 * (size_t)k], (real)seeds[2 * (size_t)k + 1]};
            const real res = (real)p->trajectory_resolution;
            idx = array_idx(p, v[0], v[1]);
            for (uint32_t s = 0; !mask[idx] && s < n - 1; s++) {
                if (mapping[idx] == 0u) mapping[idx] |= F_INTER;
                v[0] = v[0] + (real)rec[2 * s] / res;
                v[1] = v[1] + (real)rec[2 * s + 1] / res;
                idx = array_idx(p, v[0], v[1]);
            }
        }
    }
    free(rec);
}

/* ----------------------------------------------------------------- thinning */

/*
 * Zhang-Suen thinning of B = (mapping has IS_CHANNEL or IS_INTERCHANNEL);
 * pixels left in B are flagged IS_THINCHANNEL.  Neighbours in the paper's
 * order P2..P9 = N, NE, E, SE, S, SW, W, NW with N = row x-1; outside the
 * grid counts as 0.  Returns the number of full iterations.
 */
static inline int zs_get(const uint8_t* b, uint32_t nx, uint32_t ny, int64_t x, int64_t y) {
    if (x < 0 || y < 0 || x >= (int64_t)nx || y >= (int64_t)ny) return 0;
    return b[(size_t)x * ny + (size_t)y];
}

uint32_t slmo_thin_channels(const SlmParams* p, uint32_t* mapping) {
    const uint32_t nx = p->nx_padded, ny = p->ny_padded;
    const size_t nxy = (size_t)nx * ny;
    uint8_t* a = (uint8_t*)malloc(nxy);
    uint8_t* b = (uint8_t*)malloc(nxy);
    for (size_t i = 0; i < nxy; i++) a[i] = (mapping[i] & (F_CHANNEL | F_INTER)) ? 1 : 0;
    uint32_t iters = 0;
    for (;;) {
        int changed = 0;
        for (int sub = 0; sub < 2; sub++) {
            for (int64_t x = 0; x < nx; x++)
                for (int64_t y = 0; y < ny; y++) {
                    const size_t i = (size_t)x * ny + (size_t)y;
                    b[i] = a[i];
                    if (!a[i]) continue;
                    const int P[8] = {zs_get(a, nx, ny, x - 1, y),     zs_get(a, nx, ny, x - 1, y + 1),
                                      zs_get(a, nx, ny, x, y + 1),     zs_get(a, nx, ny, x + 1, y + 1),
                                      zs_get(a, nx, ny, x + 1, y),     zs_get(a, nx, ny, x + 1, y - 1),
                                      zs_get(a, nx, ny, x, y - 1),     zs_get(a, nx, ny, x - 1, y - 1)};
                    int B = 0, A = 0;
                    for (int j = 0; j < 8; j++) {
                        B += P[j];
                        A += (P[j] == 0 && P[(j + 1) & 7] == 1);
                    }
                    /* P2=P[0] P4=P[2] P6=P[4] P8=P[6] */
                    const int c1 = sub == 0 ? P[0] * P[2] * P[4] : P[0] * P[2] * P[6];
                    const int c2 = sub == 0 ? P[2] * P[4] * P[6] : P[0] * P[4] * P[6];
                    if (B >= 2 && B <= 6 && A == 1 && c1 == 0 && c2 == 0) {
                        b[i] = 0;
                        changed = 1;
                    }
                }
            uint8_t* t = a; a = b; b = t;
        }
        iters++;
        if (!changed) break;
    }
    for (size_t i = 0; i < nxy; i++)
        if (a[i]) mapping[i] |= F_THIN;
    free(a);
    free(b);
    return iters;
}

/* ---------------------------------------------------------- map_channel_heads */

/*
 * map_channel_heads (channelheads.cl:32-93): from every unmasked pixel, step
 * (channelheads_runge_kutta_step, rungekutta.cl:335-353: no stuck exit, dt
 * always updated) until the trace leaves its pixel; at the arrival pixel
 * clear IS_CHANNELHEAD unless it is thin channel reached from a non-thin
 * pixel.  Only IS_CHANNELHEAD is written and it is never read: schedule-free.
 */
void slmo_map_channel_heads(const SlmParams* p, const float* seeds, const uint8_t* mask,
                            const float* uv, uint32_t* mapping) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t kk = p->seeds_chunk_offset; kk < p->n_seed_points; kk++) {
        const uint32_t k = (uint32_t)kk;
        real vec[2] = {(real)seeds[2 * (size_t)k], (real)seeds[2 * (size_t)k + 1]};
        real dt = (real)p->dt_max;
        uint32_t n = 0, idx = array_idx(p, vec[0], vec[1]);
        const uint32_t prev_idx = idx;
        Step st;
        while (prev_idx == idx && !mask[idx] && n < p->max_n_steps) {
            compute_step(p, uv, dt, vec, NULL, 0, &st);
            idx = st.idx;
            const real err = step_error(p, &st);
            map_move(p, &st, vec, &idx);
            n++;
            dt = next_dt(p, dt, err);
        }
        if (n >= p->max_n_steps) continue;
        if (!mask[idx]) {
            if (((~mapping[idx]) & F_THIN) || (mapping[prev_idx] & F_THIN)) aand(&mapping[idx], ~F_HEAD);
        }
    }
}

/* --------------------------------------------------------- count_downchannels */

/*
 * countlink_runge_kutta_step (rungekutta.cl:382-406): returns 1 when stuck
 * (position moved, dt kept), else updates dt.
 */
static inline int countlink_step(const SlmParams* p, const float* uv, real* dt, real vec[2],
                                 uint32_t* idx) {
    Step st;
    compute_step(p, uv, *dt, vec, NULL, 0, &st);
    *idx = st.idx;
    const real err = step_error(p, &st);
    const real dl = map_move(p, &st, vec, idx);
    if (dl < (real)p->integration_halt_threshold) return 1;
    *dt = next_dt(p, *dt, err);
    return 0;
}



/*
 * count_downchannels (countlink.cl:34-114).  A trace from each channel head
 * numbers the new pixels it enters (head 1, then 3, 4, ... as the
 * reference's counter runs), flags them thin channel and not head, links
 * each pixel to the next, and stops on entering a pixel whose count exceeds
 * its own ("stepped onto a more important channel").  The reference reads
 * that count while other traces write it, and links are last-writer-wins.
 * Here:
 *   pass A  every trace runs to its end without the stop test;
 *           keyA[pixel] = max (counter, seed) over all traces;
 *   pass B  every trace runs again and stops on entering pixel q with
 *           counter < counter(keyA[q]) unless keyA[q] is its own entry;
 *           it ANDs/ORs the flags as the reference, count[q] = max(count[q],
 *           counter), and offers link keys (counter at prev, next pixel);
 *   link[p] = next pixel of the largest link key at p (largest counter,
 *           then largest next index); pixels never left keep link[p].
 * count and link hold their prior contents where untouched (the reference
 * host zeroes both first, mapping.py:730-731).
 */
void slmo_count_downchannels(const SlmParams* p, const float* seeds, const uint8_t* mask,
                             const float* uv, uint32_t* mapping, uint32_t* count, uint32_t* link) {
    const size_t nxy = p->nxy_padded;
    const uint32_t cap = 1000u * p->max_n_steps; /* uint arithmetic, as the macro product */
    uint64_t* keyA = (uint64_t*)calloc(nxy, sizeof(uint64_t));
    uint64_t* keyL = (uint64_t*)calloc(nxy, sizeof(uint64_t));
    for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t kk = p->seeds_chunk_offset; kk < p->n_seed_points; kk++) {
            const uint32_t k = (uint32_t)kk;
            real vec[2] = {(real)seeds[2 * (size_t)k], (real)seeds[2 * (size_t)k + 1]};
            real dt = (real)p->dt_max;
            uint32_t idx = array_idx(p, vec[0], vec[1]), prev_idx = idx;
            uint32_t n = 0, counter = 1, cprev = 1;
            if (pass == 0) kmax(&keyA[idx], KEY(1, k));
            else umax(&count[idx], 1);
            counter = 2;
            /* Brent state: position, dt, previous pixel */
            real sx = vec[0], sy = vec[1], sdt = dt;
            uint32_t sprev = prev_idx, sn = n, scounter = counter, power = 1;
            int ff = g_fast_forward;
            while (!mask[idx] && n < cap) {
                if (countlink_step(p, uv, &dt, vec, &idx)) break;
                n++;
                if (prev_idx != idx) {
                    ++counter;
                    if (pass == 0) {
                        kmax(&keyA[idx], KEY(counter, k));
                    } else {
                        aand(&mapping[idx], ~F_HEAD);
                        aor(&mapping[idx], F_THIN);
                        kmax(&keyL[prev_idx], KEY(cprev, idx));
                        const uint64_t a = keyA[idx];
                        if (counter < (uint32_t)(a >> 32) && (uint32_t)a != k) break;
                        umax(&count[idx], counter);
                        cprev = counter;
                    }
                    prev_idx = idx;
                }
                if (ff) {
                    if (vec[0] == sx && vec[1] == sy && dt == sdt && prev_idx == sprev) {
                        const uint32_t lam = n - sn, clap = counter - scounter, laps = (cap - n) / lam;
                        if (laps >= 2) {
                            n += (laps - 1) * lam;
                            counter += (laps - 1) * clap;
                            cprev += (laps - 1) * clap;
                        }
                        ff = 0;
                    } else if (n - sn == power) {
                        sx = vec[0]; sy = vec[1]; sdt = dt; sprev = prev_idx;
                        sn = n; scounter = counter; power <<= 1;
                    }
                }
            }
            if (pass == 1) kmax(&keyL[prev_idx], KEY(cprev, idx));
        }
    }
    for (size_t i = 0; i < nxy; i++)
        if (keyL[i]) link[i] = (uint32_t)keyL[i];
    free(keyA);
    free(keyL);
}

/* ---------------------------------------------------------- flag_downchannels */

/*
 * flag_downchannels (countlink.cl:134-210).  Each channel head sets count 1
 * at itself and walks link: counter += 1 per hop; unmasked pixels are flagged
 * thin channel and count = max(count, counter); the last unmasked pixel is
 * flagged IS_CHANNELTAIL.  The head loses IS_CHANNELHEAD if within the first
 * 18 hops it meets an IS_MINORINFLOW pixel, or its first hop lands on a pixel
 * of count > 2.  The reference reads count mid-race; here the phases are
 * ordered: (1) every head's count = 1, (2) every walk, (3) the head test on
 * the final counts.
 */
void slmo_flag_downchannels(const SlmParams* p, const float* seeds, const uint8_t* mask,
                            uint32_t* mapping, uint32_t* count, const uint32_t* link) {
    const uint32_t cap = 1000u * p->max_n_steps;
    for (uint32_t k = p->seeds_chunk_offset; k < p->n_seed_points; k++) {
        const uint32_t h = seed_pixel(p, seeds, k);
        count[h] = 1;
        mapping[h] |= F_THIN;
    }
    for (int phase = 2; phase <= 3; phase++) {
        for (uint32_t k = p->seeds_chunk_offset; k < p->n_seed_points; k++) {
            const uint32_t h = seed_pixel(p, seeds, k);
            uint32_t idx = h, prev_idx = h + 1u, counter = 1;
            uint32_t sidx = idx, scounter = counter, power = 1;
            int ff = g_fast_forward && phase == 2;
            while (!mask[idx] && prev_idx != idx && counter < cap) {
                prev_idx = idx;
                idx = lk(p, link, idx);
                counter++;
                if (!mask[idx]) {
                    if (phase == 3) {
                        if (counter >= 20) break;
                        if ((mapping[idx] & F_MININ) || (counter == 2 && count[idx] > 2u)) mapping[h] &= ~F_HEAD;
                    } else {
                        mapping[idx] |= F_THIN;
                        umax(&count[idx], counter);
                    }
                } else {
                    break;
                }
                if (ff) {   /* a walk's state is its pixel; one hop per counter */
                    if (idx == sidx && counter != scounter) {
                        const uint32_t lam = counter - scounter, laps = (cap - counter) / lam;
                        if (laps >= 2) counter += (laps - 1) * lam;
                        ff = 0;
                    } else if (counter - scounter == power) {
                        sidx = idx; scounter = counter; power <<= 1;
                    }
                }
            }
            if (phase == 2 && prev_idx < p->nxy_padded && !mask[prev_idx] && counter < cap)
                mapping[prev_idx] |= F_TAIL;
        }
    }
}

/* ---------------------------------------------------------- label_confluences */

/* label_confluences (label.cl:21-134): neighbours N,S,E,W,NE,SE,NW,SW that are
 * unmasked thin channel and link here are inflows; with more than one, flag
 * here a major confluence and each inflow minor, except among the inflows
 * with count one less than here ("equal dominant") the first of largest slt,
 * flagged major.  Writes only flags that it never reads: schedule-free. */
void slmo_label_confluences(const SlmParams* p, const float* seeds, const uint8_t* mask,
                            const float* slt, uint32_t* mapping, const uint32_t* count,
                            const uint32_t* link) {
    static const int D[8][2] = {{0, 1}, {0, -1}, {1, 0}, {-1, 0}, {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    for (uint32_t k = p->seeds_chunk_offset; k < p->n_seed_points; k++) {
        const real vx = (real)seeds[2 * (size_t)k], vy = (real)seeds[2 * (size_t)k + 1];
        const uint32_t idx = array_idx(p, vx, vy);
        uint32_t in[8], eq[8], n_in = 0, n_eq = 0, dom_idx = 0;
        float dom = -3.402823466e+38f;
        for (int d = 0; d < 8; d++) {
            const uint32_t nb = array_idx(p, vx + (real)D[d][0], vy + (real)D[d][1]);
            if (!mask[nb] && (mapping[nb] & F_THIN) && link[nb] == idx) in[n_in++] = nb;
        }
        if (n_in > 1) {
            mapping[idx] |= F_MAJCONF;
            for (uint32_t i = 0; i < n_in; i++) {
                if (count[idx] != count[in[i]] + 1u) {
                    mapping[in[i]] |= F_MININ;
                } else {
                    eq[n_eq++] = in[i];
                    if (slt[in[i]] > dom) { dom_idx = in[i]; dom = slt[dom_idx]; }
                }
            }
            for (uint32_t i = 0; i < n_eq; i++) mapping[eq[i]] |= (eq[i] == dom_idx) ? F_MAJIN : F_MININ;
        }
    }
}

/* ------------------------------------------------------- segment_downchannels */

/* Cycle skip of the full ownership walks (with the fast-forward on): Brent's
 * detection on the pixel fires once the walk has gone round the cycle from a
 * point on it; every offer after that repeats one already made.  (The owned
 * walks cannot circle: the pixel where a walk entered a cycle is owned by one
 * of its two inflows only.) */
typedef struct { uint32_t sidx, shop, power; } LinkCycle;
static LinkCycle link_cycle(uint32_t idx) { LinkCycle c = {idx, 0u, 1u}; return c; }
static int link_cycle_step(LinkCycle* c, uint32_t idx, uint32_t hop) {
    if (idx == c->sidx) return 1;
    if (hop - c->shop == c->power) {
        c->sidx = idx;
        c->shop = hop;
        c->power <<= 1;
    }
    return 0;
}

/* Ownership key of a hop prev -> link[prev]: larger count, then a major
 * inflow, then larger pixel index. */
static inline uint64_t own_key(const uint32_t* mapping, const uint32_t* count, uint32_t prev) {
    return ((uint64_t)count[prev] << 32) | ((mapping[prev] & F_MAJIN) ? 0x80000000ull : 0ull) |
           (uint64_t)(prev & 0x7fffffffu);
}

/*
 * segment_downchannels (segment.cl:38-107).  The reference walks link from
 * every channel head, labelling pixels with the current segment label (the
 * pixel index where the segment starts) and starting a segment where
 * count >= seg_counter + SEGMENTATION_THRESHOLD/2; a walk stops at a pixel
 * already flagged IS_SUBSEGMENTHEAD, so which walk labels a shared reach
 * depends on the schedule.  Here every channel pixel has one owner hop:
 *   S0  heads are flagged IS_SUBSEGMENTHEAD;
 *   S1  full walks from every head (to a mask, self link or head pixel)
 *       offer own_key(prev) for each hop prev -> q; own[q] = the max;
 *   S2  each walk follows its owned hops with the reference's counters;
 *       a segment start is marked where the test fires, including the
 *       first pixel it does not own (the confluence it joins), and the walk
 *       stops there;
 *   S3  each walk labels its owned pixels, the label switching to q at
 *       every marked q (own marks and marks of joining walks);
 *   S4  marked pixels are flagged IS_SUBSEGMENTHEAD.
 * With an acyclic link graph each channel pixel is labelled by exactly one
 * walk; labels are written as max() so cycles stay deterministic.  Walks are
 * bounded by nxy_padded hops.
 */
void slmo_segment_downchannels(const SlmParams* p, const float* seeds, const uint8_t* mask,
                               uint32_t* mapping, const uint32_t* count, const uint32_t* link,
                               uint32_t* label) {
    const size_t nxy = p->nxy_padded;
    const uint32_t T = p->segmentation_threshold;
    uint64_t* own = (uint64_t*)calloc(nxy, sizeof(uint64_t));
    uint8_t* mark = (uint8_t*)calloc(nxy, 1);
    for (uint32_t k = p->seeds_chunk_offset; k < p->n_seed_points; k++) {
        const uint32_t h = seed_pixel(p, seeds, k);
        mapping[h] |= F_SSH;
        mark[h] = 1;
    }
    for (uint32_t k = p->seeds_chunk_offset; k < p->n_seed_points; k++) {
        uint32_t prev = seed_pixel(p, seeds, k), idx = lk(p, link, prev);
        LinkCycle cyc = link_cycle(prev);
        for (size_t hop = 0; hop < nxy && !mask[idx] && prev != idx && !(mapping[idx] & F_SSH); hop++) {
            kmax(&own[idx], own_key(mapping, count, prev));
            if (g_fast_forward && link_cycle_step(&cyc, idx, (uint32_t)hop + 1u)) break;
            prev = idx;
            idx = lk(p, link, idx);
        }
    }
    for (int phase = 2; phase <= 3; phase++) {
        for (uint32_t k = p->seeds_chunk_offset; k < p->n_seed_points; k++) {
            const uint32_t h = seed_pixel(p, seeds, k);
            uint32_t prev = h, idx = lk(p, link, h), seg_label = h, segc = 0;
            if (phase == 3) umax(&label[h], h);
            for (size_t hop = 0; hop < nxy && !mask[idx] && prev != idx && !(mapping[idx] & F_SSH); hop++) {
                const int owned = own[idx] == own_key(mapping, count, prev);
                if (phase == 2) {
                    if (count[idx] >= segc + T / 2u) {
                        mark[idx] = 1;
                        segc = count[idx] + T;
                    }
                    if (!owned) break;
                } else {
                    if (!owned) break;
                    if (mark[idx]) seg_label = idx;
                    umax(&label[idx], seg_label);
                }
                prev = idx;
                idx = lk(p, link, idx);
            }
        }
    }
    for (size_t i = 0; i < nxy; i++)
        if (mark[i]) mapping[i] |= F_SSH;
    free(own);
    free(mark);
}

/* ------------------------------------------------------------ link_hillslopes */

/* hillslopes (hillslopes.cl:31-91): step until the trace leaves its pixel
 * (countlink_runge_kutta_step); if it lands unmasked, link the seed pixel
 * there.  Writes only the seed pixel: schedule-free. */
void slmo_link_hillslopes(const SlmParams* p, const float* seeds, const uint8_t* mask,
                          const float* uv, uint32_t* link) {
    const uint32_t cap = 100u * p->max_n_steps;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t kk = p->seeds_chunk_offset; kk < p->n_seed_points; kk++) {
        const uint32_t k = (uint32_t)kk;
        real vec[2] = {(real)seeds[2 * (size_t)k], (real)seeds[2 * (size_t)k + 1]};
        real dt = (real)p->dt_max;
        uint32_t idx = array_idx(p, vec[0], vec[1]), prev_idx = idx, n = 0;
        real sx = vec[0], sy = vec[1], sdt = dt;
        uint32_t sn = 0, power = 1;
        while (prev_idx == idx && !mask[idx] && n < cap) {
            prev_idx = idx;
            if (countlink_step(p, uv, &dt, vec, &idx)) break;
            n++;
            if (g_fast_forward && prev_idx == idx) {
                /* a repeated state inside the seed pixel never leaves it: no link */
                if (vec[0] == sx && vec[1] == sy && dt == sdt) { n = cap; break; }
                if (n - sn == power) { sx = vec[0]; sy = vec[1]; sdt = dt; sn = n; power <<= 1; }
            }
        }
        if (prev_idx != idx && !mask[idx]) link[prev_idx] = idx;
    }
}

/* --------------------------------------------------- hillslope segmentation */

/* segment_runge_kutta_step (rungekutta.cl:435-459): moves, counts the step,
 * returns 1 when stuck, else updates dt. */
static inline int segment_step(const SlmParams* p, const float* uv, real* dt, real vec[2],
                               uint32_t* idx, uint32_t* n) {
    Step st;
    compute_step(p, uv, *dt, vec, NULL, 0, &st);
    *idx = st.idx;
    const real err = step_error(p, &st);
    const real dl = map_move(p, &st, vec, idx);
    *n += 1u;
    if (dl < (real)p->integration_halt_threshold) return 1;
    *dt = next_dt(p, *dt, err);
    return 0;
}

/* segment_hillslopes (segment.cl:133-188): trace from a non-thin pixel to a
 * thin channel pixel and copy its label.  Channel labels are not written
 * here: schedule-free. */
void slmo_segment_hillslopes(const SlmParams* p, const float* seeds, const uint8_t* mask,
                             const float* uv, const uint32_t* mapping, uint32_t* label) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t kk = p->seeds_chunk_offset; kk < p->n_seed_points; kk++) {
        const uint32_t k = (uint32_t)kk;
        real vec[2] = {(real)seeds[2 * (size_t)k], (real)seeds[2 * (size_t)k + 1]};
        real dt = (real)p->dt_max;
        uint32_t n = 0, idx = array_idx(p, vec[0], vec[1]);
        const uint32_t hs = idx;
        while (!mask[idx] && ((~mapping[idx]) & F_THIN) && n < p->max_n_steps)
            if (segment_step(p, uv, &dt, vec, &idx, &n)) break;
        if (mapping[idx] & F_THIN) label[hs] = label[idx];
    }
}

/* rotate_left (segment.cl:195-200): turn the king move (dx,dy) 45 degrees
 * anticlockwise (char arithmetic; abs of a char is a uchar, so |-128| reads
 * back as -128 and clamps to 1); returns (prev_x+dx)*NY_PADDED + (prev_y+dy)
 * in uint arithmetic.  An index at or beyond nxy_padded, which the reference
 * would read outside its arrays, reads here as a masked pixel. */
static inline uint32_t rotate_left(const SlmParams* p, uint32_t px, uint32_t py, int8_t* dx, int8_t* dy) {
    const int8_t rx = (int8_t)(*dx - *dy), ry = (int8_t)(*dx + *dy);
    int ax = (int8_t)(uint8_t)(rx < 0 ? -rx : rx), ay = (int8_t)(uint8_t)(ry < 0 ? -ry : ry);
    ax = ax < 1 ? 1 : (ax > 2 ? 2 : ax);
    ay = ay < 1 ? 1 : (ay > 2 ? 2 : ay);
    *dx = (int8_t)(rx / ax);
    *dy = (int8_t)(ry / ay);
    return (px + (uint32_t)(int32_t)*dx) * p->ny_padded + (py + (uint32_t)(int32_t)*dy);
}

/* subsegment_channel_edges (segment.cl:226-334): walk link down from a
 * subsegment head; at each hop scan the left-hand neighbours (rotating the
 * hop vector) and flag them IS_LEFTFLANK with LEFT_FLANK_ADDITION in label,
 * up to the next thin channel pixel, unless the first thin channel pixel of
 * a wider scan is the next hop (a kink).  Writes only bits it never reads:
 * schedule-free. */
void slmo_subsegment_channel_edges(const SlmParams* p, const float* seeds, const uint8_t* mask,
                                   uint32_t* mapping, const uint32_t* link, uint32_t* label) {
    const uint32_t nyp = p->ny_padded;
    for (uint32_t k = p->seeds_chunk_offset; k < p->n_seed_points; k++) {
        uint32_t prev_idx = seed_pixel(p, seeds, k);
        uint32_t idx = lk(p, link, prev_idx), n = 0;
        while (prev_idx != idx && n++ < p->max_n_steps) {
            const uint32_t px = prev_idx / nyp, py = prev_idx % nyp, x = idx / nyp, y = idx % nyp;
            int8_t dx = (int8_t)(x - px), dy = (int8_t)(y - py);
            uint32_t left = idx, turns = 0;
            int do_left = 1;
            while (left != prev_idx && ++turns <= 7) {
                left = rotate_left(p, px, py, &dx, &dy);
                const int ok = left < p->nxy_padded && !mask[left];
                if (turns >= 2 && ok && (mapping[left] & F_THIN)) {
                    if (left == lk(p, link, idx)) do_left = 0;
                    break;
                }
            }
            if (do_left) {
                turns = 0;
                left = idx;
                dx = (int8_t)(x - px);
                dy = (int8_t)(y - py);
                while (left != prev_idx && ++turns <= 4) {
                    left = rotate_left(p, px, py, &dx, &dy);
                    const int ok = left < p->nxy_padded && !mask[left];
                    if (turns >= 2 && ok) {
                        if (mapping[left] & F_THIN) break;
                        mapping[left] |= F_LEFT;
                        label[left] |= LFA;
                    }
                }
            }
            prev_idx = idx;
            idx = lk(p, link, idx);
            if (mask[idx] || (mapping[prev_idx] & F_SSH)) break;
        }
    }
}

/* subsegment_flanks (segment.cl:360-414): trace to a left-flank or thin
 * pixel; reaching a left flank marks the seed's label LEFT_FLANK_ADDITION.
 * Reads mapping only: schedule-free. */
void slmo_subsegment_flanks(const SlmParams* p, const float* seeds, const uint8_t* mask,
                            const float* uv, const uint32_t* mapping, uint32_t* label) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t kk = p->seeds_chunk_offset; kk < p->n_seed_points; kk++) {
        const uint32_t k = (uint32_t)kk;
        real vec[2] = {(real)seeds[2 * (size_t)k], (real)seeds[2 * (size_t)k + 1]};
        real dt = (real)p->dt_max;
        uint32_t n = 0, idx = array_idx(p, vec[0], vec[1]);
        const uint32_t hs = idx;
        while (!mask[idx] && !(mapping[idx] & F_LEFT) && !(mapping[idx] & F_THIN) && n < p->max_n_steps)
            if (segment_step(p, uv, &dt, vec, &idx, &n)) break;
        if (mapping[idx] & F_LEFT) label[hs] |= LFA;
    }
}

/*
 * fix_right_flanks (segment.cl:441-504) and fix_left_flanks (:530-594):
 * trace from a seed while the pixel is on the seed's side (right: label
 * without LEFT_FLANK_ADDITION; left: with it) and not thin channel, adopting
 * the label of each probed pixel whose label (flank bit aside) differs; at
 * the end adopt the arrival label and set (right) or clear (left) the flank
 * bit from the arrival pixel.  The reference reads labels that other work
 * items rewrite in the same launch.  Here every read of another pixel sees
 * the labels as they were when the call started; the seed's own label is
 * read back as last written.
 */
static void fix_flanks(const SlmParams* p, const float* seeds, const uint8_t* mask, const float* uv,
                       const uint32_t* mapping, uint32_t* label, int left) {
    const size_t nxy = p->nxy_padded;
    uint32_t* snap = (uint32_t*)malloc(nxy * sizeof(uint32_t));
    memcpy(snap, label, nxy * sizeof(uint32_t));
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t kk = p->seeds_chunk_offset; kk < p->n_seed_points; kk++) {
        const uint32_t k = (uint32_t)kk;
        real vec[2] = {(real)seeds[2 * (size_t)k], (real)seeds[2 * (size_t)k + 1]};
        real dt = (real)p->dt_max;
        uint32_t n = 0, idx = array_idx(p, vec[0], vec[1]);
        const uint32_t hs = idx;
        uint32_t own = snap[hs];
#define LAB(q) ((q) == hs ? own : snap[q])
        for (;;) {
            const int side = left ? ((LAB(idx) & LFA) != 0) : ((LAB(idx) & LFA) == 0);
            if (!(!mask[idx] && side && !(mapping[idx] & F_THIN) && n < p->max_n_steps)) break;
            Step st;
            compute_step(p, uv, dt, vec, NULL, 0, &st);
            idx = st.idx;
            if ((own & ~LFA) != (LAB(idx) & ~LFA)) own = LAB(idx);
            const real err = step_error(p, &st);
            const real dl = map_move(p, &st, vec, &idx);
            n += 1u;
            if (dl < (real)p->integration_halt_threshold) break;
            dt = next_dt(p, dt, err);
        }
        if ((own & ~LFA) != (LAB(idx) & ~LFA)) own = LAB(idx);
        if (left) {
            if (!(LAB(idx) & LFA) && !(mapping[idx] & F_THIN)) own &= ~LFA;
        } else {
            if (LAB(idx) & LFA) own |= LFA;
        }
#undef LAB
        label[hs] = own;
    }
    free(snap);
}

void slmo_fix_right_flanks(const SlmParams* p, const float* seeds, const uint8_t* mask,
                           const float* uv, const uint32_t* mapping, uint32_t* label) {
    fix_flanks(p, seeds, mask, uv, mapping, label, 0);
}

void slmo_fix_left_flanks(const SlmParams* p, const float* seeds, const uint8_t* mask,
                          const float* uv, const uint32_t* mapping, uint32_t* label) {
    fix_flanks(p, seeds, mask, uv, mapping, label, 1);
}
