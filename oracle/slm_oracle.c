/*
 * slm_oracle.c -- CPU restatement of slm's streamline-tracing kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the
 * smoke check and bench.py's cpu_baseline leg call it to check or time the
 * HIP path.  Nothing in the product (slm_amd/) loads it.
 *
 * It restates, function by function, the OpenCL C of the reference:
 *   opencl/rng.cl, essentials.cl, computestep.cl, rungekutta.cl,
 *   updatetraj.cl, writearray.cl, jittertrajectory.cl, integratefields.cl,
 *   trajectory.cl, integratetraj.cl
 * with every -D macro taken from SlmParams (include/slm_hip.h).
 *
 * Built twice from this one source into liboracle.so:
 *   default  -> symbols slmo_*    (FP32, the parity oracle)
 *   -DSLM_F64 -> symbols slmo64_*  (FP64 trace, the error yardstick)
 * Compiled with -ffp-contract=off: the only fused multiply-adds are the
 * explicit MAD() calls, made when p->fp_contract is set, at the places where
 * clang's OpenCL default (FP_CONTRACT ON) forms llvm.fmuladd: an a*b+c or
 * a+=b*c inside one expression, the left product taken first.
 *
 * Semantics fixed here where OpenCL leaves them to the implementation:
 *   fast_length(v)    = sqrt(dot(v,v)), dot(v,v) = MAD(x,x,y*y)
 *   fast_normalize(v) = v * (1/sqrt(dot(v,v)))  (0 vector -> NaN)
 *   (char)f           = low byte of the saturating truncation to int32, NaN -> 0
 *   (uint)f           = saturating truncation, NaN -> 0
 * The HIP kernels implement exactly the same choices.
 */
#include "oracle_math.h"

uint32_t SLM_PREFIX(lehmer)(uint32_t* s) { return lehmer_u32(s); }
uint32_t SLM_PREFIX(rng_seed)(uint32_t seed_idx) { return rng_seed_u32(seed_idx); }

/* --------------------------------------------- jittered fields streamline */

/* Sub-pixel start offset (integratefields.cl:102-111): i*STEP-HALFSPAN */
static inline real sub_offset(const SlmParams* p, uint32_t i) {
    return MAD((int)p->fp_contract, (real)(float)i, (real)p->subpixel_seed_step,
               -(real)p->subpixel_seed_halfspan);
}

/*
 * One jittered streamline of integrate_fields (jittertrajectory.cl:42-88).
 * The else-branch there (Euler tail) belongs to `if (idx<NXY_PADDED)`, which
 * is always true because get_array_idx clamps; so a step whose end pixel is
 * masked is simply dropped and the loop ends.  Writes go to slc/slt when
 * non-NULL; positions go to rec (up to max_rec) when non-NULL.
 * Returns the number of accepted steps (n_steps).
 */
static uint32_t fields_streamline(const SlmParams* p, const uint8_t* mask, const float* uv,
                                  uint32_t* slc, uint32_t* slt, real x0, real y0,
                                  uint32_t rng0, float* rec, uint32_t max_rec, uint32_t* nrec,
                                  int atomic) {
    uint32_t rng = rng0, n = 0, idx;
    real l = 0, dt = (real)p->dt_max;
    real vec[2] = {x0, y0};
    uint32_t nr = 0;
    if (rec && nr < max_rec) { rec[2 * nr] = (float)vec[0]; rec[2 * nr + 1] = (float)vec[1]; }
    nr++;
    idx = array_idx(p, vec[0], vec[1]);
    if (!mask[idx] && slc) {
        if (atomic) {
#pragma omp atomic
            slt[idx] += slt_inc(l);
#pragma omp atomic
            slc[idx] += 1u;
        } else { slt[idx] += slt_inc(l); slc[idx] += 1u; }
    }
    Step st;
    while (idx < p->nxy_padded && !mask[idx] && (l < (real)p->max_length && n < p->max_n_steps)) {
        compute_step(p, uv, dt, vec, &rng, 1, &st);
        idx = st.idx;
        if (idx < p->nxy_padded) {
            if (!mask[idx]) {
                const real err = step_error(p, &st);
                const real dl = fast_length((int)p->fp_contract, st.dxy2[0], st.dxy2[1]);
                vec[0] = st.next[0]; vec[1] = st.next[1];
                const int stuck = dl < (real)p->integration_halt_threshold;
                if (!stuck) dt = next_dt(p, dt, err);
                /* update_trajectory_write_sl_data (updatetraj.cl:40-55) */
                l += dl;
                n += 1u;
                idx = array_idx(p, vec[0], vec[1]);
                if (idx < p->nxy_padded && !mask[idx] && slc) {
                    if (atomic) {
#pragma omp atomic
                        slt[idx] += slt_inc(l);
#pragma omp atomic
                        slc[idx] += 1u;
                    } else { slt[idx] += slt_inc(l); slc[idx] += 1u; }
                }
                if (rec && nr < max_rec) { rec[2 * nr] = (float)vec[0]; rec[2 * nr + 1] = (float)vec[1]; }
                nr++;
                if (stuck) break;
            }
        }
    }
    if (nrec) *nrec = nr;
    return n;
}

/*
 * integrate_fields over all seeds of one direction (integratefields.cl:60-114).
 * seeds_xy[k] is the shuffled seed k; its RNG key is seeds_chunk_offset+k.
 * slc/slt are accumulated into (not cleared).  Returns the accepted-step count.
 */
uint64_t SLM_PREFIX(trace_fields)(const SlmParams* p, const float* seeds_xy,
                                  const uint8_t* mask, const float* uv,
                                  uint32_t* slc, uint32_t* slt, int nthreads) {
    const uint32_t d = p->subpixel_seed_point_density;
    const int64_t ns = (int64_t)p->n_seed_points - (int64_t)p->seeds_chunk_offset;
    uint64_t total = 0;
    if (ns <= 0) return 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t k = 0; k < ns; k++) {
        const uint32_t seed_idx = p->seeds_chunk_offset + (uint32_t)k;
        const uint32_t rng0 = rng_seed_u32(seed_idx);
        const real sx = (real)seeds_xy[2 * (size_t)seed_idx];
        const real sy = (real)seeds_xy[2 * (size_t)seed_idx + 1];
        for (uint32_t j = 0; j < d; j++)
            for (uint32_t i = 0; i < d; i++)
                total += fields_streamline(p, mask, uv, slc, slt, sx + sub_offset(p, i),
                                           sy + sub_offset(p, j), rng0, NULL, 0, NULL,
                                           nthreads != 1);
    }
    return total;
}

/*
 * One jittered fields streamline, positions recorded (test hook).
 * sub = j*d + i selects the sub-pixel start as the loops of integratefields.cl:102-111.
 */
uint32_t SLM_PREFIX(trace_fields_record)(const SlmParams* p, const float* seeds_xy,
                                         const uint8_t* mask, const float* uv,
                                         uint32_t seed_k, uint32_t sub,
                                         float* xy, uint32_t max_rec, uint32_t* nrec) {
    const uint32_t d = p->subpixel_seed_point_density ? p->subpixel_seed_point_density : 1u;
    const uint32_t seed_idx = p->seeds_chunk_offset + seed_k;
    const uint32_t i = sub % d, j = sub / d;
    return fields_streamline(p, mask, uv, NULL, NULL,
                             (real)seeds_xy[2 * (size_t)seed_idx] + sub_offset(p, i),
                             (real)seeds_xy[2 * (size_t)seed_idx + 1] + sub_offset(p, j),
                             rng_seed_u32(seed_idx), xy, max_rec, nrec, 0);
}

/* ------------------------------------------- un-jittered trajectory record */

/* compress (essentials.cl:71-74) */
static inline void compress(const SlmParams* p, real dx, real dy, int8_t* out) {
    const real res = (real)p->trajectory_resolution;
    out[0] = cvt_i8(dx * res);
    out[1] = cvt_i8(dy * res);
}

/*
 * trajectory_record (trajectory.cl:41-80) with runge_kutta_step_record
 * (rungekutta.cl:49-77), euler_step_record (:104-120),
 * update_record_trajectory (updatetraj.cl:77-86), finalize_trajectory (:165-172).
 */
static void trajectory_record(const SlmParams* p, const uint8_t* mask, const float* uv,
                              real x0, real y0, int8_t* traj, uint16_t* nsteps_out,
                              float* length_out, real* length_out64) {
    const int c = (int)p->fp_contract;
    uint32_t n = 0, idx;
    real l = 0, dl = 0, dt = (real)p->dt_max;
    real vec[2] = {x0, y0}, prev[2] = {x0, y0};
    Step st;
    idx = array_idx(p, vec[0], vec[1]);
    while (!mask[idx] && (l < (real)p->max_length && n < p->max_n_steps)) {
        compute_step(p, uv, dt, vec, NULL, 0, &st);
        idx = st.idx;
        if (!mask[idx]) {
            const real err = step_error(p, &st);
            dl = fast_length(c, st.dxy2[0], st.dxy2[1]);
            if (dl < (real)p->integration_halt_threshold) {
                st.next[0] = vec[0] + st.dxy1[0];
                st.next[1] = vec[1] + st.dxy1[1];
                dl = fast_length(c, st.dxy1[0], st.dxy1[1]);
            }
            vec[0] = st.next[0]; vec[1] = st.next[1];
            idx = array_idx(p, st.next[0], st.next[1]);
            const int stuck = dl < (real)p->integration_halt_threshold;
            if (!stuck) dt = next_dt(p, dt, err);
            l += dl;
            compress(p, vec[0] - prev[0], vec[1] - prev[1], traj + 2 * (size_t)n);
            n += 1u;
            if (stuck) break;
            prev[0] = vec[0]; prev[1] = vec[1];
        } else {
            /* euler_step_record: DOWNUP_SIGN is 1 or (float)(uint)-1 = 2^32 */
            const real sg = (real)p->downup_sign;
            const real su = st.uv1[0] * sg, sv = st.uv1[1] * sg;
            const real dtx = dt_edge(vec[0], su);
            const real dty = dt_edge(vec[1], sv);
            dt = minmag(dtx, dty);
            vec[0] = MAD(c, su, dt, vec[0]);
            vec[1] = MAD(c, sv, dt, vec[1]);
            /* approximate(*vec) (rungekutta.cl:115) of the absolute position:
             * (char) is out of range here; approx_wrap selects the low byte,
             * otherwise the int32 truncation (the reference's recorded runs) */
            if (p->approx_wrap) {
                vec[0] = approx(p, vec[0]);
                vec[1] = approx(p, vec[1]);
            } else {
                const real res = (real)p->trajectory_resolution;
                vec[0] = (real)cvt_i32(vec[0] * res) / res;
                vec[1] = (real)cvt_i32(vec[1] * res) / res;
            }
            vec[0] = r_fmin(r_fmax(vec[0], (real)-0.5), (real)p->nxf_mp5);
            vec[1] = r_fmin(r_fmax(vec[1], (real)-0.5), (real)p->nyf_mp5);
            dl = fast_length(c, vec[0] - prev[0], vec[1] - prev[1]);
            l += dl;
            compress(p, vec[0] - prev[0], vec[1] - prev[1], traj + 2 * (size_t)n);
            n += 1u;
            break;
        }
    }
    *nsteps_out = (uint16_t)n;
    if (length_out) *length_out = (float)l;
    if (length_out64) *length_out64 = l;
}

/*
 * integrate_trajectory over seeds [seeds_chunk_offset, n_seed_points)
 * (integratetraj.cl:54-97): seed k writes traj[k*max_n_steps ...], nsteps[k], length[k].
 */
void SLM_PREFIX(trace_paths)(const SlmParams* p, const float* seeds_xy,
                             const uint8_t* mask, const float* uv,
                             int8_t* traj, uint16_t* nsteps, float* length, int nthreads) {
    const int64_t ns = (int64_t)p->n_seed_points - (int64_t)p->seeds_chunk_offset;
    if (ns <= 0) return;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t k = 0; k < ns; k++) {
        const size_t s = (size_t)p->seeds_chunk_offset + (size_t)k;
        trajectory_record(p, mask, uv, (real)seeds_xy[2 * s], (real)seeds_xy[2 * s + 1],
                          traj + 2 * (size_t)k * p->max_n_steps, nsteps + k, length + k, NULL);
    }
}

/*
 * hillslope_lengths (lengths.cl:37-116) with lengths_runge_kutta_step
 * (rungekutta.cl:492-516): seed k in [seeds_chunk_offset, n_seed_points)
 * traces downstream (un-jittered) to a thin_flag pixel and writes
 * length[k] = l * PIXEL_SIZE; masked pixels, re-crossing a from_flag pixel
 * after leaving the seed band, and stuck steps abandon the trace.
 */
void SLM_PREFIX(hillslope_lengths)(const SlmParams* p, const float* seeds_xy, const uint8_t* mask,
                                   const float* uv, const uint32_t* mapping, uint32_t thin_flag,
                                   uint32_t from_flag, float* length) {
    const int c = (int)p->fp_contract;
    for (uint32_t k = p->seeds_chunk_offset; k < p->n_seed_points; k++) {
        real vec[2] = {(real)seeds_xy[2 * (size_t)k], (real)seeds_xy[2 * (size_t)k + 1]};
        real l = 0, dt = (real)p->dt_max;
        uint32_t n = 0;
        int moved_off = 0, abandon = 0;
        uint32_t idx = array_idx(p, vec[0], vec[1]);
        Step st;
        while (((~mapping[idx]) & thin_flag) && n < p->max_n_steps) {
            compute_step(p, uv, dt, vec, NULL, 0, &st);
            idx = st.idx;
            if (mask[idx]) { abandon = 1; break; }
            if (!moved_off && ((~mapping[idx]) & from_flag)) moved_off = 1;
            else if (moved_off && (mapping[idx] & from_flag)) { abandon = 1; break; }
            const real err = step_error(p, &st);
            real dl = fast_length(c, st.dxy2[0], st.dxy2[1]);
            if (dl < (real)p->integration_halt_threshold) {
                st.next[0] = vec[0] + st.dxy1[0];
                st.next[1] = vec[1] + st.dxy1[1];
                dl = fast_length(c, st.dxy1[0], st.dxy1[1]);
            }
            vec[0] = st.next[0]; vec[1] = st.next[1];
            idx = array_idx(p, vec[0], vec[1]);
            if (dl < (real)p->integration_halt_threshold) { abandon = 1; break; }
            dt = next_dt(p, dt, err);
            l += dl;
            n += 1u;
        }
        if (!abandon && (mapping[idx] & thin_flag)) length[k] = (float)(l * (real)p->pixel_size);
    }
}

/* Lehmer draws for n seeds from idx0 (slm_rng_stream restated). */
void SLM_PREFIX(rng_stream)(uint32_t idx0, uint32_t n, uint32_t n_draws,
                            uint32_t* rng0_out, uint32_t* draws) {
    for (uint32_t i = 0; i < n; i++) {
        uint32_t s = rng_seed_u32(idx0 + i);
        if (rng0_out) rng0_out[i] = s;
        if (draws)
            for (uint32_t k = 0; k < n_draws; k++) draws[(size_t)i * n_draws + k] = lehmer_u32(&s);
    }
}
