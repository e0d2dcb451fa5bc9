/*
 * taastar serving runtime (C ABI): continuous batching of TAA frame requests from many sessions
 * of different resolutions on one GPU.
 *
 * No counterpart in the reference, whose host driver serves exactly one stream of frames
 * (source/taa.hpp:974-1192, one dispatch per frame with CF frames in flight). Here each session
 * keeps its history, depth ring and input slots resident on the device (taa_session_t); sessions
 * are admitted against an HBM budget; requests that arrive while other batches are on the GPU are
 * gathered into the next batch, and up to `max_inflight` batches overlap on separate streams.
 * Every frame is resolved by the same per-pixel arithmetic as taa_resolve_frame, so a session's
 * output does not depend on what it was batched with.
 */
#ifndef TAASTAR_TAA_SERVE_H
#define TAASTAR_TAA_SERVE_H

#include <stdint.h>

#include "taa_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct taa_server taa_server_t;

typedef struct taa_server_config {
    int32_t  device;
    uint64_t hbm_budget_bytes;  /* admission budget for sessions (0 = 90% of free memory) */
    int32_t  color_format;      /* TAA_FMT_* of every session */
    int32_t  arith;             /* TAA_ARITH_* */
    int32_t  max_batch;         /* frames per batch (<= 1024) */
    int32_t  max_inflight;      /* batches on the GPU at once (1..8) */
    float    cas_sharpness;     /* < 0: no CAS */
} taa_server_config_t;

typedef struct taa_completion {
    int32_t  session_id;
    int32_t  status;            /* TAA_OK, or the error of the batched launch that failed */
    uint64_t tag;
    double   t_submit_us;       /* host time of taa_server_submit (or the given arrival time) */
    double   t_done_us;         /* GPU completion time on the same clock */
} taa_completion_t;

typedef struct taa_server_stats {
    uint64_t batches, frames, pixels;
    uint64_t frames_failed;     /* completed with a non-zero status */
    uint64_t sessions_open, admissions_refused;
    uint64_t hbm_used, hbm_budget;
    double   busy_us;           /* sum of batch GPU spans */
} taa_server_stats_t;

int32_t taa_server_create(const taa_server_config_t* cfg, taa_server_t** out);
int32_t taa_server_destroy(taa_server_t* s);
/* Admission: TAA_ERR_BUDGET when the session (plus its output image) does not fit. */
int32_t taa_server_open(taa_server_t* s, int32_t width, int32_t height, int32_t* session_id);
/* Waits for the session's queued and in-flight frames, then frees it. */
int32_t taa_server_close(taa_server_t* s, int32_t session_id);
/* The session's input slots (see taa_session_io) and its output (screen) image. Waits (pumping)
 * until the session has no queued or running frame, so the caller may then write the next
 * frame's inputs and read the last result; both stay valid until the next taa_server_submit of
 * this session. */
int32_t taa_server_io(taa_server_t* s, int32_t session_id, taa_session_io_t* io, void** screen);
/* Queue one frame; its inputs must already be in the session's slots. arrival_us < 0 stamps the
 * call time. Frames of one session are resolved in order, one at a time, each reading the slots
 * as they are when it runs (several queued frames of a session therefore resolve the same
 * inputs). Uniforms a batched launch refuses (taa_batch_submit) are refused here with its status;
 * a frame whose batched launch fails completes with that status. */
int32_t taa_server_submit(taa_server_t* s, int32_t session_id, const taa_uniforms_t* uniforms, uint64_t tag,
                          double arrival_us);
/* Harvest finished batches and launch queued frames into free in-flight slots. */
int32_t taa_server_pump(taa_server_t* s);
/* Completed frames since the last call (at most `max`). */
int32_t taa_server_poll(taa_server_t* s, taa_completion_t* out, int32_t max, int32_t* n_out);
/* Pump until nothing is queued or in flight. */
int32_t taa_server_drain(taa_server_t* s);
int32_t taa_server_stats(const taa_server_t* s, taa_server_stats_t* out);
/* Host clock used for submit / completion stamps (microseconds). */
double taa_server_now_us(void);

/* ---- load simulation ------------------------------------------------------------------------ */

typedef struct taa_serve_sim_config {
    int32_t  device;
    uint64_t seed;
    int32_t  mode;              /* 0 open loop: Poisson arrivals + session churn
                                   1 closed loop: every session resubmits on completion
                                   2 per-frame: one taa_session_resolve launch per frame,
                                     sessions round-robin on one stream, no host waits
                                   3 per-frame synchronous: as 2, host waits for each frame
                                   4 batched graph: every session, one frame each, per
                                     taa_batch_plan_launch, back to back */
    int32_t  sessions;          /* initial sessions */
    int32_t  n_classes;         /* resolution classes */
    int32_t  class_w[8], class_h[8];
    float    class_p[8];        /* probabilities of the classes */
    double   arrival_rate;      /* open loop: frames/s over all sessions */
    double   churn_per_s;       /* open loop: share of sessions replaced per second */
    double   duration_s;        /* simulated (wall-clock) time */
    uint64_t hbm_budget_bytes;
    int32_t  color_format, arith;
    float    cas_sharpness;
    int32_t  max_batch, max_inflight;
    int32_t  warmup_frames;     /* per session, before measuring (modes 1-4) */
    int32_t  exact_mix;         /* != 0: classes assigned in proportion to class_p (largest
                                   remainder) instead of drawn at random */
    int32_t  per_session_inputs;/* != 0: every session gets an input frame of its own (own seed,
                                   jitter phase = frame index mod 8, camera pan of random
                                   direction and magnitude uniform in [0, pan_max] px/frame)
                                   instead of sharing one frame per resolution class. The frames
                                   are rendered before the measured run (the initial sessions'
                                   and a per-class reserve of 1.5x the expected churn admissions);
                                   a departing session's frame is handed to the next admission of
                                   its class, so active sessions never share an input */
    float    pan_max;
} taa_serve_sim_config_t;

typedef struct taa_serve_report {
    uint64_t frames, pixels, batches;
    uint64_t admissions, admissions_refused;
    int32_t  sessions_end, sessions_peak;
    double   wall_s;
    double   gpix_per_s;        /* resolved pixels / wall time */
    double   offered_fps, achieved_fps;
    double   lat_p50_us, lat_p90_us, lat_p99_us, lat_max_us, lat_mean_us;
    double   mean_batch;
    double   gpu_busy_frac;     /* union of batch spans / wall */
    uint64_t hbm_peak_bytes;
} taa_serve_report_t;

/* Run a load against a server on `device` with synthetic inputs (taa_scene frames of each class
 * uploaded once and copied into every session's slots when it opens). */
int32_t taa_serve_simulate(const taa_serve_sim_config_t* cfg, taa_serve_report_t* out);

#ifdef __cplusplus
}
#endif
#endif /* TAASTAR_TAA_SERVE_H */
