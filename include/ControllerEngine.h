/*
 * ControllerEngine for AMD Instinct MI355X (gfx950): public C ABI.
 *
 * Two layers are exported from libControllerEngine.so (also shipped as ControllerEngine.so):
 *
 *  1. The drop-in entry point, RunControllerEngine().  It replaces the reference library's only export
 *     (reference src/ControllerEngine.h:11, defined at src/main.cu:592-602): same name, no arguments,
 *     same file protocol in the current working directory (ControllerEngine.config, STATESPACE_*.bin,
 *     DISCRETE_*.bin, BASE_PARAMS.bin, DC/U/NL_INFO.bin, CUDA_MX.bin, INFO_MX.bin in;
 *     INFO_MX_OUT.bin, params.log and the optional stage dumps out), same return convention
 *     (0 = success, 1 = failure with the message appended to error.log).
 *     The arithmetic mode is chosen by the environment variable CE_PRECISION = fp64 (default) | mixed,
 *     because the reference ABI takes no arguments and its config parser rejects unknown keys
 *     (src/config.cu:239).
 *
 *  2. An extended API (no counterpart in the reference) for callers that keep inputs on the device:
 *     config parsing, plant loading, device-resident evaluation of a candidate block, batched
 *     evaluation of many problems in one launch chain, and a continuous-batching server.
 *
 * All pointers are plain pointers; device pointers are HIP device addresses, streams are hipStream_t
 * passed as void*.  Functions return 0 on success and nonzero on failure; on failure a message is
 * written to err (errlen bytes, NUL-terminated) when err is not NULL.
 */
#ifndef CONTROLLERENGINE_H
#define CONTROLLERENGINE_H

#include <stddef.h>
#include <stdint.h>

#if defined(_WIN32)
#define CE_EXPORT __declspec(dllexport)
#else
#define CE_EXPORT __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Drop-in entry point. Reference: src/ControllerEngine.h:11, src/main.cu:592-602. */
CE_EXPORT int RunControllerEngine(void);

/* Compile-time limits. Reference: src/main.cu:43-52. */
enum {
  CE_MAX_STATESPACE_DIM = 8,
  CE_MAX_DISCRETE_DIM = 8,
  CE_MAX_NONLIN_DIM = 4,
  CE_MAX_TOTAL_DIM = 16,
  CE_MAX_NL_DIM = 9,
  CE_MAX_DC_INSTRUCTIONS = 4,
  CE_MAX_U_INSTRUCTIONS = 4,
  CE_MAX_NL_INSTRUCTIONS = 8,
  CE_MAX_PARAMS = 34
};

/* Arithmetic modes. */
enum {
  CE_PRECISION_FP64 = 0,  /* every stage in double */
  CE_PRECISION_MIXED = 1  /* float simulation, fit accumulation and sweep; double eigensolve,
                             normal-equation solve and crossover refinement */
};

/* The parsed ControllerEngine.config. Field meaning follows reference src/config.cu:5-27; every
 * key, default and validation rule of load_config (src/config.cu:106-244) applies. */
typedef struct ce_config {
  int32_t statespace_dim, discrete_dim, nonlin_dim, n_params, input_size;
  int32_t dc_instructions, u_instructions;
  int32_t nl_instructions[CE_MAX_NONLIN_DIM];
  int32_t n_cores, n_rounds;
  int32_t serial;            /* mode = serial */
  int32_t time;              /* samples per run */
  float duration;
  double ref_step_time;
  float ref_amplitude;
  float fit_start, fit_end, steady_start;
  float bandwidth, damping, overshoot, steady_state_error, crossover_coeff;
  float target_gm, target_pm, target_b;
  float eig_fail_err, unstable_err;
  int32_t freq;
  float freq_min;
  int32_t dump_diagnostics;
} ce_config;

/* One round's plant and observer model, compact row-major, exactly the .bin contents
 * (reference src/main.cu:341-400 reads them). */
typedef struct ce_plant {
  float ss_a[CE_MAX_STATESPACE_DIM * CE_MAX_STATESPACE_DIM];    /* STATESPACE_A    ns x ns */
  float ss_a_i[CE_MAX_STATESPACE_DIM * CE_MAX_STATESPACE_DIM];  /* STATESPACE_A_I  ns x ns */
  float ss_b[CE_MAX_STATESPACE_DIM];                            /* STATESPACE_B    ns */
  float ss_c[CE_MAX_STATESPACE_DIM];                            /* STATESPACE_C    ns */
  float ss_d;                                                   /* STATESPACE_D    1 */
  float d_a[CE_MAX_DISCRETE_DIM * CE_MAX_DISCRETE_DIM];         /* DISCRETE_A      nd x nd */
  float d_a_i[CE_MAX_DISCRETE_DIM * CE_MAX_DISCRETE_DIM];       /* DISCRETE_A_I    nd x nd */
  float d_bu[CE_MAX_DISCRETE_DIM];                              /* DISCRETE_BU     nd */
  float d_bw[CE_MAX_DISCRETE_DIM * CE_MAX_NONLIN_DIM];          /* DISCRETE_BW     nd x nonlin_dim */
  float d_c[CE_MAX_DISCRETE_DIM];                               /* DISCRETE_C      nd */
  float base_params[CE_MAX_PARAMS];                             /* BASE_PARAMS     n_params */
  int32_t dc_info[CE_MAX_DC_INSTRUCTIONS * CE_MAX_DISCRETE_DIM];               /* dc_instructions x nd */
  int32_t u_info[CE_MAX_U_INSTRUCTIONS * CE_MAX_STATESPACE_DIM];               /* u_instructions x ns */
  int32_t nl_info[CE_MAX_NONLIN_DIM * CE_MAX_NL_INSTRUCTIONS * CE_MAX_NL_DIM]; /* sum(nl) x (nd+1) */
} ce_plant;

/* Quantities derive() computes from the config (reference src/config.cu:40-78, 93-104), in the
 * reference's float32 arithmetic so that sample counts and indices agree. */
typedef struct ce_derived {
  int32_t n_threads, total_dim, cuda_mx_width, info_mx_width;
  int32_t info_scoring, info_freqresp, info_allmargin, info_modal;
  int32_t fit_start_id, steady_start_id, fit_count; /* fit_count = samples the fit loop visits */
  float ts, freq_max, fit_t0, fit_t1, crossover;
} ce_derived;

/* Parse config text (file name used in messages). Reference: src/config.cu:106-244. */
CE_EXPORT int ce_config_parse(const char* text, const char* filename, ce_config* out, char* err, int errlen);
CE_EXPORT int ce_config_derive(const ce_config* cfg, ce_derived* out, char* err, int errlen);
/* Read round `round` of the plant files from directory dir (exact byte sizes enforced).
 * Reference: src/main.cu:293-315 (readbin), 341-400 (upload_plant, init64KB). */
CE_EXPORT int ce_plant_load_dir(const char* dir, const ce_config* cfg, int round, ce_plant* out, char* err, int errlen);

/* ---------------------------------------------------------------------------------------------
 * Device evaluation of one candidate block (the six-stage chain of reference src/main.cu:467-558
 * for one round).  gains: n x (2*statespace_dim+1) floats, info_in/info_out: n x info_mx_width
 * floats, all device pointers.  info64_out (optional, may be NULL) receives the same rows in double.
 * Work is enqueued on `stream`; the call returns after enqueueing unless it has to report an error
 * found on the host.  An engine holds device scratch sized by its largest call so far. */
typedef struct ce_engine ce_engine;

CE_EXPORT ce_engine* ce_engine_create(int device, char* err, int errlen);
CE_EXPORT void ce_engine_destroy(ce_engine* eng);
CE_EXPORT int ce_evaluate(ce_engine* eng, const ce_config* cfg, const ce_plant* plant, int precision,
                          const float* d_gains, const float* d_info_in, float* d_info_out, double* d_info64_out,
                          int n, void* stream, char* err, int errlen);

/* Kalman-filter closed-loop simulation: the reference's discrete_solve<ControllerType::Kalman>
 * (src/discrete_solves.cu:83-185), which the reference never launches.  Trajectories y (n x time)
 * only: in FP64 into d_y64 (double), in mixed precision into d_y32 (float, compensated state).
 * d_gains: n rows of [K (statespace_dim), N, R, Q (statespace_dim)] -- 2 statespace_dim + 2 floats;
 * the reference sizes a gains row 2 statespace_dim + 1 (src/config.cu:51) and so reads the last Q
 * entry from the next row.  Per step: predict the state and the covariance P = A P A' + diag(Q),
 * gain KG = P C' / (C P C' + R), update xhat += KG (y - C xhat) and
 * P = (I - KG C) P (I - KG C)' + R KG KG' (the product the reference's in-place update intends;
 * its fast reciprocal is an IEEE division here).  Enqueued on `stream`. */
CE_EXPORT int ce_simulate_kalman(ce_engine* eng, const ce_config* cfg, const ce_plant* plant, int precision,
                                 const float* d_gains, int n, double* d_y64, float* d_y32, void* stream, char* err,
                                 int errlen);

/* Optional per-stage device outputs of the last ce_evaluate call (the reference's dump_diagnostics
 * files, src/main.cu:483-547).  Sizes per candidate: yplant time, lina nt*nt, eigen 2*nt,
 * ata (2nt)^2 (lower triangle, row stride 2nt), atb 2nt.  All float, candidate-major like the
 * reference dumps. Any pointer may be NULL. */
CE_EXPORT int ce_engine_set_diagnostics(ce_engine* eng, float* d_yplant, float* d_lina, float* d_eigen,
                                        float* d_ata, float* d_atb);

/* Per-stage timings of the last ce_evaluate (milliseconds, from HIP events on its stream). */
CE_EXPORT int ce_engine_stage_times(ce_engine* eng, double* ms, int n_max);
/* 1 when the last ce_evaluate ran the runtime-specialised (hipRTC) simulation kernel.  The engine
 * specialises the per-step code of a problem unless CE_JIT=0 is set in the environment. */
CE_EXPORT int ce_engine_last_sim_jit(ce_engine* eng);
/* Kernel name of the last call's simulation launch (as a kernel trace lists it). */
CE_EXPORT int ce_engine_last_sim_kernel(ce_engine* eng, char* name, int len);
/* Split simulation (no counterpart in the reference): when the mixed chain's one-lane simulation
 * has more than two full rounds of one wave per SIMD, the one-lane kernel takes the list slots of
 * the full rounds and a 2-lane kernel the rest, on another stream.  ce_engine_last_sim_split
 * returns the SIMD count the last call split by (0: not split); ce_sim_split_slots(live, cap) the
 * first list slot of the remainder for live simulated candidates (live when there is none). */
CE_EXPORT int ce_engine_last_sim_split(ce_engine* eng);
CE_EXPORT int ce_sim_split_slots(int live, int cap);
/* hipRTC compilations this engine has done (no counterpart in the reference).  A problem is
 * compiled with its constants written into the source; once the engine sees a second problem of
 * the same dimensions and instruction structure (the rounds of a serial-mode run, a search over
 * plants), it compiles that structure once more with the constants read from the device
 * descriptor and uses that kernel for every later problem of the structure.  At most 32 kernels of
 * the first kind are kept (least recently used first out). */
CE_EXPORT long long ce_engine_jit_compiles(ce_engine* eng);
/* Simulation-stage history (no counterpart in the reference): with calls > 0, every ce_evaluate
 * records HIP events on its stream around its simulation kernel, kept for the last `calls` calls
 * (0 turns it off; the call waits for the device).  ce_engine_sim_history waits for the recorded
 * events and returns up to n_max durations in ms, oldest first (the count, or -1). */
CE_EXPORT int ce_engine_set_sim_history(ce_engine* eng, int calls);
CE_EXPORT int ce_engine_sim_history(ce_engine* eng, double* ms, int n_max);
/* Per-kernel timing (no counterpart in the reference).  With profiling on, ce_evaluate records an
 * event after every kernel it launches; ce_engine_kernel_times then returns, for the last call, the
 * number of kernels k (<= n_max), their durations in ms and their names, comma-separated, in
 * launch order.  Off by default (the events add small gaps between kernels). */
CE_EXPORT int ce_engine_set_profiling(ce_engine* eng, int on);
CE_EXPORT int ce_engine_kernel_times(ce_engine* eng, char* names, int names_len, double* ms, int n_max);
/* Mixed-sweep counts of the last ce_evaluate in mixed precision (no counterpart in the reference):
 * frequency points recomputed with the FP64 elimination, and candidates whose whole sweep was.
 * Waits for the device.  Returns 1 when the last call was not mixed. */
CE_EXPORT int ce_engine_sweep_counts(ce_engine* eng, int* refined_points, int* fp64_candidates);

/* ---------------------------------------------------------------------------------------------
 * Batched evaluation (no counterpart in the reference, which evaluates one plant per call; its
 * serial mode, src/main.cu:441-558, loops over rounds on the host).  A batch holds many problems,
 * each a (config, plant, candidate count) of any dimensions, horizon and frequency count.  One
 * ce_batch_run evaluates every problem with one grouped launch per stage kernel and template
 * instance.  Inputs are contiguous over the problems in the order they were added: problem j's
 * gains rows start at float offset gains_off(j) and its INFO rows at info_off(j)
 * (ce_batch_offsets); info64_out (optional) has the INFO layout in double.
 * ce_batch_update replaces a problem's config/plant values (same row widths) between runs; the
 * device layout, the compiled kernels and the captured graph are kept while the dimensions,
 * frequency count, fit window and instruction structure stay the same. */
typedef struct ce_batch ce_batch;
CE_EXPORT ce_batch* ce_batch_create(int device, char* err, int errlen);
CE_EXPORT void ce_batch_destroy(ce_batch* b);
CE_EXPORT int ce_batch_clear(ce_batch* b);
/* Returns the problem's index, or -1 with a message in err. */
CE_EXPORT int ce_batch_add(ce_batch* b, const ce_config* cfg, const ce_plant* plant, int n, char* err, int errlen);
CE_EXPORT int ce_batch_update(ce_batch* b, int index, const ce_config* cfg, const ce_plant* plant, char* err,
                              int errlen);
CE_EXPORT int ce_batch_offsets(ce_batch* b, int index, long long* gains_off, long long* info_off);
/* use_graph: capture the stage chain into a hipGraph and replay it (default on); autotune: choose
 * each stage's plan (ce_batch_set_plan) and each launch's grid size (the waves a persistent launch
 * keeps in flight) by timing whole-chain passes on the next run (results do not depend on it);
 * profiling: record an event after every launch (disables the graph). */
CE_EXPORT int ce_batch_set_options(ce_batch* b, int use_graph, int autotune, int profiling);
CE_EXPORT int ce_batch_run(ce_batch* b, int precision, const float* d_gains, const float* d_info_in, float* d_info_out,
                           double* d_info64_out, void* stream, char* err, int errlen);
/* Stage plans of the batched chain (no counterpart in the reference).  eig_lin, eig_nom (the LIN_A and
 * nominal eigensolves): 0 one launch per matrix capacity (orders <= 4, 8, 16) with the default lanes
 * per candidate (4, 4, 16), 1 the same with the alternative lanes (8 at orders 5-16), 2 one fused launch
 * over every order.  sim: 0 one simulation launch per instruction structure, 1 one fused launch.  Rows
 * do not depend on the plan.  Fixed plans are kept by the autotuner; a negative value hands the choice
 * back to it.  ce_batch_plan reads the plans in use (eig_lin, eig_nom, sim). */
CE_EXPORT int ce_batch_set_plan(ce_batch* b, int eig_lin, int eig_nom, int sim);
CE_EXPORT int ce_batch_plan(ce_batch* b, int* plans);
/* Number of grouped launches, whether the last run replayed a graph, whether the grids are tuned. */
CE_EXPORT int ce_batch_info(ce_batch* b, int* launches, int* used_graph, int* tuned);
/* As ce_engine_kernel_times, for the last run with profiling on. */
CE_EXPORT int ce_batch_kernel_times(ce_batch* b, char* names, int names_len, double* ms, int n_max);

/* ---------------------------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference).  Jobs (config, plant, candidate rows on the host)
 * are submitted from any thread; a scheduler thread admits queued jobs in arrival order while
 * their estimated device footprint fits in mem_budget_bytes (<= 0: half the free device memory)
 * minus the jobs in flight, forms a batch of up to max_jobs_per_batch jobs on one of `slots`
 * streams (batches in flight overlap on the device) and copies the rows back.  A job's rows do not
 * depend on the other jobs of its batch.  precision applies to every job of the server. */
typedef struct ce_server ce_server;
CE_EXPORT ce_server* ce_server_create(int device, int precision, long long mem_budget_bytes, int slots,
                                      int max_jobs_per_batch, char* err, int errlen);
CE_EXPORT void ce_server_destroy(ce_server* s);
/* Returns the job id (>= 0) or -1.  gains: n x (2 statespace_dim + 1), info: n x info width. */
CE_EXPORT long long ce_server_submit(ce_server* s, const ce_config* cfg, const ce_plant* plant, const float* h_gains,
                                     const float* h_info, int n, char* err, int errlen);
/* 0 pending, 1 done (rows copied to the non-NULL outputs, job forgotten), -1 unknown, -2 failed. */
CE_EXPORT int ce_server_poll(ce_server* s, long long job, float* h_info_out, double* h_info64_out, char* err,
                             int errlen);
CE_EXPORT int ce_server_drain(ce_server* s, double timeout_s);
/* jobs done, latency p50 / p99 / mean / max (ms), candidate-evals/s (first submission to last
 * completion), batches, peak jobs in flight, peak batches in flight, peak bytes admitted, budget. */
CE_EXPORT int ce_server_stats(ce_server* s, double* out, int n_max);
CE_EXPORT int ce_server_reset_stats(ce_server* s);

/* ---------------------------------------------------------------------------------------------
 * Ranking and selection of a candidate block (no counterpart in the reference: its caller ranks
 * the rows of INFO_MX_OUT.bin by the score column against the requirements of README.md:22-35).
 * d_info: n x info_mx_width float rows on the device, as ce_evaluate writes them for cfg.
 * d_order (k entries, may be NULL when k = 0): the first k candidate indices by ascending score
 * (INFO column info_scoring); ties by index, -0 and +0 tie, scores that are not a number last.
 * d_pass (n bytes, optional): 1 when the candidate passes -- stable cell 1, score below
 * unstable_err, margin error 0 (GM, PM and b at or above their targets), overshoot at most the
 * overshoot spec and |yss / ref_amplitude - 1| at most steady_state_error -- else 0.
 * d_npass (optional): the number of passing candidates.  Enqueued on `stream`; device memory for
 * the sort is allocated and released on that stream. */
CE_EXPORT int ce_rank(const ce_config* cfg, const float* d_info, int n, int k, int32_t* d_order, uint8_t* d_pass,
                      int32_t* d_npass, void* stream, char* err, int errlen);

/* ce_rank without device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  Same arguments, results and errors as ce_rank, for n <= CE_RANK_NOALLOC_MAX_N;
 * a larger n is rejected with an error.  The sort's buffers are a static device array of 20 MiB that
 * every GPU the library runs on holds until the process ends.  All calls on one GPU share that
 * GPU's array: calls that overlap on one GPU (other streams or threads) race, and the caller must
 * serialise them. */
enum { CE_RANK_NOALLOC_MAX_N = 1 << 20 };
CE_EXPORT int ce_rank_noalloc(const ce_config* cfg, const float* d_info, int n, int k, int32_t* d_order,
                              uint8_t* d_pass, int32_t* d_npass, void* stream, char* err, int errlen);

/* Host-side helpers exported for tests: deterministic double math (fn: 0 exp, 1 log, 2 sin, 3 cos,
 * 4 atan2, 5 hypot, 6 log10, 7 sqrt) on the CPU and on the device (7: the device eigensolver's
 * square root against the host's IEEE sqrt). */
CE_EXPORT int ce_detmath_host(int fn, const double* x, const double* y, double* out, int n);
CE_EXPORT int ce_detmath_device(int fn, const double* d_x, const double* d_y, double* d_out, int n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CONTROLLERENGINE_H */
