/*
 * sns_hip.h -- thin C-ABI between the C host layer and the gfx950 kernels
 * (libsns_hip.so).  Plain C types only: sizes, scalars and device pointers.
 *
 * One snsg_ctx is one simulation instance on one GPU: it owns a HIP stream,
 * the device copies of the grid metrics, the time step, the Poisson and
 * Crank-Nicolson scratch and the FFT plan.  Field arrays are device memory
 * in the reference's layouts (include/arrays/fluid.h, temperature.h of the
 * reference; see the layout comment in include/sns_arrays.h), allocated with
 * snsg_malloc and passed explicitly, so every entry point below maps onto
 * one reference function with the same arguments.
 *
 * Return values: 0 on success, otherwise a non-zero code (a hipError_t value
 * or one of SNSG_E_*); snsg_last_error() gives a message.  Work is enqueued
 * on the instance's stream; snsg_synchronize() waits for it.
 *
 * Entry points and the reference interface each one replaces:
 *   snsg_fluid_update_boundaries_{ux,uy,p}   include/fluid.h:16-18
 *   snsg_temperature_update_boundaries_temp  include/temperature.h:20
 *   snsg_temperature_compute_force           include/temperature.h:17
 *   snsg_temperature_update_temp             include/temperature.h:18
 *   snsg_fluid_update_velocity               include/fluid.h:20
 *   snsg_fluid_compute_potential             include/fluid.h:21
 *   snsg_fluid_correct_velocity              include/fluid.h:22
 *   snsg_fluid_update_pressure               include/fluid.h:23
 *   snsg_param_decide_dt                     include/param.h:59 (device reduction part)
 *   snsg_transpose                           include/parallel.h:21 (one rank)
 *   snsg_integrate                           src/main.c:79-101 (static integrate())
 *   snsg_advance                             src/main.c:28-35 (time-loop body)
 */
#ifndef SNS_HIP_H
#define SNS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SNSG_E_ARG      (-1)   /* invalid argument */
#define SNSG_E_SIZE     (-2)   /* grid size not supported by the kernels */
#define SNSG_E_NOMEM    (-3)   /* device allocation failed */
#define SNSG_E_TIMEOUT  (-4)   /* a wait ended before its job */

typedef struct snsg_ctx snsg_ctx;

typedef struct {
  int itot, jtot;                 /* cells in x (walls) and y (periodic) */
  double ly;                      /* domain length in y (lx is 1) */
  double Ra, Pr;
  int with_temperature, with_thermal_forcing;
  int implicitx, implicity;       /* diffusive treatment */
  double alpha[3], beta[3], gamma[3];  /* RK3 coefficients */
} snsg_config;

/* Device-resident state advanced by the fused step (snsg_integrate).  The
 * halo'd fields keep the reference's index order with padded rows: element
 * (I, J) of ux / ux_star is ux[J*ld_ux + (I-1)], of uy, p, psi, temp, temp_alt,
 * uy_star is uy[J*ld_c + I]; ld_ux and ld_c are multiples of 16 and cell
 * I = 1 of every row is 128-byte aligned.  The RK histories have the
 * reference layouts (no halo, no padding).  temp and temp_alt are exchanged
 * by every substage (the caller's pointers are updated).  snsg_state_upload /
 * snsg_state_download convert from / to the reference layouts. */
typedef struct {
  double *ux, *uy, *p, *psi;
  double *temp, *temp_alt;
  double *srcuxa, *srcuya, *srctempa;   /* RK history of the explicit terms */
  double *ux_star, *uy_star;            /* predictor output, (jtot+2) rows */
  double *p_alt;                        /* pressure between the substages whose
                                           correction the next predictor applies
                                           (layout of p) */
  long ld_ux, ld_c;                     /* row pitches in elements */
} snsg_state;

/* field ids of snsg_state_upload / snsg_state_download */
enum {
  SNSG_UX = 0, SNSG_UY, SNSG_P, SNSG_PSI, SNSG_TEMP, SNSG_UX_STAR, SNSG_UY_STAR,
  SNSG_SRCUXA, SNSG_SRCUYA, SNSG_SRCTEMPA
};

/* ---- instance management */
int snsg_create(int device, const snsg_config *cfg, const double *dxf, const double *dxc,
                snsg_ctx **ctx);
int snsg_destroy(snsg_ctx *ctx);
int snsg_set_treatment(snsg_ctx *ctx, int implicitx, int implicity);
int snsg_set_dt(snsg_ctx *ctx, double dt);
int snsg_synchronize(snsg_ctx *ctx);
const char *snsg_last_error(void);
int snsg_device_count(int *count);

/* ---- memory (zero-initialised device allocations) */
int snsg_malloc(snsg_ctx *ctx, size_t bytes, void **dptr);
int snsg_free(snsg_ctx *ctx, void *dptr);
int snsg_upload(snsg_ctx *ctx, void *dst, const void *src, size_t bytes);
int snsg_download(snsg_ctx *ctx, void *dst, const void *src, size_t bytes);
/* device-to-device copy on the instance's stream by a plain streaming
 * kernel; bytes and both addresses multiples of 16.  No counterpart in the
 * reference: state snapshots, and the known-size copy that calibrates
 * hardware traffic counters. */
int snsg_copy(snsg_ctx *ctx, void *dst, const void *src, size_t bytes);
/* the same copy by a tiled kernel: each thread moves `loads` (1, 2, 4 or 8)
 * 16-byte words, all loads issued before the stores; nontemporal != 0 marks
 * them nontemporal.  No reference counterpart: its shape sweep measures the
 * device's practical copy bandwidth, the ceiling quoted beside the HBM
 * peak. */
int snsg_copy_shaped(snsg_ctx *ctx, void *dst, const void *src, size_t bytes, int loads,
                     int nontemporal);

/* ---- boundary values */
int snsg_fluid_update_boundaries_ux(snsg_ctx *ctx, double *ux);
int snsg_fluid_update_boundaries_uy(snsg_ctx *ctx, double *uy);
int snsg_fluid_update_boundaries_p(snsg_ctx *ctx, double *p);
int snsg_temperature_update_boundaries_temp(snsg_ctx *ctx, double *temp);

/* ---- temperature.  The reference copies SRCTEMPB <- SRCTEMPA before
 * recomputing SRCTEMPA (src/temperature/update.c:31); here srctempb must
 * already hold the previous SRCTEMPA (the host layer exchanges the two
 * pointers) and srctempa is overwritten. */
int snsg_temperature_compute_force(snsg_ctx *ctx, const double *temp, double *tempforcex);
int snsg_temperature_update_temp(snsg_ctx *ctx, int rkstep, const double *ux, const double *uy,
                                 double *temp, double *srctempa, const double *srctempb,
                                 double *srctempg);

/* ---- fluid (same convention for the ux / uy source histories) */
int snsg_fluid_update_velocity(snsg_ctx *ctx, int rkstep, double *ux, double *uy,
                               const double *p, const double *tempforcex, double *srcuxa,
                               const double *srcuxb, double *srcuxg, double *srcuya,
                               const double *srcuyb, double *srcuyg);
int snsg_fluid_compute_potential(snsg_ctx *ctx, int rkstep, const double *ux, const double *uy,
                                 double *psi);
int snsg_fluid_correct_velocity(snsg_ctx *ctx, int rkstep, const double *psi, double *ux,
                                double *uy);
int snsg_fluid_update_pressure(snsg_ctx *ctx, int rkstep, const double *psi, double *p);

/* ---- time step: *dt_adv receives min(5e-2, min DXC(i)/max(|UX|,1e-8),
 * min dy/max(|UY|,1e-8)), the advective limit of src/param/decide_dt.c:25-45
 * before its safety factor.  Blocks until the value is on the host. */
int snsg_param_decide_dt(snsg_ctx *ctx, const double *ux, const double *uy, double *dt_adv);

/* ---- global transpose on one rank: recv[i*g_jsize + j] = send[j*g_isize + i],
 * elemsize 4, 8 or 16 bytes (src/parallel/transpose.c:6-85) */
int snsg_transpose(snsg_ctx *ctx, int g_isize, int g_jsize, size_t elemsize, const void *send,
                   void *recv);

/* ---- fused time integration on instance-owned state.
 * snsg_integrate: one RK3 step (src/main.c:79-101) with the dt set by snsg_set_dt.
 * snsg_advance: nsteps iterations of the reference's time loop body
 * (src/main.c:28-35): the time-step decision of src/param/decide_dt.c for
 * the configured treatment, evaluated on the device, then one RK3 step; the
 * run time accumulates on the device (snsg_get_time / snsg_set_time).  No
 * host synchronisation happens inside. */
int snsg_state_alloc(snsg_ctx *ctx, snsg_state *st);
int snsg_state_free(snsg_ctx *ctx, snsg_state *st);
int snsg_integrate(snsg_ctx *ctx, snsg_state *st);
int snsg_advance(snsg_ctx *ctx, snsg_state *st, int nsteps);
int snsg_get_time(snsg_ctx *ctx, double *time, double *dt);
/* copy one state field from / to a host array in the reference layout
 * (blocking) */
int snsg_state_upload(snsg_ctx *ctx, snsg_state *st, int field, const double *host);
int snsg_state_download(snsg_ctx *ctx, const snsg_state *st, int field, double *host);
/* ---- monitors (src/logging.c:36-272, src/analyses.c:8-121) evaluated on the
 * device; out[11] = { max|div|, sum div, sum ux dV, sum uy dV, energies of
 * ux, uy, T, Nusselt numbers from the wall flux, the energy injection, the
 * kinetic and the thermal dissipation }.  The maximum is exact, the sums are
 * reduced in a fixed order (reproducible, not in the reference's sequential
 * order).  temp may be NULL without the temperature equation.  Blocking. */
int snsg_diagnostics(snsg_ctx *ctx, const double *ux, const double *uy, const double *temp,
                     double *out);
int snsg_state_diagnostics(snsg_ctx *ctx, const snsg_state *st, double *out);
/* snsg_diagnostics without allocating: no counterpart in the reference;
 * for callers that cannot allocate device or pinned host memory (code under
 * HIP graph capture, tight serving loops).  Same inputs and the same 11
 * results bit for bit, written to out_dev, a device array of 11 doubles;
 * asynchronous on the instance's stream (no host synchronisation).  The
 * per-workgroup partial sums live in a file-scope device array of the
 * library, 2048 x 11 doubles (176 KiB) per GPU, held until the process
 * ends; a grid needing more partial slots is refused with SNSG_E_SIZE (the
 * monitors use at most 2048, so none is today).  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other
 * threads) race and must be serialised by the caller. */
int snsg_diagnostics_noalloc(snsg_ctx *ctx, const double *ux, const double *uy,
                             const double *temp, double *out_dev);
/* ---- statistics (src/statistics/collect.c:10-75): ux1 += ux, ux2 += ux^2
 * (I = 1..itot+1), uy1/uy2 and temp1/temp2 likewise (I = 0..itot+1), rows
 * 1..jtot; accumulators in the reference layouts (include/arrays/statistics.h
 * of the reference), fields in the reference layouts (snsg_statistics_collect)
 * or the state's (snsg_state_statistics_collect).  temp* may be NULL without
 * the temperature equation. */
int snsg_statistics_collect(snsg_ctx *ctx, const double *ux, const double *uy, const double *temp,
                            double *ux1, double *ux2, double *uy1, double *uy2, double *temp1,
                            double *temp2);
int snsg_state_statistics_collect(snsg_ctx *ctx, const snsg_state *st, double *ux1, double *ux2,
                                  double *uy1, double *uy2, double *temp1, double *temp2);
/* Seeded initial temperature of the state on the device: interior
 * T = A*(2U-1) with U from the same counter hash as the host layer's initial
 * condition (host/temperature.c: sns_initial_temperature, non-linear
 * profile; it replaces the rand() field of src/temperature/init.c:39-47),
 * then the boundary values of src/temperature/boundary_conditions.c. */
int snsg_state_init_temperature(snsg_ctx *ctx, snsg_state *st, uint64_t seed, double amplitude);
/* Device bytes that an instance of cfg's size and treatment holds with its
 * state once it has stepped (snsg_create + snsg_state_alloc + the fused
 * step's scratch), for admission against a memory budget. */
int snsg_instance_bytes(const snsg_config *cfg, size_t *bytes);
/* 1 when the grid is uniform bit for bit away from the walls (the kernels
 * then take the spacings of the interior columns as constants), else 0 */
int snsg_uniform_grid(const snsg_ctx *ctx, int *uniform);
/* Device bytes an instance holds now (its state and scratch included) */
int snsg_memory(const snsg_ctx *ctx, size_t *bytes);
/* ---- launch shapes (no counterpart in the reference, whose FFTW plans and
 * sequential solves have none; for the autotuner and callers that tune per
 * grid size).  A field left 0 keeps the built-in rule.  An instance takes
 * its shapes from snsg_set_launch, else from the process-wide table
 * (snsg_tuning_add, keyed by itot and jtot) as it stands at snsg_create,
 * else from the rules.  The shapes reorder floating-point operations, not
 * the algorithm: every choice stays within the parity tolerance of the
 * reference, and instances with equal shapes return the same bits alone,
 * batched or served (a batch's shape classes include the shapes). */
typedef struct {
  int fft_n1;           /* length of the first factor of the four-step y FFT
                           (a divisor of jtot; jtot/fft_n1 the second) */
  int ysolve_chunk;     /* rows per chunk of the partitioned CN-y solve: a
                           divisor of jtot in 4..32, with 4..513 chunks;
                           1: no partition (column solves on the transposed
                           increments) */
  int ysolve_cols;      /* columns per workgroup of its interface solve: 16,
                           32 or 64 */
  int ysolve_recompute; /* its chunk interiors: 1 solved again in the apply
                           kernel, 2 stored and read back */
  int split_corr;       /* 1: every substage ends with the correction kernel;
                           0 (the rule): with implicit x and explicit y the
                           predictors of substages 2 and 3 apply the previous
                           substage's correction to the fields as they load
                           them (the same operations, so the same bits) */
} snsg_launch;
/* SNSG_E_ARG for a value outside the ranges above, SNSG_E_SIZE for one the
 * grid does not admit (a factor that does not divide jtot or whose FFT pass
 * does not fit the kernels' LDS tile).  Host-only: needs no GPU. */
int snsg_launch_check(int itot, int jtot, const snsg_launch *l);
/* Replaces all of the instance's shapes (0: the rule).  A new chunk length
 * takes effect at the next fused step, which sizes the y solve's scratch
 * again.  SNSG_E_ARG while the instance is in a batch. */
int snsg_set_launch(snsg_ctx *ctx, const snsg_launch *l);
/* The shapes in effect, rules resolved; ysolve_cols and ysolve_recompute
 * are 0 where the rule decides at launch (it depends on the batch size). */
int snsg_get_launch(const snsg_ctx *ctx, snsg_launch *l);
/* Adds (or replaces) the table entry of (itot, jtot) after
 * snsg_launch_check; snsg_tuning_clear empties the table.  Host-only. */
int snsg_tuning_add(int itot, int jtot, const snsg_launch *l);
int snsg_tuning_clear(void);
/* ---- batches (no counterpart in the reference, which runs one instance per
 * process).  Many instances advance together: each kernel of a step is
 * launched once per shape class (the members with the same itot, jtot,
 * treatment and temperature equation) for all of the class's members, each
 * class on a stream of its own, its step replayed from a HIP graph.  Each
 * member's fields, dt and time are bitwise those of snsg_advance on the
 * instance alone.  The batch keeps the ctx and state pointers it is given:
 * the instances must outlive it, and while it exists their states are
 * advanced through snsg_batch_advance only (a member advanced on its own
 * must go an even number of steps, or the batch refuses to advance).
 * Other calls on a member (uploads, downloads, diagnostics) are ordered
 * with the batch's work through the member's own stream.  Asynchronous like
 * snsg_advance; snsg_batch_synchronize waits for the batch's work. */
typedef struct snsg_batch snsg_batch;
#define SNSG_BATCH_EAGER 1 /* flags: launch every step's kernels instead of replaying graphs */
int snsg_batch_create(int n, snsg_ctx *const *ctx, snsg_state *const *st, int flags,
                      snsg_batch **out);
/* nsteps iterations of the time loop body (as snsg_advance) for every member */
int snsg_batch_advance(snsg_batch *b, int nsteps);
int snsg_batch_synchronize(snsg_batch *b);
int snsg_batch_classes(const snsg_batch *b, int *nclasses);
int snsg_batch_destroy(snsg_batch *b);
/* ---- serving engine (no counterpart in the reference).  Simulation jobs
 * of any size arrive from any thread (snsg_server_submit); an engine thread
 * admits them in arrival order while a slot is free and their device
 * memory (snsg_instance_bytes) fits the budget, keeps every admitted job's fields
 * resident and advances all admitted jobs continuously: the jobs of one shape
 * class form a lane (as in a batch) that starts its next step as soon as the
 * previous one is done, jobs joining and leaving it between steps.  Each
 * job's fields, dt and time are bitwise those of snsg_advance(nsteps) on the
 * job alone.  Instances of finished jobs are reused by later jobs of the same
 * size and treatment, and freed when the budget needs their memory.  The
 * per-step latency of a job-step is the duration of its lane's step on the
 * device (HIP events). */
typedef struct snsg_server snsg_server;
typedef struct {
  snsg_config cfg;           /* size, parameters, treatment */
  const double *dxf, *dxc;   /* grid (as for snsg_create), copied at submission */
  uint64_t seed;             /* initial temperature (snsg_state_init_temperature) */
  double amplitude;
  int nsteps;                /* time-loop iterations (as snsg_advance) */
  int keep;                  /* keep the result for snsg_server_download until released */
} snsg_job;
typedef struct {
  long submitted, admitted, completed, active, queued;
  long active_peak;          /* most jobs on the GPU at once */
  long steps;                /* job-steps completed since the last reset */
  double cell_steps;         /* their itot*jtot summed */
  double busy_s;             /* from the first admission (or the reset) to the last step */
  double p50_ms, p99_ms, max_ms;  /* per-step latency over those job-steps */
  size_t mem_used, mem_peak; /* device bytes of the instances held (active, kept, reusable) */
  int lanes;
} snsg_server_stats;
#define SNSG_JOB_QUEUED 0
#define SNSG_JOB_ACTIVE 1
#define SNSG_JOB_DONE 2
#define SNSG_JOB_RELEASED 3
/* mem_budget: device bytes of the admitted jobs' instances; max_active:
 * at most this many jobs on the GPU at once (slots; 0: memory alone) */
int snsg_server_create(int device, size_t mem_budget, int max_active, snsg_server **out);
int snsg_server_submit(snsg_server *s, const snsg_job *job, int *id);
int snsg_server_wait(snsg_server *s, int id, double timeout_s);
int snsg_server_drain(snsg_server *s, double timeout_s);
int snsg_server_status(snsg_server *s, int id, int *state, int *steps_done);
int snsg_server_download(snsg_server *s, int id, int field, double *host);
int snsg_server_job_time(snsg_server *s, int id, double *time, double *dt);
int snsg_server_release(snsg_server *s, int id);
int snsg_server_stats_get(snsg_server *s, snsg_server_stats *out);
int snsg_server_reset_stats(snsg_server *s);
/* stops the engine (after the steps in flight) and frees every job */
int snsg_server_destroy(snsg_server *s);
/* advective time-step bound of the state's ux, uy (as snsg_param_decide_dt) */
int snsg_state_decide_dt(snsg_ctx *ctx, const snsg_state *st, double *dt_adv);
int snsg_set_time(snsg_ctx *ctx, double time);

/* ---- self-check of the fused kernels' division scheme (no counterpart in the
 * reference): q_cr[k] = a[k]/b[k] formed from r[k] = 1/b[k] by the FMA
 * correction sequence, q_ieee[k] = a[k]/b[k] by the IEEE division; device
 * arrays of n entries */
int snsg_check_division(snsg_ctx *ctx, int n, const double *a, const double *b, const double *r,
                        double *q_cr, double *q_ieee);

/* ---- per-kernel timing with HIP events on the instance's stream.  While
 * enabled, every launch of the kernel group named `filter` (NULL: all)
 * is bracketed by two events; snsg_timing_read returns the number of such
 * launches and their summed duration in milliseconds. */
int snsg_timing_enable(snsg_ctx *ctx, const char *filter);
int snsg_timing_disable(snsg_ctx *ctx);
int snsg_timing_read(snsg_ctx *ctx, int *launches, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif /* SNS_HIP_H */
