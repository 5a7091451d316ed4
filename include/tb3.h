/*
 * tb3.h -- C ABI of the MI355X three-body (Axilrod-Teller) force engine.
 *
 * One context owns one GPU's device buffers for one particle system.  All
 * calls are made from one host thread per context, are ordered on the
 * context's HIP stream and are synchronous on return unless stated otherwise.
 * No exceptions cross this boundary: every call returns a tb3_status and
 * tb3_last_error() gives the message of the last failure.
 *
 * Citations name the reference interface each entry point replaces
 * (ThreeBodyMDAlgorithms, paths relative to its src/ directory).  At one
 * process the reference runs Simulation::Start (simulation/Simulation.cpp:28-81):
 * predictor -> Algorithm::SimulationStep -> Update, one buffer triplet per
 * step (P3BCA.cpp:31-32, 89; AUTA.cpp:268-302; NATA.cpp:112-130).
 */
#ifndef TB3_H
#define TB3_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TB3_OK = 0,
    TB3_EINVAL = 1,       /* bad argument or call order */
    TB3_ENOMEM = 2,       /* device or host allocation failed */
    TB3_EHIP = 3,         /* a HIP runtime call failed */
    TB3_EUNSUPPORTED = 4, /* valid request that tb3 does not serve */
    TB3_ENODEV = 5        /* no usable GPU */
} tb3_status;

/* Same numbering as the reference's AlgorithmType (utility/enums.hpp:1). */
typedef enum { TB3_NATA = 0, TB3_P3BCA = 1, TB3_AUTA = 2 } tb3_algo;

typedef enum {
    /* Bit-identical to the reference's one-process sweep: the reference's
     * per-triplet expression (potential/AxilrodTeller.cpp:22-93) and its
     * per-particle accumulation order (algorithm/Algorithm.cpp:72-139, 246-248). */
    TB3_FP64_EXACT = 0,
    /* FP64 throughout, squared-distance formulation, any accumulation order. */
    TB3_FP64 = 1,
    /* Per-triplet arithmetic in FP32 on pair differences; FP64 differences,
     * cutoff test, accumulation and integration. */
    TB3_FP32_MIXED = 2
} tb3_precision;

typedef struct tb3_ctx tb3_ctx;

/* Library version string, e.g. "tb3 1.0.0 gfx950". */
const char *tb3_version(void);

/* Number of visible GPUs (hipGetDeviceCount). */
tb3_status tb3_device_count(int *count);

/* Replaces the construction of Simulation + topology + decomposition
 * (main.cpp:32-105, Simulation.hpp:28-31) for one GPU.  Creates a context on
 * `device` with its own stream. */
tb3_status tb3_create(int device, tb3_ctx **out);
tb3_status tb3_destroy(tb3_ctx *ctx);

/* Message of the last failed call on ctx (never NULL; "" after success).
 * tb3_last_error(NULL) returns the message of the last failed tb3_create. */
const char *tb3_last_error(const tb3_ctx *ctx);

/* Replaces createNATAContext / createP3BCAContext / createAUTAContext
 * (main.cpp:32-105) and the AxilrodTeller(nu) constructor
 * (potential/AxilrodTeller.cpp:3).
 *   algo      TB3_NATA, TB3_P3BCA or TB3_AUTA.  NATA and AUTA evaluate every
 *             triplet (Algorithm.cpp:31); P3BCA applies the periodic cutoff
 *             test (Algorithm.cpp:94-100) when cutoff > 0.
 *   nu        Axilrod-Teller coefficient (the reference passes 1.0, main.cpp:41,66,91).
 *   cutoff    P3BCA cutoff distance; <= 0 disables the test as in Algorithm.cpp:94.
 *   domain_L  periodic domain size for the cutoff test, or NULL to take the
 *             reference's rule at tb3_set_particles: the bounding box of the
 *             initial particles widened by 1e-4 on each side
 *             (RegularGridDecomposition.cpp:24-28, 135-166).
 *   precision a tb3_precision value.
 * May be called again at any time; a NULL domain_L recomputes the box from the
 * particles currently held. */
tb3_status tb3_configure(tb3_ctx *ctx, int algo, double nu, double cutoff, const double *domain_L, int precision);

/* Replaces Utility::getParticlesFromCSV + the decomposition's copy-in
 * (utility/utility.cpp:14-24; DomainDecomposition.cpp:7-8).  Host arrays,
 * caller-owned, copied to the device: pos/vel/acc are n x 3 row-major (x,y,z
 * per particle), mass is n, id is n (may be NULL: ids 0..n-1).  Forces are
 * reset to zero.  Particle order is the sweep order, as the reference keeps
 * the CSV row order (AtomDecomposition.cpp:34-36, RegularGridDecomposition.cpp:41-48). */
tb3_status tb3_set_particles(tb3_ctx *ctx, int64_t n, const int32_t *id, const double *pos, const double *vel,
                             const double *acc, const double *mass);

/* Replaces DomainDecomposition::GetMyParticles (DomainDecomposition.cpp:7) as
 * used by the CSV writer (Simulation.cpp:99-133).  Copies the current state
 * back in the input order; any pointer may be NULL.  f is the summed force of
 * the last sweep: +dE/dx, the reference's sign convention. */
tb3_status tb3_get_particles(const tb3_ctx *ctx, int64_t n, int32_t *id, double *pos, double *vel, double *acc,
                             double *mass, double *f);

/* Replaces one iteration of Simulation::Start (Simulation.cpp:38-48):
 * UpdatePredictorStage (structs.hpp:40-55) -> SimulationStep (ResetForces,
 * sweep, SumUpParticles; Algorithm.cpp:72-139, 237-271) -> Update
 * (structs.hpp:57-99).  *n_buffer receives the buffer-triplet count (1 at one
 * GPU) and *n_triplets the accepted particle triplets (Algorithm.hpp:60);
 * both may be NULL. */
tb3_status tb3_step(tb3_ctx *ctx, double dt, const double g[3], uint64_t *n_buffer, uint64_t *n_triplets);

/* `steps` iterations of tb3_step without host round trips between them
 * (Simulation.cpp:30 loop).  counts (may be NULL) receives `steps` accepted
 * triplet counts (summed over the replicas of a batch).  From two steps on,
 * the steps are replays of one captured HIP graph of a step, with one host
 * synchronisation per call; a step whose neighbour staging overflows is
 * finished on the host and the replays resume, with the same results as
 * tb3_step. */
tb3_status tb3_run(tb3_ctx *ctx, int steps, double dt, const double g[3], uint64_t *counts);

/* Replica batches: independent particle systems of uneven size swept in the
 * same launches (the cell kernel's device queue hands out cells of every
 * replica).  No counterpart in the reference, which runs one system per
 * process: each replica behaves as its own `mpiexec -n 1 ./main` run of
 * P3BCA (Simulation.cpp:28-81), with its own periodic domain (the
 * RegularGridDecomposition.cpp:135-166 rule on its own particles, or the
 * domain_L of tb3_configure).
 *
 * tb3_batch_set splits the particles of tb3_set_particles into nrep
 * replicas: replica r holds rows offsets[r] .. offsets[r+1]-1
 * (offsets[0] = 0, offsets[nrep] = n, nondecreasing).  Needs P3BCA with a
 * positive cutoff (TB3_EUNSUPPORTED otherwise); nrep = 1 restores one system,
 * and tb3_set_particles always does. */
tb3_status tb3_batch_set(tb3_ctx *ctx, int32_t nrep, const int64_t *offsets);
/* tb3_run with per-replica counts: counts (may be NULL) is steps x nrep. */
tb3_status tb3_batch_run(tb3_ctx *ctx, int steps, double dt, const double g[3], uint64_t *counts);
/* Per-replica accepted-triplet counts of the last sweep (nrep entries). */
tb3_status tb3_batch_counts(tb3_ctx *ctx, int32_t nrep, uint64_t *counts);
/* Per-replica energies E[nrep] of the current positions (the held forces are
 * untouched); counts (may be NULL) receives the per-replica triplet counts. */
tb3_status tb3_batch_energy(tb3_ctx *ctx, int32_t nrep, double *E, uint64_t *counts);

/* SimulationStep alone (no predictor or update): reset, sweep and role sum
 * (P3BCA.cpp:516-805 / AUTA.cpp:236-383 at one process).  The force is read
 * back with tb3_get_particles. */
tb3_status tb3_compute_forces(tb3_ctx *ctx, uint64_t *n_triplets);

/* Energy forward: E = sum over the accepted triplets of
 * nu (1 + 3 cos g1 cos g2 cos g3) / (r12 r13 r23)^3 at the current positions.
 * No counterpart in the reference (it never evaluates E); its gradient is the
 * reference's f (AxilrodTeller.cpp:22-130). */
tb3_status tb3_energy(tb3_ctx *ctx, double *E, uint64_t *n_triplets);

/* Energy adjoint: E, dE/dx (= f, n x 3 host buffer, may be NULL) and
 * dE/dnu (may be NULL).  E is linear in nu, so dE/dnu is the energy at
 * nu = 1: E / nu, or at nu = 0 a second (energy-only) pass at nu = 1. */
tb3_status tb3_energy_grad(tb3_ctx *ctx, double *E, double *dEdx, double *dEdnu, uint64_t *n_triplets);

/* Device-pointer forms for framework bindings (torch).  d_pos is an n x 3
 * row-major device array; results land in device memory; `stream` is a
 * hipStream_t (NULL: the context's stream).  They use the context's
 * algorithm, cutoff, domain and precision but not its particles, and leave
 * the context's particle state untouched.  The work is ordered on `stream`
 * after what the caller queued there before, and is complete on return: the
 * context's status block and scratch memory, which its next call reuses, are
 * read by then.  n_triplets (host) may be NULL. */
tb3_status tb3_energy_dev(tb3_ctx *ctx, int64_t n, const double *d_pos, double nu, double *d_E, void *stream,
                          uint64_t *n_triplets);
tb3_status tb3_energy_grad_dev(tb3_ctx *ctx, int64_t n, const double *d_pos, double nu, double *d_E,
                               double *d_dEdx, void *stream, uint64_t *n_triplets);

/* tb3_energy_grad_dev without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as tb3_energy_grad_dev (E and
 * dE/dx of n positions in device memory, ordered on `stream`, complete on
 * return), for the cutoff algorithm (P3BCA, cutoff > 0) in TB3_FP64 or
 * TB3_FP32_MIXED on one particle system with a known periodic domain
 * (tb3_configure's domain_L, or the domain of the particles last set);
 * other configurations return TB3_EUNSUPPORTED or TB3_EINVAL.  The context
 * supplies the device and those parameters only: the call calls no hipMalloc
 * or hipFree and uses none of the context's device buffers.  Its working
 * memory (cell list, staged positions, forces, status block) is a static
 * device array of TB3_NOALLOC_SCRATCH_BYTES per GPU, held from the first
 * context on that GPU until the process ends; a system whose working set
 * does not fit (about 88 bytes per particle and 16 per cell, plus ~100 KB:
 * some 350 000 particles at density 2 and cutoff 2) returns TB3_EINVAL.
 * Every call on one GPU shares that array, which keeps the last call's
 * contents; calls that overlap on one GPU (from several threads or
 * contexts) race and must be serialised by the caller. */
#define TB3_NOALLOC_SCRATCH_BYTES (32u << 20)
tb3_status tb3_energy_grad_dev_noalloc(tb3_ctx *ctx, int64_t n, const double *d_pos, double nu, double *d_E,
                                       double *d_dEdx, void *stream, uint64_t *n_triplets);

/* Autotuner.  No counterpart in the reference.  For the cutoff algorithm,
 * times a force sweep of the current particles (on a copy; the state is
 * untouched) for each cell width / cutoff ratio in {1, 1.15, 1.3, 1.5, 1.75,
 * 2} and keeps the fastest; then, when the dense-owner kernel runs (owners
 * whose neighbourhood outgrows the cell kernel's staging), times it with 4, 8
 * and 16 waves per owner and keeps the fastest.  Reports (each output may be
 * NULL) the best sweep time in ms (HIP events, minimum of `reps` runs after
 * one warm-up), the ratio, and the cell kernel's resident workgroups per CU
 * (0 when the per-owner kernel is in charge); tb3_get_tuning reports every
 * knob.  The all-triplet sweeps have nothing to tune and return at once. */
tb3_status tb3_autotune(tb3_ctx *ctx, int reps, double *best_ms, double *cell_ratio, int *blocks_per_cu);

/* Each particle's force scale S_p: the sum over p's triplets of the length of
 * p's contribution, |f_p| <= S_p (the condition of the force sum; FP64, from
 * a sweep of a copy of the held positions).  A sum of terms computed with
 * relative error e each is off by at most e S_p however much the terms
 * cancel, so S_p is the yardstick of the FP32-mixed path's force error.
 * scale holds n doubles.  P3BCA with a positive cutoff on one system;
 * TB3_EUNSUPPORTED otherwise.  No counterpart in the reference. */
tb3_status tb3_force_scale(tb3_ctx *ctx, double *scale);

/* The tuned knobs (each output may be NULL): cell width / cutoff ratio, the
 * cell kernel's resident workgroups per CU, and the dense-owner kernel's
 * waves per owner.  No counterpart in the reference. */
tb3_status tb3_get_tuning(const tb3_ctx *ctx, double *cell_ratio, int *blocks_per_cu, int *dense_waves);

/* Sets the knobs directly (to reuse a tuning, or to test each setting):
 * cell_ratio in [1, 4], dense_waves 4, 8 or 16; TB3_EINVAL otherwise.  No
 * counterpart in the reference. */
tb3_status tb3_set_tuning(tb3_ctx *ctx, double cell_ratio, int dense_waves);

/* The slotted cell layout of tb3_run's replayed steps (fast cutoff sweep on
 * one system, grids of up to 8192 cells, no dense-owner kernel): each cell
 * keeps `slots` slots that the binning pass fills directly, so that a step
 * runs no scatter pass between the binning and the sweep.  slots = -1 picks
 * three times the mean cell count (at least 32), 0 turns the layout off,
 * 1 .. 256 sets it; TB3_EINVAL otherwise.  A step in which some cell gets
 * more particles than slots is finished on the contiguous layout, which the
 * rest of the run (and later runs on the same grid) then keeps.
 * tb3_get_cell_slots reports the setting and the slots per cell that the
 * last captured step graph used (0: contiguous, or a cell has outgrown its
 * slots since).  No counterpart in the reference, whose cell lists are host
 * containers. */
tb3_status tb3_set_cell_slots(tb3_ctx *ctx, int slots);
tb3_status tb3_get_cell_slots(const tb3_ctx *ctx, int *requested, int *in_use);

/* Per-phase timing with HIP events recorded on the context's stream around
 * each phase of tb3_step, tb3_run (which runs its steps one pass at a time
 * while profiling is on), tb3_compute_forces and the device-pointer energy
 * calls.  No counterpart in
 * the reference beyond its per-method timers (PROFILE_3BMDA, Algorithm.cpp:57-63,
 * 145-162).  tb3_set_profiling(ctx, 1) enables and clears the accumulators.
 * tb3_get_profile fills up to n values:
 *   [0] passes timed, [1] predictor ms, [2] cell + neighbour-list ms,
 *   [3] force kernel ms (the sweep kernel alone), [4] role sum + update ms,
 *   [5] accepted triplets summed over the timed steps. */
tb3_status tb3_set_profiling(tb3_ctx *ctx, int enable);
tb3_status tb3_get_profile(const tb3_ctx *ctx, double *out, int64_t n);

/* Diagnostic: element-wise sqrt(a) and a / b in device FP64, as the
 * bit-exact path evaluates them, for checking that both round correctly, and
 * the raw hardware reciprocal-square-root estimate of a (v_rsq_f64, the seed
 * of the fast path's 1/sqrt; rsq_a may be NULL).  Host arrays of n doubles.
 * No counterpart in the reference. */
tb3_status tb3_selftest_fp64(tb3_ctx *ctx, int64_t n, const double *a, const double *b, double *sqrt_a,
                             double *a_div_b, double *rsq_a);

/* Diagnostic: the bit-exact path's in-range sequences for sqrt(a) and a / b
 * (the compiler's IEEE expansions without their range-scaling steps, used on
 * operands within 2^-500 .. 2^500), element-wise, for checking that they give
 * the IEEE results' bits on such operands.  No counterpart in the reference. */
tb3_status tb3_selftest_fp64_inrange(tb3_ctx *ctx, int64_t n, const double *a, const double *b, double *sqrt_a,
                                     double *a_div_b);

#ifdef __cplusplus
}
#endif

#endif
