/*
 * tb3_oracle.h -- CPU oracle for the Axilrod-Teller triplet hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain C++ restatement of the
 * reference's single-process algorithm (ThreeBodyMDAlgorithms, `mpiexec -n 1`).
 * It is used by tests/ as the parity checker, by __graft_entry__.smoke() as the
 * checker, and by bench.py's cpu_baseline leg.  The product (libtb3.so) never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference's src/ directory).  The library is compiled with
 * -ffp-contract=off so each expression rounds exactly as written.
 *
 * Pinning status (how each function is checked; see tests/test_oracle.py):
 *   orc_at_contrib     closed forms (equilateral, collinear), f = +grad E of the
 *                      independently coded textbook energy (central differences),
 *                      sum of the three contributions = 0, translation/rotation
 *                      invariance, linearity in nu.
 *   orc_sweep_brute    C(N,3) counts (reference tests.cpp:530-568, golden count
 *                      C(1000,3) = 166167000), brute force on tiny N by hand.
 *   orc_sweep_owner    bit-identical to orc_sweep_brute (random + adversarial boxes).
 *   orc_predictor /
 *   orc_update         known answers (f = 0, N < 3): closed-form positions and
 *                      velocities after 1 and 2 steps, including the aX quirk.
 *   orc_domain_size    bounding box + 2e-4 known answers.
 *   orc_energy         independent textbook formula via cosines; checked by the
 *                      closed forms (equilateral 11 nu / (8 a^9), collinear
 *                      -nu / (4 a^9)).
 *   orc_forces_fast    the (u,v,w) formulation used by the GPU fast kernels;
 *                      checked against orc_sweep_brute to 1e-13 relative.
 * Parity unpinned (no reference binary can be built here): the ulp-level order
 * of Eigen's 3-vector norm (restated as (x*x + y*y) + z*z) and GCC's FMA
 * contraction in the reference's -march=native build.
 */
#ifndef TB3_ORACLE_H
#define TB3_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* AxilrodTeller::CalculateForces (potential/AxilrodTeller.cpp:22-93): the
 * three contributions ci, cj, ck that the reference subtracts from i.f, j.f,
 * k.f (AxilrodTeller.cpp:119-129).  p* are xyz positions. */
void orc_at_contrib(const double *pi, const double *pj, const double *pk, double nu, double *ci, double *cj,
                    double *ck);

/* Utility::Particle::GetSqrDistPeriodic (utility/structs.hpp:106-115). */
double orc_sqr_dist_periodic(const double *pa, const double *pb, const double *L);

/* Algorithm::calculateInteractions at one process (algorithm/Algorithm.cpp:72-139)
 * with b0 = b1 = b2 = the particle array and owners (0,0,0): i<j<k sweep,
 * optional periodic cutoff test (cutoff > 0), three role accumulators f0 (b0),
 * f1 (b1), f2 (b2), each 3n doubles, zeroed here first (ResetForces,
 * DomainDecomposition.cpp:30-35).  Returns the number of accepted triplets;
 * *possible receives the number of candidates (Algorithm.cpp:92). */
uint64_t orc_sweep_brute(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f0,
                         double *f1, double *f2, uint64_t *possible);

/* Same result, bit for bit, as orc_sweep_brute, computed owner by owner:
 * every particle p sums its own role-0, role-1 and role-2 contributions in the
 * order the brute sweep visits them (role 0: (j,k) lexicographic; role 1:
 * (i,k) lexicographic; role 2: (i,j) lexicographic), over its index-sorted
 * neighbour list.  OpenMP-parallel over owners; deterministic. */
uint64_t orc_sweep_owner(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f0,
                         double *f1, double *f2, int nthreads);

/* Algorithm::SumUpParticles (Algorithm.cpp:245-249): f = f0 + (f1 + f2). */
void orc_sum_roles(int64_t n, const double *f0, const double *f1, const double *f2, double *f);

/* Particle::UpdatePredictorStage (structs.hpp:40-55): p = (p + v*dt) + (aX,aX,aX)*((dt*dt)*0.5). */
void orc_predictor(int64_t n, double *pos, const double *vel, const double *acc, double dt);

/* Particle::Update (structs.hpp:57-99): a = g - f/m; v = v + a*(dt*0.5). Positions untouched. */
void orc_update(int64_t n, const double *f, const double *mass, double *vel, double *acc, double dt,
                const double *g);

/* RegularGridDecomposition::getDomainMinMax + Init (RegularGridDecomposition.cpp:24-28, 135-166):
 * L = |(max + 0.0001) - (min - 0.0001)| per axis; lo = min - 0.0001. */
void orc_domain_size(int64_t n, const double *pos, double *L, double *lo);

/* Simulation::Start (simulation/Simulation.cpp:28-81) for one process:
 * predictor -> reset forces -> sweep -> sum roles -> update, `steps` times.
 * algo: 0 NATA, 1 P3BCA, 2 AUTA (utility/enums.hpp:1).  NATA and AUTA run
 * without a cutoff (Algorithm.cpp:31); P3BCA uses `cutoff` with the domain
 * size of the initial particles (P3BCA.cpp:17).  f receives the summed force of
 * the last step; counts[s] the accepted triplets of step s.
 * mode 0 = brute sweep, 1 = owner sweep (identical results).
 * If step_cb is non-NULL it is called after every step (for CSV output). */
typedef void (*orc_step_cb)(int step, void *user);
int orc_simulate(int64_t n, double *pos, double *vel, double *acc, const double *mass, double *f, int algo,
                 int steps, double dt, const double *g, double cutoff, double nu, uint64_t *counts, int mode,
                 int nthreads, orc_step_cb step_cb, void *user);

/* Total Axilrod-Teller energy over the same triplet set, from the textbook
 * form nu (1 + 3 cos g1 cos g2 cos g3) / (r12 r13 r23)^3 with the cosines taken
 * from dot products (independent of the reference's gradient expression). */
double orc_energy(int64_t n, const double *pos, double nu, double cutoff, const double *L, uint64_t *count,
                  int nthreads);

/* The condition scale of that energy sum: the sum over the same triplets of
 * |nu| (1 + |3 cos g1 cos g2 cos g3|) / (r12 r13 r23)^3, the lengths of each
 * term's two parts (the error bound of a sum whose terms carry relative
 * error e is e times this).  At least |E|, equal when every term's cosine
 * product is non-negative and nu > 0. */
double orc_energy_scale(int64_t n, const double *pos, double nu, double cutoff, const double *L, uint64_t *count,
                        int nthreads);

/* Per-particle force scale: scale[p] = sum over the accepted triplets that
 * contain p of the Euclidean length of the contribution the reference
 * subtracts from p's force (AxilrodTeller.cpp:83-93).  By the triangle
 * inequality scale[p] >= |f_p|, with equality for a single triplet; it is
 * the condition scale of p's force sum, the denominator of the mixed-precision
 * error metric.  Returns the largest scale. */
double orc_force_scale(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *scale,
                       int nthreads);

/* Gradient (the reference's f) and energy from the (u,v,w) squared-distance
 * formulation the GPU fast kernels use, owner-computes, FP64. */
double orc_forces_fast(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f,
                       uint64_t *count, int nthreads);

/* A fast CPU sweep for the cutoff algorithm (the multi-core CPU baseline):
 * cell list, owner = smallest index, (u,v,w) formulation, Newton's third
 * law, OpenMP over owners.  Same triplet set as the reference; forces agree
 * to rounding.  Expects every particle inside [min, min + L) per axis. */
double orc_sweep_cells_fast(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f,
                            uint64_t *count, int nthreads);

/* Bounded-sample timing of the reference-shaped brute sweep: runs rows
 * i in [0, rows) of the i<j<k loop only and returns the candidates visited. */
uint64_t orc_sweep_brute_rows(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f0,
                              double *f1, double *f2, int64_t rows, uint64_t *accepted);

/* A multi-rank step sweep (mpiexec -n ranks) restated in one process, for
 * NATA (algo 0) and AUTA (algo 2), which run without a cutoff: the
 * AtomDecomposition split with dummy padding, each algorithm's ring schedule
 * of buffer triplets, the owner rules of the triplet loops and the role sum
 * on each rank.  f receives the summed force of each particle (particle
 * order); buffer_counts[r] and particle_counts[r] the two counts rank r
 * returns from SimulationStep; processed, if non-NULL, up to cap rows of
 * (rank, owner of b0, of b1, of b2) in the order computed, and *nprocessed
 * their total.  Returns 0, or -1 for another algorithm or ranks < 1.  P3BCA
 * has its own entry points (orc_sweep_ranks_p3bca, orc_simulate_ranks_p3bca). */
int orc_sweep_ranks(int algo, int ranks, int64_t n, const double *pos, double nu, double *f,
                    uint64_t *buffer_counts, uint64_t *particle_counts, int32_t *processed, int64_t cap,
                    int64_t *nprocessed);

/* P3BCA's step sweep as `mpiexec -n p` runs it on the Cartesian
 * decomposition dims[0..ndims) (1 to 3 dimensions; p = their product, at
 * most 4096), restated in one process: the grid decomposition's cells, the
 * cutoff boxes and their validity check, the shift schedule of b1 and b2 with
 * its offset vectors, the owner rules and cutoff test of the triplet loops,
 * and each rank's role sum.  The reference takes dims from its decomposition
 * tables (one-dimensional {p} for 1..8 ranks in both).  Outputs as
 * orc_sweep_ranks.  Returns 0; -1 for bad arguments; -2 where the reference
 * exits with "invalid cutoff"; -3 if the ranks' schedules disagreed. */
int orc_sweep_ranks_p3bca(int ndims, const int *dims, int64_t n, const double *pos, double nu, double cutoff,
                          double *f, uint64_t *buffer_counts, uint64_t *particle_counts, int32_t *processed,
                          int64_t cap, int64_t *nprocessed);

/* A P3BCA run of `steps` steps on that decomposition: the initial binning,
 * then per step the predictor, the multi-rank sweep, the update and the
 * exchange of particles that left their rank's cell
 * (RegularGridDecomposition::Update).  buffer_counts and particle_counts
 * receive steps x p entries.  A particle that no rank binned at the start
 * (the reference leaves it out of the run) is integrated here with a zero
 * force.  Returns as orc_sweep_ranks_p3bca. */
int orc_simulate_ranks_p3bca(int ndims, const int *dims, int64_t n, double *pos, double *vel, double *acc,
                             const double *mass, double *f, int steps, double dt, const double *g, double cutoff,
                             double nu, uint64_t *buffer_counts, uint64_t *particle_counts, orc_step_cb step_cb,
                             void *user);

#ifdef __cplusplus
}
#endif

#endif
