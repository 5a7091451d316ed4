// tb3_internal.h -- kernel parameter blocks and launchers shared between the
// HIP kernels (tb3_kernels.hip) and the host side of the C ABI (tb3_host.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tb3 {

// Particle state on the device, structure of arrays, FP64.
struct DeviceState {
    int n = 0;
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *ax = nullptr, *ay = nullptr, *az = nullptr;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    double *m = nullptr;
    double *facc = nullptr;  // 9n: role r, component c at facc[(3r + c) n + p]
};

// Uniform cell grid over the periodic domain.  Cell width >= cutoff, so the
// 27 cells around a particle's cell (fewer when a dimension has < 3 cells)
// hold every particle within the periodic cutoff.
// The cell kernel's dispatch order: one contiguous run of cells per XCD, the
// runs cut to about equal work (so their lengths vary around ncell / 8),
// each sorted heaviest first.  `order` holds 8 runs of cell_run_cap(ncell)
// entries, then the 8 run lengths, then the 9 cuts (run r is cells
// [cut r, cut r+1)) that the last integrating pass's update computed, then
// two more sets of 9 that replays on the slotted layout alternate between.
__host__ __device__ inline int cell_run_cap(int ncell)
{
    const int run = (ncell + 7) / 8;
    return run + (run + 1) / 2;
}
__host__ __device__ inline long long cell_order_len(int ncell) { return 8LL * cell_run_cap(ncell) + 35; }

struct CellGrid {
    int nc[3] = {1, 1, 1};
    int ncell = 1;
    double lo[3] = {0, 0, 0};
    double L[3] = {1, 1, 1};
    int *cell_of = nullptr;     // n
    int *cell_cnt = nullptr;    // ncell (zero between steps)
    int *cell_rank = nullptr;   // n: each particle's rank in its cell (grids of at most 8192 cells)
    unsigned *scatter_ticket = nullptr;  // 1: finished scatter workgroups (returns to 0)
    int *cell_start = nullptr;  // ncell + 1
    int *order = nullptr;       // cell_order_len(ncell): the runs, their lengths and the cuts (above)
    unsigned *cell_work = nullptr;  // ncell: accepted triplets of each cell's last sweep (0: none yet)
    int *cell_idx = nullptr;    // n, particle indices grouped by cell (order within a cell unspecified)
    double *xs = nullptr, *ys = nullptr, *zs = nullptr;  // n, positions in cell_idx order

    // Independent particle systems swept together ("replicas"): replica r
    // holds particles [p_off[r], p_off[r+1]) and cells [c_off[r], c_off[r+1])
    // of the global cell numbering, with its own grid shape (rnc) and
    // periodic domain (rlo, rL).  With nrep == 1 the fields above describe
    // the one system and the tables are unused.
    int nrep = 1;
    int *p_off = nullptr, *c_off = nullptr;    // device, nrep + 1 each
    int *rnc = nullptr;                        // device, 3 nrep
    double *rlo = nullptr, *rL = nullptr;      // device, 3 nrep
    unsigned long long *rep_count = nullptr;   // device, nrep: accepted triplets per replica
    double *rep_energy = nullptr;              // device, nrep: energy per replica

    // Slotted cell layout of replayed steps (one system, fast sweep, no
    // dense-owner kernel): cell c's particles sit at slots [c pad_cap,
    // c pad_cap + count), written by the binning pass itself, so that no
    // scatter pass runs between the binning and the sweep.  The counts
    // alternate between two arrays by the parity of an epoch that each
    // binning pass advances: the sweep at epoch e reads pad_cnt[e & 1] and
    // clears its own cell's entry of pad_cnt[(e + 1) & 1], which the next
    // binning fills.  A count past pad_cap sets Status::pad_over.
    int pad_cap = 0;                 // slots per cell; 0: no slotted buffers
    int *pad_cnt = nullptr;          // 2 ncell
    unsigned *pad_state = nullptr;   // [0] epoch, [1] finished workgroups of a binning pass (returns to 0)
    double4 *pad_rec = nullptr;      // ncell pad_cap slots: x, y, z and the particle index (its bits in w)
};

struct CutoffSpec {
    int use_cut = 0;
    double c2 = 0;
    double L[3] = {-1, -1, -1};
};

// Per-wave partial sums land in NSLOT slots one L2 line apart (slot = block
// index mod NSLOT) instead of on one address, so thousands of waves do not
// serialise on a single atomic; launch_finalize folds them into Status.
constexpr int NSLOT = 256;
constexpr int SLOT_STRIDE = 16;  // 16 x 8 bytes = one 128-byte L2 line
struct Slots {
    unsigned long long count[2][NSLOT * SLOT_STRIDE];
    double energy[NSLOT * SLOT_STRIDE];
};

// Small device status block, read back with one copy per step.
struct Status {
    unsigned long long count[2];  // [0] accepted triplets, [1] discarded count of an auxiliary pass
    double energy;
    int flags[4];                 // [0] non-finite position seen, [1] longest neighbour list (recorded
                                  // when above the capacity), [2] list overflow (the update
                                  // kernel skips itself), [3] the cell kernel could not stage a
                                  // neighbourhood (re-run with the per-owner kernel)
    int dqueue[2];                // owner queue of the dense-owner kernel; returns to zero after each launch
    int dense_n;                  // owners the cell kernel handed to the dense-owner kernel
    int want_dense;               // the cell kernel had owners for a dense-owner kernel that was not launched
    Slots *slots;                 // partial sums of the sweep kernels
    // Replayed step graphs (tb3_run): a step whose sweep overflowed sets
    // halt; every later replay then leaves the state and the flags alone
    // (no predictor, no update, no status reset), so the host finds the
    // state just after that step's predictor.  Completed steps append their
    // count to counts[nstep++].
    int halt;
    int nstep;
    unsigned long long *counts;
    // replicas (nrep > 1): per-replica counts and energies, cleared with the
    // status block; a completed replayed step appends nrep counts
    int nrep;
    unsigned long long *rep_count;
    double *rep_energy;
    // a cell of the slotted layout received more particles than it has
    // slots (the sweep then fails the pass); cleared only by
    // launch_begin_steps for a slotted run, never by a status clear, since the binning that
    // sets it runs beside the clear of the replayed step's last kernel
    int pad_over;
};

// ---- launchers (all asynchronous on `s`) ----
void launch_soa_from_rows(hipStream_t s, int n, const double *rows, double *x, double *y, double *z);
void launch_rows_from_soa(hipStream_t s, int n, const double *x, const double *y, const double *z, double *rows);
// Clears the status block (and halt); keep_nonfinite preserves flags[0].
void launch_reset_status(hipStream_t s, Status *st, int keep_nonfinite);
// Starts a run of replayed step graphs: clears the status block, halt = 0,
// nstep = 0, counts = counts.
// With `slotted` (a run on the slotted layout) it also clears the slot
// counts, the epoch and Status::pad_over.
void launch_begin_steps(hipStream_t s, Status *st, unsigned long long *counts, CellGrid *slotted = nullptr);
// Folds the slot partial sums into st->count / st->energy and clears the
// slots; sets st->halt when the sweep overflowed (st->flags[2]).
void launch_finalize(hipStream_t s, Status *st);
void launch_predict(hipStream_t s, const DeviceState &d, double dt, double h);
// The last kernel of an integrating pass: launch_finalize's fold, then
// a = g - f/m, v += a dt/2 (f = b0 + (b1 + b2) first with sum_roles), the
// update skipped entirely when st->flags[2] (overflow) is set.  With record
// (inside a replayed step graph) a step that did not halt also appends its
// count to st->counts and clears the status block for the next replay.
// With `cells` (a cell list this pass swept), one more workgroup cuts the
// next sweep's XCD runs from the per-cell work.
void launch_update_fin(hipStream_t s, const DeviceState &d, double hdt, const double g[3], Status *st, int sum_roles,
                       int record, const CellGrid *cells, int slotted = 0);
void launch_sum_roles(hipStream_t s, const DeviceState &d);
void launch_zero(hipStream_t s, double *p, long long count);

// Exclusive scan of `count` int32 values into `out` (count + 1 entries, the
// last being the total).  `tmp` must hold scan_tmp_elems(count) int64 values.
long long scan_tmp_elems(long long count);
void launch_scan_i32(hipStream_t s, const int *in, int *out, int count, long long *tmp);

// Cell list, optionally fused with the predictor and with zeroing the force:
// bin (+ predict, + zero f) -> scan -> scatter.  Nothing downstream depends
// on the order within a cell: the fast sweep accumulates through atomics and
// the exact sweep sorts each owner's list by index.
void launch_cells_build(hipStream_t s, const DeviceState &d, int predict, double dt, double h, int zero_f,
                        CellGrid &g, Status *st, long long *tmp);
// The two halves of launch_cells_build for grids that scan inside the
// scatter kernel (cells_fused_scan): the binning (+ predictor, + force clear)
// and the scan-and-scatter, which does nothing in a halted step-graph replay.
bool cells_fused_scan(const CellGrid &g);
void launch_bin_ranked(hipStream_t s, const DeviceState &d, int predict, double dt, double h, int zero_f,
                       CellGrid &g, Status *st);
void launch_scatter_fused(hipStream_t s, const DeviceState &d, CellGrid &g, Status *st);
// A replayed step's end fused with the next step's start (cells_fused_scan
// grids): launch_update_fin's work with record, then the next step's
// predictor, force clear (zero_f) and binning.
void launch_update_bin(hipStream_t s, const DeviceState &d, double hdt, const double g3[3], Status *st,
                       int sum_roles, double dt, double h, int zero_f, CellGrid &g, int slotted = 0);
// The slotted layout (CellGrid::pad_cap): launch_bin_slotted is a run's
// first predictor (+ force clear) and binning into the slots, with the cell
// order of the first sweep (after launch_begin_steps with the grid).
// launch_update_bin and launch_update_fin with `slotted` read and fill the
// slotted counts instead (and order the next sweep's cells themselves).
void launch_bin_slotted(hipStream_t s, const DeviceState &d, double dt, double h, int zero_f, CellGrid &g,
                        Status *st);

// Cutoff sweeps (P3BCA).  One wave per particle builds that particle's
// neighbour list in LDS (capacity `cap`) from the cell list; a list longer
// than cap sets st->flags[2] and records its length in st->flags[1].
//   fast : each accepted triplet once (owner = smallest index), forces added
//          atomically into fx/fy/fz (zeroed by launch_cells_build), energy
//          and count into st.  mixed = FP32 per-triplet arithmetic.
//   exact: bit-exact ordered role sums into facc; count into st.
// launch_p3bca_fast with grad = 2 (FP64) writes each particle's force scale
// (the summed lengths of its triplet contributions) into fx instead.
void launch_p3bca_fast(hipStream_t s, const DeviceState &d, const CellGrid &g, const CutoffSpec &c, double nu,
                       int cap, int mixed, int grad, Status *st, int count_slot, int want_energy);
void launch_p3bca_exact(hipStream_t s, const DeviceState &d, const CellGrid &g, const CutoffSpec &c, double nu,
                        int cap, Status *st);
// Fast cutoff sweep with one workgroup per cell.  With
// dense_on, owners it cannot stage go to `owners` for launch_p3bca_dense,
// which must follow on the same stream; without, such an owner sets
// st->want_dense and st->flags[2] (the caller re-runs the pass with the
// dense-owner kernel).  Non-finite positions set st->flags[2] and [3]; the
// caller then re-runs the sweep with launch_p3bca_fast.
// `slotted`: sweep the slotted layout (needs dense_on == 0).
void launch_p3bca_cells(hipStream_t s, const DeviceState &d, const CellGrid &g, const CutoffSpec &c, double nu,
                        int mixed, int grad, Status *st, int count_slot, int want_energy, int *owners, int dense_on,
                        int slotted = 0);
// Workgroups of launch_p3bca_cells that fit on one CU at once.
int p3bca_cells_blocks_per_cu();
// The cell kernel hands owners whose neighbourhood does not fit its LDS
// staging to this kernel (one owner per workgroup, lists of up to
// p3bca_dense_cap() particles); owners beyond that set st->flags[2] and [3]
// like the cell kernel's own overflow.  `owners` holds n ints.  Call
// p3bca_dense_setup once per device before the first launch (it sets the
// kernel's LDS limit, which a captured graph cannot).
// `waves` (4, 8 or 16) is the workgroup size per owner in waves.
void p3bca_dense_setup();
int p3bca_dense_cap();
bool p3bca_dense_waves_ok(int waves);
int p3bca_dense_blocks_per_cu(int waves);
void launch_p3bca_dense(hipStream_t s, const DeviceState &d, const CellGrid &g, const CutoffSpec &c, double nu,
                        int mixed, int grad, Status *st, int count_slot, int want_energy, int *owners, int grid,
                        int waves);

// All-triplet sweeps (NATA / AUTA, or P3BCA without a cutoff).
void launch_auta_exact(hipStream_t s, const DeviceState &d, double nu, Status *st);
// Tile triples t0 <= t1 <= t2 of 64 particles; accumulates into fx/fy/fz (zero first).
void launch_forces_tiles(hipStream_t s, const DeviceState &d, const unsigned long long *triples,
                         long long ntriples, double nu, int mixed, int grad, Status *st, int count_slot,
                         int want_energy);

// Element-wise sqrt(a) and a / b (diagnostic for the FP64 rounding).
void launch_selftest_fp64_inrange(hipStream_t s, long long n, const double *a, const double *b, double *sq,
                                  double *q);
void launch_selftest_fp64(hipStream_t s, long long n, const double *a, const double *b, double *sq, double *q,
                          double *rs);

// LDS bytes the cutoff kernels need per wave for capacity `cap`.
size_t p3bca_fast_lds(int cap);
size_t p3bca_exact_lds(int cap);

}  // namespace tb3
