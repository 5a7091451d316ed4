// tb3_kernels.hip -- gfx950 kernels of the three-body (Axilrod-Teller) engine.
//
// Compiled with -ffp-contract=off: the bit-exact path depends on every FP64
// expression rounding as written; the fast paths spell their fused
// multiply-adds out with fma().
#include <climits>
#include <type_traits>

#include "at_math.h"
#include "tb3_internal.h"

namespace tb3 {

namespace {

constexpr int WAVE = 64;
constexpr int RING = 128;  // LDS queue of accepted pairs (power of two, >= 2 * WAVE)

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (WAVE - 1)); }

// v from the lane N below in the same 16-lane DPP row (row_shr:N), 0 for the
// first N lanes of each row; done on both halves of the double.
template <int N>
__device__ __forceinline__ double row_shr(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x110 + N, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x110 + N, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Inclusive sum over each 16-lane row (shift-and-add scan in the VALU, no LDS
// crossbar); the row's total lands in its last lane (15, 31, 47, 63).
__device__ __forceinline__ double row_sum_last(double v)
{
    v += row_shr<1>(v);
    v += row_shr<2>(v);
    v += row_shr<4>(v);
    v += row_shr<8>(v);
    return v;
}

template <int L>
__device__ __forceinline__ double read_lane(double v)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, L);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), L);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the wave, uniform in every lane: row sums by DPP, then the four
// row totals read as scalars.
__device__ __forceinline__ double wave_total(double v)
{
    v = row_sum_last(v);
    return (read_lane<15>(v) + read_lane<31>(v)) + (read_lane<47>(v) + read_lane<63>(v));
}

// Number of set bits of m below this lane.
__device__ __forceinline__ int lane_prefix(unsigned long long m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Orders this wave's LDS writes before its later LDS reads by other lanes
// (LDS operations of one wave execute in order; this keeps the compiler from
// moving them across the point).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void lds_add(double *p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

inline unsigned blocks_for(long long n, int per) { return (unsigned)((n + per - 1) / per); }

// Per-wave partial sums go to the block's slot (see Slots).
__device__ __forceinline__ void add_count(Status *st, int which, unsigned long long c)
{
    if (c) atomicAdd(&st->slots->count[which][(blockIdx.x & (NSLOT - 1)) * SLOT_STRIDE], c);
}

__device__ __forceinline__ void add_energy(Status *st, double e)
{
    if (e != 0.0) unsafeAtomicAdd(&st->slots->energy[(blockIdx.x & (NSLOT - 1)) * SLOT_STRIDE], e);
}

// The longest neighbour list matters only past the capacity (the host then
// grows the capacity to fit it).  Atomics on one address from every owner
// serialise in the memory system (tens of thousands of them cost
// milliseconds), so only lists over the capacity, and longer than the value
// already recorded, issue one.
__device__ __forceinline__ void note_list_length(Status *st, int len, int cap)
{
    if (len > cap && len > __hip_atomic_load(&st->flags[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMax(&st->flags[1], len);
}

// ---------------------------------------------------------------------------
// Layout conversion, integrator, status.

__global__ void k_soa_from_rows(int n, const double *__restrict__ rows, double *__restrict__ x,
                                double *__restrict__ y, double *__restrict__ z)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    x[p] = rows[3 * (long long)p + 0];
    y[p] = rows[3 * (long long)p + 1];
    z[p] = rows[3 * (long long)p + 2];
}

__global__ void k_rows_from_soa(int n, const double *__restrict__ x, const double *__restrict__ y,
                                const double *__restrict__ z, double *__restrict__ rows)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    rows[3 * (long long)p + 0] = x[p];
    rows[3 * (long long)p + 1] = y[p];
    rows[3 * (long long)p + 2] = z[p];
}

// The per-pass fields of the status block: the replica sums by the `nt`
// threads of one workgroup, the rest by thread 0.
__device__ void clear_status(Status *st, bool keep_nonfinite, int t, int nt)
{
    if (st->nrep > 1)
        for (int r = t; r < st->nrep; r += nt) {
            st->rep_count[r] = 0;
            st->rep_energy[r] = 0.0;
        }
    if (t != 0) return;
    st->count[0] = 0;
    st->count[1] = 0;
    st->energy = 0.0;
    st->dense_n = 0;
    st->want_dense = 0;
    if (!keep_nonfinite) st->flags[0] = 0;
    st->flags[1] = st->flags[2] = st->flags[3] = 0;
}

__global__ void k_reset_status(Status *st, int keep_nonfinite)
{
    if (threadIdx.x == 0) st->halt = 0;
    clear_status(st, keep_nonfinite != 0, threadIdx.x, blockDim.x);
}

// A run of replayed step graphs starts from a clear status block; each
// replay's last kernel clears it again for the next (see k_update_fin).
__global__ void k_begin_steps(Status *st, unsigned long long *counts)
{
    clear_status(st, false, threadIdx.x, blockDim.x);
    if (threadIdx.x == 0) {
        st->halt = 0;
        st->nstep = 0;
        st->counts = counts;
    }
}

// Folds the slot partial sums into the status block and clears the slots
// (all NSLOT threads of one workgroup); sets halt when the sweep overflowed.
__device__ void fold_slots(Status *st)
{
    __shared__ unsigned long long c0[NSLOT / WAVE], c1[NSLOT / WAVE];
    __shared__ double e0[NSLOT / WAVE];
    Slots *s = st->slots;
    const int t = threadIdx.x, k = t * SLOT_STRIDE;
    unsigned long long a = s->count[0][k], b = s->count[1][k];
    double e = s->energy[k];
    s->count[0][k] = 0;
    s->count[1][k] = 0;
    s->energy[k] = 0.0;
    a = wave_sum_u64(a);
    b = wave_sum_u64(b);
    e = wave_sum(e);
    if ((t & (WAVE - 1)) == 0) {
        c0[t / WAVE] = a;
        c1[t / WAVE] = b;
        e0[t / WAVE] = e;
    }
    __syncthreads();
    if (t == 0) {
        for (int w = 0; w < NSLOT / WAVE; ++w) {
            st->count[0] += c0[w];
            st->count[1] += c1[w];
            st->energy += e0[w];
        }
        if (st->flags[2]) st->halt = 1;
    }
}

__global__ __launch_bounds__(NSLOT) void k_finalize(Status *st) { fold_slots(st); }

// Particle::UpdatePredictorStage (utility/structs.hpp:40-55):
// p = (p + v*dt) + aX*h with h = (dt*dt)*0.5, aX for all three axes.
__device__ __forceinline__ double predicted(double x, double v, double aX, double dt, double h)
{
    return x + v * dt + aX * h;
}

__device__ __forceinline__ void predict_one(const DeviceState &d, int p, double dt, double h)
{
    const double aX = d.ax[p];
    d.x[p] = predicted(d.x[p], d.vx[p], aX, dt, h);
    d.y[p] = predicted(d.y[p], d.vy[p], aX, dt, h);
    d.z[p] = predicted(d.z[p], d.vz[p], aX, dt, h);
}

__global__ void k_predict(DeviceState d, double dt, double h)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    predict_one(d, p, dt, h);
}

// The dispatch order of the cell kernel (see cell_run_cap), for it to start
// the longest cells first: with under two workgroups per slot, the cells that
// start last otherwise decide when the sweep ends.  Each cell's work is its
// triplet count in the previous sweep (cell_work: most particles move little
// in a step), or 16 n^2 for a cell that has none yet (n = count(c)
// particles).
//
// The cuts: cell c joins run floor(8 (W_c + w_c / 2) / W), W_c being the
// work of the cells before it and W the total, which gives contiguous runs
// of about W / 8 each, so that the XCDs finish together (runs of equal cell
// counts left one XCD with 12% more work than the mean on a uniform
// system).  Should a run exceed the capacity, every run takes ncell / 8 cells
// instead.  One workgroup of 256 threads; the 9 cuts go to out (with
// write_out) and stay in cut[] for order_runs.  LDS: fpart (4 floats), cut
// (9 ints).
template <class Count>
__device__ void compute_cuts(int ncell, const Count &count, const unsigned *__restrict__ cell_work, int *out,
                             float *fpart, int *cut, bool write_out = true)
{
    const int t = threadIdx.x;
    const int run = (ncell + 7) >> 3, cap = cell_run_cap(ncell);
    auto work = [&](int c) {
        const unsigned w = cell_work[c];
        if (w) return (float)w;
        const float n = (float)count(c);
        return 16.0f * n * n;
    };
    // each thread sums a chunk of cells (their work kept in registers, up to
    // WCACHE of them), a scan of the chunk sums (in-wave shuffles, then the
    // four wave totals), then each cell's run
    constexpr int WCACHE = 32;  // chunks of up to 8192 cells stay in registers
    const int per = (ncell + 255) >> 8;
    const int c0 = min(t * per, ncell), c1 = min(c0 + per, ncell);
    float wc[WCACHE];
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < WCACHE; ++k) {
        wc[k] = c0 + k < c1 ? work(c0 + k) : 0.0f;
        s += wc[k];
    }
    for (int c = c0 + WCACHE; c < c1; ++c) s += work(c);
    float incl = s;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const float u = __shfl_up(incl, o);
        if ((t & (WAVE - 1)) >= o) incl += u;
    }
    if ((t & (WAVE - 1)) == WAVE - 1) fpart[t >> 6] = incl;
    if (t <= 8) cut[t] = ncell;
    __syncthreads();
    float before = 0.0f;
    for (int w = 0; w < (t >> 6); ++w) before += fpart[w];
    const float total = (fpart[0] + fpart[1]) + (fpart[2] + fpart[3]);
    if (total > 0.0f) {
        float acc = before + incl - s;
        for (int k = 0; c0 + k < c1; ++k) {
            const int c = c0 + k;
            const float w = k < WCACHE ? wc[k] : work(c);
            const int r = min(7, max(0, (int)(8.0f * (acc + 0.5f * w) / total)));
            atomicMin(&cut[r], c);  // the first cell of run r (runs are contiguous)
            acc += w;
        }
    }
    __syncthreads();
    if (t == 0) {
        // run r starts at its first cell, or where the next non-empty run does
        bool ok = total > 0.0f;
        cut[0] = 0;
        for (int r = 7; r >= 1; --r) cut[r] = min(cut[r], cut[r + 1]);
        for (int r = 0; r < 8 && ok; ++r) ok = cut[r + 1] - cut[r] <= cap;
        if (write_out)
            for (int r = 0; r <= 8; ++r) out[r] = ok ? cut[r] : min(r * run, ncell);
    }
}

// The order itself: runs r0, r0 + rstep, ... < 8 of the workgroup, each
// sorted heaviest first by a counting sort on bins of 1/16 octave of the
// work.  The cuts are those the last integrating pass's update computed
// (order[8 cap + 8 ...]), which the caller loaded into cut[0..8] early
// (overlapping that load with its own), if they fit this grid; else each
// run takes ncell / 8 cells.  256 threads; LDS: hist (256 ints), cut (9
// ints).
template <class Count>
__device__ void order_runs(int ncell, const Count &count, const unsigned *__restrict__ cell_work,
                           int *__restrict__ order, int *hist, int *cut, int r0, int rstep)
{
    if (r0 >= 8) return;
    const int t = threadIdx.x;
    const int run = (ncell + 7) >> 3, cap = cell_run_cap(ncell);
    auto bin = [&](int c) {
        const unsigned w = cell_work[c];
        const float n = (float)count(c);
        const float k = w ? (float)w : 16.0f * n * n;
        return 255 - min(255, (int)(16.0f * __log2f(1.0f + k)));
    };
    __syncthreads();  // the caller's cut[] stores
    if (t == 0) {
        bool ok = cut[0] == 0 && cut[8] == ncell;
        for (int r = 0; r < 8 && ok; ++r) ok = cut[r + 1] >= cut[r] && cut[r + 1] - cut[r] <= cap;
        if (!ok)
            for (int r = 0; r <= 8; ++r) cut[r] = min(r * run, ncell);
    }
    __syncthreads();
    for (int r = r0; r < 8; r += rstep) {
        const int lo = cut[r], hi = cut[r + 1];
        hist[t] = 0;
        __syncthreads();
        for (int c = lo + t; c < hi; c += 256) atomicAdd(&hist[bin(c)], 1);
        __syncthreads();
        // exclusive scan of the 256 bins, one per thread (Hillis-Steele)
        const int own = hist[t];
        for (int o = 1; o < 256; o <<= 1) {
            const int u = t >= o ? hist[t - o] : 0;
            __syncthreads();
            hist[t] += u;
            __syncthreads();
        }
        hist[t] -= own;
        __syncthreads();
        for (int c = lo + t; c < hi; c += 256) order[r * cap + atomicAdd(&hist[bin(c)], 1)] = c;
        if (t == 0) order[8 * cap + r] = hi - lo;
        __syncthreads();
    }
}

// Grids of at most this many cells scan their counts inside the scatter
// kernel (k_cell_scatter_fused) and may replay on the slotted layout.
constexpr int FUSED_SCAN_CELLS = 8192;

// The slotted cell layout (CellGrid::pad_cap) as the kernels see it.
struct SlotsDev {
    int cap;          // slots per cell; 0: the contiguous layout is in use
    int *cnt;         // 2 ncell counts, by epoch parity
    unsigned *state;  // [0] epoch, [1] finished workgroups of a binning pass
    double4 *rec;     // ncell cap slots: x, y, z, and the particle index's bits in w
};

// Particle p (position px, py, pz) into a slot of cell c, counted in cnt; a
// cell past its slots raises Status::pad_over instead.
__device__ __forceinline__ void bin_slotted(const SlotsDev &sl, int *cnt, int c, int p, double px, double py,
                                            double pz, Status *st)
{
    const int r = atomicAdd(&cnt[c], 1);
    if (r < sl.cap) {
        // one 32-byte record: a single store where four arrays took four
        sl.rec[(long long)c * sl.cap + r] = make_double4(px, py, pz, __longlong_as_double((long long)p));
    } else {
        atomicOr(&st->pad_over, 1);
    }
}

// Each workgroup of a slotted binning pass takes a ticket once its binning
// is done; the last one to arrive advances the epoch and leaves the ticket
// at zero.  Every workgroup's read of the old epoch has completed before its
// ticket (its value addressed the counts before the barrier), so no fence is
// needed; the next kernel sees the new epoch.  (With an agent-scope fence
// before every ticket the kernel took 3 us longer.)
__device__ __forceinline__ void slots_ticket(const SlotsDev &sl, unsigned e)
{
    __syncthreads();
    if (threadIdx.x == 0 && atomicInc(&sl.state[1], gridDim.x - 1) == gridDim.x - 1) sl.state[0] = e + 1;
}

// The last kernel of an integrating pass.  Workgroup 0 first folds the
// sweep's slot sums into the status block (k_finalize); with `record` (a
// replayed step graph) it then appends the step's count and, unless the
// step halted, clears the status block for the next replay.  Every thread
// then applies Particle::Update (structs.hpp:57-99) to its particle:
// a = g - f/m, v = v + a*(dt*0.5), with f = b0 + (b1 + b2) first when
// `sum_roles` (Algorithm::SumUpParticles, Algorithm.cpp:245-249).  Nothing
// is updated when the sweep overflowed (the host re-runs it first); the
// status clear cannot affect that test, since it happens only when
// flags[2] is already zero.
__global__ __launch_bounds__(NSLOT) void k_update_fin(DeviceState d, double hdt, double gx, double gy, double gz,
                                                     Status *st, int sum_roles, int record, int ncell,
                                                     const int *__restrict__ cell_start,
                                                     const unsigned *__restrict__ cell_work, int *__restrict__ order,
                                                     SlotsDev sl)
{
    // The particles' workgroups, then (with a cell list, ncell > 0) one that
    // cuts the next sweep's XCD runs from this sweep's per-cell work (and
    // cell counts, contiguous or slotted), then one that folds the status.
    const int nupd = (int)gridDim.x - (ncell > 0 ? 2 : 1);
    const int x = (int)blockIdx.x - nupd;
    if (x == (ncell > 0 ? 1 : 0)) {
        fold_slots(st);
        __syncthreads();  // thread 0's fold (and halt) is visible to the workgroup
        if (record && !st->halt) {
            const int t = threadIdx.x, nrep = st->nrep, ns = st->nstep;
            if (nrep > 1) {
                for (int r = t; r < nrep; r += NSLOT) st->counts[(long long)ns * nrep + r] = st->rep_count[r];
            } else if (t == 0) {
                st->counts[ns] = st->count[0];
            }
            __syncthreads();  // every read of this step's sums precedes the clear
            clear_status(st, false, t, NSLOT);
            if (t == 0) st->nstep = ns + 1;
        }
        return;
    }
    if (x == 0) {
        __shared__ float fpart[4];
        __shared__ int cut[9];
        int *out = order + 8 * cell_run_cap(ncell) + 8;
        if (sl.cap > 0) {
            const int *cur = sl.cnt + (sl.state[0] & 1u) * ncell;
            compute_cuts(ncell, [&](int c) { return min(cur[c], sl.cap); }, cell_work, out, fpart, cut);
        } else {
            compute_cuts(ncell, [&](int c) { return cell_start[c + 1] - cell_start[c]; }, cell_work, out, fpart,
                         cut);
        }
        return;
    }
    // The fold's clear happens only when flags[2] is already zero, so every
    // workgroup reads the same flags[2].
    if (st->flags[2]) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    double fx, fy, fz;
    if (sum_roles) {
        const long long n = d.n;
        fx = d.facc[0 * n + p] + (d.facc[3 * n + p] + d.facc[6 * n + p]);
        fy = d.facc[1 * n + p] + (d.facc[4 * n + p] + d.facc[7 * n + p]);
        fz = d.facc[2 * n + p] + (d.facc[5 * n + p] + d.facc[8 * n + p]);
        d.fx[p] = fx;
        d.fy[p] = fy;
        d.fz[p] = fz;
    } else {
        fx = d.fx[p];
        fy = d.fy[p];
        fz = d.fz[p];
    }
    const double m = d.m[p];
    const double a0 = gx - fx / m, a1 = gy - fy / m, a2 = gz - fz / m;
    d.vx[p] = d.vx[p] + a0 * hdt;
    d.vy[p] = d.vy[p] + a1 * hdt;
    d.vz[p] = d.vz[p] + a2 * hdt;
    d.ax[p] = a0;
    d.ay[p] = a1;
    d.az[p] = a2;
}

// Algorithm::SumUpParticles (Algorithm.cpp:245-249): f = b0 + (b1 + b2).
__global__ void k_sum_roles(DeviceState d)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    const long long n = d.n;
    d.fx[p] = d.facc[0 * n + p] + (d.facc[3 * n + p] + d.facc[6 * n + p]);
    d.fy[p] = d.facc[1 * n + p] + (d.facc[4 * n + p] + d.facc[7 * n + p]);
    d.fz[p] = d.facc[2 * n + p] + (d.facc[5 * n + p] + d.facc[8 * n + p]);
}

__global__ void k_zero(double *p, long long count)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Exclusive scan (two-level, 2048 items per 256-thread block).

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <class TIn, class TOut>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile(const TIn *__restrict__ in, long long n_in,
                                                            TOut *__restrict__ out, long long n_out,
                                                            long long *__restrict__ block_sums)
{
    __shared__ long long wtot[SCAN_THREADS / WAVE];
    const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
    long long v[SCAN_ITEMS];
    long long tsum = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const long long idx = base + i;
        const long long val = idx < n_in ? (long long)in[idx] : 0;
        v[i] = tsum;
        tsum += val;
    }
    const int lane = lane_id(), wid = threadIdx.x / WAVE;
    long long incl = tsum;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const long long y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane == WAVE - 1) wtot[wid] = incl;
    __syncthreads();
    long long wpre = 0;
    for (int w = 0; w < wid; ++w) wpre += wtot[w];
    const long long texcl = wpre + incl - tsum;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const long long idx = base + i;
        if (idx < n_out) out[idx] = (TOut)(texcl + v[i]);
    }
    if (threadIdx.x == SCAN_THREADS - 1) block_sums[blockIdx.x] = wpre + incl;
}

template <class TOut>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_add(TOut *__restrict__ out, long long n_out,
                                                           const long long *__restrict__ block_off)
{
    const long long base = (long long)blockIdx.x * SCAN_TILE;
    const long long add = block_off[blockIdx.x];
    for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_THREADS) {
        const long long idx = base + i;
        if (idx < n_out) out[idx] = (TOut)((long long)out[idx] + add);
    }
}

template <class TIn, class TOut>
void scan_rec(hipStream_t s, const TIn *in, long long n_in, TOut *out, long long n_out, long long *tmp)
{
    const long long nb = (n_out + SCAN_TILE - 1) / SCAN_TILE;
    long long *sums = tmp;
    long long *offs = tmp + nb;
    hipLaunchKernelGGL((k_scan_tile<TIn, TOut>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, in, n_in, out, n_out,
                       sums);
    if (nb > 1) {
        scan_rec<long long, long long>(s, sums, nb, offs, nb, offs + nb);
        hipLaunchKernelGGL((k_scan_add<TOut>), dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, out, n_out, offs);
    }
}

// ---------------------------------------------------------------------------
// Cell list.

struct GridDev {
    int nc[3];
    double lo[3], L[3];
    // replicas (nrep > 1): particle and cell ranges, grid shapes and domains
    int nrep;
    const int *p_off, *c_off, *rnc;
    const double *rlo, *rL;
};

// One replica's grid, resolved for a particle or a cell.
struct SysGrid {
    int r, cbase;
    int nc[3];
    double lo[3], L[3];
};

// Last r with off[r] <= v (off ascending, off[0] = 0 <= v).
__device__ __forceinline__ int find_range(const int *off, int nrep, int v)
{
    int lo = 0, hi = nrep - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ SysGrid sys_grid(const GridDev &g, int r)
{
    SysGrid s;
    s.r = r;
    if (g.nrep <= 1) {
        s.cbase = 0;
        for (int k = 0; k < 3; ++k) {
            s.nc[k] = g.nc[k];
            s.lo[k] = g.lo[k];
            s.L[k] = g.L[k];
        }
    } else {
        s.cbase = g.c_off[r];
        for (int k = 0; k < 3; ++k) {
            s.nc[k] = g.rnc[3 * r + k];
            s.lo[k] = g.rlo[3 * r + k];
            s.L[k] = g.rL[3 * r + k];
        }
    }
    return s;
}

__device__ __forceinline__ SysGrid grid_of_particle(const GridDev &g, int p)
{
    return sys_grid(g, g.nrep <= 1 ? 0 : find_range(g.p_off, g.nrep, p));
}

__device__ __forceinline__ SysGrid grid_of_cell(const GridDev &g, int c)
{
    return sys_grid(g, g.nrep <= 1 ? 0 : find_range(g.c_off, g.nrep, c));
}

__device__ __forceinline__ int cell_coord(double x, double lo, double L, int nc)
{
    if (nc == 1) return 0;
    double t = (x - lo) / L;
    t = t - __builtin_floor(t);
    const int i = (int)(t * (double)nc);
    return i < 0 ? 0 : (i >= nc ? nc - 1 : i);
}

__device__ __forceinline__ bool finite3(double x, double y, double z)
{
    return __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
}

// Wrapped cell of a position; non-finite positions go to cell 0.
__device__ __forceinline__ void cell_of_pos(const SysGrid &g, double x, double y, double z, int &cx, int &cy, int &cz)
{
    if (finite3(x, y, z)) {
        cx = cell_coord(x, g.lo[0], g.L[0], g.nc[0]);
        cy = cell_coord(y, g.lo[1], g.L[1], g.nc[1]);
        cz = cell_coord(z, g.lo[2], g.L[2], g.nc[2]);
    } else {
        cx = cy = cz = 0;
    }
}

template <bool RANK>
__global__ void k_bin(DeviceState d, int predict, double dt, double h, int zero_f, GridDev g,
                      int *__restrict__ cell_of, int *__restrict__ cell_cnt, int *__restrict__ cell_rank, Status *st)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    // the predicted position stays in registers for the binning below
    double px = d.x[p], py = d.y[p], pz = d.z[p];
    if (predict && !st->halt) {
        const double aX = d.ax[p];
        px = predicted(px, d.vx[p], aX, dt, h);
        py = predicted(py, d.vy[p], aX, dt, h);
        pz = predicted(pz, d.vz[p], aX, dt, h);
        d.x[p] = px;
        d.y[p] = py;
        d.z[p] = pz;
    }
    if (zero_f) {
        d.fx[p] = 0.0;
        d.fy[p] = 0.0;
        d.fz[p] = 0.0;
    }
    if (!finite3(px, py, pz)) atomicOr(&st->flags[0], 1);
    const SysGrid sg = grid_of_particle(g, p);
    int cx, cy, cz;
    cell_of_pos(sg, px, py, pz, cx, cy, cz);
    const int c = sg.cbase + (cz * sg.nc[1] + cy) * sg.nc[0] + cx;
    cell_of[p] = c;
    if constexpr (RANK) cell_rank[p] = atomicAdd(&cell_cnt[c], 1);
    else atomicAdd(&cell_cnt[c], 1);
}

// k_begin_steps for a run on the slotted layout, which also starts with
// every slot count at zero, epoch 0 and a clear overflow flag (one workgroup;
// a kernel, so that nothing here is a memset).
__global__ __launch_bounds__(256) void k_begin_steps_slotted(Status *st, unsigned long long *counts, int ncell,
                                                             SlotsDev sl)
{
    const int t = threadIdx.x;
    clear_status(st, false, t, blockDim.x);
    for (int i = t; i < 2 * ncell; i += blockDim.x) sl.cnt[i] = 0;
    if (t == 0) {
        st->halt = 0;
        st->nstep = 0;
        st->counts = counts;
        st->pad_over = 0;
        sl.state[0] = 0u;
        sl.state[1] = 0u;
    }
}

// The first step of a slotted run: predictor, force clear and binning of
// k_bin<true>, into the slots of the current epoch; 8 more workgroups order
// the first sweep's cells from the cuts the last integrating pass stored
// (the counts fill beside them, so cells never swept weigh alike).
__global__ __launch_bounds__(256) void k_bin_slotted(DeviceState d, double dt, double h, int zero_f, GridDev g,
                                                     SlotsDev sl, Status *st, int ncell,
                                                     const unsigned *__restrict__ cell_work, int *__restrict__ order)
{
    const int nbin = (int)gridDim.x - 8;
    if ((int)blockIdx.x >= nbin) {
        __shared__ int hist[256], cut[9];
        int *cuts = order + 8 * cell_run_cap(ncell) + 8;
        if (threadIdx.x <= 8) cut[threadIdx.x] = cuts[threadIdx.x];
        order_runs(ncell, [](int) { return 0; }, cell_work, order, hist, cut, (int)blockIdx.x - nbin, 8);
        // the cuts the first replayed update orders with (epoch 0's set)
        if (blockIdx.x == nbin && threadIdx.x <= 8) cuts[9 + threadIdx.x] = cut[threadIdx.x];
        return;
    }
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    const double aX = d.ax[p];
    const double px = predicted(d.x[p], d.vx[p], aX, dt, h);
    const double py = predicted(d.y[p], d.vy[p], aX, dt, h);
    const double pz = predicted(d.z[p], d.vz[p], aX, dt, h);
    d.x[p] = px;
    d.y[p] = py;
    d.z[p] = pz;
    if (zero_f) {
        d.fx[p] = 0.0;
        d.fy[p] = 0.0;
        d.fz[p] = 0.0;
    }
    if (!finite3(px, py, pz)) atomicOr(&st->flags[0], 1);
    const SysGrid sg = grid_of_particle(g, p);
    int cx, cy, cz;
    cell_of_pos(sg, px, py, pz, cx, cy, cz);
    const int c = sg.cbase + (cz * sg.nc[1] + cy) * sg.nc[0] + cx;
    bin_slotted(sl, sl.cnt + (sl.state[0] & 1u) * ncell, c, p, px, py, pz, st);
}

// One particle of k_update_bin: update, predictor, force clear, binning
// (contiguous: cell and rank; slotted: into the next epoch's slots).
__device__ __forceinline__ void update_bin_one(const DeviceState &d, double hdt, double gx, double gy, double gz,
                                               Status *st, int sum_roles, double dt, double h, int zero_f,
                                               const GridDev &g, int *__restrict__ cell_of,
                                               int *__restrict__ cell_cnt, int *__restrict__ cell_rank,
                                               const SlotsDev &sl, int *next_cnt, int p)
{
    double fx, fy, fz;
    if (sum_roles) {
        const long long n = d.n;
        fx = d.facc[0 * n + p] + (d.facc[3 * n + p] + d.facc[6 * n + p]);
        fy = d.facc[1 * n + p] + (d.facc[4 * n + p] + d.facc[7 * n + p]);
        fz = d.facc[2 * n + p] + (d.facc[5 * n + p] + d.facc[8 * n + p]);
        d.fx[p] = fx;
        d.fy[p] = fy;
        d.fz[p] = fz;
    } else {
        fx = d.fx[p];
        fy = d.fy[p];
        fz = d.fz[p];
    }
    const double m = d.m[p];
    const double a0 = gx - fx / m, a1 = gy - fy / m, a2 = gz - fz / m;
    const double vx = d.vx[p] + a0 * hdt, vy = d.vy[p] + a1 * hdt, vz = d.vz[p] + a2 * hdt;
    d.vx[p] = vx;
    d.vy[p] = vy;
    d.vz[p] = vz;
    d.ax[p] = a0;
    d.ay[p] = a1;
    d.az[p] = a2;
    // the next step: predictor with aX for all three axes
    const double px = predicted(d.x[p], vx, a0, dt, h);
    const double py = predicted(d.y[p], vy, a0, dt, h);
    const double pz = predicted(d.z[p], vz, a0, dt, h);
    d.x[p] = px;
    d.y[p] = py;
    d.z[p] = pz;
    if (zero_f) {
        d.fx[p] = 0.0;
        d.fy[p] = 0.0;
        d.fz[p] = 0.0;
    }
    if (!finite3(px, py, pz)) atomicOr(&st->flags[0], 1);
    const SysGrid sg = grid_of_particle(g, p);
    int cx, cy, cz;
    cell_of_pos(sg, px, py, pz, cx, cy, cz);
    const int c = sg.cbase + (cz * sg.nc[1] + cy) * sg.nc[0] + cx;
    if (sl.cap > 0) {
        bin_slotted(sl, next_cnt, c, p, px, py, pz, st);
    } else {
        cell_of[p] = c;
        cell_rank[p] = atomicAdd(&cell_cnt[c], 1);
    }
}

// The end of one replayed step fused with the start of the next (cutoff
// algorithm, grids of at most FUSED_SCAN_CELLS cells): k_update_fin's fold,
// count record, status clear and Particle::Update (structs.hpp:57-99), then,
// in the same thread, the next step's predictor (structs.hpp:40-55) on the
// updated velocity and acceleration, the force clear and the binning of
// k_bin<true>.  The same operations on the same values as the two kernels,
// one launch fewer per step.  The clear keeps flags[0] (a non-finite
// position, which the binning may raise concurrently): it can only be set
// here when the last sweep ran without overflow, i.e. with every position
// finite, or on the per-owner and exact kernels, where it is sticky anyway.
__global__ __launch_bounds__(NSLOT) void k_update_bin(DeviceState d, double hdt, double gx, double gy, double gz,
                                                     Status *st, int sum_roles, double dt, double h, int zero_f,
                                                     GridDev g, int *__restrict__ cell_of,
                                                     int *__restrict__ cell_cnt, int *__restrict__ cell_rank,
                                                     int ncell, const int *__restrict__ cell_start,
                                                     const unsigned *__restrict__ cell_work, int *__restrict__ order,
                                                     SlotsDev sl)
{
    // The particles' workgroups, then: (slotted) 8 that order the next
    // sweep's runs; one that cuts the XCD runs from this sweep's work; one
    // that folds the status.  flags[2] is the same for the whole grid: only
    // a halted replay has it, and then nothing but the fold runs and no
    // workgroup takes a ticket (the fold's status clear happens only when
    // flags[2] is already zero).
    const bool slotted = sl.cap > 0;
    const int nupd = (int)gridDim.x - (slotted ? 10 : 2);
    const int x = (int)blockIdx.x - nupd;  // >= 0: an extra workgroup
    const unsigned e = slotted ? sl.state[0] : 0u;
    if (x == (slotted ? 9 : 1)) {
        // the status fold, the step's count record and the clear for the
        // next replay, beside the particles' workgroups
        fold_slots(st);
        __syncthreads();  // thread 0's fold (and halt) is visible to the workgroup
        const bool halted = st->halt;
        if (!halted) {
            const int t = threadIdx.x, nrep = st->nrep, ns = st->nstep;
            if (nrep > 1) {
                for (int r = t; r < nrep; r += NSLOT) st->counts[(long long)ns * nrep + r] = st->rep_count[r];
            } else if (t == 0) {
                st->counts[ns] = st->count[0];
            }
            __syncthreads();  // every read of this step's sums precedes the clear
            clear_status(st, true, t, NSLOT);
            if (t == 0) st->nstep = ns + 1;
        }
        if (slotted && !halted) slots_ticket(sl, e);
        return;
    }
    if (x >= 0) {
        __shared__ float fpart[4];
        __shared__ int cut[9], hist[256];
        int *cuts = order + 8 * cell_run_cap(ncell) + 8;
        if (!slotted) {
            // the next sweep's XCD runs from this sweep's per-cell work (the
            // scatter kernel then orders them)
            compute_cuts(ncell, [&](int c) { return cell_start[c + 1] - cell_start[c]; }, cell_work, cuts, fpart,
                         cut);
            return;
        }
        if (st->flags[2]) return;
        const int *cur = sl.cnt + (e & 1u) * ncell;
        auto count = [&](int c) { return min(cur[c], sl.cap); };
        // slotted runs keep two more sets of cuts, by epoch parity: the
        // orders below use the cuts of the previous replay's work (the set of
        // this epoch), while the cut workgroup computes the next set from this
        // sweep's, so neither waits for the other
        int *par = cuts + 9;
        if (x < 8) {
            if (threadIdx.x <= 8) cut[threadIdx.x] = par[(e & 1u) * 9 + threadIdx.x];
            order_runs(ncell, count, cell_work, order, hist, cut, x, 8);
        } else {
            compute_cuts(ncell, count, cell_work, par + ((e + 1u) & 1u) * 9, fpart, cut);
        }
        slots_ticket(sl, e);
        return;
    }
    if (st->flags[2]) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < d.n) update_bin_one(d, hdt, gx, gy, gz, st, sum_roles, dt, h, zero_f, g, cell_of, cell_cnt, cell_rank,
                                sl, slotted ? sl.cnt + ((e + 1u) & 1u) * ncell : nullptr, p);
    if (slotted) slots_ticket(sl, e);
}

// Places each particle in its cell's range (index and a cell-ordered copy of
// its position), counting the cell's counter back down to zero so the
// counters are clear for the next step; also orders the cells (order_runs).
__global__ __launch_bounds__(256) void k_cell_scatter(int n, const double *__restrict__ x,
                                                      const double *__restrict__ y, const double *__restrict__ z,
                                                      const int *__restrict__ cell_of,
                                                      const int *__restrict__ cell_start, int *__restrict__ cell_cnt,
                                                      int *__restrict__ cell_idx, double *__restrict__ xs,
                                                      double *__restrict__ ys, double *__restrict__ zs, int ncell,
                                                      int *__restrict__ order, const unsigned *__restrict__ cell_work)
{
    __shared__ int hist[256], cut[9];
    if (threadIdx.x <= 8) cut[threadIdx.x] = order[8 * cell_run_cap(ncell) + 8 + threadIdx.x];
    order_runs(ncell, [&](int c) { return cell_start[c + 1] - cell_start[c]; }, cell_work, order, hist, cut,
               blockIdx.x, gridDim.x);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int c = cell_of[p];
    const int slot = cell_start[c] + atomicSub(&cell_cnt[c], 1) - 1;
    cell_idx[slot] = p;
    xs[slot] = x[p];
    ys[slot] = y[p];
    zs[slot] = z[p];
}

// Grids of at most FUSED_SCAN_CELLS cells: the scan of the counts is done by
// every workgroup of the scatter in its own LDS, which saves the scan kernel
// and a kernel boundary per step.  The binning pass recorded each particle's
// rank in its cell, so the slot is start + rank and the counters are only
// read here; the last workgroup to finish (a ticket that wraps back to 0)
// clears them for the next step, after every workgroup has read them.
// Workgroup 0 publishes the starts for the sweep kernels.  The grid is the
// nsc = ceil(n / 256) scatter workgroups and 8 more that each order one
// XCD's run (order_runs), so the ordering runs beside the scatter instead of
// after it.  Each scatter thread loads its particle before the scan, which
// overlaps the two memory round trips.
__global__ __launch_bounds__(256) void k_cell_scatter_fused(
    int n, const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const int *__restrict__ cell_of, const int *__restrict__ cell_rank, int *__restrict__ cell_cnt,
    int *__restrict__ cell_start, int *__restrict__ cell_idx, double *__restrict__ xs, double *__restrict__ ys,
    double *__restrict__ zs, int ncell, int *__restrict__ order, const unsigned *__restrict__ cell_work,
    unsigned *__restrict__ ticket, int nsc, const Status *__restrict__ st)
{
    __shared__ int lst[FUSED_SCAN_CELLS + 1];
    __shared__ int part[256];
    __shared__ int hist[256], cut[9];
    __shared__ int s_last;
    // A halted step-graph replay keeps the cell list of the step that
    // overflowed (the host finishes that step on it); every workgroup reads
    // the same halt, so the ticket stays consistent.
    if (st->halt) return;
    const int t = threadIdx.x;
    const int p = blockIdx.x * blockDim.x + t;  // >= n in the ordering workgroups
    int pcell = 0, prank = 0, gcut = 0;
    double px = 0.0, py = 0.0, pz = 0.0;
    if (blockIdx.x >= nsc && t <= 8) gcut = order[8 * cell_run_cap(ncell) + 8 + t];  // the given cuts
    if (p < n) {
        pcell = cell_of[p];
        prank = cell_rank[p];
        px = x[p];
        py = y[p];
        pz = z[p];
    }
    // exclusive scan of the counts: coalesced load, a serial scan of each
    // thread's chunk, then a scan of the 256 chunk sums
    for (int c = t; c < ncell; c += 256) lst[c] = cell_cnt[c];
    __syncthreads();
    const int per = (ncell + 255) >> 8;
    const int c0 = min(t * per, ncell), c1 = min(c0 + per, ncell);
    int sum = 0;
    for (int c = c0; c < c1; ++c) {
        const int k = lst[c];
        lst[c] = sum;
        sum += k;
    }
    part[t] = sum;
    int v = sum;
    for (int o = 1; o < 256; o <<= 1) {
        __syncthreads();
        const int u = t >= o ? part[t - o] : 0;
        __syncthreads();
        v += u;
        part[t] = v;
    }
    const int off = v - sum;
    for (int c = c0; c < c1; ++c) lst[c] += off;
    if (t == 255) lst[ncell] = v;
    __syncthreads();
    if (blockIdx.x == 0)
        for (int c = t; c <= ncell; c += 256) cell_start[c] = lst[c];
    if (blockIdx.x >= nsc) {
        if (t <= 8) cut[t] = gcut;
        order_runs(ncell, [&](int c) { return lst[c + 1] - lst[c]; }, cell_work, order, hist, cut, blockIdx.x - nsc,
                   8);
    }

    if (p < n) {
        const int slot = lst[pcell] + prank;
        cell_idx[slot] = p;
        xs[slot] = px;
        ys[slot] = py;
        zs[slot] = pz;
    }
    __syncthreads();
    if (t == 0) s_last = atomicInc(ticket, gridDim.x - 1) == gridDim.x - 1;
    __syncthreads();
    if (s_last)
        for (int c = t; c < ncell; c += 256) cell_cnt[c] = 0;
}

// ---------------------------------------------------------------------------
// Cutoff sweeps: shared parameters and the in-wave neighbour search.

struct CutDev {
    int n;
    const double *x, *y, *z;
    GridDev g;
    const int *cell_start, *cell_idx;
    const double *xs, *ys, *zs;  // positions in cell order (slot s holds particle cell_idx[s])
    double c2;
    double nu;
    int cap;
    double *fx, *fy, *fz;
    double *facc;
    Status *st;
    int count_slot;
    int want_energy;
    unsigned long long *rep_count;  // nrep > 1: per-replica counts (count_slot 0 only)
    double *rep_energy;
};

// Per-replica partial sums (replica batches only).
__device__ __forceinline__ void add_rep(const CutDev &a, int r, unsigned long long c, double e)
{
    if (a.g.nrep <= 1) return;
    if (c && a.count_slot == 0) atomicAdd(&a.rep_count[r], c);
    if (a.want_energy && e != 0.0) atomicAdd(&a.rep_energy[r], e);
}

// The candidate set of owner p: the particles of the 27 (or fewer) cells
// around p's cell, or every particle when some position is not finite (the
// reference's brute sweep pairs NaN positions with everything).  Lanes
// 0..ncells-1 publish one cell's (start, exclusive prefix) in LDS; the
// flattened candidate t maps back to a particle with a short search.
struct Candidates {
    int total;
    int ncells;
    bool brute;
    int base;  // brute: first particle of the owner's replica
};

__device__ __forceinline__ Candidates gather_candidates(const CutDev &a, const SysGrid &sg, double px, double py,
                                                        double pz, int *sh_start, int *sh_pre)
{
    Candidates c;
    c.brute = a.st->flags[0] != 0;
    if (c.brute) {
        c.base = a.g.nrep <= 1 ? 0 : a.g.p_off[sg.r];
        c.total = (a.g.nrep <= 1 ? a.n : a.g.p_off[sg.r + 1]) - c.base;
        c.ncells = 0;
        return c;
    }
    c.base = 0;
    const int lane = lane_id();
    int cx, cy, cz;
    cell_of_pos(sg, px, py, pz, cx, cy, cz);
    const int ncx = sg.nc[0], ncy = sg.nc[1], ncz = sg.nc[2];
    const int wx = ncx >= 3 ? 3 : 1, wy = ncy >= 3 ? 3 : 1, wz = ncz >= 3 ? 3 : 1;
    c.ncells = wx * wy * wz;
    int cnt = 0, st = 0;
    if (lane < c.ncells) {
        const int ox = lane % wx, oy = (lane / wx) % wy, oz = lane / (wx * wy);
        int xx = cx + (wx == 3 ? ox - 1 : 0), yy = cy + (wy == 3 ? oy - 1 : 0), zz = cz + (wz == 3 ? oz - 1 : 0);
        xx = xx < 0 ? xx + ncx : (xx >= ncx ? xx - ncx : xx);
        yy = yy < 0 ? yy + ncy : (yy >= ncy ? yy - ncy : yy);
        zz = zz < 0 ? zz + ncz : (zz >= ncz ? zz - ncz : zz);
        const int cc = sg.cbase + (zz * ncy + yy) * ncx + xx;
        st = a.cell_start[cc];
        cnt = a.cell_start[cc + 1] - st;
    }
    int incl = cnt;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    if (lane < c.ncells) {
        sh_start[lane] = st;
        sh_pre[lane] = incl - cnt;
    }
    c.total = __shfl(incl, WAVE - 1);
    __syncthreads();
    return c;
}

__device__ __forceinline__ int candidate(const CutDev &a, const Candidates &c, const int *sh_start,
                                         const int *sh_pre, int t)
{
    if (c.brute) return c.base + t;
    int lo = 0, hi = c.ncells - 1;  // last cell k with sh_pre[k] <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sh_pre[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return a.cell_idx[sh_start[lo] + (t - sh_pre[lo])];
}

__device__ __forceinline__ bool within_cutoff(const CutDev &a, const SysGrid &sg, double ax, double ay, double az,
                                              double bx, double by, double bz)
{
    // negation of the rejection test of Algorithm.cpp:95-97
    return !(sqr_dist_periodic(ax, ay, az, bx, by, bz, sg.L[0], sg.L[1], sg.L[2]) > a.c2);
}

// Owner slot for block b (one wave per owner): blocks are spread
// round-robin over the 8 XCDs, so block b runs on XCD b % 8.  Each XCD takes
// every 8th stretch of OWNER_STRETCH consecutive cell-ordered particles:
// a stretch keeps a cell neighbourhood in one XCD's L2, and interleaving the
// stretches spreads the density's variations (and so the work) evenly over
// the XCDs, where one contiguous eighth of the particles per XCD left the
// XCDs' work up to ~12% apart (the exact sweep ran 5% faster interleaved).
// Slots from n up to the grid's end are empty (owner_grid).
constexpr int OWNER_STRETCH = 256;

__device__ __forceinline__ int owner_slot(int b)
{
    const int k = b >> 3, r = b & 7;
    return ((k / OWNER_STRETCH) * 8 + r) * OWNER_STRETCH + k % OWNER_STRETCH;
}

// ---------------------------------------------------------------------------
// Fast cutoff sweep: every accepted triplet once.
//
// Wave per owner p.  The owner collects N+(p), its neighbours with a larger
// index, in LDS, walks the unordered pairs {a, b} of N+(p) in a folded order
// (row d pairs a with (a + d) mod len, so all 64 lanes test distinct pairs),
// applies the periodic cutoff test to the third side, queues accepted pairs
// and evaluates 64 queued triplets at a time.  The owner's gradient stays in
// registers; the two partners' gradients go to per-neighbour LDS
// accumulators (LDS atomics), flushed to global memory with one atomic per
// neighbour and component at the end.

// SCALE (with T = double and GRAD): instead of the gradient, each particle's
// force scale, the sum over its triplets of the length of its contribution
// (the condition of its force sum), into fx (fy and fz stay zero).
template <class T, bool GRAD, bool SCALE = false>
__global__ __launch_bounds__(WAVE) void k_p3bca_fast(CutDev a)
{
    extern __shared__ double smem[];
    const int cap = a.cap;
    double *sx = smem;
    double *sy = sx + cap;
    double *sz = sy + cap;
    double *gx_ = sz + cap;
    double *gy_ = gx_ + cap;
    double *gz_ = gy_ + cap;
    int *sid = reinterpret_cast<int *>(gz_ + cap);
    unsigned *ring = reinterpret_cast<unsigned *>(sid + cap);
    int *sh_start = reinterpret_cast<int *>(ring + RING);
    int *sh_pre = sh_start + 32;

    // a halted step-graph replay leaves the step to the host
    if (a.st->halt) return;
    const int slot = owner_slot((int)blockIdx.x);
    if (slot >= a.n) return;
    const int p = a.cell_idx[slot];
    const int lane = lane_id();
    const double xp = a.x[p], yp = a.y[p], zp = a.z[p];
    const SysGrid sg = grid_of_particle(a.g, p);

    // 1. neighbour list N+(p)
    const Candidates cand = gather_candidates(a, sg, xp, yp, zp, sh_start, sh_pre);
    int len = 0;
    for (int t0 = 0; t0 < cand.total; t0 += WAVE) {
        const int t = t0 + lane;
        bool ok = t < cand.total;
        int q = 0;
        double qx = 0, qy = 0, qz = 0;
        if (ok) {
            q = candidate(a, cand, sh_start, sh_pre, t);
            ok = q > p;
            if (ok) {
                qx = a.x[q];
                qy = a.y[q];
                qz = a.z[q];
                ok = within_cutoff(a, sg, xp, yp, zp, qx, qy, qz);
            }
        }
        const unsigned long long m = __ballot(ok);
        const int pos = len + __popcll(m & __lanemask_lt());
        if (ok && pos < cap) {
            sx[pos] = qx;
            sy[pos] = qy;
            sz[pos] = qz;
            gx_[pos] = 0.0;
            gy_[pos] = 0.0;
            gz_[pos] = 0.0;
            sid[pos] = q;
        }
        len += __popcll(m);
    }
    if (lane == 0) note_list_length(a.st, len, cap);
    if (len > cap) {
        if (lane == 0) atomicOr(&a.st->flags[2], 1);
        return;
    }
    __syncthreads();

    // 2. pairs of N+(p)
    const T nu = (T)a.nu, K = (T)(0.375 * a.nu);
    double gpx = 0.0, gpy = 0.0, gpz = 0.0, esum = 0.0;
    unsigned cnt = 0;
    int head = 0, qn = 0;

    auto flush = [&](int nq) {
        __syncthreads();
        if (lane < nq) {
            const unsigned e = ring[(head + lane) & (RING - 1)];
            const int ia = (int)(e & 0xffffu), ib = (int)(e >> 16);
            const double xj = sx[ia], yj = sy[ia], zj = sz[ia];
            const double xk = sx[ib], yk = sy[ib], zk = sz[ib];
            const double djx = xj - xp, djy = yj - yp, djz = zj - zp;
            const double dkx = xk - xp, dky = yk - yp, dkz = zk - zp;
            const double ex = xk - xj, ey = yk - yj, ez = zk - zj;
            const double u = __builtin_fma(djx, djx, __builtin_fma(djy, djy, djz * djz));
            const double v = __builtin_fma(dkx, dkx, __builtin_fma(dky, dky, dkz * dkz));
            const double w = __builtin_fma(ex, ex, __builtin_fma(ey, ey, ez * ez));
            double cu, cv, cw;
            T et;
            at_coeffs_from64<T, GRAD>(u, v, w, nu, K, cu, cv, cw, et);
            if (GRAD && SCALE) {
                // lengths of the three gradients (each particle's contribution)
                const double tx = cw * ex, ty = cw * ey, tz = cw * ez;
                const double jx = __builtin_fma(cu, djx, -tx), jy = __builtin_fma(cu, djy, -ty),
                             jz = __builtin_fma(cu, djz, -tz);
                const double kx = __builtin_fma(cv, dkx, tx), ky = __builtin_fma(cv, dky, ty),
                             kz = __builtin_fma(cv, dkz, tz);
                const double ix = jx + kx, iy = jy + ky, iz = jz + kz;
                gpx += __builtin_sqrt(ix * ix + iy * iy + iz * iz);
                lds_add(&gx_[ia], __builtin_sqrt(jx * jx + jy * jy + jz * jz));
                lds_add(&gx_[ib], __builtin_sqrt(kx * kx + ky * ky + kz * kz));
            } else if (GRAD) {
                gpx = __builtin_fma(-cu, djx, __builtin_fma(-cv, dkx, gpx));
                gpy = __builtin_fma(-cu, djy, __builtin_fma(-cv, dky, gpy));
                gpz = __builtin_fma(-cu, djz, __builtin_fma(-cv, dkz, gpz));
                // the j and k terms share the products cw * e: (-cw) e = -(cw e) exactly
                const double tx = cw * ex, ty = cw * ey, tz = cw * ez;
                lds_add(&gx_[ia], __builtin_fma(cu, djx, -tx));
                lds_add(&gy_[ia], __builtin_fma(cu, djy, -ty));
                lds_add(&gz_[ia], __builtin_fma(cu, djz, -tz));
                lds_add(&gx_[ib], __builtin_fma(cv, dkx, tx));
                lds_add(&gy_[ib], __builtin_fma(cv, dky, ty));
                lds_add(&gz_[ib], __builtin_fma(cv, dkz, tz));
            }
            esum += (double)et;
            ++cnt;
        }
        __syncthreads();
        head = (head + nq) & (RING - 1);
        qn -= nq;
    };

    const int H = (len - 1) / 2;
    const int full = H * len;
    const int total = len * (len - 1) / 2;
    const float inv_len = len > 0 ? 1.0f / (float)len : 0.0f;
    for (int t0 = 0; t0 < total; t0 += WAVE) {
        const int t = t0 + lane;
        bool ok = t < total;
        int ia = 0, ib = 0;
        if (ok) {
            if (t < full) {
                // t = d * len + r with 0 <= r < len: float estimate, exact fix-up
                int d = (int)((float)t * inv_len);
                int r = t - d * len;
                while (r < 0) {
                    --d;
                    r += len;
                }
                while (r >= len) {
                    ++d;
                    r -= len;
                }
                ia = r;
                ib = ia + d + 1;
                if (ib >= len) ib -= len;
            } else {
                ia = t - full;
                ib = ia + len / 2;
            }
            ok = within_cutoff(a, sg, sx[ia], sy[ia], sz[ia], sx[ib], sy[ib], sz[ib]);
        }
        const unsigned long long m = __ballot(ok);
        if (ok) ring[(head + qn + __popcll(m & __lanemask_lt())) & (RING - 1)] = (unsigned)ia | ((unsigned)ib << 16);
        qn += __popcll(m);
        if (qn >= WAVE) flush(WAVE);
    }
    if (qn > 0) flush(qn);

    // 3. write back
    if (GRAD) {
        for (int t = lane; t < len; t += WAVE) {
            const int q = sid[t];
            unsafeAtomicAdd(&a.fx[q], gx_[t]);
            unsafeAtomicAdd(&a.fy[q], gy_[t]);
            unsafeAtomicAdd(&a.fz[q], gz_[t]);
        }
        gpx = wave_sum(gpx);
        gpy = wave_sum(gpy);
        gpz = wave_sum(gpz);
        if (lane == 0) {
            unsafeAtomicAdd(&a.fx[p], gpx);
            unsafeAtomicAdd(&a.fy[p], gpy);
            unsafeAtomicAdd(&a.fz[p], gpz);
        }
    }
    if (a.want_energy) esum = wave_sum(esum);
    const unsigned long long c = wave_sum_u64(cnt);
    if (lane == 0) {
        if (a.want_energy) add_energy(a.st, esum);
        add_count(a.st, a.count_slot, c);
        add_rep(a, sg.r, c, esum);
    }
}

// ---------------------------------------------------------------------------
// Fast cutoff sweep, one cell per workgroup (the default fast kernel).
//
// Ownership: each accepted triplet belongs to the one particle that sees both
// others "forward": in its own cell at a later slot, or in a forward
// neighbour cell (the half shell of positive stencil offsets, or of larger
// cell index on grids with a 3-cell dimension; see below).  One workgroup
// per cell c stages in LDS c's particles followed by those particles of c's
// forward cells that lie within the cutoff of c's box: every particle
// forward of an owner in c and within the cutoff of it is in this range.
// Each wave then claims owners i of cell c, collects i's candidates t > i
// within the cutoff (a contiguous LDS range, no search) into a contiguous
// per-wave list, then walks the list's unordered pairs in folded order, 64
// at a time, testing the third side and evaluating each passing triplet in
// place.  The cutoff tests use the periodic form, or plain differences where
// the staged particles are close enough for the two to agree bit for bit.
// The j and k gradients go to per-wave list accumulators (LDS atomics),
// folded into the staged particles' accumulators once per owner; the
// workgroup adds those to global memory once per staged particle that
// received any.  The owners of a neighbourhood larger than NBCAP, and an
// owner whose list is longer than LCAP, are queued for the dense-owner
// kernel below instead.

constexpr int CELL_WAVES = 4;
constexpr int NBCAP = 256;
constexpr int LCAP = 96;   // list positions fit in the 8-bit halves of a queue entry
constexpr int RQ = 128;    // per-wave queue of passing pairs (>= 2 x 64)

struct CellsDev {
    CutDev c;
    const double *xs, *ys, *zs;
    int ncell;
    int *owners;   // owners handed to the dense-owner kernel (st->dense_n of them)
    int dense_on;  // 0: the dense-owner kernel does not follow; an owner that needs it fails the pass
    const int *order;  // each XCD's run of cells, heaviest first (k_cell_scatter)
    unsigned *cell_work;  // each cell's accepted triplets, the next sweep's dispatch key
    SlotsDev sl;          // the slotted layout (cap > 0: staged from its slot records)
};

// An owner that needs the dense-owner kernel when none follows this launch:
// the pass fails (the update skips itself) and asks the host for a re-run
// with the dense-owner kernel.
__device__ __forceinline__ void want_dense_kernel(Status *st)
{
    atomicOr(&st->want_dense, 1);
    atomicOr(&st->flags[2], 1);
}

// The staged cells of cell c (one wave): c first, then its forward
// neighbour cells.  Publishes each cell's (start, staged prefix) in cst/cpre,
// cpre[nk] = total, and info = {own_start, own_cnt, nk, total}.  With
// scnt (the slotted layout, scap slots per cell) cell k starts at slot
// k scap and holds scnt[k] particles.
//
// Which neighbour cells are "forward".  With >= 4 cells (or 1) in every
// dimension, pairwise-adjacent cells lie in one unwrapped 2x2x2 block, so
// orienting each pair of cells by the lexicographic sign of the stencil
// offset (dz, dy, dx) is acyclic on every triplet: the 13 positive offsets
// (the half shell) give each triplet exactly one owner.  With exactly 3 cells
// in a dimension three cells can wrap into a cycle, so the cell index orders
// them instead.
__device__ void stage_cells(const CutDev &A, const SysGrid &sg, int c, int *cst, int *cpre, int *info,
                            const int *scnt = nullptr, int scap = 0)
{
    const int lane = lane_id();
    const int ncx = sg.nc[0], ncy = sg.nc[1], ncz = sg.nc[2];
    const int lc = c - sg.cbase;
    const int cx = lc % ncx, cy = (lc / ncx) % ncy, cz = lc / (ncx * ncy);
    const int wx = ncx >= 3 ? 3 : 1, wy = ncy >= 3 ? 3 : 1, wz = ncz >= 3 ? 3 : 1;
    const int nst = wx * wy * wz;
    int cc = -1;
    if (lane < nst) {
        const int ox = lane % wx, oy = (lane / wx) % wy, oz = lane / (wx * wy);
        int xx = cx + (wx == 3 ? ox - 1 : 0), yy = cy + (wy == 3 ? oy - 1 : 0);
        int zz = cz + (wz == 3 ? oz - 1 : 0);
        xx = xx < 0 ? xx + ncx : (xx >= ncx ? xx - ncx : xx);
        yy = yy < 0 ? yy + ncy : (yy >= ncy ? yy - ncy : yy);
        zz = zz < 0 ? zz + ncz : (zz >= ncz ? zz - ncz : zz);
        cc = sg.cbase + (zz * ncy + yy) * ncx + xx;
    }
    const bool lexi = (ncx == 1 || ncx >= 4) && (ncy == 1 || ncy >= 4) && (ncz == 1 || ncz >= 4);
    const int centre = nst / 2;
    const bool fwd = cc >= 0 && (lexi ? lane > centre : cc > c);
    int own_start, own_cnt, st = 0, n_in = 0;
    if (scnt) {
        own_start = c * scap;
        own_cnt = min(scnt[c], scap);
        if (fwd) {
            st = cc * scap;
            n_in = min(scnt[cc], scap);
        }
    } else {
        own_start = A.cell_start[c];
        own_cnt = A.cell_start[c + 1] - own_start;
        if (fwd) {
            st = A.cell_start[cc];
            n_in = A.cell_start[cc + 1] - st;
        }
    }
    const unsigned long long m = __ballot(fwd);
    int incl = n_in;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    const int nfwd = __popcll(m);
    if (fwd) {
        const int k = 1 + lane_prefix(m);
        cst[k] = st;
        cpre[k] = own_cnt + incl - n_in;
    }
    const int total = own_cnt + __shfl(incl, WAVE - 1);
    if (lane == 0) {
        cst[0] = own_start;
        cpre[0] = 0;
        cpre[1 + nfwd] = total;
        info[0] = own_start;
        info[1] = own_cnt;
        info[2] = 1 + nfwd;
        info[3] = total;
    }
}

// Squared distance for the cutoff test.  PER: the periodic form
// min(t, L - t) per component (Algorithm.cpp's cutoff test).  !PER: the plain
// difference, which gives the same bits whenever every |difference| is below
// L/2, because min(t, L - t) is then t itself and t * t = d * d.
template <bool PER>
__device__ __forceinline__ double cut_sq(double ax, double ay, double az, double bx, double by, double bz, double Lx,
                                         double Ly, double Lz)
{
    if constexpr (PER) {
        return sqr_dist_periodic_min(ax, ay, az, bx, by, bz, Lx, Ly, Lz);
    } else {
        const double dx = bx - ax, dy = by - ay, dz = bz - az;
        return dx * dx + dy * dy + dz * dz;
    }
}

// Global slot (cell order) of staged index t: cpre is nondecreasing with
// cpre[0] = 0, so the staged cell holding t is the number of later cells
// starting at or before t (independent LDS reads, no dependent chain).
__device__ __forceinline__ int staged_slot(const int *cst, const int *cpre, int nk, int t)
{
    int k = 0;
    for (int j = 1; j < nk; ++j) k += cpre[j] <= t ? 1 : 0;
    return cst[k] + (t - cpre[k]);
}

// Cell for workgroup b, or -1: workgroups are dispatched round-robin over
// the 8 XCDs (b % 8), so each XCD gets a contiguous run of cells, whose
// neighbourhoods overlap in its own L2, of about an eighth of the work, and
// takes them heaviest first (order_runs).
__device__ __forceinline__ int cell_for_block(int b, int ncell, const int *order)
{
    const int cap = cell_run_cap(ncell);
    const int r = b & 7, k = b >> 3;
    if (k >= cap || k >= order[8 * cap + r]) return -1;
    return order[r * cap + k];
}

template <class T, bool GRAD, bool WANT_E>
__global__ __launch_bounds__(CELL_WAVES *WAVE, 4) void k_p3bca_cells(CellsDev a)
{
    // one block per array family, so that the y, z and gradient arrays sit
    // at constant offsets from x and their addresses fold into the LDS
    // instructions' immediate offsets
    __shared__ double stg[6][NBCAP];
    double *nx = stg[0], *ny = stg[1], *nz = stg[2];
    double *gxs = stg[3], *gys = stg[4], *gzs = stg[5];
    __shared__ int nid[NBCAP];
    __shared__ unsigned short wl[CELL_WAVES][LCAP];
    __shared__ double wls[CELL_WAVES][6][LCAP];
    __shared__ unsigned short ring[CELL_WAVES][RQ];
    __shared__ int cst[32], cpre[33], sinfo[4];
    __shared__ int s_next, s_far, s_wcnt[2][CELL_WAVES];
    __shared__ unsigned s_work;

    const CutDev &A = a.c;
    const int c = cell_for_block(blockIdx.x, a.ncell, a.order);
    if (c < 0) return;
    const int tid = threadIdx.x, lane = lane_id(), wv = tid / WAVE;
    // A halted step-graph replay leaves the step to the host.  Non-finite
    // positions pair with everything: leave the sweep to the per-owner
    // kernel.  (Earlier launches set both, so every thread reads the same
    // values.)
    if (A.st->halt) return;
    // The slotted layout: this sweep's counts, and cell c's counter of the
    // next binning cleared (every cell has one workgroup).  A cell that had
    // more particles than slots fails the pass; the host re-runs the step
    // on the contiguous layout.
    const int *scnt = nullptr;
    if (a.sl.cap > 0) {
        const unsigned e = a.sl.state[0];
        scnt = a.sl.cnt + (e & 1u) * a.ncell;
        if (tid == 0) a.sl.cnt[((e + 1u) & 1u) * a.ncell + c] = 0;
        if (A.st->pad_over) {
            if (tid == 0) atomicOr(&A.st->flags[2], 1);  // the update kernel skips itself
            return;
        }
    }
    if (A.st->flags[0] != 0) {
        if (tid == 0) {
            atomicOr(&A.st->flags[2], 1);  // the update kernel skips itself
            atomicOr(&A.st->flags[3], 1);  // the host re-runs the sweep with the per-owner kernel
        }
        return;
    }
    const T nu = (T)A.nu, K = (T)(0.375 * A.nu);
    const double c2 = A.c2;
    const SysGrid sg = grid_of_cell(A.g, c);
    const double Lx = sg.L[0], Ly = sg.L[1], Lz = sg.L[2];
    double esum = 0.0;
    unsigned cnt = 0;

    // 1. the staged cells: c first, then its forward neighbour cells.
    // Staging is a chain of memory round trips; at the top wave priority
    // (until the owners' loop) its loads go out ahead of the other
    // workgroups' arithmetic (C1 sweep 120.7 -> 117.5 us; the same priority
    // for the evaluations made it 130 us, for the owners' list builds 121 us).
    __builtin_amdgcn_s_setprio(3);
    if (wv == 0) stage_cells(A, sg, c, cst, cpre, sinfo, scnt, a.sl.cap);
    __syncthreads();
    const int cand = sinfo[3], own = sinfo[1], nk = sinfo[2];

    // 2. stage, in order, c's particles and those particles of the forward
    // cells that lie within the cutoff of c's box (periodic box distance,
    // with a relative margin of 1e-6: no owner in c reaches the others), and
    // clear their accumulators.  c's own particles come first, so owner i is
    // staged particle i.  With at least 7 cells per dimension, staged
    // particles whose unwrapped cell coordinates all lie within
    // [c - 1.01, c + 2.01) are less than 3.02 cell widths (< L/2) apart in
    // every dimension, so the cutoff tests may use plain differences
    // (cut_sq<false>); a particle outside (a wrapped neighbour cell, or one
    // outside the box) sends the workgroup to the periodic form.
    if (tid == 0) {
        s_next = 0;
        s_work = 0;
        s_far = (sg.nc[0] < 7 || sg.nc[1] < 7 || sg.nc[2] < 7) ? 1 : 0;
    }
    int total = 0;
    {
        const int lc = c - sg.cbase;
        const double ccx = (double)(lc % sg.nc[0]), ccy = (double)((lc / sg.nc[0]) % sg.nc[1]),
                     ccz = (double)(lc / (sg.nc[0] * sg.nc[1]));
        const double wx = Lx / sg.nc[0], wy = Ly / sg.nc[1], wz = Lz / sg.nc[2];
        const double ix = 1.0 / wx, iy = 1.0 / wy, iz = 1.0 / wz;
        const double mx = sg.lo[0] + (ccx + 0.5) * wx, my = sg.lo[1] + (ccy + 0.5) * wy,
                     mz = sg.lo[2] + (ccz + 0.5) * wz;
        const double reach2 = c2 * (1.0 + 1e-6);
        auto gap = [](double d, double L, double w) {
            d -= L * __builtin_rint(d / L);
            const double g = __builtin_fabs(d) - 0.5 * w;
            return g > 0.0 ? g : 0.0;
        };
        // two candidates per thread and round (a cell's neighbourhood has
        // ~265 on average), with every load issued before the first use
        bool far = false;
        constexpr int NT = CELL_WAVES * WAVE;
        for (int t0 = 0; t0 < cand; t0 += 2 * NT) {
            bool keep[2];
            double x[2], y[2], z[2];
            int q[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int t = t0 + h * NT + tid;
                keep[h] = false;
                x[h] = y[h] = z[h] = 0.0;
                q[h] = 0;
                if (t < cand) {
                    const int slot = staged_slot(cst, cpre, nk, t);
                    if (scnt) {
                        const double4 rc = a.sl.rec[slot];
                        x[h] = rc.x;
                        y[h] = rc.y;
                        z[h] = rc.z;
                        q[h] = (int)__double_as_longlong(rc.w);
                    } else {
                        x[h] = a.xs[slot];
                        y[h] = a.ys[slot];
                        z[h] = a.zs[slot];
                        q[h] = A.cell_idx[slot];
                    }
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int t = t0 + h * NT + tid;
                if (t < cand) {
                    const double gx = gap(x[h] - mx, Lx, wx), gy = gap(y[h] - my, Ly, wy),
                                 gz = gap(z[h] - mz, Lz, wz);
                    keep[h] = t < own || gx * gx + gy * gy + gz * gz <= reach2;
                }
            }
            const unsigned long long m0 = __ballot(keep[0]), m1 = __ballot(keep[1]);
            if (lane == 0) {
                s_wcnt[0][wv] = __popcll(m0);
                s_wcnt[1][wv] = __popcll(m1);
            }
            __syncthreads();
            // the kept candidates of the first half, then of the second, in order
            int pos[2] = {total + lane_prefix(m0), lane_prefix(m1)};
            for (int h = 0; h < 2; ++h) {
                if (h == 1) pos[1] += total;
                for (int w = 0; w < CELL_WAVES; ++w) {
                    const int k = s_wcnt[h][w];
                    if (w < wv) pos[h] += k;
                    total += k;
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (keep[h] && pos[h] < NBCAP) {
                    const int p = pos[h];
                    nx[p] = x[h];
                    ny[p] = y[h];
                    nz[p] = z[h];
                    nid[p] = q[h];
                    gxs[p] = 0.0;
                    gys[p] = 0.0;
                    gzs[p] = 0.0;
                    const double ux = (x[h] - sg.lo[0]) * ix - ccx, uy = (y[h] - sg.lo[1]) * iy - ccy,
                                 uz = (z[h] - sg.lo[2]) * iz - ccz;
                    far = far || !(ux >= -1.01 && ux < 2.01 && uy >= -1.01 && uy < 2.01 && uz >= -1.01 && uz < 2.01);
                }
            }
            __syncthreads();  // s_wcnt is reused
        }
        if (far) s_far = 1;
    }
    if (total > NBCAP) {
        // hand every owner of c to the dense-owner kernel (which makes c a
        // light workgroup: it goes last in the next sweep's order)
        if (tid == 0) a.cell_work[c] = 1;
        if (!a.dense_on) {
            if (tid == 0) want_dense_kernel(A.st);
            return;
        }
        if (tid == 0) s_next = atomicAdd(&A.st->dense_n, own);
        __syncthreads();
        for (int t = tid; t < own; t += CELL_WAVES * WAVE) a.owners[s_next + t] = sinfo[0] + t;
        return;
    }
    __syncthreads();

    // 3. owners of cell c, claimed one at a time by the waves; the cutoff
    // tests take the plain form unless the staging found a far particle
    unsigned short *list = wl[wv];
    double *qx = wls[wv][0], *qy = wls[wv][1], *qz = wls[wv][2];
    double *hx = wls[wv][3], *hy = wls[wv][4], *hz = wls[wv][5];
    auto owners = [&](auto per) {
        constexpr bool PER = decltype(per)::value;
        for (;;) {
            int i = 0;
            if (lane == 0) i = atomicAdd(&s_next, 1);
            i = __builtin_amdgcn_readfirstlane(i);
            if (i >= own) break;
            const double xi = nx[i], yi = ny[i], zi = nz[i];

            // the owner's forward candidates within the cutoff, in staging
            // order, copied to this wave's contiguous list
            int len = 0;
            for (int t0 = i + 1; t0 < total; t0 += WAVE) {
                const int t = t0 + lane;
                double xt = 0.0, yt = 0.0, zt = 0.0;
                bool ok = t < total;
                if (ok) {
                    xt = nx[t];
                    yt = ny[t];
                    zt = nz[t];
                    ok = !(cut_sq<PER>(xi, yi, zi, xt, yt, zt, Lx, Ly, Lz) > c2);
                }
                const unsigned long long m = __ballot(ok);
                const int pos = len + lane_prefix(m);
                if (ok && pos < LCAP) {
                    list[pos] = (unsigned short)t;
                    qx[pos] = xt;
                    qy[pos] = yt;
                    qz[pos] = zt;
                    hx[pos] = 0.0;
                    hy[pos] = 0.0;
                    hz[pos] = 0.0;
                }
                len += __popcll(m);
            }
            if (len > LCAP) {
                // too long for this wave's list: the dense-owner kernel takes i
                if (lane == 0) {
                    if (a.dense_on) a.owners[atomicAdd(&A.st->dense_n, 1)] = sinfo[0] + i;
                    else want_dense_kernel(A.st);
                }
                continue;
            }
            wave_sync();

            // Every unordered pair of the list once, in folded order: item
            // t = (o - 1) len + r pairs r with (r + o) mod len, o = 1 .. H;
            // for even len the last len / 2 items pair r with r + len / 2.
            // Lane l takes items l, l + 64, ...; (r, o) advance by
            // (64 mod len, 64 / len) with a carry.  Pairs that pass the test of
            // their third side are queued (in item order, so the queue walks the
            // list nearly in sequence) and evaluated 64 at a time.  Only the j
            // and k gradients are accumulated: the owner's is minus their sum
            // (the triplet's gradients sum to zero), taken from the list at the
            // end.
            const int npairs = len * (len - 1) / 2;
            if (npairs > 0) {
                unsigned short *rq = ring[wv];
                int head = 0, qn = 0;
                auto flush = [&](int nq) {
                    wave_sync();
                    if (lane < nq) {
                        const unsigned e = rq[(head + lane) & (RQ - 1)];
                        const int r = (int)(e & 0xffu), kk = (int)(e >> 8);
                        const double xj = qx[r], yj = qy[r], zj = qz[r];
                        const double xk = qx[kk], yk = qy[kk], zk = qz[kk];
                        const double djx = xj - xi, djy = yj - yi, djz = zj - zi;
                        const double dkx = xk - xi, dky = yk - yi, dkz = zk - zi;
                        const double ex = xk - xj, ey = yk - yj, ez = zk - zj;
                        const double u64 = __builtin_fma(djx, djx, __builtin_fma(djy, djy, djz * djz));
                        const double v64 = __builtin_fma(dkx, dkx, __builtin_fma(dky, dky, dkz * dkz));
                        const double w64 = __builtin_fma(ex, ex, __builtin_fma(ey, ey, ez * ez));
                        double cu, cv, cw;
                        T et;
                        at_coeffs_from64<T, GRAD>(u64, v64, w64, nu, K, cu, cv, cw, et);
                        if (GRAD) {
                            // the j and k terms share the products cw * e: (-cw) e = -(cw e) exactly
                            const double tx = cw * ex, ty = cw * ey, tz = cw * ez;
                            lds_add(&hx[r], __builtin_fma(cu, djx, -tx));
                            lds_add(&hy[r], __builtin_fma(cu, djy, -ty));
                            lds_add(&hz[r], __builtin_fma(cu, djz, -tz));
                            lds_add(&hx[kk], __builtin_fma(cv, dkx, tx));
                            lds_add(&hy[kk], __builtin_fma(cv, dky, ty));
                            lds_add(&hz[kk], __builtin_fma(cv, dkz, tz));
                        }
                        if (WANT_E) esum += (double)et;
                        ++cnt;
                    }
                    wave_sync();
                    head = (head + nq) & (RQ - 1);
                    qn -= nq;
                };
                // 64 / len and lane / len from one FP32 reciprocal: for
                // len <= LCAP the fraction of (k + 0.5) / len stays 0.5 / len
                // away from an integer, far beyond the FP32 rounding
                const float inv = 1.0f / (float)len;
                const int q64 = (int)(64.5f * inv), r64 = WAVE - q64 * len;
                int o = (int)(((float)lane + 0.5f) * inv);
                int r = lane - o * len;
                ++o;
                auto test = [&](bool live) {
                    int kk = r + o;
                    kk = kk >= len ? kk - len : kk;
                    kk = live ? kk : r;  // lanes past the last item read their own entry
                    const bool ok = live && !(cut_sq<PER>(qx[r], qy[r], qz[r], qx[kk], qy[kk], qz[kk], Lx, Ly, Lz) > c2);
                    const unsigned long long m = __ballot(ok);
                    if (ok) rq[(head + qn + lane_prefix(m)) & (RQ - 1)] = (unsigned short)(r | (kk << 8));
                    qn += __popcll(m);
                    if (qn >= WAVE) flush(WAVE);
                    r += r64;
                    o += q64;
                    if (r >= len) {
                        r -= len;
                        ++o;
                    }
                };
                // whole rounds of 64 items, then the last partial round
                const int full = npairs & ~(WAVE - 1);
                for (int t0 = 0; t0 < full; t0 += WAVE) test(true);
                if (full < npairs) test(full + lane < npairs);
                if (qn > 0) flush(qn);
            }
            if (GRAD) {
                // the list's gradients go to the staged accumulators; the owner's
                // is minus their sum (16-lane row sums in the VALU, then one LDS
                // atomic per row)
                wave_sync();
                double gix = 0.0, giy = 0.0, giz = 0.0;
                for (int t = lane; t < len; t += WAVE) {
                    const int q = list[t];
                    const double hxv = hx[t], hyv = hy[t], hzv = hz[t];
                    lds_add(&gxs[q], hxv);
                    lds_add(&gys[q], hyv);
                    lds_add(&gzs[q], hzv);
                    gix -= hxv;
                    giy -= hyv;
                    giz -= hzv;
                }
                gix = row_sum_last(gix);
                giy = row_sum_last(giy);
                giz = row_sum_last(giz);
                if ((lane & 15) == 15) {
                    lds_add(&gxs[i], gix);
                    lds_add(&gys[i], giy);
                    lds_add(&gzs[i], giz);
                }
            }
            wave_sync();
        }
    };
    __builtin_amdgcn_s_setprio(0);
    if (s_far) owners(std::true_type{});
    else owners(std::false_type{});
    __syncthreads();

    // 4. add the staged particles' gradients to global memory.  Staged
    // particles that no owner of c paired with (many in the edge and corner
    // cells) hold exact zeros and are skipped: f starts at +0.0, and adding
    // +-0.0 to it never changes it.
    if (GRAD) {
        for (int t = tid; t < total; t += CELL_WAVES * WAVE) {
            const int q = nid[t];
            const double gx = gxs[t], gy = gys[t], gz = gzs[t];
            if (gx != 0.0 || gy != 0.0 || gz != 0.0) {
                unsafeAtomicAdd(&A.fx[q], gx);
                unsafeAtomicAdd(&A.fy[q], gy);
                unsafeAtomicAdd(&A.fz[q], gz);
            }
        }
    }
    const unsigned long long cw = wave_sum_u64(cnt);
    const double ee = WANT_E ? wave_sum(esum) : 0.0;
    if (lane == 0) {
        if (WANT_E) add_energy(A.st, ee);
        add_count(A.st, A.count_slot, cw);
        add_rep(A, sg.r, cw, ee);
        atomicAdd(&s_work, (unsigned)cw);
    }
    __syncthreads();
    if (tid == 0) a.cell_work[c] = max(s_work, 1u);
}


// ---------------------------------------------------------------------------
// Dense-owner sweep: the owners the cell kernel could not stage.
//
// One owner per workgroup of DW waves (4, 8 or 16: a tuning knob), taken from the queue the
// cell kernel filled.  The owner's list is the same forward set the cell
// kernel would have used (its own cell's later slots, then the forward
// neighbour cells, within the cutoff), gathered from the cell-ordered
// positions into LDS by all waves; the list's unordered pairs are then split
// over the workgroup's lanes in folded order and evaluated in place as in
// the cell kernel.  A list longer than DCAP raises the overflow flags (the
// host then re-runs the sweep with the per-owner kernel).

constexpr int DCAP = 1440;  // two workgroups of 4 or 8 waves (one of 16) fit a CU's 160 KB of LDS

struct DenseDev {
    CutDev c;
    const double *xs, *ys, *zs;
    const int *cell_of;
    const int *owners;
    int *queue;  // [0] next owner ticket, [1] finished workgroups (both return to 0)
};

size_t dense_lds() { return (size_t)DCAP * (6 * sizeof(double) + sizeof(int)); }

template <class T, bool GRAD, int DW>
__global__ __launch_bounds__(DW *WAVE) void k_p3bca_dense(DenseDev a)
{
    extern __shared__ double dsm[];
    double *qx = dsm, *qy = qx + DCAP, *qz = qy + DCAP;
    double *hx = qz + DCAP, *hy = hx + DCAP, *hz = hy + DCAP;
    int *qid = reinterpret_cast<int *>(hz + DCAP);
    __shared__ int cst[32], cpre[33], sinfo[4];
    __shared__ int s_item, s_len;
    __shared__ double s_gi[3];
    __shared__ unsigned dring[DW][RQ];

    const CutDev &A = a.c;
    const int tid = threadIdx.x, lane = lane_id(), wv = tid / WAVE;
    constexpr int NT = DW * WAVE;
    const T nu = (T)A.nu, K = (T)(0.375 * A.nu);
    const double c2 = A.c2;
    const int nowners = A.st->halt ? 0 : A.st->dense_n;  // written by the cell kernel before this launch
    if (nowners == 0) return;                            // the common case: no queue traffic at all
    double esum = 0.0;
    unsigned cnt = 0;
    bool overflow = false;

    for (;;) {
        if (tid == 0) s_item = atomicAdd(&a.queue[0], 1);
        __syncthreads();
        const int item = s_item;
        if (item >= nowners) break;
        const int slot = a.owners[item];
        const int p = A.cell_idx[slot];
        const int c = a.cell_of[p];
        const SysGrid sg = grid_of_cell(A.g, c);
        const double Lx = sg.L[0], Ly = sg.L[1], Lz = sg.L[2];
        if (wv == 0) stage_cells(A, sg, c, cst, cpre, sinfo);
        if (tid == 0) {
            s_len = 0;
            s_gi[0] = s_gi[1] = s_gi[2] = 0.0;
        }
        __syncthreads();
        const int i = slot - sinfo[0], nk = sinfo[2], total = sinfo[3];
        const double xi = a.xs[slot], yi = a.ys[slot], zi = a.zs[slot];

        // the owner's forward list
        for (int t0 = i + 1; t0 < total; t0 += NT) {
            const int t = t0 + tid;
            bool ok = t < total;
            double xt = 0.0, yt = 0.0, zt = 0.0;
            int st = 0;
            if (ok) {
                st = staged_slot(cst, cpre, nk, t);
                xt = a.xs[st];
                yt = a.ys[st];
                zt = a.zs[st];
                ok = !(sqr_dist_periodic_min(xi, yi, zi, xt, yt, zt, Lx, Ly, Lz) > c2);
            }
            const unsigned long long m = __ballot(ok);
            int base = 0;
            if (lane == 0 && m) base = atomicAdd(&s_len, __popcll(m));
            base = __shfl(base, 0);
            const int pos = base + lane_prefix(m);
            if (ok && pos < DCAP) {
                qx[pos] = xt;
                qy[pos] = yt;
                qz[pos] = zt;
                hx[pos] = 0.0;
                hy[pos] = 0.0;
                hz[pos] = 0.0;
                qid[pos] = A.cell_idx[st];
            }
        }
        __syncthreads();
        const int len = s_len;
        if (len > DCAP) {
            overflow = true;
            __syncthreads();
            continue;
        }

        // the list's unordered pairs, folded, split over the workgroup's
        // lanes; each wave queues its passing pairs and evaluates them 64 at
        // a time (as the cell kernel does)
        double gix = 0.0, giy = 0.0, giz = 0.0;
        const int npairs = len * (len - 1) / 2;
        if (npairs > 0) {
            unsigned *rq = dring[wv];
            int head = 0, qw = 0;
            auto flush = [&](int nq) {
                wave_sync();
                if (lane < nq) {
                    const unsigned e = rq[(head + lane) & (RQ - 1)];
                    const int r = (int)(e & 0xffffu), kk = (int)(e >> 16);
                    const double xj = qx[r], yj = qy[r], zj = qz[r];
                    const double xk = qx[kk], yk = qy[kk], zk = qz[kk];
                    const double djx = xj - xi, djy = yj - yi, djz = zj - zi;
                    const double dkx = xk - xi, dky = yk - yi, dkz = zk - zi;
                    const double ex = xk - xj, ey = yk - yj, ez = zk - zj;
                    const double u64 = __builtin_fma(djx, djx, __builtin_fma(djy, djy, djz * djz));
                    const double v64 = __builtin_fma(dkx, dkx, __builtin_fma(dky, dky, dkz * dkz));
                    const double w64 = __builtin_fma(ex, ex, __builtin_fma(ey, ey, ez * ez));
                    double cu, cv, cw;
                    T et;
                    at_coeffs_from64<T, GRAD>(u64, v64, w64, nu, K, cu, cv, cw, et);
                    if (GRAD) {
                        gix = __builtin_fma(-cu, djx, __builtin_fma(-cv, dkx, gix));
                        giy = __builtin_fma(-cu, djy, __builtin_fma(-cv, dky, giy));
                        giz = __builtin_fma(-cu, djz, __builtin_fma(-cv, dkz, giz));
                        // the j and k terms share the products cw * e: (-cw) e = -(cw e) exactly
                        const double tx = cw * ex, ty = cw * ey, tz = cw * ez;
                        lds_add(&hx[r], __builtin_fma(cu, djx, -tx));
                        lds_add(&hy[r], __builtin_fma(cu, djy, -ty));
                        lds_add(&hz[r], __builtin_fma(cu, djz, -tz));
                        lds_add(&hx[kk], __builtin_fma(cv, dkx, tx));
                        lds_add(&hy[kk], __builtin_fma(cv, dky, ty));
                        lds_add(&hz[kk], __builtin_fma(cv, dkz, tz));
                    }
                    esum += (double)et;
                    ++cnt;
                }
                wave_sync();
                head = (head + nq) & (RQ - 1);
                qw -= nq;
            };
            const int qn = NT / len, rn = NT - qn * len;
            int o = tid / len;
            int r = tid - o * len;
            ++o;
            for (int t0 = 0; t0 < npairs; t0 += NT) {
                bool ok = false;
                int kk = 0;
                if (t0 + tid < npairs) {
                    kk = r + o;
                    kk = kk >= len ? kk - len : kk;
                    ok = !(sqr_dist_periodic_min(qx[r], qy[r], qz[r], qx[kk], qy[kk], qz[kk], Lx, Ly, Lz) > c2);
                }
                const unsigned long long m = __ballot(ok);
                if (ok) rq[(head + qw + lane_prefix(m)) & (RQ - 1)] = (unsigned)r | ((unsigned)kk << 16);
                qw += __popcll(m);
                if (qw >= WAVE) flush(WAVE);
                r += rn;
                o += qn;
                if (r >= len) {
                    r -= len;
                    ++o;
                }
            }
            if (qw > 0) flush(qw);
        }
        if (GRAD) {
            gix = wave_total(gix);
            giy = wave_total(giy);
            giz = wave_total(giz);
            if (lane == 0) {
                lds_add(&s_gi[0], gix);
                lds_add(&s_gi[1], giy);
                lds_add(&s_gi[2], giz);
            }
        }
        __syncthreads();
        if (GRAD) {
            for (int t = tid; t < len; t += NT) {
                const int q = qid[t];
                unsafeAtomicAdd(&A.fx[q], hx[t]);
                unsafeAtomicAdd(&A.fy[q], hy[t]);
                unsafeAtomicAdd(&A.fz[q], hz[t]);
            }
            if (tid == 0) {
                unsafeAtomicAdd(&A.fx[p], s_gi[0]);
                unsafeAtomicAdd(&A.fy[p], s_gi[1]);
                unsafeAtomicAdd(&A.fz[p], s_gi[2]);
            }
        }
        if (A.g.nrep > 1) {
            const unsigned long long cc = wave_sum_u64(cnt);
            const double ee = A.want_energy ? wave_sum(esum) : 0.0;
            if (lane == 0) {
                if (A.want_energy) add_energy(A.st, ee);
                add_count(A.st, A.count_slot, cc);
                add_rep(A, sg.r, cc, ee);
            }
            cnt = 0;
            esum = 0.0;
        }
        __syncthreads();
    }

    if (overflow && tid == 0) {
        atomicOr(&A.st->flags[2], 1);
        atomicOr(&A.st->flags[3], 1);
    }
    if (A.want_energy) esum = wave_sum(esum);
    const unsigned long long cw = wave_sum_u64(cnt);
    if (lane == 0) {
        if (A.want_energy) add_energy(A.st, esum);
        add_count(A.st, A.count_slot, cw);
    }
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        const int done = atomicAdd(&a.queue[1], 1);
        if (done == (int)gridDim.x - 1) {
            atomicExch(&a.queue[0], 0);
            atomicExch(&a.queue[1], 0);
            atomicExch(&A.st->dense_n, 0);
        }
    }
}

// ---------------------------------------------------------------------------
// Bit-exact ordered sweep.
//
// Reproduces the sequential accumulation the reference performs into
// b_r[p].f (Algorithm.cpp:72-139) for owner p and role r:
//   role 0 (p = i): pairs (j, k), p < j < k, lexicographic;
//   role 1 (p = j): pairs (i, k), i < p < k, lexicographic;
//   role 2 (p = k): pairs (i, j), i < j < p, lexicographic;
// restricted to pairs passing the cutoff test.  Lanes test 64 consecutive
// pairs of the stream at a time, accepted pairs are appended in stream order
// to an LDS ring, 64 queued triplets are evaluated in parallel with the
// reference's expression, and three lanes fold the contributions into the
// x, y, z accumulators strictly in stream order: acc = acc - c.
// The owner's neighbour list is index-sorted: explicit (LDS) for the cutoff
// algorithm, implicit (every other particle) otherwise.

struct ExactShared {
    int qa[RING], qb[RING];
    double cb[WAVE * 3];
};

struct ImplicitList {
    int p;
    __device__ int operator()(int t) const { return t < p ? t : t + 1; }
};

struct LdsList {
    const int *sid;
    __device__ int operator()(int t) const { return sid[t]; }
};

template <int ROLE, class List>
__device__ void exact_role(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
                           int p, const List &elem, int len, int split, bool use_cut, double c2, double Lx,
                           double Ly, double Lz, double nu, ExactShared &sh, double &acc,
                           unsigned long long &accepted)
{
    const int lane = lane_id();
    const bool nu_ok = nu_inrange(nu);
    const double xp = x[p], yp = y[p], zp = z[p];
    int a0, a1, b1;
    bool tri;
    if constexpr (ROLE == 0) {
        a0 = split; a1 = len; b1 = len; tri = true;
    } else if constexpr (ROLE == 1) {
        a0 = 0; a1 = split; b1 = len; tri = false;
    } else {
        a0 = 0; a1 = split; b1 = split; tri = true;
    }
    acc = 0.0;  // lanes 0, 1, 2 hold x, y, z
    int head = 0, qn = 0;

    auto flush = [&](int cnt) {
        __syncthreads();
        if (lane < cnt) {
            const int s = (head + lane) & (RING - 1);
            const int ea = sh.qa[s], eb = sh.qb[s];
            const double xa = x[ea], ya = y[ea], za = z[ea];
            const double xb = x[eb], yb = y[eb], zb = z[eb];
            double cx, cy, cz;
            if constexpr (ROLE == 0)
                ref_contrib<0>(xp, yp, zp, xa, ya, za, xb, yb, zb, nu, nu_ok, cx, cy, cz);
            else if constexpr (ROLE == 1)
                ref_contrib<1>(xa, ya, za, xp, yp, zp, xb, yb, zb, nu, nu_ok, cx, cy, cz);
            else
                ref_contrib<2>(xa, ya, za, xb, yb, zb, xp, yp, zp, nu, nu_ok, cx, cy, cz);
            sh.cb[lane * 3 + 0] = cx;
            sh.cb[lane * 3 + 1] = cy;
            sh.cb[lane * 3 + 2] = cz;
        }
        __syncthreads();
        if (lane < 3) {
            for (int t = 0; t < cnt; ++t) acc = acc - sh.cb[t * 3 + lane];
        }
        __syncthreads();
        head = (head + cnt) & (RING - 1);
        qn -= cnt;
    };

    for (int ia = a0; ia < a1; ++ia) {
        const int ea = elem(ia);
        const double xa = x[ea], ya = y[ea], za = z[ea];
        const int bstart = tri ? ia + 1 : split;
        for (int b = bstart; b < b1; b += WAVE) {
            const int bl = b + lane;
            bool ok = bl < b1;
            int eb = 0;
            if (ok) {
                eb = elem(bl);
                if (use_cut) ok = !(sqr_dist_periodic(xa, ya, za, x[eb], y[eb], z[eb], Lx, Ly, Lz) > c2);
            }
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const int s = (head + qn + __popcll(m & __lanemask_lt())) & (RING - 1);
                sh.qa[s] = ea;
                sh.qb[s] = eb;
            }
            const int got = __popcll(m);
            qn += got;
            accepted += (unsigned long long)got;
            if (qn >= WAVE) flush(WAVE);
        }
    }
    if (qn > 0) flush(qn);
}

// NATA / AUTA (every triplet): wave w handles role w % 3 of particle w / 3
// (role 2 in reverse so the longest streams of every role start first).
struct AutaExactDev {
    int n;
    const double *x, *y, *z;
    double nu;
    double *facc;
    Status *st;
};

__global__ __launch_bounds__(WAVE) void k_auta_exact(AutaExactDev a)
{
    __shared__ ExactShared sh;
    const int w = blockIdx.x;
    const int role = w % 3, q = w / 3;
    if (q >= a.n) return;
    const int lane = lane_id();
    const int p = role == 2 ? a.n - 1 - q : q;
    const ImplicitList list{p};
    double acc = 0.0;
    unsigned long long accepted = 0;
    if (role == 0)
        exact_role<0>(a.x, a.y, a.z, p, list, a.n - 1, p, false, 0, 0, 0, 0, a.nu, sh, acc, accepted);
    else if (role == 1)
        exact_role<1>(a.x, a.y, a.z, p, list, a.n - 1, p, false, 0, 0, 0, 0, a.nu, sh, acc, accepted);
    else
        exact_role<2>(a.x, a.y, a.z, p, list, a.n - 1, p, false, 0, 0, 0, 0, a.nu, sh, acc, accepted);
    if (lane < 3) a.facc[(long long)(3 * role + lane) * a.n + p] = acc;
    if (role == 0 && lane == 0) add_count(a.st, 0, accepted);
}

// The same role streams over an owner list staged in LDS (sorted ids and
// their positions), with the stream's candidate pairs numbered in stream
// order and tested 64 consecutive items at a time, so every lane tests a
// pair even when the rows are short.  Item t of a triangular stream over m
// list entries is row a, column b > a with t = a (2m - a - 1) / 2 + (b - a - 1);
// of the rectangular stream (role 1), t = a m2 + (b - split).  Accepted
// pairs enter the queue in item order, which is the stream order, so the
// fold below adds the contributions exactly as the nested loops did.
struct StagedList {
    const double *p;  // entry t at p[3t], p[3t + 1], p[3t + 2] (x, y, z side by side: two LDS reads)
};

template <int ROLE>
__device__ void exact_role_staged(const StagedList &L, double xp, double yp, double zp, int len, int split,
                                  bool plain, double c2, double Lx, double Ly, double Lz, double nu, ExactShared &sh,
                                  double &acc, unsigned long long &accepted)
{
    const int lane = lane_id();
    const bool nu_ok = nu_inrange(nu);
    // stream geometry: rows [s, s + m), columns after the row (triangular) or [split, len)
    int s0, m, m2;
    if constexpr (ROLE == 0) {
        s0 = split;
        m = len - split;
        m2 = 0;
    } else if constexpr (ROLE == 1) {
        s0 = 0;
        m = split;
        m2 = len - split;
    } else {
        s0 = 0;
        m = split;
        m2 = 0;
    }
    // Item numbers and row starts fit an int (lists hold at most 5851
    // entries: 160 KB of LDS).  Up to 2048 entries (64 KB) they and
    // (2m - 1)^2 stay below 2^24, exact in FP32, so the row estimates below
    // are off by less than 0.001; on longer lists the correction loops may
    // run more than once.
    const int nitems = ROLE == 1 ? m * m2 : m * (m - 1) / 2;
    acc = 0.0;  // lanes 0, 1, 2 hold x, y, z
    int head = 0, qn = 0;

    auto flush = [&](int cnt) {
        __syncthreads();
        if (lane < cnt) {
            const int q = (head + lane) & (RING - 1);
            const int ia = sh.qa[q] & 0xffff, ib = sh.qa[q] >> 16;
            const double xa = L.p[3 * ia], ya = L.p[3 * ia + 1], za = L.p[3 * ia + 2];
            const double xb = L.p[3 * ib], yb = L.p[3 * ib + 1], zb = L.p[3 * ib + 2];
            double cx, cy, cz;
            if constexpr (ROLE == 0)
                ref_contrib<0>(xp, yp, zp, xa, ya, za, xb, yb, zb, nu, nu_ok, cx, cy, cz);
            else if constexpr (ROLE == 1)
                ref_contrib<1>(xa, ya, za, xp, yp, zp, xb, yb, zb, nu, nu_ok, cx, cy, cz);
            else
                ref_contrib<2>(xa, ya, za, xb, yb, zb, xp, yp, zp, nu, nu_ok, cx, cy, cz);
            sh.cb[lane * 3 + 0] = cx;
            sh.cb[lane * 3 + 1] = cy;
            sh.cb[lane * 3 + 2] = cz;
        }
        __syncthreads();
        if (lane < 3) {
            for (int t = 0; t < cnt; ++t) acc = acc - sh.cb[t * 3 + lane];
        }
        __syncthreads();
        head = (head + cnt) & (RING - 1);
        qn -= cnt;
    };

    const float twom1 = (float)(2 * m - 1), disc0 = (float)((2 * m - 1) * (2 * m - 1));
    const float inv_m2 = m2 > 0 ? 1.0f / (float)m2 : 0.0f;
    for (int t0 = 0; t0 < nitems; t0 += WAVE) {
        const int t = t0 + lane;
        bool ok = t < nitems;
        int ia = 0, ib = 0;
        if (ok) {
            // the row from an FP32 estimate (within one of the true row),
            // then corrected with exact integer comparisons
            // (the loops below correct the estimate exactly, whatever its
            // error; they are kept as loops, not unrolled: they rarely run)
            if constexpr (ROLE == 1) {
                int ra = (int)((float)t * inv_m2);
                int rem = t - ra * m2;
#pragma clang loop unroll(disable)
                while (rem < 0) {
                    --ra;
                    rem += m2;
                }
#pragma clang loop unroll(disable)
                while (rem >= m2) {
                    ++ra;
                    rem -= m2;
                }
                ia = ra;
                ib = split + rem;
            } else {
                // largest row a with a (2m - a - 1) / 2 <= t
                const float disc = disc0 - 8.0f * (float)t;
                int ra = (int)((twom1 - __builtin_amdgcn_sqrtf(disc)) * 0.5f);
                ra = ra < 0 ? 0 : (ra > m - 2 ? m - 2 : ra);
                auto start = [&](int r) { return r * (2 * m - r - 1) / 2; };
                int st = start(ra);
#pragma clang loop unroll(disable)
                while (ra > 0 && st > t) st = start(--ra);
#pragma clang loop unroll(disable)
                while (ra < m - 2 && start(ra + 1) <= t) st = start(++ra);
                ia = s0 + ra;
                ib = ia + 1 + (t - st);
            }
            const double ax = L.p[3 * ia], ay = L.p[3 * ia + 1], az = L.p[3 * ia + 2];
            const double bx = L.p[3 * ib], by = L.p[3 * ib + 1], bz = L.p[3 * ib + 2];
            double d2;
            if (plain) {
                // the periodic form's bits whenever every |difference| is
                // below L/2 (min(t, L - t) is then t, and t * t = d * d)
                const double dx = bx - ax, dy = by - ay, dz = bz - az;
                d2 = dx * dx + dy * dy + dz * dz;
            } else {
                d2 = sqr_dist_periodic_min(ax, ay, az, bx, by, bz, Lx, Ly, Lz);
            }
            ok = !(d2 > c2);
        }
        const unsigned long long mk = __ballot(ok);
        if (ok) {
            const int q = (head + qn + lane_prefix(mk)) & (RING - 1);
            sh.qa[q] = ia | (ib << 16);  // list indices are below 2^16
        }
        const int got = __popcll(mk);
        qn += got;
        accepted += (unsigned long long)got;
        if (qn >= WAVE) flush(WAVE);
    }
    if (qn > 0) flush(qn);
}

// Cutoff algorithm: one wave per owner collects N(p) from the cells with its
// positions, sorts it by index in LDS, then runs the three role streams in
// turn over that one list.
//
// The candidates' ids and positions are read by cell-order slot (cell_idx
// and the cell-ordered copies xs, ys, zs, which hold the same values as x, y,
// z), four per lane before any is used, so the gather costs one memory round
// trip per 256 candidates.  Lists of up to 256 entries are sorted by rank
// (each entry's rank is the number of smaller ids), carrying their positions
// along; longer ones take a bitonic sort of the ids and read the positions
// again.
constexpr int EXACT_GATHER = 4;  // candidates per lane and round

// 5 waves per SIMD (at most 96 VGPRs): 4% faster than the 101-VGPR build.
__global__ __launch_bounds__(WAVE, 5) void k_p3bca_exact(CutDev a)
{
    extern __shared__ __attribute__((aligned(16))) int sid_ex[];
    __shared__ ExactShared sh;
    __shared__ int sh_start[32], sh_pre[32];
    // a halted step-graph replay leaves the step to the host
    if (a.st->halt) return;
    const int slot = owner_slot((int)blockIdx.x);
    if (slot >= a.n) return;
    const int p = a.cell_idx[slot];
    const int lane = lane_id();
    const int cap = a.cap;
    const double xp = a.xs[slot], yp = a.ys[slot], zp = a.zs[slot];
    const SysGrid sg = grid_of_particle(a.g, p);
    double *sp = reinterpret_cast<double *>(sid_ex + cap);  // 3 cap doubles: x, y, z of entry t at 3t

    const Candidates cand = gather_candidates(a, sg, xp, yp, zp, sh_start, sh_pre);
    int len = 0;
    for (int t0 = 0; t0 < cand.total; t0 += EXACT_GATHER * WAVE) {
        int q[EXACT_GATHER];
        double qx[EXACT_GATHER], qy[EXACT_GATHER], qz[EXACT_GATHER];
#pragma unroll
        for (int h = 0; h < EXACT_GATHER; ++h) {
            const int t = t0 + h * WAVE + lane;
            q[h] = -1;
            qx[h] = qy[h] = qz[h] = 0.0;
            if (t < cand.total) {
                if (cand.brute) {
                    q[h] = cand.base + t;
                    qx[h] = a.x[q[h]];
                    qy[h] = a.y[q[h]];
                    qz[h] = a.z[q[h]];
                } else {
                    int lo = 0, hi = cand.ncells - 1;  // last cell k with sh_pre[k] <= t
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (sh_pre[mid] <= t) lo = mid;
                        else hi = mid - 1;
                    }
                    const int s = sh_start[lo] + (t - sh_pre[lo]);
                    q[h] = a.cell_idx[s];
                    qx[h] = a.xs[s];
                    qy[h] = a.ys[s];
                    qz[h] = a.zs[s];
                }
            }
        }
#pragma unroll
        for (int h = 0; h < EXACT_GATHER; ++h) {
            const bool ok = q[h] >= 0 && q[h] != p && within_cutoff(a, sg, xp, yp, zp, qx[h], qy[h], qz[h]);
            const unsigned long long m = __ballot(ok);
            const int pos = len + lane_prefix(m);
            if (ok && pos < cap) {
                sid_ex[pos] = q[h];
                sp[3 * pos] = qx[h];
                sp[3 * pos + 1] = qy[h];
                sp[3 * pos + 2] = qz[h];
            }
            len += __popcll(m);
        }
    }
    if (lane == 0) note_list_length(a.st, len, cap);
    if (len > cap) {
        if (lane == 0) atomicOr(&a.st->flags[2], 1);
        return;
    }
    __syncthreads();

    // When every neighbour lies within 0.24 L of p in every dimension, any
    // two of them differ by less than L / 2, and the pair tests may take the
    // plain form.
    bool far = false;
    auto note_far = [&](double qx, double qy, double qz) {
        far = far || !(__builtin_fabs(qx - xp) <= 0.24 * sg.L[0] && __builtin_fabs(qy - yp) <= 0.24 * sg.L[1] &&
                       __builtin_fabs(qz - zp) <= 0.24 * sg.L[2]);
    };
    int below = 0;
    if (len <= EXACT_GATHER * WAVE) {
        // rank sort: ids are distinct, so the ranks are a permutation.  The
        // ids are read four at a time; the last group is padded with INT_MAX
        // (the capacity is a power of two of at least 64, so it has room).
        if (lane < ((4 - (len & 3)) & 3)) sid_ex[len + lane] = INT_MAX;
        __syncthreads();
        int id[EXACT_GATHER], rk[EXACT_GATHER];
        double ex[EXACT_GATHER], ey[EXACT_GATHER], ez[EXACT_GATHER];
#pragma unroll
        for (int h = 0; h < EXACT_GATHER; ++h) {
            const int t = h * WAVE + lane;
            id[h] = INT_MAX;
            rk[h] = 0;
            ex[h] = ey[h] = ez[h] = 0.0;
            if (t < len) {
                id[h] = sid_ex[t];
                ex[h] = sp[3 * t];
                ey[h] = sp[3 * t + 1];
                ez[h] = sp[3 * t + 2];
                note_far(ex[h], ey[h], ez[h]);
                below += id[h] < p ? 1 : 0;
            }
        }
        const int4 *ids4 = reinterpret_cast<const int4 *>(sid_ex);
        for (int j = 0; j < (len + 3) >> 2; ++j) {
            const int4 v = ids4[j];
#pragma unroll
            for (int h = 0; h < EXACT_GATHER; ++h)
                rk[h] += (v.x < id[h] ? 1 : 0) + (v.y < id[h] ? 1 : 0) + (v.z < id[h] ? 1 : 0) + (v.w < id[h] ? 1 : 0);
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < EXACT_GATHER; ++h) {
            if (h * WAVE + lane < len) {
                sid_ex[rk[h]] = id[h];
                sp[3 * rk[h]] = ex[h];
                sp[3 * rk[h] + 1] = ey[h];
                sp[3 * rk[h] + 2] = ez[h];
            }
        }
    } else {
        // bitonic sort of sid_ex[0, P), padded with INT_MAX, then the
        // positions read again in sorted order
        int P = WAVE;
        while (P < len) P <<= 1;
        for (int t = len + lane; t < P; t += WAVE) sid_ex[t] = INT_MAX;
        __syncthreads();
        for (int k = 2; k <= P; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < P; t += WAVE) {
                    const int u = t ^ j;
                    if (u > t) {
                        const int va = sid_ex[t], vb = sid_ex[u];
                        if ((va > vb) == ((t & k) == 0)) {
                            sid_ex[t] = vb;
                            sid_ex[u] = va;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int t = lane; t < len; t += WAVE) {
            const int q = sid_ex[t];
            below += q < p ? 1 : 0;
            const double qx = a.x[q], qy = a.y[q], qz = a.z[q];
            sp[3 * t] = qx;
            sp[3 * t + 1] = qy;
            sp[3 * t + 2] = qz;
            note_far(qx, qy, qz);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) below += __shfl_xor(below, o);
    const int split = below;
    const bool plain = __ballot(far) == 0;
    __syncthreads();
    const StagedList list{sp};
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    unsigned long long accepted = 0, unused = 0;
    exact_role_staged<0>(list, xp, yp, zp, len, split, plain, a.c2, sg.L[0], sg.L[1], sg.L[2], a.nu, sh, acc0,
                         accepted);
    exact_role_staged<1>(list, xp, yp, zp, len, split, plain, a.c2, sg.L[0], sg.L[1], sg.L[2], a.nu, sh, acc1,
                         unused);
    exact_role_staged<2>(list, xp, yp, zp, len, split, plain, a.c2, sg.L[0], sg.L[1], sg.L[2], a.nu, sh, acc2,
                         unused);
    if (lane < 3) {
        a.facc[(long long)lane * a.n + p] = acc0;
        a.facc[(long long)(3 + lane) * a.n + p] = acc1;
        a.facc[(long long)(6 + lane) * a.n + p] = acc2;
    }
    if (lane == 0) {
        add_count(a.st, 0, accepted);
        add_rep(a, sg.r, accepted, 0.0);
    }
}

// ---------------------------------------------------------------------------
// Fast all-triplets sweep over tile triples (NATA / AUTA, or P3BCA without a
// cutoff).
//
// A wave takes a triple of 64-particle tiles t0 <= t1 <= t2 and every
// triplet (i, j, k) in it with i < j < k.  Lane l owns i = 64 t0 + l; j runs
// over tile t1 in lock step (broadcast from LDS); k runs over tile t2 skewed
// by lane, k = 64 t2 + ((l + s) mod 64), so at every step the 64 lanes touch
// 64 distinct k and can update k's force accumulator in LDS without
// conflicts.  j's force is reduced across lanes once per j.  The three tiles'
// forces are added to global memory with one FP64 atomic per particle and
// component at the end.

struct TileDev {
    int n;
    const double *x, *y, *z;
    const unsigned long long *triples;
    double nu;
    double *fx, *fy, *fz;
    Status *st;
    int count_slot;
    int want_energy;
};

__device__ __forceinline__ void unpack_triple(unsigned long long t, int &t0, int &t1, int &t2)
{
    t0 = (int)(t & 0x1fffffull);
    t1 = (int)((t >> 21) & 0x1fffffull);
    t2 = (int)((t >> 42) & 0x1fffffull);
}

template <class T, bool GRAD>
__global__ __launch_bounds__(WAVE) void k_forces_tiles(TileDev a)
{
    __shared__ double jx[WAVE], jy[WAVE], jz[WAVE];
    __shared__ double kx[WAVE], ky[WAVE], kz[WAVE];
    __shared__ double fkx[WAVE], fky[WAVE], fkz[WAVE];
    __shared__ double fjx[WAVE], fjy[WAVE], fjz[WAVE];
    int t0, t1, t2;
    unpack_triple(a.triples[blockIdx.x], t0, t1, t2);
    const int lane = lane_id();
    const int n = a.n;
    const int i = t0 * WAVE + lane;
    {
        const int j = t1 * WAVE + lane, k = t2 * WAVE + lane;
        // Padding slots get distinct far-away coordinates so the (masked)
        // arithmetic on them stays finite.
        const double far = 1.0e6 + 16.0 * lane;
        jx[lane] = j < n ? a.x[j] : far;
        jy[lane] = j < n ? a.y[j] : far;
        jz[lane] = j < n ? a.z[j] : far;
        kx[lane] = k < n ? a.x[k] : -far;
        ky[lane] = k < n ? a.y[k] : -far;
        kz[lane] = k < n ? a.z[k] : -far;
        fkx[lane] = fky[lane] = fkz[lane] = 0.0;
        fjx[lane] = fjy[lane] = fjz[lane] = 0.0;
    }
    const double xi = i < n ? a.x[i] : 3.0e6, yi = i < n ? a.y[i] : 3.0e6, zi = i < n ? a.z[i] : 3.0e6;
    __syncthreads();

    const T nu = (T)a.nu, K = (T)(0.375 * a.nu);
    double gix = 0.0, giy = 0.0, giz = 0.0, esum = 0.0;
    unsigned cnt = 0;
    const int jbase = t1 * WAVE, kbase = t2 * WAVE;

    for (int jj = 0; jj < WAVE; ++jj) {
        const int j = jbase + jj;
        if (j >= n) break;
        const double xj = jx[jj], yj = jy[jj], zj = jz[jj];
        const double dijx = xj - xi, dijy = yj - yi, dijz = zj - zi;
        const double u64 = __builtin_fma(dijx, dijx, __builtin_fma(dijy, dijy, dijz * dijz));
        const bool ij_ok = i < j;
        double gjx = 0.0, gjy = 0.0, gjz = 0.0;
        for (int s = 0; s < WAVE; ++s) {
            const int kk = (lane + s) & (WAVE - 1);
            const int k = kbase + kk;
            const bool ok = ij_ok && j < k && k < n;
            const double xk = kx[kk], yk = ky[kk], zk = kz[kk];
            const double dikx = xk - xi, diky = yk - yi, dikz = zk - zi;
            const double djkx = xk - xj, djky = yk - yj, djkz = zk - zj;
            const double v = __builtin_fma(dikx, dikx, __builtin_fma(diky, diky, dikz * dikz));
            const double w = __builtin_fma(djkx, djkx, __builtin_fma(djky, djky, djkz * djkz));
            double cu, cv, cw;
            T et;
            at_coeffs_from64<T, true>(u64, v, w, nu, K, cu, cv, cw, et);
            cu = ok ? cu : 0.0;
            cv = ok ? cv : 0.0;
            cw = ok ? cw : 0.0;
            if (GRAD) {
                gix = __builtin_fma(-cu, dijx, __builtin_fma(-cv, dikx, gix));
                giy = __builtin_fma(-cu, dijy, __builtin_fma(-cv, diky, giy));
                giz = __builtin_fma(-cu, dijz, __builtin_fma(-cv, dikz, giz));
                gjx = __builtin_fma(cu, dijx, __builtin_fma(-cw, djkx, gjx));
                gjy = __builtin_fma(cu, dijy, __builtin_fma(-cw, djky, gjy));
                gjz = __builtin_fma(cu, dijz, __builtin_fma(-cw, djkz, gjz));
                fkx[kk] = __builtin_fma(cv, dikx, __builtin_fma(cw, djkx, fkx[kk]));
                fky[kk] = __builtin_fma(cv, diky, __builtin_fma(cw, djky, fky[kk]));
                fkz[kk] = __builtin_fma(cv, dikz, __builtin_fma(cw, djkz, fkz[kk]));
            }
            if (ok) {
                esum += (double)et;
                ++cnt;
            }
        }
        if (GRAD) {
            gjx = wave_sum(gjx);
            gjy = wave_sum(gjy);
            gjz = wave_sum(gjz);
            if (lane == 0) {
                fjx[jj] += gjx;
                fjy[jj] += gjy;
                fjz[jj] += gjz;
            }
        }
    }
    __syncthreads();
    if (GRAD) {
        if (i < n) {
            unsafeAtomicAdd(&a.fx[i], gix);
            unsafeAtomicAdd(&a.fy[i], giy);
            unsafeAtomicAdd(&a.fz[i], giz);
        }
        const int j = jbase + lane, k = kbase + lane;
        if (j < n) {
            unsafeAtomicAdd(&a.fx[j], fjx[lane]);
            unsafeAtomicAdd(&a.fy[j], fjy[lane]);
            unsafeAtomicAdd(&a.fz[j], fjz[lane]);
        }
        if (k < n) {
            unsafeAtomicAdd(&a.fx[k], fkx[lane]);
            unsafeAtomicAdd(&a.fy[k], fky[lane]);
            unsafeAtomicAdd(&a.fz[k], fkz[lane]);
        }
    }
    if (a.want_energy) esum = wave_sum(esum);
    const unsigned long long c = wave_sum_u64(cnt);
    if (lane == 0) {
        if (a.want_energy) add_energy(a.st, esum);
        add_count(a.st, a.count_slot, c);
    }
}

__global__ void k_selftest_fp64(long long n, const double *__restrict__ a, const double *__restrict__ b,
                                double *__restrict__ sq, double *__restrict__ q, double *__restrict__ rs)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sq[i] = __builtin_sqrt(a[i]);
    q[i] = a[i] / b[i];
    rs[i] = __builtin_amdgcn_rsq(a[i]);
}

// The in-range sequences of the bit-exact path, element-wise.
__global__ void k_selftest_fp64_inrange(long long n, const double *__restrict__ a, const double *__restrict__ b,
                                        double *__restrict__ sq, double *__restrict__ q)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sq[i] = sqrt_rn_inrange(a[i]);
    q[i] = div_rn_inrange(a[i], b[i]);
}

GridDev grid_dev(const CellGrid &g)
{
    GridDev gd;
    for (int c = 0; c < 3; ++c) {
        gd.nc[c] = g.nc[c];
        gd.lo[c] = g.lo[c];
        gd.L[c] = g.L[c];
    }
    gd.nrep = g.nrep;
    gd.p_off = g.p_off;
    gd.c_off = g.c_off;
    gd.rnc = g.rnc;
    gd.rlo = g.rlo;
    gd.rL = g.rL;
    return gd;
}

CutDev cut_dev(const DeviceState &d, const CellGrid &g, const CutoffSpec &c, double nu, int cap, Status *st,
               int count_slot)
{
    CutDev a;
    a.n = d.n;
    a.x = d.x;
    a.y = d.y;
    a.z = d.z;
    a.g = grid_dev(g);
    a.cell_start = g.cell_start;
    a.cell_idx = g.cell_idx;
    a.xs = g.xs;
    a.ys = g.ys;
    a.zs = g.zs;
    a.c2 = c.c2;
    a.nu = nu;
    a.cap = cap;
    a.fx = d.fx;
    a.fy = d.fy;
    a.fz = d.fz;
    a.facc = d.facc;
    a.st = st;
    a.count_slot = count_slot;
    a.want_energy = 0;
    a.rep_count = g.rep_count;
    a.rep_energy = g.rep_energy;
    return a;
}

// The slotted layout's device view (cap 0 when off).
SlotsDev slots_dev(const CellGrid *g, int on)
{
    SlotsDev v{};
    if (on && g && g->pad_cap > 0) {
        v.cap = g->pad_cap;
        v.cnt = g->pad_cnt;
        v.state = g->pad_state;
        v.rec = g->pad_rec;
    }
    return v;
}

unsigned owner_grid(int n) { return 8u * OWNER_STRETCH * (unsigned)((n + 8 * OWNER_STRETCH - 1) / (8 * OWNER_STRETCH)); }

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.

size_t p3bca_fast_lds(int cap)
{
    return (size_t)cap * (6 * sizeof(double) + sizeof(int)) + RING * sizeof(unsigned) + 64 * sizeof(int);
}

// sorted ids, then (8-byte aligned, cap is a multiple of 64) the positions
size_t p3bca_exact_lds(int cap) { return (size_t)cap * (sizeof(int) + 3 * sizeof(double)); }

void launch_soa_from_rows(hipStream_t s, int n, const double *rows, double *x, double *y, double *z)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_soa_from_rows, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, rows, x, y, z);
}

void launch_rows_from_soa(hipStream_t s, int n, const double *x, const double *y, const double *z, double *rows)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_rows_from_soa, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, y, z, rows);
}

void launch_reset_status(hipStream_t s, Status *st, int keep_nonfinite)
{
    hipLaunchKernelGGL(k_reset_status, dim3(1), dim3(64), 0, s, st, keep_nonfinite);
}

void launch_begin_steps(hipStream_t s, Status *st, unsigned long long *counts, CellGrid *slotted)
{
    if (slotted) {
        hipLaunchKernelGGL(k_begin_steps_slotted, dim3(1), dim3(256), 0, s, st, counts, slotted->ncell,
                           slots_dev(slotted, 1));
        return;
    }
    hipLaunchKernelGGL(k_begin_steps, dim3(1), dim3(64), 0, s, st, counts);
}


void launch_finalize(hipStream_t s, Status *st) { hipLaunchKernelGGL(k_finalize, dim3(1), dim3(NSLOT), 0, s, st); }

void launch_predict(hipStream_t s, const DeviceState &d, double dt, double h)
{
    if (d.n <= 0) return;
    hipLaunchKernelGGL(k_predict, dim3(blocks_for(d.n, 256)), dim3(256), 0, s, d, dt, h);
}

void launch_update_fin(hipStream_t s, const DeviceState &d, double hdt, const double g[3], Status *st, int sum_roles,
                       int record, const CellGrid *cells, int slotted)
{
    // the particles' workgroups, then (with a cell list) one that cuts the
    // next sweep's XCD runs, and one that folds the status
    const unsigned nb = (d.n > 0 ? blocks_for(d.n, NSLOT) : 1) + (cells ? 1 : 0) + 1;
    hipLaunchKernelGGL(k_update_fin, dim3(nb), dim3(NSLOT), 0, s, d, hdt, g[0], g[1], g[2], st, sum_roles, record,
                       cells ? cells->ncell : 0, cells ? cells->cell_start : nullptr,
                       cells ? cells->cell_work : nullptr, cells ? cells->order : nullptr,
                       slots_dev(cells, slotted && cells));
}

void launch_sum_roles(hipStream_t s, const DeviceState &d)
{
    if (d.n <= 0) return;
    hipLaunchKernelGGL(k_sum_roles, dim3(blocks_for(d.n, 256)), dim3(256), 0, s, d);
}

void launch_zero(hipStream_t s, double *p, long long count)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(k_zero, dim3(blocks_for(count, 256)), dim3(256), 0, s, p, count);
}

long long scan_tmp_elems(long long count)
{
    long long total = 0, m = count + 1;
    while (true) {
        const long long nb = (m + SCAN_TILE - 1) / SCAN_TILE;
        total += nb + (nb > 1 ? nb + 1 : 0);
        if (nb <= 1) break;
        m = nb;
    }
    return total + 64;
}

void launch_scan_i32(hipStream_t s, const int *in, int *out, int count, long long *tmp)
{
    scan_rec<int, int>(s, in, count, out, (long long)count + 1, tmp);
}

bool cells_fused_scan(const CellGrid &g) { return g.ncell <= FUSED_SCAN_CELLS; }

void launch_bin_ranked(hipStream_t s, const DeviceState &d, int predict, double dt, double h, int zero_f,
                       CellGrid &g, Status *st)
{
    if (d.n <= 0) return;
    hipLaunchKernelGGL(k_bin<true>, dim3(blocks_for(d.n, 256)), dim3(256), 0, s, d, predict, dt, h, zero_f,
                       grid_dev(g), g.cell_of, g.cell_cnt, g.cell_rank, st);
}

void launch_scatter_fused(hipStream_t s, const DeviceState &d, CellGrid &g, Status *st)
{
    if (d.n <= 0) return;
    const int nsc = blocks_for(d.n, 256);
    hipLaunchKernelGGL(k_cell_scatter_fused, dim3(nsc + 8), dim3(256), 0, s, d.n, d.x, d.y, d.z, g.cell_of,
                       g.cell_rank, g.cell_cnt, g.cell_start, g.cell_idx, g.xs, g.ys, g.zs, g.ncell, g.order,
                       g.cell_work, g.scatter_ticket, nsc, st);
}

void launch_update_bin(hipStream_t s, const DeviceState &d, double hdt, const double g3[3], Status *st,
                       int sum_roles, double dt, double h, int zero_f, CellGrid &g, int slotted)
{
    // the particles' workgroups, then (slotted) 8 that order the next
    // sweep's runs, one that cuts the runs and one that folds the status
    const unsigned nb = (d.n > 0 ? blocks_for(d.n, NSLOT) : 1) + (slotted ? 10 : 2);
    hipLaunchKernelGGL(k_update_bin, dim3(nb), dim3(NSLOT), 0, s, d, hdt, g3[0], g3[1], g3[2], st, sum_roles, dt,
                       h, zero_f, grid_dev(g), g.cell_of, g.cell_cnt, g.cell_rank, g.ncell, g.cell_start,
                       g.cell_work, g.order, slots_dev(&g, slotted));
}

void launch_bin_slotted(hipStream_t s, const DeviceState &d, double dt, double h, int zero_f, CellGrid &g,
                        Status *st)
{
    // the particles' workgroups, then 8 that order the first sweep's cells
    const unsigned nb = (d.n > 0 ? blocks_for(d.n, 256) : 0) + 8;
    hipLaunchKernelGGL(k_bin_slotted, dim3(nb), dim3(256), 0, s, d, dt, h, zero_f, grid_dev(g), slots_dev(&g, 1),
                       st, g.ncell, g.cell_work, g.order);
}

void launch_cells_build(hipStream_t s, const DeviceState &d, int predict, double dt, double h, int zero_f,
                        CellGrid &g, Status *st, long long *tmp)
{
    const int n = d.n;
    if (n > 0 && g.ncell <= FUSED_SCAN_CELLS) {
        hipLaunchKernelGGL(k_bin<true>, dim3(blocks_for(n, 256)), dim3(256), 0, s, d, predict, dt, h, zero_f,
                           grid_dev(g), g.cell_of, g.cell_cnt, g.cell_rank, st);
        const int nsc = blocks_for(n, 256);
        hipLaunchKernelGGL(k_cell_scatter_fused, dim3(nsc + 8), dim3(256), 0, s, n, d.x, d.y, d.z, g.cell_of,
                           g.cell_rank, g.cell_cnt, g.cell_start, g.cell_idx, g.xs, g.ys, g.zs, g.ncell, g.order,
                           g.cell_work, g.scatter_ticket, nsc, st);
        return;
    }
    if (n > 0)
        hipLaunchKernelGGL(k_bin<false>, dim3(blocks_for(n, 256)), dim3(256), 0, s, d, predict, dt, h, zero_f,
                           grid_dev(g), g.cell_of, g.cell_cnt, nullptr, st);
    launch_scan_i32(s, g.cell_cnt, g.cell_start, g.ncell, tmp);
    if (n > 0)
        hipLaunchKernelGGL(k_cell_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, d.x, d.y, d.z, g.cell_of,
                           g.cell_start, g.cell_cnt, g.cell_idx, g.xs, g.ys, g.zs, g.ncell, g.order, g.cell_work);
}

template <class K>
static void allow_lds(K kernel, size_t bytes)
{
    if (bytes > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes);
}

void launch_p3bca_fast(hipStream_t s, const DeviceState &d, const CellGrid &g, const CutoffSpec &c, double nu,
                       int cap, int mixed, int grad, Status *st, int count_slot, int want_energy)
{
    if (d.n <= 0) return;
    CutDev a = cut_dev(d, g, c, nu, cap, st, count_slot);
    a.want_energy = want_energy;
    const size_t shm = p3bca_fast_lds(cap);
    const dim3 grid(owner_grid(d.n)), blk(WAVE);
    if (mixed) {
        if (grad) {
            allow_lds(k_p3bca_fast<float, true>, shm);
            hipLaunchKernelGGL((k_p3bca_fast<float, true>), grid, blk, shm, s, a);
        } else {
            allow_lds(k_p3bca_fast<float, false>, shm);
            hipLaunchKernelGGL((k_p3bca_fast<float, false>), grid, blk, shm, s, a);
        }
    } else if (grad == 2) {
        allow_lds(k_p3bca_fast<double, true, true>, shm);
        hipLaunchKernelGGL((k_p3bca_fast<double, true, true>), grid, blk, shm, s, a);
    } else {
        if (grad) {
            allow_lds(k_p3bca_fast<double, true>, shm);
            hipLaunchKernelGGL((k_p3bca_fast<double, true>), grid, blk, shm, s, a);
        } else {
            allow_lds(k_p3bca_fast<double, false>, shm);
            hipLaunchKernelGGL((k_p3bca_fast<double, false>), grid, blk, shm, s, a);
        }
    }
}

void launch_p3bca_cells(hipStream_t s, const DeviceState &d, const CellGrid &g, const CutoffSpec &c, double nu,
                        int mixed, int grad, Status *st, int count_slot, int want_energy, int *owners, int dense_on,
                        int slotted)
{
    if (d.n <= 0) return;
    CellsDev a;
    a.dense_on = dense_on;
    a.c = cut_dev(d, g, c, nu, 0, st, count_slot);
    a.c.want_energy = want_energy;
    a.xs = g.xs;
    a.ys = g.ys;
    a.zs = g.zs;
    a.ncell = g.ncell;
    a.owners = owners;
    a.order = g.order;
    a.cell_work = g.cell_work;
    a.sl = slots_dev(&g, slotted);
    // the 8 XCDs' runs at their capacity (the workgroups past a run's length exit at once)
    const dim3 gr((unsigned)(8 * cell_run_cap(g.ncell))), blk(CELL_WAVES * WAVE);
#define TB3_CELLS(T, G, E) hipLaunchKernelGGL((k_p3bca_cells<T, G, E>), gr, blk, 0, s, a)
    if (mixed) {
        if (grad) {
            if (want_energy) TB3_CELLS(float, true, true);
            else TB3_CELLS(float, true, false);
        } else {
            TB3_CELLS(float, false, true);
        }
    } else {
        if (grad) {
            if (want_energy) TB3_CELLS(double, true, true);
            else TB3_CELLS(double, true, false);
        } else {
            TB3_CELLS(double, false, true);
        }
    }
#undef TB3_CELLS
}

int p3bca_cells_blocks_per_cu()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb,
                                                     reinterpret_cast<const void *>(k_p3bca_cells<double, true, false>),
                                                     CELL_WAVES * WAVE, 0) != hipSuccess || nb < 1)
        nb = 1;
    return nb;
}

template <int DW>
void dense_setup_waves(size_t shm)
{
    allow_lds(k_p3bca_dense<double, true, DW>, shm);
    allow_lds(k_p3bca_dense<double, false, DW>, shm);
    allow_lds(k_p3bca_dense<float, true, DW>, shm);
    allow_lds(k_p3bca_dense<float, false, DW>, shm);
}

void p3bca_dense_setup()
{
    const size_t shm = dense_lds();
    dense_setup_waves<4>(shm);
    dense_setup_waves<8>(shm);
    dense_setup_waves<16>(shm);
}

int p3bca_dense_cap() { return DCAP; }

bool p3bca_dense_waves_ok(int waves) { return waves == 4 || waves == 8 || waves == 16; }

int p3bca_dense_blocks_per_cu(int waves)
{
    const void *k = waves == 4    ? reinterpret_cast<const void *>(k_p3bca_dense<double, true, 4>)
                    : waves == 16 ? reinterpret_cast<const void *>(k_p3bca_dense<double, true, 16>)
                                  : reinterpret_cast<const void *>(k_p3bca_dense<double, true, 8>);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, waves * WAVE, dense_lds()) != hipSuccess || nb < 1)
        nb = 1;
    return nb;
}

template <int DW>
void launch_dense_waves(hipStream_t s, const DenseDev &a, int mixed, int grad, unsigned grid)
{
    const dim3 gr(grid), blk(DW * WAVE);
    const size_t shm = dense_lds();
    if (mixed) {
        if (grad) hipLaunchKernelGGL((k_p3bca_dense<float, true, DW>), gr, blk, shm, s, a);
        else hipLaunchKernelGGL((k_p3bca_dense<float, false, DW>), gr, blk, shm, s, a);
    } else {
        if (grad) hipLaunchKernelGGL((k_p3bca_dense<double, true, DW>), gr, blk, shm, s, a);
        else hipLaunchKernelGGL((k_p3bca_dense<double, false, DW>), gr, blk, shm, s, a);
    }
}

void launch_p3bca_dense(hipStream_t s, const DeviceState &d, const CellGrid &g, const CutoffSpec &c, double nu,
                        int mixed, int grad, Status *st, int count_slot, int want_energy, int *owners, int grid,
                        int waves)
{
    if (d.n <= 0) return;
    DenseDev a;
    a.c = cut_dev(d, g, c, nu, 0, st, count_slot);
    a.c.want_energy = want_energy;
    a.xs = g.xs;
    a.ys = g.ys;
    a.zs = g.zs;
    a.cell_of = g.cell_of;
    a.owners = owners;
    a.queue = &st->dqueue[0];
    const unsigned gr = (unsigned)(grid < 1 ? 1 : grid);
    if (waves == 4) launch_dense_waves<4>(s, a, mixed, grad, gr);
    else if (waves == 16) launch_dense_waves<16>(s, a, mixed, grad, gr);
    else launch_dense_waves<8>(s, a, mixed, grad, gr);
}

void launch_p3bca_exact(hipStream_t s, const DeviceState &d, const CellGrid &g, const CutoffSpec &c, double nu,
                        int cap, Status *st)
{
    if (d.n <= 0) return;
    const CutDev a = cut_dev(d, g, c, nu, cap, st, 0);
    const size_t shm = p3bca_exact_lds(cap);
    allow_lds(k_p3bca_exact, shm);
    hipLaunchKernelGGL(k_p3bca_exact, dim3(owner_grid(d.n)), dim3(WAVE), shm, s, a);
}

void launch_auta_exact(hipStream_t s, const DeviceState &d, double nu, Status *st)
{
    if (d.n <= 0) return;
    AutaExactDev a;
    a.n = d.n;
    a.x = d.x;
    a.y = d.y;
    a.z = d.z;
    a.nu = nu;
    a.facc = d.facc;
    a.st = st;
    hipLaunchKernelGGL(k_auta_exact, dim3((unsigned)(3 * d.n)), dim3(WAVE), 0, s, a);
}

void launch_forces_tiles(hipStream_t s, const DeviceState &d, const unsigned long long *triples,
                         long long ntriples, double nu, int mixed, int grad, Status *st, int count_slot,
                         int want_energy)
{
    if (d.n <= 0 || ntriples <= 0) return;
    TileDev a;
    a.n = d.n;
    a.x = d.x;
    a.y = d.y;
    a.z = d.z;
    a.triples = triples;
    a.nu = nu;
    a.fx = d.fx;
    a.fy = d.fy;
    a.fz = d.fz;
    a.st = st;
    a.count_slot = count_slot;
    a.want_energy = want_energy;
    const dim3 grid((unsigned)ntriples), blk(WAVE);
    if (mixed) {
        if (grad) hipLaunchKernelGGL((k_forces_tiles<float, true>), grid, blk, 0, s, a);
        else hipLaunchKernelGGL((k_forces_tiles<float, false>), grid, blk, 0, s, a);
    } else {
        if (grad) hipLaunchKernelGGL((k_forces_tiles<double, true>), grid, blk, 0, s, a);
        else hipLaunchKernelGGL((k_forces_tiles<double, false>), grid, blk, 0, s, a);
    }
}

void launch_selftest_fp64(hipStream_t s, long long n, const double *a, const double *b, double *sq, double *q,
                          double *rs)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_selftest_fp64, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, b, sq, q, rs);
}

void launch_selftest_fp64_inrange(hipStream_t s, long long n, const double *a, const double *b, double *sq,
                                  double *q)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_selftest_fp64_inrange, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, b, sq, q);
}

}  // namespace tb3
