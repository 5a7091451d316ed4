// tb3_host.cpp -- host side of the C ABI declared in include/tb3.h.
//
// Owns the device buffers of one particle system, keeps the cell grid of the
// cutoff algorithm, and dispatches the sweep kernels by algorithm and
// precision.  No exception crosses the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tb3.h"
#include "tb3_internal.h"

using namespace tb3;

namespace {

std::string g_create_error;

struct Fail : std::runtime_error {
    tb3_status code;
    Fail(tb3_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw Fail(e_ == hipErrorOutOfMemory ? TB3_ENOMEM : TB3_EHIP,                             \
                       std::string(#expr) + ": " + hipGetErrorString(e_));                            \
    } while (0)

template <class T>
void dev_alloc(T *&p, long long count)
{
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
    if (count <= 0) return;
    HIPCHK(hipMalloc((void **)&p, (size_t)count * sizeof(T)));
}

template <class T>
void dev_free(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Device arrays of one particle system (structure of arrays).
struct Buffers {
    long long cap = 0;
    double *block = nullptr;  // 22 * cap doubles
    DeviceState d;

    void ensure(long long n)
    {
        if (n <= cap && block) {
            d.n = (int)n;
            return;
        }
        const long long c = std::max<long long>(n, 64);
        dev_alloc(block, 22 * c);
        cap = c;
        double *b = block;
        d.x = b; b += c; d.y = b; b += c; d.z = b; b += c;
        d.vx = b; b += c; d.vy = b; b += c; d.vz = b; b += c;
        d.ax = b; b += c; d.ay = b; b += c; d.az = b; b += c;
        d.fx = b; b += c; d.fy = b; b += c; d.fz = b; b += c;
        d.m = b; b += c;
        d.facc = b;  // 9 * c
        d.n = (int)n;
    }
    void release()
    {
        dev_free(block);
        cap = 0;
        d = DeviceState();
    }
};

constexpr int MAX_LIST = 16384;  // longest neighbour list the LDS staging accepts

// Working memory of tb3_energy_grad_dev_noalloc: one static array per GPU,
// loaded with the code object and held until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device memory: its
// device address comes from hipGetSymbolAddress, once per device (at
// tb3_create, with that device current), and only that address is used.
constexpr size_t NOALLOC_BYTES = TB3_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
unsigned char *g_scratch_addr[MAX_DEVICES] = {};

// Records this device's address of `scratch` (the device must be current).
void init_scratch_address(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_addr[device]) return;
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Fail(TB3_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_addr[device] = static_cast<unsigned char *>(ptr);
}

unsigned char *scratch_address(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_addr[device];
}

}  // namespace

struct tb3_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    int algo = TB3_NATA;
    double nu = 1.0;
    double cutoff = 0.0;
    int precision = TB3_FP64_EXACT;
    bool user_L = false;
    double L[3] = {-1, -1, -1};
    double lo[3] = {0, 0, 0};
    bool have_L = false;

    long long n = 0;
    std::vector<int32_t> ids;
    Buffers state;    // the particle system
    Buffers scratch;  // staging and the device-pointer API's working set

    // replica batch (tb3_batch_set): particle offsets and per-replica domains
    int nrep = 1;
    std::vector<int64_t> rep_off;
    std::vector<double> rep_L, rep_lo;  // 3 nrep each
    int rep_cap = 0;                    // capacity of the device replica tables

    // cell grid of the cutoff algorithm
    CellGrid grid;
    int grid_cell_cap = 0, grid_n_cap = 0;
    long long *scan_tmp = nullptr;
    long long scan_tmp_cap = 0;
    int list_cap = 128;    // LDS neighbour-list capacity per wave (power of two)
    bool cells_ok = true;  // the cell-per-workgroup kernel staged every neighbourhood so far
    bool dense_on = false;  // the dense-owner kernel follows the cell kernel (once some owner needed it)
    // the slotted cell layout of replayed steps (CellGrid::pad_cap): slots
    // per cell asked for (-1: three times the mean cell count, 0: off), and
    // whether a cell outgrew its slots in this particle system's grid (the
    // layout then stays off until the grid changes)
    int slots_req = -1;
    bool slots_failed = false;
    long long slot_buf_cap = 0;  // slot entries allocated
    int slot_cnt_cap = 0;        // slotted count entries allocated (2 ncell)
    int dense_grid = 1;    // persistent grid of the dense-owner kernel (workgroups that fit the GPU)
    int dense_waves = 8;   // waves per owner in the dense-owner kernel (tuned: 4, 8 or 16)
    int ncu = 1;           // compute units of the device
    // tuned by tb3_autotune: cell width / cutoff, resident cell-kernel workgroups per CU
    double cell_ratio = 1.0;
    int cells_per_cu = 0;  // resident cell-kernel workgroups per CU at the tuned width
    int *d_owners = nullptr;  // owners the cell kernel hands to the dense-owner kernel (n)

    // tile triples for the all-triplets fast sweep
    unsigned long long *triples = nullptr;
    long long ntriples = 0;
    long long triples_n = -1;

    Status *d_status = nullptr;
    Status *h_status = nullptr;  // pinned
    Slots *d_slots = nullptr;

    // the replayed step graph of tb3_run, and what its capture baked in
    hipGraphExec_t step_exec = nullptr;
    // cutoff grids that scan inside the scatter kernel replay the steps as
    // one launch of the binning, then step_exec_mid (scatter, sweep, the
    // update fused with the next step's binning) for all steps but the
    // last, 2^k of them per launch of step_exec_batch[k] (k = 1 .. 4,
    // largest first), and step_exec (scatter, sweep, update) for the last
    hipGraphExec_t step_exec_mid = nullptr;
    hipGraphExec_t step_exec_batch[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bool graph_fused = false;
    unsigned long long epoch = 1;  // bumped whenever buffers, grid or parameters change
    unsigned long long graph_epoch = 0;
    double graph_dt = 0.0, graph_g[3] = {0, 0, 0};
    int graph_cap = -1;
    bool graph_cells_ok = false;
    bool graph_dense_on = false;
    int graph_slots = 0;  // the slots per cell the capture used (0: contiguous layout)
    unsigned long long *d_counts = nullptr;
    unsigned long long *h_counts = nullptr;  // pinned, counts_cap entries
    long long counts_cap = 0;

    // phase timing (tb3_set_profiling)
    bool prof = false;
    bool recording = false;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double prof_acc[6] = {0, 0, 0, 0, 0, 0};
};

namespace {

tb3_status fail(tb3_ctx *c, tb3_status code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

template <class F>
tb3_status guarded(tb3_ctx *c, F &&f)
{
    if (!c) return TB3_EINVAL;
    try {
        HIPCHK(hipSetDevice(c->device));
        f();
        c->err.clear();
        return TB3_OK;
    } catch (const Fail &e) {
        return fail(c, e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(c, TB3_ENOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(c, TB3_EHIP, e.what());
    }
}

// RegularGridDecomposition::getDomainMinMax / Init (RegularGridDecomposition.cpp:24-28, 135-166)
void domain_from_rows(long long n, const double *rows, double *L, double *lo)
{
    double mn[3], mx[3];
    for (int c = 0; c < 3; ++c) {
        mn[c] = 1.7976931348623157e308;
        mx[c] = -1.7976931348623157e308;
    }
    for (long long i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const double v = rows[3 * i + c];
            if (v < mn[c]) mn[c] = v;
            if (v > mx[c]) mx[c] = v;
        }
    for (int c = 0; c < 3; ++c) {
        const double a = mn[c] - 0.0001, b = mx[c] + 0.0001;
        L[c] = std::fabs(b - a);
        lo[c] = a;
    }
}

bool uses_cutoff(const tb3_ctx *c) { return c->algo == TB3_P3BCA && c->cutoff > 0; }

CutoffSpec cutoff_spec(const tb3_ctx *c)
{
    CutoffSpec s;
    s.use_cut = uses_cutoff(c) ? 1 : 0;
    s.c2 = c->cutoff * c->cutoff;  // Algorithm.cpp:67
    for (int k = 0; k < 3; ++k) s.L[k] = c->L[k];
    return s;
}

// Cell width h >= cutoff (1 + 1e-6), so every pair within the periodic
// cutoff lies in the same or adjacent cells (mod the grid) even after the
// rounding of the wrapped coordinates; dimensions with fewer than three cells
// collapse to one.
void grid_shape(const tb3_ctx *c, const double *L, int nc[3])
{
    long long total = 1;
    for (int k = 0; k < 3; ++k) {
        nc[k] = 1;
        if (uses_cutoff(c) && L[k] > 0 && std::isfinite(L[k])) {
            const double q = std::floor(L[k] / (c->cutoff * c->cell_ratio * (1.0 + 1e-6)));
            nc[k] = q >= 3.0 ? (int)std::min(q, 1024.0) : 1;
        }
        total *= nc[k];
    }
    while (total > (1LL << 22)) {  // keep the grid bounded: widen the largest dimension
        int kmax = 0;
        for (int k = 1; k < 3; ++k)
            if (nc[k] > nc[kmax]) kmax = k;
        nc[kmax] = std::max(3, nc[kmax] / 2);
        total = (long long)nc[0] * nc[1] * nc[2];
    }
}

// Publishes the replica tables (or their absence) to the status block.
void publish_replicas(tb3_ctx *c)
{
    Status fields;
    std::memset(&fields, 0, sizeof(fields));
    fields.nrep = c->nrep;
    fields.rep_count = c->nrep > 1 ? c->grid.rep_count : nullptr;
    fields.rep_energy = c->nrep > 1 ? c->grid.rep_energy : nullptr;
    const size_t off = offsetof(Status, nrep);
    HIPCHK(hipMemcpy(reinterpret_cast<char *>(c->d_status) + off, reinterpret_cast<char *>(&fields) + off,
                     sizeof(Status) - off, hipMemcpyHostToDevice));
}

void configure_grid(tb3_ctx *c)
{
    CellGrid &g = c->grid;
    const int old_nrep = g.nrep, old_ncell = g.ncell;
    int old_nc[3];
    double old_lo[3], old_L[3];
    for (int k = 0; k < 3; ++k) {
        old_nc[k] = g.nc[k];
        old_lo[k] = g.lo[k];
        old_L[k] = g.L[k];
    }
    if (c->nrep <= 1) {
        grid_shape(c, c->L, g.nc);
        g.ncell = g.nc[0] * g.nc[1] * g.nc[2];
        for (int k = 0; k < 3; ++k) {
            g.lo[k] = c->lo[k];
            g.L[k] = c->L[k];
        }
        g.nrep = 1;
    } else {
        const int R = c->nrep;
        std::vector<int> rnc((size_t)3 * R), coff((size_t)R + 1, 0), poff((size_t)R + 1, 0);
        for (int r = 0; r < R; ++r) {
            grid_shape(c, &c->rep_L[3 * (size_t)r], &rnc[3 * (size_t)r]);
            const long long cells = (long long)rnc[3 * r] * rnc[3 * r + 1] * rnc[3 * r + 2];
            if (coff[r] + cells > (1LL << 24)) throw Fail(TB3_EUNSUPPORTED, "the replicas need too many cells");
            coff[r + 1] = coff[r] + (int)cells;
            poff[r + 1] = (int)c->rep_off[r + 1];
        }
        if (R > c->rep_cap) {
            dev_alloc(g.p_off, R + 1);
            dev_alloc(g.c_off, R + 1);
            dev_alloc(g.rnc, 3LL * R);
            dev_alloc(g.rlo, 3LL * R);
            dev_alloc(g.rL, 3LL * R);
            dev_alloc(g.rep_count, R);
            dev_alloc(g.rep_energy, R);
            c->rep_cap = R;
        }
        HIPCHK(hipMemcpy(g.p_off, poff.data(), poff.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(g.c_off, coff.data(), coff.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(g.rnc, rnc.data(), rnc.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(g.rlo, c->rep_lo.data(), c->rep_lo.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(g.rL, c->rep_L.data(), c->rep_L.size() * sizeof(double), hipMemcpyHostToDevice));
        g.ncell = coff[R];
        for (int k = 0; k < 3; ++k) {
            g.nc[k] = rnc[k];
            g.lo[k] = c->rep_lo[k];
            g.L[k] = c->rep_L[k];
        }
        g.nrep = R;
    }
    publish_replicas(c);
    // A single system whose grid comes out identical (the same particle set
    // again, say) keeps its step graph and its per-cell work.
    bool same = c->nrep <= 1 && old_nrep == 1 && g.ncell == old_ncell;
    for (int k = 0; k < 3; ++k)
        same = same && g.nc[k] == old_nc[k] && g.lo[k] == old_lo[k] && g.L[k] == old_L[k];
    // the per-cell work of the last sweep belongs to the old cell numbering
    if (!same && g.cell_work && g.ncell <= c->grid_cell_cap)
        HIPCHK(hipMemset(g.cell_work, 0, (size_t)g.ncell * sizeof(unsigned)));
    c->cells_ok = true;  // a new grid gets a new chance at the cell kernel
    c->dense_on = false;
    c->slots_failed = false;
    if (!same) ++c->epoch;
}

// Per-replica domains: the bbox rule on each replica's particles, or the
// caller's domain size for every replica.
void replica_domains(tb3_ctx *c, const double *rows)
{
    const int R = c->nrep;
    c->rep_L.assign((size_t)3 * R, 1.0);
    c->rep_lo.assign((size_t)3 * R, 0.0);
    for (int r = 0; r < R; ++r) {
        const long long a = c->rep_off[r], b = c->rep_off[r + 1];
        if (c->user_L) {
            for (int k = 0; k < 3; ++k) c->rep_L[3 * r + k] = c->L[k];
        } else if (b > a) {
            domain_from_rows(b - a, rows + 3 * a, &c->rep_L[3 * (size_t)r], &c->rep_lo[3 * (size_t)r]);
        }
    }
}

void ensure_grid_buffers(tb3_ctx *c, int n)
{
    CellGrid &g = c->grid;
    if (g.ncell > c->grid_cell_cap) {
        ++c->epoch;
        dev_alloc(g.cell_cnt, g.ncell);
        HIPCHK(hipMemset(g.cell_cnt, 0, (size_t)g.ncell * sizeof(int)));
        dev_alloc(g.cell_start, (long long)g.ncell + 1);
        dev_alloc(g.order, cell_order_len(g.ncell));
        dev_alloc(g.cell_work, g.ncell);
        HIPCHK(hipMemset(g.cell_work, 0, (size_t)g.ncell * sizeof(unsigned)));
        dev_alloc(g.scatter_ticket, 1);
        HIPCHK(hipMemset(g.scatter_ticket, 0, sizeof(unsigned)));
        c->grid_cell_cap = g.ncell;
    }
    if (n > c->grid_n_cap) {
        ++c->epoch;
        dev_alloc(c->d_owners, n);
        dev_alloc(g.cell_of, n);
        dev_alloc(g.cell_rank, n);
        dev_alloc(g.cell_idx, n);
        dev_alloc(g.xs, n);
        dev_alloc(g.ys, n);
        dev_alloc(g.zs, n);
        c->grid_n_cap = n;
    }
    const long long need = scan_tmp_elems(std::max<long long>(n, g.ncell));
    if (need > c->scan_tmp_cap) {
        ++c->epoch;
        dev_alloc(c->scan_tmp, need);
        c->scan_tmp_cap = need;
    }
}

// Slots per cell of the slotted layout for n particles: three times the
// mean cell count (at least 32, a multiple of 16) unless set by
// tb3_set_cell_slots; 0 where the layout does not apply (off, failed, the
// exact path, a replica batch, or cells too full to stage).
int slots_for(const tb3_ctx *c, int n)
{
    const CellGrid &g = c->grid;
    if (c->slots_req == 0 || c->slots_failed || c->precision == TB3_FP64_EXACT || g.nrep > 1 || g.ncell <= 0 ||
        n < 3)
        return 0;
    if (c->slots_req > 0) return c->slots_req;
    const int mean = (int)((n + (long long)g.ncell - 1) / g.ncell);
    const int cap = std::max(32, (3 * mean + 15) / 16 * 16);
    return cap <= 256 ? cap : 0;
}

// Sets grid.pad_cap for the current particle count and allocates the slot
// buffers it needs.
void ensure_slot_buffers(tb3_ctx *c, int n)
{
    CellGrid &g = c->grid;
    g.pad_cap = slots_for(c, n);
    if (g.pad_cap == 0) return;
    const long long need = (long long)g.ncell * g.pad_cap;
    if (need > c->slot_buf_cap) {
        ++c->epoch;
        dev_alloc(g.pad_rec, need);
        c->slot_buf_cap = need;
    }
    if (2 * g.ncell > c->slot_cnt_cap) {
        ++c->epoch;
        dev_alloc(g.pad_cnt, 2LL * g.ncell);
        dev_alloc(g.pad_state, 2);
        c->slot_cnt_cap = 2 * g.ncell;
    }
}

void ensure_triples(tb3_ctx *c, long long n)
{
    if (c->triples_n == n) return;
    ++c->epoch;
    const long long nt = (n + 63) / 64;
    std::vector<unsigned long long> t;
    t.reserve((size_t)(nt * (nt + 1) * (nt + 2) / 6));
    for (long long a = 0; a < nt; ++a)
        for (long long b = a; b < nt; ++b)
            for (long long d = b; d < nt; ++d)
                t.push_back((unsigned long long)a | ((unsigned long long)b << 21) | ((unsigned long long)d << 42));
    dev_alloc(c->triples, (long long)t.size());
    if (!t.empty())
        HIPCHK(hipMemcpy(c->triples, t.data(), t.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    c->ntriples = (long long)t.size();
    c->triples_n = n;
}

void read_status(tb3_ctx *c, hipStream_t s)
{
    HIPCHK(hipMemcpyAsync(c->h_status, c->d_status, sizeof(Status), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
}

int next_pow2(int v)
{
    int p = 64;
    while (p < v) p <<= 1;
    return p;
}

// The force sweep proper on positions already in d (and, for the cutoff
// algorithm, on a cell list already built for them).
//   grad   : dE/dx (the reference's f) into d.fx/fy/fz
//   energy : E into the status block
// The sweep of a pass.  With `tail` it ends with the role sum (exact path)
// and the status fold; an integrating pass leaves both to launch_tail.
// `slotted`: the cell kernel sweeps the slotted layout (replayed steps).
void launch_sweep(tb3_ctx *c, hipStream_t s, DeviceState &d, bool grad, bool energy, bool tail,
                  bool slotted = false)
{
    const bool cut = uses_cutoff(c);
    const bool exact = c->precision == TB3_FP64_EXACT;
    const int mixed = (!exact && c->precision == TB3_FP32_MIXED) ? 1 : 0;
    const CutoffSpec cs = cutoff_spec(c);
    const bool fast_grad = grad && !exact;
    if (!cut && d.n >= 3 && (fast_grad || energy)) ensure_triples(c, d.n);
    if (c->recording) HIPCHK(hipEventRecord(c->ev[2], s));
    if (d.n >= 3) {
        if (exact && grad) {
            if (cut) launch_p3bca_exact(s, d, c->grid, cs, c->nu, c->list_cap, c->d_status);
            else launch_auta_exact(s, d, c->nu, c->d_status);
        }
        if (fast_grad || energy) {
            const int slot = (exact && grad) ? 1 : 0;  // the exact pass owns count[0]
            if (cut && c->cells_ok) {
                launch_p3bca_cells(s, d, c->grid, cs, c->nu, mixed, fast_grad ? 1 : 0, c->d_status, slot,
                                   energy ? 1 : 0, c->d_owners, c->dense_on ? 1 : 0, slotted ? 1 : 0);
                if (c->dense_on)
                    launch_p3bca_dense(s, d, c->grid, cs, c->nu, mixed, fast_grad ? 1 : 0, c->d_status, slot,
                                       energy ? 1 : 0, c->d_owners, c->dense_grid, c->dense_waves);
            }
            else if (cut)
                launch_p3bca_fast(s, d, c->grid, cs, c->nu, c->list_cap, mixed, fast_grad ? 1 : 0, c->d_status,
                                  slot, energy ? 1 : 0);
            else launch_forces_tiles(s, d, c->triples, c->ntriples, c->nu, mixed, fast_grad ? 1 : 0, c->d_status,
                                     slot, energy ? 1 : 0);
        }
    }
    if (c->recording) HIPCHK(hipEventRecord(c->ev[3], s));
    if (tail) {
        if (exact && grad && d.n >= 3) launch_sum_roles(s, d);
        launch_finalize(s, c->d_status);
    }
    HIPCHK(hipGetLastError());
}

// The end of an integrating pass after launch_sweep(..., tail = false): one
// kernel for the role sum, the status fold and the update (and, in a step
// graph, the count record and the status clear for the next replay).
void launch_tail(tb3_ctx *c, hipStream_t s, DeviceState &d, double dt, const double *g, bool grad, bool record,
                 bool slotted = false)
{
    const bool sum_roles = c->precision == TB3_FP64_EXACT && grad && d.n >= 3;
    const bool cells = uses_cutoff(c) && d.n >= 3 && c->grid.order;
    launch_update_fin(s, d, dt * 0.5, g, c->d_status, sum_roles ? 1 : 0, record ? 1 : 0, cells ? &c->grid : nullptr,
                      slotted ? 1 : 0);
    HIPCHK(hipGetLastError());
}

// Prepares a sweep over d: resets the status block (not inside a step
// graph, whose previous replay, or k_begin_steps, cleared it), runs the
// predictor when asked (fused into the binning pass for the cutoff
// algorithm), zeroes the force where the sweep accumulates atomically and
// builds the cell list.
void launch_prepare(tb3_ctx *c, hipStream_t s, DeviceState &d, bool predict, double dt, bool grad,
                    bool in_graph = false)
{
    const bool cut = uses_cutoff(c);
    const bool exact = c->precision == TB3_FP64_EXACT;
    const bool zero_f = grad && (!exact || d.n < 3);
    if (!in_graph) launch_reset_status(s, c->d_status, 0);
    if (c->recording) HIPCHK(hipEventRecord(c->ev[0], s));
    if (cut && d.n >= 3) {
        // the predictor is fused into the binning pass: its time is counted
        // with the cell-list phase
        ensure_grid_buffers(c, d.n);
        if (c->recording) HIPCHK(hipEventRecord(c->ev[1], s));
        launch_cells_build(s, d, predict ? 1 : 0, dt, dt * dt * 0.5, zero_f ? 1 : 0, c->grid, c->d_status,
                           c->scan_tmp);
    } else {
        if (predict) launch_predict(s, d, dt, dt * dt * 0.5);
        if (c->recording) HIPCHK(hipEventRecord(c->ev[1], s));
        if (zero_f) {
            launch_zero(s, d.fx, d.n);
            launch_zero(s, d.fy, d.n);
            launch_zero(s, d.fz, d.n);
        }
    }
    HIPCHK(hipGetLastError());
}

// Re-runs an overflowed cutoff sweep with a larger LDS list capacity.  The
// update kernel skipped itself, so the state is just after the predictor.
bool grow_after_overflow(tb3_ctx *c)
{
    if (!c->h_status->flags[2]) return false;
    // a neighbourhood too large for the cell kernel's staging: use the
    // per-owner kernel for this particle system from now on
    if (c->h_status->flags[3]) c->cells_ok = false;
    // owners that need the dense-owner kernel: launch it from now on
    if (c->h_status->want_dense) c->dense_on = true;
    const int need = c->h_status->flags[1];
    if (need > c->list_cap) {
        if (need > MAX_LIST)
            throw Fail(TB3_EUNSUPPORTED, "a particle has " + std::to_string(need) +
                                             " neighbours within the cutoff; at most " + std::to_string(MAX_LIST) +
                                             " are supported");
        c->list_cap = std::min(next_pow2(need + need / 8), MAX_LIST);
    }
    return true;
}

// Phase times of one pass: ev[0..2] bracket the status reset and the cell
// list (read once the first attempt has finished), ev[2..4] the sweep and
// the update of the attempt that produced the result.
void accumulate_profile(tb3_ctx *c, unsigned long long triplets, const float *prep)
{
    float ms[4] = {prep[0], prep[1], 0, 0};
    for (int k = 2; k < 4; ++k) HIPCHK(hipEventElapsedTime(&ms[k], c->ev[k], c->ev[k + 1]));
    c->prof_acc[0] += 1.0;
    for (int k = 0; k < 4; ++k) c->prof_acc[1 + k] += ms[k];
    c->prof_acc[5] += (double)triplets;
}

// Completes a pass whose sweep overflowed (the status block is on the host):
// grows the capacity or switches kernels, then sweeps and updates again.
void finish_overflowed(tb3_ctx *c, hipStream_t s, DeviceState &d, bool integrate, double dt, const double *g,
                       bool grad, bool energy)
{
    const bool exact = c->precision == TB3_FP64_EXACT;
    for (int retry = 0; grow_after_overflow(c); ++retry) {
        if (retry >= 8) throw Fail(TB3_EHIP, "the cutoff sweep kept overflowing its LDS staging");
        launch_reset_status(s, c->d_status, 1);  // keeps the non-finite flag of the binning pass
        if (grad && !exact) {
            launch_zero(s, d.fx, d.n);
            launch_zero(s, d.fy, d.n);
            launch_zero(s, d.fz, d.n);
        }
        launch_sweep(c, s, d, grad, energy, !integrate);
        if (integrate) launch_tail(c, s, d, dt, g, grad, false);
        if (c->recording) HIPCHK(hipEventRecord(c->ev[4], s));
        read_status(c, s);
    }
}

// One pass over d: optional predictor, sweep, optional update.  Handles the
// neighbour-list overflow protocol and returns the accepted-triplet count.
unsigned long long run_pass(tb3_ctx *c, hipStream_t s, DeviceState &d, bool integrate, double dt, const double *g,
                            bool grad, bool energy, bool record)
{
    c->recording = record && c->prof;
    struct Reset {
        tb3_ctx *c;
        ~Reset() { c->recording = false; }
    } reset{c};
    launch_prepare(c, s, d, integrate, dt, grad);
    launch_sweep(c, s, d, grad, energy, !integrate);
    if (integrate) launch_tail(c, s, d, dt, g, grad, false);
    if (c->recording) HIPCHK(hipEventRecord(c->ev[4], s));
    read_status(c, s);
    float prep[2] = {0.0f, 0.0f};
    if (c->recording)
        for (int k = 0; k < 2; ++k) HIPCHK(hipEventElapsedTime(&prep[k], c->ev[k], c->ev[k + 1]));
    finish_overflowed(c, s, d, integrate, dt, g, grad, energy);
    if (c->recording) accumulate_profile(c, c->h_status->count[0], prep);
    return c->h_status->count[0];
}

// The batched middle-step graphs hold 2, 4, 8 and 16 steps.
constexpr int GRAPH_BATCH_LOG2 = 4;

// Replays of one captured step graph run `steps` steps of tb3_run with one
// launch per step (per up to 16 steps where the batched graphs apply)
// and one host synchronisation per run.  A step whose sweep
// overflows halts the replays that follow (they leave state and flags
// alone); the host completes that step with the overflow protocol and
// captures again for the rest.
bool graph_ok(const tb3_ctx *c)
{
    if (c->prof || !uses_cutoff(c)) return !c->prof;
    // a dynamic LDS size above 64 KB needs a function attribute set at launch
    if (c->precision == TB3_FP64_EXACT) return p3bca_exact_lds(c->list_cap) <= 64 * 1024;
    return c->cells_ok || p3bca_fast_lds(c->list_cap) <= 64 * 1024;
}

// The fused replay sequence serves the cutoff algorithm on grids whose
// count scan runs inside the scatter kernel.
bool fused_steps(const tb3_ctx *c) { return uses_cutoff(c) && c->state.d.n >= 3 && cells_fused_scan(c->grid); }

// Fused replays of the cell kernel alone (fast sweep, no dense-owner
// kernel, one system) run on the slotted layout: the binning writes each
// cell's slots itself and no scatter kernel runs.  ensure_slot_buffers sets
// grid.pad_cap first.
bool slotted_steps(const tb3_ctx *c)
{
    return fused_steps(c) && c->precision != TB3_FP64_EXACT && c->cells_ok && !c->dense_on && c->grid.pad_cap > 0;
}

template <class F>
hipGraphExec_t capture(hipStream_t s, F &&body)
{
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        body();
    } catch (...) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    HIPCHK(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCHK(e);
    return exec;
}

void capture_step_graph(tb3_ctx *c, double dt, const double *g)
{
    const bool fused = fused_steps(c);
    const int slots = slotted_steps(c) ? c->grid.pad_cap : 0;
    if (c->step_exec && c->graph_epoch == c->epoch && c->graph_dt == dt && c->graph_g[0] == g[0] &&
        c->graph_g[1] == g[1] && c->graph_g[2] == g[2] && c->graph_cap == c->list_cap &&
        c->graph_cells_ok == c->cells_ok && c->graph_dense_on == c->dense_on && c->graph_fused == fused &&
        c->graph_slots == slots)
        return;
    if (c->step_exec) {
        (void)hipGraphExecDestroy(c->step_exec);
        c->step_exec = nullptr;
    }
    if (c->step_exec_mid) {
        (void)hipGraphExecDestroy(c->step_exec_mid);
        c->step_exec_mid = nullptr;
    }
    for (hipGraphExec_t &b : c->step_exec_batch) {
        if (b) (void)hipGraphExecDestroy(b);
        b = nullptr;
    }

    hipStream_t s = c->stream;
    DeviceState &d = c->state.d;
    if (fused) {
        const bool exact = c->precision == TB3_FP64_EXACT;
        const bool slotted = slots > 0;  // then the binning fills the slots and no scatter runs
        auto mid = [&] {
            if (!slotted) launch_scatter_fused(s, d, c->grid, c->d_status);
            launch_sweep(c, s, d, true, false, false, slotted);
            launch_update_bin(s, d, dt * 0.5, g, c->d_status, exact ? 1 : 0, dt, dt * dt * 0.5, exact ? 0 : 1,
                              c->grid, slotted ? 1 : 0);
        };
        c->step_exec_mid = capture(s, mid);
        // about 8 us pass between one graph launch's last kernel and the
        // next one's first, none between the kernels of one graph
        for (int k = 1; k <= GRAPH_BATCH_LOG2; ++k)
            c->step_exec_batch[k] = capture(s, [&] {
                for (int j = 0; j < (1 << k); ++j) mid();
            });
        c->step_exec = capture(s, [&] {
            if (!slotted) launch_scatter_fused(s, d, c->grid, c->d_status);
            launch_sweep(c, s, d, true, false, false, slotted);
            launch_tail(c, s, d, dt, g, true, true, slotted);
        });
    } else {
        c->step_exec = capture(s, [&] {
            launch_prepare(c, s, d, true, dt, true, true);
            launch_sweep(c, s, d, true, false, false);
            launch_tail(c, s, d, dt, g, true, true);
        });
    }
    c->graph_fused = fused;
    c->graph_slots = slots;
    c->graph_epoch = c->epoch;
    c->graph_dt = dt;
    for (int k = 0; k < 3; ++k) c->graph_g[k] = g[k];
    c->graph_cap = c->list_cap;
    c->graph_cells_ok = c->cells_ok;
    c->graph_dense_on = c->dense_on;
}

// The last pass's accepted-triplet count, or per-replica counts for a batch.
void step_counts(tb3_ctx *c, uint64_t *out)
{
    if (c->nrep > 1) {
        HIPCHK(hipMemcpy(out, c->grid.rep_count, (size_t)c->nrep * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
    } else {
        out[0] = c->h_status->count[0];
    }
}

// counts: `steps` totals, or steps x nrep per-replica counts for a batch.
void run_steps_graph(tb3_ctx *c, int steps, double dt, const double *g, uint64_t *counts)
{
    hipStream_t s = c->stream;
    DeviceState &d = c->state.d;
    const int per = c->nrep > 1 ? c->nrep : 1;
    if ((long long)steps * per > c->counts_cap) {
        // room for 4096 counts at least, and twice what is asked: a run
        // longer than the last one must not pay for a device and a pinned
        // host allocation on its own clock (they cost a 20-step run about
        // 4 us per step)
        const long long cap = std::max<long long>(2LL * steps * per, 4096);
        dev_alloc(c->d_counts, cap);
        if (c->h_counts) (void)hipHostFree(c->h_counts);
        c->h_counts = nullptr;
        HIPCHK(hipHostMalloc((void **)&c->h_counts, (size_t)cap * sizeof(unsigned long long), hipHostMallocDefault));
        c->counts_cap = cap;
    }
    int done = 0;
    while (done < steps) {
        if (!graph_ok(c)) {  // e.g. the overflow protocol grew the lists past 64 KB of LDS
            for (; done < steps; ++done) {
                run_pass(c, s, d, true, dt, g, true, false, false);
                if (counts) step_counts(c, counts + (size_t)done * per);
            }
            return;
        }
        // what a step would allocate is allocated before the capture
        if (uses_cutoff(c) && d.n >= 3) {
            ensure_grid_buffers(c, d.n);
            ensure_slot_buffers(c, d.n);
        }
        if (!uses_cutoff(c) && d.n >= 3 && c->precision != TB3_FP64_EXACT) ensure_triples(c, d.n);
        capture_step_graph(c, dt, g);
        launch_begin_steps(s, c->d_status, c->d_counts,
                           c->graph_fused && c->graph_slots > 0 ? &c->grid : nullptr);
        if (c->graph_fused) {
            // the first step's predictor and binning; each step_exec_mid
            // replay ends with the next step's
            if (c->graph_slots > 0) {
                launch_bin_slotted(s, d, dt, dt * dt * 0.5, 1, c->grid, c->d_status);
            } else {
                launch_bin_ranked(s, d, 1, dt, dt * dt * 0.5, c->precision == TB3_FP64_EXACT ? 0 : 1, c->grid,
                                  c->d_status);
            }
            int k = done;
            // the middle steps in batches of 16, 8, 4, 2, then one
            for (int b = GRAPH_BATCH_LOG2; b >= 1; --b)
                for (; k + (1 << b) < steps; k += 1 << b) HIPCHK(hipGraphLaunch(c->step_exec_batch[b], s));
            for (; k + 1 < steps; ++k) HIPCHK(hipGraphLaunch(c->step_exec_mid, s));
            HIPCHK(hipGraphLaunch(c->step_exec, s));
        } else {
            for (int k = done; k < steps; ++k) HIPCHK(hipGraphLaunch(c->step_exec, s));
        }
        // the status and every slot the replays could have recorded come back
        // behind one synchronisation
        HIPCHK(hipMemcpyAsync(c->h_status, c->d_status, sizeof(Status), hipMemcpyDeviceToHost, s));
        if (counts)
            HIPCHK(hipMemcpyAsync(c->h_counts, c->d_counts, (size_t)(steps - done) * per * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        const int ok = c->h_status->nstep;
        if (ok < 0 || ok > steps - done) throw Fail(TB3_EHIP, "replayed steps reported a bad step count");
        if (ok > 0 && counts)
            std::memcpy(counts + (size_t)done * per, c->h_counts, (size_t)ok * per * sizeof(unsigned long long));
        done += ok;
        if (done < steps) {
            // step `done` overflowed: its predictor ran, its update did not
            if (!c->h_status->flags[2]) throw Fail(TB3_EHIP, "replayed steps halted without an overflow");
            if (c->graph_fused && c->graph_slots > 0) {
                // that step was binned into the slots: the host's sweeps
                // below need its contiguous cell list (the counters of the
                // contiguous binning are at zero, untouched by the slotted
                // steps), built once the halt is cleared (the scatter
                // kernel skips halted replays; the status read above keeps
                // what the host needs); a cell that outgrew its slots turns
                // the layout off
                if (c->h_status->pad_over) c->slots_failed = true;
                launch_reset_status(s, c->d_status, 1);
                launch_cells_build(s, d, 0, dt, dt * dt * 0.5, 0, c->grid, c->d_status, c->scan_tmp);
            }
            finish_overflowed(c, s, d, true, dt, g, true, false);
            if (counts) step_counts(c, counts + (size_t)done * per);
            ++done;
        }
    }
}

void recompute_domain(tb3_ctx *c)
{
    if (c->user_L) {
        if (c->nrep > 1) replica_domains(c, nullptr);
        c->have_L = true;
        return;
    }
    if (c->n > 0) {
        std::vector<double> rows((size_t)(3 * c->n));
        c->scratch.ensure(c->n);
        double *stage = c->scratch.d.facc;
        launch_rows_from_soa(c->stream, (int)c->n, c->state.d.x, c->state.d.y, c->state.d.z, stage);
        HIPCHK(hipMemcpyAsync(rows.data(), stage, rows.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        domain_from_rows(c->n, rows.data(), c->L, c->lo);
        if (c->nrep > 1) replica_domains(c, rows.data());
        c->have_L = true;
    } else {
        c->have_L = false;
    }
}

// tb3_run / tb3_batch_run.  per_replica: counts are steps x nrep; otherwise
// one total per step.
void run_steps(tb3_ctx *c, int steps, double dt, const double *g, uint64_t *counts, bool per_replica)
{
    const int per = c->nrep > 1 ? c->nrep : 1;
    std::vector<uint64_t> tmp;
    uint64_t *out = counts;
    if (counts && !per_replica && per > 1) {
        tmp.assign((size_t)steps * per, 0);
        out = tmp.data();
    }
    if (steps >= 2 && graph_ok(c)) {
        run_steps_graph(c, steps, dt, g, out);
    } else {
        for (int s = 0; s < steps; ++s) {
            run_pass(c, c->stream, c->state.d, true, dt, g, true, false, true);
            if (out) step_counts(c, out + (size_t)s * per);
        }
    }
    if (out != counts)
        for (int s = 0; s < steps; ++s) {
            uint64_t t = 0;
            for (int r = 0; r < per; ++r) t += tmp[(size_t)s * per + r];
            counts[s] = t;
        }
}

}  // namespace

extern "C" {

const char *tb3_version(void) { return "tb3 1.0.0 gfx950"; }

tb3_status tb3_device_count(int *count)
{
    if (!count) return TB3_EINVAL;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return TB3_OK;
}

tb3_status tb3_create(int device, tb3_ctx **out)
{
    if (!out) return TB3_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_create_error = "no HIP device is visible";
        return TB3_ENODEV;
    }
    if (device < 0 || device >= ndev) {
        g_create_error = "device " + std::to_string(device) + " out of range (" + std::to_string(ndev) + " visible)";
        return TB3_EINVAL;
    }
    tb3_ctx *c = new (std::nothrow) tb3_ctx();
    if (!c) return TB3_ENOMEM;
    c->device = device;
    tb3_status st = guarded(c, [&] {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        int ncu = 0;
        HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
        c->ncu = std::max(1, ncu);
        p3bca_dense_setup();
        init_scratch_address(device);
        c->dense_grid = std::max(1, ncu) * p3bca_dense_blocks_per_cu(c->dense_waves);
        dev_alloc(c->d_slots, 1);
        HIPCHK(hipMemset(c->d_slots, 0, sizeof(Slots)));
        dev_alloc(c->d_status, 1);
        Status init;
        std::memset(&init, 0, sizeof(init));
        init.slots = c->d_slots;
        HIPCHK(hipMemcpy(c->d_status, &init, sizeof(Status), hipMemcpyHostToDevice));
        HIPCHK(hipHostMalloc((void **)&c->h_status, sizeof(Status), hipHostMallocDefault));
        std::memset(c->h_status, 0, sizeof(Status));
    });
    if (st != TB3_OK) {
        g_create_error = c->err;
        tb3_destroy(c);
        return st;
    }
    g_create_error.clear();
    *out = c;
    return TB3_OK;
}

tb3_status tb3_destroy(tb3_ctx *c)
{
    if (!c) return TB3_EINVAL;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->state.release();
    c->scratch.release();
    dev_free(c->grid.cell_of);
    dev_free(c->grid.cell_cnt);
    dev_free(c->grid.cell_start);
    dev_free(c->grid.order);
    dev_free(c->grid.cell_work);
    dev_free(c->grid.cell_rank);
    dev_free(c->grid.scatter_ticket);
    dev_free(c->grid.cell_idx);
    dev_free(c->grid.xs);
    dev_free(c->grid.ys);
    dev_free(c->grid.zs);
    dev_free(c->grid.pad_cnt);
    dev_free(c->grid.pad_state);
    dev_free(c->grid.pad_rec);
    dev_free(c->scan_tmp);
    dev_free(c->triples);
    dev_free(c->d_status);
    dev_free(c->d_slots);
    dev_free(c->d_counts);
    dev_free(c->d_owners);
    dev_free(c->grid.p_off);
    dev_free(c->grid.c_off);
    dev_free(c->grid.rnc);
    dev_free(c->grid.rlo);
    dev_free(c->grid.rL);
    dev_free(c->grid.rep_count);
    dev_free(c->grid.rep_energy);
    if (c->step_exec) (void)hipGraphExecDestroy(c->step_exec);
    if (c->step_exec_mid) (void)hipGraphExecDestroy(c->step_exec_mid);
    for (hipGraphExec_t b : c->step_exec_batch)
        if (b) (void)hipGraphExecDestroy(b);
    for (hipEvent_t &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->h_status) (void)hipHostFree(c->h_status);
    if (c->h_counts) (void)hipHostFree(c->h_counts);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TB3_OK;
}

const char *tb3_last_error(const tb3_ctx *c)
{
    if (!c) return g_create_error.c_str();
    return c->err.c_str();
}

tb3_status tb3_configure(tb3_ctx *c, int algo, double nu, double cutoff, const double *domain_L, int precision)
{
    if (!c) return TB3_EINVAL;
    if (algo < TB3_NATA || algo > TB3_AUTA) return fail(c, TB3_EINVAL, "unknown algorithm " + std::to_string(algo));
    if (precision < TB3_FP64_EXACT || precision > TB3_FP32_MIXED)
        return fail(c, TB3_EINVAL, "unknown precision " + std::to_string(precision));
    if (!std::isfinite(nu)) return fail(c, TB3_EINVAL, "nu must be finite");
    if (std::isnan(cutoff)) return fail(c, TB3_EINVAL, "cutoff is NaN");
    if (domain_L)
        for (int k = 0; k < 3; ++k)
            if (!(domain_L[k] > 0) || !std::isfinite(domain_L[k]))
                return fail(c, TB3_EINVAL, "domain_L entries must be positive and finite");
    return guarded(c, [&] {
        c->algo = algo;
        c->nu = nu;
        c->cutoff = cutoff;
        c->precision = precision;
        c->user_L = domain_L != nullptr;
        if (domain_L) {
            for (int k = 0; k < 3; ++k) {
                c->L[k] = domain_L[k];
                c->lo[k] = 0.0;
            }
        }
        recompute_domain(c);
        configure_grid(c);
        ++c->epoch;
    });
}

tb3_status tb3_set_particles(tb3_ctx *c, int64_t n, const int32_t *id, const double *pos, const double *vel,
                             const double *acc, const double *mass)
{
    if (!c) return TB3_EINVAL;
    if (n < 0 || n > (int64_t)INT32_MAX / 16) return fail(c, TB3_EINVAL, "particle count out of range");
    if (n > 0 && (!pos || !vel || !acc || !mass)) return fail(c, TB3_EINVAL, "pos, vel, acc and mass are required");
    return guarded(c, [&] {
        // the step graph bakes in n and the state arrays
        if (n != c->n || c->nrep != 1) ++c->epoch;
        c->n = n;
        c->nrep = 1;  // a new particle set is one system until tb3_batch_set splits it
        c->ids.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) c->ids[(size_t)i] = id ? id[i] : (int32_t)i;
        const double *old_block = c->state.block;
        c->state.ensure(n);
        if (c->state.block != old_block) ++c->epoch;
        DeviceState &d = c->state.d;
        if (n > 0) {
            // stage n x 3 rows through the role accumulators, then split into SoA
            const size_t bytes = (size_t)n * 3 * sizeof(double);
            HIPCHK(hipMemcpyAsync(d.facc, pos, bytes, hipMemcpyHostToDevice, c->stream));
            launch_soa_from_rows(c->stream, (int)n, d.facc, d.x, d.y, d.z);
            HIPCHK(hipMemcpyAsync(d.facc + 3 * n, vel, bytes, hipMemcpyHostToDevice, c->stream));
            launch_soa_from_rows(c->stream, (int)n, d.facc + 3 * n, d.vx, d.vy, d.vz);
            HIPCHK(hipMemcpyAsync(d.facc + 6 * n, acc, bytes, hipMemcpyHostToDevice, c->stream));
            launch_soa_from_rows(c->stream, (int)n, d.facc + 6 * n, d.ax, d.ay, d.az);
            HIPCHK(hipMemcpyAsync(d.m, mass, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            launch_zero(c->stream, d.fx, n);
            launch_zero(c->stream, d.fy, n);
            launch_zero(c->stream, d.fz, n);
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        if (!c->user_L) {
            if (n > 0) domain_from_rows(n, pos, c->L, c->lo);
            c->have_L = n > 0;
        }
        configure_grid(c);
    });
}

tb3_status tb3_get_particles(const tb3_ctx *cc, int64_t n, int32_t *id, double *pos, double *vel, double *acc,
                             double *mass, double *f)
{
    tb3_ctx *c = const_cast<tb3_ctx *>(cc);
    if (!c) return TB3_EINVAL;
    if (n != c->n) return fail(c, TB3_EINVAL, "n does not match the context's particle count");
    return guarded(c, [&] {
        if (n == 0) return;
        DeviceState &d = c->state.d;
        c->scratch.ensure(n);
        double *stage = c->scratch.d.facc;  // 9n doubles of scratch
        auto rows_out = [&](const double *x, const double *y, const double *z, double *host) {
            if (!host) return;
            launch_rows_from_soa(c->stream, (int)n, x, y, z, stage);
            HIPCHK(hipMemcpyAsync(host, stage, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        };
        rows_out(d.x, d.y, d.z, pos);
        rows_out(d.vx, d.vy, d.vz, vel);
        rows_out(d.ax, d.ay, d.az, acc);
        rows_out(d.fx, d.fy, d.fz, f);
        if (mass) {
            HIPCHK(hipMemcpyAsync(mass, d.m, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        if (id) std::memcpy(id, c->ids.data(), (size_t)n * sizeof(int32_t));
    });
}

static tb3_status check_ready(tb3_ctx *c)
{
    if (uses_cutoff(c) && !c->have_L)
        return fail(c, TB3_EINVAL, "P3BCA with a cutoff needs particles or an explicit domain size");
    if (c->nrep > 1 && !uses_cutoff(c))
        return fail(c, TB3_EUNSUPPORTED, "replica batches run P3BCA with a positive cutoff only");
    return TB3_OK;
}

tb3_status tb3_step(tb3_ctx *c, double dt, const double g[3], uint64_t *n_buffer, uint64_t *n_triplets)
{
    if (!c || !g) return fail(c, TB3_EINVAL, "g is required");
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] {
        const unsigned long long cnt = run_pass(c, c->stream, c->state.d, true, dt, g, true, false, true);
        if (n_buffer) *n_buffer = 1;
        if (n_triplets) *n_triplets = cnt;
    });
}

tb3_status tb3_run(tb3_ctx *c, int steps, double dt, const double g[3], uint64_t *counts)
{
    if (!c || !g || steps < 0) return fail(c, TB3_EINVAL, "bad arguments");
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] { run_steps(c, steps, dt, g, counts, false); });
}

tb3_status tb3_batch_run(tb3_ctx *c, int steps, double dt, const double g[3], uint64_t *counts)
{
    if (!c || !g || steps < 0) return fail(c, TB3_EINVAL, "bad arguments");
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] { run_steps(c, steps, dt, g, counts, true); });
}

tb3_status tb3_batch_set(tb3_ctx *c, int32_t nrep, const int64_t *offsets)
{
    if (!c) return TB3_EINVAL;
    if (nrep < 1 || !offsets) return fail(c, TB3_EINVAL, "nrep >= 1 and offsets are required");
    if (offsets[0] != 0 || offsets[nrep] != c->n)
        return fail(c, TB3_EINVAL, "offsets must run from 0 to the particle count");
    for (int32_t r = 0; r < nrep; ++r)
        if (offsets[r + 1] < offsets[r]) return fail(c, TB3_EINVAL, "offsets must not decrease");
    if (nrep > 1 && !uses_cutoff(c))
        return fail(c, TB3_EUNSUPPORTED, "replica batches run P3BCA with a positive cutoff: configure it first");
    return guarded(c, [&] {
        c->nrep = nrep;
        c->rep_off.assign(offsets, offsets + nrep + 1);
        recompute_domain(c);
        configure_grid(c);
    });
}

tb3_status tb3_batch_counts(tb3_ctx *c, int32_t nrep, uint64_t *counts)
{
    if (!c || !counts) return TB3_EINVAL;
    if (nrep != c->nrep) return fail(c, TB3_EINVAL, "nrep does not match the batch");
    return guarded(c, [&] { step_counts(c, counts); });
}

tb3_status tb3_batch_energy(tb3_ctx *c, int32_t nrep, double *E, uint64_t *counts)
{
    if (!c || !E) return TB3_EINVAL;
    if (nrep != c->nrep) return fail(c, TB3_EINVAL, "nrep does not match the batch");
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] {
        c->scratch.ensure(c->n);
        DeviceState sd = c->scratch.d;
        const DeviceState &d = c->state.d;
        if (c->n > 0) {
            const size_t b = (size_t)c->n * sizeof(double);
            HIPCHK(hipMemcpyAsync(sd.x, d.x, b, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(sd.y, d.y, b, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(sd.z, d.z, b, hipMemcpyDeviceToDevice, c->stream));
        }
        run_pass(c, c->stream, sd, false, 0.0, nullptr, false, true, false);
        if (nrep > 1) {
            HIPCHK(hipMemcpy(E, c->grid.rep_energy, (size_t)nrep * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            E[0] = c->h_status->energy;
        }
        if (counts) step_counts(c, counts);
    });
}

tb3_status tb3_compute_forces(tb3_ctx *c, uint64_t *n_triplets)
{
    if (!c) return TB3_EINVAL;
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] {
        const unsigned long long cnt = run_pass(c, c->stream, c->state.d, false, 0.0, nullptr, true, false, true);
        if (n_triplets) *n_triplets = cnt;
    });
}

// The energy of the held positions at the configured nu, from a pass over a
// copy of the positions so that the held force stays as it is.
static double energy_on_copy(tb3_ctx *c, unsigned long long *count)
{
    c->scratch.ensure(c->n);
    DeviceState sd = c->scratch.d;
    const DeviceState &d = c->state.d;
    if (c->n > 0) {
        const size_t b = (size_t)c->n * sizeof(double);
        HIPCHK(hipMemcpyAsync(sd.x, d.x, b, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(sd.y, d.y, b, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(sd.z, d.z, b, hipMemcpyDeviceToDevice, c->stream));
    }
    const unsigned long long cnt = run_pass(c, c->stream, sd, false, 0.0, nullptr, false, true, false);
    if (count) *count = cnt;
    return c->h_status->energy;
}

tb3_status tb3_force_scale(tb3_ctx *c, double *scale)
{
    if (!c) return TB3_EINVAL;
    if (!scale && c->n > 0) return fail(c, TB3_EINVAL, "scale is required");
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    if (!uses_cutoff(c) || c->nrep > 1)
        return fail(c, TB3_EUNSUPPORTED, "the force scale serves one P3BCA system with a positive cutoff");
    return guarded(c, [&] {
        if (c->n == 0) return;
        hipStream_t s = c->stream;
        c->scratch.ensure(c->n);
        DeviceState sd = c->scratch.d;
        const DeviceState &d = c->state.d;
        const size_t b = (size_t)c->n * sizeof(double);
        HIPCHK(hipMemcpyAsync(sd.x, d.x, b, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(sd.y, d.y, b, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(sd.z, d.z, b, hipMemcpyDeviceToDevice, s));
        // the per-owner kernel with the list capacity grown until no list overflows
        for (int attempt = 0;; ++attempt) {
            launch_prepare(c, s, sd, false, 0.0, false);
            launch_zero(s, sd.fx, sd.n);
            launch_zero(s, sd.fy, sd.n);
            launch_zero(s, sd.fz, sd.n);
            if (sd.n >= 3)
                launch_p3bca_fast(s, sd, c->grid, cutoff_spec(c), c->nu, c->list_cap, 0, 2, c->d_status, 0, 0);
            launch_finalize(s, c->d_status);
            read_status(c, s);
            if (!c->h_status->flags[2]) break;
            const int need = c->h_status->flags[1];
            if (attempt >= 8 || need <= c->list_cap || need > MAX_LIST ||
                p3bca_fast_lds(next_pow2(need + need / 8)) > 160 * 1024)
                throw Fail(TB3_EUNSUPPORTED, "a particle has too many neighbours within the cutoff for the force scale");
            c->list_cap = std::min(next_pow2(need + need / 8), MAX_LIST);
        }
        HIPCHK(hipMemcpyAsync(scale, sd.fx, b, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    });
}

tb3_status tb3_energy(tb3_ctx *c, double *E, uint64_t *n_triplets)
{
    if (!c || !E) return fail(c, TB3_EINVAL, "E is required");
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] {
        unsigned long long cnt = 0;
        *E = energy_on_copy(c, &cnt);
        if (n_triplets) *n_triplets = cnt;
    });
}

tb3_status tb3_energy_grad(tb3_ctx *c, double *E, double *dEdx, double *dEdnu, uint64_t *n_triplets)
{
    if (!c) return TB3_EINVAL;
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] {
        DeviceState &d = c->state.d;
        const unsigned long long cnt = run_pass(c, c->stream, d, false, 0.0, nullptr, true, true, false);
        const double e = c->h_status->energy;
        if (E) *E = e;
        if (n_triplets) *n_triplets = cnt;
        if (dEdx && c->n > 0) {
            c->scratch.ensure(c->n);
            launch_rows_from_soa(c->stream, (int)c->n, d.fx, d.fy, d.fz, c->scratch.d.facc);
            HIPCHK(hipMemcpyAsync(dEdx, c->scratch.d.facc, (size_t)c->n * 3 * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        if (dEdnu) {
            // E is linear in nu, so dE/dnu is the energy at nu = 1: E / nu,
            // or at nu = 0 (where E / nu is undefined) a pass at nu = 1
            if (c->nu != 0.0) {
                *dEdnu = e / c->nu;
            } else {
                c->nu = 1.0;
                double e1 = 0.0;
                try {
                    e1 = energy_on_copy(c, nullptr);
                } catch (...) {
                    c->nu = 0.0;
                    throw;
                }
                c->nu = 0.0;
                *dEdnu = e1;
            }
        }
    });
}

static tb3_status dev_common(tb3_ctx *c, int64_t n, const double *d_pos, double nu, double *d_E, double *d_dEdx,
                             void *stream, uint64_t *n_triplets, bool grad)
{
    if (!c) return TB3_EINVAL;
    if (n < 0 || (n > 0 && !d_pos) || !d_E || (grad && n > 0 && !d_dEdx))
        return fail(c, TB3_EINVAL, "bad device pointers");
    if (!std::isfinite(nu)) return fail(c, TB3_EINVAL, "nu must be finite");
    if (c->nrep > 1 && n != c->n)
        return fail(c, TB3_EINVAL, "a replica batch takes positions for all of its particles");
    return guarded(c, [&] {
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        c->scratch.ensure(n);
        DeviceState sd = c->scratch.d;
        sd.n = (int)n;
        if (n > 0) launch_soa_from_rows(s, (int)n, d_pos, sd.x, sd.y, sd.z);
        if (uses_cutoff(c) && !c->have_L && n > 0) {
            std::vector<double> rows((size_t)(3 * n));
            HIPCHK(hipMemcpyAsync(rows.data(), d_pos, rows.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            domain_from_rows(n, rows.data(), c->L, c->lo);
            c->have_L = true;
            configure_grid(c);
        }
        const double saved_nu = c->nu;
        c->nu = nu;
        try {
            run_pass(c, s, sd, false, 0.0, nullptr, grad, true, true);
        } catch (...) {
            c->nu = saved_nu;
            throw;
        }
        c->nu = saved_nu;
        HIPCHK(hipMemcpyAsync(d_E, &c->d_status->energy, sizeof(double), hipMemcpyDeviceToDevice, s));
        if (grad && n > 0) launch_rows_from_soa(s, (int)n, sd.fx, sd.fy, sd.fz, d_dEdx);
        HIPCHK(hipGetLastError());
        // both reads above use the context's status block and scratch, which
        // the context's next call (on its own stream) clears and reuses:
        // they complete before the call returns
        HIPCHK(hipStreamSynchronize(s));
        if (n_triplets) *n_triplets = c->h_status->count[0];
    });
}

tb3_status tb3_energy_dev(tb3_ctx *c, int64_t n, const double *d_pos, double nu, double *d_E, void *stream,
                          uint64_t *n_triplets)
{
    return dev_common(c, n, d_pos, nu, d_E, nullptr, stream, n_triplets, false);
}

tb3_status tb3_energy_grad_dev(tb3_ctx *c, int64_t n, const double *d_pos, double nu, double *d_E,
                               double *d_dEdx, void *stream, uint64_t *n_triplets)
{
    return dev_common(c, n, d_pos, nu, d_E, d_dEdx, stream, n_triplets, true);
}

}  // extern "C"

namespace {

// Start state of a tb3_energy_grad_dev_noalloc call, written by a kernel
// (no memset): a clear status block pointing at its slots, clear slots, and
// the cell counters, per-cell work and scatter ticket at zero.
__global__ void k_noalloc_clear(Status *st, Slots *slots, int *cell_cnt, unsigned *cell_work, unsigned *ticket,
                                int ncell)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned long long *sw = reinterpret_cast<unsigned long long *>(slots);
    const long long nsw = (long long)(sizeof(Slots) / sizeof(unsigned long long));
    for (long long i = t; i < nsw; i += stride) sw[i] = 0ull;
    for (long long i = t; i < ncell; i += stride) {
        cell_cnt[i] = 0;
        cell_work[i] = 0u;
    }
    if (t == 0) {
        *ticket = 0u;
        st->count[0] = st->count[1] = 0ull;
        st->energy = 0.0;
        for (int k = 0; k < 4; ++k) st->flags[k] = 0;
        st->dqueue[0] = st->dqueue[1] = 0;
        st->dense_n = 0;
        st->want_dense = 0;
        st->slots = slots;
        st->halt = 0;
        st->nstep = 0;
        st->counts = nullptr;
        st->nrep = 1;
        st->rep_count = nullptr;
        st->rep_energy = nullptr;
    }
}

// Carves the working set of one call out of the scratch array: 256-byte
// aligned pieces; returns the bytes needed (which may exceed the array).
struct NoallocLayout {
    Status *st;
    Slots *slots;
    DeviceState d;
    CellGrid g;
    int *owners;
    long long *scan_tmp;
};

size_t noalloc_layout(unsigned char *base, long long n, int ncell, NoallocLayout *L)
{
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const size_t dn = (size_t)std::max<long long>(n, 1) * sizeof(double);
    const size_t in = (size_t)std::max<long long>(n, 1) * sizeof(int);
    const size_t cn = (size_t)std::max(ncell, 1) * sizeof(int);
    NoallocLayout l;
    l.st = reinterpret_cast<Status *>(take(sizeof(Status)));
    l.slots = reinterpret_cast<Slots *>(take(sizeof(Slots)));
    l.d.n = (int)n;
    l.d.x = reinterpret_cast<double *>(take(dn));
    l.d.y = reinterpret_cast<double *>(take(dn));
    l.d.z = reinterpret_cast<double *>(take(dn));
    l.d.fx = reinterpret_cast<double *>(take(dn));
    l.d.fy = reinterpret_cast<double *>(take(dn));
    l.d.fz = reinterpret_cast<double *>(take(dn));
    l.g.cell_of = reinterpret_cast<int *>(take(in));
    l.g.cell_rank = reinterpret_cast<int *>(take(in));
    l.g.cell_idx = reinterpret_cast<int *>(take(in));
    l.owners = reinterpret_cast<int *>(take(in));
    l.g.xs = reinterpret_cast<double *>(take(dn));
    l.g.ys = reinterpret_cast<double *>(take(dn));
    l.g.zs = reinterpret_cast<double *>(take(dn));
    l.g.cell_cnt = reinterpret_cast<int *>(take(cn));
    l.g.cell_start = reinterpret_cast<int *>(take(cn + sizeof(int)));
    l.g.order = reinterpret_cast<int *>(take((size_t)cell_order_len(std::max(ncell, 1)) * sizeof(int)));
    l.g.cell_work = reinterpret_cast<unsigned *>(take(cn));
    l.g.scatter_ticket = reinterpret_cast<unsigned *>(take(sizeof(unsigned)));
    l.scan_tmp = reinterpret_cast<long long *>(
        take((size_t)scan_tmp_elems(std::max<long long>(n, ncell)) * sizeof(long long)));
    if (L) *L = l;
    return off;
}

}  // namespace

extern "C" {

tb3_status tb3_energy_grad_dev_noalloc(tb3_ctx *c, int64_t n, const double *d_pos, double nu, double *d_E,
                                       double *d_dEdx, void *stream, uint64_t *n_triplets)
{
    if (!c) return TB3_EINVAL;
    if (n < 0 || (n > 0 && (!d_pos || !d_dEdx)) || !d_E) return fail(c, TB3_EINVAL, "bad device pointers");
    if (!std::isfinite(nu)) return fail(c, TB3_EINVAL, "nu must be finite");
    if (!uses_cutoff(c))
        return fail(c, TB3_EUNSUPPORTED, "the allocation-free call serves P3BCA with a positive cutoff");
    if (c->precision == TB3_FP64_EXACT)
        return fail(c, TB3_EUNSUPPORTED, "the allocation-free call serves the fp64 and fp32-mixed precisions");
    if (c->nrep > 1) return fail(c, TB3_EUNSUPPORTED, "the allocation-free call serves one particle system");
    if (!c->have_L)
        return fail(c, TB3_EINVAL, "the allocation-free call needs the periodic domain: pass domain_L to "
                                   "tb3_configure or set particles first");
    return guarded(c, [&] {
        unsigned char *base = scratch_address(c->device);
        if (!base) throw Fail(TB3_EHIP, "the scratch array's device address was not recorded");
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        NoallocLayout l;
        l.g.nc[0] = l.g.nc[1] = l.g.nc[2] = 1;
        int nc[3];
        grid_shape(c, c->L, nc);
        const int ncell = nc[0] * nc[1] * nc[2];
        const size_t need = noalloc_layout(base, n, ncell, &l);
        if (need > NOALLOC_BYTES)
            throw Fail(TB3_EINVAL, "the allocation-free call needs " + std::to_string(need) + " bytes of scratch "
                                   "for " + std::to_string(n) + " particles in " + std::to_string(ncell) +
                                   " cells; the scratch array holds " + std::to_string(NOALLOC_BYTES));
        CellGrid &g = l.g;
        for (int k = 0; k < 3; ++k) {
            g.nc[k] = nc[k];
            g.lo[k] = c->lo[k];
            g.L[k] = c->L[k];
        }
        g.ncell = ncell;
        g.nrep = 1;
        DeviceState &d = l.d;
        const int clear_blocks = (int)std::min<long long>(1024, (std::max<long long>(ncell, 12288) + 255) / 256);
        hipLaunchKernelGGL(k_noalloc_clear, dim3(clear_blocks), dim3(256), 0, s, l.st, l.slots, g.cell_cnt,
                           g.cell_work, g.scatter_ticket, ncell);
        HIPCHK(hipGetLastError());
        if (n >= 3) {
            launch_soa_from_rows(s, (int)n, d_pos, d.x, d.y, d.z);
            launch_cells_build(s, d, 0, 0.0, 0.0, 1, g, l.st, l.scan_tmp);
        }
        const CutoffSpec cs = cutoff_spec(c);
        const int mixed = c->precision == TB3_FP32_MIXED ? 1 : 0;
        // the overflow protocol of run_pass, with the context's device
        // buffers replaced by the scratch array: the dense-owner kernel once
        // an owner needs it, the per-owner kernel with a growing list
        // capacity once a neighbourhood is too large for the cell kernel
        bool cells = true, dense = false;
        int cap = c->list_cap;
        Status h;
        for (int attempt = 0;; ++attempt) {
            if (attempt > 0) {
                launch_reset_status(s, l.st, 1);
                launch_zero(s, d.fx, n);
                launch_zero(s, d.fy, n);
                launch_zero(s, d.fz, n);
            }
            if (n >= 3) {
                if (cells) {
                    launch_p3bca_cells(s, d, g, cs, nu, mixed, 1, l.st, 0, 1, l.owners, dense ? 1 : 0);
                    if (dense)
                        launch_p3bca_dense(s, d, g, cs, nu, mixed, 1, l.st, 0, 1, l.owners, c->dense_grid,
                                           c->dense_waves);
                } else {
                    launch_p3bca_fast(s, d, g, cs, nu, cap, mixed, 1, l.st, 0, 1);
                }
            }
            launch_finalize(s, l.st);
            HIPCHK(hipMemcpyAsync(&h, l.st, sizeof(Status), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (!h.flags[2]) break;
            if (attempt >= 8) throw Fail(TB3_EHIP, "the cutoff sweep kept overflowing its LDS staging");
            if (h.flags[3]) cells = false;
            if (h.want_dense) dense = true;
            if (!cells && h.flags[1] > cap) {
                if (h.flags[1] > MAX_LIST)
                    throw Fail(TB3_EUNSUPPORTED, "a particle has " + std::to_string(h.flags[1]) +
                                                     " neighbours within the cutoff; at most " +
                                                     std::to_string(MAX_LIST) + " are supported");
                cap = std::min(next_pow2(h.flags[1] + h.flags[1] / 8), MAX_LIST);
            }
        }
        HIPCHK(hipMemcpyAsync(d_E, &l.st->energy, sizeof(double), hipMemcpyDeviceToDevice, s));
        if (n > 0) {
            if (n >= 3) launch_rows_from_soa(s, (int)n, d.fx, d.fy, d.fz, d_dEdx);
            else launch_zero(s, d_dEdx, 3 * n);
        }
        HIPCHK(hipGetLastError());
        // the scratch array is the next call's: the reads above complete first
        HIPCHK(hipStreamSynchronize(s));
        if (n_triplets) *n_triplets = h.count[0];
    });
}

namespace {

// Waves per owner of the dense-owner kernel, and its persistent grid.
void set_dense_waves(tb3_ctx *c, int waves)
{
    if (waves == c->dense_waves) return;
    c->dense_waves = waves;
    c->dense_grid = c->ncu * p3bca_dense_blocks_per_cu(waves);
    ++c->epoch;  // captured step graphs hold the old launch
}

}  // namespace

tb3_status tb3_autotune(tb3_ctx *c, int reps, double *best_ms, double *cell_ratio, int *blocks_per_cu)
{
    if (!c) return TB3_EINVAL;
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    if (!uses_cutoff(c)) {
        // the all-triplet sweeps have no tunable geometry
        if (best_ms) *best_ms = 0.0;
        if (cell_ratio) *cell_ratio = c->cell_ratio;
        if (blocks_per_cu) *blocks_per_cu = c->cells_per_cu;
        return TB3_OK;
    }
    if (reps < 1) reps = 3;
    return guarded(c, [&] {
        hipStream_t s = c->stream;
        c->scratch.ensure(c->n);
        DeviceState sd = c->scratch.d;
        const DeviceState &d = c->state.d;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        struct Events {
            hipEvent_t a, b;
            ~Events()
            {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
        } guard{e0, e1};
        const bool was_prof = c->prof;
        c->prof = false;
        // best of `reps` force sweeps of the current particles on a scratch copy
        auto time_sweep = [&]() {
            double t = 1e300;
            for (int r = 0; r < reps + 1; ++r) {  // the first run warms up
                if (c->n > 0) {
                    const size_t b = (size_t)c->n * sizeof(double);
                    HIPCHK(hipMemcpyAsync(sd.x, d.x, b, hipMemcpyDeviceToDevice, s));
                    HIPCHK(hipMemcpyAsync(sd.y, d.y, b, hipMemcpyDeviceToDevice, s));
                    HIPCHK(hipMemcpyAsync(sd.z, d.z, b, hipMemcpyDeviceToDevice, s));
                }
                HIPCHK(hipEventRecord(e0, s));
                run_pass(c, s, sd, false, 0.0, nullptr, true, false, false);
                HIPCHK(hipEventRecord(e1, s));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                if (r > 0) t = std::min(t, (double)ms);
            }
            return t;
        };
        // 1. the cell width: the cell kernel runs one workgroup per cell, so
        // the width is its one free parameter (the hardware keeps
        // p3bca_cells_blocks_per_cu of them resident on each CU)
        const int max_per_cu = p3bca_cells_blocks_per_cu();
        const double ratios[] = {1.0, 1.15, 1.3, 1.5, 1.75, 2.0};
        double best = 1e300, best_ratio = c->cell_ratio;
        int best_blocks = c->cells_per_cu;
        for (double ratio : ratios) {
            c->cell_ratio = ratio;
            configure_grid(c);
            const double t = time_sweep();
            if (t < best) {
                best = t;
                best_ratio = ratio;
                best_blocks = c->cells_ok ? max_per_cu : 0;
            }
        }
        c->cell_ratio = best_ratio;
        c->cells_per_cu = best_blocks;
        configure_grid(c);
        // 2. the dense-owner kernel's waves per owner, where that kernel runs
        // (owners whose neighbourhood outgrows the cell kernel's staging)
        if (c->dense_on && c->cells_ok) {
            const int waves[] = {4, 8, 16};
            int best_waves = c->dense_waves;
            double best_w = 1e300;
            for (int w : waves) {
                set_dense_waves(c, w);
                const double t = time_sweep();
                if (t < best_w) {
                    best_w = t;
                    best_waves = w;
                }
            }
            set_dense_waves(c, best_waves);
            best = std::min(best, best_w);
        }
        c->prof = was_prof;
        if (best_ms) *best_ms = best;
        if (cell_ratio) *cell_ratio = best_ratio;
        if (blocks_per_cu) *blocks_per_cu = best_blocks;
    });
}

tb3_status tb3_get_tuning(const tb3_ctx *c, double *cell_ratio, int *blocks_per_cu, int *dense_waves)
{
    if (!c) return TB3_EINVAL;
    if (cell_ratio) *cell_ratio = c->cell_ratio;
    if (blocks_per_cu) *blocks_per_cu = c->cells_per_cu;
    if (dense_waves) *dense_waves = c->dense_waves;
    return TB3_OK;
}

tb3_status tb3_set_tuning(tb3_ctx *c, double cell_ratio, int dense_waves)
{
    if (!c) return TB3_EINVAL;
    if (!(cell_ratio >= 1.0 && cell_ratio <= 4.0))
        return fail(c, TB3_EINVAL, "cell_ratio must lie in [1, 4] (cells at least one cutoff wide)");
    if (!p3bca_dense_waves_ok(dense_waves)) return fail(c, TB3_EINVAL, "dense_waves must be 4, 8 or 16");
    return guarded(c, [&] {
        set_dense_waves(c, dense_waves);
        if (cell_ratio != c->cell_ratio) {
            c->cell_ratio = cell_ratio;
            if (c->n > 0 && uses_cutoff(c)) configure_grid(c);
        }
    });
}

tb3_status tb3_set_cell_slots(tb3_ctx *c, int slots)
{
    if (!c) return TB3_EINVAL;
    if (!(slots == -1 || slots == 0 || (slots >= 1 && slots <= 256)))
        return fail(c, TB3_EINVAL, "slots must be -1 (automatic), 0 (off) or 1 .. 256");
    c->slots_req = slots;
    c->slots_failed = false;
    ++c->epoch;
    return TB3_OK;
}

tb3_status tb3_get_cell_slots(const tb3_ctx *c, int *requested, int *in_use)
{
    if (!c) return TB3_EINVAL;
    if (requested) *requested = c->slots_req;
    if (in_use) *in_use = c->graph_fused && !c->slots_failed ? c->graph_slots : 0;
    return TB3_OK;
}

tb3_status tb3_set_profiling(tb3_ctx *c, int enable)
{
    if (!c) return TB3_EINVAL;
    return guarded(c, [&] {
        if (enable && !c->ev[0])
            for (int k = 0; k < 5; ++k) HIPCHK(hipEventCreate(&c->ev[k]));
        c->prof = enable != 0;
        for (double &v : c->prof_acc) v = 0.0;
    });
}

tb3_status tb3_get_profile(const tb3_ctx *cc, double *out, int64_t n)
{
    tb3_ctx *c = const_cast<tb3_ctx *>(cc);
    if (!c || !out || n < 0) return fail(c, TB3_EINVAL, "bad arguments");
    for (int64_t k = 0; k < n && k < 6; ++k) out[k] = c->prof_acc[k];
    return TB3_OK;
}

tb3_status tb3_selftest_fp64(tb3_ctx *c, int64_t n, const double *a, const double *b, double *sqrt_a,
                             double *a_div_b, double *rsq_a)
{
    if (!c) return TB3_EINVAL;
    if (n < 0 || (n > 0 && (!a || !b || !sqrt_a || !a_div_b))) return fail(c, TB3_EINVAL, "bad arguments");
    return guarded(c, [&] {
        if (n == 0) return;
        double *buf = nullptr;
        HIPCHK(hipMalloc((void **)&buf, (size_t)n * 5 * sizeof(double)));
        try {
            const size_t bytes = (size_t)n * sizeof(double);
            HIPCHK(hipMemcpyAsync(buf, a, bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(buf + n, b, bytes, hipMemcpyHostToDevice, c->stream));
            launch_selftest_fp64(c->stream, n, buf, buf + n, buf + 2 * n, buf + 3 * n, buf + 4 * n);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(sqrt_a, buf + 2 * n, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(a_div_b, buf + 3 * n, bytes, hipMemcpyDeviceToHost, c->stream));
            if (rsq_a) HIPCHK(hipMemcpyAsync(rsq_a, buf + 4 * n, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        } catch (...) {
            (void)hipFree(buf);
            throw;
        }
        HIPCHK(hipFree(buf));
    });
}

tb3_status tb3_selftest_fp64_inrange(tb3_ctx *c, int64_t n, const double *a, const double *b, double *sqrt_a,
                                     double *a_div_b)
{
    if (!c) return TB3_EINVAL;
    if (n < 0 || (n > 0 && (!a || !b || !sqrt_a || !a_div_b))) return fail(c, TB3_EINVAL, "bad arguments");
    return guarded(c, [&] {
        if (n == 0) return;
        double *buf = nullptr;
        HIPCHK(hipMalloc((void **)&buf, (size_t)n * 4 * sizeof(double)));
        try {
            const size_t bytes = (size_t)n * sizeof(double);
            HIPCHK(hipMemcpyAsync(buf, a, bytes, hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(buf + n, b, bytes, hipMemcpyHostToDevice, c->stream));
            launch_selftest_fp64_inrange(c->stream, n, buf, buf + n, buf + 2 * n, buf + 3 * n);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(sqrt_a, buf + 2 * n, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(a_div_b, buf + 3 * n, bytes, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        } catch (...) {
            (void)hipFree(buf);
            throw;
        }
        HIPCHK(hipFree(buf));
    });
}

}  // extern "C"
