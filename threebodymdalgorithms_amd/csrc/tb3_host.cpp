// tb3_host.cpp -- host side of the C ABI declared in include/tb3.h.
//
// Owns the device buffers of one particle system, builds the cell grid and
// neighbour lists for the cutoff algorithm, and dispatches the sweep kernels
// by algorithm and precision.  No exception crosses the ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
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

// Working set for one sweep: positions in, forces out, plus role accumulators.
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
    Buffers scratch;  // device-pointer API working set

    // cell grid and neighbour lists
    CellGrid grid;
    int grid_cell_cap = 0, grid_n_cap = 0;
    int *nbr_cnt = nullptr;
    long long *nbr_off = nullptr;
    int *nbr = nullptr;
    long long nbr_n_cap = 0, nbr_cap = 0;
    int nbr_maxlen = 0;
    long long *scan_tmp = nullptr;
    long long scan_tmp_cap = 0;

    // tile triples for the all-triplets fast sweep
    unsigned long long *triples = nullptr;
    long long ntriples = 0;
    long long triples_n = -1;

    // reductions and host staging
    unsigned long long *d_count = nullptr;  // [0] sweep, [1] discard
    double *d_energy = nullptr;
    int *d_flags = nullptr;
    void *h_pinned = nullptr;  // 64 bytes

    // phase timing (tb3_set_profiling)
    bool prof = false;       // enabled by the caller
    bool recording = false;  // inside a profiled step
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

void domain_from_rows(long long n, const double *rows, double *L, double *lo)
{
    // RegularGridDecomposition::getDomainMinMax / Init (RegularGridDecomposition.cpp:24-28, 135-166)
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
void configure_grid(tb3_ctx *c)
{
    long long total = 1;
    for (int k = 0; k < 3; ++k) {
        int nc = 1;
        if (uses_cutoff(c) && c->L[k] > 0 && std::isfinite(c->L[k])) {
            const double q = std::floor(c->L[k] / (c->cutoff * (1.0 + 1e-6)));
            nc = q >= 3.0 ? (int)std::min(q, 1024.0) : 1;
        }
        c->grid.nc[k] = nc;
        total *= nc;
    }
    while (total > (1LL << 22)) {  // keep the grid bounded: widen the largest dimension
        int kmax = 0;
        for (int k = 1; k < 3; ++k)
            if (c->grid.nc[k] > c->grid.nc[kmax]) kmax = k;
        c->grid.nc[kmax] = std::max(3, c->grid.nc[kmax] / 2);
        total = (long long)c->grid.nc[0] * c->grid.nc[1] * c->grid.nc[2];
    }
    c->grid.ncell = (int)total;
    for (int k = 0; k < 3; ++k) {
        c->grid.lo[k] = c->lo[k];
        c->grid.L[k] = c->L[k];
    }
}

void ensure_scan_tmp(tb3_ctx *c, long long count)
{
    const long long need = scan_tmp_elems(count);
    if (need > c->scan_tmp_cap) {
        dev_alloc(c->scan_tmp, need);
        c->scan_tmp_cap = need;
    }
}

template <class T>
void readback(tb3_ctx *c, hipStream_t s, T *host, const T *dev, size_t count)
{
    HIPCHK(hipMemcpyAsync(c->h_pinned, dev, count * sizeof(T), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::memcpy(host, c->h_pinned, count * sizeof(T));
}

// Cell list + CSR neighbour lists for the positions in d.
void build_neighbours(tb3_ctx *c, hipStream_t s, const DeviceState &d, bool sorted)
{
    const int n = d.n;
    CellGrid &g = c->grid;
    if (g.ncell > c->grid_cell_cap) {
        dev_alloc(g.cell_cnt, g.ncell);
        dev_alloc(g.cell_start, (long long)g.ncell + 1);
        dev_alloc(g.cell_cur, g.ncell);
        c->grid_cell_cap = g.ncell;
    }
    if (n > c->grid_n_cap) {
        dev_alloc(g.cell_of, n);
        dev_alloc(g.cell_idx, n);
        c->grid_n_cap = n;
    }
    if (n > c->nbr_n_cap) {
        dev_alloc(c->nbr_cnt, n);
        dev_alloc(c->nbr_off, (long long)n + 1);
        c->nbr_n_cap = n;
    }
    ensure_scan_tmp(c, std::max<long long>(n, g.ncell));
    HIPCHK(hipMemsetAsync(c->d_flags, 0, 2 * sizeof(int), s));
    const CutoffSpec cs = cutoff_spec(c);
    launch_cells_build(s, n, d.x, d.y, d.z, g, c->d_flags, c->scan_tmp);
    launch_nbr_count(s, n, d.x, d.y, d.z, g, cs, 0, c->nbr_cnt, c->d_flags);
    launch_scan_i32_to_i64(s, c->nbr_cnt, c->nbr_off, n, c->scan_tmp);
    HIPCHK(hipGetLastError());
    long long total = 0;
    int flags[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(c->h_pinned, c->nbr_off + n, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync((char *)c->h_pinned + 16, c->d_flags, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::memcpy(&total, c->h_pinned, sizeof(long long));
    std::memcpy(flags, (char *)c->h_pinned + 16, sizeof(flags));
    int brute = 0;
    if (flags[0]) {
        // Non-finite positions defeat binning; the reference's brute sweep
        // still pairs them (NaN compares false), so scan all pairs instead.
        brute = 1;
        HIPCHK(hipMemsetAsync(c->d_flags + 1, 0, sizeof(int), s));
        launch_nbr_count(s, n, d.x, d.y, d.z, g, cs, 1, c->nbr_cnt, c->d_flags);
        launch_scan_i32_to_i64(s, c->nbr_cnt, c->nbr_off, n, c->scan_tmp);
        HIPCHK(hipMemcpyAsync(c->h_pinned, c->nbr_off + n, sizeof(long long), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync((char *)c->h_pinned + 16, c->d_flags, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::memcpy(&total, c->h_pinned, sizeof(long long));
        std::memcpy(flags, (char *)c->h_pinned + 16, sizeof(flags));
    }
    if (flags[1] > 65535)
        throw Fail(TB3_EUNSUPPORTED, "a particle has " + std::to_string(flags[1]) +
                                         " neighbours within the cutoff; at most 65535 are supported");
    if (total > c->nbr_cap) {
        const long long cap = std::max<long long>(total + total / 4, 1024);
        dev_alloc(c->nbr, cap);
        c->nbr_cap = cap;
    }
    c->nbr_maxlen = flags[1];
    launch_nbr_fill(s, n, d.x, d.y, d.z, g, cs, brute, c->nbr_off, c->nbr);
    if (sorted) launch_nbr_sort(s, n, c->nbr_off, c->nbr, c->nbr_maxlen);
    HIPCHK(hipGetLastError());
}

void ensure_triples(tb3_ctx *c, long long n)
{
    if (c->triples_n == n) return;
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

// One sweep over the positions in d.
//   grad   : write dE/dx (the reference's f) into d.fx/fy/fz
//   energy : accumulate E into c->d_energy
// The accepted-triplet count lands in c->d_count[0].
void sweep(tb3_ctx *c, hipStream_t s, DeviceState &d, bool grad, bool energy)
{
    HIPCHK(hipMemsetAsync(c->d_count, 0, 2 * sizeof(unsigned long long), s));
    HIPCHK(hipMemsetAsync(c->d_energy, 0, sizeof(double), s));
    const bool cut = uses_cutoff(c);
    const bool exact = c->precision == TB3_FP64_EXACT;
    const int mixed = c->precision == TB3_FP32_MIXED ? 1 : 0;
    const CutoffSpec cs = cutoff_spec(c);
    if (d.n < 3) {
        if (grad) {
            launch_zero(s, d.fx, d.n);
            launch_zero(s, d.fy, d.n);
            launch_zero(s, d.fz, d.n);
        }
        return;
    }
    if (cut) {
        build_neighbours(c, s, d, exact && grad);
        const size_t lds = (size_t)std::max(c->nbr_maxlen, 1) * (4 * sizeof(double) + sizeof(int)) + 512;
        if (lds > 64 * 1024)
            throw Fail(TB3_EUNSUPPORTED, "neighbour lists of " + std::to_string(c->nbr_maxlen) +
                                             " particles exceed the 64 KiB LDS staging of the cutoff sweep");
    }
    const bool fast_grad = grad && !exact;
    if (!cut && (fast_grad || energy)) {
        ensure_triples(c, d.n);
        if (fast_grad) {
            launch_zero(s, d.fx, d.n);
            launch_zero(s, d.fy, d.n);
            launch_zero(s, d.fz, d.n);
        }
    }
    if (c->recording) HIPCHK(hipEventRecord(c->ev[2], s));
    if (exact && grad)
        launch_forces_exact(s, d, cut ? c->nbr_off : nullptr, cut ? c->nbr : nullptr, cs, c->nu, c->d_count);
    if (fast_grad || energy) {
        unsigned long long *cnt = (exact && grad) ? c->d_count + 1 : c->d_count;
        double *e = energy ? c->d_energy : nullptr;
        if (cut)
            launch_forces_fast_nbr(s, d, c->nbr_off, c->nbr, c->nbr_maxlen, cs, c->nu, exact ? 0 : mixed,
                                   fast_grad ? 1 : 0, e, cnt);
        else
            launch_forces_tiles(s, d, c->triples, c->ntriples, c->nu, exact ? 0 : mixed, fast_grad ? 1 : 0, e,
                                cnt);
    }
    if (c->recording) HIPCHK(hipEventRecord(c->ev[3], s));
    if (exact && grad) launch_sum_roles(s, d);
    HIPCHK(hipGetLastError());
}

// Accumulates the phase times of the step just recorded (the stream has
// been synchronised by the caller's readback).
void accumulate_profile(tb3_ctx *c, unsigned long long triplets)
{
    float ms[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) HIPCHK(hipEventElapsedTime(&ms[k], c->ev[k], c->ev[k + 1]));
    c->prof_acc[0] += 1.0;
    for (int k = 0; k < 4; ++k) c->prof_acc[1 + k] += ms[k];
    c->prof_acc[5] += (double)triplets;
}

void recompute_domain(tb3_ctx *c)
{
    if (c->user_L) {
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
        c->have_L = true;
    } else {
        c->have_L = false;
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
        dev_alloc(c->d_count, 2);
        dev_alloc(c->d_energy, 1);
        dev_alloc(c->d_flags, 2);
        HIPCHK(hipHostMalloc(&c->h_pinned, 64, hipHostMallocDefault));
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
    dev_free(c->grid.cell_cur);
    dev_free(c->grid.cell_idx);
    dev_free(c->nbr_cnt);
    dev_free(c->nbr_off);
    dev_free(c->nbr);
    dev_free(c->scan_tmp);
    dev_free(c->triples);
    dev_free(c->d_count);
    dev_free(c->d_energy);
    dev_free(c->d_flags);
    for (hipEvent_t &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
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
    });
}

tb3_status tb3_set_particles(tb3_ctx *c, int64_t n, const int32_t *id, const double *pos, const double *vel,
                             const double *acc, const double *mass)
{
    if (!c) return TB3_EINVAL;
    if (n < 0 || n > (int64_t)INT32_MAX / 16) return fail(c, TB3_EINVAL, "particle count out of range");
    if (n > 0 && (!pos || !vel || !acc || !mass)) return fail(c, TB3_EINVAL, "pos, vel, acc and mass are required");
    return guarded(c, [&] {
        c->n = n;
        c->ids.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i) c->ids[(size_t)i] = id ? id[i] : (int32_t)i;
        c->state.ensure(n);
        DeviceState &d = c->state.d;
        if (n > 0) {
            // stage n x 3 rows through facc, then split into SoA
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
    return TB3_OK;
}

// Predictor (optional) -> sweep -> update (optional) on the context's own
// particles, with phase events when profiling.  Returns the accepted count.
static unsigned long long step_once(tb3_ctx *c, bool integrate, double dt, const double *g)
{
    DeviceState &d = c->state.d;
    c->recording = c->prof;
    struct Reset {
        tb3_ctx *c;
        ~Reset() { c->recording = false; }
    } reset{c};
    if (c->recording) HIPCHK(hipEventRecord(c->ev[0], c->stream));
    if (integrate) launch_predict(c->stream, d, dt, dt * dt * 0.5);
    if (c->recording) HIPCHK(hipEventRecord(c->ev[1], c->stream));
    sweep(c, c->stream, d, true, false);
    if (integrate) launch_update(c->stream, d, 0, dt * 0.5, g);
    if (c->recording) HIPCHK(hipEventRecord(c->ev[4], c->stream));
    HIPCHK(hipGetLastError());
    unsigned long long cnt = 0;
    readback(c, c->stream, &cnt, c->d_count, 1);
    if (c->recording) accumulate_profile(c, cnt);
    return cnt;
}

tb3_status tb3_step(tb3_ctx *c, double dt, const double g[3], uint64_t *n_buffer, uint64_t *n_triplets)
{
    if (!c || !g) return fail(c, TB3_EINVAL, "g is required");
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] {
        const unsigned long long cnt = step_once(c, true, dt, g);
        if (n_buffer) *n_buffer = 1;
        if (n_triplets) *n_triplets = cnt;
    });
}

tb3_status tb3_run(tb3_ctx *c, int steps, double dt, const double g[3], uint64_t *counts)
{
    if (!c || !g || steps < 0) return fail(c, TB3_EINVAL, "bad arguments");
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] {
        for (int s = 0; s < steps; ++s) {
            const unsigned long long cnt = step_once(c, true, dt, g);
            if (counts) counts[s] = cnt;
        }
    });
}

tb3_status tb3_compute_forces(tb3_ctx *c, uint64_t *n_triplets)
{
    if (!c) return TB3_EINVAL;
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] {
        const unsigned long long cnt = step_once(c, false, 0.0, nullptr);
        if (n_triplets) *n_triplets = cnt;
    });
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

tb3_status tb3_energy(tb3_ctx *c, double *E, uint64_t *n_triplets)
{
    if (!c || !E) return fail(c, TB3_EINVAL, "E is required");
    tb3_status st = check_ready(c);
    if (st != TB3_OK) return st;
    return guarded(c, [&] {
        // the energy pass must not disturb the held force: run on scratch
        c->scratch.ensure(c->n);
        DeviceState sd = c->scratch.d;
        const DeviceState &d = c->state.d;
        if (c->n > 0) {
            const size_t b = (size_t)c->n * sizeof(double);
            HIPCHK(hipMemcpyAsync(sd.x, d.x, b, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(sd.y, d.y, b, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(sd.z, d.z, b, hipMemcpyDeviceToDevice, c->stream));
        }
        sweep(c, c->stream, sd, false, true);
        unsigned long long cnt = 0;
        readback(c, c->stream, E, c->d_energy, 1);
        readback(c, c->stream, &cnt, c->d_count, 1);
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
        sweep(c, c->stream, d, true, true);
        double e = 0.0;
        unsigned long long cnt = 0;
        readback(c, c->stream, &e, c->d_energy, 1);
        readback(c, c->stream, &cnt, c->d_count, 1);
        if (E) *E = e;
        if (dEdnu) *dEdnu = c->nu != 0.0 ? e / c->nu : 0.0;
        if (n_triplets) *n_triplets = cnt;
        if (dEdx && c->n > 0) {
            c->scratch.ensure(c->n);
            launch_rows_from_soa(c->stream, (int)c->n, d.fx, d.fy, d.fz, c->scratch.d.facc);
            HIPCHK(hipMemcpyAsync(dEdx, c->scratch.d.facc, (size_t)c->n * 3 * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
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
    return guarded(c, [&] {
        hipStream_t s = stream ? (hipStream_t)stream : c->stream;
        c->scratch.ensure(n);
        DeviceState sd = c->scratch.d;
        sd.n = (int)n;
        if (n > 0) launch_soa_from_rows(s, (int)n, d_pos, sd.x, sd.y, sd.z);
        if (uses_cutoff(c) && !c->user_L && !c->have_L) {
            std::vector<double> rows((size_t)(3 * n));
            if (n > 0) {
                HIPCHK(hipMemcpyAsync(rows.data(), d_pos, rows.size() * sizeof(double), hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                domain_from_rows(n, rows.data(), c->L, c->lo);
                c->have_L = true;
                configure_grid(c);
            }
        }
        const double saved_nu = c->nu;
        c->nu = nu;
        try {
            sweep(c, s, sd, grad, true);
        } catch (...) {
            c->nu = saved_nu;
            throw;
        }
        c->nu = saved_nu;
        HIPCHK(hipMemcpyAsync(d_E, c->d_energy, sizeof(double), hipMemcpyDeviceToDevice, s));
        if (grad && n > 0) launch_rows_from_soa(s, (int)n, sd.fx, sd.fy, sd.fz, d_dEdx);
        HIPCHK(hipGetLastError());
        if (n_triplets) {
            unsigned long long cnt = 0;
            readback(c, s, &cnt, c->d_count, 1);
            *n_triplets = cnt;
        }
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

tb3_status tb3_selftest_fp64(tb3_ctx *c, int64_t n, const double *a, const double *b, double *sqrt_a,
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
            launch_selftest_fp64(c->stream, n, buf, buf + n, buf + 2 * n, buf + 3 * n);
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
