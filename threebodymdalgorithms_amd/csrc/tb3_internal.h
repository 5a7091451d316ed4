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

// Uniform cell grid over the periodic domain, used to build neighbour lists.
struct CellGrid {
    int nc[3] = {1, 1, 1};
    int ncell = 1;
    double lo[3] = {0, 0, 0};
    double L[3] = {1, 1, 1};
    int *cell_of = nullptr;     // n
    int *cell_cnt = nullptr;    // ncell
    int *cell_start = nullptr;  // ncell + 1
    int *cell_cur = nullptr;    // ncell
    int *cell_idx = nullptr;    // n
};

struct CutoffSpec {
    int use_cut = 0;
    double c2 = 0;
    double L[3] = {-1, -1, -1};
};

// Device scratch for reductions.
struct Reduce {
    unsigned long long *count = nullptr;  // accepted triplets
    double *energy = nullptr;             // energy sum
    int *flags = nullptr;                 // [0] non-finite position seen, [1] max neighbour count
};

// ---- launchers (all asynchronous on `s`) ----
void launch_soa_from_rows(hipStream_t s, int n, const double *rows, double *x, double *y, double *z);
void launch_rows_from_soa(hipStream_t s, int n, const double *x, const double *y, const double *z, double *rows);
void launch_predict(hipStream_t s, const DeviceState &d, double dt, double h);
void launch_update(hipStream_t s, const DeviceState &d, int sum_roles, double hdt, const double g[3]);
void launch_sum_roles(hipStream_t s, const DeviceState &d);
void launch_zero(hipStream_t s, double *p, long long count);

// Exclusive scan of `count` int32 values into `out` (count + 1 entries, the
// last being the total).  `tmp` must hold scan_tmp_elems(count) int64 values.
long long scan_tmp_elems(long long count);
void launch_scan_i32_to_i64(hipStream_t s, const int *in, long long *out, long long count, long long *tmp);
void launch_scan_i32(hipStream_t s, const int *in, int *out, int count, long long *tmp);

// Cell list: assign -> count -> scan -> scatter -> per-cell index sort.
void launch_cells_build(hipStream_t s, int n, const double *x, const double *y, const double *z, CellGrid &g,
                        int *flags, long long *tmp);

// Neighbour lists (CSR): q != p with !(d2_periodic(p, q) > c2).
// brute = 1 scans all particles (used when positions are not finite).
void launch_nbr_count(hipStream_t s, int n, const double *x, const double *y, const double *z, const CellGrid &g,
                      const CutoffSpec &c, int brute, int *cnt, int *flags);
void launch_nbr_fill(hipStream_t s, int n, const double *x, const double *y, const double *z, const CellGrid &g,
                     const CutoffSpec &c, int brute, const long long *off, int *nbr);
// Sort each particle's list by index (wave-level bitonic sort in LDS).
void launch_nbr_sort(hipStream_t s, int n, const long long *off, int *nbr, int maxlen);

// Bit-exact ordered sweep.  nbr_off == nullptr means every particle is a
// neighbour of every other (NATA/AUTA or no cutoff).
void launch_forces_exact(hipStream_t s, const DeviceState &d, const long long *nbr_off, const int *nbr,
                         const CutoffSpec &c, double nu, unsigned long long *count);

// Fast owner-computes sweep over neighbour lists.  Writes fx/fy/fz (if grad)
// and adds to energy/count.  mixed = FP32 per-triplet arithmetic.
void launch_forces_fast_nbr(hipStream_t s, const DeviceState &d, const long long *nbr_off, const int *nbr,
                            int maxlen, const CutoffSpec &c, double nu, int mixed, int grad, double *energy,
                            unsigned long long *count);

// Fast all-triplets sweep over 64-particle tile triples t0 <= t1 <= t2.
// Accumulates (atomically) into fx/fy/fz, which must be zeroed first.
void launch_forces_tiles(hipStream_t s, const DeviceState &d, const unsigned long long *triples,
                         long long ntriples, double nu, int mixed, int grad, double *energy,
                         unsigned long long *count);

// Element-wise sqrt(a) and a / b (diagnostic for the FP64 rounding).
void launch_selftest_fp64(hipStream_t s, long long n, const double *a, const double *b, double *sq, double *q);

}  // namespace tb3
