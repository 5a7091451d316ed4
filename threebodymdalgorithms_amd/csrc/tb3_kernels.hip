// tb3_kernels.hip -- gfx950 kernels of the three-body (Axilrod-Teller) engine.
//
// Compiled with -ffp-contract=off: the bit-exact path depends on every FP64
// expression rounding as written; the fast paths spell their fused
// multiply-adds out with fma().
#include <climits>

#include "at_math.h"
#include "tb3_internal.h"

namespace tb3 {

namespace {

constexpr int WAVE = 64;

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

inline unsigned blocks_for(long long n, int per) { return (unsigned)((n + per - 1) / per); }

// ---------------------------------------------------------------------------
// Layout conversion and integrator.

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

// Particle::UpdatePredictorStage (utility/structs.hpp:40-55):
// p = (p + v*dt) + aX*h with h = (dt*dt)*0.5, aX for all three axes.
__global__ void k_predict(DeviceState d, double dt, double h)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    const double aX = d.ax[p];
    d.x[p] = d.x[p] + d.vx[p] * dt + aX * h;
    d.y[p] = d.y[p] + d.vy[p] * dt + aX * h;
    d.z[p] = d.z[p] + d.vz[p] * dt + aX * h;
}

// Algorithm::SumUpParticles (Algorithm.cpp:245-249) then Particle::Update
// (structs.hpp:57-99): a = g - f/m, v = v + a*(dt*0.5).
__global__ void k_update(DeviceState d, int sum_roles, double hdt, double gx, double gy, double gz)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.n) return;
    double f[3];
    if (sum_roles) {
        const long long n = d.n;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            f[c] = d.facc[(0 + c) * n + p] + (d.facc[(3 + c) * n + p] + d.facc[(6 + c) * n + p]);
        d.fx[p] = f[0];
        d.fy[p] = f[1];
        d.fz[p] = f[2];
    } else {
        f[0] = d.fx[p];
        f[1] = d.fy[p];
        f[2] = d.fz[p];
    }
    const double m = d.m[p];
    const double a0 = gx - f[0] / m, a1 = gy - f[1] / m, a2 = gz - f[2] / m;
    d.vx[p] = d.vx[p] + a0 * hdt;
    d.vy[p] = d.vy[p] + a1 * hdt;
    d.vz[p] = d.vz[p] + a2 * hdt;
    d.ax[p] = a0;
    d.ay[p] = a1;
    d.az[p] = a2;
}

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

__global__ void k_zero_i32(int *p, long long count)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
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
};

__device__ __forceinline__ int cell_coord(double x, double lo, double L, int nc)
{
    if (nc == 1) return 0;
    double t = (x - lo) / L;
    t = t - __builtin_floor(t);
    int i = (int)(t * (double)nc);
    return i < 0 ? 0 : (i >= nc ? nc - 1 : i);
}

__global__ void k_cell_assign(int n, const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, GridDev g, int *__restrict__ cell_of,
                              int *__restrict__ cell_cnt, int *__restrict__ flags)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double px = x[p], py = y[p], pz = z[p];
    int c = 0;
    if (__builtin_isfinite(px) && __builtin_isfinite(py) && __builtin_isfinite(pz)) {
        const int ix = cell_coord(px, g.lo[0], g.L[0], g.nc[0]);
        const int iy = cell_coord(py, g.lo[1], g.L[1], g.nc[1]);
        const int iz = cell_coord(pz, g.lo[2], g.L[2], g.nc[2]);
        c = (iz * g.nc[1] + iy) * g.nc[0] + ix;
    } else {
        atomicOr(&flags[0], 1);
    }
    cell_of[p] = c;
    atomicAdd(&cell_cnt[c], 1);
}

__global__ void k_cell_scatter(int n, const int *__restrict__ cell_of, const int *__restrict__ cell_start,
                               int *__restrict__ cell_cur, int *__restrict__ cell_idx)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int c = cell_of[p];
    const int slot = atomicAdd(&cell_cur[c], 1);
    cell_idx[cell_start[c] + slot] = p;
}

// Sort each cell's particle indices ascending so the cell order, and every
// list built from it, is independent of atomic scheduling.
__global__ void k_cell_sort(int ncell, const int *__restrict__ cell_start, int *__restrict__ cell_idx)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int b = cell_start[c], e = cell_start[c + 1];
    for (int i = b + 1; i < e; ++i) {
        const int v = cell_idx[i];
        int j = i - 1;
        while (j >= b && cell_idx[j] > v) {
            cell_idx[j + 1] = cell_idx[j];
            --j;
        }
        cell_idx[j + 1] = v;
    }
}

// ---------------------------------------------------------------------------
// Neighbour lists.

struct NbrDev {
    int n;
    const double *x, *y, *z;
    GridDev g;
    const int *cell_of, *cell_start, *cell_idx;
    int use_cut;
    double c2, Lx, Ly, Lz;
};

__device__ __forceinline__ bool is_neighbour(const NbrDev &a, double px, double py, double pz, int q)
{
    if (!a.use_cut) return true;
    const double d2 = sqr_dist_periodic(px, py, pz, a.x[q], a.y[q], a.z[q], a.Lx, a.Ly, a.Lz);
    return !(d2 > a.c2);
}

// Visits every q != p in the 27 (or fewer) cells around p's cell.
template <class F>
__device__ __forceinline__ void for_each_candidate(const NbrDev &a, int p, F &&f)
{
    const int c = a.cell_of[p];
    const int ncx = a.g.nc[0], ncy = a.g.nc[1], ncz = a.g.nc[2];
    const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
    const int rx = ncx >= 3 ? 1 : 0, ry = ncy >= 3 ? 1 : 0, rz = ncz >= 3 ? 1 : 0;
    for (int dz = -rz; dz <= rz; ++dz) {
        int zz = cz + dz;
        zz = zz < 0 ? zz + ncz : (zz >= ncz ? zz - ncz : zz);
        for (int dy = -ry; dy <= ry; ++dy) {
            int yy = cy + dy;
            yy = yy < 0 ? yy + ncy : (yy >= ncy ? yy - ncy : yy);
            for (int dx = -rx; dx <= rx; ++dx) {
                int xx = cx + dx;
                xx = xx < 0 ? xx + ncx : (xx >= ncx ? xx - ncx : xx);
                const int cc = (zz * ncy + yy) * ncx + xx;
                const int e = a.cell_start[cc + 1];
                for (int t = a.cell_start[cc]; t < e; ++t) {
                    const int q = a.cell_idx[t];
                    if (q != p) f(q);
                }
            }
        }
    }
}

__global__ void k_nbr_count(NbrDev a, int brute, int *__restrict__ cnt, int *__restrict__ flags)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const double px = a.x[p], py = a.y[p], pz = a.z[p];
    int k = 0;
    if (brute) {
        for (int q = 0; q < a.n; ++q)
            if (q != p && is_neighbour(a, px, py, pz, q)) ++k;
    } else {
        for_each_candidate(a, p, [&](int q) {
            if (is_neighbour(a, px, py, pz, q)) ++k;
        });
    }
    cnt[p] = k;
    atomicMax(&flags[1], k);
}

__global__ void k_nbr_fill(NbrDev a, int brute, const long long *__restrict__ off, int *__restrict__ nbr)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const double px = a.x[p], py = a.y[p], pz = a.z[p];
    long long w = off[p];
    if (brute) {
        for (int q = 0; q < a.n; ++q)
            if (q != p && is_neighbour(a, px, py, pz, q)) nbr[w++] = q;
    } else {
        for_each_candidate(a, p, [&](int q) {
            if (is_neighbour(a, px, py, pz, q)) nbr[w++] = q;
        });
    }
}

// One wave per particle: bitonic sort of its list in LDS.
__global__ __launch_bounds__(WAVE) void k_nbr_sort(int n, const long long *__restrict__ off, int *__restrict__ nbr)
{
    extern __shared__ int sh_sort[];
    const int p = blockIdx.x;
    const long long base = off[p];
    const int len = (int)(off[p + 1] - base);
    if (len <= 1) return;
    const int lane = lane_id();
    int P = WAVE;
    while (P < len) P <<= 1;
    for (int t = lane; t < P; t += WAVE) sh_sort[t] = t < len ? nbr[base + t] : INT_MAX;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = lane; t < P; t += WAVE) {
                const int u = t ^ j;
                if (u > t) {
                    const int va = sh_sort[t], vb = sh_sort[u];
                    const bool up = (t & k) == 0;
                    if ((va > vb) == up) {
                        sh_sort[t] = vb;
                        sh_sort[u] = va;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int t = lane; t < len; t += WAVE) nbr[base + t] = sh_sort[t];
}

// ---------------------------------------------------------------------------
// Bit-exact ordered sweep.
//
// One wave owns one (particle p, role r) pair and reproduces the sequential
// accumulation the reference performs into b_r[p].f (Algorithm.cpp:72-139):
//   role 0 (p = i): pairs (j, k), p < j < k, lexicographic;
//   role 1 (p = j): pairs (i, k), i < p < k, lexicographic;
//   role 2 (p = k): pairs (i, j), i < j < p, lexicographic;
// restricted to pairs passing the cutoff test.  Lanes test 64 consecutive
// pairs of the stream at a time, accepted pairs are appended in stream order
// to an LDS ring, 64 queued triplets are evaluated in parallel with the
// reference's expression, and three lanes fold the contributions into the
// x, y, z accumulators strictly in stream order: acc = acc - c.

struct ExactDev {
    int n;
    const double *x, *y, *z;
    const long long *nbr_off;
    const int *nbr;
    int use_cut;
    double c2, Lx, Ly, Lz;
    double nu;
    double *facc;
    unsigned long long *count;
};

constexpr int RING = 128;

template <int ROLE>
__device__ void exact_owner(const ExactDev &a, int p)
{
    __shared__ int qa[RING], qb[RING];
    __shared__ double cb[WAVE * 3];
    const int lane = lane_id();
    const int n = a.n;
    const double xp = a.x[p], yp = a.y[p], zp = a.z[p];

    // The owner's neighbour list, sorted by index: explicit (CSR) or implicit
    // (every other particle).  split = number of entries below p.
    long long base = 0;
    int len, split;
    const bool implicit = a.nbr_off == nullptr;
    if (implicit) {
        len = n - 1;
        split = p;
    } else {
        base = a.nbr_off[p];
        len = (int)(a.nbr_off[p + 1] - base);
        int below = 0;
        for (int t = lane; t < len; t += WAVE) below += a.nbr[base + t] < p ? 1 : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) below += __shfl_xor(below, o);
        split = below;
    }
    auto elem = [&](int t) -> int { return implicit ? (t < p ? t : t + 1) : a.nbr[base + t]; };

    int a0, a1, b1;
    bool tri;
    if constexpr (ROLE == 0) {
        a0 = split; a1 = len; b1 = len; tri = true;
    } else if constexpr (ROLE == 1) {
        a0 = 0; a1 = split; b1 = len; tri = false;
    } else {
        a0 = 0; a1 = split; b1 = split; tri = true;
    }

    double acc = 0.0;  // lanes 0, 1, 2 hold x, y, z
    int head = 0, qn = 0;
    unsigned long long accepted = 0;

    auto flush = [&](int cnt) {
        __syncthreads();
        if (lane < cnt) {
            const int slot = (head + lane) & (RING - 1);
            const int ea = qa[slot], eb = qb[slot];
            const double xa = a.x[ea], ya = a.y[ea], za = a.z[ea];
            const double xb = a.x[eb], yb = a.y[eb], zb = a.z[eb];
            double cx, cy, cz;
            if constexpr (ROLE == 0)
                ref_contrib<0>(xp, yp, zp, xa, ya, za, xb, yb, zb, a.nu, cx, cy, cz);
            else if constexpr (ROLE == 1)
                ref_contrib<1>(xa, ya, za, xp, yp, zp, xb, yb, zb, a.nu, cx, cy, cz);
            else
                ref_contrib<2>(xa, ya, za, xb, yb, zb, xp, yp, zp, a.nu, cx, cy, cz);
            cb[lane * 3 + 0] = cx;
            cb[lane * 3 + 1] = cy;
            cb[lane * 3 + 2] = cz;
        }
        __syncthreads();
        if (lane < 3) {
            for (int t = 0; t < cnt; ++t) acc = acc - cb[t * 3 + lane];
        }
        __syncthreads();
        head = (head + cnt) & (RING - 1);
        qn -= cnt;
    };

    for (int ia = a0; ia < a1; ++ia) {
        const int ea = elem(ia);
        const double xa = a.x[ea], ya = a.y[ea], za = a.z[ea];
        const int bstart = tri ? ia + 1 : split;
        for (int b = bstart; b < b1; b += WAVE) {
            const int bl = b + lane;
            bool ok = bl < b1;
            int eb = 0;
            if (ok) {
                eb = elem(bl);
                if (a.use_cut) {
                    const double d2 = sqr_dist_periodic(xa, ya, za, a.x[eb], a.y[eb], a.z[eb], a.Lx, a.Ly, a.Lz);
                    ok = !(d2 > a.c2);
                }
            }
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const int slot = (head + qn + __popcll(m & __lanemask_lt())) & (RING - 1);
                qa[slot] = ea;
                qb[slot] = eb;
            }
            const int got = __popcll(m);
            qn += got;
            accepted += (unsigned long long)got;
            if (qn >= WAVE) flush(WAVE);
        }
    }
    if (qn > 0) flush(qn);

    if (lane < 3) a.facc[(long long)(3 * ROLE + lane) * n + p] = acc;
    if (ROLE == 0 && lane == 0 && accepted) atomicAdd(a.count, accepted);
}

// Wave w handles role w % 3 of particle (w / 3) (role 2 in reverse so the
// longest streams of every role start first).
__global__ __launch_bounds__(WAVE) void k_forces_exact(ExactDev a)
{
    const int w = blockIdx.x;
    const int role = w % 3, q = w / 3;
    if (q >= a.n) return;
    if (role == 0) exact_owner<0>(a, q);
    else if (role == 1) exact_owner<1>(a, q);
    else exact_owner<2>(a, a.n - 1 - q);
}

// ---------------------------------------------------------------------------
// Fast owner-computes sweep over neighbour lists (cutoff algorithm).
//
// One wave per owner p.  Its neighbours are staged in LDS; the wave walks the
// unordered neighbour pairs {a, b} in a folded order (row d pairs a with
// (a + d) mod len) so every lane tests a distinct pair, applies the reference's
// periodic cutoff test to the third side, appends accepted pairs to an LDS
// ring and evaluates 64 queued triplets at a time.  Each triplet contributes
// only dE/dx_p here; the triplet's other two particles compute their own
// share in their own waves.  Energy and the unique-triplet count are taken
// by the owner with the smallest index.

struct FastNbrDev {
    int n;
    const double *x, *y, *z;
    const long long *nbr_off;
    const int *nbr;
    int cap;
    int use_cut;
    double c2, Lx, Ly, Lz;
    double nu;
    double *fx, *fy, *fz;
    double *energy;
    unsigned long long *count;
};

template <class T, bool GRAD>
__global__ __launch_bounds__(WAVE) void k_forces_fast_nbr(FastNbrDev a)
{
    extern __shared__ double smem[];
    const int cap = a.cap;
    double *sx = smem;
    double *sy = sx + cap;
    double *sz = sy + cap;
    double *su = sz + cap;
    int *sid = reinterpret_cast<int *>(su + cap);
    unsigned *ring = reinterpret_cast<unsigned *>(sid + cap);

    const int p = blockIdx.x;
    const int lane = lane_id();
    const long long base = a.nbr_off[p];
    const int len = (int)(a.nbr_off[p + 1] - base);
    const double xp = a.x[p], yp = a.y[p], zp = a.z[p];
    for (int t = lane; t < len; t += WAVE) {
        const int q = a.nbr[base + t];
        const double qx = a.x[q], qy = a.y[q], qz = a.z[q];
        sx[t] = qx;
        sy[t] = qy;
        sz[t] = qz;
        const double dx = qx - xp, dy = qy - yp, dz = qz - zp;
        su[t] = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
        sid[t] = q;
    }
    __syncthreads();

    const T nu = (T)a.nu, K = (T)(0.375 * a.nu);
    double gx = 0.0, gy = 0.0, gz = 0.0, esum = 0.0;
    unsigned cnt = 0;
    int head = 0, qn = 0;

    auto flush = [&](int nq) {
        __syncthreads();
        if (lane < nq) {
            const unsigned e = ring[(head + lane) & (RING - 1)];
            const int ia = (int)(e & 0xffffu), ib = (int)(e >> 16);
            const double djx = sx[ia] - xp, djy = sy[ia] - yp, djz = sz[ia] - zp;
            const double dkx = sx[ib] - xp, dky = sy[ib] - yp, dkz = sz[ib] - zp;
            const double ex = sx[ib] - sx[ia], ey = sy[ib] - sy[ia], ez = sz[ib] - sz[ia];
            T u, v, w;
            if constexpr (sizeof(T) == 8) {
                u = su[ia];
                v = su[ib];
                w = __builtin_fma(ex, ex, __builtin_fma(ey, ey, ez * ez));
            } else {
                u = (T)su[ia];
                v = (T)su[ib];
                const T fx_ = (T)ex, fy_ = (T)ey, fz_ = (T)ez;
                w = fma_t<T>(fx_, fx_, fma_t<T>(fy_, fy_, fz_ * fz_));
            }
            T eu, ev, ew, et;
            at_coeffs<T, false>(u, v, w, nu, K, eu, ev, ew, et);
            if (GRAD) {
                const double du = -(double)eu, dv = -(double)ev;
                gx = __builtin_fma(du, djx, __builtin_fma(dv, dkx, gx));
                gy = __builtin_fma(du, djy, __builtin_fma(dv, dky, gy));
                gz = __builtin_fma(du, djz, __builtin_fma(dv, dkz, gz));
            }
            if (p < sid[ia] && p < sid[ib]) {
                esum += (double)et;
                ++cnt;
            }
        }
        __syncthreads();
        head = (head + nq) & (RING - 1);
        qn -= nq;
    };

    const int H = (len - 1) / 2;
    const long long full = (long long)H * len;
    const long long total = (long long)len * (len - 1) / 2;
    const float inv_len = len > 0 ? 1.0f / (float)len : 0.0f;
    for (long long t0 = 0; t0 < total; t0 += WAVE) {
        const long long t = t0 + lane;
        bool ok = t < total;
        int ia = 0, ib = 0;
        if (ok) {
            if (t < full) {
                // t = d * len + r with 0 <= r < len: float estimate, exact fix-up
                int d = (int)((float)t * inv_len);
                long long r = t - (long long)d * len;
                while (r < 0) {
                    --d;
                    r += len;
                }
                while (r >= len) {
                    ++d;
                    r -= len;
                }
                ia = (int)r;
                ib = ia + d + 1;
                if (ib >= len) ib -= len;
            } else {
                ia = (int)(t - full);
                ib = ia + len / 2;
            }
            if (a.use_cut) {
                const double d2 = sqr_dist_periodic(sx[ia], sy[ia], sz[ia], sx[ib], sy[ib], sz[ib], a.Lx, a.Ly, a.Lz);
                ok = !(d2 > a.c2);
            }
        }
        const unsigned long long m = __ballot(ok);
        if (ok) ring[(head + qn + __popcll(m & __lanemask_lt())) & (RING - 1)] = (unsigned)ia | ((unsigned)ib << 16);
        qn += __popcll(m);
        if (qn >= WAVE) flush(WAVE);
    }
    if (qn > 0) flush(qn);

    if (GRAD) {
        gx = wave_sum(gx);
        gy = wave_sum(gy);
        gz = wave_sum(gz);
        if (lane == 0) {
            a.fx[p] = gx;
            a.fy[p] = gy;
            a.fz[p] = gz;
        }
    }
    esum = wave_sum(esum);
    const unsigned long long c = wave_sum_u64(cnt);
    if (lane == 0) {
        if (a.energy && esum != 0.0) unsafeAtomicAdd(a.energy, esum);
        if (c) atomicAdd(a.count, c);
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
    double *energy;
    unsigned long long *count;
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
            T u, v, w;
            if constexpr (sizeof(T) == 8) {
                u = u64;
                v = __builtin_fma(dikx, dikx, __builtin_fma(diky, diky, dikz * dikz));
                w = __builtin_fma(djkx, djkx, __builtin_fma(djky, djky, djkz * djkz));
            } else {
                u = (T)u64;
                const T ax_ = (T)dikx, ay_ = (T)diky, az_ = (T)dikz;
                const T bx_ = (T)djkx, by_ = (T)djky, bz_ = (T)djkz;
                v = fma_t<T>(ax_, ax_, fma_t<T>(ay_, ay_, az_ * az_));
                w = fma_t<T>(bx_, bx_, fma_t<T>(by_, by_, bz_ * bz_));
            }
            T eu, ev, ew, et;
            at_coeffs<T, true>(u, v, w, nu, K, eu, ev, ew, et);
            const double cu = ok ? (double)eu : 0.0;
            const double cv = ok ? (double)ev : 0.0;
            const double cw = ok ? (double)ew : 0.0;
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
    esum = wave_sum(esum);
    const unsigned long long c = wave_sum_u64(cnt);
    if (lane == 0) {
        if (a.energy && esum != 0.0) unsafeAtomicAdd(a.energy, esum);
        if (c) atomicAdd(a.count, c);
    }
}

__global__ void k_selftest_fp64(long long n, const double *__restrict__ a, const double *__restrict__ b,
                                double *__restrict__ sq, double *__restrict__ q)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sq[i] = __builtin_sqrt(a[i]);
    q[i] = a[i] / b[i];
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.

void launch_selftest_fp64(hipStream_t s, long long n, const double *a, const double *b, double *sq, double *q)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_selftest_fp64, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, a, b, sq, q);
}

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

void launch_predict(hipStream_t s, const DeviceState &d, double dt, double h)
{
    if (d.n <= 0) return;
    hipLaunchKernelGGL(k_predict, dim3(blocks_for(d.n, 256)), dim3(256), 0, s, d, dt, h);
}

void launch_update(hipStream_t s, const DeviceState &d, int sum_roles, double hdt, const double g[3])
{
    if (d.n <= 0) return;
    hipLaunchKernelGGL(k_update, dim3(blocks_for(d.n, 256)), dim3(256), 0, s, d, sum_roles, hdt, g[0], g[1], g[2]);
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

void launch_scan_i32_to_i64(hipStream_t s, const int *in, long long *out, long long count, long long *tmp)
{
    scan_rec<int, long long>(s, in, count, out, count + 1, tmp);
}

void launch_scan_i32(hipStream_t s, const int *in, int *out, int count, long long *tmp)
{
    scan_rec<int, int>(s, in, count, out, (long long)count + 1, tmp);
}

void launch_cells_build(hipStream_t s, int n, const double *x, const double *y, const double *z, CellGrid &g,
                        int *flags, long long *tmp)
{
    GridDev gd;
    for (int c = 0; c < 3; ++c) {
        gd.nc[c] = g.nc[c];
        gd.lo[c] = g.lo[c];
        gd.L[c] = g.L[c];
    }
    hipLaunchKernelGGL(k_zero_i32, dim3(blocks_for(g.ncell, 256)), dim3(256), 0, s, g.cell_cnt, (long long)g.ncell);
    hipLaunchKernelGGL(k_zero_i32, dim3(blocks_for(g.ncell, 256)), dim3(256), 0, s, g.cell_cur, (long long)g.ncell);
    if (n > 0)
        hipLaunchKernelGGL(k_cell_assign, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, x, y, z, gd, g.cell_of,
                           g.cell_cnt, flags);
    launch_scan_i32(s, g.cell_cnt, g.cell_start, g.ncell, tmp);
    if (n > 0)
        hipLaunchKernelGGL(k_cell_scatter, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, g.cell_of, g.cell_start,
                           g.cell_cur, g.cell_idx);
    hipLaunchKernelGGL(k_cell_sort, dim3(blocks_for(g.ncell, 64)), dim3(64), 0, s, g.ncell, g.cell_start, g.cell_idx);
}

static NbrDev make_nbr_dev(int n, const double *x, const double *y, const double *z, const CellGrid &g,
                           const CutoffSpec &c)
{
    NbrDev a;
    a.n = n;
    a.x = x;
    a.y = y;
    a.z = z;
    for (int k = 0; k < 3; ++k) {
        a.g.nc[k] = g.nc[k];
        a.g.lo[k] = g.lo[k];
        a.g.L[k] = g.L[k];
    }
    a.cell_of = g.cell_of;
    a.cell_start = g.cell_start;
    a.cell_idx = g.cell_idx;
    a.use_cut = c.use_cut;
    a.c2 = c.c2;
    a.Lx = c.L[0];
    a.Ly = c.L[1];
    a.Lz = c.L[2];
    return a;
}

void launch_nbr_count(hipStream_t s, int n, const double *x, const double *y, const double *z, const CellGrid &g,
                      const CutoffSpec &c, int brute, int *cnt, int *flags)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_nbr_count, dim3(blocks_for(n, 128)), dim3(128), 0, s, make_nbr_dev(n, x, y, z, g, c), brute,
                       cnt, flags);
}

void launch_nbr_fill(hipStream_t s, int n, const double *x, const double *y, const double *z, const CellGrid &g,
                     const CutoffSpec &c, int brute, const long long *off, int *nbr)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_nbr_fill, dim3(blocks_for(n, 128)), dim3(128), 0, s, make_nbr_dev(n, x, y, z, g, c), brute,
                       off, nbr);
}

void launch_nbr_sort(hipStream_t s, int n, const long long *off, int *nbr, int maxlen)
{
    if (n <= 0 || maxlen <= 1) return;
    int P = WAVE;
    while (P < maxlen) P <<= 1;
    hipLaunchKernelGGL(k_nbr_sort, dim3((unsigned)n), dim3(WAVE), (size_t)P * sizeof(int), s, n, off, nbr);
}

void launch_forces_exact(hipStream_t s, const DeviceState &d, const long long *nbr_off, const int *nbr,
                         const CutoffSpec &c, double nu, unsigned long long *count)
{
    if (d.n <= 0) return;
    ExactDev a;
    a.n = d.n;
    a.x = d.x;
    a.y = d.y;
    a.z = d.z;
    a.nbr_off = nbr_off;
    a.nbr = nbr;
    a.use_cut = c.use_cut;
    a.c2 = c.c2;
    a.Lx = c.L[0];
    a.Ly = c.L[1];
    a.Lz = c.L[2];
    a.nu = nu;
    a.facc = d.facc;
    a.count = count;
    hipLaunchKernelGGL(k_forces_exact, dim3((unsigned)(3 * d.n)), dim3(WAVE), 0, s, a);
}

void launch_forces_fast_nbr(hipStream_t s, const DeviceState &d, const long long *nbr_off, const int *nbr,
                            int maxlen, const CutoffSpec &c, double nu, int mixed, int grad, double *energy,
                            unsigned long long *count)
{
    if (d.n <= 0) return;
    FastNbrDev a;
    a.n = d.n;
    a.x = d.x;
    a.y = d.y;
    a.z = d.z;
    a.nbr_off = nbr_off;
    a.nbr = nbr;
    a.cap = maxlen < 1 ? 1 : maxlen;
    a.use_cut = c.use_cut;
    a.c2 = c.c2;
    a.Lx = c.L[0];
    a.Ly = c.L[1];
    a.Lz = c.L[2];
    a.nu = nu;
    a.fx = d.fx;
    a.fy = d.fy;
    a.fz = d.fz;
    a.energy = energy;
    a.count = count;
    const size_t shm = (size_t)a.cap * (4 * sizeof(double) + sizeof(int)) + RING * sizeof(unsigned);
    const dim3 grid((unsigned)d.n), blk(WAVE);
    if (mixed) {
        if (grad) hipLaunchKernelGGL((k_forces_fast_nbr<float, true>), grid, blk, shm, s, a);
        else hipLaunchKernelGGL((k_forces_fast_nbr<float, false>), grid, blk, shm, s, a);
    } else {
        if (grad) hipLaunchKernelGGL((k_forces_fast_nbr<double, true>), grid, blk, shm, s, a);
        else hipLaunchKernelGGL((k_forces_fast_nbr<double, false>), grid, blk, shm, s, a);
    }
}

void launch_forces_tiles(hipStream_t s, const DeviceState &d, const unsigned long long *triples,
                         long long ntriples, double nu, int mixed, int grad, double *energy,
                         unsigned long long *count)
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
    a.energy = energy;
    a.count = count;
    const dim3 grid((unsigned)ntriples), blk(WAVE);
    if (mixed) {
        if (grad) hipLaunchKernelGGL((k_forces_tiles<float, true>), grid, blk, 0, s, a);
        else hipLaunchKernelGGL((k_forces_tiles<float, false>), grid, blk, 0, s, a);
    } else {
        if (grad) hipLaunchKernelGGL((k_forces_tiles<double, true>), grid, blk, 0, s, a);
        else hipLaunchKernelGGL((k_forces_tiles<double, false>), grid, blk, 0, s, a);
    }
}

}  // namespace tb3
