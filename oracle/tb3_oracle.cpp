/*
 * tb3_oracle.cpp -- CPU oracle (test infrastructure only; see tb3_oracle.h).
 *
 * Restates the reference's one-process hot path.  Compiled with
 * -ffp-contract=off -fno-fast-math so that every expression rounds exactly as
 * written; citations are relative to the reference's src/ directory.
 */
#include "tb3_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* 120-byte particle record, field for field as Utility::Particle
 * (utility/structs.hpp:11-18).  The brute sweep runs over this AoS layout so
 * the CPU baseline has the reference's memory shape. */
struct Particle {
    int ID;
    double pX, pY, pZ;
    double vX, vY, vZ;
    double aX, aY, aZ;
    double fX, fY, fZ;
    double mass;
    bool isDummy;
};
static_assert(sizeof(Particle) == 120, "reference particle record is 120 bytes");

/* Eigen's Vector3d::norm() restated as sqrt((x*x + y*y) + z*z)
 * (AxilrodTeller.cpp:29-31).  The summation order inside Eigen is not
 * observable without the reference binary: parity unpinned at ulp level. */
inline double norm3(double x, double y, double z) { return std::sqrt((x * x + y * y) + z * z); }

/* The bracket of dVdRa/b/c (AxilrodTeller.cpp:68-81) written once for an own
 * side x and the two other sides y, z in the order each of the three
 * expressions names them:  a -> (x=a, y=b, z=c), b -> (x=b, y=a, z=c),
 * c -> (x=c, y=b, z=a).  Products associate left to right as in the source. */
inline double at_bracket(double nu, const double *dx, const double *dy, const double *dz)
{
    /* d?[0..5] hold d1..d6 */
    const double d1x = dx[0], d2x = dx[1], d4x = dx[3], d6x = dx[5];
    const double d1y = dy[0], d3y = dy[2], d5y = dy[4];
    const double d1z = dz[0], d3z = dz[2], d5z = dz[4];
    double t = -8. / (d4x * d3y * d3z);
    t = t - 1. / (d5y * d5z);
    t = t + 5. * d1y / (d6x * d5z);
    t = t + 5. * d1z / (d6x * d5y);
    t = t - 1. / (d2x * d3y * d5z);
    t = t - 1. / (d2x * d5y * d3z);
    t = t - 3. / (d4x * d1y * d5z);
    t = t - 3. / (d4x * d5y * d1z);
    t = t - 5. / (d6x * d1y * d3z);
    t = t - 5. / (d6x * d3y * d1z);
    t = t + 6. / (d4x * d3y * d3z);
    return (3. * nu / (8. * d1x)) * t;
}

inline void powers(double d1, double *d)
{
    /* AxilrodTeller.cpp:33-51: d2 = d1*d1, d3 = d2*d1, ... d6 = d5*d1 */
    d[0] = d1;
    d[1] = d[0] * d1;
    d[2] = d[1] * d1;
    d[3] = d[2] * d1;
    d[4] = d[3] * d1;
    d[5] = d[4] * d1;
}

/* AxilrodTeller::CalculateForces body (AxilrodTeller.cpp:22-93). */
inline void at_contrib_xyz(double xi, double yi, double zi, double xj, double yj, double zj, double xk, double yk,
                           double zk, double nu, double *ci, double *cj, double *ck)
{
    /* sides a = rj - ri, b = ri - rk, c = rj - rk (23-25) */
    const double d1a = norm3(xj - xi, yj - yi, zj - zi);
    const double d1b = norm3(xi - xk, yi - yk, zi - zk);
    const double d1c = norm3(xj - xk, yj - yk, zj - zk);
    double da[6], db[6], dc[6];
    powers(d1a, da);
    powers(d1b, db);
    powers(d1c, dc);

    const double dXa = xj - xi, dYa = yj - yi, dZa = zj - zi; /* 53-55 */
    const double dXb = xk - xi, dYb = yk - yi, dZb = zk - zi; /* 57-59 */
    const double dXc = xk - xj, dYc = yk - yj, dZc = zk - zj; /* 61-63 */

    const double dVdRa = at_bracket(nu, da, db, dc); /* 68-71 */
    const double dVdRb = at_bracket(nu, db, da, dc); /* 73-76 */
    const double dVdRc = at_bracket(nu, dc, db, da); /* 78-81 */

    ci[0] = dXa * dVdRa + dXb * dVdRb; /* 83-85 */
    ci[1] = dYa * dVdRa + dYb * dVdRb;
    ci[2] = dZa * dVdRa + dZb * dVdRb;
    cj[0] = dXa * (-dVdRa) + dXc * dVdRc; /* 87-89 */
    cj[1] = dYa * (-dVdRa) + dYc * dVdRc;
    cj[2] = dZa * (-dVdRa) + dZc * dVdRc;
    ck[0] = dXb * (-dVdRb) + dXc * (-dVdRc); /* 91-93 */
    ck[1] = dYb * (-dVdRb) + dYc * (-dVdRc);
    ck[2] = dZb * (-dVdRb) + dZc * (-dVdRc);
}

/* GetSqrDistPeriodic (structs.hpp:106-115); std::min(a, b) is (b < a) ? b : a. */
inline double sqr_dist_periodic(double ax, double ay, double az, double bx, double by, double bz, const double *L)
{
    const double absX = std::fabs(bx - ax);
    const double absY = std::fabs(by - ay);
    const double absZ = std::fabs(bz - az);
    const double lx = L[0] - absX, ly = L[1] - absY, lz = L[2] - absZ;
    const double xP = (lx < absX) ? lx : absX;
    const double yP = (ly < absY) ? ly : absY;
    const double zP = (lz < absZ) ? lz : absZ;
    return xP * xP + yP * yP + zP * zP;
}

/* The per-triplet plug-in point, called through a virtual function as the
 * reference calls Potential::CalculateForces (potential/Potential.hpp:25). */
struct PotentialBase {
    virtual ~PotentialBase() {}
    virtual void CalculateForces(Particle &i, Particle &j, Particle &k) = 0;
};
struct AxilrodTellerOracle final : PotentialBase {
    double nu;
    explicit AxilrodTellerOracle(double v) : nu(v) {}
    void CalculateForces(Particle &i, Particle &j, Particle &k) override
    {
        double ci[3], cj[3], ck[3];
        at_contrib_xyz(i.pX, i.pY, i.pZ, j.pX, j.pY, j.pZ, k.pX, k.pY, k.pZ, nu, ci, cj, ck);
        /* AxilrodTeller.cpp:119-129 */
        i.fX -= ci[0];
        i.fY -= ci[1];
        i.fZ -= ci[2];
        j.fX -= cj[0];
        j.fY -= cj[1];
        j.fZ -= cj[2];
        k.fX -= ck[0];
        k.fY -= ck[1];
        k.fZ -= ck[2];
    }
};
PotentialBase *volatile g_potential_slot = nullptr; /* keeps the call virtual */

std::vector<Particle> make_buffer(int64_t n, const double *pos)
{
    std::vector<Particle> b((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        Particle &p = b[(size_t)i];
        std::memset(&p, 0, sizeof(p));
        p.ID = (int)i;
        p.pX = pos[3 * i + 0];
        p.pY = pos[3 * i + 1];
        p.pZ = pos[3 * i + 2];
        p.isDummy = false;
    }
    return b;
}

/* Algorithm::calculateInteractions loop nest (Algorithm.cpp:72-139) at one
 * process: owners (0,0,0) so j starts at i+1 and k at j+1 (77-87). */
uint64_t sweep_aos(std::vector<Particle> &b0, std::vector<Particle> &b1, std::vector<Particle> &b2, double cutoff,
                   const double *L, int64_t row_end, uint64_t *possible_out, PotentialBase *pot)
{
    uint64_t act = 0, possible = 0;
    const double sqrCutoff = cutoff * cutoff; /* 67 */
    const size_t n = b0.size();
    const size_t iend = row_end < 0 ? n : std::min(n, (size_t)row_end);
    for (size_t i = 0; i < iend; ++i) {
        if (b0[i].isDummy) continue;
        for (size_t j = i + 1; j < b1.size(); ++j) {
            if (b1[j].isDummy) continue;
            for (size_t k = j + 1; k < b2.size(); ++k) {
                if (b2[k].isDummy) continue;
                possible++;
                if (cutoff > 0) { /* 94-100 */
                    if (sqr_dist_periodic(b0[i].pX, b0[i].pY, b0[i].pZ, b1[j].pX, b1[j].pY, b1[j].pZ, L) >
                            sqrCutoff ||
                        sqr_dist_periodic(b0[i].pX, b0[i].pY, b0[i].pZ, b2[k].pX, b2[k].pY, b2[k].pZ, L) >
                            sqrCutoff ||
                        sqr_dist_periodic(b1[j].pX, b1[j].pY, b1[j].pZ, b2[k].pX, b2[k].pY, b2[k].pZ, L) >
                            sqrCutoff) {
                        continue;
                    }
                }
                pot->CalculateForces(b0[i], b1[j], b2[k]); /* 109 */
                act++;                                      /* 138 */
            }
        }
    }
    if (possible_out) *possible_out = possible;
    return act;
}

int resolve_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    return nthreads;
#else
    (void)nthreads;
    return 1;
#endif
}

/* Index-sorted neighbour lists: q != p with !(d2(p,q) > c^2) when cutoff > 0
 * (the negated test of Algorithm.cpp:95-97), every q != p otherwise. */
void neighbour_lists(int64_t n, const double *pos, double cutoff, const double *L,
                     std::vector<std::vector<int>> &nb, int nthreads)
{
    nb.assign((size_t)n, std::vector<int>());
    const double c2 = cutoff * cutoff;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t p = 0; p < n; ++p) {
        std::vector<int> &l = nb[(size_t)p];
        for (int64_t q = 0; q < n; ++q) {
            if (q == p) continue;
            if (cutoff > 0) {
                const double d2 = sqr_dist_periodic(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2], pos[3 * q],
                                                    pos[3 * q + 1], pos[3 * q + 2], L);
                if (d2 > c2) continue;
            }
            l.push_back((int)q);
        }
    }
}

inline bool pair_ok(const double *pos, int a, int b, double cutoff, double c2, const double *L)
{
    if (!(cutoff > 0)) return true;
    return !(sqr_dist_periodic(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], pos[3 * b], pos[3 * b + 1],
                               pos[3 * b + 2], L) > c2);
}

/* Algorithm::calculateInteractions with the owner rules of a multi-rank run
 * (Algorithm.cpp:51-140): j starts at i+1 when b1 and b0 have one owner,
 * else at 0 (77); k starts at j+1 when b2 and b1 have one owner, else at i+1
 * when b2 and b0 do, else at 0 (81-87); b0 rows [start, start + steps), or
 * every row when steps is -1 (72).  NATA and AUTA call the form without a
 * cutoff (Algorithm.cpp:23-36, cutoff -1); P3BCA passes its cutoff and the
 * physical domain size (94-100). */
uint64_t sweep_aos_owned(std::vector<Particle> &b0, std::vector<Particle> &b1, std::vector<Particle> &b2, int o0,
                         int o1, int o2, int64_t start, int64_t steps, PotentialBase *pot, double cutoff = -1.0,
                         const double *L = nullptr)
{
    const double sqrCutoff = cutoff * cutoff;
    uint64_t act = 0;
    const size_t iend = steps != -1 ? (size_t)(start + steps) : b0.size();
    for (size_t i = (size_t)start; i < iend; ++i) {
        if (b0[i].isDummy) continue;
        const size_t jstart = o1 == o0 ? i + 1 : 0;
        for (size_t j = jstart; j < b1.size(); ++j) {
            if (b1[j].isDummy) continue;
            size_t kstart = 0;
            if (o2 == o1)
                kstart = j + 1;
            else if (o2 == o0)
                kstart = i + 1;
            for (size_t k = kstart; k < b2.size(); ++k) {
                if (b2[k].isDummy) continue;
                if (cutoff > 0) {
                    if (sqr_dist_periodic(b0[i].pX, b0[i].pY, b0[i].pZ, b1[j].pX, b1[j].pY, b1[j].pZ, L) >
                            sqrCutoff ||
                        sqr_dist_periodic(b0[i].pX, b0[i].pY, b0[i].pZ, b2[k].pX, b2[k].pY, b2[k].pZ, L) >
                            sqrCutoff ||
                        sqr_dist_periodic(b1[j].pX, b1[j].pY, b1[j].pZ, b2[k].pX, b2[k].pY, b2[k].pZ, L) >
                            sqrCutoff) {
                        continue;
                    }
                }
                pot->CalculateForces(b0[i], b1[j], b2[k]);
                act++;
            }
        }
    }
    return act;
}

/* AtomDecomposition::Init (AtomDecomposition.cpp:6-40): the first
 * p - (n mod p) ranks hold floor(n/p) consecutive particles plus one dummy
 * (when p does not divide n), the others floor(n/p) + 1, so every buffer has
 * the same length. */
std::vector<Particle> atom_share(int64_t n, const double *pos, int p, int r, int64_t *offset, int64_t *count)
{
    const int64_t normal = n / p, procs_normal = p - (int)(n % p);
    int64_t cnt, off;
    if (r < procs_normal) {
        cnt = normal;
        off = r * normal;
    } else {
        cnt = normal + 1;
        off = procs_normal * normal + (r - procs_normal) * cnt;
    }
    std::vector<Particle> b = make_buffer(cnt, pos + 3 * off);
    for (int64_t t = 0; t < cnt; ++t) b[(size_t)t].ID = (int)(off + t);
    if (r < procs_normal && n % p != 0) {
        Particle d;
        std::memset(&d, 0, sizeof(d));
        d.isDummy = true;
        b.push_back(d);
    }
    *offset = off;
    *count = cnt;
    return b;
}

inline int ring_mod(int a, int p) { return ((a % p) + p) % p; }

/* Utility::Triplet::operator== (utility/structs.hpp:164-169): equal as
 * unordered triples. */
inline bool same_triplet(const int *a, int b0, int b1, int b2)
{
    int x[3] = {a[0], a[1], a[2]}, y[3] = {b0, b1, b2};
    std::sort(x, x + 3);
    std::sort(y, y + 3);
    return x[0] == y[0] && x[1] == y[1] && x[2] == y[2];
}

/* The grid decomposition of a P3BCA run on a Cartesian decomposition dims
 * (1 to 3 dimensions, periodic; ranks numbered row-major as MPI_Cart_create
 * numbers them without reordering): RegularGridDecomposition::Init
 * (RegularGridDecomposition.cpp:6-39) with the domain of getDomainMinMax
 * (135-166), cell widths diff / dims, rank r's cell from min + coords w to
 * that plus w.  members[r] lists rank r's particles in buffer order. */
struct P3bcaRanks {
    int p = 1, nd = 1, dims[3] = {1, 1, 1}, ncb[3] = {0, 0, 0}, numSteps = 1;
    double L[3], lo[3], w[3];
    std::vector<std::array<double, 3>> cmin, cmax;
    std::vector<std::vector<int64_t>> members;

    std::array<int, 3> coords(int r) const
    {
        return {r / (dims[1] * dims[2]), (r / dims[2]) % dims[1], r % dims[2]};
    }
    int rank(const std::array<int, 3> &c) const { return (c[0] * dims[1] + c[1]) * dims[2] + c[2]; }
    /* schedule1D / schedule2D / schedule3D (P3BCA.cpp:100, 130-139, 183-195):
     * the i-th source of the rank at c */
    std::array<int, 3> sched(int i, const std::array<int, 3> &c) const
    {
        if (nd == 1) return {ring_mod(i + c[0], dims[0]), c[1], c[2]};
        const int lenY = 2 * ncb[1] + 1;
        if (nd == 2) {
            const int gy = ((i + ncb[1]) % lenY) - ncb[1];
            const int gx = (i + ncb[1]) / lenY;
            return {ring_mod(gx + c[0], dims[0]), ring_mod(gy + c[1], dims[1]), c[2]};
        }
        const int lenZ = 2 * ncb[2] + 1;
        const int gz = ((i + ncb[2]) % lenZ) - ncb[2];
        const int gy = ((((i + ncb[2]) / lenZ) + ncb[1]) % lenY) - ncb[1];
        const int gx = (i + ((lenZ * lenY) / 2)) / (lenZ * lenY);
        return {ring_mod(gx + c[0], dims[0]), ring_mod(gy + c[1], dims[1]), ring_mod(gz + c[2], dims[2])};
    }
    /* calcDiff1D/2D/3D (120-128, 209-245): the (i-1)-th source minus src,
     * component by component, unreduced */
    std::array<int, 3> diff(const std::array<int, 3> &c, const std::array<int, 3> &src, int i) const
    {
        const std::array<int, 3> old = sched(i - 1, c);
        return {old[0] - src[0], old[1] - src[1], old[2] - src[2]};
    }
};

/* Init: binParticles keeps, in input order, the particles with
 * min < x <= max in every dimension (isInsideLocalCell, 125-133); then
 * P3BCA::Init's cutoff boxes and the validity check the reference exits on
 * (P3BCA.cpp:7-83), and calcSteps (85-98).  Returns 0 or -2. */
int p3bca_ranks_init(P3bcaRanks &g, int nd, const int *dims, int64_t n, const double *pos, double cutoff)
{
    g.nd = nd;
    g.p = 1;
    for (int k = 0; k < 3; ++k) {
        g.dims[k] = k < nd ? dims[k] : 1;
        g.p *= g.dims[k];
    }
    orc_domain_size(n, pos, g.L, g.lo);
    for (int k = 0; k < 3; ++k) g.w[k] = g.L[k] / (double)g.dims[k];
    g.members.assign((size_t)g.p, std::vector<int64_t>());
    g.cmin.resize((size_t)g.p);
    g.cmax.resize((size_t)g.p);
    for (int r = 0; r < g.p; ++r) {
        const std::array<int, 3> c = g.coords(r);
        for (int k = 0; k < 3; ++k) {
            g.cmin[(size_t)r][k] = g.lo[k] + (double)c[k] * g.w[k];
            g.cmax[(size_t)r][k] = g.cmin[(size_t)r][k] + g.w[k];
        }
        const std::array<double, 3> &mn = g.cmin[(size_t)r], &mx = g.cmax[(size_t)r];
        for (int64_t i = 0; i < n; ++i) {
            const double *q = pos + 3 * i;
            if (q[0] <= mn[0] || q[1] <= mn[1] || q[2] <= mn[2] || q[0] > mx[0] || q[1] > mx[1] || q[2] > mx[2])
                continue;
            g.members[(size_t)r].push_back(i);
        }
    }
    g.ncb[0] = g.p == 1 ? 0 : (g.p == 2 ? 1 : (int)std::ceil(cutoff / g.w[0]));
    g.ncb[1] = nd > 1 ? (int)std::ceil(cutoff / g.w[1]) : 0;
    g.ncb[2] = nd > 2 ? (int)std::ceil(cutoff / g.w[2]) : 0;
    if (g.p > 2 && (g.ncb[0] < 1 || (double)g.ncb[0] >= (double)g.dims[0] / 2.0)) return -2;
    if (nd > 1 && (g.ncb[1] < 1 || (double)g.ncb[1] >= (double)g.dims[1] / 2.0)) return -2;
    if (nd > 2 && (g.ncb[2] < 1 || (double)g.ncb[2] >= (double)g.dims[2] / 2.0)) return -2;
    const int x = g.ncb[0], y = g.ncb[1], z = g.ncb[2];
    g.numSteps = nd == 1 ? x + 1 : (nd == 2 ? 2 * x * y + x + y + 1 : 1 + z + y * (z * 2 + 1) + x * ((y * 2 + 1) * (z * 2 + 1)));
    return 0;
}

/* RegularGridDecomposition::Update after the integrator (168-177):
 * exchangeParticlesDim for each dimension in turn (59-123).  Each rank
 * sends, in buffer order, the particles below its cell to the lower
 * neighbour and those above it to the upper neighbour (periodic ranks),
 * keeps the rest in order, and appends what it receives from the lower
 * neighbour, then from the upper one. */
void p3bca_ranks_exchange(P3bcaRanks &g, const double *pos)
{
    const int p = g.p;
    if (p <= 1) return;
    for (int d = 0; d < g.nd; ++d) {
        std::vector<std::vector<int64_t>> toL((size_t)p), toR((size_t)p);
        for (int r = 0; r < p; ++r) {
            std::vector<int64_t> keep;
            for (int64_t i : g.members[(size_t)r]) {
                const double x = pos[3 * i + d];
                if (x < g.cmin[(size_t)r][d]) toL[(size_t)r].push_back(i);
                else if (x > g.cmax[(size_t)r][d]) toR[(size_t)r].push_back(i);
                else keep.push_back(i);
            }
            g.members[(size_t)r] = keep;
        }
        for (int r = 0; r < p; ++r) {
            std::array<int, 3> lc = g.coords(r), rc = lc;
            lc[d] = ring_mod(lc[d] - 1, g.dims[d]);
            rc[d] = ring_mod(rc[d] + 1, g.dims[d]);
            for (int64_t i : toR[(size_t)g.rank(lc)]) g.members[(size_t)r].push_back(i);
            for (int64_t i : toL[(size_t)g.rank(rc)]) g.members[(size_t)r].push_back(i);
        }
    }
}

/* P3BCA::SimulationStep (P3BCA.cpp:516-805) over the current members. */
int p3bca_ranks_sweep(const P3bcaRanks &g, int64_t n, const double *pos, double nu, double cutoff, double *f,
                       uint64_t *buffer_counts, uint64_t *particle_counts, int32_t *processed, int64_t cap,
                       int64_t *nprocessed)
{
    const int p = g.p, numSteps = g.numSteps;
    const double *L = g.L;
    const std::vector<std::vector<int64_t>> &members = g.members;
    std::vector<std::vector<Particle>> own((size_t)p), c1((size_t)p), c2((size_t)p);
    for (int r = 0; r < p; ++r) {
        std::vector<Particle> b((size_t)members[(size_t)r].size());
        for (size_t t = 0; t < b.size(); ++t) {
            const int64_t i = members[(size_t)r][t];
            std::memset(&b[t], 0, sizeof(Particle));
            b[t].ID = (int)i;
            b[t].pX = pos[3 * i];
            b[t].pY = pos[3 * i + 1];
            b[t].pZ = pos[3 * i + 2];
        }
        own[(size_t)r] = b;
        c1[(size_t)r] = b;
        c2[(size_t)r] = b;
    }
    AxilrodTellerOracle at(nu);
    g_potential_slot = &at;
    /* SimulationStep (516-805): b1 and b2 start as copies of b0 and travel;
     * shiftLeft (321-367) hands rank r the buffer that the rank at
     * (next source + offset vector) held, then adds the step's difference
     * to the offset vector (handleOffsetVector*, 278-319).  Every rank runs
     * the same loop; the 2-D and 3-D helpers may skip inner steps, by an
     * amount that does not depend on the rank. */
    using A3 = std::array<int, 3>;
    std::vector<int> hold1((size_t)p), hold2((size_t)p);
    std::vector<A3> nsi((size_t)p), dfi((size_t)p), nso((size_t)p), dfo((size_t)p), ovi((size_t)p, A3{0, 0, 0}),
        ovo((size_t)p, A3{0, 0, 0});
    for (int r = 0; r < p; ++r) hold1[(size_t)r] = hold2[(size_t)r] = r;
    auto shift = [&](std::vector<int> &hold, const std::vector<A3> &src, std::vector<A3> &ov, const std::vector<A3> &df) {
        const std::vector<int> old = hold;
        for (int r = 0; r < p; ++r) {
            A3 cs;
            for (int k = 0; k < 3; ++k) cs[k] = ring_mod(src[(size_t)r][k] + ov[(size_t)r][k], g.dims[k]);
            hold[(size_t)r] = old[(size_t)g.rank(cs)];
            for (int k = 0; k < 3; ++k) ov[(size_t)r][k] += df[(size_t)r][k];
        }
    };
    for (int r = 0; r < p; ++r) buffer_counts[r] = particle_counts[r] = 0;
    int64_t np = 0;
    const int outer = p == 2 ? 1 : numSteps; /* the edge case of two ranks (577) */
    for (int i2 = 0; i2 < outer; ++i2) {
        for (int i3 = i2; i3 < numSteps; ++i3) {
            for (int r = 0; r < p; ++r) {
                const int o1 = hold1[(size_t)r], o2 = hold2[(size_t)r];
                particle_counts[r] += sweep_aos_owned(own[(size_t)r], c1[(size_t)o1], c2[(size_t)o2], r, o1, o2, 0, -1,
                                                      g_potential_slot, cutoff, L);
                buffer_counts[r]++;
                if (processed && np < cap) {
                    processed[4 * np] = r;
                    processed[4 * np + 1] = r;
                    processed[4 * np + 2] = o1;
                    processed[4 * np + 3] = o2;
                }
                ++np;
            }
            if (i3 < numSteps - 1) {
                /* schedule1DHelper (110-118) / schedule2DHelper, schedule3DHelper
                 * (151-181, 247-276): the next source, advancing i3 while it
                 * does not continue the current b1's sequence (the loop's
                 * post-increment stops it at i3 = 0) */
                int i3n = i3;
                for (int r = 0; r < p; ++r) {
                    const A3 c = g.coords(r);
                    int j = i3;
                    A3 src;
                    if (g.nd == 1) {
                        src = g.sched(j + 1, c);
                    } else {
                        for (;;) {
                            const A3 b1Init = g.sched(i2, c);
                            src = g.sched(j + 1, c);
                            const A3 src2 = g.sched(j + 1 - i2, b1Init);
                            if (src == src2 || !(j < numSteps)) break;
                            if (j++ == 0) break;
                        }
                    }
                    if (r == 0) i3n = j;
                    else if (j != i3n) return -3;
                    nsi[(size_t)r] = src;
                    dfi[(size_t)r] = g.diff(c, src, i3 + 1);
                }
                i3 = i3n;
                if (i3 < numSteps - 1) shift(hold2, nsi, ovi, dfi);
            }
        }
        if (i2 < outer - 1) {
            /* schedule and calcDiff for both buffers; b1 moves with the outer
             * offset vector, b2 with the inner one, which then takes the
             * outer one's value */
            for (int r = 0; r < p; ++r) {
                const A3 c = g.coords(r);
                nso[(size_t)r] = nsi[(size_t)r] = g.sched(i2 + 1, c);
                dfo[(size_t)r] = dfi[(size_t)r] = g.diff(c, nso[(size_t)r], i2 + 1);
            }
            shift(hold1, nso, ovo, dfo);
            shift(hold2, nsi, ovi, dfi);
            ovi = ovo;
        }
    }
    /* sendBackParticles (415-504), then SumUpParticles on each rank's own
     * buffers; particles no rank binned keep a zero force */
    for (int64_t i = 0; i < 3 * n; ++i) f[i] = 0.0;
    for (int q = 0; q < p; ++q)
        for (size_t t = 0; t < members[(size_t)q].size(); ++t) {
            const Particle &a = own[(size_t)q][t], &b = c1[(size_t)q][t], &c = c2[(size_t)q][t];
            double *fp = f + 3 * members[(size_t)q][t];
            fp[0] = a.fX + (b.fX + c.fX);
            fp[1] = a.fY + (b.fY + c.fY);
            fp[2] = a.fZ + (b.fZ + c.fZ);
        }
    if (nprocessed) *nprocessed = np;
    return 0;
}


}  // namespace

extern "C" {

void orc_at_contrib(const double *pi, const double *pj, const double *pk, double nu, double *ci, double *cj,
                    double *ck)
{
    at_contrib_xyz(pi[0], pi[1], pi[2], pj[0], pj[1], pj[2], pk[0], pk[1], pk[2], nu, ci, cj, ck);
}

double orc_sqr_dist_periodic(const double *pa, const double *pb, const double *L)
{
    return sqr_dist_periodic(pa[0], pa[1], pa[2], pb[0], pb[1], pb[2], L);
}

uint64_t orc_sweep_brute_rows(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f0,
                              double *f1, double *f2, int64_t rows, uint64_t *accepted)
{
    /* ResetForces then b0 = copy, b1 = b0, b2 = b0 (P3BCA.cpp:518-538; AUTA.cpp:239-247) */
    std::vector<Particle> b0 = make_buffer(n, pos);
    std::vector<Particle> b1 = b0;
    std::vector<Particle> b2 = b0;
    AxilrodTellerOracle at(nu);
    g_potential_slot = &at;
    const double Lz[3] = {-1, -1, -1};
    uint64_t possible = 0;
    uint64_t act = sweep_aos(b0, b1, b2, cutoff, L ? L : Lz, rows, &possible, g_potential_slot);
    for (int64_t i = 0; i < n; ++i) {
        const Particle &p0 = b0[(size_t)i], &p1 = b1[(size_t)i], &p2 = b2[(size_t)i];
        if (f0) {
            f0[3 * i] = p0.fX;
            f0[3 * i + 1] = p0.fY;
            f0[3 * i + 2] = p0.fZ;
        }
        if (f1) {
            f1[3 * i] = p1.fX;
            f1[3 * i + 1] = p1.fY;
            f1[3 * i + 2] = p1.fZ;
        }
        if (f2) {
            f2[3 * i] = p2.fX;
            f2[3 * i + 1] = p2.fY;
            f2[3 * i + 2] = p2.fZ;
        }
    }
    if (accepted) *accepted = act;
    return possible;
}

uint64_t orc_sweep_brute(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f0,
                         double *f1, double *f2, uint64_t *possible)
{
    uint64_t act = 0;
    uint64_t poss = orc_sweep_brute_rows(n, pos, nu, cutoff, L, f0, f1, f2, -1, &act);
    if (possible) *possible = poss;
    return act;
}

uint64_t orc_sweep_owner(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f0,
                         double *f1, double *f2, int nthreads)
{
    nthreads = resolve_threads(nthreads);
    const double Lz[3] = {-1, -1, -1};
    if (!L) L = Lz;
    std::vector<std::vector<int>> nb;
    neighbour_lists(n, pos, cutoff, L, nb, nthreads);
    const double c2 = cutoff * cutoff;
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total) num_threads(nthreads)
    for (int64_t p = 0; p < n; ++p) {
        const std::vector<int> &l = nb[(size_t)p];
        /* split the sorted list into N-(p) = {q < p} and N+(p) = {q > p} */
        const size_t split = (size_t)(std::lower_bound(l.begin(), l.end(), (int)p) - l.begin());
        double a0[3] = {0.0, 0.0, 0.0}, a1[3] = {0.0, 0.0, 0.0}, a2[3] = {0.0, 0.0, 0.0};
        double ci[3], cj[3], ck[3];
        const double *P = pos + 3 * p;
        uint64_t cnt = 0;
        /* role 0: p = i, (j,k) lexicographic over N+ */
        for (size_t a = split; a < l.size(); ++a) {
            const int j = l[a];
            for (size_t b = a + 1; b < l.size(); ++b) {
                const int k = l[b];
                if (!pair_ok(pos, j, k, cutoff, c2, L)) continue;
                at_contrib_xyz(P[0], P[1], P[2], pos[3 * j], pos[3 * j + 1], pos[3 * j + 2], pos[3 * k],
                               pos[3 * k + 1], pos[3 * k + 2], nu, ci, cj, ck);
                a0[0] -= ci[0];
                a0[1] -= ci[1];
                a0[2] -= ci[2];
                cnt++;
            }
        }
        /* role 1: p = j, (i,k) lexicographic, i in N-, k in N+ */
        for (size_t a = 0; a < split; ++a) {
            const int i = l[a];
            for (size_t b = split; b < l.size(); ++b) {
                const int k = l[b];
                if (!pair_ok(pos, i, k, cutoff, c2, L)) continue;
                at_contrib_xyz(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], P[0], P[1], P[2], pos[3 * k],
                               pos[3 * k + 1], pos[3 * k + 2], nu, ci, cj, ck);
                a1[0] -= cj[0];
                a1[1] -= cj[1];
                a1[2] -= cj[2];
            }
        }
        /* role 2: p = k, (i,j) lexicographic over N- */
        for (size_t a = 0; a < split; ++a) {
            const int i = l[a];
            for (size_t b = a + 1; b < split; ++b) {
                const int j = l[b];
                if (!pair_ok(pos, i, j, cutoff, c2, L)) continue;
                at_contrib_xyz(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], pos[3 * j], pos[3 * j + 1],
                               pos[3 * j + 2], P[0], P[1], P[2], nu, ci, cj, ck);
                a2[0] -= ck[0];
                a2[1] -= ck[1];
                a2[2] -= ck[2];
            }
        }
        for (int c = 0; c < 3; ++c) {
            f0[3 * p + c] = a0[c];
            f1[3 * p + c] = a1[c];
            f2[3 * p + c] = a2[c];
        }
        total += cnt;
    }
    return total;
}

void orc_sum_roles(int64_t n, const double *f0, const double *f1, const double *f2, double *f)
{
    for (int64_t i = 0; i < 3 * n; ++i) f[i] = f0[i] + (f1[i] + f2[i]); /* Algorithm.cpp:246-248 */
}

void orc_predictor(int64_t n, double *pos, const double *vel, const double *acc, double dt)
{
    const double h = dt * dt * 0.5; /* structs.hpp:45 */
    for (int64_t i = 0; i < n; ++i) {
        const double aX = acc[3 * i]; /* structs.hpp:44: acc(aX, aX, aX) */
        pos[3 * i + 0] = pos[3 * i + 0] + vel[3 * i + 0] * dt + aX * h;
        pos[3 * i + 1] = pos[3 * i + 1] + vel[3 * i + 1] * dt + aX * h;
        pos[3 * i + 2] = pos[3 * i + 2] + vel[3 * i + 2] * dt + aX * h;
    }
}

void orc_update(int64_t n, const double *f, const double *mass, double *vel, double *acc, double dt,
                const double *g)
{
    const double h = dt * 0.5; /* structs.hpp:75 */
    for (int64_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            const double dragAcc = f[3 * i + c] / mass[i]; /* 69 */
            const double newAcc = g[c] - dragAcc;         /* 71 */
            vel[3 * i + c] = vel[3 * i + c] + newAcc * h;  /* 75, 92-94 */
            acc[3 * i + c] = newAcc;                       /* 96-98 */
        }
    }
}

void orc_domain_size(int64_t n, const double *pos, double *L, double *lo)
{
    double mn[3], mx[3];
    for (int c = 0; c < 3; ++c) {
        mn[c] = 1.7976931348623157e308;   /* numeric_limits<double>::max() */
        mx[c] = -1.7976931348623157e308;  /* numeric_limits<double>::lowest() */
    }
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const double v = pos[3 * i + c];
            if (v < mn[c]) mn[c] = v;
            if (v > mx[c]) mx[c] = v;
        }
    for (int c = 0; c < 3; ++c) {
        const double a = mn[c] - 0.0001, b = mx[c] + 0.0001;
        L[c] = std::fabs(b - a);
        if (lo) lo[c] = a;
    }
}

int orc_simulate(int64_t n, double *pos, double *vel, double *acc, const double *mass, double *f, int algo,
                 int steps, double dt, const double *g, double cutoff, double nu, uint64_t *counts, int mode,
                 int nthreads, orc_step_cb step_cb, void *user)
{
    double L[3] = {-1, -1, -1};
    double cut = -1.0;
    if (algo == 1) {
        orc_domain_size(n, pos, L, nullptr);
        cut = cutoff;
    }
    std::vector<double> f0((size_t)(3 * n)), f1((size_t)(3 * n)), f2((size_t)(3 * n));
    for (int s = 0; s < steps; ++s) {
        orc_predictor(n, pos, vel, acc, dt);
        uint64_t c;
        if (mode == 0)
            c = orc_sweep_brute(n, pos, nu, cut, L, f0.data(), f1.data(), f2.data(), nullptr);
        else
            c = orc_sweep_owner(n, pos, nu, cut, L, f0.data(), f1.data(), f2.data(), nthreads);
        orc_sum_roles(n, f0.data(), f1.data(), f2.data(), f);
        orc_update(n, f, mass, vel, acc, dt, g);
        if (counts) counts[s] = c;
        if (step_cb) step_cb(s, user);
    }
    return 0;
}

/* The textbook energy sum over the accepted triplets; *scale, if non-NULL,
 * receives the sum over the same triplets of |nu| (1 + |3 cos cos cos|) / r^3,
 * the lengths of each term's two parts. */
static double energy_sum(int64_t n, const double *pos, double nu, double cutoff, const double *L, uint64_t *count,
                         double *scale, int nthreads)
{
    nthreads = resolve_threads(nthreads);
    const double Lz[3] = {-1, -1, -1};
    if (!L) L = Lz;
    std::vector<std::vector<int>> nb;
    neighbour_lists(n, pos, cutoff, L, nb, nthreads);
    const double c2 = cutoff * cutoff;
    double E = 0.0, Es = 0.0;
    uint64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : E, Es, cnt) num_threads(nthreads)
    for (int64_t p = 0; p < n; ++p) {
        const std::vector<int> &l = nb[(size_t)p];
        const size_t split = (size_t)(std::lower_bound(l.begin(), l.end(), (int)p) - l.begin());
        const double *P = pos + 3 * p;
        for (size_t a = split; a < l.size(); ++a) {
            const int j = l[a];
            for (size_t b = a + 1; b < l.size(); ++b) {
                const int k = l[b];
                if (!pair_ok(pos, j, k, cutoff, c2, L)) continue;
                const double *J = pos + 3 * j, *K = pos + 3 * k;
                /* edge vectors and the three interior angles */
                const double ij[3] = {J[0] - P[0], J[1] - P[1], J[2] - P[2]};
                const double ik[3] = {K[0] - P[0], K[1] - P[1], K[2] - P[2]};
                const double jk[3] = {K[0] - J[0], K[1] - J[1], K[2] - J[2]};
                const double rij = std::sqrt(ij[0] * ij[0] + ij[1] * ij[1] + ij[2] * ij[2]);
                const double rik = std::sqrt(ik[0] * ik[0] + ik[1] * ik[1] + ik[2] * ik[2]);
                const double rjk = std::sqrt(jk[0] * jk[0] + jk[1] * jk[1] + jk[2] * jk[2]);
                const double cosi = (ij[0] * ik[0] + ij[1] * ik[1] + ij[2] * ik[2]) / (rij * rik);
                const double cosj = (-ij[0] * jk[0] - ij[1] * jk[1] - ij[2] * jk[2]) / (rij * rjk);
                const double cosk = (ik[0] * jk[0] + ik[1] * jk[1] + ik[2] * jk[2]) / (rik * rjk);
                const double r = rij * rik * rjk;
                E += nu * (1.0 + 3.0 * cosi * cosj * cosk) / (r * r * r);
                Es += std::fabs(nu) * (1.0 + std::fabs(3.0 * cosi * cosj * cosk)) / (r * r * r);
                cnt++;
            }
        }
    }
    if (count) *count = cnt;
    if (scale) *scale = Es;
    return E;
}

double orc_energy(int64_t n, const double *pos, double nu, double cutoff, const double *L, uint64_t *count,
                  int nthreads)
{
    return energy_sum(n, pos, nu, cutoff, L, count, nullptr, nthreads);
}

double orc_energy_scale(int64_t n, const double *pos, double nu, double cutoff, const double *L, uint64_t *count,
                        int nthreads)
{
    double s = 0.0;
    energy_sum(n, pos, nu, cutoff, L, count, &s, nthreads);
    return s;
}

double orc_force_scale(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *scale,
                       int nthreads)
{
    /* The same three role streams per particle as orc_sweep_owner; every
     * contribution the reference subtracts from p's force (AxilrodTeller.cpp
     * 83-93) adds its Euclidean length to scale[p]. */
    nthreads = resolve_threads(nthreads);
    const double Lz[3] = {-1, -1, -1};
    if (!L) L = Lz;
    std::vector<std::vector<int>> nb;
    neighbour_lists(n, pos, cutoff, L, nb, nthreads);
    const double c2 = cutoff * cutoff;
    double top = 0.0;
#pragma omp parallel for schedule(dynamic, 1) reduction(max : top) num_threads(nthreads)
    for (int64_t p = 0; p < n; ++p) {
        const std::vector<int> &l = nb[(size_t)p];
        const size_t split = (size_t)(std::lower_bound(l.begin(), l.end(), (int)p) - l.begin());
        const double *P = pos + 3 * p;
        double ci[3], cj[3], ck[3];
        double s = 0.0;
        auto len3 = [](const double *c) { return std::sqrt((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]); };
        for (size_t a = 0; a < l.size(); ++a) {
            const int q = l[a];
            const double *Q = pos + 3 * q;
            for (size_t b = a + 1; b < l.size(); ++b) {
                const int r = l[b];
                if (!pair_ok(pos, q, r, cutoff, c2, L)) continue;
                const double *R = pos + 3 * r;
                if (a >= split) { /* p < q < r: p is i */
                    at_contrib_xyz(P[0], P[1], P[2], Q[0], Q[1], Q[2], R[0], R[1], R[2], nu, ci, cj, ck);
                    s += len3(ci);
                } else if (b >= split) { /* q < p < r: p is j */
                    at_contrib_xyz(Q[0], Q[1], Q[2], P[0], P[1], P[2], R[0], R[1], R[2], nu, ci, cj, ck);
                    s += len3(cj);
                } else { /* q < r < p: p is k */
                    at_contrib_xyz(Q[0], Q[1], Q[2], R[0], R[1], R[2], P[0], P[1], P[2], nu, ci, cj, ck);
                    s += len3(ck);
                }
            }
        }
        scale[p] = s;
        top = s > top ? s : top;
    }
    return top;
}

double orc_forces_fast(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f,
                       uint64_t *count, int nthreads)
{
    nthreads = resolve_threads(nthreads);
    const double Lz[3] = {-1, -1, -1};
    if (!L) L = Lz;
    std::vector<std::vector<int>> nb;
    neighbour_lists(n, pos, cutoff, L, nb, nthreads);
    const double c2 = cutoff * cutoff;
    const double K = 0.375 * nu;
    double E = 0.0;
    uint64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : E, cnt) num_threads(nthreads)
    for (int64_t p = 0; p < n; ++p) {
        const std::vector<int> &l = nb[(size_t)p];
        const double *P = pos + 3 * p;
        double g[3] = {0, 0, 0};
        for (size_t a = 0; a < l.size(); ++a) {
            const int j = l[a];
            const double *J = pos + 3 * j;
            const double dj[3] = {J[0] - P[0], J[1] - P[1], J[2] - P[2]};
            const double u = dj[0] * dj[0] + dj[1] * dj[1] + dj[2] * dj[2];
            for (size_t b = a + 1; b < l.size(); ++b) {
                const int k = l[b];
                if (!pair_ok(pos, j, k, cutoff, c2, L)) continue;
                const double *Kp = pos + 3 * k;
                const double dk[3] = {Kp[0] - P[0], Kp[1] - P[1], Kp[2] - P[2]};
                const double djk[3] = {Kp[0] - J[0], Kp[1] - J[1], Kp[2] - J[2]};
                const double v = dk[0] * dk[0] + dk[1] * dk[1] + dk[2] * dk[2];
                const double w = djk[0] * djk[0] + djk[1] * djk[1] + djk[2] * djk[2];
                /* E = nu s^3 + K A B C s^5, s = (uvw)^-1/2, A = u+v-w, B = u-v+w, C = -u+v+w */
                const double s = 1.0 / std::sqrt(u * v * w);
                const double s2 = s * s;
                const double A = u + v - w, B = u - v + w, C = -u + v + w;
                const double X = B * C, Y = A * C, Z = A * B;
                const double ABC = A * X;
                const double Q = ABC * s2;
                const double R = -1.5 * nu - 2.5 * K * Q;
                const double S = X + Y + Z;
                const double s5 = s2 * s2 * s;
                const double Eu = s5 * (R * v * w + K * (S - 2.0 * Z));
                const double Ev = s5 * (R * u * w + K * (S - 2.0 * Y));
                /* dE/dx_p = -2 Eu (xj - xp) - 2 Ev (xk - xp) */
                for (int c = 0; c < 3; ++c) g[c] += -2.0 * Eu * dj[c] - 2.0 * Ev * dk[c];
                if (p < j && p < k) {
                    E += s2 * s * (nu + K * Q);
                    cnt++;
                }
            }
        }
        f[3 * p] = g[0];
        f[3 * p + 1] = g[1];
        f[3 * p + 2] = g[2];
    }
    if (count) *count = cnt;
    return E;
}

double orc_sweep_cells_fast(int64_t n, const double *pos, double nu, double cutoff, const double *L, double *f,
                            uint64_t *count, int nthreads)
{
    /* A CPU implementation written for speed rather than for the reference's
     * order: cells of width >= cutoff (27-cell stencil, periodic wrap),
     * owner = smallest index, the (u,v,w) formulation with Newton's third
     * law, thread-private force arrays reduced at the end. */
    nthreads = resolve_threads(nthreads);
    const double c2 = cutoff * cutoff;
    const double K = 0.375 * nu;
    int nc[3];
    double lo[3];
    for (int d = 0; d < 3; ++d) {
        double mn = 1e300;
        for (int64_t i = 0; i < n; ++i) mn = std::min(mn, pos[3 * i + d]);
        lo[d] = n > 0 ? mn : 0.0;
        const double q = std::floor(L[d] / (cutoff * (1.0 + 1e-6)));
        nc[d] = q >= 3.0 ? (int)std::min(q, 512.0) : 1;
    }
    const int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    std::vector<int> cell_of((size_t)n), start((size_t)ncell + 1, 0), idx((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        int cc[3];
        for (int d = 0; d < 3; ++d) {
            double t = (pos[3 * i + d] - lo[d]) / L[d];
            t -= std::floor(t);
            cc[d] = std::min(nc[d] - 1, std::max(0, (int)(t * nc[d])));
        }
        cell_of[(size_t)i] = (cc[2] * nc[1] + cc[1]) * nc[0] + cc[0];
        start[(size_t)cell_of[(size_t)i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) start[(size_t)c + 1] += start[(size_t)c];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < n; ++i) idx[(size_t)fill[(size_t)cell_of[(size_t)i]]++] = (int)i;
    }
    std::vector<std::vector<double>> fpriv((size_t)nthreads, std::vector<double>((size_t)(3 * n), 0.0));
    double E = 0.0;
    uint64_t cnt = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : E, cnt)
    {
#ifdef _OPENMP
        std::vector<double> &fp = fpriv[(size_t)omp_get_thread_num()];
#else
        std::vector<double> &fp = fpriv[0];
#endif
        std::vector<int> l;
#pragma omp for schedule(dynamic, 16)
        for (int64_t p = 0; p < n; ++p) {
            const double *P = pos + 3 * p;
            l.clear();
            const int c = cell_of[(size_t)p];
            const int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / (nc[0] * nc[1]);
            const int wx = nc[0] >= 3 ? 1 : 0, wy = nc[1] >= 3 ? 1 : 0, wz = nc[2] >= 3 ? 1 : 0;
            for (int oz = -wz; oz <= wz; ++oz)
                for (int oy = -wy; oy <= wy; ++oy)
                    for (int ox = -wx; ox <= wx; ++ox) {
                        const int xx = (cx + ox + nc[0]) % nc[0], yy = (cy + oy + nc[1]) % nc[1];
                        const int zz = (cz + oz + nc[2]) % nc[2];
                        const int cc = (zz * nc[1] + yy) * nc[0] + xx;
                        for (int t = start[(size_t)cc]; t < start[(size_t)cc + 1]; ++t) {
                            const int q = idx[(size_t)t];
                            if (q <= p) continue;
                            if (sqr_dist_periodic(P[0], P[1], P[2], pos[3 * q], pos[3 * q + 1], pos[3 * q + 2], L) > c2)
                                continue;
                            l.push_back(q);
                        }
                    }
            double gp[3] = {0, 0, 0};
            for (size_t a = 0; a < l.size(); ++a) {
                const int j = l[a];
                const double *J = pos + 3 * j;
                const double dj[3] = {J[0] - P[0], J[1] - P[1], J[2] - P[2]};
                const double u = dj[0] * dj[0] + dj[1] * dj[1] + dj[2] * dj[2];
                double gj[3] = {0, 0, 0};
                for (size_t b = a + 1; b < l.size(); ++b) {
                    const int k = l[b];
                    const double *Kp = pos + 3 * k;
                    if (sqr_dist_periodic(J[0], J[1], J[2], Kp[0], Kp[1], Kp[2], L) > c2) continue;
                    const double dk[3] = {Kp[0] - P[0], Kp[1] - P[1], Kp[2] - P[2]};
                    const double e[3] = {Kp[0] - J[0], Kp[1] - J[1], Kp[2] - J[2]};
                    const double v = dk[0] * dk[0] + dk[1] * dk[1] + dk[2] * dk[2];
                    const double w = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
                    const double s = 1.0 / std::sqrt(u * v * w);
                    const double s2 = s * s;
                    const double A = u + v - w, B = u - v + w, C = -u + v + w;
                    const double X = B * C, Y = A * C, Z = A * B;
                    const double Q = A * X * s2;
                    const double R = -1.5 * nu - 2.5 * K * Q;
                    const double S = X + Y + Z;
                    const double c2s5 = 2.0 * s2 * s2 * s;
                    const double eu = c2s5 * (R * v * w + K * (S - 2.0 * Z));
                    const double ev = c2s5 * (R * u * w + K * (S - 2.0 * Y));
                    const double ew = c2s5 * (R * u * v + K * (S - 2.0 * X));
                    for (int d = 0; d < 3; ++d) {
                        gp[d] -= eu * dj[d] + ev * dk[d];
                        gj[d] += eu * dj[d] - ew * e[d];
                        fp[3 * (size_t)k + d] += ev * dk[d] + ew * e[d];
                    }
                    E += s2 * s * (nu + K * Q);
                    cnt++;
                }
                for (int d = 0; d < 3; ++d) fp[3 * (size_t)j + d] += gj[d];
            }
            for (int d = 0; d < 3; ++d) fp[3 * (size_t)p + d] += gp[d];
        }
    }
    for (int64_t i = 0; i < 3 * n; ++i) {
        double a = 0.0;
        for (int t = 0; t < nthreads; ++t) a += fpriv[(size_t)t][(size_t)i];
        f[i] = a;
    }
    if (count) *count = cnt;
    return E;
}

int orc_sweep_ranks(int algo, int ranks, int64_t n, const double *pos, double nu, double *f,
                    uint64_t *buffer_counts, uint64_t *particle_counts, int32_t *processed, int64_t cap,
                    int64_t *nprocessed)
{
    if ((algo != 0 && algo != 2) || ranks < 1 || n < 0) return -1;
    const int p = ranks;
    AxilrodTellerOracle at(nu);
    g_potential_slot = &at;
    /* copies[k][q]: rank q's particles in role buffer k (b0, b1, b2), each a
     * separate accumulator that travels with its buffer (SimulationStep:
     * b1 = b0, b2 = b0, NATA.cpp:87-99, AUTA.cpp:236-256) */
    std::vector<std::vector<Particle>> copies[3];
    std::vector<int64_t> off((size_t)p), cnt((size_t)p);
    for (int k = 0; k < 3; ++k) copies[k].resize((size_t)p);
    for (int q = 0; q < p; ++q) {
        copies[0][(size_t)q] = atom_share(n, pos, p, q, &off[(size_t)q], &cnt[(size_t)q]);
        copies[1][(size_t)q] = copies[0][(size_t)q];
        copies[2][(size_t)q] = copies[0][(size_t)q];
    }
    /* hold[k][r]: whose role-k buffer rank r holds; a shift to the right
     * neighbour (MPI_Sendrecv_replace, NATA.cpp:62-63, AUTA.cpp:30-31) hands
     * rank r the buffer rank r-1 held */
    std::vector<int> hold[3];
    for (int k = 0; k < 3; ++k) {
        hold[k].resize((size_t)p);
        for (int r = 0; r < p; ++r) hold[k][(size_t)r] = r;
    }
    auto shift = [&](int k) {
        std::vector<int> h = hold[k];
        for (int r = 0; r < p; ++r) hold[k][(size_t)r] = h[(size_t)ring_mod(r - 1, p)];
    };
    for (int r = 0; r < p; ++r) buffer_counts[r] = particle_counts[r] = 0;
    int64_t np = 0;
    auto record = [&](int r, int a, int b, int c) {
        if (processed && np < cap) {
            processed[4 * np] = r;
            processed[4 * np + 1] = a;
            processed[4 * np + 2] = b;
            processed[4 * np + 3] = c;
        }
        ++np;
    };
    if (algo == 0) {
        /* NATA::SimulationStep (NATA.cpp:84-196): b0 stays; b2 shifts after
         * every inner step, b1 after every outer one; a buffer triplet is
         * computed by the first rank (in rank order) whose triplet is new
         * among those of all ranks at that step (calculateProcessed, 27-52) */
        std::vector<std::array<int, 3>> done;
        int step = 0;
        for (int i = 0; i < p; ++i) {
            for (int j = 0; j < p; ++j) {
                std::vector<char> calc((size_t)p, 0);
                for (int x = 0; x < p; ++x) {
                    const int b1 = ring_mod(x - step / p, p), b2 = ring_mod(x - step, p);
                    bool seen = false;
                    for (const auto &t : done) seen = seen || same_triplet(t.data(), x, b1, b2);
                    if (!seen) {
                        done.push_back({x, b1, b2});
                        calc[(size_t)x] = 1;
                    }
                }
                for (int r = 0; r < p; ++r) {
                    if (!calc[(size_t)r]) continue;
                    const int o1 = ring_mod(r - step / p, p), o2 = ring_mod(r - step, p);
                    particle_counts[r] +=
                        sweep_aos_owned(copies[0][(size_t)r], copies[1][(size_t)hold[1][(size_t)r]],
                                        copies[2][(size_t)hold[2][(size_t)r]], r, o1, o2, 0, -1, g_potential_slot);
                    buffer_counts[r]++;
                    record(r, r, o1, o2);
                }
                if (p > 1) {
                    shift(2);
                    ++step;
                }
            }
            if (p > 1) shift(1);
        }
    } else {
        /* AUTA::SimulationStep (AUTA.cpp:236-383): all three buffers travel;
         * rounds s = p, p-3, ... shift buffer i (2, 0, 1, ...) before every
         * step but the very first, and every rank computes every step; with
         * 3 | p one more shift, and each rank computes a third of the rows
         * of the buffers sorted by owner (calculateOneThirdOfInteractions,
         * 48-119); the buffers then go home (sendBackParticles) */
        int i = 2;
        for (int s = p; s > 0; s -= 3) {
            for (int j = 0; j < s; ++j) {
                if (j != 0 || s != p) shift(i);
                for (int r = 0; r < p; ++r) {
                    const int o0 = hold[0][(size_t)r], o1 = hold[1][(size_t)r], o2 = hold[2][(size_t)r];
                    particle_counts[r] += sweep_aos_owned(copies[0][(size_t)o0], copies[1][(size_t)o1],
                                                          copies[2][(size_t)o2], o0, o1, o2, 0, -1,
                                                          g_potential_slot);
                    buffer_counts[r]++;
                    record(r, o0, o1, o2);
                }
            }
            i = (i + 1) % 3;
        }
        if (p % 3 == 0) {
            shift(i);
            for (int r = 0; r < p; ++r) {
                int o[3] = {hold[0][(size_t)r], hold[1][(size_t)r], hold[2][(size_t)r]};
                if (o[0] == o[1] || o[0] == o[2] || o[1] == o[2]) return -1; /* the reference sorts distinct owners only */
                int slot[3] = {0, 1, 2};
                std::sort(slot, slot + 3, [&](int a, int b) { return o[a] < o[b]; });
                std::vector<Particle> &s0 = copies[slot[0]][(size_t)o[slot[0]]];
                std::vector<Particle> &s1 = copies[slot[1]][(size_t)o[slot[1]]];
                std::vector<Particle> &s2 = copies[slot[2]][(size_t)o[slot[2]]];
                const int third = r / (p / 3);
                const int64_t sz = (int64_t)s0.size(), start = third * (sz / 3);
                const int64_t steps = third == 2 ? sz - 2 * (sz / 3) : sz / 3;
                particle_counts[r] += sweep_aos_owned(s0, s1, s2, o[slot[0]], o[slot[1]], o[slot[2]], start, steps,
                                                      g_potential_slot);
                buffer_counts[r]++;
                record(r, o[0], o[1], o[2]);
            }
        }
    }
    /* SumUpParticles on each rank's own buffers (Algorithm.cpp:245-249) */
    for (int q = 0; q < p; ++q)
        for (int64_t t = 0; t < cnt[(size_t)q]; ++t) {
            const Particle &a = copies[0][(size_t)q][(size_t)t], &b = copies[1][(size_t)q][(size_t)t],
                           &c = copies[2][(size_t)q][(size_t)t];
            double *fp = f + 3 * (off[(size_t)q] + t);
            fp[0] = a.fX + (b.fX + c.fX);
            fp[1] = a.fY + (b.fY + c.fY);
            fp[2] = a.fZ + (b.fZ + c.fZ);
        }
    if (nprocessed) *nprocessed = np;
    return 0;
}

int orc_sweep_ranks_p3bca(int ndims, const int *dims, int64_t n, const double *pos, double nu, double cutoff,
                          double *f, uint64_t *buffer_counts, uint64_t *particle_counts, int32_t *processed,
                          int64_t cap, int64_t *nprocessed)
{
    if (ndims < 1 || ndims > 3 || !dims || n < 0 || !(cutoff > 0)) return -1;
    long long p = 1;
    for (int k = 0; k < ndims; ++k) {
        if (dims[k] < 1) return -1;
        p *= dims[k];
    }
    if (p > 4096) return -1;
    P3bcaRanks g;
    if (p3bca_ranks_init(g, ndims, dims, n, pos, cutoff) != 0) return -2;
    return p3bca_ranks_sweep(g, n, pos, nu, cutoff, f, buffer_counts, particle_counts, processed, cap, nprocessed);
}

int orc_simulate_ranks_p3bca(int ndims, const int *dims, int64_t n, double *pos, double *vel, double *acc,
                             const double *mass, double *f, int steps, double dt, const double *g, double cutoff,
                             double nu, uint64_t *buffer_counts, uint64_t *particle_counts, orc_step_cb step_cb,
                             void *user)
{
    if (ndims < 1 || ndims > 3 || !dims || n < 0 || !(cutoff > 0)) return -1;
    long long p = 1;
    for (int k = 0; k < ndims; ++k) {
        if (dims[k] < 1) return -1;
        p *= dims[k];
    }
    if (p > 4096) return -1;
    /* Simulation::Init bins the initial particles; Start then runs predictor,
     * sweep and update (with the exchange) per step (Simulation.cpp:28-81) */
    P3bcaRanks grid;
    if (p3bca_ranks_init(grid, ndims, dims, n, pos, cutoff) != 0) return -2;
    for (int s = 0; s < steps; ++s) {
        orc_predictor(n, pos, vel, acc, dt);
        const int rc = p3bca_ranks_sweep(grid, n, pos, nu, cutoff, f, buffer_counts + (size_t)s * p,
                                         particle_counts + (size_t)s * p, nullptr, 0, nullptr);
        if (rc != 0) return rc;
        orc_update(n, f, mass, vel, acc, dt, g);
        p3bca_ranks_exchange(grid, pos);
        if (step_cb) step_cb(s, user);
    }
    return 0;
}

}  // extern "C"
