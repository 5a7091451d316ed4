// at_math.h -- device-side Axilrod-Teller arithmetic for gfx950.
//
// Two formulations of the same potential:
//  * the reference's gradient expression (potential/AxilrodTeller.cpp:22-93),
//    evaluated operation for operation in IEEE FP64 (this translation unit is
//    compiled with -ffp-contract=off, so nothing is fused), used by the
//    bit-exact path;
//  * the squared-distance form used by the fast paths:
//      u = |xj-xi|^2, v = |xk-xi|^2, w = |xk-xj|^2, s = (u v w)^-1/2,
//      A = u+v-w, B = u-v+w, C = -u+v+w, K = 3 nu / 8,
//      E = s^3 (nu + K Q),  Q = A B C s^2,
//      dE/du = s^5 (R v w + K (S - 2 A B)), dE/dv = s^5 (R u w + K (S - 2 A C)),
//      dE/dw = s^5 (R u v + K (S - 2 B C)),
//      S = BC + AC + AB, R = -3 nu / 2 - 5 K Q / 2.
//    Fused multiply-adds are written explicitly with fma() so the rounding of
//    this form is fixed by the source, not by the compiler.
#pragma once

#include <hip/hip_runtime.h>

namespace tb3 {

// Utility::Particle::GetSqrDistPeriodic (utility/structs.hpp:106-115):
// per axis t = |b - a|, e = min(t, L - t) with std::min's (b < a) ? b : a,
// result ex*ex + ey*ey + ez*ez.  No wrapping of positions: for t > L the
// second operand is negative and squares to (t - L)^2, exactly as the reference.
__device__ __forceinline__ double sqr_dist_periodic(double ax, double ay, double az, double bx, double by,
                                                    double bz, double Lx, double Ly, double Lz)
{
    const double tx = __builtin_fabs(bx - ax);
    const double ty = __builtin_fabs(by - ay);
    const double tz = __builtin_fabs(bz - az);
    const double lx = Lx - tx, ly = Ly - ty, lz = Lz - tz;
    const double ex = (lx < tx) ? lx : tx;
    const double ey = (ly < ty) ? ly : ty;
    const double ez = (lz < tz) ? lz : tz;
    return ex * ex + ey * ey + ez * ez;
}

// The same value with one v_min_f64 per axis instead of a compare and two
// selects.  It is bit-identical to sqr_dist_periodic for every input: t is
// |b - a| >= +0 and L > 0, so t and L - t are never both zero (no +0/-0
// tie), t is NaN exactly when L - t is (both operands NaN), and t = +inf gives
// L - t = -inf, which both forms return.
__device__ __forceinline__ double sqr_dist_periodic_min(double ax, double ay, double az, double bx, double by,
                                                        double bz, double Lx, double Ly, double Lz)
{
    const double tx = __builtin_fabs(bx - ax);
    const double ty = __builtin_fabs(by - ay);
    const double tz = __builtin_fabs(bz - az);
    const double ex = __builtin_fmin(tx, Lx - tx);
    const double ey = __builtin_fmin(ty, Ly - ty);
    const double ez = __builtin_fmin(tz, Lz - tz);
    return ex * ex + ey * ey + ez * ez;
}

// Correctly rounded FP64 division and square root for operands in range.
//
// gfx950 has no FP64 divide or square-root instruction; the compiler expands
// a / b into a Newton-Raphson sequence (reciprocal estimate, two refinements
// y <- y + y (1 - b y), quotient q = a y, and the Markstein correction
// q + y (a - b q), each step one FMA) wrapped in three range steps:
// v_div_scale scales an operand by a power of two when the quotient or the
// reciprocal would leave the normal range or the numerator is near the
// bottom of it, v_div_fmas undoes that scaling in the last FMA, and
// v_div_fixup handles zeros, infinities, NaNs and overflow.  For finite
// operands far inside the normal range (every caller here keeps numerators,
// denominators and quotients within 2^-500 .. 2^500, or a zero numerator)
// the scale step leaves both operands as they are, the last FMA is a plain
// FMA, and the fixup returns the quotient unchanged.  div_rn_inrange is the
// sequence without the three range steps: on such operands it yields the
// bits of a / b in 8 instructions instead of 11.  The square root likewise:
// sqrt(x) rescales x when it is tiny and passes +-0 and +inf through around
// a Newton iteration on the reciprocal square-root estimate; for x within
// 2^-500 .. 2^500 those steps are the identity and sqrt_rn_inrange is the
// iteration alone (10 instructions instead of 17).  Both are checked bit for
// bit against the compiler's IEEE expansions on the device
// (tb3_selftest_fp64_inrange).
__device__ __forceinline__ double div_rn_inrange(double a, double b)
{
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    const double q = a * r;
    const double rem = __builtin_fma(-b, q, a);
    return __builtin_fma(rem, r, q);
}

__device__ __forceinline__ double sqrt_rn_inrange(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

struct RefPowers {
    double d1, d2, d3, d4, d5, d6;  // AxilrodTeller.cpp:33-51
};

__device__ __forceinline__ RefPowers ref_powers(double d1)
{
    RefPowers p;
    p.d1 = d1;
    p.d2 = p.d1 * d1;
    p.d3 = p.d2 * d1;
    p.d4 = p.d3 * d1;
    p.d5 = p.d4 * d1;
    p.d6 = p.d5 * d1;
    return p;
}

// IEEE division, or its in-range sequence when the caller has checked the
// operands' range.
template <bool INRANGE>
__device__ __forceinline__ double ref_div(double a, double b)
{
    if constexpr (INRANGE) return div_rn_inrange(a, b);
    else return a / b;
}

// One dVdR bracket (AxilrodTeller.cpp:68-81) for own side x and the two other
// sides y, z in the order the source names them: dVdRa = (a; b, c),
// dVdRb = (b; a, c), dVdRc = (c; b, a).  Left-to-right sums and products.
template <bool INRANGE = false>
__device__ __forceinline__ double ref_bracket(double nu, const RefPowers &x, const RefPowers &y,
                                              const RefPowers &z)
{
    double t = ref_div<INRANGE>(-8., x.d4 * y.d3 * z.d3);
    t = t - ref_div<INRANGE>(1., y.d5 * z.d5);
    t = t + ref_div<INRANGE>(5. * y.d1, x.d6 * z.d5);
    t = t + ref_div<INRANGE>(5. * z.d1, x.d6 * y.d5);
    t = t - ref_div<INRANGE>(1., x.d2 * y.d3 * z.d5);
    t = t - ref_div<INRANGE>(1., x.d2 * y.d5 * z.d3);
    t = t - ref_div<INRANGE>(3., x.d4 * y.d1 * z.d5);
    t = t - ref_div<INRANGE>(3., x.d4 * y.d5 * z.d1);
    t = t - ref_div<INRANGE>(5., x.d6 * y.d1 * z.d3);
    t = t - ref_div<INRANGE>(5., x.d6 * y.d3 * z.d1);
    t = t + ref_div<INRANGE>(6., x.d4 * y.d3 * z.d3);
    return ref_div<INRANGE>(3. * nu, 8. * x.d1) * t;
}

// The range the in-range sequences need, per triplet: every squared side in
// [2^-80, 2^80] puts the sides in [2^-40, 2^40], so every product of powers
// in the brackets (at most the 11th power of a side in total) lies within
// 2^-440 .. 2^440, every quotient within 2^-500 .. 2^500 and the square
// roots' arguments within 2^-80 .. 2^80; nu must be 0 or within 2^-200 ..
// 2^200 (nu_inrange, checked once per launch), which keeps 3 nu / (8 d1)
// within 2^-250 .. 2^250.
__device__ __forceinline__ bool sides_inrange(double sa, double sb, double sc)
{
    constexpr double lo = 0x1p-80, hi = 0x1p80;
    return sa >= lo && sa <= hi && sb >= lo && sb <= hi && sc >= lo && sc <= hi;
}

__host__ __device__ inline bool nu_inrange(double nu)
{
    const double m = nu < 0 ? -nu : nu;
    return nu == 0.0 || (m >= 0x1p-200 && m <= 0x1p200);
}

// The sides of the triplet (AxilrodTeller.cpp:23-31, 53-63): the difference
// vectors dXa = rj - ri, dXb = rk - ri, dXc = rk - rj and the squared norms
// of a = rj - ri, b = ri - rk, c = rj - rk.  Negating a difference is exact,
// so |b|^2 and |c|^2 from -dXb and -dXc are the reference's bits.
struct RefSides {
    double dXa, dYa, dZa, dXb, dYb, dZb, dXc, dYc, dZc;
    double sa, sb, sc;
};

__device__ __forceinline__ RefSides ref_sides(double xi, double yi, double zi, double xj, double yj, double zj,
                                              double xk, double yk, double zk)
{
    RefSides t;
    t.dXa = xj - xi, t.dYa = yj - yi, t.dZa = zj - zi;
    t.dXb = xk - xi, t.dYb = yk - yi, t.dZb = zk - zi;
    t.dXc = xk - xj, t.dYc = yk - yj, t.dZc = zk - zj;
    t.sa = (t.dXa * t.dXa + t.dYa * t.dYa) + t.dZa * t.dZa;
    t.sb = (t.dXb * t.dXb + t.dYb * t.dYb) + t.dZb * t.dZb;
    t.sc = (t.dXc * t.dXc + t.dYc * t.dYc) + t.dZc * t.dZc;
    return t;
}

template <bool INRANGE>
__device__ __forceinline__ double ref_sqrt(double x)
{
    if constexpr (INRANGE) return sqrt_rn_inrange(x);
    else return __builtin_sqrt(x);
}

// The brackets dVdRa, dVdRb, dVdRc that role ROLE needs (-1: all three).
template <int ROLE, bool INRANGE>
__device__ __forceinline__ void ref_brackets(const RefSides &t, double nu, double &A, double &B, double &C)
{
    const RefPowers a = ref_powers(ref_sqrt<INRANGE>(t.sa));
    const RefPowers b = ref_powers(ref_sqrt<INRANGE>(t.sb));
    const RefPowers c = ref_powers(ref_sqrt<INRANGE>(t.sc));
    A = B = C = 0.0;
    if constexpr (ROLE != 2) A = ref_bracket<INRANGE>(nu, a, b, c);
    if constexpr (ROLE != 1) B = ref_bracket<INRANGE>(nu, b, a, c);
    if constexpr (ROLE != 0) C = ref_bracket<INRANGE>(nu, c, b, a);
}

// The brackets from the in-range sequences where the triplet's sides and nu
// allow them (nu_ok = nu_inrange(nu), hoisted by the caller), else from the
// IEEE expansions: the same bits either way.
template <int ROLE>
__device__ __forceinline__ void ref_brackets_any(const RefSides &t, double nu, bool nu_ok, double &A, double &B,
                                                 double &C)
{
    if (nu_ok && sides_inrange(t.sa, t.sb, t.sc)) ref_brackets<ROLE, true>(t, nu, A, B, C);
    else ref_brackets<ROLE, false>(t, nu, A, B, C);
}

// Contribution that AxilrodTeller::CalculateForces subtracts from the force
// of the particle in role ROLE (0: i, 1: j, 2: k) of the triplet (i, j, k),
// AxilrodTeller.cpp:83-93.  Only the two brackets that role needs are
// evaluated; each is the same IEEE operation sequence as in the full call, so
// the result is bit-identical to the reference's.
template <int ROLE>
__device__ __forceinline__ void ref_contrib(double xi, double yi, double zi, double xj, double yj, double zj,
                                            double xk, double yk, double zk, double nu, bool nu_ok, double &cx,
                                            double &cy, double &cz)
{
    const RefSides t = ref_sides(xi, yi, zi, xj, yj, zj, xk, yk, zk);
    double A, B, C;
    ref_brackets_any<ROLE>(t, nu, nu_ok, A, B, C);
    if constexpr (ROLE == 0) {
        cx = t.dXa * A + t.dXb * B;
        cy = t.dYa * A + t.dYb * B;
        cz = t.dZa * A + t.dZb * B;
    } else if constexpr (ROLE == 1) {
        cx = t.dXa * (-A) + t.dXc * C;
        cy = t.dYa * (-A) + t.dYc * C;
        cz = t.dZa * (-A) + t.dZc * C;
    } else {
        cx = t.dXb * (-B) + t.dXc * (-C);
        cy = t.dYb * (-B) + t.dYc * (-C);
        cz = t.dZb * (-B) + t.dZc * (-C);
    }
}

// All three contributions of one triplet from one evaluation of the three
// brackets (AxilrodTeller.cpp:65-93): ci, cj, ck as the reference subtracts
// them from i.f, j.f and k.f.
__device__ __forceinline__ void ref_contrib_all(double xi, double yi, double zi, double xj, double yj, double zj,
                                                double xk, double yk, double zk, double nu, bool nu_ok, double *ci,
                                                double *cj, double *ck)
{
    const RefSides t = ref_sides(xi, yi, zi, xj, yj, zj, xk, yk, zk);
    double A, B, C;
    ref_brackets_any<-1>(t, nu, nu_ok, A, B, C);
    ci[0] = t.dXa * A + t.dXb * B;
    ci[1] = t.dYa * A + t.dYb * B;
    ci[2] = t.dZa * A + t.dZb * B;
    cj[0] = t.dXa * (-A) + t.dXc * C;
    cj[1] = t.dYa * (-A) + t.dYc * C;
    cj[2] = t.dZa * (-A) + t.dZc * C;
    ck[0] = t.dXb * (-B) + t.dXc * (-C);
    ck[1] = t.dYb * (-B) + t.dYc * (-C);
    ck[2] = t.dZb * (-B) + t.dZc * (-C);
}

// ---------------------------------------------------------------------------
// Fast formulation.

// 1/sqrt(x) in FP64 from the hardware estimate.
__device__ __forceinline__ double rsqrt_f64(double x)
{
    // v_rsq_f64 is good to ~2^-24 (checked on device); one third-order step
    // y (1 + r/2 + 3 r^2/8) with r = 1 - x y^2 leaves an error of order r^3,
    // below the rounding of the FP64 result, in five operations.
    const double y = __builtin_amdgcn_rsq(x);
    const double r = __builtin_fma(-(x * y), y, 1.0);
    const double t = __builtin_fma(0.375, r, 0.5);
    return __builtin_fma(y * r, t, y);
}

// 1/sqrt(x) in FP32: hardware estimate refined by one Newton step.
__device__ __forceinline__ float rsqrt_f32(float x)
{
    float y = __builtin_amdgcn_rsqf(x);
    const float r = __builtin_fmaf(-(x * y), y, 1.0f);
    return __builtin_fmaf(0.5f * y, r, y);
}

template <class T>
__device__ __forceinline__ T fma_t(T a, T b, T c);
template <>
__device__ __forceinline__ double fma_t<double>(double a, double b, double c)
{
    return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fma_t<float>(float a, float b, float c)
{
    return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double rsqrt_t(double x) { return rsqrt_f64(x); }
__device__ __forceinline__ float rsqrt_t(float x) { return rsqrt_f32(x); }

// Mixed-precision form: u, v, w arrive in FP64 and A = u+v-w, B = u-v+w,
// C = -u+v+w are formed in FP64 before rounding to FP32, because for
// elongated triangles (two long sides of nearly equal length) the difference
// cancels and FP32 inputs would leave only a few correct bits; everything
// after that is FP32, and the coefficients come back in FP64.
//
// A triplet with one short side (a close pair) gives its third particle two
// huge, nearly opposite terms, -(eu dij + ev dik) for particle i when w is
// the short side: the FP32 rounding of eu and ev alone would leave that
// gradient with ~1e-6 relative error.  Two exact identities avoid this:
// X + Y = 2uC, X + Z = 2vB and Y + Z = 2wA give S = X + Y + Z and the third
// particle's coefficient sum (eu + ev for w, ev + ew for u, eu + ew for v)
// without cancellation; the coefficient paired with the sum is then rebuilt
// as sum - partner in FP64 (exact for FP32 operands), so that the third
// particle sees sum * (its vector) plus partner * (the short side).
__device__ __forceinline__ void at_coeffs_mixed(double u64, double v64, double w64, float nu, float K, double &cu,
                                                double &cv, double &cw, float &e, bool need_w)
{
    const float u = (float)u64, v = (float)v64, w = (float)w64;
    const float A = (float)((u64 + v64) - w64);
    const float B = (float)((u64 - v64) + w64);
    const float C = (float)((v64 - u64) + w64);
    const float uv = u * v;
    const float s = rsqrt_f32(uv * w);
    const float s2 = s * s;
    const float X = B * C, Y = A * C, Z = A * B;
    const float Q = (A * X) * s2;
    const float R = __builtin_fmaf(-2.5f * K, Q, -1.5f * nu);
    // the shortest side m, the sum o of the other two, the factor P of 2 m P
    // in S, and the product G that completes S = G + 2 m P
    const bool wmin = w64 <= u64 && w64 <= v64;
    const bool umin = !wmin && u64 <= v64;
    const float m = wmin ? w : (umin ? u : v);
    const float o = wmin ? u + v : (umin ? v + w : w + u);
    const float P = wmin ? A : (umin ? C : B);
    const float G = wmin ? X : (umin ? Z : Y);
    const float S = __builtin_fmaf(2.0f * m, P, G);
    const float KS = K * S;
    const float m2K = -2.0f * K;
    const float c2s5 = 2.0f * (s2 * s2 * s);
    const double du = (double)(c2s5 * __builtin_fmaf(R, v * w, __builtin_fmaf(m2K, Z, KS)));
    const double dv = (double)(c2s5 * __builtin_fmaf(R, u * w, __builtin_fmaf(m2K, Y, KS)));
    e = (s2 * s) * __builtin_fmaf(K, Q, nu);
    if (need_w) {
        const double dw = (double)(c2s5 * __builtin_fmaf(R, uv, __builtin_fmaf(m2K, X, KS)));
        const double sum = (double)(c2s5 * __builtin_fmaf(R, m * o, __builtin_fmaf(m2K, 2.0f * m * P, 2.0f * KS)));
        // w shortest, third particle i: -(cu dij + cv dik), rebuild cv;
        // u shortest, k: cv dik + cw djk, rebuild cw; v shortest, j:
        // cu dij - cw djk with djk ~ -dij, rebuild cw
        const double rebuilt = sum - (umin ? dv : du);
        cu = du;
        cv = wmin ? rebuilt : dv;
        cw = wmin ? dw : rebuilt;
    } else {
        cu = du;
        cv = dv;
        cw = 0.0;
    }
}

// Dispatch on the per-triplet type: FP64 throughout, or the mixed form; the
// gradient coefficients come back in FP64 either way.
template <class T, bool NEED_W>
__device__ __forceinline__ void at_coeffs_from64(double u, double v, double w, T nu, T K, double &cu, double &cv,
                                                 double &cw, T &e);

// Gradient coefficients of one triplet in the squared-distance form.
// Returns eu = 2 dE/du, ev = 2 dE/dv, ew = 2 dE/dw and the energy e, so that
//   dE/dxi = -(eu dij + ev dik), dE/dxj = eu dij - ew djk, dE/dxk = ev dik + ew djk
// with dij = xj - xi, dik = xk - xi, djk = xk - xj.
template <class T, bool NEED_W>
__device__ __forceinline__ void at_coeffs(T u, T v, T w, T nu, T K, T &eu, T &ev, T &ew, T &e)
{
    const T uv = u * v;
    const T s = rsqrt_t(uv * w);
    const T s2 = s * s;
    const T A = (u + v) - w;
    const T B = (u - v) + w;
    const T C = (v - u) + w;
    const T X = B * C, Y = A * C, Z = A * B;
    const T Q = (A * X) * s2;
    const T R = fma_t<T>(T(-2.5) * K, Q, T(-1.5) * nu);
    const T S = (X + Y) + Z;
    const T KS = K * S;
    const T m2K = T(-2) * K;
    const T c2s5 = T(2) * (s2 * s2 * s);
    eu = c2s5 * fma_t<T>(R, v * w, fma_t<T>(m2K, Z, KS));
    ev = c2s5 * fma_t<T>(R, u * w, fma_t<T>(m2K, Y, KS));
    if constexpr (NEED_W) ew = c2s5 * fma_t<T>(R, uv, fma_t<T>(m2K, X, KS));
    else ew = T(0);
    e = (s2 * s) * fma_t<T>(K, Q, nu);
}

template <class T, bool NEED_W>
__device__ __forceinline__ void at_coeffs_from64(double u, double v, double w, T nu, T K, double &cu, double &cv,
                                                 double &cw, T &e)
{
    if constexpr (sizeof(T) == 8) at_coeffs<double, NEED_W>(u, v, w, nu, K, cu, cv, cw, e);
    else at_coeffs_mixed(u, v, w, nu, K, cu, cv, cw, e, NEED_W);
}

}  // namespace tb3
