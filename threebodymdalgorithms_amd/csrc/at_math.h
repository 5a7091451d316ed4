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

// Eigen's Vector3d::norm() as sqrt((x*x + y*y) + z*z) (AxilrodTeller.cpp:29-31).
__device__ __forceinline__ double ref_norm3(double x, double y, double z)
{
    return __builtin_sqrt((x * x + y * y) + z * z);
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

// One dVdR bracket (AxilrodTeller.cpp:68-81) for own side x and the two other
// sides y, z in the order the source names them: dVdRa = (a; b, c),
// dVdRb = (b; a, c), dVdRc = (c; b, a).  Left-to-right sums and products.
__device__ __forceinline__ double ref_bracket(double nu, const RefPowers &x, const RefPowers &y,
                                              const RefPowers &z)
{
    double t = -8. / (x.d4 * y.d3 * z.d3);
    t = t - 1. / (y.d5 * z.d5);
    t = t + 5. * y.d1 / (x.d6 * z.d5);
    t = t + 5. * z.d1 / (x.d6 * y.d5);
    t = t - 1. / (x.d2 * y.d3 * z.d5);
    t = t - 1. / (x.d2 * y.d5 * z.d3);
    t = t - 3. / (x.d4 * y.d1 * z.d5);
    t = t - 3. / (x.d4 * y.d5 * z.d1);
    t = t - 5. / (x.d6 * y.d1 * z.d3);
    t = t - 5. / (x.d6 * y.d3 * z.d1);
    t = t + 6. / (x.d4 * y.d3 * z.d3);
    return (3. * nu / (8. * x.d1)) * t;
}

// Contribution that AxilrodTeller::CalculateForces subtracts from the force
// of the particle in role ROLE (0: i, 1: j, 2: k) of the triplet (i, j, k),
// AxilrodTeller.cpp:83-93.  Only the two brackets that role needs are
// evaluated; each is the same IEEE operation sequence as in the full call, so
// the result is bit-identical to the reference's.
template <int ROLE>
__device__ __forceinline__ void ref_contrib(double xi, double yi, double zi, double xj, double yj, double zj,
                                            double xk, double yk, double zk, double nu, double &cx, double &cy,
                                            double &cz)
{
    const RefPowers a = ref_powers(ref_norm3(xj - xi, yj - yi, zj - zi));  // a = rj - ri
    const RefPowers b = ref_powers(ref_norm3(xi - xk, yi - yk, zi - zk));  // b = ri - rk
    const RefPowers c = ref_powers(ref_norm3(xj - xk, yj - yk, zj - zk));  // c = rj - rk
    const double dXa = xj - xi, dYa = yj - yi, dZa = zj - zi;
    const double dXb = xk - xi, dYb = yk - yi, dZb = zk - zi;
    const double dXc = xk - xj, dYc = yk - yj, dZc = zk - zj;
    if constexpr (ROLE == 0) {
        const double A = ref_bracket(nu, a, b, c);
        const double B = ref_bracket(nu, b, a, c);
        cx = dXa * A + dXb * B;
        cy = dYa * A + dYb * B;
        cz = dZa * A + dZb * B;
    } else if constexpr (ROLE == 1) {
        const double A = ref_bracket(nu, a, b, c);
        const double C = ref_bracket(nu, c, b, a);
        cx = dXa * (-A) + dXc * C;
        cy = dYa * (-A) + dYc * C;
        cz = dZa * (-A) + dZc * C;
    } else {
        const double B = ref_bracket(nu, b, a, c);
        const double C = ref_bracket(nu, c, b, a);
        cx = dXb * (-B) + dXc * (-C);
        cy = dYb * (-B) + dYc * (-C);
        cz = dZb * (-B) + dZc * (-C);
    }
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
