// Per-cell and per-face numerics of the MUSCL-Hancock sweep, shared by the x- and y-sweep
// kernels (and compilable for the host, which the CPU unit tests of the line scanner use).
//
// Every expression keeps the reference's left-to-right C++ evaluation order so that, built
// with -ffp-contract=off, the results of the exact arithmetic (template argument F = false, the
// default) are bitwise those of the reference arithmetic. The fast arithmetic (F = true) differs
// in ddiv and dsqrt and, in its `if constexpr (F)` branches, skips work that only reproduces the
// exact arithmetic's roundings (each such shortcut changes values by rounding only, and a
// positivity decision only where its quantity is within rounding of the threshold). Reference:
//   EOS            E/headers/stiffened_gas_eos.hpp:57-70, A/headers/mixturemodel.hpp:18-52
//   state algebra  E/headers/misc_euler.hpp:20-131, A/headers/misc.hpp:31-266
//   limiter        E/headers/flux_solver_MUSCLhancock.hpp:67-89 (+ van Leer, not in the reference)
//   face           E/headers/flux_solver_MUSCLhancock.hpp:128-227, A/headers/flux_solver_MUSCLhancock.hpp:272-376
//   HLLC           E/headers/riemann_solver_HLLC.hpp:50-145, A/headers/riemann_solver_HLLC.hpp:49-150
//   z update       A/headers/zupdate.hpp:58-73
// std::min(a,b) is (b<a)?b:a and std::max(a,b) is (a<b)?b:a; the exact arithmetic never uses
// fmin/fmax, which differ on signed zeros and NaN.
//
// Restructuring relative to the reference (values unchanged): the reference evaluates each
// face from a 4-cell stencil, so each cell's primitives, slope, half-step states and their
// positivity tests are computed four or two times. Here they are computed once per cell
// ("cell stage") and each face combines the right-hand state of its left cell with the
// left-hand state of its right cell ("face stage"). The per-face slope fallback of the
// reference (both slopes of a face are zeroed when either extrapolation is unphysical) is
// reproduced by re-evaluating the two unsloped states for that face only.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define CFD_HD __host__ __device__ __forceinline__
#else
#define CFD_HD inline
#endif

namespace cfd {

enum { EULER = 0, ALLAIRE = 1 };

// FP64 division and square root. F = false (the exact arithmetic, the default): correctly
// rounded. On the device these are the steps of the compiler's IEEE sequences without their
// operand scaling: v_rcp_f64 refined by two Newton steps, the product, one residual correction
// and v_div_fixup for a / b; v_rsq_f64 refined by Goldschmidt steps for sqrt. The scaling only
// acts on operands the solver's states never reach (a divide whose numerator is below 2^-969 in
// magnitude, a subnormal divisor or quotient, an exponent gap of 768 or more; a square root of a
// number below 2^-767), so every quotient and root equals the IEEE one; v_div_fixup and the class
// test keep zeros, infinities and NaNs exact.
// F = true (the fast arithmetic, device only): a * r with r = v_rcp_f64 refined by one Newton
// step, and one Goldschmidt step after v_rsq_f64 -- within a few ulp of the correctly rounded
// results, at 1 instead of 4 instructions per numerator once a divisor's reciprocal is formed.
// The host always computes exactly.
template <bool F = false> CFD_HD double ddiv(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    if constexpr (F) {
        return a * r;
    } else {
        e = __builtin_fma(-b, r, 1.0);
        r = __builtin_fma(r, e, r);
        double q = a * r;
        const double t = __builtin_fma(-b, q, a);
        q = __builtin_fma(t, r, q);
        return __builtin_amdgcn_div_fixup(q, b, a);
    }
#else
    return a / b;
#endif
}

template <bool F = false> CFD_HD double dsqrt(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;    // sqrt(x)
    double h = y * 0.5;  // 1 / (2 sqrt(x))
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    if constexpr (!F) {
        h = __builtin_fma(h, r, h);
        double d = __builtin_fma(-g, g, x);
        g = __builtin_fma(d, h, g);
        d = __builtin_fma(-g, g, x);
        g = __builtin_fma(d, h, g);
    }
    // +0, -0 and +inf are their own roots (class mask: -0, +0, +inf)
    return __builtin_amdgcn_class(x, 0x260) ? x : g;
#else
    return sqrt(x);
#endif
}

// sqrt(x / rho), the form of every sound and shock speed here. F = false: dsqrt(ddiv(x, rho)).
// F = true (device): x * y with y = v_rsq_f64(x rho) refined by one Newton step,
// x y = sqrt(x/rho) (x >= 0, rho > 0) -- one reciprocal square root in place of a reciprocal and
// a square root. x = +0 and x = +inf give themselves, as the root of x / rho does for rho > 0.
template <bool F = false> CFD_HD double sqrt_ratio(double x, double rho)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (F) {
        const double a = x * rho;
        const double y = __builtin_amdgcn_rsq(a);
        const double r = __builtin_fma(-(a * y), y, 1.0);  // 1 - a y^2
        const double c = x * y;
        const double cr = __builtin_fma(c * 0.5, r, c);
        return __builtin_amdgcn_class(x, 0x260) ? x : cr;
    }
#endif
    return dsqrt<F>(ddiv<F>(x, rho));
}

template <int M> struct Model;
template <> struct Model<EULER> {
    static constexpr int NC = 4;
    static constexpr int MN = 1;  // normal momentum component (x frame)
    static constexpr int MT = 2;  // tangential momentum component
};
template <> struct Model<ALLAIRE> {
    static constexpr int NC = 6;
    static constexpr int MN = 2;
    static constexpr int MT = 3;
};

// Material constants. g1m1 = gamma1 - 1.0 and g1pinf1 = gamma1 * pinf1 are exactly the
// sub-expressions the reference recomputes on every call. kmin and epsmin are lower bounds of
// the Allaire mixture k and eps over z in [0, 1], used only by the conservative positivity
// screens below (never in a value that is stored or compared bitwise).
// The fast arithmetic evaluates the mixture's eps(z) = z e1 + (1-z) e2 as e2 + z (e1 - e2) and
// k(z) eps(z) = z q1 + (1-z) q2 as q2 + z (q1 - q2), one fma each, with ei = 1/(gi - 1) and
// qi = gi pinfi ei held as fe2 = e2, fe12 = e1 - e2, fk2 = q2, fk12 = q1 - q2. Its Euler wave
// speeds take gamma (p + pinf) = wa w + wb from w = E - (mu^2 + mv^2)/(2 rho), with
// wa = gamma (gamma - 1) and wb = gamma (pinf - gamma pinf).
struct Material {
    double g1, pinf1, g2, pinf2;
    double g1m1, g2m1, g1pinf1;
    double kmin, epsmin;
    double fe2, fe12, fk2, fk12;
    double wa, wb;
};

CFD_HD Material make_material(double g1, double pinf1, double g2, double pinf2)
{
    Material m;
    m.g1 = g1; m.pinf1 = pinf1; m.g2 = g2; m.pinf2 = pinf2;
    m.g1m1 = g1 - 1.0; m.g2m1 = g2 - 1.0; m.g1pinf1 = g1 * pinf1;
    double k1 = g1 * pinf1, k2 = g2 * pinf2;
    m.kmin = k1 < k2 ? k1 : k2;
    double e1 = 1.0 / (g1 - 1.0), e2 = 1.0 / (g2 - 1.0);
    m.epsmin = (e1 < e2 ? e1 : e2) * (1.0 - 1e-12);
    m.fe2 = e2;
    m.fe12 = e1 - e2;
    m.fk2 = k2 * e2;
    m.fk12 = k1 * e1 - k2 * e2;
    m.wa = g1 * (g1 - 1.0);
    m.wb = g1 * (pinf1 - g1 * pinf1);
    return m;
}

struct SchemeOpts {
    int limiter;      // 0 minmod (reference), 1 van Leer
    int wave_speeds;  // 0 Garrick-Owkes-Regele (reference), 1 PVRS (Allaire)
    double z_lo, z_hi;
    int godunov;      // 1: first-order Godunov faces and upwind z update instead of MUSCL-Hancock
    int fast;         // 1: the sweeps' fast arithmetic (ddiv<true>, dsqrt<true>)
};

CFD_HD double smin(double a, double b) { return (b < a) ? b : a; }
CFD_HD double smax(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------- EOS

// Mixture eps and k (A/headers/mixturemodel.hpp:20-22), recomputed per call there; here per z.
struct Mix { double eps, k; };
template <bool F = false> CFD_HD Mix mixture(const Material& m, double z)
{
    Mix r;
    if constexpr (F) {
        r.eps = __builtin_fma(z, m.fe12, m.fe2);
        r.k = ddiv<F>(__builtin_fma(z, m.fk12, m.fk2), r.eps);
        return r;
    }
    r.eps = ddiv<F>(z, m.g1m1) + ddiv<F>(1.0 - z, m.g2m1);
    r.k = ddiv<F>(ddiv<F>(z * m.g1 * m.pinf1, m.g1m1) + ddiv<F>((1.0 - z) * m.g2 * m.pinf2, m.g2m1), r.eps);
    return r;
}

template <int M> struct Eos;
template <> struct Eos<EULER> {
    // E/headers/stiffened_gas_eos.hpp:57-70
    CFD_HD static double pressure(const Material& m, double rho, double e) { return m.g1m1 * rho * e - m.g1pinf1; }
    template <bool F = false> CFD_HD static double specific_ie(const Material& m, double rho, double p) { return ddiv<F>(p + m.g1pinf1, m.g1m1 * rho); }
    template <bool F = false> CFD_HD static double soundspeed(const Material& m, double rho, double p) { return sqrt_ratio<F>(m.g1 * (p + m.pinf1), rho); }
};
struct MixEos {
    // A/headers/mixturemodel.hpp:18-52
    template <bool F = false> CFD_HD static double pressure(const Mix& x, double rho, double e) { return ddiv<F>(rho * e, x.eps) - x.k; }
    template <bool F = false> CFD_HD static double specific_ie(const Mix& x, double rho, double p) { return ddiv<F>((p + x.k) * x.eps, rho); }
    template <bool F = false> CFD_HD static double soundspeed(const Mix& x, double rho, double p) { return sqrt_ratio<F>(x.k + p * (ddiv<F>(1.0, x.eps) + 1.0), rho); }
    template <bool F = false> CFD_HD static double shock_jump(const Mix& x, double p, double rho, double pstar)
    {
        double p_avg = 0.5 * (pstar + p);
        return sqrt_ratio<F>(pstar + x.k + ddiv<F>(p_avg, x.eps), rho);
    }
};

// ---------------------------------------------------------------- state algebra

// Conserved -> primitive (E/headers/misc_euler.hpp:63-78, A/headers/misc.hpp:130-147).
// Also returns the sound speed, which the Jacobian of the same state needs
// (E/headers/misc_euler.hpp:110, A/headers/misc.hpp:198 evaluate it from the same inputs); the
// fast arithmetic returns rho c^2 in its place, the Jacobian's only use of it, without the root.
template <int M> struct Prim;
template <> struct Prim<EULER> {
    double W[4];
    double c;
    template <bool F = false> CFD_HD void from_cons(const Material& m, const double* U)
    {
        double rho = U[0];
        double u = ddiv<F>(U[1], rho);
        double v = ddiv<F>(U[2], rho);
        double e = ddiv<F>(U[3], rho) - 0.5 * (u * u + v * v);
        double p = Eos<EULER>::pressure(m, rho, e);
        W[0] = rho; W[1] = u; W[2] = v; W[3] = p;
        if constexpr (F) c = m.g1 * (p + m.pinf1);  // rho c^2
        else c = Eos<EULER>::soundspeed<F>(m, rho, p);
    }
};
template <> struct Prim<ALLAIRE> {
    double W[6];
    double c;
    template <bool F = false> CFD_HD void from_cons(const Material& m, const double* U)
    {
        double rho = U[0] + U[1];
        double u = ddiv<F>(U[2], rho);
        double v = ddiv<F>(U[3], rho);
        double e = ddiv<F>(U[4], rho) - 0.5 * (u * u + v * v);
        Mix x = mixture<F>(m, U[5]);
        double p = MixEos::pressure<F>(x, rho, e);
        W[0] = U[0]; W[1] = U[1]; W[2] = u; W[3] = v; W[4] = p; W[5] = U[5];
        if constexpr (F) c = x.k + p * (ddiv<F>(1.0, x.eps) + 1.0);  // rho c^2
        else c = MixEos::soundspeed<F>(x, rho, p);
    }
};

// Primitive -> conserved (E/headers/misc_euler.hpp:82-97, A/headers/misc.hpp:149-163).
template <bool F = false> CFD_HD void p2c(const Material& m, const double* W, double* U, Model<EULER>*)
{
    double rho = W[0], u = W[1], v = W[2], p = W[3];
    double e = Eos<EULER>::specific_ie<F>(m, rho, p);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * e + 0.5 * rho * (u * u + v * v);
}
template <bool F = false> CFD_HD void p2c(const Material& m, const double* W, double* U, Model<ALLAIRE>*)
{
    double rho = W[0] + W[1];
    Mix x = mixture<F>(m, W[5]);
    double e = MixEos::specific_ie<F>(x, rho, W[4]);
    U[0] = W[0];
    U[1] = W[1];
    U[2] = rho * W[2];
    U[3] = rho * W[3];
    U[4] = rho * e + 0.5 * rho * (W[2] * W[2] + W[3] * W[3]);
    U[5] = W[5];
}
template <int M, bool F = false> CFD_HD void p2c(const Material& m, const double* W, double* U)
{
    p2c<F>(m, W, U, (Model<M>*)nullptr);
}

// Allaire snapping of the non-const is_physical_state overload (A/headers/misc.hpp:253-256).
CFD_HD void snap(double* U)
{
    if (fabs(U[0]) < 1e-12) U[0] = 0.0;
    if (fabs(U[1]) < 1e-12) U[1] = 0.0;
    if (fabs(U[5]) < 1e-12) U[5] = 0.0;
    if (fabs(U[5] - 1.0) < 1e-12) U[5] = 1.0;
}

// Physical-state test (E/headers/misc_euler.hpp:123-131, A/headers/misc.hpp:228-237). Euler
// has no snapping; for Allaire the caller snaps first where the reference does.
template <bool F = false> CFD_HD bool is_physical(const double* U, Model<EULER>*)
{
    double rho = U[0];
    double u = ddiv<F>(U[1], rho);
    double v = ddiv<F>(U[2], rho);
    double e = ddiv<F>(U[3], rho) - 0.5 * (u * u + v * v);
    return rho >= 0.0 && e >= 0.0;
}
template <bool F = false> CFD_HD bool is_physical(const double* U, Model<ALLAIRE>*)
{
    double rho = U[0] + U[1];
    double u = ddiv<F>(U[2], rho);
    double v = ddiv<F>(U[3], rho);
    double e = ddiv<F>(U[4], rho) - 0.5 * (u * u + v * v);
    double z = U[5];
    return U[0] >= 0.0 && U[1] >= 0.0 && e >= 0.0 && z >= 0.0 && z <= 1.0;
}
template <int M, bool F = false> CFD_HD bool is_physical(const double* U) { return is_physical<F>(U, (Model<M>*)nullptr); }

// Division-free screens for the physical-state tests. Each returns true only when the exact
// test (with its divisions and roundings) is certain to return true; otherwise the caller runs
// the exact test, so the boolean is always the reference's.
//
// Exact test on U = p2c(T): rho >= 0 and fl(E/rho) - fl(0.5(u''^2+v''^2)) >= 0, where
// E = rho e_T + 0.5 rho (u^2+v^2), e_T = (p+K)/(G rho) and u'' = fl(fl(rho u)/rho). With every
// value in the normal range each side carries a relative error of at most ~8 ulp, so the
// computed difference is >= 0 whenever e_T >= 16 eps K_kin (K_kin = 0.5(u^2+v^2)), i.e. whenever
// (p+K) >= 16 eps G rho K_kin. The screen demands (p+K) > 1e-12 G rho (u^2+v^2), a margin of
// ~10^3, and bounds every magnitude so no intermediate leaves the normal range.
CFD_HD bool prim_clearly_physical(double rho, double u, double v, double p, double K, double G)
{
    // K = gamma pinf (Euler) or a lower bound of the mixture k (Allaire); G = gamma - 1 (Euler)
    // or 1/epsmin (Allaire). The relative margin on p + K keeps its rounding (and that of the
    // mixture k) far below p + K itself.
    double w = p + K;
    double q = u * u + v * v;
    return rho > 1e-100 && rho < 1e100 && q < 1e100 && w > 1e-100 * rho && w < 1e100 * rho &&
           w > 1e-9 * (fabs(p) + K) && w > 1e-12 * G * rho * q;
}

// Exact test on a conserved state with density rho, energy E and momenta mu, mv:
// fl(E/rho) - fl(0.5((mu/rho)^2 + (mv/rho)^2)) >= 0 holds whenever
// E rho - 0.5(mu^2 + mv^2) >= 6 eps E rho; the computed difference q has an absolute error
// below 6 eps E rho, so q >= 1e-12 fl(E rho) is sufficient.
CFD_HD bool cons_clearly_physical(double rho, double mu, double mv, double E)
{
    double er = E * rho;
    double q = er - 0.5 * (mu * mu + mv * mv);
    return rho > 1e-100 && rho < 1e100 && er < 1e200 && fabs(mu) < 1e100 && fabs(mv) < 1e100 &&
           er > 1e-200 && q >= 1e-12 * er;
}

// Allaire post-update test (after snapping): a division-free screen, then the exact test.
// The fast arithmetic decides e >= 0 as E rho - (mu^2 + mv^2)/2 >= 0 with rho > 0.
template <bool F = false> CFD_HD bool allaire_physical_screened(const double* U)
{
    if constexpr (F) {
        const double rho = U[0] + U[1];
        if (U[0] >= 0.0 && U[1] >= 0.0 && U[5] >= 0.0 && U[5] <= 1.0 && rho > 0.0 &&
            U[4] * rho - 0.5 * (U[2] * U[2] + U[3] * U[3]) >= 0.0)
            return true;
    } else {
        if (U[0] >= 0.0 && U[1] >= 0.0 && U[5] >= 0.0 && U[5] <= 1.0 &&
            cons_clearly_physical(U[0] + U[1], U[2], U[3], U[4]))
            return true;
    }
    return is_physical<ALLAIRE, F>(U);
}

// ---------------------------------------------------------------- limiter

// The fast arithmetic forms minmod as max(0, min(a, b)) + min(0, max(a, b)) (one term is zero:
// the same value, without the data-dependent selects; a zero may carry the other sign) and van
// Leer's 2ab/(a+b) from the product a*b its sign test forms.
template <bool F = false> CFD_HD double limit(int limiter, double a, double b)
{
    if constexpr (F) {
        if (limiter == 1) {
            const double t = a * b;
            return t > 0.0 ? ddiv<F>(t + t, a + b) : 0.0;
        }
        return fmax(0.0, fmin(a, b)) + fmin(0.0, fmax(a, b));
    }
    if (limiter == 1) return (a * b > 0.0) ? ddiv<F>(2.0 * a * b, a + b) : 0.0;
    if (b > 0.0) return smax(0.0, smin(a, b));
    return smin(0.0, smax(a, b));
}

// ---------------------------------------------------------------- one side of a face

// A Riemann-problem input state with the quantities HLLC derives from it
// (E/headers/riemann_solver_HLLC.hpp:55-67): rho, normal velocity u, p, c; z for Allaire.
// 'ok' is the physical-state test of the evolved state (flux solver step M5).
template <int M> struct Side {
    static constexpr int NC = Model<M>::NC;
    double U[NC];
    double rho, u, p, c;
    double v;       // fast Euler evolved sides: the tangential velocity their U is formed from
    double eps, k;  // Allaire: the mixture coefficients of z = U[5] (the PVRS speeds reuse them)
    bool ok;
};

// The conserved state of a fast-arithmetic Euler evolved side from its rho, u, v, p; returns
// rho e = (p + gamma pinf)/(gamma - 1), without the division by rho that specific_ie's product
// undoes. Such a side carries its primitives only; U is formed by this function where a face
// needs it, from the same four values wherever that is.
CFD_HD double euler_fast_cons(const Material& m, Side<EULER>& s)
{
    const double rho = s.rho, u = s.u, v = s.v;
    const double rhoe = ddiv<true>(s.p + m.g1pinf1, m.g1m1);
    s.U[0] = rho;
    s.U[1] = rho * u;
    s.U[2] = rho * v;
    s.U[3] = rhoe + 0.5 * rho * (u * u + v * v);
    return rhoe;
}

// From a conserved state: HLLC's own primitive recovery (identical to the physical-state
// test's divisions, which the compiler shares).
template <bool F = false> CFD_HD void side_from_cons(const Material& m, Side<EULER>& s)
{
    const double* U = s.U;
    double rho = U[0];
    double u = ddiv<F>(U[1], rho);
    double v = ddiv<F>(U[2], rho);
    double e = ddiv<F>(U[3], rho) - 0.5 * (u * u + v * v);
    double p = Eos<EULER>::pressure(m, rho, e);
    s.rho = rho; s.u = u; s.p = p;
    s.c = Eos<EULER>::soundspeed<F>(m, rho, p);
    s.ok = rho >= 0.0 && e >= 0.0;
}
template <bool F = false> CFD_HD void side_from_cons(const Material& m, Side<ALLAIRE>& s)
{
    const double* U = s.U;
    double rho = U[0] + U[1];
    double u = ddiv<F>(U[2], rho);
    double v = ddiv<F>(U[3], rho);
    double e = ddiv<F>(U[4], rho) - 0.5 * (u * u + v * v);
    double z = U[5];
    Mix x = mixture<F>(m, z);
    double p = MixEos::pressure<F>(x, rho, e);
    s.rho = rho; s.u = u; s.p = p;
    s.c = MixEos::soundspeed<F>(x, rho, p);
    s.eps = x.eps; s.k = x.k;
    s.ok = U[0] >= 0.0 && U[1] >= 0.0 && e >= 0.0 && z >= 0.0 && z <= 1.0;
}

// HLLC inputs of a conserved Allaire state whose mixture coefficients are already known.
template <bool F = false> CFD_HD void side_from_cons_mix(const Mix& x, Side<ALLAIRE>& s)
{
    const double* U = s.U;
    double rho = U[0] + U[1];
    double u = ddiv<F>(U[2], rho);
    double v = ddiv<F>(U[3], rho);
    double e = ddiv<F>(U[4], rho) - 0.5 * (u * u + v * v);
    double z = U[5];
    double p = MixEos::pressure<F>(x, rho, e);
    s.rho = rho; s.u = u; s.p = p;
    s.c = MixEos::soundspeed<F>(x, rho, p);
    s.eps = x.eps; s.k = x.k;
    s.ok = U[0] >= 0.0 && U[1] >= 0.0 && e >= 0.0 && z >= 0.0 && z <= 1.0;
}

// Evolved state from its primitive vector T (step M4 result): p2c, then the M5 test; the
// Allaire test snaps the state in place and the snapped state is what HLLC receives. The exact
// arithmetic recovers HLLC's rho, u, p and c from the conserved state, as the reference does;
// the fast arithmetic takes them from T (the recovery returns T up to its rounding), and the
// test's e >= 0 becomes rho e >= 0 with rho > 0 (the recovery's e is NaN at rho = 0).
template <bool F = false> CFD_HD void side_from_prim_(const Material& m, const double* T, Side<EULER>& s)
{
    if constexpr (F) {
        // U is formed (euler_fast_cons) only where a face needs it; rho e = (p + gamma pinf)/(gamma - 1)
        // has the sign of p + gamma pinf
        s.rho = T[0]; s.u = T[1]; s.v = T[2]; s.p = T[3];
        s.c = Eos<EULER>::soundspeed<F>(m, s.rho, s.p);
        s.ok = s.rho > 0.0 && s.p + m.g1pinf1 >= 0.0;
        return;
    }
    p2c<EULER, F>(m, T, s.U);
    side_from_cons<F>(m, s);
}
template <bool F = false> CFD_HD void side_from_prim_(const Material& m, const double* T, Side<ALLAIRE>& s)
{
    // p2c with the mixture of z = T[5]; the same coefficients serve HLLC unless snapping
    // changes z (then they are recomputed from the snapped z, as the reference does)
    double rho = T[0] + T[1];
    Mix x = mixture<F>(m, T[5]);
    // rho e; the fast arithmetic without the division by rho that specific_ie's product undoes
    double rhoe = F ? (T[4] + x.k) * x.eps : rho * MixEos::specific_ie<F>(x, rho, T[4]);
    if constexpr (F) {
        // no component within 1e-12 of a snapping bound: snapping is the identity (the
        // general path below reaches the same result with the snapping's tests)
        if (!(fabs(T[0]) < 1e-12) && !(fabs(T[1]) < 1e-12) && !(fabs(T[5]) < 1e-12) && !(fabs(T[5] - 1.0) < 1e-12)) {
            s.U[0] = T[0];
            s.U[1] = T[1];
            s.U[2] = rho * T[2];
            s.U[3] = rho * T[3];
            s.U[4] = rhoe + 0.5 * rho * (T[2] * T[2] + T[3] * T[3]);
            s.U[5] = T[5];
            s.rho = rho; s.u = T[2]; s.p = T[4];
            s.c = MixEos::soundspeed<F>(x, rho, T[4]);
            s.eps = x.eps; s.k = x.k;
            s.ok = T[0] >= 0.0 && T[1] >= 0.0 && rho > 0.0 && rhoe >= 0.0 && T[5] >= 0.0 && T[5] <= 1.0;
            return;
        }
    }
    s.U[0] = T[0];
    s.U[1] = T[1];
    s.U[2] = rho * T[2];
    s.U[3] = rho * T[3];
    s.U[4] = rhoe + 0.5 * rho * (T[2] * T[2] + T[3] * T[3]);
    s.U[5] = T[5];
    snap(s.U);
    if constexpr (F) {
        if (s.U[0] == T[0] && s.U[1] == T[1] && s.U[5] == T[5]) {  // snapping changed nothing
            double p = T[4];
            s.rho = rho; s.u = T[2]; s.p = p;
            s.c = MixEos::soundspeed<F>(x, rho, p);
            s.eps = x.eps; s.k = x.k;
            s.ok = T[0] >= 0.0 && T[1] >= 0.0 && rho > 0.0 && rhoe >= 0.0 && T[5] >= 0.0 && T[5] <= 1.0;
            return;
        }
    }
    if (s.U[5] != T[5]) x = mixture<F>(m, s.U[5]);
    side_from_cons_mix<F>(x, s);
}
template <int M, bool F = false> CFD_HD void side_from_prim(const Material& m, const double* T, Side<M>& s)
{
    side_from_prim_<F>(m, T, s);
}

// The conserved state of a side made by side_from_prim: a fast-arithmetic Euler side holds only
// its primitives until a face needs U (every other side already has it).
template <int M, bool F = false> CFD_HD void form_cons(const Material& m, Side<M>& s)
{
    if constexpr (F && M == EULER) (void)euler_fast_cons(m, s);
}

// The first-order fallback state (M5 failed): the raw stencil cell, unsnapped.
template <int M, bool F = false> CFD_HD void side_from_raw(const Material& m, const double* U, Side<M>& s)
{
    for (int k = 0; k < Model<M>::NC; ++k) s.U[k] = U[k];
    side_from_cons<F>(m, s);
}

// ---------------------------------------------------------------- cell stage

// Everything about one cell that the reference recomputes per face: slope-extrapolation
// positivity (M3) for both faces, and the two sloped half-step states (M4, M5).
template <int M> struct CellStage {
    static constexpr int NC = Model<M>::NC;
    bool fR, fL;       // M3: p2c(W + del/2), p2c(W - del/2) physical
    Side<M> R, L;      // sloped evolved states at the cell's right and left faces
};

// tmp = (s * A(W)) * del, summed over k = 0.. in order (Eigen's unrolled lazy product of a
// fixed-size matrix and a vector: E/headers/flux_solver_MUSCLhancock.hpp:197,201), zero entries
// included. The fast arithmetic leaves out the zero entries' products (each adds a signed zero
// to a finite sum) and takes rho c^2 from the primitive state's c field.
template <bool F = false> CFD_HD void jac_times(const double* W, double c, double s, const double* d, double* tmp, Model<EULER>*)
{
    double rho = W[0], u = W[1];
    double a00 = s * u, a01 = s * rho, z0 = s * 0.0;
    double a13 = s * ddiv<F>(1.0, rho);
    if constexpr (F) {
        double a31 = s * c;
        tmp[0] = a01 * d[1] + a00 * d[0];
        tmp[1] = a13 * d[3] + a00 * d[1];
        tmp[2] = a00 * d[2];
        tmp[3] = a00 * d[3] + a31 * d[1];
        return;
    }
    double a31 = s * (rho * c * c);
    tmp[0] = z0 * d[3] + (z0 * d[2] + (a01 * d[1] + a00 * d[0]));
    tmp[1] = a13 * d[3] + (z0 * d[2] + (a00 * d[1] + z0 * d[0]));
    tmp[2] = z0 * d[3] + (a00 * d[2] + (z0 * d[1] + z0 * d[0]));
    tmp[3] = a00 * d[3] + (z0 * d[2] + (a31 * d[1] + z0 * d[0]));
}
template <bool F = false> CFD_HD void jac_times(const double* W, double c, double s, const double* d, double* tmp, Model<ALLAIRE>*)
{
    double rho = W[0] + W[1];
    double su = s * W[2], z0 = s * 0.0;
    double a02 = s * W[0], a12 = s * W[1];
    double a24 = s * ddiv<F>(1.0, rho);
    if constexpr (F) {
        double a42 = s * c;
        tmp[0] = a02 * d[2] + su * d[0];
        tmp[1] = a12 * d[2] + su * d[1];
        tmp[2] = a24 * d[4] + su * d[2];
        tmp[3] = su * d[3];
        tmp[4] = su * d[4] + a42 * d[2];
        tmp[5] = su * d[5];
        return;
    }
    double a42 = s * (rho * c * c);
    tmp[0] = z0 * d[5] + (z0 * d[4] + (z0 * d[3] + (a02 * d[2] + (z0 * d[1] + su * d[0]))));
    tmp[1] = z0 * d[5] + (z0 * d[4] + (z0 * d[3] + (a12 * d[2] + (su * d[1] + z0 * d[0]))));
    tmp[2] = z0 * d[5] + (a24 * d[4] + (z0 * d[3] + (su * d[2] + (z0 * d[1] + z0 * d[0]))));
    tmp[3] = z0 * d[5] + (z0 * d[4] + (su * d[3] + (z0 * d[2] + (z0 * d[1] + z0 * d[0]))));
    tmp[4] = z0 * d[5] + (su * d[4] + (z0 * d[3] + (a42 * d[2] + (z0 * d[1] + z0 * d[0]))));
    tmp[5] = su * d[5] + (z0 * d[4] + (z0 * d[3] + (z0 * d[2] + (z0 * d[1] + z0 * d[0]))));
}

// M3 test of one extrapolated state (Allaire: the snapping overload on a temporary).
// The fast arithmetic decides the Euler test on T itself: for rho > 0, e of p2c(T) has the sign
// of p + gamma pinf (at rho = 0 the recovery's e is NaN, so the state fails).
template <bool F = false> CFD_HD bool extrap_physical(const Material& m, const double* T, Model<EULER>*)
{
    if constexpr (F) return T[0] > 0.0 && T[3] + m.g1pinf1 >= 0.0;
    if (prim_clearly_physical(T[0], T[1], T[2], T[3], m.g1pinf1, m.g1m1)) return true;
    double U[4];
    p2c<EULER, F>(m, T, U);
    return is_physical<EULER, F>(U);
}
template <bool F = false> CFD_HD bool extrap_physical(const Material& m, const double* T, Model<ALLAIRE>*)
{
    // no snapping can occur when both partial densities and z are away from 0 (and z from 1);
    // then the fast arithmetic decides on T itself: e of p2c(T) has the sign of p + k(z), and the
    // mixture k(z) is a weighted mean of gamma_i pinf_i, so p + kmin >= 0 is sufficient
    if constexpr (F) {
        if (T[0] >= 1e-12 && T[1] >= 1e-12 && T[5] >= 1e-12 && T[5] <= 1.0 - 1e-12 && T[4] + m.kmin >= 0.0)
            return true;
    } else {
        if (T[0] >= 1e-12 && T[1] >= 1e-12 && T[5] >= 1e-12 && T[5] <= 1.0 - 1e-12 &&
            prim_clearly_physical(T[0] + T[1], T[2], T[3], T[4], m.kmin, ddiv<F>(1.0, m.epsmin)))
            return true;
    }
    double U[6];
    p2c<ALLAIRE, F>(m, T, U);
    snap(U);
    return is_physical<ALLAIRE, F>(U);
}
template <int M, bool F = false> CFD_HD bool extrap_physical(const Material& m, const double* T)
{
    return extrap_physical<F>(m, T, (Model<M>*)nullptr);
}

// Cell stage from the cell's primitives W (with c) and its two neighbours' primitives.
template <int M, bool F = false>
CFD_HD void cell_stage(const Material& m, int limiter, double s, const Prim<M>& P,
                       const double* Wm, const double* Wp, CellStage<M>& cs)
{
    constexpr int NC = Model<M>::NC;
    const double* W = P.W;
    double del[NC], TR[NC], TL[NC], tmp[NC];
    for (int k = 0; k < NC; ++k) del[k] = limit<F>(limiter, W[k] - Wm[k], Wp[k] - W[k]);
    for (int k = 0; k < NC; ++k) { TR[k] = W[k] + 0.5 * del[k]; TL[k] = W[k] - 0.5 * del[k]; }
    cs.fR = extrap_physical<M, F>(m, TR);
    cs.fL = extrap_physical<M, F>(m, TL);
    jac_times<F>(W, P.c, s, del, tmp, (Model<M>*)nullptr);
    for (int k = 0; k < NC; ++k) { TR[k] = TR[k] - tmp[k]; TL[k] = TL[k] - tmp[k]; }
    side_from_prim<M, F>(m, TR, cs.R);
    side_from_prim<M, F>(m, TL, cs.L);
}

// Unsloped half-step state (slopes zeroed by M3): W + 0.5*0 - (s*A)*0, evaluated literally.
template <int M, bool F = false>
CFD_HD void unsloped_side(const Material& m, double s, const Prim<M>& P, bool right, Side<M>& out)
{
    constexpr int NC = Model<M>::NC;
    double del[NC], T[NC], tmp[NC];
    for (int k = 0; k < NC; ++k) del[k] = 0.0;
    jac_times<F>(P.W, P.c, s, del, tmp, (Model<M>*)nullptr);
    for (int k = 0; k < NC; ++k) T[k] = (right ? P.W[k] + 0.5 * del[k] : P.W[k] - 0.5 * del[k]) - tmp[k];
    side_from_prim<M, F>(m, T, out);
    form_cons<M, F>(m, out);
}

// ---------------------------------------------------------------- HLLC

template <int M> struct FaceOut {
    double F[Model<M>::NC];
    double ustar, zstar;
};

// HLLC flux from two prepared sides (E/headers/riemann_solver_HLLC.hpp:69-144,
// A/headers/riemann_solver_HLLC.hpp:70-149) and z* (A/headers/flux_solver_MUSCLhancock.hpp:375).
// PRIM_U (fast-arithmetic Euler sides that carry only rho, u, v, p): the upwind side's U is
// formed here from its primitives, after the upwind choice, instead of being read from the side.
template <int M, bool F = false, bool PRIM_U = false>
CFD_HD void hllc(const Material& m, int wave_speeds, const Side<M>& L, const Side<M>& R, FaceOut<M>& o)
{
    constexpr int NC = Model<M>::NC;
    constexpr int MN = Model<M>::MN;
    double rhoL = L.rho, uL = L.u, pL = L.p, cL = L.c;
    double rhoR = R.rho, uR = R.u, pR = R.p, cR = R.c;
    double SL, SR;
    if (M == ALLAIRE && wave_speeds == 1) {
        double rho_avg = 0.5 * (rhoL + rhoR);
        double c_avg = 0.5 * (cL + cR);
        double p_pvrs = smax(0.0, 0.5 * (pL + pR) + 0.5 * (uL - uR) * rho_avg * c_avg);
        // shock_speed_jump evaluates the mixture of each side's z (A/headers/mixturemodel.hpp:42-52):
        // the side already holds exactly those coefficients
        Mix xL, xR;
        xL.eps = L.eps; xL.k = L.k;
        xR.eps = R.eps; xR.k = R.k;
        if (p_pvrs > pL) SL = uL - MixEos::shock_jump<F>(xL, pL, rhoL, p_pvrs);
        else SL = uL - cL;
        if (p_pvrs > pR) SR = uR + MixEos::shock_jump<F>(xR, pR, rhoR, p_pvrs);
        else SR = uR + cR;
    } else {
        double u_avg = 0.5 * (uL + uR);
        double c_avg = 0.5 * (cL + cR);
        SL = smin(u_avg - c_avg, uL - cL);
        SR = smax(u_avg + c_avg, uR + cR);
    }
    double us = ddiv<F>(pR - pL + rhoL * uL * (SL - uL) - rhoR * uR * (SR - uR),
                     rhoL * (SL - uL) - rhoR * (SR - uR));
    if constexpr (F) {
        // The weights (1 + sg)/2 and (1 - sg)/2 below are exactly 1 and 0: the fast arithmetic
        // forms the upwind side's star state and flux only (the other term adds a signed zero).
        const bool left = !signbit(us);
        const Side<M>& W = left ? L : R;
        const double S = left ? SL : SR, rho = left ? rhoL : rhoR, u = left ? uL : uR, p = left ? pL : pR;
        const double mS = left ? smin(0.0, SL) : smax(0.0, SR);
        double Uw[NC];
        if constexpr (PRIM_U && M == EULER) {
            Side<EULER> t;
            t.rho = rho; t.u = u; t.v = left ? L.v : R.v; t.p = p;
            (void)euler_fast_cons(m, t);
            for (int k = 0; k < NC; ++k) Uw[k] = t.U[k];
        } else {
            for (int k = 0; k < NC; ++k) Uw[k] = W.U[k];
        }
        // one reciprocal, 1/(S - us): f = (S - u)/(S - us), and the star energy
        // f (E + (us - u)(rho us + p/(S - u))) as f (E + (us - u) rho us) + (us - u) p/(S - us)
        const double rd = ddiv<F>(1.0, S - us);
        const double f = (S - u) * rd;
        const double dus = us - u, rus = rho * us;
        const double fe = __builtin_fma(f, __builtin_fma(dus, rus, Uw[NC == 4 ? 3 : 4]), dus * p * rd);
        for (int k = 0; k < NC; ++k) {
            double st, fl;
            if (M == ALLAIRE && k == 5) {
                o.F[k] = 0.0;
                continue;
            } else if (k == MN) {
                st = f * rus;
                fl = Uw[k] * u + p;
            } else if (k == NC - 1 - (M == ALLAIRE ? 1 : 0)) {
                st = fe;
                fl = (Uw[k] + p) * u;
            } else {
                st = f * Uw[k];
                fl = Uw[k] * u;
            }
            o.F[k] = fl + mS * (st - Uw[k]);
        }
        o.ustar = us;
        o.zstar = M == ALLAIRE ? Uw[NC - 1] : 0.0;
        return;
    }
    double fL = ddiv<F>(SL - uL, SL - us);
    double fR = ddiv<F>(SR - uR, SR - us);
    double sg = copysign(1.0, us);
    double wL = (1.0 + sg) / 2.0;
    double wR = (1.0 - sg) / 2.0;
    double mSL = smin(0.0, SL);
    double MSR = smax(0.0, SR);
    double eL = L.U[NC == 4 ? 3 : 4] + (us - uL) * (rhoL * us + ddiv<F>(pL, SL - uL));
    double eR = R.U[NC == 4 ? 3 : 4] + (us - uR) * (rhoR * us + ddiv<F>(pR, SR - uR));
    for (int k = 0; k < NC; ++k) {
        double usl, usr, fl, fr;
        if (M == ALLAIRE && k == 5) {
            usl = 0.0; usr = 0.0; fl = 0.0; fr = 0.0;
        } else if (k == MN) {
            usl = fL * rhoL * us;
            usr = fR * rhoR * us;
            fl = L.U[k] * uL + pL;
            fr = R.U[k] * uR + pR;
        } else if (k == NC - 1 - (M == ALLAIRE ? 1 : 0)) {  // energy
            usl = fL * eL;
            usr = fR * eR;
            fl = (L.U[k] + pL) * uL;
            fr = (R.U[k] + pR) * uR;
        } else {
            usl = fL * L.U[k];
            usr = fR * R.U[k];
            fl = L.U[k] * uL;
            fr = R.U[k] * uR;
        }
        o.F[k] = wL * (fl + mSL * (usl - L.U[k])) + wR * (fr + MSR * (usr - R.U[k]));
    }
    if (M == ALLAIRE) o.F[5] = 0.0;
    o.ustar = us;
    if (M == ALLAIRE) o.zstar = 0.5 * (1.0 + sg) * L.U[5] + 0.5 * (1.0 - sg) * R.U[5];
    else o.zstar = 0.0;
}

// ---------------------------------------------------------------- face stage

// Face between cell a (left) and cell b (right): M3 combination, M5 fallback, HLLC.
// flags: bit0 slope fallback taken, bit1 evolve fallback taken. The two fallbacks only choose
// the Riemann-problem inputs; there is one HLLC evaluation per face.
template <int M, bool F = false>
CFD_HD int face_stage(const Material& m, const SchemeOpts& o, double s,
                      const Prim<M>& Pa, const double* Ua, const CellStage<M>& a,
                      const Prim<M>& Pb, const double* Ub, const CellStage<M>& b, FaceOut<M>& out)
{
    int flags = 0;
    Side<M> L = a.R, R = b.L;
    form_cons<M, F>(m, L);
    form_cons<M, F>(m, R);
    if (!(a.fR && b.fL)) {
        flags |= 1;
        unsloped_side<M, F>(m, s, Pa, true, L);
        unsloped_side<M, F>(m, s, Pb, false, R);
    }
    if (!(L.ok && R.ok)) {
        flags |= 2;
        side_from_raw<M, F>(m, Ua, L);
        side_from_raw<M, F>(m, Ub, R);
    }
    hllc<M, F>(m, o.wave_speeds, L, R, out);
    return flags;
}

// ---------------------------------------------------------------- update

// Euler repair (E/sourcecode/euler_rowsweep.cpp:47-82): 'old' and 'fut' in the unrotated frame.
template <bool F = false> CFD_HD bool euler_repair(const Material& m, const double* old, double* fut)
{
    if constexpr (F) {
        // e >= 0 with rho > 0 is E rho - (mu^2 + mv^2)/2 >= 0; other states take the full test
        if (fut[0] > 0.0 && fut[3] * fut[0] - 0.5 * (fut[1] * fut[1] + fut[2] * fut[2]) >= 0.0) return false;
    } else if (cons_clearly_physical(fut[0], fut[1], fut[2], fut[3])) {
        return false;
    }
    double rho = fut[0];
    double u = ddiv<F>(fut[1], rho);
    double v = ddiv<F>(fut[2], rho);
    double e = ddiv<F>(fut[3], rho) - 0.5 * (u * u + v * v);
    if (rho >= 0.0 && e >= 0.0) return false;
    double rho_old = old[0];
    double u_old = ddiv<F>(old[1], rho_old);
    double v_old = ddiv<F>(old[2], rho_old);
    double e_old = ddiv<F>(old[3], rho_old) - 0.5 * (u_old * u_old + v_old * v_old);
    if (rho <= 1e-12) rho = rho_old;
    if (e <= 1e-12) e = e_old;
    double p = Eos<EULER>::pressure(m, rho, e);
    double W[4] = {rho, u, v, p};
    p2c<EULER, F>(m, W, fut);
    return true;
}

// A/headers/zupdate.hpp:62-67 with the clamp bounds parameterised.
CFD_HD double zupdate(double r, double z, double usL, double usR, double zsL, double zsR,
                      double z_lo, double z_hi)
{
    double newz = z - r * (usR * zsR - usL * zsL - z * (usR - usL));
    return smax(smin(z_hi, newz), z_lo);
}

// First-order upwind z update (A/headers/zupdate.hpp:44-50), used with the Godunov flux.
CFD_HD double zupdate_upwind(double r, double zL, double z, double zR, double usL, double usR, double z_lo,
                             double z_hi)
{
    double newz = z - r * ((zR - z) * smin(0.0, usR) + (z - zL) * smax(0.0, usL));
    return smax(smin(z_hi, newz), z_lo);
}

// Godunov face (E/headers/flux_solver_godunov.hpp:30-36, A/headers/flux_solver_godunov.hpp:30-39):
// HLLC of the two cells' own conserved states; z* as in hllc.
template <int M, bool F = false>
CFD_HD void godunov_face(const Material& m, int wave_speeds, const double* Ua, const double* Ub, FaceOut<M>& out)
{
    Side<M> L, R;
    side_from_raw<M, F>(m, Ua, L);
    side_from_raw<M, F>(m, Ub, R);
    hllc<M, F>(m, wave_speeds, L, R, out);
}

// Wave speed |u|+c, |v|+c of a final (unrotated) state: compute_dt's per-cell work
// (E/sourcecode/euler_timestep.cpp:16-24, A/sourcecode/allaire_diffuse_timestep.cpp:17-26).
// The fast arithmetic forms the Euler speeds from one reciprocal of rho: |u| + c = |mu| (1/rho) + c
// with c = sqrt(gamma (p + pinf) / rho) and gamma (p + pinf) = wa w + wb (the pressure's
// gamma pinf and the sound speed's pinf folded into one constant).
template <bool F = false> CFD_HD void wave_speeds(const Material& m, const double* U, double& su, double& sv, Model<EULER>*)
{
    if constexpr (F) {
        const double r = ddiv<F>(1.0, U[0]);
        const double w = __builtin_fma(-0.5 * (U[1] * U[1] + U[2] * U[2]), r, U[3]);
        const double c = sqrt_ratio<F>(__builtin_fma(w, m.wa, m.wb), U[0]);
        su = __builtin_fma(fabs(U[1]), r, c);
        sv = __builtin_fma(fabs(U[2]), r, c);
        return;
    }
    double rho = U[0];
    double u = ddiv<F>(U[1], rho);
    double v = ddiv<F>(U[2], rho);
    double e = ddiv<F>(U[3], rho) - 0.5 * (u * u + v * v);
    double p = Eos<EULER>::pressure(m, rho, e);
    double c = Eos<EULER>::soundspeed<F>(m, rho, p);
    su = fabs(u) + c;
    sv = fabs(v) + c;
}
// The fast arithmetic forms the Allaire speeds from the reciprocals of rho and of eps(z):
// rho e = E - (mu^2 + mv^2)/(2 rho), p = rho e / eps - k and rho c^2 = k + p (1/eps + 1).
template <bool F = false> CFD_HD void wave_speeds(const Material& m, const double* U, double& su, double& sv, Model<ALLAIRE>*)
{
    if constexpr (F) {
        const double rho = U[0] + U[1];
        const double r = ddiv<F>(1.0, rho);
        const double w = __builtin_fma(-0.5 * (U[2] * U[2] + U[3] * U[3]), r, U[4]);
        const double reps = ddiv<F>(1.0, __builtin_fma(U[5], m.fe12, m.fe2));
        const double k = __builtin_fma(U[5], m.fk12, m.fk2) * reps;
        const double p = __builtin_fma(w, reps, -k);
        const double c = sqrt_ratio<F>(k + p * (reps + 1.0), rho);
        su = __builtin_fma(fabs(U[2]), r, c);
        sv = __builtin_fma(fabs(U[3]), r, c);
        return;
    }
    double rho = U[0] + U[1];
    double u = ddiv<F>(U[2], rho);
    double v = ddiv<F>(U[3], rho);
    double e = ddiv<F>(U[4], rho) - 0.5 * (u * u + v * v);
    double z = U[5];
    Mix x = mixture<F>(m, z);
    double p = MixEos::pressure<F>(x, rho, e);
    double c = MixEos::soundspeed<F>(x, rho, p);
    su = fabs(u) + c;
    sv = fabs(v) + c;
}

}  // namespace cfd
