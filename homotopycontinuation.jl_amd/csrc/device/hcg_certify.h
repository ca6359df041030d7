// Krawczyk certification of approximate zeros of a square system F(x; p), one candidate
// per lane: the ε-inflation test of certify_solution / ε_inflation_krawczyk
// (src/certification.jl:1029-1114, 1225-1287) in FP64 interval arithmetic.
//
// Per candidate: Newton refinement with double-double residuals (newton(...; rtol = 8 eps,
// atol = 0, extended_precision = true, max_iters = 8), src/newton.jl:107-175), an approximate
// inverse C of the last Jacobian, then
//   r₀ = F([x̃]),  Δx₀ = C r₀,  x₀ = x̃ + [-εᵢ, εᵢ](1 + i) with εᵢ = 512 max(mag(Δx₀ᵢ), eps),
//   M = C J(x₀) - I,  x₁ = (x̃ - Δx₀) - M (x₀ - x̃)
// in interval arithmetic.  x₁ ⊂ interior(x₀) together with 1.41422 maxᵢ Σⱼ mag(Mᵢⱼ) < 0.7071
// proves a unique zero in x₀ (Krawczyk), contained in x₁.  It is real if conj(x₁) ⊂
// interior(x₀) too and F maps real points to real values; it is not real if some
// coordinate of x₁ excludes the real axis.  When the test fails and extended precision is
// on, it is repeated with x̃ refined in double-double and r₀ enclosed by double-double balls,
// and, if that fails too, with C recomputed from a double-double Jacobian (the reference
// retries with Arb balls of growing precision instead).
//
// The generated source defines HCG_N, HCG_NPAR and
//   cert_eval_jac(x, P, u, U)   F and ∂F/∂x at a point (U column-major)
//   cert_eval_dd(x, P, u)       F with a double-double argument
//   cert_ieval(X, IP, u)        F on a box,  cert_ijac(X, IP, U)  ∂F/∂x on a box
//   cert_beval(x, BP, u)        F at a double-double point in double-double balls
//   cert_bjac(x, BP, U)         ∂F/∂x at that point, the same way
// with P the parameter point (HCG_NPAR complex), IP its box (HCG_NPAR × 4 doubles) and BP
// its double-double ball (HCG_NPAR × 6 doubles).
#pragma once
#include "hcg_cert_kargs.h"

#define CN HCG_N
#define CAT(A, i, j) A[(i) + (j) * CN]

// lu!, src/linear_algebra.jl:130-184: partial pivoting on abs2, rows swapped in place
HD void cert_lu(cplx* A, int* ipiv) {
    for (int k = 0; k < CN; k++) {
        int kp = k;
        double amax = c_abs2(CAT(A, k, k));
        for (int i = k + 1; i < CN; i++) {
            const double absi = c_abs2(CAT(A, i, k));
            if (absi > amax) { kp = i; amax = absi; }
        }
        ipiv[k] = kp;
        if (amax != 0.0) {
            if (k != kp)
                for (int j = 0; j < CN; j++) { cplx t = CAT(A, k, j); CAT(A, k, j) = CAT(A, kp, j); CAT(A, kp, j) = t; }
            const cplx akk = CAT(A, k, k);
            for (int i = k + 1; i < CN; i++) CAT(A, i, k) = c_div(CAT(A, i, k), akk);
        }
        for (int j = k + 1; j < CN; j++) {
            const cplx akj = CAT(A, k, j);
            for (int i = k + 1; i < CN; i++) CAT(A, i, j) = c_submul(CAT(A, i, k), akj, CAT(A, i, j));
        }
    }
}
// lu_ldiv!, src/linear_algebra.jl:268-316 (in place)
HD void cert_lu_solve(const cplx* A, const int* ipiv, cplx* x) {
    for (int i = 0; i < CN; i++)
        if (i != ipiv[i]) { cplx t = x[i]; x[i] = x[ipiv[i]]; x[ipiv[i]] = t; }
    for (int j = 0; j < CN; j++) {
        const cplx xj = x[j];
        for (int i = j + 1; i < CN; i++) x[i] = c_submul(CAT(A, i, j), xj, x[i]);
    }
    for (int j = CN - 1; j >= 0; j--) {
        const cplx xj = c_div(x[j], CAT(A, j, j));
        x[j] = xj;
        for (int i = 0; i < j; i++) x[i] = c_submul(CAT(A, i, j), xj, x[i]);
    }
}
// norm(x, InfNorm()), src/norm.jl:194-201
HD double cert_inf_norm(const cplx* x) {
    double m = c_abs2(x[0]);
    for (int i = 1; i < CN; i++) m = max_fast(m, c_abs2(x[i]));
    return sqrt(m);
}

struct KrawczykOut {
    bool certified, conj_inside, not_real;
    double bound, inflation;
};
// Given the point x̃ (a double-double vector; its low parts are zero in the FP64 test), the
// approximate inverse C and an enclosure R0 of F(x̃): Δx₀ = C r₀, the inflated box x₀,
// M = C J(x₀) - I, x₁ = (x̃ - Δx₀) - M (x₀ - x̃); writes x₁ and x₀.  The FP64 test
// (ext = false) is the reference's arithmetic step for step; the extended test encloses
// x̃ and x₀ - x̃ through double-double balls and floors εᵢ at 512 eps (1 + |x̃ᵢ|) instead of
// 512 eps, since the box must stay wider than the FP64 enclosure of x̃.
__device__ __noinline__ KrawczykOut cert_krawczyk(const cdd* xd, const cplx* C, const icplx* R0, const double* IP,
                                                  bool ext, double* box1, double* box0) {
    icplx D[CN], X0[CN], dX[CN], XT[CN];
    KrawczykOut o;
    o.inflation = 0.0;
    for (int i = 0; i < CN; i++) {
        icplx d = ic_zero();
        for (int kk = 0; kk < CN; kk++) d = ic_muladd_pt(CAT(C, i, kk), R0[kk], d);
        D[i] = d;  // Δx₀ = C r₀
        const cplx h = cdd_round(xd[i]);
        if (!ext) {
            const double e = 512.0 * jmax(ic_mag(d), DEPS);
            o.inflation = jmax(o.inflation, e);
            X0[i] = mkic(mki(h.re - e, h.re + e), mki(h.im - e, h.im + e));
            dX[i] = mkic(iv_sub_r(X0[i].re, h.re), iv_sub_r(X0[i].im, h.im));  // x₀ - x̃
            XT[i] = ic_pt(h);
        } else {
            const double e = 512.0 * jmax(ic_mag(d), DEPS * (1.0 + jmax(fabs(h.re), fabs(h.im))));
            o.inflation = jmax(o.inflation, e);
            X0[i] = mkic(mki(iv_prev(iv_prev(h.re) - e), iv_next(iv_next(h.re) + e)),
                         mki(iv_prev(iv_prev(h.im) - e), iv_next(iv_next(h.im) + e)));
            const dball xr = mkdb(xd[i].re, 0.0), xi = mkdb(xd[i].im, 0.0);
            dX[i] = mkic(mki(db_to_ival(db_sub(mkdb(mkdd(X0[i].re.lo, 0.0), 0.0), xr)).lo,
                             db_to_ival(db_sub(mkdb(mkdd(X0[i].re.hi, 0.0), 0.0), xr)).hi),
                         mki(db_to_ival(db_sub(mkdb(mkdd(X0[i].im.lo, 0.0), 0.0), xi)).lo,
                             db_to_ival(db_sub(mkdb(mkdd(X0[i].im.hi, 0.0), 0.0), xi)).hi));
            XT[i] = mkic(db_to_ival(xr), db_to_ival(xi));
        }
    }
    icplx JI[CN * CN];
    cert_ijac(X0, IP, JI);
    double bound = 0.0;
    bool inside = true;
    o.conj_inside = true;
    o.not_real = false;
    for (int i = 0; i < CN; i++) {
        ival b = mki(0.0, 0.0);
        icplx x1 = ic_zero();
        for (int j = 0; j < CN; j++) {
            icplx m = ic_zero();  // M = C J(x₀) - I, row i
            for (int kk = 0; kk < CN; kk++) m = ic_muladd_pt(CAT(C, i, kk), CAT(JI, kk, j), m);
            if (i == j) m.re = iv_sub_r(m.re, 1.0);
            b = iv_add_r(b, ic_mag(m));
            x1 = ic_muladd(m, dX[j], x1);
        }
        bound = jmax(bound, iv_mag(b));
        x1 = ci_sub(ci_sub(XT[i], D[i]), x1);  // x₁ = (x̃ - Δx₀) - M (x₀ - x̃)
        inside = inside && ic_interior(x1, X0[i]);
        o.conj_inside = o.conj_inside && ic_interior(ic_conj(x1), X0[i]);
        o.not_real = o.not_real || !iv_contains(x1.im, 0.0);
        double* o1 = box1 + 4 * i;
        o1[0] = x1.re.lo; o1[1] = x1.re.hi; o1[2] = x1.im.lo; o1[3] = x1.im.hi;
        double* o0 = box0 + 4 * i;
        o0[0] = X0[i].re.lo; o0[1] = X0[i].re.hi; o0[2] = X0[i].im.lo; o0[3] = X0[i].im.hi;
    }
    o.bound = 1.41422 * bound;
    o.certified = o.bound < 0.7071 && inside;
    return o;
}

// Double-double LU with partial pivoting (pivots chosen on the rounded entries) and solve
HD void cert_lu_dd(cdd* A, int* ipiv) {
    for (int k = 0; k < CN; k++) {
        int kp = k;
        double amax = c_abs2(cdd_round(CAT(A, k, k)));
        for (int i = k + 1; i < CN; i++) {
            const double absi = c_abs2(cdd_round(CAT(A, i, k)));
            if (absi > amax) { kp = i; amax = absi; }
        }
        ipiv[k] = kp;
        if (amax != 0.0) {
            if (k != kp)
                for (int j = 0; j < CN; j++) { cdd t = CAT(A, k, j); CAT(A, k, j) = CAT(A, kp, j); CAT(A, kp, j) = t; }
            const cdd akk = CAT(A, k, k);
            for (int i = k + 1; i < CN; i++) CAT(A, i, k) = cdd_div(CAT(A, i, k), akk);
        }
        for (int j = k + 1; j < CN; j++) {
            const cdd akj = CAT(A, k, j);
            for (int i = k + 1; i < CN; i++) CAT(A, i, j) = cdd_submul(CAT(A, i, k), akj, CAT(A, i, j));
        }
    }
}
HD void cert_lu_solve_dd(const cdd* A, const int* ipiv, cdd* x) {
    for (int i = 0; i < CN; i++)
        if (i != ipiv[i]) { cdd t = x[i]; x[i] = x[ipiv[i]]; x[ipiv[i]] = t; }
    for (int j = 0; j < CN; j++) {
        const cdd xj = x[j];
        for (int i = j + 1; i < CN; i++) x[i] = cdd_submul(CAT(A, i, j), xj, x[i]);
    }
    for (int j = CN - 1; j >= 0; j--) {
        const cdd xj = cdd_div(x[j], CAT(A, j, j));
        x[j] = xj;
        for (int i = 0; i < j; i++) x[i] = cdd_submul(CAT(A, i, j), xj, x[i]);
    }
}
// C ← the inverse of the Jacobian at the double-double point x̃, computed in double-double
// and rounded (the role of approx_inv! on the Arb Jacobian, src/certification.jl:1384-1391)
__device__ __noinline__ void cert_dd_inverse(const cdd* xd, const double* BP, cplx* C) {
    cball JB[CN * CN];
    cert_bjac(xd, BP, JB);
    cdd A[CN * CN];
    for (int i = 0; i < CN * CN; i++) A[i] = mkcdd(JB[i].re.m, JB[i].im.m);
    int ipiv[CN];
    cert_lu_dd(A, ipiv);
    for (int j = 0; j < CN; j++) {
        cdd e[CN];
        for (int i = 0; i < CN; i++) e[i] = mkcdd(mkdd(i == j ? 1.0 : 0.0, 0.0), mkdd(0.0, 0.0));
        cert_lu_solve_dd(A, ipiv, e);
        for (int i = 0; i < CN; i++) CAT(C, i, j) = cdd_round(e[i]);
    }
}
// x̃ ← x̃ - C mid(F(x̃)) in double-double with ball residuals, at most 10 steps, until
// ‖C F(x̃)‖ stops decreasing (the refinement of arb_ε_inflation_krawczyk,
// src/certification.jl:1303-1320); leaves the enclosure of F at the final x̃ in R0.
__device__ __noinline__ void cert_refine_dd(cdd* xd, const cplx* C, const double* BP, icplx* R0) {
    double acc = HCG_INF;
    for (int it = 1; it <= 10; it++) {
        cball rb[CN];
        cert_beval(xd, BP, rb);
        double nrm = 0.0;
        cplx dl[CN];
        for (int i = 0; i < CN; i++) R0[i] = cb_to_ic(rb[i]);
        for (int i = 0; i < CN; i++) {
            icplx d = ic_zero();
            for (int kk = 0; kk < CN; kk++) d = ic_muladd_pt(CAT(C, i, kk), R0[kk], d);
            nrm = jmax(nrm, ic_mag(d));
            dl[i] = mkc(0.5 * d.re.lo + 0.5 * d.re.hi, 0.5 * d.im.lo + 0.5 * d.im.hi);
        }
        if (!(nrm <= acc) || it == 10) break;
        acc = nrm;
        for (int i = 0; i < CN; i++)
            xd[i] = mkcdd(dd_add_acc(xd[i].re, mkdd(-dl[i].re, 0.0)), dd_add_acc(xd[i].im, mkdd(-dl[i].im, 0.0)));
    }
}

// Four waves per SIMD (128 registers per lane): the kernel is latency-bound on its private
// memory, and occupancy beats registers here.  Measured on 262,144 candidates at n = 10: 176 ms
// at one wave per SIMD, 109 ms at two, 81 ms at four, 127 ms at eight.
extern "C" __global__ void __launch_bounds__(64, 4) hcg_certify_kernel(hcg_cert_kargs a) {
    const long k = (long)blockIdx.x * 64 + threadIdx.x;
    if (k >= a.nsol) return;
    const int g = a.group_of ? a.group_of[k] : 0;
    const double* P = a.params ? a.params + 2 * HCG_NPAR * g : nullptr;
    const double* IP = a.params_ival ? a.params_ival + 4 * HCG_NPAR * g : nullptr;
    const double* BP = a.params_ball ? a.params_ball + 6 * HCG_NPAR * g : nullptr;
    cplx x[CN], r[CN], LU[CN * CN];
    int ipiv[CN];
    for (int i = 0; i < CN; i++) x[i] = mkc(a.candidates[2 * (k * CN + i)], a.candidates[2 * (k * CN + i) + 1]);

    // ---- Newton refinement (newton, src/newton.jl:107-175; contraction factor 1) ----
    int code = -1, iters = 0;
    if (a.refine) {
        const double norm_x = cert_inf_norm(x), tol = 8.0 * DEPS * norm_x;
        double prev = HCG_NAN, ac = 1.0;
        code = 2;
        iters = 8;
        for (int it = 1; it <= 8; it++) {
            cert_eval_jac(x, P, r, LU);
            cdd xe[CN], re[CN];
            for (int i = 0; i < CN; i++) xe[i] = cdd_from(x[i]);
            cert_eval_dd(xe, P, re);
            for (int i = 0; i < CN; i++) r[i] = cdd_round(re[i]);
            cert_lu(LU, ipiv);
            cert_lu_solve(LU, ipiv, r);
            const double nrm = cert_inf_norm(r);
            for (int i = 0; i < CN; i++) x[i] = c_sub(x[i], r[i]);
            if (it > 1) {
                if (nrm < ac * prev) {
                    prev = nrm;
                    ac *= ac;
                } else {
                    code = nrm < jmax(0.0, tol) ? 0 : 1;
                    iters = it;
                    break;
                }
            } else {
                prev = nrm;
            }
        }
    } else {
        cert_eval_jac(x, P, r, LU);
        cert_lu(LU, ipiv);
    }
    // approximate inverse C of the last Jacobian, column by column from its LU factors
    cplx C[CN * CN];
    for (int j = 0; j < CN; j++) {
        cplx e[CN];
        for (int i = 0; i < CN; i++) e[i] = mkc(i == j ? 1.0 : 0.0, 0.0);
        cert_lu_solve(LU, ipiv, e);
        for (int i = 0; i < CN; i++) CAT(C, i, j) = e[i];
    }

    // ---- ε-inflation Krawczyk test (src/certification.jl:1225-1287) ----
    icplx X[CN], R0[CN];
    cdd xd[CN];
    for (int i = 0; i < CN; i++) {
        X[i] = ic_pt(x[i]);
        xd[i] = cdd_from(x[i]);
    }
    cert_ieval(X, IP, R0);  // r₀ = F([x̃]) in FP64 intervals
    double* b1 = a.box1 + 4 * k * CN;
    double* b0 = a.box0 + 4 * k * CN;
    KrawczykOut o = cert_krawczyk(xd, C, R0, IP, false, b1, b0);
    int prec = 53;
    if (!o.certified && a.extended) {
        // Extended precision (the role of the reference's Arb retry, extended_prec_certify_solution
        // and arb_ε_inflation_krawczyk, src/certification.jl:1116-1218, 1289-1394): x̃ becomes a
        // double-double point refined with double-double ball residuals and the test runs again;
        // if it still fails, C is recomputed from a double-double Jacobian and it runs once more.
        cert_refine_dd(xd, C, BP, R0);
        o = cert_krawczyk(xd, C, R0, IP, true, b1, b0);
        prec = 106;
        if (!o.certified) {
            cert_dd_inverse(xd, BP, C);
            cert_refine_dd(xd, C, BP, R0);
            o = cert_krawczyk(xd, C, R0, IP, true, b1, b0);
        }
    }
    for (int i = 0; i < CN; i++) {
        const cplx h = cdd_round(xd[i]);
        a.refined[2 * (k * CN + i)] = h.re;
        a.refined[2 * (k * CN + i) + 1] = h.im;
    }
    const bool is_complex = o.certified && o.not_real;
    hcg_certificate c;
    c.certified = o.certified;
    c.real = o.certified && o.conj_inside && a.real_system && !is_complex;
    c.complex_ = is_complex;
    c.newton_code = code;
    c.newton_iters = iters;
    c.prec = prec;
    c.contraction = o.bound;
    c.inflation = o.inflation;
    a.out[k] = c;
}
