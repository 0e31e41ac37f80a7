// adjoint_math.hpp -- per-lane pieces of the discrete adjoint of the WHFast
// democratic-heliocentric step with tides, rotational flattening and Kidder
// 1995 general relativity.
//
// The includer defines POS_HD (the function qualifiers: `__device__ inline`
// in the HIP library, `inline` in the host-side derivative check) before
// including this header.  Everything here is plain FP64 arithmetic on one
// host/planet pair or one body; the reductions over the bodies of a system
// live in the kernel.
//
// The functions restate the equations of the reference's effect kernels in a
// form suited to differentiation.  They are evaluated at states the forward
// step has already produced and are not meant to reproduce its rounding:
//   tides (constant time lag)   src/effects/tides/constant_time_lag.rs:171-332,
//                               src/effects/tides/common.rs:143-170, 223-261, 312-345
//   rotational flattening       src/effects/rotational_flattening/oblate_spheroid.rs:21-97,
//                               src/effects/rotational_flattening/common.rs:93-120, 165-237
//   Kidder 1995 GR              src/effects/general_relativity.rs:177-456
//   universal-variable Kepler   src/integrator/whfast.rs:676-876
//   Newtonian pair              src/particles/universe.rs:198-303
#pragma once

namespace posadj {

constexpr double AU_M = 1.49597870700e11;
constexpr double DAY_S = 24. * 3600.;
constexpr double G_K2 = 6.67428e-11 / (AU_M * AU_M * AU_M) * 1.9818e30 * (DAY_S * DAY_S);
constexpr double C_LIGHT = (2.99792458e8 / AU_M) * DAY_S;
constexpr double C2 = C_LIGHT * C_LIGHT;

struct V3 {
    double x, y, z;
};
POS_HD V3 v3(double x, double y, double z) { return V3{x, y, z}; }
POS_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
POS_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
POS_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
POS_HD V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
POS_HD V3 &operator+=(V3 &a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
POS_HD V3 &operator-=(V3 &a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
POS_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
POS_HD V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
POS_HD V3 hadamard(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
POS_HD V3 zero3() { return V3{0., 0., 0.}; }
// adjoint of c = a x b:  a_b += b x c_b,  b_b += c_b x a
POS_HD void cross_adj(V3 a, V3 b, V3 cb, V3 &ab, V3 &bb) { ab += cross(b, cb); bb += cross(cb, a); }

// ---------------------------------------------------------------------------
// Host/planet pair: the additional forces and torques of one orbiting body
// ---------------------------------------------------------------------------
// Constants of a pair (masses, radii, moments of inertia) and its
// differentiable parameters (dissipation factors sigma = scale * factor,
// tidal Love numbers k2, fluid Love numbers k2f), host first.
struct PairConst {
    double mh, mp, Rh, Rp, Ih, Ip;
    bool tides, flat, gr;
};
struct PairPar {
    double sh, sp, k2h, k2p, kfh, kfp;
};
// State of a pair: heliocentric position r and velocity u of the planet,
// spins W = L / I of host and planet.
struct PairIn {
    V3 r, u, Wh, Wp;
};
// Outputs: F = tidal + flattening force on the planet (the host receives -F),
// g = GR acceleration of the planet (the host receives -(mp/mh) g),
// Np, Nh = dL/dt of planet and host due to this pair.
struct PairOut {
    V3 F, g, Np, Nh;
};
struct PairParBar {
    double sh, sp, k2h, k2p, kfh, kfp;
};

POS_HD double pw5(double x) { double x2 = x * x; return x2 * x2 * x; }
POS_HD double pw10(double x) { double x5 = pw5(x); return x5 * x5; }

POS_HD PairOut pair_forward(const PairConst &C, const PairPar &P, const PairIn &I) {
    PairOut o{zero3(), zero3(), zero3(), zero3()};
    const V3 r = I.r, u = I.u;
    const double d2 = dot(r, r), d = sqrt(d2), id = 1. / d;
    const double u2 = dot(u, u);
    const double vr = dot(r, u) * id;
    const double rsh = dot(r, I.Wh), rsp = dot(r, I.Wp);
    const double d5 = d2 * d2 * d, d7 = d5 * d2, d8 = d7 * d;
    if (C.tides) {
        const double Ah = 4.5 * C.mp * C.mp * pw10(C.Rh) * P.sh;
        const double Ap = 4.5 * C.mh * C.mh * pw10(C.Rp) * P.sp;
        const double os = Ah / d7, op = Ap / d7;
        const double Ck = C.mp * C.mp * pw5(C.Rh) * P.k2h + C.mh * C.mh * pw5(C.Rp) * P.k2p;
        const double f3 = -3. * G_K2 * Ck / d7 - 2. * vr * (Ah + Ap) / d8;
        const V3 c = cross(r, u);
        const V3 wh = cross(I.Wh, r) - u, wp = cross(I.Wp, r) - u;
        o.F += (f3 * id) * r + (os * id) * wh + (op * id) * wp;
        o.Np += (-op) * (d * I.Wp - (rsp * id) * r - id * c);
        o.Nh += (-os) * (d * I.Wh - (rsh * id) * r - id * c);
    }
    if (C.flat) {
        const double Bh = C.mp * P.kfh * pw5(C.Rh) / 6., Bp = C.mh * P.kfp * pw5(C.Rp) / 6.;
        const double Wh2 = dot(I.Wh, I.Wh), Wp2 = dot(I.Wp, I.Wp);
        const double ofs = -6. * Bh * rsh / d5, ofp = -6. * Bp * rsp / d5;
        const double radf = -3. * (Bp * Wp2 + Bh * Wh2) / d5 + 15. * (Bh * rsh * rsh + Bp * rsp * rsp) / d7;
        o.F += radf * r + ofp * I.Wp + ofs * I.Wh;
        o.Np += (-ofp) * cross(r, I.Wp);
        o.Nh += (-ofs) * cross(r, I.Wh);
    }
    if (C.gr) {
        const double mgh = G_K2 * C.mh, mgp = G_K2 * C.mp, mg = mgh + mgp;
        const double f = mgh * mgp / (mg * mg);
        const double mgc = -mg / (d2 * C2);
        const double mgd = mg / d;
        const double vr2 = vr * vr;
        // 1PN and 2PN (general_relativity.rs:187-297)
        const double P1 = (1. + 3. * f) * u2 - 2. * (2. + f) * mgd - 1.5 * f * vr2;
        const double P2 = 0.75 * (12. + 29. * f) * mg * mg / d2 + f * (3. - 4. * f) * u2 * u2
                        + 15. / 8. * f * (1. - 3. * f) * vr2 * vr2 - 1.5 * f * (3. - 4. * f) * vr2 * u2
                        - 0.5 * f * (13. - 4. * f) * mgd * u2 - (2. + 25. * f + 2. * f * f) * mgd * vr2;
        const double Q2 = f * (15. + 4. * f) * u2 - (4. + 41. * f + 8. * f * f) * mgd - 3. * f * (3. + 2. * f) * vr2;
        const double rad = mgc * P1 + mgc * P2;
        const double orth = -mgc * 2. * (2. - f) * vr - 0.5 * mgc * vr * Q2;
        o.g += (rad * id) * r + orth * u;
        // 1.5PN spin-orbit and the GR torques (general_relativity.rs:300-456)
        const double fa = G_K2 / C2;
        const V3 Ls = C.Ih * I.Wh, Lp = C.Ip * I.Wp;
        const double dm = C.mh - C.mp;
        const V3 msf = (dm / C.mp) * Lp - (dm / C.mh) * Ls;
        const V3 S = Ls + Lp;
        const V3 e7 = 7. * S + 3. * msf, e3 = 3. * S + msf, e2 = 2. * S + msf;
        const V3 n = id * r;
        const V3 nu = cross(n, u);
        const V3 e1 = 6. * hadamard(hadamard(n, nu), e2);
        o.g += fa * (e1 - cross(u, e7) + (3. * vr) * cross(n, e3));
        const double mu = C.mh * C.mp / (C.mh + C.mp);
        const V3 Lo = mu * cross(r, u);
        const double fmh = 2. + 1.5 * C.mp / C.mh, fmp = 2. + 1.5 * C.mh / C.mp;
        o.Nh += fa * (fmh * cross(Lo, Ls) - cross(Lp, Ls) + (3. * dot(n, Lp)) * cross(n, Ls));
        o.Np += fa * (fmp * cross(Lo, Lp) - cross(Ls, Lp) + (3. * dot(n, Ls)) * cross(n, Lp));
    }
    return o;
}

// Vector-Jacobian product of pair_forward: given the cotangents of the
// outputs, accumulate the cotangents of the inputs (into Ib) and of the
// parameters (into Pb).
POS_HD void pair_vjp(const PairConst &C, const PairPar &P, const PairIn &I, const PairOut &ob, PairIn &Ib,
                     PairParBar &Pb) {
    const V3 r = I.r, u = I.u;
    const double d2 = dot(r, r), d = sqrt(d2), id = 1. / d;
    const double u2 = dot(u, u);
    const double ru = dot(r, u);
    const double vr = ru * id;
    const double rsh = dot(r, I.Wh), rsp = dot(r, I.Wp);
    const double d5 = d2 * d2 * d, d7 = d5 * d2, d8 = d7 * d;
    // Reciprocals of the linearisation point's distances: every quotient of
    // the reverse part below is a cotangent times a coefficient that depends
    // on the linearisation point only, so a caller that repeats the product
    // at a fixed point (the implicit-midpoint fixed point) evaluates the
    // coefficients once.
    const double id5 = 1. / d5, id7 = 1. / d7, id8 = 1. / d8;
    // adjoints of the shared intermediates
    double d_b = 0., d2_b = 0., id_b = 0., u2_b = 0., vr_b = 0., rsh_b = 0., rsp_b = 0.;
    double d5_b = 0., d7_b = 0., d8_b = 0.;
    V3 r_b = zero3(), u_b = zero3(), Wh_b = zero3(), Wp_b = zero3();
    if (C.tides) {
        const double Ah_c = 4.5 * C.mp * C.mp * pw10(C.Rh), Ap_c = 4.5 * C.mh * C.mh * pw10(C.Rp);
        const double Ah = Ah_c * P.sh, Ap = Ap_c * P.sp;
        const double os = Ah / d7, op = Ap / d7;
        const double Kh = C.mp * C.mp * pw5(C.Rh), Kp = C.mh * C.mh * pw5(C.Rp);
        const double Ck = Kh * P.k2h + Kp * P.k2p;
        const double SA = Ah + Ap;
        const double f3 = -3. * G_K2 * Ck / d7 - 2. * vr * SA / d8;
        const V3 c = cross(r, u);
        const V3 wh = cross(I.Wh, r) - u, wp = cross(I.Wp, r) - u;
        const V3 Fb = ob.F;
        // F_t = f3 id r + os id wh + op id wp
        const double f3_b = id * dot(Fb, r);
        id_b += f3 * dot(Fb, r) + os * dot(Fb, wh) + op * dot(Fb, wp);
        r_b += (f3 * id) * Fb;
        double os_b = id * dot(Fb, wh), op_b = id * dot(Fb, wp);
        const V3 wh_b = (os * id) * Fb, wp_b = (op * id) * Fb;
        // N_tp = -op T_p,  T_p = d Wp - rsp id r - id c
        V3 c_b = zero3();
        {
            const V3 T = d * I.Wp - (rsp * id) * r - id * c;
            op_b += -dot(ob.Np, T);
            const V3 Tb = (-op) * ob.Np;
            d_b += dot(Tb, I.Wp);
            Wp_b += d * Tb;
            rsp_b += -id * dot(Tb, r);
            id_b += -rsp * dot(Tb, r) - dot(Tb, c);
            r_b += (-rsp * id) * Tb;
            c_b += (-id) * Tb;
        }
        {
            const V3 T = d * I.Wh - (rsh * id) * r - id * c;
            os_b += -dot(ob.Nh, T);
            const V3 Tb = (-os) * ob.Nh;
            d_b += dot(Tb, I.Wh);
            Wh_b += d * Tb;
            rsh_b += -id * dot(Tb, r);
            id_b += -rsh * dot(Tb, r) - dot(Tb, c);
            r_b += (-rsh * id) * Tb;
            c_b += (-id) * Tb;
        }
        // wh = Wh x r - u, wp = Wp x r - u, c = r x u
        cross_adj(I.Wh, r, wh_b, Wh_b, r_b);
        u_b -= wh_b;
        cross_adj(I.Wp, r, wp_b, Wp_b, r_b);
        u_b -= wp_b;
        cross_adj(r, u, c_b, r_b, u_b);
        // f3 = -3 K2 Ck / d7 - 2 vr SA / d8
        const double Ck_b = f3_b * (-3. * G_K2 / d7);
        d7_b += f3_b * (3. * G_K2 * Ck * id7 * id7);
        vr_b += f3_b * (-2. * SA / d8);
        const double SA_b = f3_b * (-2. * vr / d8);
        d8_b += f3_b * (2. * vr * SA * id8 * id8);
        // os = Ah / d7, op = Ap / d7
        const double Ah_b = os_b * id7 + SA_b, Ap_b = op_b * id7 + SA_b;
        d7_b += -(os_b * (Ah * id7 * id7)) - op_b * (Ap * id7 * id7);
        Pb.sh += Ah_b * Ah_c;
        Pb.sp += Ap_b * Ap_c;
        Pb.k2h += Ck_b * Kh;
        Pb.k2p += Ck_b * Kp;
    }
    if (C.flat) {
        const double Bh_c = C.mp * pw5(C.Rh) / 6., Bp_c = C.mh * pw5(C.Rp) / 6.;
        const double Bh = Bh_c * P.kfh, Bp = Bp_c * P.kfp;
        const double Wh2 = dot(I.Wh, I.Wh), Wp2 = dot(I.Wp, I.Wp);
        const double ofs = -6. * Bh * rsh / d5, ofp = -6. * Bp * rsp / d5;
        const double radf = -3. * (Bp * Wp2 + Bh * Wh2) / d5 + 15. * (Bh * rsh * rsh + Bp * rsp * rsp) / d7;
        const V3 Fb = ob.F;
        // F_f = radf r + ofp Wp + ofs Wh
        const double radf_b = dot(Fb, r);
        r_b += radf * Fb;
        double ofp_b = dot(Fb, I.Wp), ofs_b = dot(Fb, I.Wh);
        Wp_b += ofp * Fb;
        Wh_b += ofs * Fb;
        // N_fp = -ofp (r x Wp), N_fh = -ofs (r x Wh)
        ofp_b += -dot(ob.Np, cross(r, I.Wp));
        cross_adj(r, I.Wp, (-ofp) * ob.Np, r_b, Wp_b);
        ofs_b += -dot(ob.Nh, cross(r, I.Wh));
        cross_adj(r, I.Wh, (-ofs) * ob.Nh, r_b, Wh_b);
        // radf
        const double Wp2_b = radf_b * (-3. * Bp / d5), Wh2_b = radf_b * (-3. * Bh / d5);
        d5_b += radf_b * (3. * (Bp * Wp2 + Bh * Wh2) * id5 * id5);
        rsh_b += radf_b * (30. * Bh * rsh * id7);
        rsp_b += radf_b * (30. * Bp * rsp * id7);
        d7_b += -radf_b * (15. * (Bh * rsh * rsh + Bp * rsp * rsp) * id7 * id7);
        double Bp_b = radf_b * (-3. * Wp2 / d5 + 15. * rsp * rsp / d7);
        double Bh_b = radf_b * (-3. * Wh2 / d5 + 15. * rsh * rsh / d7);
        // ofs, ofp
        Bh_b += ofs_b * (-6. * rsh / d5);
        rsh_b += ofs_b * (-6. * Bh / d5);
        d5_b += ofs_b * (6. * Bh * rsh * id5 * id5);
        Bp_b += ofp_b * (-6. * rsp / d5);
        rsp_b += ofp_b * (-6. * Bp / d5);
        d5_b += ofp_b * (6. * Bp * rsp * id5 * id5);
        Wh_b += (2. * Wh2_b) * I.Wh;
        Wp_b += (2. * Wp2_b) * I.Wp;
        Pb.kfh += Bh_b * Bh_c;
        Pb.kfp += Bp_b * Bp_c;
    }
    if (C.gr) {
        const double mgh = G_K2 * C.mh, mgp = G_K2 * C.mp, mg = mgh + mgp;
        const double f = mgh * mgp / (mg * mg);
        const double mgc = -mg / (d2 * C2);
        const double mgd = mg / d;
        const double vr2 = vr * vr;
        const double P1 = (1. + 3. * f) * u2 - 2. * (2. + f) * mgd - 1.5 * f * vr2;
        const double P2 = 0.75 * (12. + 29. * f) * mg * mg / d2 + f * (3. - 4. * f) * u2 * u2
                        + 15. / 8. * f * (1. - 3. * f) * vr2 * vr2 - 1.5 * f * (3. - 4. * f) * vr2 * u2
                        - 0.5 * f * (13. - 4. * f) * mgd * u2 - (2. + 25. * f + 2. * f * f) * mgd * vr2;
        const double Q2 = f * (15. + 4. * f) * u2 - (4. + 41. * f + 8. * f * f) * mgd - 3. * f * (3. + 2. * f) * vr2;
        const double rad = mgc * P1 + mgc * P2;
        const double orth = -mgc * 2. * (2. - f) * vr - 0.5 * mgc * vr * Q2;
        const V3 gb = ob.g;
        // g12 = rad id r + orth u
        const double rdot = dot(gb, r);
        const double rad_b = id * rdot;
        id_b += rad * rdot;
        r_b += (rad * id) * gb;
        const double orth_b = dot(gb, u);
        u_b += orth * gb;
        double mgc_b = rad_b * (P1 + P2) + orth_b * (-2. * (2. - f) * vr - 0.5 * vr * Q2);
        vr_b += orth_b * (-mgc * 2. * (2. - f) - 0.5 * mgc * Q2);
        const double P1_b = rad_b * mgc, P2_b = rad_b * mgc, Q2_b = orth_b * (-0.5 * mgc * vr);
        double mgd_b = 0., vr2_b = 0.;
        // P1
        u2_b += P1_b * (1. + 3. * f);
        mgd_b += P1_b * (-2. * (2. + f));
        vr2_b += P1_b * (-1.5 * f);
        // P2
        d2_b += P2_b * (-0.75 * (12. + 29. * f) * mg * mg / (d2 * d2));
        u2_b += P2_b * (f * (3. - 4. * f) * 2. * u2 - 1.5 * f * (3. - 4. * f) * vr2 - 0.5 * f * (13. - 4. * f) * mgd);
        vr2_b += P2_b * (15. / 8. * f * (1. - 3. * f) * 2. * vr2 - 1.5 * f * (3. - 4. * f) * u2
                         - (2. + 25. * f + 2. * f * f) * mgd);
        mgd_b += P2_b * (-0.5 * f * (13. - 4. * f) * u2 - (2. + 25. * f + 2. * f * f) * vr2);
        // Q2
        u2_b += Q2_b * f * (15. + 4. * f);
        mgd_b += Q2_b * (-(4. + 41. * f + 8. * f * f));
        vr2_b += Q2_b * (-3. * f * (3. + 2. * f));
        // spin-orbit
        const double fa = G_K2 / C2;
        const V3 Ls = C.Ih * I.Wh, Lp = C.Ip * I.Wp;
        const double dm = C.mh - C.mp;
        const V3 msf = (dm / C.mp) * Lp - (dm / C.mh) * Ls;
        const V3 S = Ls + Lp;
        const V3 e7 = 7. * S + 3. * msf, e3 = 3. * S + msf, e2 = 2. * S + msf;
        const V3 n = id * r;
        const V3 nu = cross(n, u);
        V3 n_b = zero3(), nu_b = zero3(), e7_b = zero3(), e3_b = zero3(), e2_b = zero3();
        V3 Ls_b = zero3(), Lp_b = zero3(), Lo_b = zero3();
        // a3 = fa (e1 - u x e7 + 3 vr (n x e3)),  e1 = 6 n * nu * e2 (component-wise)
        const V3 e1_b = fa * gb;
        cross_adj(u, e7, (-fa) * gb, u_b, e7_b);
        cross_adj(n, e3, (3. * fa * vr) * gb, n_b, e3_b);
        vr_b += 3. * fa * dot(gb, cross(n, e3));
        n_b += 6. * hadamard(hadamard(e1_b, nu), e2);
        nu_b += 6. * hadamard(hadamard(e1_b, n), e2);
        e2_b += 6. * hadamard(hadamard(e1_b, n), nu);
        cross_adj(n, u, nu_b, n_b, u_b);
        // torques
        const double mu = C.mh * C.mp / (C.mh + C.mp);
        const V3 Lo = mu * cross(r, u);
        const double fmh = 2. + 1.5 * C.mp / C.mh, fmp = 2. + 1.5 * C.mh / C.mp;
        {
            const V3 q = fa * ob.Nh;
            cross_adj(Lo, Ls, fmh * q, Lo_b, Ls_b);
            cross_adj(Lp, Ls, -q, Lp_b, Ls_b);
            const double s1 = dot(n, Lp);
            cross_adj(n, Ls, (3. * s1) * q, n_b, Ls_b);
            const double s1_b = 3. * dot(q, cross(n, Ls));
            n_b += s1_b * Lp;
            Lp_b += s1_b * n;
        }
        {
            const V3 q = fa * ob.Np;
            cross_adj(Lo, Lp, fmp * q, Lo_b, Lp_b);
            cross_adj(Ls, Lp, -q, Ls_b, Lp_b);
            const double s2 = dot(n, Ls);
            cross_adj(n, Lp, (3. * s2) * q, n_b, Lp_b);
            const double s2_b = 3. * dot(q, cross(n, Lp));
            n_b += s2_b * Ls;
            Ls_b += s2_b * n;
        }
        cross_adj(r, u, mu * Lo_b, r_b, u_b);
        const V3 S_b = 7. * e7_b + 3. * e3_b + 2. * e2_b;
        const V3 msf_b = 3. * e7_b + e3_b + e2_b;
        Lp_b += (dm / C.mp) * msf_b + S_b;
        Ls_b += (-(dm / C.mh)) * msf_b + S_b;
        Wh_b += C.Ih * Ls_b;
        Wp_b += C.Ip * Lp_b;
        // n = id r
        id_b += dot(n_b, r);
        r_b += id * n_b;
        // shared GR scalars
        vr_b += 2. * vr * vr2_b;
        d_b += -mgd_b * (mg * id * id);
        d2_b += mgc_b * (mg / (d2 * d2 * C2));
    }
    // shared intermediates
    d_b += 5. * d5 / d * d5_b + 7. * d7 / d * d7_b + 8. * d8 / d * d8_b;
    // vr = ru id
    const double ru_b = vr_b * id;
    id_b += vr_b * ru;
    r_b += ru_b * u;
    u_b += ru_b * r;
    u_b += (2. * u2_b) * u;
    r_b += rsh_b * I.Wh;
    Wh_b += rsh_b * r;
    r_b += rsp_b * I.Wp;
    Wp_b += rsp_b * r;
    d_b += -id_b * (id * id);
    d2_b += d_b * (0.5 * id);
    r_b += (2. * d2_b) * r;
    Ib.r += r_b;
    Ib.u += u_b;
    Ib.Wh += Wh_b;
    Ib.Wp += Wp_b;
}

// ---------------------------------------------------------------------------
// Universal-variable Kepler drift with the universal anomaly X given
// ---------------------------------------------------------------------------
// Stiefel functions G_k(beta, X) = X^k c_k(beta X^2), k = 0..5.
POS_HD void stiefel_g05(double beta, double X, double g[6]) {
    const double z = beta * X * X;
    double c[6];
    if (fabs(z) < 1.) {
        // c_k(z) = sum_j (-z)^j / (k + 2j)! for k = 4, 5 (12 terms: the
        // remainder is below 1/28! for |z| < 1), then c_k = 1/k! - z c_{k+2}
        // downwards, which only multiplies by |z| < 1.
        double t4 = 1. / 24., t5 = 1. / 120., s4 = 0., s5 = 0.;
        for (int j = 0; j < 12; ++j) {
            s4 += t4;
            s5 += t5;
            t4 *= -z / ((double)(5 + 2 * j) * (double)(6 + 2 * j));
            t5 *= -z / ((double)(6 + 2 * j) * (double)(7 + 2 * j));
        }
        c[4] = s4;
        c[5] = s5;
        c[3] = 1. / 6. - z * c[5];
        c[2] = 0.5 - z * c[4];
        c[1] = 1. - z * c[3];
        c[0] = 1. - z * c[2];
    } else if (z > 0.) {
        const double s = sqrt(z);
        c[0] = cos(s);
        c[1] = sin(s) / s;
        c[2] = (1. - c[0]) / z;
        c[3] = (1. - c[1]) / z;
        c[4] = (0.5 - c[2]) / z;
        c[5] = (1. / 6. - c[3]) / z;
    } else {
        const double s = sqrt(-z);
        c[0] = cosh(s);
        c[1] = sinh(s) / s;
        c[2] = (1. - c[0]) / z;
        c[3] = (1. - c[1]) / z;
        c[4] = (0.5 - c[2]) / z;
        c[5] = (1. / 6. - c[3]) / z;
    }
    double xp = 1.;
    for (int k = 0; k < 6; ++k) {
        g[k] = xp * c[k];
        xp *= X;
    }
}

// One drift of a body (position x0, velocity v0 relative to a centre of
// gravitational parameter mu) over dt, for a given converged X.
POS_HD void kepler_forward_x(double mu, double dt, double X, V3 x0, V3 v0, V3 &x1, V3 &v1) {
    const double r0 = sqrt(dot(x0, x0));
    const double v2 = dot(v0, v0);
    const double beta = 2. * mu / r0 - v2;
    const double eta0 = dot(x0, v0);
    const double zeta0 = mu - beta * r0;
    double g[6];
    stiefel_g05(beta, X, g);
    const double r = r0 + eta0 * g[1] + zeta0 * g[2];
    const double f = -mu * g[2] / r0, gg = dt - mu * g[3];
    const double fd = -mu * g[1] / (r0 * r), gd = -mu * g[2] / r;
    x1 = x0 + f * x0 + gg * v0;
    v1 = v0 + fd * x0 + gd * v0;
}

// Residual of the universal Kepler equation, r0 X + eta0 G2 + zeta0 G3 - dt.
POS_HD double kepler_residual(double mu, double dt, double X, V3 x0, V3 v0) {
    const double r0 = sqrt(dot(x0, x0));
    const double beta = 2. * mu / r0 - dot(v0, v0);
    double g[6];
    stiefel_g05(beta, X, g);
    return r0 * X + dot(x0, v0) * g[2] + (mu - beta * r0) * g[3] - dt;
}

// Vector-Jacobian product of the drift, X being the solution of the
// universal Kepler equation (implicit differentiation): accumulates the
// cotangents of x0, v0 from those of x1, v1.
POS_HD void kepler_vjp(double mu, double dt, double X, V3 x0, V3 v0, V3 x1b, V3 v1b, V3 &x0b, V3 &v0b) {
    const double r0 = sqrt(dot(x0, x0));
    const double v2 = dot(v0, v0);
    const double beta = 2. * mu / r0 - v2;
    const double eta0 = dot(x0, v0);
    const double zeta0 = mu - beta * r0;
    double g[6];
    stiefel_g05(beta, X, g);
    const double r = r0 + eta0 * g[1] + zeta0 * g[2];
    const double f = -mu * g[2] / r0, gg = dt - mu * g[3];
    const double fd = -mu * g[1] / (r0 * r), gd = -mu * g[2] / r;
    // dG_k/dbeta = (k G_{k+2} - X G_{k+1}) / 2, dG_k/dX = G_{k-1}
    const double g1b = (1. * g[3] - X * g[2]) * 0.5;
    const double g2b = (2. * g[4] - X * g[3]) * 0.5;
    const double g3b = (3. * g[5] - X * g[4]) * 0.5;
    // x1 = x0 + f x0 + gg v0 ; v1 = v0 + fd x0 + gd v0
    V3 xb = (1. + f) * x1b + fd * v1b;
    V3 vb = gg * x1b + (1. + gd) * v1b;
    const double f_b = dot(x1b, x0), gg_b = dot(x1b, v0);
    const double fd_b = dot(v1b, x0), gd_b = dot(v1b, v0);
    double G1_b = 0., G2_b = 0., G3_b = 0., r0_b = 0., r_b = 0.;
    G2_b += -mu * f_b / r0;
    r0_b += mu * g[2] * f_b / (r0 * r0);
    G3_b += -mu * gg_b;
    G1_b += -mu * fd_b / (r0 * r);
    r0_b += mu * g[1] * fd_b / (r0 * r0 * r);
    r_b += mu * g[1] * fd_b / (r0 * r * r);
    G2_b += -mu * gd_b / r;
    r_b += mu * g[2] * gd_b / (r * r);
    // r = r0 + eta0 G1 + zeta0 G2
    r0_b += r_b;
    double eta0_b = r_b * g[1], zeta0_b = r_b * g[2];
    G1_b += r_b * eta0;
    G2_b += r_b * zeta0;
    // G_k(beta, X)
    double beta_b = G1_b * g1b + G2_b * g2b + G3_b * g3b;
    const double X_b = G1_b * g[0] + G2_b * g[1] + G3_b * g[2];
    // X(r0, eta0, zeta0, beta): r0 X + eta0 G2 + zeta0 G3 = dt, dF/dX = r
    r0_b += -X_b * X / r;
    eta0_b += -X_b * g[2] / r;
    zeta0_b += -X_b * g[3] / r;
    beta_b += -X_b * (eta0 * g2b + zeta0 * g3b) / r;
    // zeta0 = mu - beta r0
    beta_b += -zeta0_b * r0;
    r0_b += -zeta0_b * beta;
    // eta0 = x0 . v0
    xb += eta0_b * v0;
    vb += eta0_b * x0;
    // beta = 2 mu / r0 - v2
    r0_b += -2. * mu * beta_b / (r0 * r0);
    vb += (-2. * beta_b) * v0;
    xb += (r0_b / r0) * x0;
    x0b += xb;
    v0b += vb;
}

// ---------------------------------------------------------------------------
// Newtonian pair (planet-planet interaction of the democratic-heliocentric
// splitting): a_i += -G m_j (x_i - x_j) / |x_i - x_j|^3.  Given the
// cotangent ab of a_i, returns the cotangent of d = x_i - x_j.
// ---------------------------------------------------------------------------
POS_HD V3 gravity_pair_vjp(double mj, V3 d, V3 ab) {
    const double d2 = dot(d, d);
    const double id = 1. / sqrt(d2);
    const double id3 = id * id * id;
    const double id5 = id3 * id * id;
    return (-G_K2 * mj) * (id3 * ab - (3. * id5 * dot(d, ab)) * d);
}

}  // namespace posadj
