// Pointwise MHD face physics for the O2 operator, usable on host and device.
// Each function cites the reference arithmetic it restates (src/ = the
// HelioCubed reference repository). States are 8 components
// [rho, v(3), p, B(3)] in Gaussian units (magnetic pressure B^2 / 8 pi).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <type_traits>

#define HC_HD __host__ __device__ __forceinline__
// Scheduling fence of device code: instructions are not moved across it, which
// bounds the live registers of long unrolled sequences (no-op on the host).
#if defined(__HIP_DEVICE_COMPILE__)
#define HC_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define HC_SCHED_FENCE() ((void)0)
#endif

// Plasma beta below which mixed precision evaluates the Roe eigen-system in FP64.
#ifndef HC_MIXED_FP64_BETA
#define HC_MIXED_FP64_BETA 0.5
#endif
// Relative density or pressure jump across a face above which mixed precision
// evaluates the Roe eigen-system in FP64 (0 disables the test).
#ifndef HC_MIXED_FP64_JUMP
#define HC_MIXED_FP64_JUMP 0.0
#endif

namespace hc
{
constexpr double kRhoFloor = 1.67262192e-24/100.;      // RHO00 = c_MP/100
constexpr double kInv4Pi = 0.079577471545947667884441881686257181;
constexpr double kInv8Pi = 0.039788735772973833942220940843128591;

// Conserved (times inverse Jacobian) -> Cartesian primitive with the density
// floor and non-negative pressure, src/MHD_StateConversion.H:11-32.
template <typename R>
HC_HD void conservedToPrimitive(R w[8], const R u0[8], R invJ, R gamma)
{
    R u[8];
    for (int c = 0; c < 8; ++c) u[c] = u0[c]*invJ;
    const R rho = u[0] < R(kRhoFloor) ? R(kRhoFloor) : u[0];
    R kinetic = 0, magnetic = 0;
    w[0] = rho;
    for (int d = 0; d < 3; ++d)
    {
        w[1+d] = u[1+d]/rho;
        w[5+d] = u[5+d];
        kinetic += R(0.5)*rho*w[1+d]*w[1+d];
        magnetic += u[5+d]*u[5+d]/R(8*3.14159265358979323846);
    }
    const R pressure = (gamma-R(1))*(u[4]-kinetic-magnetic);
    w[4] = pressure < R(0) ? R(0) : pressure;
}

// PLM slopes in mapped coordinates (limiter 0 none, 1/2 minmod, 3 MC) with
// the complete-state fallback, src/BoxOp_SecondOrder.H:7-37. lo is the state
// on the low side of the face (extrapolated from cell m), hi from cell c.
template <typename R>
HC_HD void plmFaceStates(R lo[8], R hi[8], const R mm[8], const R m[8],
                         const R c[8], const R p[8], int limiter)
{
    for (int k = 0; k < 8; ++k)
    {
        const R a = m[k]-mm[k], b = c[k]-m[k], d = p[k]-c[k];
        R sl = R(0.5)*(a+b), sr = R(0.5)*(b+d);
        if (limiter != 0)
        {
            sl = a*b <= R(0) ? R(0) : (a > R(0) ? fmin(a,b) : fmax(a,b));
            sr = b*d <= R(0) ? R(0) : (b > R(0) ? fmin(b,d) : fmax(b,d));
            if (limiter == 3)
            {
                if (a*b > R(0)) sl = (a > R(0) ? R(1) : R(-1))*fmin(R(0.5)*fabs(a+b), R(2)*fmin(fabs(a),fabs(b)));
                if (b*d > R(0)) sr = (b > R(0) ? R(1) : R(-1))*fmin(R(0.5)*fabs(b+d), R(2)*fmin(fabs(b),fabs(d)));
            }
        }
        lo[k] = m[k]+R(0.5)*sl;
        hi[k] = c[k]-R(0.5)*sr;
    }
    if (!(lo[0] >= R(kRhoFloor) && lo[4] >= R(0))) for (int k = 0; k < 8; ++k) lo[k] = m[k];
    if (!(hi[0] >= R(kRhoFloor) && hi[4] >= R(0))) for (int k = 0; k < 8; ++k) hi[k] = c[k];
}

// Same reconstruction with the slope limiting done in precision S: the
// one-sided differences are formed in FP64 and rounded to S, the limited
// slope is added to the FP64 cell value. S = double reproduces plmFaceStates.
template <typename S>
HC_HD void plmFaceStatesMixed(double lo[8], double hi[8], const double mm[8], const double m[8],
                              const double c[8], const double p[8], int limiter)
{
    for (int k = 0; k < 8; ++k)
    {
        const S a = S(m[k]-mm[k]), b = S(c[k]-m[k]), d = S(p[k]-c[k]);
        S sl = S(0.5)*(a+b), sr = S(0.5)*(b+d);
        if (limiter != 0)
        {
            sl = a*b <= S(0) ? S(0) : (a > S(0) ? fmin(a,b) : fmax(a,b));
            sr = b*d <= S(0) ? S(0) : (b > S(0) ? fmin(b,d) : fmax(b,d));
            if (limiter == 3)
            {
                if (a*b > S(0)) sl = (a > S(0) ? S(1) : S(-1))*fmin(S(0.5)*fabs(a+b), S(2)*fmin(fabs(a),fabs(b)));
                if (b*d > S(0)) sr = (b > S(0) ? S(1) : S(-1))*fmin(S(0.5)*fabs(b+d), S(2)*fmin(fabs(b),fabs(d)));
            }
        }
        lo[k] = m[k]+0.5*double(sl);
        hi[k] = c[k]-0.5*double(sr);
    }
    if (!(lo[0] >= kRhoFloor && lo[4] >= 0.0)) for (int k = 0; k < 8; ++k) lo[k] = m[k];
    if (!(hi[0] >= kRhoFloor && hi[4] >= 0.0)) for (int k = 0; k < 8; ++k) hi[k] = c[k];
}

// Only the state on one side of the face, with that side's fallback.
template <typename R>
HC_HD void plmLowSide(R lo[8], const R mm[8], const R m[8], const R c[8], int limiter)
{
    for (int k = 0; k < 8; ++k)
    {
        const R a = m[k]-mm[k], b = c[k]-m[k];
        R sl = R(0.5)*(a+b);
        if (limiter != 0)
        {
            sl = a*b <= R(0) ? R(0) : (a > R(0) ? fmin(a,b) : fmax(a,b));
            if (limiter == 3 && a*b > R(0))
                sl = (a > R(0) ? R(1) : R(-1))*fmin(R(0.5)*fabs(a+b), R(2)*fmin(fabs(a),fabs(b)));
        }
        lo[k] = m[k]+R(0.5)*sl;
    }
    if (!(lo[0] >= R(kRhoFloor) && lo[4] >= R(0))) for (int k = 0; k < 8; ++k) lo[k] = m[k];
}

template <typename R>
HC_HD void plmHighSide(R hi[8], const R m[8], const R c[8], const R p[8], int limiter)
{
    for (int k = 0; k < 8; ++k)
    {
        const R b = c[k]-m[k], d = p[k]-c[k];
        R sr = R(0.5)*(b+d);
        if (limiter != 0)
        {
            sr = b*d <= R(0) ? R(0) : (b > R(0) ? fmin(b,d) : fmax(b,d));
            if (limiter == 3 && b*d > R(0))
                sr = (b > R(0) ? R(1) : R(-1))*fmin(R(0.5)*fabs(b+d), R(2)*fmin(fabs(b),fabs(d)));
        }
        hi[k] = c[k]-R(0.5)*sr;
    }
    if (!(hi[0] >= R(kRhoFloor) && hi[4] >= R(0))) for (int k = 0; k < 8; ++k) hi[k] = c[k];
}

// Local Lax-Friedrichs dissipation speed * (U_R - U_L) in the face frame,
// src/MHD_Roe8Wave.H:604-629. Conserved states, their difference and the
// product are formed in FP64; the wave speed is evaluated in S.
template <typename S>
HC_HD void rusanovDissipation(double out[8], const double l[8], const double r[8], double gamma)
{
    double cons[2][8];
    S speed = 0;
    for (int s = 0; s < 2; ++s)
    {
        const double* w = s == 0 ? l : r;
        double v2 = 0, b2 = 0;
        cons[s][0] = w[0];
        for (int d = 0; d < 3; ++d)
        {
            v2 += w[1+d]*w[1+d];
            b2 += w[5+d]*w[5+d];
            cons[s][1+d] = w[0]*w[1+d];
            cons[s][5+d] = w[5+d];
        }
        cons[s][4] = w[4]/(gamma-1.0)+0.5*w[0]*v2+0.5*kInv4Pi*b2;
        const S wn = S(w[1]), g = S(gamma), p = S(w[4]), bb = S(b2), rho = S(w[0]);
        speed = fmax(speed, fabs(wn)+sqrt((g*p+S(kInv4Pi)*bb)/rho));
    }
    for (int c = 0; c < 8; ++c) out[c] = double(speed)*(cons[1][c]-cons[0][c]);
}

// Mixed precision only: FP64 square root, reciprocal square root and
// quotient from the FP32 hardware estimate and one FP64 Newton step. The
// estimate's relative error e (below 2^-22) becomes about e^2/2 + 1e-16, so
// the results lie within ~1e-13 of the correctly rounded ones. Arguments
// outside [2^-100, 2^100] (zero, FP32-subnormal or FP32-infinite estimates,
// non-finite values) take the exact FP64 form.
HC_HD bool fp32Seedable(double x) { return x >= 7.888609052210118e-31 && x <= 1.2676506002282294e30; }
HC_HD double rsqrtN(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (!fp32Seedable(x)) return 1.0/sqrt(x);
    const double r = double(__builtin_amdgcn_rsqf(float(x)));
    // r (1 + (1 - x r^2)/2)
    return r + (0.5*r)*(1.0 - x*r*r);
#else
    return 1.0/sqrt(x);
#endif
}
HC_HD double sqrtN(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if (!fp32Seedable(x)) return sqrt(x);
    const double r = double(__builtin_amdgcn_rsqf(float(x)));
    const double y = x*r;
    // y + (r/2) (x - y^2)
    return y + (0.5*r)*(x - y*y);
#else
    return sqrt(x);
#endif
}
HC_HD double quotN(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double m = fabs(b);
    if (!fp32Seedable(m)) return a/b;
    const double r = double(__builtin_amdgcn_rcpf(float(b)));
    const double q = a*r;
    // q + r (a - b q)
    return q + r*(a - b*q);
#else
    return a/b;
#endif
}

// Quotient a / b in the eigen-system precision S: an FP64 division, or in
// FP32 the product with the hardware reciprocal (within 2 ulp, the size of
// the FP32 roundings the eigen-system already carries).
HC_HD double quotS(double a, double b) { return a/b; }
HC_HD float quotS(float a, float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return a*__builtin_amdgcn_rcpf(b);
#else
    return a/b;
#endif
}

template <typename S>
HC_HD S entropyFixedSpeed(S lambda, S del)
{
    const S z = fabs(lambda);
    return z >= del ? z : quotS(S(0.5)*(z*z+del*del), del);
}

// Roe 8-wave dissipation D = |A| (U_R - U_L), arithmetic averages, eight-wave
// entropy fix and the wave-fan admissibility walk, src/MHD_Roe8Wave.H:16-600.
// Returns 0 when the caller must fall back to Rusanov, 1 with the result in
// out. The averages, the conserved jump, the walked fan states and the result
// are FP64; the eigen-system and the projections are evaluated in S (S =
// double: all FP64). The result stays FP64 in both cases, so a face evaluated
// in FP64 by the mixed path carries no FP32 rounding. With DEFER, a face the
// mixed path must evaluate in FP64 returns 2 (nothing written) for the caller
// to evaluate later, which keeps the FP64 solver out of the caller's code.
template <typename S, bool DEFER = false>
HC_HD int roe8WaveDissipation(double out[8], const double WL[8], const double WR[8], double gamma_d, double efix_d)
{
    const S rs2 = S(0.707106781186547524400844362104849039);
    if (!(WL[0] > 0.0 && WR[0] > 0.0 && WL[4] > 0.0 && WR[4] > 0.0)) return 0;
    const double hgd = gamma_d-1.0;
    // mixed precision (S = float): the FP64 divisions by gamma-1 and by the
    // average density become products with their reciprocals
    constexpr bool FAST = !std::is_same<S, double>::value;
    const double ihgd = FAST ? quotN(1.0, hgd) : 0.0;
    auto byH = [&](double x) { return FAST ? x*ihgd : x/hgd; };
    const double rL = WL[0], rR = WR[0];
    const double mL0 = rL*WL[1], mL1 = rL*WL[2], mL2 = rL*WL[3];
    const double mR0 = rR*WR[1], mR1 = rR*WR[2], mR2 = rR*WR[3];
    const double bbL = kInv8Pi*(WL[5]*WL[5]+WL[6]*WL[6]+WL[7]*WL[7]);
    const double keL = 0.5*(mL0*WL[1]+mL1*WL[2]+mL2*WL[3]);
    const double EL = byH(WL[4])+keL+bbL, a2L = FAST ? quotN(gamma_d*WL[4], rL) : gamma_d*WL[4]/rL;
    const double bbR = kInv8Pi*(WR[5]*WR[5]+WR[6]*WR[6]+WR[7]*WR[7]);
    const double keR = 0.5*(mR0*WR[1]+mR1*WR[2]+mR2*WR[3]);
    const double ER = byH(WR[4])+keR+bbR, a2R = FAST ? quotN(gamma_d*WR[4], rR) : gamma_d*WR[4]/rR;
    const double Rhod = 0.5*(rL+rR);
    const double iRhod = FAST ? quotN(1.0, Rhod) : 0.0;
    auto byR = [&](double x) { return FAST ? x*iRhod : x/Rhod; };
    const double Ud = 0.5*(WL[1]+WR[1]), Vd = 0.5*(WL[2]+WR[2]), Wd = 0.5*(WL[3]+WR[3]);
    const double BXd = 0.5*(WL[5]+WR[5]), BYd = 0.5*(WL[6]+WR[6]), BZd = 0.5*(WL[7]+WR[7]);
    const double Pd = 0.5*((WL[4]+bbL)+(WR[4]+bbR));
    const double BX2d = BXd*BXd, BYZ2d = BYd*BYd+BZd*BZd;
    const double BxBHd = kInv8Pi*(BX2d+BYZ2d);
    const double kkd = 0.5*(Ud*Ud+Vd*Vd+Wd*Wd);
    double ccd = byR(gamma_d*(Pd-BxBHd));
    const double ccmi = a2L > a2R ? a2R : a2L, ccma = a2L > a2R ? a2L : a2R;
    if (ccd < ccmi) ccd = ccmi;
    if (ccd > ccma) ccd = ccma;
    if (!(ccd > 0.0) || !isfinite(ccd)) return 0;
    if constexpr (!std::is_same<S, double>::value)
    {
        // Below plasma beta 1/2 (of either side or of the average state) the
        // FP32 eigen-system loses digits to cancellation in the fast/slow
        // recombination: evaluate such faces in FP64.
        constexpr double betaFp64 = HC_MIXED_FP64_BETA, jumpFp64 = HC_MIXED_FP64_JUMP;
        const bool jump = jumpFp64 > 0.0 && (fabs(rL-rR) > jumpFp64*fmin(rL, rR) ||
                                             fabs(WL[4]-WR[4]) > jumpFp64*fmin(WL[4], WR[4]));
        if (jump || ccd*Rhod < betaFp64*gamma_d*BxBHd || WL[4] < betaFp64*bbL || WR[4] < betaFp64*bbR)
        {
            if constexpr (DEFER) return 2;
            else return roe8WaveDissipation<double>(out, WL, WR, gamma_d, efix_d);
        }
    }
    // Jump in primitive-like variables (left minus right), FP64.
    double dud[8] = {rL-rR, mL0-mR0, mL1-mR1, mL2-mR2, EL-ER, WL[5]-WR[5], WL[6]-WR[6], WL[7]-WR[7]};
    dud[4] = -hgd*(-kkd*dud[0]+Ud*dud[1]+Vd*dud[2]+Wd*dud[3]-dud[4]+(BXd*dud[5]+BYd*dud[6]+BZd*dud[7])*kInv4Pi);
    dud[1] = byR(-Ud*dud[0]+dud[1]);
    dud[2] = byR(-Vd*dud[0]+dud[2]);
    dud[3] = byR(-Wd*dud[0]+dud[3]);
    // Wave speeds and eigenvector coefficients: FP64 (the slow speed and the
    // fast/slow mixing factors cancel catastrophically at low beta); the
    // eigenvector entries, projections and products below are in S.
    const double rs2d = 0.707106781186547524400844362104849039, smallBd = 1.0e-9;
    // mixed precision: square roots and quotients from FP32 estimates and a
    // Newton step (within ~1e-13)
    auto sqrtS = [](double x) { return FAST ? sqrtN(x) : sqrt(x); };
    auto quotD = [](double a, double b) { return FAST ? quotN(a, b) : a/b; };
    const double bbd = byR(kInv4Pi*BX2d), bd = sqrtS(bbd);
    const double cd = sqrtS(ccd);
    double hyyd, hzzd, aafd, aasd, afd, asd;
    if (BYZ2d < smallBd*smallBd)
    {
        hyyd = rs2d; hzzd = rs2d;
        if (bbd > ccd) { aafd = bbd; aasd = ccd; afd = bd; asd = cd; }
        else { aafd = ccd; aasd = bbd; afd = cd; asd = bd; }
    }
    else
    {
        if constexpr (FAST)
        {
            const double ihyz = rsqrtN(BYZ2d);
            hyyd = BYd*ihyz; hzzd = BZd*ihyz;
        }
        else
        {
            const double hyz = sqrt(BYZ2d);
            hyyd = BYd/hyz; hzzd = BZd/hyz;
        }
        const double szb = ccd+byR(BxBHd+BxBHd);
        const double hh = sqrtS(fmax(0.0, szb*szb-4.0*ccd*bbd));
        aasd = 0.5*(szb-hh);
        if (aasd <= 0.0) { aasd = 0; asd = 0; }
        else if (aasd >= ccd) { aasd = ccd; asd = cd; }
        else asd = sqrtS(aasd);
        aafd = 0.5*(szb+hh);
        if (aafd <= ccd) { aafd = ccd; afd = cd; }
        else afd = sqrtS(aafd);
    }
    const double deld = efix_d*(fabs(Ud)+fabs(Vd)+fabs(Wd)+afd);
    const double dfsd = aafd-aasd;
    double alfd, alsd;
    if (dfsd < 1.0e-8) { alfd = 1; alsd = 0; }
    else if ((ccd-aasd) <= 0.0) { alfd = 0; alsd = 1; }
    else if ((aafd-ccd) <= 0.0) { alfd = 1; alsd = 0; }
    else { alfd = sqrtS(quotD(ccd-aasd, dfsd)); alsd = sqrtS(quotD(aafd-ccd, dfsd)); }
    const S Rho = S(Rhod), BX = S(BXd);
    const S cc = S(ccd), c = S(cd), hyy = S(hyyd), hzz = S(hzzd), af = S(afd), as = S(asd);
    const S alf = S(alfd), als = S(alsd);
    const S lam[8] = {S(Ud), S(Ud), S(Ud+bd), S(Ud-bd), S(Ud+afd), S(Ud-afd), S(Ud+asd), S(Ud-asd)};
    const S del = S(deld);
    const S sih = BX >= S(0) ? S(1) : S(-1);
    S du[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) du[q] = S(dud[q]);
    const S sqPiRho = sqrt(S(3.14159265358979323846264338327950288)*Rho);
    const S sq2PiRho = sqrt(S(6.28318530717958647692528676655900576)*Rho);
    // Eigenvector entries (src/MHD_Roe8Wave.H:356-498) from a few scalars;
    // rows of the left matrix and columns of the right matrix are formed on
    // the fly so no 8x8 matrix is held in registers. Non-zero terms are summed
    // in the same order as the dense products.
    const S aa = alf*af, sa = als*as, rc = Rho*cc;
    const S fy = S(2)*sqPiRho*c*hyy, fz = S(2)*sqPiRho*c*hzz;
    const S ih = quotS(S(0.5)*sih, sq2PiRho), qy = quotS(S(0.25)*hyy, c*sqPiRho), qz = quotS(S(0.25)*hzz, c*sqPiRho);
    const S hzr = -hzz*rs2, hyr = hyy*rs2;
    const S zq = hzz*sq2PiRho*sih, yq = hyy*sq2PiRho*sih;
    const S ayh = aa*hyy*sih, azh = aa*hzz*sih, syh = sa*hyy*sih, szh = sa*hzz*sih;
    auto rcol = [&](int w, S col[8]) {
        for (int q = 0; q < 8; ++q) col[q] = 0;
        switch (w)
        {
        case 0: col[0] = S(1); break;
        case 1: col[5] = S(1); break;
        case 2: col[2] = hzr; col[3] = hyr; col[6] = zq; col[7] = -yq; break;
        case 3: col[2] = hzr; col[3] = hyr; col[6] = -zq; col[7] = yq; break;
        case 4: col[0] = Rho*alf; col[1] = aa; col[2] = -syh; col[3] = -szh; col[4] = alf*rc; col[6] = als*fy; col[7] = als*fz; break;
        case 5: col[0] = Rho*alf; col[1] = -aa; col[2] = syh; col[3] = szh; col[4] = alf*rc; col[6] = als*fy; col[7] = als*fz; break;
        case 6: col[0] = Rho*als; col[1] = sa; col[2] = ayh; col[3] = azh; col[4] = als*rc; col[6] = -alf*fy; col[7] = -alf*fz; break;
        default: col[0] = Rho*als; col[1] = -sa; col[2] = -ayh; col[3] = -azh; col[4] = als*rc; col[6] = -alf*fy; col[7] = -alf*fz; break;
        }
    };
    // Projections of the jump on the left eigenvectors.
    S Ldu[8];
    {
        const S a2 = quotS(S(0.5)*aa, cc), s2 = quotS(S(0.5)*sa, cc);
        const S sy2 = quotS(S(0.5)*sa*hyy*sih, cc), sz2 = quotS(S(0.5)*sa*hzz*sih, cc);
        const S ay2 = quotS(S(0.5)*aa*hyy*sih, cc), az2 = quotS(S(0.5)*aa*hzz*sih, cc);
        const S f4 = quotS(S(0.5)*alf, cc*Rho), f6 = quotS(S(0.5)*als, Rho*cc);
        Ldu[0] = du[0] + quotS(-S(1), cc)*du[4];
        Ldu[1] = du[5];
        Ldu[2] = hzr*du[2] + hyr*du[3] + (hzz*ih)*du[6] + (-hyy*ih)*du[7];
        Ldu[3] = hzr*du[2] + hyr*du[3] + (-hzz*ih)*du[6] + (hyy*ih)*du[7];
        Ldu[4] = a2*du[1] + (-sy2)*du[2] + (-sz2)*du[3] + f4*du[4] + (als*qy)*du[6] + (als*qz)*du[7];
        Ldu[5] = (-a2)*du[1] + sy2*du[2] + sz2*du[3] + f4*du[4] + (als*qy)*du[6] + (als*qz)*du[7];
        Ldu[6] = s2*du[1] + ay2*du[2] + az2*du[3] + f6*du[4] + (-alf*qy)*du[6] + (-alf*qz)*du[7];
        Ldu[7] = (-s2)*du[1] + (-ay2)*du[2] + (-az2)*du[3] + f6*du[4] + (-alf*qy)*du[6] + (-alf*qz)*du[7];
    }
    // Walk the conservative wave fan in speed order; reject a fan whose
    // intermediate state has non-positive density or negative thermal energy.
    {
        const int order[8] = {5,3,7,0,1,6,2,4};
        double st[8] = {rL, mL0, mL1, mL2, EL, WL[5], WL[6], WL[7]};
#pragma unroll
        for (int wv = 0; wv < 7; ++wv)
        {
            const int k = order[wv];
            const S amp = -Ldu[k];
            S col[8];
            rcol(k, col);
            double dl[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) dl[q] = double(amp*col[q]);
            dl[4] = kkd*dl[0]+Rhod*(Ud*dl[1]+Vd*dl[2]+Wd*dl[3])+byH(dl[4])+kInv4Pi*(BXd*dl[5]+BYd*dl[6]+BZd*dl[7]);
            dl[1] = Ud*dl[0]+Rhod*dl[1];
            dl[2] = Vd*dl[0]+Rhod*dl[2];
            dl[3] = Wd*dl[0]+Rhod*dl[3];
#pragma unroll
            for (int q = 0; q < 8; ++q) st[q] += dl[q];
            if (lam[order[wv+1]] > lam[k])
            {
                if (!(st[0] > 0.0) || !isfinite(st[0])) return 0;
                if constexpr (FAST)
                {
                    // the same test times the positive density: no FP64
                    // division (the fan states carry FP32 roundings here)
                    const double th = st[4]*st[0]-0.5*(st[1]*st[1]+st[2]*st[2]+st[3]*st[3])
                        -(kInv8Pi*(st[5]*st[5]+st[6]*st[6]+st[7]*st[7]))*st[0];
                    if (!(th >= 0.0) || !isfinite(th)) return 0;
                }
                else
                {
                    const double th = st[4]-0.5*(st[1]*st[1]+st[2]*st[2]+st[3]*st[3])/st[0]
                        -kInv8Pi*(st[5]*st[5]+st[6]*st[6]+st[7]*st[7]);
                    if (!(th >= 0.0) || !isfinite(th)) return 0;
                }
            }
            HC_SCHED_FENCE();
        }
    }
    S lu[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) lu[w] = entropyFixedSpeed(lam[w], del)*Ldu[w];
    S ra[8];
    ra[0] = lu[0] + (Rho*alf)*lu[4] + (Rho*alf)*lu[5] + (Rho*als)*lu[6] + (Rho*als)*lu[7];
    ra[1] = aa*lu[4] + (-aa)*lu[5] + sa*lu[6] + (-sa)*lu[7];
    ra[2] = hzr*lu[2] + hzr*lu[3] + (-syh)*lu[4] + syh*lu[5] + ayh*lu[6] + (-ayh)*lu[7];
    ra[3] = hyr*lu[2] + hyr*lu[3] + (-szh)*lu[4] + szh*lu[5] + azh*lu[6] + (-azh)*lu[7];
    ra[4] = (alf*rc)*lu[4] + (alf*rc)*lu[5] + (als*rc)*lu[6] + (als*rc)*lu[7];
    ra[5] = lu[1];
    ra[6] = zq*lu[2] + (-zq)*lu[3] + (als*fy)*lu[4] + (als*fy)*lu[5] + (-alf*fy)*lu[6] + (-alf*fy)*lu[7];
    ra[7] = (-yq)*lu[2] + yq*lu[3] + (als*fz)*lu[4] + (als*fz)*lu[5] + (-alf*fz)*lu[6] + (-alf*fz)*lu[7];
    // Back to conserved variables in FP64: the energy row sums kinetic,
    // thermal and magnetic parts that cancel at low beta.
    double rd[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) rd[q] = double(ra[q]);
    rd[4] = rd[0]*kkd+Rhod*(Ud*rd[1]+Vd*rd[2]+Wd*rd[3])+byH(rd[4])+kInv4Pi*(BXd*rd[5]+BYd*rd[6]+BZd*rd[7]);
    rd[1] = Ud*rd[0]+Rhod*rd[1];
    rd[2] = Vd*rd[0]+Rhod*rd[2];
    rd[3] = Wd*rd[0]+Rhod*rd[3];
    bool finite = true;
#pragma unroll
    for (int q = 0; q < 8; ++q) { out[q] = -rd[q]; finite = finite && isfinite(rd[q]); }
    return finite ? 1 : 0;
}

// Face frame: rows are the unit normal and two orthonormal tangents, plus
// the normal's length, src/MHD_SecondOrderFlux.H:11-30.
struct FaceFrame { double f[3][3]; double norm; };

HC_HD FaceFrame makeFaceFrame(const double n[3])
{
    FaceFrame F;
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) F.f[a][b] = 0;
    double nn = 0;
    for (int i = 0; i < 3; ++i) nn += n[i]*n[i];
    F.norm = sqrt(nn);
    for (int i = 0; i < 3; ++i) F.f[0][i] = n[i]/F.norm;
    int axis = 0;
    for (int i = 1; i < 3; ++i) if (fabs(F.f[0][i]) < fabs(F.f[0][axis])) axis = i;
    F.f[1][(axis+1)%3] = F.f[0][(axis+2)%3];
    F.f[1][(axis+2)%3] = -F.f[0][(axis+1)%3];
    double t = 0;
    for (int i = 0; i < 3; ++i) t += F.f[1][i]*F.f[1][i];
    const double st = sqrt(t);
    for (int i = 0; i < 3; ++i) F.f[1][i] /= st;
    for (int i = 0; i < 3; ++i)
    {
        const int j = (i+1)%3, k = (i+2)%3;
        F.f[2][i] = F.f[0][j]*F.f[1][k]-F.f[0][k]*F.f[1][j];
    }
    return F;
}

// Rotated MHD face flux times face area, src/MHD_SecondOrderFlux.H:34-74,
// with the Riemann dissipation evaluated in precision S: frame
// rotation, central flux, area and the back-rotation stay FP64. S = double
// reproduces o2FaceFlux operation for operation.
// RS selects the dissipation at compile time: 1 Rusanov, 2 Roe (Rusanov
// where the Roe fan is inadmissible), 0 by the run-time flag roe.
template <typename S, int RS = 0>
HC_HD void o2FaceFluxMixed(double out[8], const double l[8], const double r[8], const double frame[9],
                           double area, double gamma, double entropy, bool roe)
{
    const double inv4pi = 1.0/(4.0*3.14159265358979323846);
    double loc[2][8];
#pragma unroll
    for (int s = 0; s < 2; ++s)
    {
        const double* w = s == 0 ? l : r;
        double* q = loc[s];
        q[0] = w[0]; q[4] = w[4];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            double v = 0, bsum = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) { v += frame[3*i+j]*w[1+j]; bsum += frame[3*i+j]*w[5+j]; }
            q[1+i] = v; q[5+i] = bsum;
        }
    }
    // the Rusanov fallback of an inadmissible Roe fan is FP64 in both
    // precisions (such faces carry the largest jumps); the Rusanov solver
    // itself (riemann_solver 1) takes its speed in S
    double ds[8];
    if constexpr (RS == 1) rusanovDissipation<S>(ds, loc[0], loc[1], gamma);
    else if constexpr (RS == 2)
    {
        if (!roe8WaveDissipation<S>(ds, loc[0], loc[1], gamma, entropy)) rusanovDissipation<double>(ds, loc[0], loc[1], gamma);
    }
    else if (!roe) rusanovDissipation<S>(ds, loc[0], loc[1], gamma);
    else if (!roe8WaveDissipation<S>(ds, loc[0], loc[1], gamma, entropy))
        rusanovDissipation<double>(ds, loc[0], loc[1], gamma);
    double num[8];
#pragma unroll
    for (int s = 0; s < 2; ++s)
    {
        const double* q = loc[s];
        double v2 = 0, b2 = 0, vb = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) { v2 += q[1+i]*q[1+i]; b2 += q[5+i]*q[5+i]; vb += q[1+i]*q[5+i]; }
        const double pressure = q[4]+0.5*inv4pi*b2;
        const double energy = q[4]/(gamma-1.0)+0.5*q[0]*v2+0.5*inv4pi*b2;
        double f[8];
        f[0] = q[0]*q[1];
        f[4] = (energy+pressure)*q[1]-inv4pi*q[5]*vb;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            f[1+i] = q[0]*q[1]*q[1+i]-inv4pi*q[5]*q[5+i]+(i == 0 ? pressure : 0.0);
            f[5+i] = q[1]*q[5+i]-q[1+i]*q[5];
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) num[c] = s == 0 ? f[c] : num[c]+f[c];
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) num[c] = 0.5*area*(num[c]-ds[c]);
    out[0] = num[0]; out[4] = num[4];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        double v = 0, bsum = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) { v += frame[3*j+i]*num[1+j]; bsum += frame[3*j+i]*num[5+j]; }
        out[1+i] = v; out[5+i] = bsum;
    }
}
} // namespace hc
