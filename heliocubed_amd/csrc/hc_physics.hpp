// Pointwise MHD face physics for the O2 operator, usable on host and device.
// Each function cites the reference arithmetic it restates (src/ = the
// HelioCubed reference repository). States are 8 components
// [rho, v(3), p, B(3)] in Gaussian units (magnetic pressure B^2 / 8 pi).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#define HC_HD __host__ __device__ __forceinline__

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
// src/MHD_Roe8Wave.H:604-629.
template <typename R>
HC_HD void rusanovDissipation(R out[8], const R l[8], const R r[8], R gamma)
{
    R cons[2][8];
    R speed = 0;
    for (int s = 0; s < 2; ++s)
    {
        const R* w = s == 0 ? l : r;
        R v2 = 0, b2 = 0;
        cons[s][0] = w[0];
        for (int d = 0; d < 3; ++d)
        {
            v2 += w[1+d]*w[1+d];
            b2 += w[5+d]*w[5+d];
            cons[s][1+d] = w[0]*w[1+d];
            cons[s][5+d] = w[5+d];
        }
        cons[s][4] = w[4]/(gamma-R(1))+R(0.5)*w[0]*v2+R(0.5)*R(kInv4Pi)*b2;
        speed = fmax(speed, fabs(w[1])+sqrt((gamma*w[4]+R(kInv4Pi)*b2)/w[0]));
    }
    for (int c = 0; c < 8; ++c) out[c] = speed*(cons[1][c]-cons[0][c]);
}

template <typename R>
HC_HD R entropyFixedSpeed(R lambda, R del)
{
    const R z = fabs(lambda);
    return z >= del ? z : R(0.5)*(z*z+del*del)/del;
}

// Roe 8-wave dissipation D = |A| (U_R - U_L), arithmetic averages, eight-wave
// entropy fix and the wave-fan admissibility walk, src/MHD_Roe8Wave.H:16-600.
// Returns false when the caller must fall back to Rusanov.
template <typename R>
HC_HD bool roe8WaveDissipation(R out[8], const R WL[8], const R WR[8], R gamma, R efix)
{
    const R rs2 = R(0.707106781186547524400844362104849039);
    const R smallB = R(1.0e-9);
    if (!(WL[0] > R(0) && WR[0] > R(0) && WL[4] > R(0) && WR[4] > R(0))) return false;
    const R hg = gamma-R(1);
    const R rL = WL[0], rR = WR[0];
    const R mL0 = rL*WL[1], mL1 = rL*WL[2], mL2 = rL*WL[3];
    const R mR0 = rR*WR[1], mR1 = rR*WR[2], mR2 = rR*WR[3];
    const R bbL = R(kInv8Pi)*(WL[5]*WL[5]+WL[6]*WL[6]+WL[7]*WL[7]);
    const R keL = R(0.5)*(mL0*WL[1]+mL1*WL[2]+mL2*WL[3]);
    const R EL = WL[4]/hg+keL+bbL, a2L = gamma*WL[4]/rL;
    const R bbR = R(kInv8Pi)*(WR[5]*WR[5]+WR[6]*WR[6]+WR[7]*WR[7]);
    const R keR = R(0.5)*(mR0*WR[1]+mR1*WR[2]+mR2*WR[3]);
    const R ER = WR[4]/hg+keR+bbR, a2R = gamma*WR[4]/rR;
    const R ccmi = a2L > a2R ? a2R : a2L, ccma = a2L > a2R ? a2L : a2R;
    const R Rho = R(0.5)*(rL+rR);
    const R U = R(0.5)*(WL[1]+WR[1]), V = R(0.5)*(WL[2]+WR[2]), W = R(0.5)*(WL[3]+WR[3]);
    const R BX = R(0.5)*(WL[5]+WR[5]), BY = R(0.5)*(WL[6]+WR[6]), BZ = R(0.5)*(WL[7]+WR[7]);
    const R P = R(0.5)*((WL[4]+bbL)+(WR[4]+bbR));
    const R BX2 = BX*BX, BYZ2 = BY*BY+BZ*BZ;
    const R BxBH = R(kInv8Pi)*(BX2+BYZ2);
    const R kk = R(0.5)*(U*U+V*V+W*W);
    R cc = gamma*(P-BxBH)/Rho;
    const R bb = R(kInv4Pi)*BX2/Rho, b = sqrt(bb);
    if (cc < ccmi) cc = ccmi;
    if (cc > ccma) cc = ccma;
    if (!(cc > R(0)) || !isfinite(cc)) return false;
    const R c = sqrt(cc);
    R hyy, hzz, aaf, aas, af, as;
    if (BYZ2 < smallB*smallB)
    {
        hyy = rs2; hzz = rs2;
        if (bb > cc) { aaf = bb; aas = cc; af = b; as = c; }
        else { aaf = cc; aas = bb; af = c; as = b; }
    }
    else
    {
        const R hyz = sqrt(BYZ2);
        hyy = BY/hyz; hzz = BZ/hyz;
        const R szb = cc+(BxBH+BxBH)/Rho;
        const R hh = sqrt(fmax(R(0), szb*szb-R(4)*cc*bb));
        aas = R(0.5)*(szb-hh);
        if (aas <= R(0)) { aas = 0; as = 0; }
        else if (aas >= cc) { aas = cc; as = c; }
        else as = sqrt(aas);
        aaf = R(0.5)*(szb+hh);
        if (aaf <= cc) { aaf = cc; af = c; }
        else af = sqrt(aaf);
    }
    const R lam[8] = {U, U, U+b, U-b, U+af, U-af, U+as, U-as};
    const R del = efix*(fabs(U)+fabs(V)+fabs(W)+af);
    const R dfs = aaf-aas;
    R alf, als;
    if (dfs < R(1.0e-8)) { alf = 1; als = 0; }
    else if ((cc-aas) <= R(0)) { alf = 0; als = 1; }
    else if ((aaf-cc) <= R(0)) { alf = 1; als = 0; }
    else { alf = sqrt((cc-aas)/dfs); als = sqrt((aaf-cc)/dfs); }
    const R sih = BX >= R(0) ? R(1) : R(-1);
    R du[8] = {rL-rR, mL0-mR0, mL1-mR1, mL2-mR2, EL-ER, WL[5]-WR[5], WL[6]-WR[6], WL[7]-WR[7]};
    du[4] = -hg*(-kk*du[0]+U*du[1]+V*du[2]+W*du[3]-du[4]+(BX*du[5]+BY*du[6]+BZ*du[7])*R(kInv4Pi));
    du[1] = (-U*du[0]+du[1])/Rho;
    du[2] = (-V*du[0]+du[2])/Rho;
    du[3] = (-W*du[0]+du[3])/Rho;
    const R sqPiRho = sqrt(R(3.14159265358979323846264338327950288)*Rho);
    const R sq2PiRho = sqrt(R(6.28318530717958647692528676655900576)*Rho);
    // Right eigenvectors (rows: primitive-like variables; columns: waves).
    const R aa = alf*af, sa = als*as, rc = Rho*cc;
    const R fy = R(2)*sqPiRho*c*hyy, fz = R(2)*sqPiRho*c*hzz;
    const R Rm[8][8] = {
        {R(1), 0, 0, 0, Rho*alf, Rho*alf, Rho*als, Rho*als},
        {0, 0, 0, 0, aa, -aa, sa, -sa},
        {0, 0, -hzz*rs2, -hzz*rs2, -sa*hyy*sih, sa*hyy*sih, aa*hyy*sih, -aa*hyy*sih},
        {0, 0, hyy*rs2, hyy*rs2, -sa*hzz*sih, sa*hzz*sih, aa*hzz*sih, -aa*hzz*sih},
        {0, 0, 0, 0, alf*rc, alf*rc, als*rc, als*rc},
        {0, R(1), 0, 0, 0, 0, 0, 0},
        {0, 0, hzz*sq2PiRho*sih, -hzz*sq2PiRho*sih, als*fy, als*fy, -alf*fy, -alf*fy},
        {0, 0, -hyy*sq2PiRho*sih, hyy*sq2PiRho*sih, als*fz, als*fz, -alf*fz, -alf*fz}};
    // Left eigenvectors (rows: waves).
    const R ih = R(0.5)*sih/sq2PiRho, qy = R(0.25)*hyy/(c*sqPiRho), qz = R(0.25)*hzz/(c*sqPiRho);
    const R Lm[8][8] = {
        {R(1), 0, 0, 0, -R(1)/cc, 0, 0, 0},
        {0, 0, 0, 0, 0, R(1), 0, 0},
        {0, 0, -hzz*rs2, hyy*rs2, 0, 0, hzz*ih, -hyy*ih},
        {0, 0, -hzz*rs2, hyy*rs2, 0, 0, -hzz*ih, hyy*ih},
        {0, R(0.5)*aa/cc, -R(0.5)*sa*hyy*sih/cc, -R(0.5)*sa*hzz*sih/cc, R(0.5)*alf/(cc*Rho), 0, als*qy, als*qz},
        {0, -R(0.5)*aa/cc, R(0.5)*sa*hyy*sih/cc, R(0.5)*sa*hzz*sih/cc, R(0.5)*alf/(cc*Rho), 0, als*qy, als*qz},
        {0, R(0.5)*sa/cc, R(0.5)*aa*hyy*sih/cc, R(0.5)*aa*hzz*sih/cc, R(0.5)*als/(Rho*cc), 0, -alf*qy, -alf*qz},
        {0, -R(0.5)*sa/cc, -R(0.5)*aa*hyy*sih/cc, -R(0.5)*aa*hzz*sih/cc, R(0.5)*als/(Rho*cc), 0, -alf*qy, -alf*qz}};
    // Walk the conservative wave fan in speed order; reject a fan whose
    // intermediate state has non-positive density or negative thermal energy.
    {
        const int order[8] = {5,3,7,0,1,6,2,4};
        R st[8] = {rL, mL0, mL1, mL2, EL, WL[5], WL[6], WL[7]};
#pragma unroll
        for (int wv = 0; wv < 7; ++wv)
        {
            const int k = order[wv];
            R amp = 0;
#pragma unroll
            for (int q = 0; q < 8; ++q) amp -= Lm[k][q]*du[q];
            R dl[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) dl[q] = amp*Rm[q][k];
            dl[4] = kk*dl[0]+Rho*(U*dl[1]+V*dl[2]+W*dl[3])+dl[4]/hg+R(kInv4Pi)*(BX*dl[5]+BY*dl[6]+BZ*dl[7]);
            dl[1] = U*dl[0]+Rho*dl[1];
            dl[2] = V*dl[0]+Rho*dl[2];
            dl[3] = W*dl[0]+Rho*dl[3];
#pragma unroll
            for (int q = 0; q < 8; ++q) st[q] += dl[q];
            if (lam[order[wv+1]] > lam[k])
            {
                if (!(st[0] > R(0)) || !isfinite(st[0])) return false;
                const R th = st[4]-R(0.5)*(st[1]*st[1]+st[2]*st[2]+st[3]*st[3])/st[0]
                    -R(kInv8Pi)*(st[5]*st[5]+st[6]*st[6]+st[7]*st[7]);
                if (!(th >= R(0)) || !isfinite(th)) return false;
            }
        }
    }
    R lu[8];
#pragma unroll
    for (int w = 0; w < 8; ++w)
    {
        R s = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) s += Lm[w][q]*du[q];
        lu[w] = entropyFixedSpeed(lam[w], del)*s;
    }
    R ra[8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
    {
        R s = 0;
#pragma unroll
        for (int w = 0; w < 8; ++w) s += Rm[r][w]*lu[w];
        ra[r] = s;
    }
    ra[4] = ra[0]*kk+Rho*(U*ra[1]+V*ra[2]+W*ra[3])+ra[4]/hg+R(kInv4Pi)*(BX*ra[5]+BY*ra[6]+BZ*ra[7]);
    ra[1] = U*ra[0]+Rho*ra[1];
    ra[2] = V*ra[0]+Rho*ra[2];
    ra[3] = W*ra[0]+Rho*ra[3];
    bool finite = true;
#pragma unroll
    for (int q = 0; q < 8; ++q) { out[q] = -ra[q]; finite = finite && isfinite(out[q]); }
    return finite;
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

// Rotated MHD face flux times face area, src/MHD_SecondOrderFlux.H:34-74.
// frame: 9 row-major entries; area = radial area * |normal|.
template <typename R>
HC_HD void o2FaceFlux(R out[8], const R l[8], const R r[8], const R frame[9],
                      R area, R gamma, R entropy, bool roe)
{
    const R inv4pi = R(1.0/(4.0*3.14159265358979323846));
    R loc[2][8], fl[2][8];
#pragma unroll
    for (int s = 0; s < 2; ++s)
    {
        const R* w = s == 0 ? l : r;
        R* q = loc[s];
        q[0] = w[0]; q[4] = w[4];
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            R v = 0, bsum = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) { v += frame[3*i+j]*w[1+j]; bsum += frame[3*i+j]*w[5+j]; }
            q[1+i] = v; q[5+i] = bsum;
        }
        R v2 = 0, b2 = 0, vb = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) { v2 += q[1+i]*q[1+i]; b2 += q[5+i]*q[5+i]; vb += q[1+i]*q[5+i]; }
        const R pressure = q[4]+R(0.5)*inv4pi*b2;
        const R energy = q[4]/(gamma-R(1))+R(0.5)*q[0]*v2+R(0.5)*inv4pi*b2;
        fl[s][0] = q[0]*q[1];
        fl[s][4] = (energy+pressure)*q[1]-inv4pi*q[5]*vb;
#pragma unroll
        for (int i = 0; i < 3; ++i)
        {
            fl[s][1+i] = q[0]*q[1]*q[1+i]-inv4pi*q[5]*q[5+i]+(i == 0 ? pressure : R(0));
            fl[s][5+i] = q[1]*q[5+i]-q[1+i]*q[5];
        }
    }
    R diss[8];
    if (!roe || !roe8WaveDissipation(diss, loc[0], loc[1], gamma, entropy))
        rusanovDissipation(diss, loc[0], loc[1], gamma);
    R num[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) num[c] = R(0.5)*area*(fl[0][c]+fl[1][c]-diss[c]);
    out[0] = num[0]; out[4] = num[4];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        R v = 0, bsum = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) { v += frame[3*j+i]*num[1+j]; bsum += frame[3*j+i]*num[5+j]; }
        out[1+i] = v; out[5+i] = bsum;
    }
}
} // namespace hc
