// Host-side geometry of the six-panel cubed-sphere shell: closed-form metric
// averages, cell-centre bases, seam gather tables and the seam edge map.
//
// Reference interfaces replaced (src/ = HelioCubed reference repository):
//   * radial metrics r, dVol_r, adjDr: src/BoxOp_Utilities.H:6-61;
//   * pointwise basis A (columns: radial unit vector, d xhat/d xi_1,
//     d xhat/d xi_2): src/MHD_BoundaryAverages.H:62-70, panel permutation and
//     sign src/BoxOp_EulerCubedSphereImplem.H:187-188;
//   * cell-averaged det A, CubedSphereShell::detA(box,h,0): the exact
//     spherical-quadrilateral solid angle over h^2 (same quantity as
//     src/MHD_AMR.H:29-45);
//   * face-averaged cofactor rows, CubedSphereShell::adjA(box,h,block,d):
//     lines of constant alpha or beta are great-circle arcs, so the average of
//     the cofactor row over an angular face is the arc's plane normal times its
//     angle over h. The radial-face row follows from the closed cone surface,
//     which makes the discrete metric identity (sum of face area vectors = 0)
//     hold to round-off;
//   * seam ghost interpolation (the role of Proto's MBInterpOp,
//     src/MHD_SecondOrderSeamPlan.H:103-123): every ghost column of a panel is
//     a weighted sum of valid columns of the panel its centre lies in, at the
//     same radial index. Along each neighbour index the weights are one point
//     when the continuous index is an integer (edge ghosts land on a grid line
//     of the neighbour) or a 4-point Lagrange window clamped into the valid
//     cells.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace hc
{
constexpr double kPi = 3.14159265358979323846;

struct ShellDims
{
    int nr, na, g;
    double r0, r1, stretch;
};

// Panel frame: global Cartesian a = sign[p][a] * local[perm[p][a]] with
// local = (1, tan alpha, tan beta) / |.|.
inline const int (&panelPerm(int p))[3]
{
    static const int perm[6][3] = {{1,2,0},{1,2,0},{1,0,2},{0,1,2},{1,0,2},{0,1,2}};
    return perm[p];
}
inline const int (&panelSign(int p))[3]
{
    static const int sign[6][3] = {{1,-1,-1},{1,1,1},{1,-1,1},{1,1,1},{-1,1,1},{-1,-1,1}};
    return sign[p];
}

inline double xiToAngle(double xi) { return -kPi/4 + (kPi/2)*xi; }

struct Vec3 { double x[3]; };
inline Vec3 cross(const Vec3& a, const Vec3& b)
{
    return {{a.x[1]*b.x[2]-a.x[2]*b.x[1], a.x[2]*b.x[0]-a.x[0]*b.x[2], a.x[0]*b.x[1]-a.x[1]*b.x[0]}};
}
inline double dot(const Vec3& a, const Vec3& b) { return a.x[0]*b.x[0]+a.x[1]*b.x[1]+a.x[2]*b.x[2]; }
inline double norm(const Vec3& a) { return std::sqrt(dot(a,a)); }
inline Vec3 sub(const Vec3& a, const Vec3& b) { return {{a.x[0]-b.x[0], a.x[1]-b.x[1], a.x[2]-b.x[2]}}; }

// Unit radial direction at panel angles (alpha, beta).
inline Vec3 panelPoint(int p, double alpha, double beta)
{
    const double t1 = std::tan(alpha), t2 = std::tan(beta);
    const double inv = 1.0/std::sqrt(1.0+t1*t1+t2*t2);
    const double local[3] = {inv, t1*inv, t2*inv};
    Vec3 v;
    for (int a = 0; a < 3; ++a) v.x[a] = panelSign(p)[a]*local[panelPerm(p)[a]];
    return v;
}

// Components of a global direction in panel q's local frame.
inline Vec3 toPanelLocal(int q, const Vec3& X)
{
    Vec3 L;
    for (int a = 0; a < 3; ++a) L.x[panelPerm(q)[a]] = panelSign(q)[a]*X.x[a];
    return L;
}

// Pointwise basis A at panel angles (row-major A[a][b]).
inline void pointBasis(int p, double alpha, double beta, double A[3][3])
{
    const double x = std::tan(alpha), y = std::tan(beta);
    const double z = 1.0/std::sqrt(1.0+x*x+y*y);
    const double c = z*z*z*kPi/2;
    const double ax = 1.0+x*x, ay = 1.0+y*y;
    const double local[3][3] = {{z, -x*ax*c, -y*ay*c},
                                {x*z, ax*ay*c, -x*y*ay*c},
                                {y*z, -x*y*ax*c, ax*ay*c}};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) A[a][b] = panelSign(p)[a]*local[panelPerm(p)[a]][b];
}

inline void invert3(const double A[3][3], double B[3][3])
{
    const double c00 = A[1][1]*A[2][2]-A[1][2]*A[2][1];
    const double c01 = A[1][2]*A[2][0]-A[1][0]*A[2][2];
    const double c02 = A[1][0]*A[2][1]-A[1][1]*A[2][0];
    const double det = A[0][0]*c00+A[0][1]*c01+A[0][2]*c02;
    const double r = 1.0/det;
    B[0][0] = c00*r; B[1][0] = c01*r; B[2][0] = c02*r;
    B[0][1] = (A[0][2]*A[2][1]-A[0][1]*A[2][2])*r;
    B[1][1] = (A[0][0]*A[2][2]-A[0][2]*A[2][0])*r;
    B[2][1] = (A[0][1]*A[2][0]-A[0][0]*A[2][1])*r;
    B[0][2] = (A[0][1]*A[1][2]-A[0][2]*A[1][1])*r;
    B[1][2] = (A[0][2]*A[1][0]-A[0][0]*A[1][2])*r;
    B[2][2] = (A[0][0]*A[1][1]-A[0][1]*A[1][0])*r;
}

// Radial metric line, src/BoxOp_Utilities.H:36-59.
struct RadialCell { double rlo, rhi, rmid, dvol, area0, areaT; };
inline RadialCell radialCell(const ShellDims& s, int i)
{
    const double dxi = 1.0/s.nr, C = s.stretch;
    const double span = C == 0.0 ? (s.r1-s.r0) : (s.r1-s.r0)/std::expm1(C);
    auto at = [&](double eta) { return s.r0 + span*(C == 0.0 ? eta : std::expm1(C*eta)); };
    const double lo = i*dxi, hi = (i+1.0)*dxi;
    RadialCell r;
    r.rlo = at(lo); r.rhi = at(hi); r.rmid = at((lo+hi)/2.0);
    r.dvol = (r.rhi*r.rhi*r.rhi - r.rlo*r.rlo*r.rlo)/3.0/dxi;
    r.area0 = r.rlo*r.rlo;
    r.areaT = (r.rhi*r.rhi - r.rlo*r.rlo)/dxi/2.0;
    return r;
}

// Solid angle of the spherical triangle (a,b,c) of unit vectors.
inline double triangleSolidAngle(const Vec3& a, const Vec3& b, const Vec3& c)
{
    const double triple = dot(a, cross(sub(b,a), sub(c,a)));
    const double denom = 1.0 + dot(a,b) + dot(b,c) + dot(c,a);
    return 2.0*std::atan2(std::fabs(triple), denom);
}

// Cell-averaged det A of angular cell (j,k) (same for every panel).
inline double cellDetA(int na, int j, int k)
{
    const double h = 1.0/na;
    const double a0 = xiToAngle(j*h), a1 = xiToAngle((j+1)*h);
    const double b0 = xiToAngle(k*h), b1 = xiToAngle((k+1)*h);
    const Vec3 c00 = panelPoint(3,a0,b0), c10 = panelPoint(3,a1,b0);
    const Vec3 c11 = panelPoint(3,a1,b1), c01 = panelPoint(3,a0,b1);
    const double omega = triangleSolidAngle(c00,c10,c11) + triangleSolidAngle(c00,c11,c01);
    return omega/(h*h);
}

// Average over an angular face of the cofactor row along that face's normal:
// e0 -> e1 is the arc traversed with increasing tangential angle; the row is
// normalize(e1 x e0) * angle / h for d = 1 and normalize(e0 x e1) * angle / h
// for d = 2.
inline Vec3 arcNormal(const Vec3& e0, const Vec3& e1, double h, bool flip)
{
    Vec3 n = flip ? cross(e1,e0) : cross(e0,e1);
    const double s = norm(n);
    const double ang = std::atan2(s, dot(e0,e1));
    const double f = ang/(s*h);
    return {{n.x[0]*f, n.x[1]*f, n.x[2]*f}};
}

// Face-averaged cofactor row of face (d; j,k) of panel p; (j,k) is the angular
// index of the face's low corner as stored in a face-centred array.
inline Vec3 faceNormal(int na, int p, int d, int j, int k)
{
    const double h = 1.0/na;
    if (d == 1)
    {
        const double a = xiToAngle(j*h);
        return arcNormal(panelPoint(p,a,xiToAngle(k*h)), panelPoint(p,a,xiToAngle((k+1)*h)), h, true);
    }
    if (d == 2)
    {
        const double b = xiToAngle(k*h);
        return arcNormal(panelPoint(p,xiToAngle(j*h),b), panelPoint(p,xiToAngle((j+1)*h),b), h, false);
    }
    const Vec3 n1lo = faceNormal(na,p,1,j,k), n1hi = faceNormal(na,p,1,j+1,k);
    const Vec3 n2lo = faceNormal(na,p,2,j,k), n2hi = faceNormal(na,p,2,j,k+1);
    Vec3 r;
    for (int a = 0; a < 3; ++a)
        r.x[a] = -((n1hi.x[a]-n1lo.x[a]) + (n2hi.x[a]-n2lo.x[a]))/(2.0*h);
    return r;
}

// ---- Seam ghost interpolation ----------------------------------------------
struct GhostDonor { int q, j, k; double w; };

inline int closestPanel(const Vec3& X)
{
    int q = 0; double best = -2.0;
    for (int c = 0; c < 6; ++c)
    {
        const double v = toPanelLocal(c, X).x[0];
        if (v > best) { best = v; q = c; }
    }
    return q;
}

inline void stencil1d(double u, int na, int& base, int& n, double w[4])
{
    const double nearest = std::round(u);
    if (std::fabs(u-nearest) < 1e-8) { base = (int)nearest; n = 1; w[0] = 1.0; return; }
    base = std::max(0, std::min(na-4, (int)std::floor(u)-1));
    n = 4;
    for (int a = 0; a < 4; ++a)
    {
        double v = 1.0;
        for (int b = 0; b < 4; ++b)
            if (b != a) v *= (u-(base+b))/double(a-b);
        w[a] = v;
    }
}

inline std::vector<GhostDonor> ghostColumnDonors(int na, int p, int j, int k)
{
    const double h = 1.0/na;
    const Vec3 X = panelPoint(p, xiToAngle((j+0.5)*h), xiToAngle((k+0.5)*h));
    const int q = closestPanel(X);
    const Vec3 L = toPanelLocal(q, X);
    const double u = (std::atan2(L.x[1],L.x[0])+kPi/4)/(h*kPi/2)-0.5;
    const double v = (std::atan2(L.x[2],L.x[0])+kPi/4)/(h*kPi/2)-0.5;
    int bu, nu, bv, nv; double wu[4], wv[4];
    stencil1d(u, na, bu, nu, wu);
    stencil1d(v, na, bv, nv, wv);
    std::vector<GhostDonor> out;
    for (int s = 0; s < nv; ++s)
        for (int t = 0; t < nu; ++t) out.push_back({q, bu+t, bv+s, wu[t]*wv[s]});
    return out;
}

// Seam edge map: the angular face layer of panel p normal to d (1 or 2) on
// side s (0 low, 1 high) is the layer of panel q normal to dq on side sq;
// tangential index t of p is index (rev ? na-1-t : t) of q.
struct SeamEdge { int q, dq, sq, rev; };

inline SeamEdge seamEdge(int na, int p, int d, int s)
{
    const double h = 1.0/na;
    SeamEdge e{};
    int tq[2] = {0, 0};
    for (int pass = 0; pass < 2; ++pass)
    {
        const int t = pass == 0 ? 0 : na-1;
        const int jo = d == 1 ? (s ? na : -1) : t;
        const int ko = d == 1 ? t : (s ? na : -1);
        const int q = closestPanel(panelPoint(p, xiToAngle((jo+0.5)*h), xiToAngle((ko+0.5)*h)));
        const double fa = d == 1 ? (s ? na : 0)*h : (t+0.5)*h;
        const double fb = d == 1 ? (t+0.5)*h : (s ? na : 0)*h;
        const Vec3 L = toPanelLocal(q, panelPoint(p, xiToAngle(fa), xiToAngle(fb)));
        const double u = (std::atan2(L.x[1],L.x[0])+kPi/4)/(h*kPi/2);
        const double v = (std::atan2(L.x[2],L.x[0])+kPi/4)/(h*kPi/2);
        const long ur = std::lround(u), vr = std::lround(v);
        if (std::fabs(u-ur) < 1e-6 && (ur == 0 || ur == na))
        { e.dq = 1; e.sq = ur == 0 ? 0 : 1; tq[pass] = (int)std::floor(v); }
        else
        { e.dq = 2; e.sq = vr == 0 ? 0 : 1; tq[pass] = (int)std::floor(u); }
        e.q = q;
    }
    e.rev = tq[0] != 0 ? 1 : 0;
    return e;
}

// ---- Face averages for the fourth-order operator -------------------------
// Gauss-Legendre nodes/weights on [0,1] (n points) by Newton iteration on
// the Legendre polynomial.
inline void gaussLegendre01(int n, std::vector<double>& x, std::vector<double>& w)
{
    x.resize(n); w.resize(n);
    for (int i = 0; i < n; ++i)
    {
        double z = std::cos(kPi*(i+0.75)/(n+0.5)), dp = 0;
        for (int it = 0; it < 100; ++it)
        {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) { const double p2 = ((2*k-1)*z*p1-(k-1)*p0)/k; p0 = p1; p1 = p2; }
            dp = n*(z*p1-p0)/(z*z-1);
            const double dz = p1/dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = 0.5*(1-z);
        w[i] = 1.0/((1-z*z)*dp*dp);
    }
}

// Averages over face (d; j,k) of det A (out[0]), det(A) A (out[1..9],
// row-major) and the cofactor row d (out[10..12]), all in block coordinates
// xi (A = d xhat / d xi). d = 0 is the angular cell (j,k), d = 1 the line
// alpha = j h over cell k, d = 2 the line beta = k h over cell j. The det A
// average of d = 0 is the exact solid angle; d = 1, 2 use the closed form
// int det A dxi = (pi/2) [t / sqrt(1 + s^2 + t^2)] along the line (s the fixed
// tangent); the cofactor rows use the great-circle closed forms.
inline void faceMetricAverages(int na, int p, int d, int j, int k, double out[13])
{
    static std::vector<double> gx, gw;
    if (gx.empty()) gaussLegendre01(12, gx, gw);
    const double h = 1.0/na;
    for (int a = 0; a < 13; ++a) out[a] = 0;
    auto add = [&](double wgt, double al, double be) {
        double A[3][3];
        pointBasis(p, al, be, A);
        const double det = A[0][0]*(A[1][1]*A[2][2]-A[1][2]*A[2][1])
                         - A[0][1]*(A[1][0]*A[2][2]-A[1][2]*A[2][0])
                         + A[0][2]*(A[1][0]*A[2][1]-A[1][1]*A[2][0]);
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) out[1+3*a+b] += wgt*det*A[a][b];
    };
    const int n = (int)gx.size();
    if (d == 0)
    {
        for (int q = 0; q < n; ++q) for (int r = 0; r < n; ++r)
            add(gw[q]*gw[r], xiToAngle((j+gx[q])*h), xiToAngle((k+gx[r])*h));
        out[0] = cellDetA(na, j, k);
    }
    else
    {
        for (int q = 0; q < n; ++q)
        {
            if (d == 1) add(gw[q], xiToAngle(j*h), xiToAngle((k+gx[q])*h));
            else add(gw[q], xiToAngle((j+gx[q])*h), xiToAngle(k*h));
        }
        const double s = std::tan(xiToAngle((d == 1 ? j : k)*h));
        const double t0 = std::tan(xiToAngle((d == 1 ? k : j)*h)), t1 = std::tan(xiToAngle(((d == 1 ? k : j)+1)*h));
        out[0] = (kPi/2)/h*(t1/std::sqrt(1+s*s+t1*t1) - t0/std::sqrt(1+s*s+t0*t0));
    }
    const Vec3 nrm = faceNormal(na, p, d, j, k);
    for (int a = 0; a < 3; ++a) out[10+a] = nrm.x[a];
}
} // namespace hc
