// Inner radial boundary ghosts from boundary-file data (the reader's primitive
// strip; init_condition_type 0 with BC_conserved_averages 0),
// src/BoxOp_PreStage.H:161-317, and at second order the conversion of those
// native conserved ghosts to Cartesian primitives,
// src/BoxOp_SecondOrderPreStage.H:194-214.
//
// Per ghost cell (radial layer i = -g..-1, every ghosted angular cell):
//   * the file values of the cell's angular column (one radial strip, in the
//     file's units and order rho, v_r, v_phi, v_theta, p, B_r, B_phi, B_theta)
//     are copied to every ghost layer (:231-236);
//   * the spherical vectors become Cartesian at the cell centre's polar and
//     azimuthal angles (:238-262), then native with A^-1 (:264-278): the
//     inverse of the point basis at the cell centre at second order, of
//     <det(A) A>/<det A> at fourth order (:166-191);
//   * units and radial scaling with s = r_BC / r (:280-316): rho x 1.67262192e-24
//     s^2, p x 1e-12 s^(2 gamma), native B x 1e-6 (s^2 radial, s angular),
//     energy from the Cartesian speed and the scaled field magnitude.
// The fourth-order level takes these native conserved values as its inner
// ghosts; the second-order level converts them back with the point basis to
// Cartesian primitives (density floor, pressure clipped at 0). The
// radial-outflow test boundary (init_condition_type 1) uses the same ghost
// geometry and conversion.
#include "hc_common.hpp"
#include "hc_geometry.hpp"
#include "hc_physics.hpp"

#include <cmath>
#include <vector>

namespace hc
{
namespace
{
constexpr int kBcBlock = 256;
constexpr int kBcCol = 21;   // unit direction (3), A^-1 (9), A (9) per ghosted angular cell

struct BcView
{
    int na, g, order;
    int NJ;                  // na + 2g
    double gamma, rBC;
    const double* col;       // [6][NJ][NJ][kBcCol]
    const double* rad;       // [g] radius of ghost layer i = -g + ig
};

// Writes the native conserved ghost u (fourth order), or its Cartesian
// primitives through the point basis (second order).
__device__ inline void storeGhost(const BcView& v, const double* c, double u[8], double* __restrict__ out,
                                  long long p, long long col, int ig)
{
    if (v.order == 2)
    {
        double cart[8];
        cart[0] = u[0]; cart[4] = u[4];
#pragma unroll
        for (int grp = 0; grp < 2; ++grp)
        {
            const int st0 = grp == 0 ? 1 : 5;
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                double a = 0.0;
#pragma unroll
                for (int j = 0; j < 3; ++j) a += c[12+3*d+j]*u[st0+j];
                cart[st0+d] = a;
            }
        }
        conservedToPrimitive<double>(u, cart, 1.0, v.gamma);
    }
    const long long nn = (long long)v.NJ*v.NJ;
    const long long o = ((p*8)*nn + col)*v.g + ig;
#pragma unroll
    for (int q = 0; q < 8; ++q) out[o + (long long)q*nn*v.g] = u[q];
}

__global__ void __launch_bounds__(kBcBlock) k_inner_bc(BcView v, const double* __restrict__ bc,
                                                       double* __restrict__ out)
{
    const long long nn = (long long)v.NJ*v.NJ;
    const long long n = 6LL*nn*v.g;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int ig = (int)(t % v.g);
    const long long colAll = t / v.g;              // p*nn + k*NJ + j
    const long long p = colAll / nn, col = colAll - p*nn;
    const double* c = v.col + colAll*kBcCol;
    double w[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) w[q] = bc[(p*8 + q)*nn + col];
    const double rad = v.rad[ig];
    const double X[3] = {rad*c[0], rad*c[1], rad*c[2]};
    const double r = sqrt(X[0]*X[0] + X[1]*X[1] + X[2]*X[2]);
    const double theta = acos(X[2]/r);
    const double phi = atan2(X[1], X[0]);
    const double st = sin(theta), ct = cos(theta), sp = sin(phi), cp = cos(phi);
    double Vc[3], Bc[3];
    {
        const double ur = w[1], uth = w[3], uph = w[2];
        Vc[0] = st*cp*ur + ct*cp*uth - sp*uph;
        Vc[1] = st*sp*ur + ct*sp*uth + cp*uph;
        Vc[2] = ct*ur - st*uth;
        const double br = w[5], bth = w[7], bph = w[6];
        Bc[0] = st*cp*br + ct*cp*bth - sp*bph;
        Bc[1] = st*sp*br + ct*sp*bth + cp*bph;
        Bc[2] = ct*br - st*bth;
    }
    double Vs[3], Bs[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        double sv = 0.0, sb = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) { sv += c[3+3*d+j]*Vc[j]; sb += c[3+3*d+j]*Bc[j]; }
        Vs[d] = sv; Bs[d] = sb;
    }
    const double s = v.rBC/rad;
    const double rho = w[0]*1.67262192e-24*pow(s, 2.0);
    const double pres = w[4]*1.0e-12*pow(s, 2.0*v.gamma);
    const double Bx = Bs[0]*1.0e-6*pow(s, 2.0);
    const double By = Bs[1]*1.0e-6*pow(s, 1.0);
    const double Bz = Bs[2]*1.0e-6*pow(s, 1.0);
    const double V2 = Vc[0]*Vc[0] + Vc[1]*Vc[1] + Vc[2]*Vc[2];
    const double B2 = Bc[0]*Bc[0] + Bc[1]*Bc[1] + Bc[2]*Bc[2];
    const double Brad2 = Bs[0]*Bs[0];
    const double Bnon2 = B2 - Brad2;
    const double Bsc2 = 1.0e-12*Bnon2*pow(s, 2.0) + 1.0e-12*Brad2*pow(s, 4.0);
    const double e = pres/(v.gamma - 1.0) + rho*V2/2.0 + Bsc2/8.0/kPi;
    double u[8] = {rho, rho*Vs[0], rho*Vs[1], rho*Vs[2], e, Bx, By, Bz};
    storeGhost(v, c, u, out, p, col, ig);
}

// Radial-outflow inner boundary (init_condition_type 1, f_radialoutflowIBCs,
// src/BoxOp_PreStage.H:5-48): rho = 700 m_p s^2, p = 1e-7 s^(2 gamma) with
// s = r_BC / r, a radial speed 500 km/s less strength/100/3 of it times the
// co-rotating pattern sin^5(theta) sin(5 phi_t) + sin^10(theta) sin(10 phi_t),
// phi_t = phi - Omega t (Carrington period 25.38 days), and no field.
__global__ void __launch_bounds__(kBcBlock) k_inner_outflow(BcView v, double strength, double time,
                                                            double* __restrict__ out)
{
    const long long nn = (long long)v.NJ*v.NJ;
    const long long n = 6LL*nn*v.g;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int ig = (int)(t % v.g);
    const long long colAll = t / v.g;
    const long long p = colAll / nn, col = colAll - p*nn;
    const double* c = v.col + colAll*kBcCol;
    const double rad = v.rad[ig];
    const double X[3] = {rad*c[0], rad*c[1], rad*c[2]};
    const double r = sqrt(X[0]*X[0] + X[1]*X[1] + X[2]*X[2]);
    const double theta = atan2(sqrt(X[0]*X[0] + X[1]*X[1]), X[2]);
    const double phi = atan2(X[1], X[0]);
    const double p0 = 1.0e-7, rho0 = 700*1.67262192e-24, u0 = 500.0*1e5;
    const double rho = rho0*pow(v.rBC/r, 2.0);
    const double pres = p0*pow(v.rBC/r, 2.0*v.gamma);
    const double carr = 25.38*24*60*60;
    const double omega = 2.0*kPi/carr;
    const double phit = phi - omega*time;
    const double scale = strength/100.0;
    const double st = sin(theta);
    const double pattern = pow(st, 5)*sin(5*phit) + pow(st, 10)*sin(10*phit);
    const double ur = u0 - scale*u0/3*pattern;
    double u[8] = {rho, ur*rho, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double umag = sqrt(u[1]*u[1] + u[2]*u[2] + u[3]*u[3])/u[0];
    const double bmag = sqrt(u[5]*u[5] + u[6]*u[6] + u[7]*u[7]);
    u[4] = pres/(v.gamma - 1.0) + 0.5*u[0]*umag*umag + bmag*bmag/8.0/kPi;
    storeGhost(v, c, u, out, p, col, ig);
}
} // namespace

struct BcTable
{
    ShellDims dims{};
    int order = 0;
    double gamma = 0;
    double* col = nullptr;
    double* rad = nullptr;
    ~BcTable()
    {
        if (col) (void)hipFree(col);
        if (rad) (void)hipFree(rad);
    }
};

BcTable* bcCreate(const ShellDims& dims, int order, double gamma)
{
    const int na = dims.na, g = dims.g, NJ = na + 2*g;
    const double h = 1.0/na;
    std::vector<double> col((size_t)6*NJ*NJ*kBcCol), rad(g);
    for (int p = 0; p < 6; ++p)
        for (int k = -g; k < na+g; ++k)
            for (int j = -g; j < na+g; ++j)
            {
                double* e = &col[(((size_t)p*NJ + (k+g))*NJ + (j+g))*kBcCol];
                const double al = xiToAngle((j+0.5)*h), be = xiToAngle((k+0.5)*h);
                const Vec3 u = panelPoint(p, al, be);
                for (int a = 0; a < 3; ++a) e[a] = u.x[a];
                double A[3][3], M[3][3], Mi[3][3];
                pointBasis(p, al, be, A);
                if (order == 2)
                {
                    invert3(A, Mi);
                }
                else
                {
                    double m0[13];
                    faceMetricAverages(na, p, 0, j, k, m0);
                    const double invDet = 1.0/m0[0];
                    for (int x = 0; x < 3; ++x) for (int y = 0; y < 3; ++y) M[x][y] = m0[1+3*x+y]*invDet;
                    invert3(M, Mi);
                }
                for (int x = 0; x < 9; ++x) { e[3+x] = Mi[x/3][x%3]; e[12+x] = A[x/3][x%3]; }
            }
    for (int ig = 0; ig < g; ++ig) rad[ig] = radialCell(dims, ig - g).rmid;
    BcTable* T = new BcTable;
    T->dims = dims;
    T->order = order;
    T->gamma = gamma;
    try
    {
        HC_THROW(hipMalloc(&T->col, col.size()*sizeof(double)));
        HC_THROW(hipMalloc(&T->rad, rad.size()*sizeof(double)));
        HC_THROW(hipMemcpy(T->col, col.data(), col.size()*sizeof(double), hipMemcpyHostToDevice));
        HC_THROW(hipMemcpy(T->rad, rad.data(), rad.size()*sizeof(double), hipMemcpyHostToDevice));
    }
    catch (...)
    {
        delete T;
        throw;
    }
    return T;
}

void bcDestroy(BcTable* T) { delete T; }

static BcView bcView(const BcTable& T)
{
    BcView v;
    v.na = T.dims.na;
    v.g = T.dims.g;
    v.order = T.order;
    v.NJ = v.na + 2*v.g;
    v.gamma = T.gamma;
    v.rBC = T.dims.r0;
    v.col = T.col;
    v.rad = T.rad;
    return v;
}

void outflowLaunch(BcTable& T, double strength, double time, double* inner, hipStream_t s)
{
    const BcView v = bcView(T);
    const long long n = 6LL*v.NJ*v.NJ*v.g;
    hipLaunchKernelGGL(k_inner_outflow, dim3((unsigned)((n + kBcBlock - 1)/kBcBlock)), dim3(kBcBlock), 0, s,
                       v, strength, time, inner);
    HC_THROW(hipGetLastError());
}

void bcLaunch(BcTable& T, const double* bc, double* inner, hipStream_t s)
{
    const BcView v = bcView(T);
    const long long n = 6LL*v.NJ*v.NJ*v.g;
    hipLaunchKernelGGL(k_inner_bc, dim3((unsigned)((n + kBcBlock - 1)/kBcBlock)), dim3(kBcBlock), 0, s, v, bc, inner);
    HC_THROW(hipGetLastError());
}
} // namespace hc
