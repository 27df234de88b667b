// Boundary-file frames to the inner radial boundary (the arithmetic of
// MHDReader::BoxData_to_BC, src/MHDReader.H:221-331, and of
// MHD_BoundaryAverages, src/MHD_BoundaryAverages.H:17-266).
//
//  * hc_bc_frame_at: the frame step. A frame is one longitude-latitude map of
//    the 8 file variables (n, v_r, v_phi, v_theta, p, B_r, B_phi, B_theta) on
//    ntheta x nphi nodes, [8][nphi][ntheta]. At time t the map is the frame
//    of t / cadence (or the last frame past the end, or the only one)
//    rotated in longitude by the sidereal rotation (25.38 days) and, between
//    two frames, interpolated linearly in time; the rotation shifts by whole
//    cells plus a linear blend of the two neighbouring columns. The cell
//    count per degree uses the integer 360 / nphi, as the reference does.
//  * hc_level_bc_points (fillPoints): the map's value at every strip cell
//    centre of the level's inner boundary (angular cells -g .. na+g-1), by
//    bilinear interpolation in (theta, phi): theta the polar angle from +z
//    (clamped to the node range), phi the azimuth, periodic.
//  * hc_level_bc_averages (fill): native conserved averages of the inner
//    ghost cells (BC_conserved_averages 1): per angular cell an 8 x 8
//    Gauss-Legendre rule in xi over the bilinear map, the spherical
//    tangential components taken to native contravariant components
//    through the cell's basis, units to cgs, and the radial ghost layers
//    scaled by the cell averages of (r0/r)^k (4-point rule), kinetic,
//    thermal and magnetic energy terms each with its own radial power; the
//    midpoint rule at second order.
#include "hc_common.hpp"
#include "hc_geometry.hpp"

#include <cmath>
#include <vector>

namespace hc
{
namespace
{
constexpr int kBcfBlock = 256;
constexpr int kBcfNodes = 8;      // angular Gauss-Legendre nodes per direction

struct BcfView
{
    int na, g, nr, order;
    int NJ;                        // na + 2g
    int ntheta, nphi;
    double r0, span, stretch, gamma;
    const double* theta;           // [ntheta] ascending
    const double* phi;             // [nphi] ascending within [phi0, phi0 + 2 pi)
    const double* data0;           // [8][nphi][ntheta]
    double nodes[kBcfNodes], weights[kBcfNodes];
    double rnodes[4], rweights[4];
};

struct BcfSample
{
    int it, jt, ip, jp;
    double ft, fp;
    double tangent[2][2];          // (phi, theta) -> native angular components
};

__device__ inline int upperBound(const double* a, int n, double x)
{
    int lo = 0, hi = n;
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// One sample of the bilinear source map at xi offsets (ox, oy) inside
// angular cell (j, k) of panel p.
__device__ inline void bcfSample(const BcfView& v, int p, int j, int k, double ox, double oy, BcfSample& s)
{
    const int perm[6][3] = {{1,2,0},{1,2,0},{1,0,2},{0,1,2},{1,0,2},{0,1,2}};
    const int sign[6][3] = {{1,-1,-1},{1,1,1},{1,-1,1},{1,1,1},{-1,1,1},{-1,-1,1}};
    const int n = v.na;
    const double x = tan(-kPi/4 + (j + ox)*kPi/(2*n));
    const double y = tan(-kPi/4 + (k + oy)*kPi/(2*n));
    const double z = 1/sqrt(1 + x*x + y*y), c = z*z*z*kPi/2;
    const double local[3][3] = {{z, -x*(1+x*x)*c, -y*(1+y*y)*c},
                                {x*z, (1+x*x)*(1+y*y)*c, -x*y*(1+y*y)*c},
                                {y*z, -x*y*(1+x*x)*c, (1+x*x)*(1+y*y)*c}};
    double A[3][3];
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) A[a][b] = sign[p][a]*local[perm[p][a]][b];
    const double t = acos(fmax(-1.0, fmin(1.0, A[2][0])));
    double ph = atan2(A[1][0], A[0][0]);
    const double phi0 = v.phi[0];
    ph = phi0 + fmod(ph - phi0 + 4*kPi, 2*kPi);
    const double ep[3] = {-sin(ph), cos(ph), 0};
    const double et[3] = {cos(t)*cos(ph), cos(t)*sin(ph), -sin(t)};
    double g11 = 0, g12 = 0, g22 = 0;
    for (int a = 0; a < 3; ++a)
    {
        g11 += A[a][1]*A[a][1];
        g12 += A[a][1]*A[a][2];
        g22 += A[a][2]*A[a][2];
    }
    const double det = g11*g22 - g12*g12;
    for (int b = 0; b < 2; ++b)
    {
        double d1 = 0, d2 = 0;
        const double* e = b == 0 ? ep : et;
        for (int a = 0; a < 3; ++a) { d1 += A[a][1]*e[a]; d2 += A[a][2]*e[a]; }
        s.tangent[0][b] = (g22*d1 - g12*d2)/det;
        s.tangent[1][b] = (g11*d2 - g12*d1)/det;
    }
    int it = upperBound(v.theta, v.ntheta, t) - 1;
    int ip = upperBound(v.phi, v.nphi, ph) - 1;
    ip = ip > 0 ? ip : 0;
    const int jp = (ip + 1) % v.nphi;
    s.fp = (ph - v.phi[ip])/((jp == 0 ? phi0 + 2*kPi : v.phi[jp]) - v.phi[ip]);
    s.ft = 0;
    if (it < 0) it = 0;
    else if (it >= v.ntheta - 1) it = v.ntheta - 1;
    else s.ft = (t - v.theta[it])/(v.theta[it+1] - v.theta[it]);
    s.it = it;
    s.jt = it + 1 < v.ntheta - 1 ? it + 1 : v.ntheta - 1;
    s.ip = ip;
    s.jp = jp;
}

// The map's 8 values at a sample (corners in the order (it,ip), (it,jp),
// (jt,ip), (jt,jp)).
__device__ inline void bcfValues(const BcfView& v, const BcfSample& s, double w[8])
{
    const int th[4] = {s.it, s.it, s.jt, s.jt}, ph[4] = {s.ip, s.jp, s.ip, s.jp};
    double coef[4];
    for (int b = 0; b < 4; ++b) coef[b] = (b >= 2 ? s.ft : 1 - s.ft)*(b % 2 ? s.fp : 1 - s.fp);
    const long long cs = (long long)v.nphi*v.ntheta;
    for (int c = 0; c < 8; ++c)
    {
        double value = 0;
        for (int b = 0; b < 4; ++b) value += coef[b]*v.data0[c*cs + (long long)ph[b]*v.ntheta + th[b]];
        w[c] = value;
    }
}

// Frame step: out = single ? (1-fL) low(phi+sL) + fL low(phi+sL-1)
//                          : ft ((1-fH) high(phi+sH) + fH high(phi+sH-1)) + (1-ft) (low ...),
// longitudes periodic. One thread per (component, node).
__global__ void __launch_bounds__(kBcfBlock) k_bc_frame(const double* __restrict__ low, const double* __restrict__ high,
                                                        int ntheta, int nphi, int sL, int sH, double fL, double fH,
                                                        double ft, int single, double* __restrict__ out)
{
    const long long n = 8LL*nphi*ntheta;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int it = (int)(t % ntheta);
    const int ip = (int)((t / ntheta) % nphi);
    const int c = (int)(t / ((long long)ntheta*nphi));
    auto at = [&](const double* f, int q) {
        return f[((long long)c*nphi + ((q % nphi) + nphi) % nphi)*ntheta + it]; };
    const double l = (1.0 - fL)*at(low, ip + sL) + fL*at(low, ip + sL - 1);
    if (single) { out[t] = l; return; }
    const double h = (1.0 - fH)*at(high, ip + sH) + fH*at(high, ip + sH - 1);
    out[t] = ft*h + (1.0 - ft)*l;
}

// fillPoints: the map at every strip cell centre, bc[6][8][NJ][NJ].
__global__ void __launch_bounds__(kBcfBlock) k_bc_points(BcfView v, double* __restrict__ bc)
{
    const long long nn = (long long)v.NJ*v.NJ;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= 6*nn) return;
    const int jj = (int)(t % v.NJ), kk = (int)((t / v.NJ) % v.NJ), p = (int)(t / nn);
    BcfSample s;
    bcfSample(v, p, jj - v.g, kk - v.g, 0.5, 0.5, s);
    double w[8];
    bcfValues(v, s, w);
    for (int c = 0; c < 8; ++c) bc[(p*8LL + c)*nn + (long long)kk*v.NJ + jj] = w[c];
}

// fill: native conserved averages of the inner ghost cells, inner[6][8][NJ][NJ][g].
__global__ void __launch_bounds__(kBcfBlock) k_bc_averages(BcfView v, double* __restrict__ inner)
{
    const long long nn = (long long)v.NJ*v.NJ;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= 6*nn) return;
    const int jj = (int)(t % v.NJ), kk = (int)((t / v.NJ) % v.NJ), p = (int)(t / nn);
    const int j = jj - v.g, k = kk - v.g;
    const bool midpoint = v.order == 2;
    double mean[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int nq = midpoint ? 1 : kBcfNodes;
    for (int q = 0; q < nq; ++q) for (int r = 0; r < nq; ++r)
    {
        BcfSample s;
        bcfSample(v, p, j, k, midpoint ? 0.5 : v.nodes[q], midpoint ? 0.5 : v.nodes[r], s);
        const double weight = midpoint ? 1.0 : v.weights[q]*v.weights[r];
        double w[8];
        bcfValues(v, s, w);
        const double rho = w[0]*1.67262192e-24;
        mean[0] += weight*rho;
        mean[1] += weight*rho*w[1];
        mean[4] += weight*w[4]*1e-12/(v.gamma-1);
        mean[5] += weight*(.5*rho*(w[1]*w[1]+w[2]*w[2]+w[3]*w[3]) + 1e-12*(w[6]*w[6]+w[7]*w[7])/(8*kPi));
        mean[6] += weight*1e-12*w[5]*w[5]/(8*kPi);
        mean[7] += weight*1e-6*w[5];
        for (int d = 0; d < 2; ++d)
        {
            mean[2+d] += weight*rho*(s.tangent[d][0]*w[2] + s.tangent[d][1]*w[3]);
            mean[8+d] += weight*1e-6*(s.tangent[d][0]*w[6] + s.tangent[d][1]*w[7]);
        }
    }
    for (int ig = 0; ig < v.g; ++ig)
    {
        const int ir = ig - v.g;
        double sc[4] = {0, 0, 0, 0};
        for (int q = 0; q < (midpoint ? 1 : 4); ++q)
        {
            const double xi = (ir + (midpoint ? .5 : v.rnodes[q]))/v.nr;
            const double radius = v.r0 + v.span*(v.stretch == 0 ? xi : expm1(v.stretch*xi));
            const double ratio = v.r0/radius, r2 = ratio*ratio;
            const double weight = midpoint ? 1.0 : v.rweights[q];
            sc[0] += weight*ratio;
            sc[1] += weight*r2;
            sc[2] += weight*r2*r2;
            sc[3] += weight*pow(ratio, 2*v.gamma);
        }
        double out[8];
        for (int c = 0; c < 4; ++c) out[c] = mean[c]*sc[1];
        out[4] = mean[4]*sc[3] + mean[5]*sc[1] + mean[6]*sc[2];
        out[5] = mean[7]*sc[1];
        out[6] = mean[8]*sc[0];
        out[7] = mean[9]*sc[0];
        for (int c = 0; c < 8; ++c) inner[(((p*8LL + c)*nn) + (long long)kk*v.NJ + jj)*v.g + ig] = out[c];
    }
}
} // namespace

// Frame selection, rotation and temporal weights of BoxData_to_BC
// (src/MHDReader.H:247-331).
struct BcFrameStep { int low, high, sL, sH, single; double fL, fH, ft; };
BcFrameStep bcFrameStep(int nframes, int nphi, double cadence_h, int frame_rotate, double time)
{
    BcFrameStep st{};
    const double carr = 25.38*24*60*60;
    const int perCell = 360/nphi;      // integer degrees per cell, as the reference computes it
    const double elapsed = time/(cadence_h*3600.0);   // frames elapsed; compared before any int conversion
    auto cells = [&](double angle, int& shift, double& frac) {
        int c = (int)(angle/perCell);
        frac = angle/perCell - c;
        c = c % nphi;
        shift = nphi - c;
    };
    if (nframes == 1 || elapsed >= nframes - 1)
    {
        const double angle = std::fmod(360.0*time/carr, 360.0);
        cells(angle, st.sL, st.fL);
        st.low = st.high = nframes == 1 ? 0 : nframes - 1;
        st.single = 1;
        return st;
    }
    const int low = (int)(time/(cadence_h*3600)), next = low + 1;
    st.ft = (time - (low*cadence_h*3600))/(cadence_h*3600);
    double a1, a2;
    if (frame_rotate == 1)
    {
        a1 = std::fmod(360*(time + low*cadence_h*3600)/carr, 360);
        a2 = std::fmod(360*(time + next*cadence_h*3600)/carr, 360);
    }
    else
    {
        a1 = std::fmod(360*time/carr, 360);
        a2 = std::fmod(360*time/carr, 360);
    }
    cells(a1, st.sL, st.fL);
    cells(a2, st.sH, st.fH);
    st.low = low; st.high = next;
    return st;
}

static void bcfView(BcfView& v, const ShellDims& d, int order, double gamma, const double* data0, int ntheta,
                    int nphi, const double* theta, const double* phi)
{
    v.na = d.na; v.g = d.g; v.nr = d.nr; v.order = order; v.NJ = d.na + 2*d.g;
    v.ntheta = ntheta; v.nphi = nphi;
    v.r0 = d.r0; v.stretch = d.stretch;
    v.span = d.stretch == 0 ? d.r1 - d.r0 : (d.r1 - d.r0)/std::expm1(d.stretch);
    v.gamma = gamma;
    v.theta = theta; v.phi = phi; v.data0 = data0;
    std::vector<double> x, w;
    gaussLegendre01(kBcfNodes, x, w);
    for (int q = 0; q < kBcfNodes; ++q) { v.nodes[q] = x[q]; v.weights[q] = w[q]; }
    gaussLegendre01(4, x, w);
    for (int q = 0; q < 4; ++q) { v.rnodes[q] = x[q]; v.rweights[q] = w[q]; }
}

void bcPoints(const ShellDims& d, const double* data0, int ntheta, int nphi, const double* theta, const double* phi,
              double* bc, hipStream_t s)
{
    BcfView v{};
    bcfView(v, d, 2, 5.0/3.0, data0, ntheta, nphi, theta, phi);
    const long long n = 6LL*v.NJ*v.NJ;
    hipLaunchKernelGGL(k_bc_points, dim3((unsigned)((n + kBcfBlock - 1)/kBcfBlock)), dim3(kBcfBlock), 0, s, v, bc);
    HC_THROW(hipGetLastError());
}

void bcAverages(const ShellDims& d, int order, double gamma, const double* data0, int ntheta, int nphi,
                const double* theta, const double* phi, double* inner, hipStream_t s)
{
    BcfView v{};
    bcfView(v, d, order, gamma, data0, ntheta, nphi, theta, phi);
    const long long n = 6LL*v.NJ*v.NJ;
    hipLaunchKernelGGL(k_bc_averages, dim3((unsigned)((n + kBcfBlock - 1)/kBcfBlock)), dim3(kBcfBlock), 0, s, v, inner);
    HC_THROW(hipGetLastError());
}
} // namespace hc

extern "C" {

int hc_bc_frame_at(const double* frames, int nframes, int ntheta, int nphi, double cadence_hours, int frame_rotate,
                   double time, double* data0, hc_stream stream)
{
    if (!frames || !data0 || nframes < 1 || ntheta < 1 || nphi < 1) return hc::fail("hc_bc_frame_at: bad argument");
    if (nphi > 360) return hc::fail("hc_bc_frame_at: nphi must be <= 360 (whole degrees per longitude cell)");
    if (!(cadence_hours > 0) || !std::isfinite(cadence_hours))
        return hc::fail("hc_bc_frame_at: cadence must be positive and finite");
    if (frame_rotate != 0 && frame_rotate != 1) return hc::fail("hc_bc_frame_at: frame_rotate must be 0 or 1");
    if (!(time >= 0) || !std::isfinite(time)) return hc::fail("hc_bc_frame_at: time must be finite and >= 0");
    try
    {
        const hc::BcFrameStep st = hc::bcFrameStep(nframes, nphi, cadence_hours, frame_rotate, time);
        const long long per = 8LL*nphi*ntheta;
        const long long n = per;
        hipLaunchKernelGGL(hc::k_bc_frame, dim3((unsigned)((n + hc::kBcfBlock - 1)/hc::kBcfBlock)),
                           dim3(hc::kBcfBlock), 0, static_cast<hipStream_t>(stream), frames + st.low*per,
                           frames + st.high*per, ntheta, nphi, st.sL, st.sH, st.fL, st.fH, st.ft, st.single, data0);
        HC_THROW(hipGetLastError());
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_bc_frame_at: ") + e.what()); }
}

} // extern "C"
