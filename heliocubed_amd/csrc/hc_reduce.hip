// Level reductions: the CFL time step and the conserved-state check.
//
// CFL time step of a level (replaces dtCFLSecondOrderHost,
// src/BoxOp_SecondOrderCFL.H:99-210, and the fourth-order dtCFL + dtInv +
// f_speedmax, src/BoxOp_EulerCubedSphereImplem.H:878-954,1013-1131,
// src/BoxOp_EulerCubedSphere.H:127-194; both orders share the arithmetic).
//
// Per valid cell of JU: second-order conversion to native primitives with the
// cell-averaged det A, the inverse of the cell-averaged det(A) A and the radial
// volume; pressure from the cell-centred basis, clipped at 0; then
//   inverseDt = sum_d (sqrt(gamma p/rho + |A_d b|^2/(4 pi rho_m)) + |A_d v|) / width_d
// with A_0 the cell-centred basis, A_1, A_2 the basis at the centre of the low
// face of that direction, width_0 = Dr dxi and width_d = dVolr <det A>_d h /
// (adjDr_d |<adjA>_d|). rho_m = min(rho, cone density) when the second-order
// caller passes a positive cone density, else rho. The level result is
// 1 / max over cells; the caller applies its CFL number.
//
// Layout: one table row per angular column (p,k,j) holding the conversion and
// face geometry (kAng doubles), one row per radial index (kRad doubles); the
// cell loop reads JU once (64 B per cell) and reduces the maximum per block,
// a one-block kernel reduces the block maxima.
#include "hc_common.hpp"
#include "hc_geometry.hpp"
#include "hc_physics.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace hc
{
namespace
{
constexpr int kCflBlock = 256;
constexpr int kCflMaxBlocks = 2048;
// angular row: invDet, invDetA[9], cell[9], face1[9], face2[9], det1, det2, normal1, normal2
constexpr int kAng = 41;
enum { A_INVDET = 0, A_INV = 1, A_CELL = 10, A_FACE1 = 19, A_FACE2 = 28, A_DET1 = 37, A_DET2 = 38,
       A_NRM1 = 39, A_NRM2 = 40 };
// radial row: volume, 1/volume, width0, adjDr (1 and 2 are equal)
constexpr int kRad = 4;

struct CflView
{
    int nr, na;
    long long ncell;
    double h, gamma, cone;
    const double* ang;
    const double* rad;
};

__device__ inline double waveMax(double a, double b) { return a < b ? b : a; }

__device__ inline double blockMax(double m)
{
    for (int off = 32; off > 0; off >>= 1) m = waveMax(m, __shfl_xor(m, off, 64));
    __shared__ double part[kCflBlock/64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int w = 1; w < kCflBlock/64; ++w) m = waveMax(m, part[w]);
    }
    return m;
}

__global__ void __launch_bounds__(kCflBlock) k_cfl_cells(CflView v, const double* __restrict__ JU,
                                                         double* __restrict__ partial)
{
    const long long nn = (long long)v.na*v.na;
    const long long stride = (long long)gridDim.x*blockDim.x;
    double localMaximum = 0.0;
    for (long long c = (long long)blockIdx.x*blockDim.x + threadIdx.x; c < v.ncell; c += stride)
    {
        const int i = (int)(c % v.nr);
        const long long col = c / v.nr;            // (p*na + k)*na + j
        const long long p = col / nn, colp = col - p*nn;
        const double* a = v.ang + col*kAng;
        const double* r = v.rad + (long long)i*kRad;
        const double* src = JU + (p*8*nn + colp)*v.nr + i;
        double u[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) u[q] = src[(long long)q*nn*v.nr];
        const double inverseDet = a[A_INVDET], inverseVolume = r[1];
        double rho = u[0]*inverseDet;
        rho *= inverseVolume;
        rho = rho < kRhoFloor ? kRhoFloor : rho;
        double energy = u[4]*inverseDet;
        energy *= inverseVolume;
        double vel[3], b[3];
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            double sv = 0.0, sb = 0.0;
#pragma unroll
            for (int q = 0; q < 3; ++q) { sv += a[A_INV+3*d+q]*u[1+q]; sb += a[A_INV+3*d+q]*u[5+q]; }
            sv *= inverseVolume; sb *= inverseVolume;
            vel[d] = sv/rho; b[d] = sb;
        }
        double kinetic = 0.0, magnetic = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            double vc = 0.0, bc = 0.0;
#pragma unroll
            for (int q = 0; q < 3; ++q) { vc += a[A_CELL+3*d+q]*vel[q]; bc += a[A_CELL+3*d+q]*b[q]; }
            kinetic += vc*vc/2;
            magnetic += bc*bc/8.0/kPi;
        }
        kinetic *= rho;
        double pressure = (energy-kinetic-magnetic)*(v.gamma-1);
        if (pressure < 0) pressure = 0;
        const double magneticDensity = v.cone > 0 ? (rho < v.cone ? rho : v.cone) : rho;
        double inverseDt = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const double* basis = a + (d == 0 ? A_CELL : d == 1 ? A_FACE1 : A_FACE2);
            double v2 = 0.0, fullB2 = 0.0;
#pragma unroll
            for (int q = 0; q < 3; ++q)
            {
                double velocity = 0.0, field = 0.0;
#pragma unroll
                for (int s = 0; s < 3; ++s) { velocity += basis[3*q+s]*vel[s]; field += basis[3*q+s]*b[s]; }
                v2 += velocity*velocity; fullB2 += field*field;
            }
            const double width = d == 0 ? r[2]
                : r[0]*a[d == 1 ? A_DET1 : A_DET2]*v.h/r[3]/a[d == 1 ? A_NRM1 : A_NRM2];
            const double wave = sqrt(v.gamma*pressure/rho + fullB2/4.0/kPi/magneticDensity);
            inverseDt += (wave + sqrt(v2))/width;
        }
        localMaximum = waveMax(localMaximum, inverseDt);
    }
    const double m = blockMax(localMaximum);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

// Conserved-state check (validConservedState, src/BoxOp_EulerCubedSphereImplem.H:813-876):
// a value is bad when it is not finite, or when it is a density that is not
// > 0. Each block keeps the smallest flat index of a bad value.
__device__ inline long long indexMin(long long m)
{
    for (int off = 32; off > 0; off >>= 1)
    {
        const long long o = __shfl_xor(m, off, 64);
        m = o < m ? o : m;
    }
    __shared__ long long part[kCflBlock/64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kCflBlock/64; ++w) m = part[w] < m ? part[w] : m;
    return m;
}

constexpr long long kNoIndex = 0x7fffffffffffffffLL;

__global__ void __launch_bounds__(kCflBlock) k_state_check(const double* __restrict__ U, long long n,
                                                           long long compStride, long long* __restrict__ partial)
{
    const long long stride = (long long)gridDim.x*blockDim.x;
    long long first = kNoIndex;
    for (long long x = (long long)blockIdx.x*blockDim.x + threadIdx.x; x < n; x += stride)
    {
        const double value = U[x];
        const bool density = (x / compStride) % 8 == 0;
        const bool bad = !isfinite(value) || (density && !(value > 0.0));
        if (bad && x < first) first = x;
    }
    const long long m = indexMin(first);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

__global__ void __launch_bounds__(kCflBlock) k_state_final(const long long* __restrict__ partial, int n,
                                                           long long* __restrict__ result)
{
    long long m = kNoIndex;
    for (int b = threadIdx.x; b < n; b += blockDim.x) m = partial[b] < m ? partial[b] : m;
    m = indexMin(m);
    if (threadIdx.x == 0) result[0] = m == kNoIndex ? -1 : m;
}

__global__ void __launch_bounds__(kCflBlock) k_cfl_final(const double* __restrict__ partial, int n,
                                                         double* __restrict__ result)
{
    double m = 0.0;
    for (int b = threadIdx.x; b < n; b += blockDim.x) m = waveMax(m, partial[b]);
    m = blockMax(m);
    if (threadIdx.x == 0) result[0] = 1.0/m;
}
} // namespace

struct CflTable
{
    ShellDims dims{};
    double gamma = 0;
    double* ang = nullptr;
    double* rad = nullptr;
    double* partial = nullptr;   // block maxima, then the result at [kCflMaxBlocks]
    long long* firstBad = nullptr; // block minima of the state check, result at [kCflMaxBlocks]
    bool ready = false;            // ang and rad hold the tables
    ~CflTable()
    {
        if (ang) (void)hipFree(ang);
        if (rad) (void)hipFree(rad);
        if (partial) (void)hipFree(partial);
        if (firstBad) (void)hipFree(firstBad);
    }
};

// Column and radial tables of the CFL arithmetic, built on the first call.
void cflGeometry(CflTable& T)
{
    const ShellDims& dims = T.dims;
    const int na = dims.na, nr = dims.nr;
    const double h = 1.0/na, dxi0 = 1.0/nr;
    std::vector<double> ang((size_t)6*na*na*kAng), rad((size_t)nr*kRad);
    for (int p = 0; p < 6; ++p)
        for (int k = 0; k < na; ++k)
            for (int j = 0; j < na; ++j)
            {
                double* e = &ang[(((size_t)p*na + k)*na + j)*kAng];
                double m0[13], m1[13], m2[13];
                faceMetricAverages(na, p, 0, j, k, m0);
                faceMetricAverages(na, p, 1, j, k, m1);
                faceMetricAverages(na, p, 2, j, k, m2);
                e[A_INVDET] = 1.0/m0[0];
                double detAA[3][3], inv[3][3], A[3][3];
                for (int x = 0; x < 3; ++x) for (int y = 0; y < 3; ++y) detAA[x][y] = m0[1+3*x+y];
                invert3(detAA, inv);
                for (int x = 0; x < 9; ++x) e[A_INV+x] = inv[x/3][x%3];
                pointBasis(p, xiToAngle((j+0.5)*h), xiToAngle((k+0.5)*h), A);
                for (int x = 0; x < 9; ++x) e[A_CELL+x] = A[x/3][x%3];
                pointBasis(p, xiToAngle(j*h), xiToAngle((k+0.5)*h), A);
                for (int x = 0; x < 9; ++x) e[A_FACE1+x] = A[x/3][x%3];
                pointBasis(p, xiToAngle((j+0.5)*h), xiToAngle(k*h), A);
                for (int x = 0; x < 9; ++x) e[A_FACE2+x] = A[x/3][x%3];
                e[A_DET1] = m1[0];
                e[A_DET2] = m2[0];
                const double* rows[2] = {&m1[10], &m2[10]};
                for (int d = 0; d < 2; ++d)
                {
                    double n2 = 0.0;
                    for (int x = 0; x < 3; ++x) n2 += rows[d][x]*rows[d][x];
                    e[A_NRM1+d] = std::sqrt(n2);
                }
            }
    for (int i = 0; i < nr; ++i)
    {
        const RadialCell rc = radialCell(dims, i);
        rad[(size_t)i*kRad + 0] = rc.dvol;
        rad[(size_t)i*kRad + 1] = 1.0/rc.dvol;
        rad[(size_t)i*kRad + 2] = ((rc.rhi - rc.rlo)/dxi0)*dxi0;
        rad[(size_t)i*kRad + 3] = rc.areaT;
    }
    if (!T.ang) HC_THROW(hipMalloc(&T.ang, ang.size()*sizeof(double)));
    if (!T.rad) HC_THROW(hipMalloc(&T.rad, rad.size()*sizeof(double)));
    HC_THROW(hipMemcpy(T.ang, ang.data(), ang.size()*sizeof(double), hipMemcpyHostToDevice));
    HC_THROW(hipMemcpy(T.rad, rad.data(), rad.size()*sizeof(double), hipMemcpyHostToDevice));
    T.ready = true;
}

CflTable* cflCreate(const ShellDims& dims, double gamma)
{
    CflTable* T = new CflTable;
    T->dims = dims;
    T->gamma = gamma;
    try
    {
        HC_THROW(hipMalloc(&T->partial, (kCflMaxBlocks + 1)*sizeof(double)));
        HC_THROW(hipMalloc(&T->firstBad, (kCflMaxBlocks + 1)*sizeof(long long)));
    }
    catch (...)
    {
        delete T;
        throw;
    }
    return T;
}

void cflDestroy(CflTable* T) { delete T; }

// Enqueues the reduction on s; the result (1/max) lands in device memory
// returned by the call, valid once s has reached it.
const double* cflLaunch(CflTable& T, const double* JU, double coneDensity, hipStream_t s)
{
    if (!T.ready) cflGeometry(T);
    CflView v;
    v.nr = T.dims.nr;
    v.na = T.dims.na;
    v.ncell = 6LL*v.na*v.na*v.nr;
    v.h = 1.0/v.na;
    v.gamma = T.gamma;
    v.cone = coneDensity;
    v.ang = T.ang;
    v.rad = T.rad;
    const long long want = (v.ncell + kCflBlock - 1)/kCflBlock;
    const int nb = (int)std::min<long long>(want, kCflMaxBlocks);
    hipLaunchKernelGGL(k_cfl_cells, dim3(nb), dim3(kCflBlock), 0, s, v, JU, T.partial);
    HC_THROW(hipGetLastError());
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(kCflBlock), 0, s, T.partial, nb, T.partial + kCflMaxBlocks);
    HC_THROW(hipGetLastError());
    return T.partial + kCflMaxBlocks;
}

// Enqueues the conserved-state check of the valid level U; the device result
// is the flat index of the first bad value, or -1.
const long long* stateCheckLaunch(CflTable& T, const double* U, hipStream_t s)
{
    const long long compStride = (long long)T.dims.na*T.dims.na*T.dims.nr;
    const long long n = 6LL*8*compStride;
    const int nb = (int)std::min<long long>((n + kCflBlock - 1)/kCflBlock, kCflMaxBlocks);
    hipLaunchKernelGGL(k_state_check, dim3(nb), dim3(kCflBlock), 0, s, U, n, compStride, T.firstBad);
    HC_THROW(hipGetLastError());
    hipLaunchKernelGGL(k_state_final, dim3(1), dim3(kCflBlock), 0, s, T.firstBad, nb, T.firstBad + kCflMaxBlocks);
    HC_THROW(hipGetLastError());
    return T.firstBad + kCflMaxBlocks;
}
} // namespace hc
