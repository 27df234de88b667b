// Level driver and O2 kernels of the HelioCubed MHD right-hand side on
// MI355X (gfx950). The C ABI is declared in include/heliocubed.h.
//
// Data: one ghosted FP64 array per stage for the whole shell (panel-major,
// component-major, radial index unit-stride), so intra-panel box halos need
// no copies; only the seam ghost ring of each panel is gathered.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <algorithm>
#include <array>
#include <vector>

#include "../../include/heliocubed.h"
#include "hc_common.hpp"
#include "hc_geometry.hpp"
#include "hc_physics.hpp"

std::string& hc::lastError()
{
    thread_local std::string message;
    return message;
}

namespace
{
using hc::fail;
#define HC_CHECK HC_THROW

constexpr int kBlock = 256;

// Geometry and sizes handed to every kernel by value.
struct LevelView
{
    int nr, na, g, NI, NJ, NK;
    long long cs, ps;      // stage component / panel strides
    long long vcs, vps;    // valid component / panel strides
    int limiter, roe;
    int powell;            // divergence_cleaning 1: Powell source terms
    double gamma, entropy;
    const double* radial;  // [nr+2g+1][4]: rmid, dvol, area0, areaT for i = -g ..
    const double* detA;    // [na][na]
    const double* frame[3];// [6][fk][fj][10]: 9 frame entries + |normal|
    const double* normal[3];// [6][fk][fj][3]: integrated face normal (Powell source only)
    const double* cellA;   // [6][NK][NJ][9]
    const double* cellAi;  // [6][NK][NJ][9]
    const int* edge;       // [6][2][2][4]: q, dq, sq, rev
};

__device__ __forceinline__ long long stageIndex(const LevelView& v, int p, int c, int i, int j, int k)
{
    return p*v.ps + c*v.cs + ((long long)(k+v.g)*v.NJ + (j+v.g))*v.NI + (i+v.g);
}

__device__ __forceinline__ long long validIndex(const LevelView& v, int p, int c, int i, int j, int k)
{
    return p*v.vps + c*v.vcs + ((long long)k*v.na + j)*v.nr + i;
}

__device__ __forceinline__ const double* rad(const LevelView& v, int i) { return v.radial + 4*(i+v.g); }

__device__ __forceinline__ void loadStage(const LevelView& v, const double* W, int p, int i, int j, int k, double w[8])
{
    const long long base = stageIndex(v, p, 0, i, j, k);
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = W[base + c*v.cs];
}

// 1. Valid cells: stage state, conversion to Cartesian primitives.
// Plain copy (the RK4 stage base state), a kernel so the step graph holds no
// memory-copy or memset nodes.
__global__ void __launch_bounds__(kBlock) k_copy(const double* __restrict__ src, double* __restrict__ dst,
                                                 long long n)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

__global__ void __launch_bounds__(kBlock)
k_prepare_valid(LevelView v, const double* __restrict__ U0, const double* __restrict__ K,
                double a_dt, double* __restrict__ W)
{
    const long long n = 6LL*v.na*v.na*v.nr;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = t % v.nr;
    long long r = t / v.nr;
    const int j = r % v.na; r /= v.na;
    const int k = r % v.na;
    const int p = r / v.na;
    double u[8], w[8];
    const long long vb = validIndex(v, p, 0, i, j, k);
#pragma unroll
    for (int c = 0; c < 8; ++c)
        u[c] = K ? U0[vb + c*v.vcs] + a_dt*K[vb + c*v.vcs] : U0[vb + c*v.vcs];
    const double invJ = 1.0/(rad(v,i)[1]*v.detA[k*v.na + j]);
    hc::conservedToPrimitive(w, u, invJ, v.gamma);
    const long long sb = stageIndex(v, p, 0, i, j, k);
#pragma unroll
    for (int c = 0; c < 8; ++c) W[sb + c*v.cs] = w[c];
}

// 2. Seam ghost columns: weighted sums of valid donor columns, then floors.
__global__ void __launch_bounds__(kBlock)
k_seam_gather(LevelView v, int ncol, const long long* __restrict__ colDst,
              const int* __restrict__ colStart, const long long* __restrict__ donorSrc,
              const double* __restrict__ donorW, double* __restrict__ W)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= (long long)ncol*v.nr) return;
    const int i = t % v.nr;
    const int col = t / v.nr;
    const int d0 = colStart[col], d1 = colStart[col+1];
    const long long dst = colDst[col] + i;
    // donors outer, components inner (each component sums in table order)
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int d = d0; d < d1; ++d)
    {
        const double wd = donorW[d];
        const double* sp = W + donorSrc[d] + i;
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] += wd*sp[c*v.cs];
    }
    acc[0] = fmax(acc[0], hc::kRhoFloor);
    acc[4] = fmax(acc[4], 0.0);
#pragma unroll
    for (int c = 0; c < 8; ++c) W[dst + c*v.cs] = acc[c];
}

// 3. Radial boundary ghosts: inner fixed primitives, outer copy of the last
// layer scaled by (r_last/r)^2 (r for the two angular B components) in the
// native basis, src/BoxOp_PreStage.H:105-141, src/BoxOp_SecondOrderPreStage.H:169-215.
__global__ void __launch_bounds__(kBlock)
k_radial_bc(LevelView v, const double* __restrict__ inner, double* __restrict__ W)
{
    const int n = 6*v.NJ*v.NK;
    const int t = blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int jj = t % v.NJ;
    const int kk = (t / v.NJ) % v.NK;
    const int p = t / (v.NJ*v.NK);
    const int j = jj - v.g, k = kk - v.g;
    for (int ig = 0; ig < v.g; ++ig)
        for (int c = 0; c < 8; ++c)
            W[stageIndex(v, p, c, -v.g+ig, j, k)] = inner[((((long long)p*8+c)*v.NK+kk)*v.NJ+jj)*v.g+ig];
    const double* A = v.cellA + 9LL*(((long long)p*v.NK+kk)*v.NJ+jj);
    const double* Ai = v.cellAi + 9LL*(((long long)p*v.NK+kk)*v.NJ+jj);
    double w[8], u[8];
    loadStage(v, W, p, v.nr-1, j, k, w);
    double kinetic = 0, magnetic = 0;
    u[0] = w[0];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        kinetic += 0.5*w[0]*w[1+d]*w[1+d];
        magnetic += w[5+d]*w[5+d]/(8*hc::kPi);
        double m = 0, b = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) { m += Ai[3*d+q]*w[0]*w[1+q]; b += Ai[3*d+q]*w[5+q]; }
        u[1+d] = m; u[5+d] = b;
    }
    u[4] = w[4]/(v.gamma-1)+kinetic+magnetic;
    const double rlast = rad(v, v.nr-1)[0];
    for (int i = v.nr; i < v.nr+v.g; ++i)
    {
        const double ratio = rlast/rad(v,i)[0];
        double ug[8], cg[8], wg[8];
#pragma unroll
        for (int c = 0; c < 6; ++c) ug[c] = u[c]*(ratio*ratio);
        ug[6] = u[6]*ratio; ug[7] = u[7]*ratio;
        cg[0] = ug[0]; cg[4] = ug[4];
#pragma unroll
        for (int grp = 0; grp < 2; ++grp)
        {
            const int st = grp == 0 ? 1 : 5;
#pragma unroll
            for (int d = 0; d < 3; ++d)
            {
                double s = 0;
#pragma unroll
                for (int q = 0; q < 3; ++q) s += A[3*d+q]*ug[st+q];
                cg[st+d] = s;
            }
        }
        hc::conservedToPrimitive(wg, cg, 1.0, v.gamma);
        const long long sb = stageIndex(v, p, 0, i, j, k);
#pragma unroll
        for (int c = 0; c < 8; ++c) W[sb + c*v.cs] = wg[c];
    }
}

// 4. Interior reconstructed state of every panel at its own seam faces,
// src/BoxOp_SecondOrder.H:80-97. PS[p][d-1][s][c][t][i].
__global__ void __launch_bounds__(kBlock)
k_panel_states(LevelView v, const double* __restrict__ W, double* __restrict__ PS)
{
    const long long n = 24LL*v.na*v.nr;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = t % v.nr;
    long long r = t / v.nr;
    const int tt = r % v.na; r /= v.na;
    const int s = r % 2; r /= 2;
    const int d = 1 + (r % 2);
    const int p = r / 2;
    const int f = s ? v.na : 0;
    double mm[8], m[8], c[8], pp[8], out[8];
    auto at = [&](int off, double o[8]) {
        if (d == 1) loadStage(v, W, p, i, f+off, tt, o); else loadStage(v, W, p, i, tt, f+off, o);
    };
    if (s == 0) { at(-1, m); at(0, c); at(1, pp); hc::plmHighSide(out, m, c, pp, v.limiter); }
    else { at(-2, mm); at(-1, m); at(0, c); hc::plmLowSide(out, mm, m, c, v.limiter); }
    const long long base = ((((long long)(p*2+(d-1))*2+s)*8)*v.na + tt)*v.nr + i;
    const long long cstride = (long long)v.na*v.nr;
#pragma unroll
    for (int q = 0; q < 8; ++q) PS[base + q*cstride] = out[q];
}

// 5. Face fluxes of direction D for every face of the level.
template <int D, bool ROE>
__global__ void __launch_bounds__(kBlock)
k_face_flux(LevelView v, const double* __restrict__ W, const double* __restrict__ PS,
            double* __restrict__ F)
{
    const int fi = v.nr + (D == 0), fj = v.na + (D == 1), fk = v.na + (D == 2);
    const long long n = 6LL*fi*fj*fk;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = t % fi;
    long long r = t / fi;
    const int j = r % fj; r /= fj;
    const int k = r % fk;
    const int p = r / fk;
    constexpr int e0 = D == 0, e1 = D == 1, e2 = D == 2;
    double mm[8], m[8], c[8], pp[8], l[8], rr[8];
    loadStage(v, W, p, i-2*e0, j-2*e1, k-2*e2, mm);
    loadStage(v, W, p, i-e0, j-e1, k-e2, m);
    loadStage(v, W, p, i, j, k, c);
    loadStage(v, W, p, i+e0, j+e1, k+e2, pp);
    hc::plmFaceStates(l, rr, mm, m, c, pp, v.limiter);
    if (D > 0)
    {
        const int f = D == 1 ? j : k, tt = D == 1 ? k : j;
        if (f == 0 || f == v.na)
        {
            const int s = f == 0 ? 0 : 1;
            const int* e = v.edge + 4*((p*2 + (D-1))*2 + s);
            const int tq = e[3] ? v.na-1-tt : tt;
            const long long base = ((((long long)(e[0]*2+(e[1]-1))*2+e[2])*8)*v.na + tq)*v.nr + i;
            const long long cstride = (long long)v.na*v.nr;
            double* dst = s == 0 ? l : rr;
#pragma unroll
            for (int q = 0; q < 8; ++q) dst[q] = PS[base + q*cstride];
        }
    }
    const double* fr = v.frame[D] + 10LL*(((long long)p*fk + k)*fj + j);
    const double area = (D == 0 ? rad(v,i)[2] : rad(v,i)[3])*fr[9];
    double out[8];
    hc::o2FaceFluxMixed<double, ROE ? 2 : 1>(out, l, rr, fr, area, v.gamma, v.entropy, ROE);
    const long long cstride = (long long)fi*fj*fk;
    const long long base = (long long)p*8*cstride + ((long long)k*fj + j)*fi + i;
#pragma unroll
    for (int q = 0; q < 8; ++q) F[base + q*cstride] = out[q];
}

// 6. Seam flux reconciliation 1/2 (F_local + sign F_adjacent), each physical
// seam face once, src/MHD_AMR.H:1484-1496.
__global__ void __launch_bounds__(kBlock)
k_seam_reconcile(LevelView v, double* __restrict__ F1, double* __restrict__ F2)
{
    const long long n = 24LL*v.na*v.nr;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = t % v.nr;
    long long r = t / v.nr;
    const int tt = r % v.na; r /= v.na;
    const int s = r % 2; r /= 2;
    const int d = 1 + (r % 2);
    const int p = r / 2;
    const int* e = v.edge + 4*((p*2 + (d-1))*2 + s);
    const int q = e[0], dq = e[1], sq = e[2];
    const int self = (p*2 + (d-1))*2 + s, other = (q*2 + (dq-1))*2 + sq;
    if (self > other) return;
    const double sign = (s != sq) ? 1.0 : -1.0;
    const int tq = e[3] ? v.na-1-tt : tt;
    auto faceBase = [&](int dd, int pp, int f, int tg, double*& arr, long long& cstr) {
        const int fj = v.na + (dd == 1), fk = v.na + (dd == 2);
        cstr = (long long)v.nr*fj*fk;
        const int jj = dd == 1 ? f : tg, kk = dd == 1 ? tg : f;
        arr = dd == 1 ? F1 : F2;
        return (long long)pp*8*cstr + ((long long)kk*fj + jj)*v.nr + i;
    };
    double* Fa; double* Fb; long long csa, csb;
    const long long a = faceBase(d, p, s ? v.na : 0, tt, Fa, csa);
    const long long b = faceBase(dq, q, sq ? v.na : 0, tq, Fb, csb);
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        const double fa = Fa[a + c*csa], fb = Fb[b + c*csb];
        Fa[a + c*csa] = 0.5*(fa + sign*fb);
        Fb[b + c*csb] = 0.5*(fb + sign*fa);
    }
}

// 7. Flux divergence into K, fused with the RK accumulate U += b_dt*K.
__global__ void __launch_bounds__(kBlock)
k_divergence(LevelView v, const double* __restrict__ F0, const double* __restrict__ F1,
             const double* __restrict__ F2, double* __restrict__ K, double* __restrict__ U, double b_dt)
{
    const long long n = 6LL*v.na*v.na*v.nr;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = t % v.nr;
    long long r = t / v.nr;
    const int j = r % v.na; r /= v.na;
    const int k = r % v.na;
    const int p = r / v.na;
    const double inv[3] = {-1.0/(1.0/v.nr), -1.0/(1.0/v.na), -1.0/(1.0/v.na)};
    const long long cs0 = (long long)(v.nr+1)*v.na*v.na, cs1 = (long long)v.nr*(v.na+1)*v.na;
    const long long b0 = p*8*cs0 + ((long long)k*v.na + j)*(v.nr+1) + i;
    const long long b1 = p*8*cs1 + ((long long)k*(v.na+1) + j)*v.nr + i;
    const long long b2 = p*8*cs1 + ((long long)k*v.na + j)*v.nr + i;
    const long long vb = validIndex(v, p, 0, i, j, k);
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        double s = 0;
        s += inv[0]*(F0[b0 + c*cs0 + 1] - F0[b0 + c*cs0]);
        s += inv[1]*(F1[b1 + c*cs1 + v.nr] - F1[b1 + c*cs1]);
        s += inv[2]*(F2[b2 + c*cs1 + (long long)v.na*v.nr] - F2[b2 + c*cs1]);
        K[vb + c*v.vcs] = s;
        if (U) U[vb + c*v.vcs] += b_dt*s;
    }
}

// 8. Seam faces: both panels' fluxes of every physical seam face from the
// two interior states, reconciled 1/2 (F_local + sign F_adjacent), stored for
// both panels in SF[p][d-1][s][c][t][i]. Each physical face is handled once.
template <typename S, bool ROE>
__global__ void __launch_bounds__(kBlock)
k_seam_flux(LevelView v, const double* __restrict__ PS, double* __restrict__ SF)
{
    const long long n = 24LL*v.na*v.nr;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = t % v.nr;
    long long r = t / v.nr;
    const int tt = r % v.na; r /= v.na;
    const int s = r % 2; r /= 2;
    const int d = 1 + (r % 2);
    const int p = r / 2;
    const int* e = v.edge + 4*((p*2 + (d-1))*2 + s);
    const int q = e[0], dq = e[1], sq = e[2];
    const int self = (p*2 + (d-1))*2 + s, other = (q*2 + (dq-1))*2 + sq;
    if (self > other) return;
    const int tq = e[3] ? v.na-1-tt : tt;
    const long long cstride = (long long)v.na*v.nr;
    const long long bp = ((long long)self*8*v.na + tt)*v.nr + i;
    const long long bq = ((long long)other*8*v.na + tq)*v.nr + i;
    double sp[8], sqs[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) { sp[c] = PS[bp + c*cstride]; sqs[c] = PS[bq + c*cstride]; }
    auto flux = [&](int pp, int dd, int ss, int tg, const double* own, const double* adj, double out[8]) {
        const int fj = v.na + (dd == 1), fk = v.na + (dd == 2);
        const int f = ss ? v.na : 0;
        const int jj = dd == 1 ? f : tg, kk = dd == 1 ? tg : f;
        const double* fr = v.frame[dd] + 10LL*(((long long)pp*fk + kk)*fj + jj);
        const double area = rad(v,i)[3]*fr[9];
        if (ss == 0) hc::o2FaceFluxMixed<S, ROE ? 2 : 1>(out, adj, own, fr, area, v.gamma, v.entropy, ROE);
        else hc::o2FaceFluxMixed<S, ROE ? 2 : 1>(out, own, adj, fr, area, v.gamma, v.entropy, ROE);
    };
    double fp[8], fq[8];
    flux(p, d, s, tt, sp, sqs, fp);
    flux(q, dq, sq, tq, sqs, sp, fq);
    const double sign = (s != sq) ? 1.0 : -1.0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        SF[bp + c*cstride] = 0.5*(fp[c] + sign*fq[c]);
        SF[bq + c*cstride] = 0.5*(fq[c] + sign*fp[c]);
    }
}

// 9. Fused RHS: one block per tile of TI x TJ x TK cells of one panel. For
// each direction the block evaluates every face of the tile once into LDS
// (seam faces come reconciled from SF), then each thread accumulates the
// divergence of its cell. K = d<JU>/dt; U += b_dt*K when U is given.
constexpr int kTI = 32, kTJ = 4, kTK = 2;
constexpr int kTileCells = kTI*kTJ*kTK;
constexpr int kMax2(int a, int b) { return a > b ? a : b; }
constexpr int kMaxFaces = kMax2(kMax2((kTI+1)*kTJ*kTK, kTI*(kTJ+1)*kTK), kTI*kTJ*(kTK+1));

template <typename S, bool ROE>
__global__ void __launch_bounds__(kTileCells)
k_rhs_fused(LevelView v, const double* __restrict__ W, const double* __restrict__ SF,
            double* __restrict__ K, double* __restrict__ U, double b_dt, int tilesI, int tilesJ, int tilesK)
{
    __shared__ double flux[8*kMaxFaces];
    int b = blockIdx.x;
    const int ti = b % tilesI; b /= tilesI;
    const int tj = b % tilesJ; b /= tilesJ;
    const int tk = b % tilesK;
    const int p = b / tilesK;
    const int i0 = ti*kTI, j0 = tj*kTJ, k0 = tk*kTK;
    const int tid = threadIdx.x;
    const int ci = tid % kTI, cj = (tid / kTI) % kTJ, ck = tid / (kTI*kTJ);
    const bool active = i0+ci < v.nr && j0+cj < v.na && k0+ck < v.na;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const double inv[3] = {-1.0/(1.0/v.nr), -1.0/(1.0/v.na), -1.0/(1.0/v.na)};
#pragma unroll 1
    for (int d = 0; d < 3; ++d)
    {
        const int ni = kTI + (d == 0), nj = kTJ + (d == 1), nk = kTK + (d == 2);
        const int nf = ni*nj*nk;
        for (int f = tid; f < nf; f += kTileCells)
        {
            const int fi = f % ni, fj = (f / ni) % nj, fk = f / (ni*nj);
            const int i = i0+fi, j = j0+fj, k = k0+fk;
            double out[8];
            const bool inside = i <= v.nr - (d != 0) && j <= v.na - (d != 1) && k <= v.na - (d != 2);
            if (inside)
            {
                const int fa = d == 1 ? j : k, tg = d == 1 ? k : j;
                if (d > 0 && (fa == 0 || fa == v.na))
                {
                    const int s = fa == 0 ? 0 : 1;
                    const long long base = ((((long long)(p*2+(d-1))*2+s)*8)*v.na + tg)*v.nr + i;
                    const long long cstride = (long long)v.na*v.nr;
#pragma unroll
                    for (int c = 0; c < 8; ++c) out[c] = SF[base + c*cstride];
                }
                else
                {
                    const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
                    double mm[8], m[8], c[8], pp[8], l[8], rr[8];
                    loadStage(v, W, p, i-2*e0, j-2*e1, k-2*e2, mm);
                    loadStage(v, W, p, i-e0, j-e1, k-e2, m);
                    loadStage(v, W, p, i, j, k, c);
                    loadStage(v, W, p, i+e0, j+e1, k+e2, pp);
                    hc::plmFaceStatesMixed<S>(l, rr, mm, m, c, pp, v.limiter);
                    const int fjd = v.na + (d == 1), fkd = v.na + (d == 2);
                    const double* fr = v.frame[d] + 10LL*(((long long)p*fkd + k)*fjd + j);
                    const double area = (d == 0 ? rad(v,i)[2] : rad(v,i)[3])*fr[9];
                    hc::o2FaceFluxMixed<S, ROE ? 2 : 1>(out, l, rr, fr, area, v.gamma, v.entropy, ROE);
                }
            }
            else
            {
#pragma unroll
                for (int c = 0; c < 8; ++c) out[c] = 0;
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) flux[c*kMaxFaces + f] = out[c];
        }
        __syncthreads();
        const int flo = (ck*nj + cj)*ni + ci;
        const int fhi = flo + (d == 0 ? 1 : d == 1 ? ni : ni*nj);
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] += inv[d]*(flux[c*kMaxFaces + fhi] - flux[c*kMaxFaces + flo]);
        __syncthreads();
    }
    if (!active) return;
    const long long vb = validIndex(v, p, 0, i0+ci, j0+cj, k0+ck);
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        K[vb + c*v.vcs] = acc[c];
        if (U) U[vb + c*v.vcs] += b_dt*acc[c];
    }
}

// 10. Powell source (divergence_cleaning 1), src/BoxOp_SecondOrder.H:317-369.
// (a) Magnetic flux 1/2 (B_l + B_r).n area of every face of direction D, n the
// integrated face normal, from the face states of the flux kernels: PLM in
// precision S, and at a seam face the adjacent panel's state from PS on that
// side. Each panel keeps its own value at a seam face (not reconciled).
template <int D, typename S>
__global__ void __launch_bounds__(kBlock)
k_powell_bflux(LevelView v, const double* __restrict__ W, const double* __restrict__ PS, double* __restrict__ BF)
{
    const int fi = v.nr + (D == 0), fj = v.na + (D == 1), fk = v.na + (D == 2);
    const long long n = 6LL*fi*fj*fk;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = t % fi;
    long long r = t / fi;
    const int j = r % fj; r /= fj;
    const int k = r % fk;
    const int p = r / fk;
    constexpr int e0 = D == 0, e1 = D == 1, e2 = D == 2;
    double mm[8], m[8], c[8], pp[8], l[8], rr[8];
    loadStage(v, W, p, i-2*e0, j-2*e1, k-2*e2, mm);
    loadStage(v, W, p, i-e0, j-e1, k-e2, m);
    loadStage(v, W, p, i, j, k, c);
    loadStage(v, W, p, i+e0, j+e1, k+e2, pp);
    hc::plmFaceStatesMixed<S>(l, rr, mm, m, c, pp, v.limiter);
    if (D > 0)
    {
        const int f = D == 1 ? j : k, tt = D == 1 ? k : j;
        if (f == 0 || f == v.na)
        {
            const int s = f == 0 ? 0 : 1;
            const int* e = v.edge + 4*((p*2 + (D-1))*2 + s);
            const int tq = e[3] ? v.na-1-tt : tt;
            const long long base = ((((long long)(e[0]*2+(e[1]-1))*2+e[2])*8)*v.na + tq)*v.nr + i;
            const long long cstride = (long long)v.na*v.nr;
            double* dst = s == 0 ? l : rr;
#pragma unroll
            for (int q = 0; q < 8; ++q) dst[q] = PS[base + q*cstride];
        }
    }
    const double* nr = v.normal[D] + 3LL*(((long long)p*fk + k)*fj + j);
    const double area = D == 0 ? rad(v,i)[2] : rad(v,i)[3];
    double f = 0;
#pragma unroll
    for (int x = 0; x < 3; ++x) f += 0.5*(l[5+x]+rr[5+x])*nr[x]*area;
    BF[((long long)p*fk + k)*fj*fi + (long long)j*fi + i] = f;
}

// (b) -J div B of the magnetic face fluxes times (B/4pi, v, v.B/4pi) of the
// cell's Cartesian primitives, added to K; then the RK stage accumulate
// U += b_dt*K when U is given (the divergence kernel leaves U alone when the
// Powell source is on, so U takes the complete K in one rounding).
__global__ void __launch_bounds__(kBlock)
k_powell_source(LevelView v, const double* __restrict__ W, const double* __restrict__ B0,
                const double* __restrict__ B1, const double* __restrict__ B2, double* __restrict__ K,
                double* __restrict__ U, double b_dt)
{
    const long long n = 6LL*v.na*v.na*v.nr;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = t % v.nr;
    long long r = t / v.nr;
    const int j = r % v.na; r /= v.na;
    const int k = r % v.na;
    const int p = r / v.na;
    const double inv[3] = {-1.0/(1.0/v.nr), -1.0/(1.0/v.na), -1.0/(1.0/v.na)};
    const long long b0 = (((long long)p*v.na + k)*v.na + j)*(v.nr+1) + i;
    const long long b1 = (((long long)p*v.na + k)*(v.na+1) + j)*v.nr + i;
    const long long b2 = (((long long)p*(v.na+1) + k)*v.na + j)*v.nr + i;
    double md = 0;
    md += inv[0]*(B0[b0 + 1] - B0[b0]);
    md += inv[1]*(B1[b1 + v.nr] - B1[b1]);
    md += inv[2]*(B2[b2 + (long long)v.na*v.nr] - B2[b2]);
    double w[8];
    loadStage(v, W, p, i, j, k, w);
    double src[8], vb = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const double b = w[5+d];
        src[1+d] = md*b/(4.0*hc::kPi);
        src[5+d] = md*w[1+d];
        vb += w[1+d]*b;
    }
    src[4] = md*vb/(4.0*hc::kPi);
    const long long vbase = validIndex(v, p, 0, i, j, k);
#pragma unroll
    for (int c = 1; c < 8; ++c) K[vbase + c*v.vcs] += src[c];
    if (U)
#pragma unroll
        for (int c = 0; c < 8; ++c) U[vbase + c*v.vcs] += b_dt*K[vbase + c*v.vcs];
}

inline unsigned blocksFor(long long n) { return (unsigned)((n + kBlock - 1)/kBlock); }

template <typename T>
T* deviceCopy(const std::vector<T>& host)
{
    T* p = nullptr;
    HC_CHECK(hipMalloc(&p, std::max<size_t>(1, host.size())*sizeof(T)));
    if (!host.empty()) HC_CHECK(hipMemcpy(p, host.data(), host.size()*sizeof(T), hipMemcpyHostToDevice));
    return p;
}
} // namespace

// Optional per-kernel HIP-event timing on the launch stream.
struct KernelProfile
{
    bool enabled = false;
    std::vector<std::string> names;
    std::vector<std::vector<std::pair<hipEvent_t, hipEvent_t>>> events;

    int slot(const char* name)
    {
        for (size_t i = 0; i < names.size(); ++i) if (names[i] == name) return (int)i;
        names.emplace_back(name);
        events.emplace_back();
        return (int)names.size()-1;
    }
    template <typename Launch>
    void run(const char* name, hipStream_t s, Launch&& launch)
    {
        if (!enabled) { launch(); return; }
        hipEvent_t a, b;
        HC_CHECK(hipEventCreate(&a));
        HC_CHECK(hipEventCreate(&b));
        HC_CHECK(hipEventRecord(a, s));
        launch();
        HC_CHECK(hipEventRecord(b, s));
        events[slot(name)].emplace_back(a, b);
    }
    void clear()
    {
        for (auto& v : events) for (auto& e : v) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        for (auto& v : events) v.clear();
    }
    ~KernelProfile() { clear(); }
};

struct hc_level
{
    KernelProfile profile;
    hc::Level4* l4 = nullptr;     // fourth-order level (spatial_order 4)
    hc::CflTable* cfl = nullptr;  // CFL geometry, built on the first time-step query
    hc::BcTable* bcTable = nullptr; // boundary-file ghost conversion, built on first use
    hc::ShellDims dims{};
    hc_options opt{};
    LevelView view{};
    int ncol = 0;
    long long* colDst = nullptr;
    int* colStart = nullptr;
    long long* donorSrc = nullptr;
    double* donorW = nullptr;
    std::vector<void*> owned;
    // workspace
    double* W = nullptr;
    double* F[3] = {nullptr, nullptr, nullptr};
    double* PS = nullptr;
    double* SF = nullptr;
    double* BF[3] = {nullptr, nullptr, nullptr};   // Powell magnetic face fluxes
    double* U0 = nullptr;
    double* K = nullptr;
    // hipGraph of one RK4 step (state copy + 4 x (prepare, RHS)), captured on a
    // private stream and replayed on the caller's; re-captured when the state,
    // inner-ghost pointers or dt change.
    bool useGraph = true;
    hipStream_t capStream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    const double* gU = nullptr;
    const double* gInner = nullptr;
    double gDt = 0.0;
    double* consSum = nullptr;      // radialCons accumulator (device double[8]) or none
    double* gCons = nullptr;

    void dropGraph()
    {
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (graph) (void)hipGraphDestroy(graph);
        graphExec = nullptr; graph = nullptr; gU = nullptr; gInner = nullptr;
    }

    ~hc_level()
    {
        dropGraph();
        if (capStream) (void)hipStreamDestroy(capStream);
        hc::level4Destroy(l4);
        hc::cflDestroy(cfl);
        hc::bcDestroy(bcTable);
        for (void* p : owned) (void)hipFree(p);
    }
    template <typename T> T* own(T* p) { owned.push_back(p); return p; }

    size_t validCount() const { return (size_t)6*8*dims.na*dims.na*dims.nr; }
    size_t stageCount() const { return (size_t)6*8*view.NI*view.NJ*view.NK; }
    size_t innerCount() const { return (size_t)6*8*view.NJ*view.NK*dims.g; }
    size_t faceCount(int d) const
    { return (size_t)6*8*(dims.nr+(d==0))*(dims.na+(d==1))*(dims.na+(d==2)); }

    void ensureWorkspace()
    {
        if (W) return;
        HC_CHECK(hipMalloc(&W, stageCount()*sizeof(double))); owned.push_back(W);
        HC_CHECK(hipMemset(W, 0, stageCount()*sizeof(double)));
        if (l4)
        {
            HC_CHECK(hipMalloc(&U0, validCount()*sizeof(double))); owned.push_back(U0);
            HC_CHECK(hipMalloc(&K, validCount()*sizeof(double))); owned.push_back(K);
            return;
        }
        for (int d = 0; d < 3; ++d) { HC_CHECK(hipMalloc(&F[d], faceCount(d)*sizeof(double))); owned.push_back(F[d]); }
        HC_CHECK(hipMalloc(&PS, 24ull*8*dims.na*dims.nr*sizeof(double))); owned.push_back(PS);
        HC_CHECK(hipMalloc(&SF, 24ull*8*dims.na*dims.nr*sizeof(double))); owned.push_back(SF);
        HC_CHECK(hipMalloc(&U0, validCount()*sizeof(double))); owned.push_back(U0);
        HC_CHECK(hipMalloc(&K, validCount()*sizeof(double))); owned.push_back(K);
        if (view.powell)
            for (int d = 0; d < 3; ++d) { HC_CHECK(hipMalloc(&BF[d], faceCount(d)/8*sizeof(double))); owned.push_back(BF[d]); }
    }
};

namespace
{
void buildLevel(hc_level& L)
{
    const int nr = L.dims.nr, na = L.dims.na, g = L.dims.g;
    LevelView& v = L.view;
    v.nr = nr; v.na = na; v.g = g;
    v.NI = nr+2*g; v.NJ = na+2*g; v.NK = na+2*g;
    v.cs = (long long)v.NI*v.NJ*v.NK; v.ps = 8*v.cs;
    v.vcs = (long long)nr*na*na; v.vps = 8*v.vcs;
    v.limiter = L.opt.limiter;
    v.roe = L.opt.riemann_solver == 2;
    v.powell = L.opt.divergence_cleaning == 1;
    v.gamma = L.opt.gamma;
    v.entropy = L.opt.entropy_fix;

    std::vector<double> radial(4*(nr+2*g+1));
    for (int i = -g; i <= nr+g; ++i)
    {
        const hc::RadialCell c = hc::radialCell(L.dims, i);
        double* r = &radial[4*(i+g)];
        r[0] = c.rmid; r[1] = c.dvol; r[2] = c.area0; r[3] = c.areaT;
    }
    v.radial = L.own(deviceCopy(radial));

    std::vector<double> detA((size_t)na*na);
    for (int k = 0; k < na; ++k) for (int j = 0; j < na; ++j) detA[(size_t)k*na+j] = hc::cellDetA(na, j, k);
    v.detA = L.own(deviceCopy(detA));

    for (int d = 0; d < 3; ++d)
    {
        const int fj = na + (d == 1), fk = na + (d == 2);
        std::vector<double> fr(10ull*6*fj*fk), nrm(3ull*6*fj*fk);
        for (int p = 0; p < 6; ++p) for (int k = 0; k < fk; ++k) for (int j = 0; j < fj; ++j)
        {
            const hc::Vec3 n = hc::faceNormal(na, p, d, j, k);
            const hc::FaceFrame F = hc::makeFaceFrame(n.x);
            double* o = &fr[10*(((size_t)p*fk+k)*fj+j)];
            for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) o[3*a+b] = F.f[a][b];
            o[9] = F.norm;
            for (int a = 0; a < 3; ++a) nrm[3*(((size_t)p*fk+k)*fj+j)+a] = n.x[a];
        }
        v.frame[d] = L.own(deviceCopy(fr));
        v.normal[d] = v.powell ? L.own(deviceCopy(nrm)) : nullptr;
    }

    const int NJ = v.NJ, NK = v.NK;
    std::vector<double> A(9ull*6*NJ*NK), Ai(9ull*6*NJ*NK);
    const double h = 1.0/na;
    for (int p = 0; p < 6; ++p) for (int kk = 0; kk < NK; ++kk) for (int jj = 0; jj < NJ; ++jj)
    {
        double a[3][3], ai[3][3];
        hc::pointBasis(p, hc::xiToAngle((jj-g+0.5)*h), hc::xiToAngle((kk-g+0.5)*h), a);
        hc::invert3(a, ai);
        const size_t x = 9*(((size_t)p*NK+kk)*NJ+jj);
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) { A[x+3*r+c] = a[r][c]; Ai[x+3*r+c] = ai[r][c]; }
    }
    v.cellA = L.own(deviceCopy(A));
    v.cellAi = L.own(deviceCopy(Ai));

    std::vector<int> edge(6*2*2*4);
    for (int p = 0; p < 6; ++p) for (int d = 1; d < 3; ++d) for (int s = 0; s < 2; ++s)
    {
        const hc::SeamEdge e = hc::seamEdge(na, p, d, s);
        int* o = &edge[4*((p*2+(d-1))*2+s)];
        o[0] = e.q; o[1] = e.dq; o[2] = e.sq; o[3] = e.rev;
    }
    v.edge = L.own(deviceCopy(edge));

    std::vector<long long> colDst, donorSrc;
    std::vector<int> colStart{0};
    std::vector<double> donorW;
    auto stageOff = [&](int p, int j, int k) {
        return (long long)p*v.ps + ((long long)(k+g)*NJ + (j+g))*v.NI + g;
    };
    for (int p = 0; p < 6; ++p) for (int k = -g; k < na+g; ++k) for (int j = -g; j < na+g; ++j)
    {
        if (j >= 0 && j < na && k >= 0 && k < na) continue;
        for (const auto& d : hc::ghostColumnDonors(na, p, j, k))
        {
            donorSrc.push_back(stageOff(d.q, d.j, d.k));
            donorW.push_back(d.w);
        }
        colDst.push_back(stageOff(p, j, k));
        colStart.push_back((int)donorW.size());
    }
    L.ncol = (int)colDst.size();
    L.colDst = L.own(deviceCopy(colDst));
    L.colStart = L.own(deviceCopy(colStart));
    L.donorSrc = L.own(deviceCopy(donorSrc));
    L.donorW = L.own(deviceCopy(donorW));
}

void prepareStage(hc_level& L, const double* U0, const double* K, double a_dt,
                  const double* inner, double* W, hipStream_t s)
{
    if (L.l4)
    {
        L.profile.run("k4_prepare", s, [&] { hc::level4Prepare(*L.l4, U0, K, a_dt, inner, W, s); });
        return;
    }
    const LevelView& v = L.view;
    KernelProfile& P = L.profile;
    P.run("k_prepare_valid", s, [&] {
        hipLaunchKernelGGL(k_prepare_valid, dim3(blocksFor(6LL*v.na*v.na*v.nr)), dim3(kBlock), 0, s, v, U0, K, a_dt, W); });
    P.run("k_seam_gather", s, [&] {
        hipLaunchKernelGGL(k_seam_gather, dim3(blocksFor((long long)L.ncol*v.nr)), dim3(kBlock), 0, s,
                           v, L.ncol, L.colDst, L.colStart, L.donorSrc, L.donorW, W); });
    P.run("k_radial_bc", s, [&] {
        hipLaunchKernelGGL(k_radial_bc, dim3(blocksFor(6LL*v.NJ*v.NK)), dim3(kBlock), 0, s, v, inner, W); });
    HC_CHECK(hipGetLastError());
}

// Powell source after the flux divergence (needs the panel states in L.PS).
template <typename S>
void powellSource(hc_level& L, const double* W, double* K, double* U, double b_dt, hipStream_t s)
{
    const LevelView& v = L.view;
    if (!v.powell) return;
    KernelProfile& P = L.profile;
    P.run("k_powell_bflux", s, [&] {
        hipLaunchKernelGGL((k_powell_bflux<0, S>), dim3(blocksFor(6LL*(v.nr+1)*v.na*v.na)), dim3(kBlock), 0, s, v, W, L.PS, L.BF[0]);
        hipLaunchKernelGGL((k_powell_bflux<1, S>), dim3(blocksFor(6LL*v.nr*(v.na+1)*v.na)), dim3(kBlock), 0, s, v, W, L.PS, L.BF[1]);
        hipLaunchKernelGGL((k_powell_bflux<2, S>), dim3(blocksFor(6LL*v.nr*v.na*(v.na+1))), dim3(kBlock), 0, s, v, W, L.PS, L.BF[2]); });
    P.run("k_powell_source", s, [&] {
        hipLaunchKernelGGL(k_powell_source, dim3(blocksFor(6LL*v.na*v.na*v.nr)), dim3(kBlock), 0, s,
                           v, W, L.BF[0], L.BF[1], L.BF[2], K, U, b_dt); });
}

template <typename S>
void stageRhsFused(hc_level& L, const double* W, double* K, double* U, double b_dt, hipStream_t s)
{
    const LevelView& v = L.view;
    KernelProfile& P = L.profile;
    P.run("k_panel_states", s, [&] {
        hipLaunchKernelGGL(k_panel_states, dim3(blocksFor(24LL*v.na*v.nr)), dim3(kBlock), 0, s, v, W, L.PS); });
    P.run("k_seam_flux", s, [&] {
        if (v.roe) hipLaunchKernelGGL((k_seam_flux<S, true>), dim3(blocksFor(24LL*v.na*v.nr)), dim3(kBlock), 0, s, v, L.PS, L.SF);
        else hipLaunchKernelGGL((k_seam_flux<S, false>), dim3(blocksFor(24LL*v.na*v.nr)), dim3(kBlock), 0, s, v, L.PS, L.SF); });
    const int tI = (v.nr + kTI - 1)/kTI, tJ = (v.na + kTJ - 1)/kTJ, tK = (v.na + kTK - 1)/kTK;
    double* const Ud = v.powell ? nullptr : U;   // with Powell, U is accumulated after the source
    P.run("k_rhs_fused", s, [&] {
        if (v.roe) hipLaunchKernelGGL((k_rhs_fused<S, true>), dim3(6u*tI*tJ*tK), dim3(kTileCells), 0, s,
                           v, W, L.SF, K, Ud, b_dt, tI, tJ, tK);
        else hipLaunchKernelGGL((k_rhs_fused<S, false>), dim3(6u*tI*tJ*tK), dim3(kTileCells), 0, s,
                           v, W, L.SF, K, Ud, b_dt, tI, tJ, tK); });
    powellSource<S>(L, W, K, U, b_dt, s);
    HC_CHECK(hipGetLastError());
}

void stageRhs(hc_level& L, const double* W, double* K, double* U, double b_dt,
              double* F0, double* F1, double* F2, hipStream_t s)
{
    L.ensureWorkspace();
    const LevelView& v = L.view;
    if (L.l4)
    {
        if (F0 || F1 || F2) throw std::runtime_error("face fluxes of the fourth-order level: use hc_boxop_apply");
        L.profile.run("k4_rhs", s, [&] { hc::level4Rhs(*L.l4, W, K, U, b_dt, s); });
        if (L.consSum)
            L.profile.run("k4_radcons", s, [&] {
                hc::level4RadialCons(*L.l4, b_dt/((double)v.na*v.na), L.consSum, s); });
        return;
    }
    if (!F0 && !F1 && !F2)
    {
        if (L.opt.precision == 1) stageRhsFused<float>(L, W, K, U, b_dt, s);
        else stageRhsFused<double>(L, W, K, U, b_dt, s);
        return;
    }
    double* F[3] = {F0 ? F0 : L.F[0], F1 ? F1 : L.F[1], F2 ? F2 : L.F[2]};
    KernelProfile& P = L.profile;
    P.run("k_panel_states", s, [&] {
        hipLaunchKernelGGL(k_panel_states, dim3(blocksFor(24LL*v.na*v.nr)), dim3(kBlock), 0, s, v, W, L.PS); });
    P.run("k_face_flux_0", s, [&] {
        if (v.roe) hipLaunchKernelGGL((k_face_flux<0, true>), dim3(blocksFor(6LL*(v.nr+1)*v.na*v.na)), dim3(kBlock), 0, s, v, W, L.PS, F[0]);
        else hipLaunchKernelGGL((k_face_flux<0, false>), dim3(blocksFor(6LL*(v.nr+1)*v.na*v.na)), dim3(kBlock), 0, s, v, W, L.PS, F[0]); });
    P.run("k_face_flux_1", s, [&] {
        if (v.roe) hipLaunchKernelGGL((k_face_flux<1, true>), dim3(blocksFor(6LL*v.nr*(v.na+1)*v.na)), dim3(kBlock), 0, s, v, W, L.PS, F[1]);
        else hipLaunchKernelGGL((k_face_flux<1, false>), dim3(blocksFor(6LL*v.nr*(v.na+1)*v.na)), dim3(kBlock), 0, s, v, W, L.PS, F[1]); });
    P.run("k_face_flux_2", s, [&] {
        if (v.roe) hipLaunchKernelGGL((k_face_flux<2, true>), dim3(blocksFor(6LL*v.nr*v.na*(v.na+1))), dim3(kBlock), 0, s, v, W, L.PS, F[2]);
        else hipLaunchKernelGGL((k_face_flux<2, false>), dim3(blocksFor(6LL*v.nr*v.na*(v.na+1))), dim3(kBlock), 0, s, v, W, L.PS, F[2]); });
    P.run("k_seam_reconcile", s, [&] {
        hipLaunchKernelGGL(k_seam_reconcile, dim3(blocksFor(24LL*v.na*v.nr)), dim3(kBlock), 0, s, v, F[1], F[2]); });
    P.run("k_divergence", s, [&] {
        hipLaunchKernelGGL(k_divergence, dim3(blocksFor(6LL*v.na*v.na*v.nr)), dim3(kBlock), 0, s,
                           v, F[0], F[1], F[2], K, v.powell ? nullptr : U, b_dt); });
    powellSource<double>(L, W, K, U, b_dt, s);
    HC_CHECK(hipGetLastError());
}

// ---- Seeded initial data ---------------------------------------------------
struct SplitMix
{
    unsigned long long s;
    unsigned long long next()
    {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30))*0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27))*0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11)*(1.0/9007199254740992.0); }
};

struct WindState
{
    double r0, r1, gamma;
    double center[3], width, pulse;
    bool hemispheric;
    int nmodes = 0;
    double kvec[16][3], evec[16][3], phase[16], amp = 0;
    bool cone = false;             // low-beta cone: p = 1e-4 |B|^2/2 inside it
    double coneAxis[3] = {0, 0, 1}, coneCos = 1.0;
    bool sectors = false;          // B_r sign by quadrant of the azimuth (4 sectors)
    bool blob = false;             // magnetized blob: rho x4, v_r = 2, toroidal B = 3 |B_r|
    double blobCenter[3] = {0, 0, 0}, blobRadius = 0;
    // Cartesian primitive state at position X.
    void at(const double X[3], double w[8]) const
    {
        base(X, w);
        if (nmodes != 0) modes(X, w);
        if (cone)
        {
            const double r = std::sqrt(X[0]*X[0]+X[1]*X[1]+X[2]*X[2]);
            const double c = (X[0]*coneAxis[0]+X[1]*coneAxis[1]+X[2]*coneAxis[2])/r;
            if (c > coneCos) w[4] = 1.0e-4*0.5*(w[5]*w[5]+w[6]*w[6]+w[7]*w[7]);
        }
        if (blob) blobState(X, w);
    }
    // Blob of radius R around C moving radially: density x4, v = 2 r_hat, and a
    // toroidal field about the blob's radial axis n = C/|C| of magnitude
    // 3 |B_r| sin(pi d/R) s/d (d distance to C, s distance to the axis), zero
    // on the axis and at the blob's surface.
    void blobState(const double X[3], double w[8]) const
    {
        double rel[3], d2 = 0, cn = 0;
        for (int a = 0; a < 3; ++a) { rel[a] = X[a]-blobCenter[a]; d2 += rel[a]*rel[a]; cn += blobCenter[a]*blobCenter[a]; }
        const double d = std::sqrt(d2);
        if (!(d < blobRadius)) return;
        const double r = std::sqrt(X[0]*X[0]+X[1]*X[1]+X[2]*X[2]);
        const double br = std::fabs((w[5]*X[0]+w[6]*X[1]+w[7]*X[2])/r);
        w[0] *= 4.0;
        for (int a = 0; a < 3; ++a) w[1+a] = 2.0*X[a]/r;
        cn = std::sqrt(cn);
        const double n[3] = {blobCenter[0]/cn, blobCenter[1]/cn, blobCenter[2]/cn};
        const double along = rel[0]*n[0]+rel[1]*n[1]+rel[2]*n[2];
        double sv[3], s2 = 0;
        for (int a = 0; a < 3; ++a) { sv[a] = rel[a]-along*n[a]; s2 += sv[a]*sv[a]; }
        const double sd = std::sqrt(s2);
        if (!(sd > 1e-12*blobRadius)) return;
        const double e[3] = {(n[1]*sv[2]-n[2]*sv[1])/sd, (n[2]*sv[0]-n[0]*sv[2])/sd, (n[0]*sv[1]-n[1]*sv[0])/sd};
        const double mag = 3.0*br*std::sin(hc::kPi*d/blobRadius)*(sd/d);
        for (int a = 0; a < 3; ++a) w[5+a] += mag*e[a];
    }
    void modes(const double X[3], double w[8]) const
    {
        const double r = std::sqrt(X[0]*X[0]+X[1]*X[1]+X[2]*X[2]);
        const double rh[3] = {X[0]/r, X[1]/r, X[2]/r};
        double dv[3] = {0, 0, 0};
        for (int m = 0; m < nmodes; ++m)
        {
            const double s = std::sin(kvec[m][0]*rh[0]+kvec[m][1]*rh[1]+kvec[m][2]*rh[2]+phase[m]);
            const double en = evec[m][0]*rh[0]+evec[m][1]*rh[1]+evec[m][2]*rh[2];
            for (int a = 0; a < 3; ++a) dv[a] += s*(evec[m][a]-en*rh[a]);
        }
        const double sq = std::sqrt(4*hc::kPi*w[0]);
        for (int a = 0; a < 3; ++a) { w[1+a] += amp*dv[a]; w[5+a] -= amp*dv[a]*sq; }
    }
    void base(const double X[3], double w[8]) const
    {
        const double r = std::sqrt(X[0]*X[0]+X[1]*X[1]+X[2]*X[2]);
        const double rho = (r0/r)*(r0/r);
        double sign = (!hemispheric || X[2] > 0) ? 1.0 : -1.0;
        if (sectors)
        {
            const double phi = std::atan2(X[1], X[0]);                     // (-pi, pi]
            const int q = (int)std::floor((phi + hc::kPi)/(0.5*hc::kPi)) & 3;
            sign = (q & 1) ? -1.0 : 1.0;
        }
        double d2 = 0;
        for (int a = 0; a < 3; ++a) d2 += (X[a]-center[a])*(X[a]-center[a]);
        const double p = 0.1*rho*(1.0 + pulse*std::exp(-d2/(width*width)));
        w[0] = rho; w[4] = p;
        for (int a = 0; a < 3; ++a) { w[1+a] = X[a]/r; w[5+a] = sign*rho*X[a]/r; }
    }
};

// Seeded test states of the fourth-order paths: config 0 radial wind with a
// Gaussian pressure pulse and hemispheric B_r, 1 monopole B_r, 2 monopole
// without the pulse, 3 = 0 plus Alfvenic modes (|k| <= 8, amplitude 0.05) and
// (applied by the callers) lognormal density noise, 4 = 3 plus a low-beta cone
// (p = 1e-4 |B|^2/2 within 15 degrees of a seeded axis), 5 the radial wind
// without pulse, modes or noise, with B_r signed by four azimuthal sectors and
// a magnetized blob (radius 0.05 (r1 - r0) at a seeded point of the middle
// radius: density x4, v_r = 2, toroidal B of 3 |B_r|).
WindState makeWind(const hc_shell& shell, const hc_options& options, int config, unsigned long long seed)
{
    SplitMix rng{seed};
    WindState ws;
    ws.r0 = shell.inner_radius; ws.r1 = shell.outer_radius; ws.gamma = options.gamma;
    const double cz = 2.0*rng.uniform()-1.0, phi = 2.0*hc::kPi*rng.uniform();
    const double rc = 0.5*(ws.r0+ws.r1), sz = std::sqrt(std::max(0.0, 1.0-cz*cz));
    ws.center[0] = rc*sz*std::cos(phi); ws.center[1] = rc*sz*std::sin(phi); ws.center[2] = rc*cz;
    ws.width = 0.1*(ws.r1-ws.r0);
    ws.pulse = (config == 2 || config == 5) ? 0.0 : 0.2;
    ws.hemispheric = config == 0 || config == 3 || config == 4;
    if (config == 5)
    {
        ws.sectors = true;
        ws.blob = true;
        ws.blobRadius = 0.05*(ws.r1-ws.r0);
        for (int a = 0; a < 3; ++a) ws.blobCenter[a] = ws.center[a];
    }
    if (config == 3 || config == 4)
    {
        ws.nmodes = 16; ws.amp = 0.05/std::sqrt(16.0);
        for (int m = 0; m < 16; ++m)
        {
            double kk[3], n2 = 0;
            for (int a = 0; a < 3; ++a) { kk[a] = 2.0*rng.uniform()-1.0; n2 += kk[a]*kk[a]; }
            const double kmag = 1.0 + 7.0*rng.uniform();
            for (int a = 0; a < 3; ++a) ws.kvec[m][a] = kk[a]/std::sqrt(n2)*kmag;
            double e[3], e2 = 0;
            for (int a = 0; a < 3; ++a) { e[a] = 2.0*rng.uniform()-1.0; e2 += e[a]*e[a]; }
            for (int a = 0; a < 3; ++a) ws.evec[m][a] = e[a]/std::sqrt(e2);
            ws.phase[m] = 2.0*hc::kPi*rng.uniform();
        }
    }
    if (config == 4)
    {
        const double az = 2.0*rng.uniform()-1.0, aphi = 2.0*hc::kPi*rng.uniform();
        const double as = std::sqrt(std::max(0.0, 1.0-az*az));
        ws.cone = true;
        ws.coneAxis[0] = as*std::cos(aphi); ws.coneAxis[1] = as*std::sin(aphi); ws.coneAxis[2] = az;
        ws.coneCos = std::cos(15.0*hc::kPi/180.0);
    }
    return ws;
}
} // namespace

extern "C" {

const char* hc_last_error(void) { return hc::lastError().c_str(); }

int hc_level_create(const hc_shell* shell, const hc_options* options, hc_level** out)
{
    if (!shell || !options || !out) return fail("hc_level_create: null argument");
    if (shell->n_rad < 2 || shell->n_ang < 4) return fail("hc_level_create: need n_rad >= 2 and n_ang >= 4");
    if (!(shell->outer_radius > shell->inner_radius && shell->inner_radius > 0))
        return fail("hc_level_create: need 0 < inner_radius < outer_radius");
    if (options->spatial_order != 2 && options->spatial_order != 4)
        return fail("hc_level_create: spatial_order must be 2 or 4");
    if (options->riemann_solver != 1 && options->riemann_solver != 2)
        return fail("hc_level_create: riemann_solver must be 1 (Rusanov) or 2 (Roe)");
    if (options->limiter < 0 || options->limiter > 3) return fail("hc_level_create: limiter must be 0..3");
    if (options->spatial_order == 4 && options->limiter > 2)
        return fail("hc_level_create: the fourth-order operator takes limiter 0, 1 or 2");
    if (options->spatial_order == 4 && (shell->n_ang < 16 || shell->n_rad < 4))
        return fail("hc_level_create: the fourth-order level needs n_ang >= 16 (five ghost layers must stay "
                    "inside the panel's horizon) and n_rad >= 4");
    if (options->precision != 0 && options->precision != 1) return fail("hc_level_create: precision must be 0 (fp64) or 1 (mixed)");
    if (const char* e = hc::checkOperatorOptions(*options)) return fail(std::string("hc_level_create: ") + e);
    try
    {
        std::unique_ptr<hc_level> L(new hc_level());
        const int g = options->spatial_order == 4 ? 5 : 2;
        L->dims = {shell->n_rad, shell->n_ang, g, shell->inner_radius, shell->outer_radius, shell->radial_stretch};
        L->opt = *options;
        if (options->spatial_order == 4)
        {
            LevelView& v = L->view;
            v.nr = shell->n_rad; v.na = shell->n_ang; v.g = g;
            v.NI = v.nr+2*g; v.NJ = v.na+2*g; v.NK = v.na+2*g;
            v.cs = (long long)v.NI*v.NJ*v.NK; v.ps = 8*v.cs;
            v.vcs = (long long)v.nr*v.na*v.na; v.vps = 8*v.vcs;
            const int tile[3] = {32, 32, 32};
            L->l4 = hc::level4Create(*shell, *options, tile);
        }
        else buildLevel(*L);
        *out = L.release();
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_create: ") + e.what()); }
}

void hc_level_destroy(hc_level* level) { delete level; }

int hc_level_ghost(const hc_level* level) { return level ? level->dims.g : -1; }
size_t hc_level_valid_count(const hc_level* level) { return level ? level->validCount() : 0; }
size_t hc_level_stage_count(const hc_level* level) { return level ? level->stageCount() : 0; }
size_t hc_level_inner_count(const hc_level* level) { return level ? level->innerCount() : 0; }
size_t hc_level_face_count(const hc_level* level, int d)
{ return level && d >= 0 && d < 3 ? level->faceCount(d) : 0; }

int hc_prepare_stage(hc_level* level, const double* U0, const double* K, double a_dt,
                     const double* inner, double* W, hc_stream stream)
{
    if (!level || !U0 || !inner || !W) return fail("hc_prepare_stage: null argument");
    try { prepareStage(*level, U0, K, a_dt, inner, W, static_cast<hipStream_t>(stream)); return 0; }
    catch (const std::exception& e) { return fail(std::string("hc_prepare_stage: ") + e.what()); }
}

int hc_stage_rhs(hc_level* level, const double* W, double* K, double* U, double b_dt,
                 double* F0, double* F1, double* F2, hc_stream stream)
{
    if (!level || !W || !K) return fail("hc_stage_rhs: null argument");
    try { stageRhs(*level, W, K, U, b_dt, F0, F1, F2, static_cast<hipStream_t>(stream)); return 0; }
    catch (const std::exception& e) { return fail(std::string("hc_stage_rhs: ") + e.what()); }
}

int hc_rk4_advance(hc_level* level, double* U, const double* inner, double dt, int nsteps, hc_stream stream)
{
    if (!level || !U || !inner) return fail("hc_rk4_advance: null argument");
    if (nsteps < 0) return fail("hc_rk4_advance: nsteps must be >= 0");
    try
    {
        hc_level& L = *level;
        L.ensureWorkspace();
        hipStream_t s = static_cast<hipStream_t>(stream);
        const double a[4] = {0.0, 0.5, 0.5, 1.0};
        const double b[4] = {1.0/6, 1.0/3, 1.0/3, 1.0/6};
        const long long nv = (long long)L.validCount();
        auto step = [&](hipStream_t q) {
            hipLaunchKernelGGL(k_copy, dim3((unsigned)((nv+kBlock-1)/kBlock)), dim3(kBlock), 0, q, U, L.U0, nv);
            for (int st = 0; st < 4; ++st)
            {
                prepareStage(L, L.U0, st ? L.K : nullptr, dt*a[st], inner, L.W, q);
                stageRhs(L, L.W, L.K, U, dt*b[st], nullptr, nullptr, nullptr, q);
            }
        };
        if (!L.useGraph || L.profile.enabled || nsteps == 0)
        {
            for (int n = 0; n < nsteps; ++n) step(s);
            HC_CHECK(hipGetLastError());
            return 0;
        }
        if (!L.graphExec || L.gU != U || L.gInner != inner || L.gDt != dt || L.gCons != L.consSum)
        {
            L.dropGraph();
            if (!L.capStream) HC_CHECK(hipStreamCreateWithFlags(&L.capStream, hipStreamNonBlocking));
            // order the capture stream after the caller's pending work on s
            hipEvent_t ready;
            HC_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
            HC_CHECK(hipEventRecord(ready, s));
            HC_CHECK(hipStreamWaitEvent(L.capStream, ready, 0));
            HC_CHECK(hipStreamSynchronize(L.capStream));
            (void)hipEventDestroy(ready);
            HC_CHECK(hipStreamBeginCapture(L.capStream, hipStreamCaptureModeThreadLocal));
            try { step(L.capStream); }
            catch (...)
            {
                // end the capture so the stream is usable again, drop the
                // partial graph, then report the error
                hipGraph_t partial = nullptr;
                (void)hipStreamEndCapture(L.capStream, &partial);
                if (partial) (void)hipGraphDestroy(partial);
                (void)hipGetLastError();
                throw;
            }
            HC_CHECK(hipStreamEndCapture(L.capStream, &L.graph));
            HC_CHECK(hipGraphInstantiate(&L.graphExec, L.graph, nullptr, nullptr, 0));
            L.gU = U; L.gInner = inner; L.gDt = dt; L.gCons = L.consSum;
        }
        for (int n = 0; n < nsteps; ++n) HC_CHECK(hipGraphLaunch(L.graphExec, s));
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_rk4_advance: ") + e.what()); }
}

int hc_level_set_cons_sum(hc_level* level, double* cons_sum)
{
    if (!level) return fail("hc_level_set_cons_sum: null level");
    if (cons_sum && !level->l4)
        return fail("hc_level_set_cons_sum: the conservation sums use the fourth-order level's face fluxes "
                    "(spatial_order 4)");
    level->consSum = cons_sum;
    return 0;
}

int hc_level_use_graph(hc_level* level, int enable)
{
    if (!level) return fail("hc_level_use_graph: null level");
    level->useGraph = enable != 0;
    if (!level->useGraph) level->dropGraph();
    return 0;
}

int hc_level_set_tile_shape(hc_level* level, int t0, int t1, int t2)
{
    if (!level) return fail("hc_level_set_tile_shape: null level");
    if (!level->l4) return fail("hc_level_set_tile_shape: only a fourth-order level has a box table");
    if (t0 < 8 || t1 < 8 || t2 < 8) return fail("hc_level_set_tile_shape: every tile edge must be >= 8 cells");
    try
    {
        int cur[3];
        hc::level4TileShape(*level->l4, cur);
        if (cur[0] == t0 && cur[1] == t1 && cur[2] == t2) return 0;
        level->dropGraph();
        HC_CHECK(hipDeviceSynchronize());
        const int tile[3] = {t0, t1, t2};
        hc::level4Retile(*level->l4, tile);
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_set_tile_shape: ") + e.what()); }
}

int hc_level_set_tile(hc_level* level, int tile)
{
    if (!level) return fail("hc_level_set_tile: null level");
    if (!level->l4) return fail("hc_level_set_tile: only a fourth-order level has a box table");
    if (tile < 8) return fail("hc_level_set_tile: tile must be >= 8 cells");
    return hc_level_set_tile_shape(level, tile, tile, tile);
}

int hc_level_tile(const hc_level* level)
{
    if (!level || !level->l4) return 0;
    int t[3];
    hc::level4TileShape(*level->l4, t);
    return t[0] == t[1] && t[1] == t[2] ? t[0] : -1;
}

int hc_level_tile_shape(const hc_level* level, int* tile3)
{
    if (!level || !tile3) return fail("hc_level_tile_shape: null argument");
    if (!level->l4) return fail("hc_level_tile_shape: only a fourth-order level has a box table");
    hc::level4TileShape(*level->l4, tile3);
    return 0;
}

int hc_level_autotune(hc_level* level, const double* U, const double* inner, int* tile3, double* ms_per_stage,
                      hc_stream stream)
{
    if (!level || !U || !inner || !tile3) return fail("hc_level_autotune: null argument");
    if (!level->l4) return fail("hc_level_autotune: only a fourth-order level has a box table");
    try
    {
        hc_level& L = *level;
        L.ensureWorkspace();
        hipStream_t s = static_cast<hipStream_t>(stream);
        const int nr = L.dims.nr, na = L.dims.na;
        // cubes and radial pencils (long along the unit-stride radius, which
        // halves the ghost fraction per radial line), clipped to the panel
        const int shapes[][3] = {{32, 32, 32}, {64, 64, 64}, {128, 128, 128}, {64, 32, 32}, {128, 32, 32},
                                 {nr, 32, 32}, {128, 64, 64}, {nr, 64, 64}, {nr, na, na}};
        std::vector<std::array<int, 3>> cand;
        for (const auto& sh : shapes)
        {
            const std::array<int, 3> c = {std::min(sh[0], nr), std::min(sh[1], na), std::min(sh[2], na)};
            if (std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
        }
        hipEvent_t e0, e1;
        HC_CHECK(hipEventCreate(&e0));
        HC_CHECK(hipEventCreate(&e1));
        std::array<int, 3> best = cand.front();
        double bestMs = 1e300;
        for (const auto& t : cand)
        {
            int cur[3];
            hc::level4TileShape(*L.l4, cur);
            if (cur[0] != t[0] || cur[1] != t[1] || cur[2] != t[2])
            {
                L.dropGraph();
                HC_CHECK(hipDeviceSynchronize());
                hc::level4Retile(*L.l4, t.data());
            }
            // one stage (preparation + RHS into the level's K; U is not
            // modified), warm, then the RHS timed twice
            prepareStage(L, U, nullptr, 0.0, inner, L.W, s);
            stageRhs(L, L.W, L.K, nullptr, 0.0, nullptr, nullptr, nullptr, s);
            HC_CHECK(hipEventRecord(e0, s));
            for (int r = 0; r < 2; ++r) stageRhs(L, L.W, L.K, nullptr, 0.0, nullptr, nullptr, nullptr, s);
            HC_CHECK(hipEventRecord(e1, s));
            HC_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HC_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms/2 < bestMs) { bestMs = ms/2; best = t; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        int cur[3];
        hc::level4TileShape(*L.l4, cur);
        if (cur[0] != best[0] || cur[1] != best[1] || cur[2] != best[2])
        {
            L.dropGraph();
            HC_CHECK(hipDeviceSynchronize());
            hc::level4Retile(*L.l4, best.data());
        }
        for (int d = 0; d < 3; ++d) tile3[d] = best[d];
        if (ms_per_stage) *ms_per_stage = bestMs;
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_autotune: ") + e.what()); }
}

int hc_level_dt_cfl(hc_level* level, const double* U, double cone_boundary_density, double* dt, hc_stream stream)
{
    if (!level || !U || !dt) return fail("hc_level_dt_cfl: null argument");
    if (!(cone_boundary_density >= 0.0)) return fail("hc_level_dt_cfl: cone_boundary_density must be >= 0");
    try
    {
        hc_level& L = *level;
        if (!L.cfl) L.cfl = hc::cflCreate(L.dims, L.opt.gamma);
        hipStream_t s = static_cast<hipStream_t>(stream);
        // the fourth-order dtCFL takes no cone bound (src/BoxOp_EulerCubedSphereImplem.H:886-887)
        const double cone = L.opt.spatial_order == 2 ? cone_boundary_density : 0.0;
        const double* r = hc::cflLaunch(*L.cfl, U, cone, s);
        HC_CHECK(hipMemcpyAsync(dt, r, sizeof(double), hipMemcpyDeviceToHost, s));
        HC_CHECK(hipStreamSynchronize(s));
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_dt_cfl: ") + e.what()); }
}

int hc_level_bc_points(hc_level* level, const double* data0, int ntheta, int nphi, const double* theta,
                       const double* phi, double* bc, hc_stream stream)
{
    if (!level || !data0 || !theta || !phi || !bc || ntheta < 1 || nphi < 1)
        return fail("hc_level_bc_points: bad argument");
    try
    {
        hc::bcPoints(level->dims, data0, ntheta, nphi, theta, phi, bc, static_cast<hipStream_t>(stream));
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_bc_points: ") + e.what()); }
}

int hc_level_bc_averages(hc_level* level, const double* data0, int ntheta, int nphi, const double* theta,
                         const double* phi, double* inner, hc_stream stream)
{
    if (!level || !data0 || !theta || !phi || !inner || ntheta < 1 || nphi < 1)
        return fail("hc_level_bc_averages: bad argument");
    try
    {
        hc::bcAverages(level->dims, level->opt.spatial_order, level->opt.gamma, data0, ntheta, nphi, theta, phi, inner,
                       static_cast<hipStream_t>(stream));
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_bc_averages: ") + e.what()); }
}

int hc_level_inner_ghosts_from_bc(hc_level* level, const double* bc, double* inner, hc_stream stream)
{
    if (!level || !bc || !inner) return fail("hc_level_inner_ghosts_from_bc: null argument");
    try
    {
        hc_level& L = *level;
        if (!L.bcTable) L.bcTable = hc::bcCreate(L.dims, L.opt.spatial_order, L.opt.gamma);
        hc::bcLaunch(*L.bcTable, bc, inner, static_cast<hipStream_t>(stream));
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_inner_ghosts_from_bc: ") + e.what()); }
}

int hc_level_inner_ghosts_outflow(hc_level* level, double strength, double time, double* inner, hc_stream stream)
{
    if (!level || !inner) return fail("hc_level_inner_ghosts_outflow: null argument");
    try
    {
        hc_level& L = *level;
        if (!L.bcTable) L.bcTable = hc::bcCreate(L.dims, L.opt.spatial_order, L.opt.gamma);
        hc::outflowLaunch(*L.bcTable, strength, time, inner, static_cast<hipStream_t>(stream));
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_inner_ghosts_outflow: ") + e.what()); }
}

int hc_level_valid_state(hc_level* level, const double* U, int* valid, long long* first_invalid, hc_stream stream)
{
    if (!level || !U || !valid) return fail("hc_level_valid_state: null argument");
    try
    {
        hc_level& L = *level;
        if (!L.cfl) L.cfl = hc::cflCreate(L.dims, L.opt.gamma);
        hipStream_t s = static_cast<hipStream_t>(stream);
        const long long* r = hc::stateCheckLaunch(*L.cfl, U, s);
        long long first = -1;
        HC_CHECK(hipMemcpyAsync(&first, r, sizeof(long long), hipMemcpyDeviceToHost, s));
        HC_CHECK(hipStreamSynchronize(s));
        *valid = first < 0 ? 1 : 0;
        if (first_invalid) *first_invalid = first;
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_valid_state: ") + e.what()); }
}

int hc_init_state(const hc_shell* shell, const hc_options* options, int config,
                  unsigned long long seed, double* JU, double* inner)
{
    if (!shell || !options || !JU || !inner) return fail("hc_init_state: null argument");
    if (config < 0 || config > 2) return fail("hc_init_state: config must be 0, 1 or 2");
    const int nr = shell->n_rad, na = shell->n_ang, g = 2;
    const hc::ShellDims dims{nr, na, g, shell->inner_radius, shell->outer_radius, shell->radial_stretch};
    SplitMix rng{seed};
    WindState ws;
    ws.r0 = shell->inner_radius; ws.r1 = shell->outer_radius; ws.gamma = options->gamma;
    const double cz = 2.0*rng.uniform()-1.0, phi = 2.0*hc::kPi*rng.uniform();
    const double rc = 0.5*(ws.r0+ws.r1), sz = std::sqrt(std::max(0.0, 1.0-cz*cz));
    ws.center[0] = rc*sz*std::cos(phi); ws.center[1] = rc*sz*std::sin(phi); ws.center[2] = rc*cz;
    ws.width = 0.1*(ws.r1-ws.r0);
    ws.pulse = config == 2 ? 0.0 : 0.2;
    ws.hemispheric = config == 0;
    const double h = 1.0/na, gamma = options->gamma;
    const int NJ = na+2*g, NK = na+2*g;
    for (int p = 0; p < 6; ++p) for (int k = -g; k < na+g; ++k) for (int j = -g; j < na+g; ++j)
    {
        const hc::Vec3 dir = hc::panelPoint(p, hc::xiToAngle((j+0.5)*h), hc::xiToAngle((k+0.5)*h));
        const bool valid = j >= 0 && j < na && k >= 0 && k < na;
        const double detA = valid ? hc::cellDetA(na, j, k) : 0.0;
        for (int i = -g; i < nr; ++i)
        {
            if (i >= 0 && !valid) continue;
            const hc::RadialCell rc2 = hc::radialCell(dims, i);
            const double X[3] = {rc2.rmid*dir.x[0], rc2.rmid*dir.x[1], rc2.rmid*dir.x[2]};
            double w[8];
            ws.at(X, w);
            if (i < 0)
            {
                for (int c = 0; c < 8; ++c)
                    inner[((((size_t)p*8+c)*NK+(k+g))*NJ+(j+g))*g+(i+g)] = w[c];
                continue;
            }
            double u[8];
            double v2 = 0, b2 = 0;
            u[0] = w[0];
            for (int a = 0; a < 3; ++a)
            {
                u[1+a] = w[0]*w[1+a]; u[5+a] = w[5+a];
                v2 += w[1+a]*w[1+a]; b2 += w[5+a]*w[5+a];
            }
            u[4] = w[4]/(gamma-1.0) + 0.5*w[0]*v2 + b2/(8*hc::kPi);
            const double J = rc2.dvol*detA;
            for (int c = 0; c < 8; ++c)
                JU[((((size_t)p*8+c)*na+k)*na+j)*nr+i] = J*u[c];
        }
    }
    return 0;
}

int hc_level_profile(hc_level* level, int enable)
{
    if (!level) return fail("hc_level_profile: null level");
    level->profile.clear();
    level->profile.enabled = enable != 0;
    return 0;
}

int hc_level_profile_read(hc_level* level, int slot, char* name, int name_len, double* total_ms, long long* launches)
{
    if (!level) return fail("hc_level_profile_read: null level");
    KernelProfile& P = level->profile;
    if (slot < 0 || slot >= (int)P.names.size()) return 1;
    try
    {
        double ms = 0;
        for (auto& e : P.events[slot])
        {
            HC_CHECK(hipEventSynchronize(e.second));
            float t = 0;
            HC_CHECK(hipEventElapsedTime(&t, e.first, e.second));
            ms += t;
        }
        if (name && name_len > 0) { std::strncpy(name, P.names[slot].c_str(), name_len-1); name[name_len-1] = 0; }
        if (total_ms) *total_ms = ms;
        if (launches) *launches = (long long)P.events[slot].size();
        return 0;
    }
    catch (const std::exception& e) { return fail(std::string("hc_level_profile_read: ") + e.what()); }
}

int hc_init_native_boxes(const hc_shell* shell, const hc_options* options, int config, unsigned long long seed,
                         const hc_box* boxes, int nbox, int quad, double* USph)
{
    if (!shell || !options || !boxes || !USph || nbox <= 0) return fail("hc_init_native_boxes: bad argument");
    if (config < 0 || config > 5) return fail("hc_init_native_boxes: config must be 0..5");
    if (quad < 1 || quad > 8) return fail("hc_init_native_boxes: quad must be 1..8");
    const int nr = shell->n_rad, na = shell->n_ang, g = 5;
    const hc::ShellDims dims{nr, na, g, shell->inner_radius, shell->outer_radius, shell->radial_stretch};
    const WindState ws = makeWind(*shell, *options, config, seed);
    std::vector<double> gx, gw;
    hc::gaussLegendre01(quad, gx, gw);
    const double h = 1.0/na, gamma = options->gamma;
    size_t off = 0;
    for (int n = 0; n < nbox; ++n)
    {
        const hc_box& b = boxes[n];
        int lo[3], e[3];
        for (int d = 0; d < 3; ++d) { lo[d] = b.lo[d]-g; e[d] = b.hi[d]-b.lo[d]+1+2*g; }
        const size_t ec = (size_t)e[0]*e[1]*e[2];
        for (int kk = 0; kk < e[2]; ++kk) for (int jj = 0; jj < e[1]; ++jj) for (int ii = 0; ii < e[0]; ++ii)
        {
            const int i = lo[0]+ii, j = lo[1]+jj, k = lo[2]+kk;
            double avg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            double noise = 1.0;
            if (config == 3 || config == 4)
            {
                // lognormal density factor per cell, sigma 0.05, mean-preserving
                SplitMix cr{seed ^ (0x9E3779B97F4A7C15ull*((((unsigned long long)b.block*4096+(k+8))*4096+(j+8))*65536+(i+8)))};
                const double u1 = std::max(cr.uniform(), 1e-300), u2 = cr.uniform();
                const double gn = std::sqrt(-2.0*std::log(u1))*std::cos(2*hc::kPi*u2);
                noise = std::exp(0.05*gn-0.00125);
            }
            for (int qk = 0; qk < quad; ++qk) for (int qj = 0; qj < quad; ++qj)
            {
                const double al = hc::xiToAngle((j+gx[qj])*h), be = hc::xiToAngle((k+gx[qk])*h);
                double A[3][3], Ai[3][3];
                hc::pointBasis(b.block, al, be, A);
                hc::invert3(A, Ai);
                const hc::Vec3 dir = hc::panelPoint(b.block, al, be);
                for (int qi = 0; qi < quad; ++qi)
                {
                    const double eta = (i+gx[qi])/double(nr);
                    const double C = dims.stretch;
                    const double span = C == 0.0 ? (dims.r1-dims.r0) : (dims.r1-dims.r0)/std::expm1(C);
                    const double r = dims.r0 + span*(C == 0.0 ? eta : std::expm1(C*eta));
                    const double X[3] = {r*dir.x[0], r*dir.x[1], r*dir.x[2]};
                    double w[8];
                    ws.at(X, w);
                    w[0] *= noise;
                    double u[8], v2 = 0, b2 = 0;
                    u[0] = w[0];
                    for (int a = 0; a < 3; ++a) { v2 += w[1+a]*w[1+a]; b2 += w[5+a]*w[5+a]; }
                    u[4] = w[4]/(gamma-1.0) + 0.5*w[0]*v2 + b2/(8*hc::kPi);
                    for (int a = 0; a < 3; ++a)
                    {
                        double m = 0, bb = 0;
                        for (int c = 0; c < 3; ++c) { m += Ai[a][c]*w[0]*w[1+c]; bb += Ai[a][c]*w[5+c]; }
                        u[1+a] = m; u[5+a] = bb;
                    }
                    const double wt = gw[qi]*gw[qj]*gw[qk];
                    for (int c = 0; c < 8; ++c) avg[c] += wt*u[c];
                }
            }
            for (int c = 0; c < 8; ++c) USph[off + c*ec + ((size_t)kk*e[1]+jj)*e[0]+ii] = avg[c];
        }
        off += 8*ec;
    }
    return 0;
}

int hc_init_level4(const hc_shell* shell, const hc_options* options, int config, unsigned long long seed,
                   int quad, double* JU, double* inner)
{
    if (!shell || !options || !JU || !inner) return fail("hc_init_level4: null argument");
    if (config < 0 || config > 5) return fail("hc_init_level4: config must be 0..5");
    if (quad < 1 || quad > 8) return fail("hc_init_level4: quad must be 1..8");
    const int nr = shell->n_rad, na = shell->n_ang, g = 5;
    // inner radial ghosts: native averages of the analytic state
    std::vector<hc_box> boxes;
    for (int p = 0; p < 6; ++p) { hc_box b{}; b.block = p; b.lo[0] = -g; b.hi[0] = -1;
        b.lo[1] = 0; b.lo[2] = 0; b.hi[1] = na-1; b.hi[2] = na-1; boxes.push_back(b); }
    // hc_init_native_boxes writes each box grown by 5: take the ghosted rows
    // i in [-5, -1] of the angular range [-5, na+5) from a thin box
    const int e0 = g + 2*g, e1 = na + 2*g;
    std::vector<double> tmp((size_t)6*8*e0*e1*e1);
    const int rc = hc_init_native_boxes(shell, options, config, seed, boxes.data(), 6, quad, tmp.data());
    if (rc != 0) return rc;
    for (int p = 0; p < 6; ++p) for (int c = 0; c < 8; ++c) for (int kk = 0; kk < e1; ++kk) for (int jj = 0; jj < e1; ++jj)
        for (int ig = 0; ig < g; ++ig)
            inner[((((size_t)p*8+c)*e1+kk)*e1+jj)*g+ig] =
                tmp[(size_t)p*8*e0*e1*e1 + (((size_t)c*e1+kk)*e1+jj)*e0 + (ig+g)];
    // <J U> of the six whole panels (the valid level layout is the panel
    // boxes concatenated)
    std::vector<hc_box> panels;
    for (int p = 0; p < 6; ++p) { hc_box b{}; b.block = p; b.hi[0] = nr-1; b.hi[1] = na-1; b.hi[2] = na-1;
        panels.push_back(b); }
    return hc_init_ju_boxes(shell, options, config, seed, panels.data(), 6, quad, JU);
}

int hc_init_ju_boxes(const hc_shell* shell, const hc_options* options, int config, unsigned long long seed,
                     const hc_box* boxes, int nbox, int quad, double* JU)
{
    if (!shell || !options || !boxes || !JU || nbox <= 0) return fail("hc_init_ju_boxes: bad argument");
    if (config < 0 || config > 5) return fail("hc_init_ju_boxes: config must be 0..5");
    if (quad < 1 || quad > 8) return fail("hc_init_ju_boxes: quad must be 1..8");
    const int nr = shell->n_rad, na = shell->n_ang, g = 5;
    for (int n = 0; n < nbox; ++n)
    {
        const hc_box& b = boxes[n];
        if (b.block < 0 || b.block > 5) return fail("hc_init_ju_boxes: block out of range");
        for (int d = 0; d < 3; ++d)
            if (b.lo[d] < 0 || b.hi[d] < b.lo[d] || b.hi[d] >= (d == 0 ? nr : na))
                return fail("hc_init_ju_boxes: box outside the level");
    }
    // <J U>: J-weighted averages with Cartesian vector components
    const hc::ShellDims dims{nr, na, g, shell->inner_radius, shell->outer_radius, shell->radial_stretch};
    const WindState ws = makeWind(*shell, *options, config, seed);
    std::vector<double> gx, gw;
    hc::gaussLegendre01(quad, gx, gw);
    const double h = 1.0/na, gamma = options->gamma;
    const double C = dims.stretch;
    const double span = C == 0.0 ? (dims.r1-dims.r0) : (dims.r1-dims.r0)/std::expm1(C);
    size_t off = 0;
    for (int n = 0; n < nbox; ++n)
    {
        const hc_box& bx = boxes[n];
        const int p = bx.block;
        const int n0 = bx.hi[0]-bx.lo[0]+1, n1 = bx.hi[1]-bx.lo[1]+1, n2 = bx.hi[2]-bx.lo[2]+1;
        for (int k = bx.lo[2]; k <= bx.hi[2]; ++k) for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
        {
            // per-column angular samples
            std::vector<double> dirs(3*quad*quad), dets(quad*quad);
            for (int qk = 0; qk < quad; ++qk) for (int qj = 0; qj < quad; ++qj)
            {
                const double al = hc::xiToAngle((j+gx[qj])*h), be = hc::xiToAngle((k+gx[qk])*h);
                double A[3][3];
                hc::pointBasis(p, al, be, A);
                dets[qk*quad+qj] = A[0][0]*(A[1][1]*A[2][2]-A[1][2]*A[2][1]) - A[0][1]*(A[1][0]*A[2][2]-A[1][2]*A[2][0])
                                 + A[0][2]*(A[1][0]*A[2][1]-A[1][1]*A[2][0]);
                const hc::Vec3 d = hc::panelPoint(p, al, be);
                for (int a = 0; a < 3; ++a) dirs[3*(qk*quad+qj)+a] = d.x[a];
            }
            for (int i = bx.lo[0]; i <= bx.hi[0]; ++i)
            {
                double noise = 1.0;
                if (config == 3 || config == 4)
                {
                    SplitMix cr{seed ^ (0x9E3779B97F4A7C15ull*((((unsigned long long)p*4096+(k+8))*4096+(j+8))*65536+(i+8)))};
                    const double u1 = std::max(cr.uniform(), 1e-300), u2 = cr.uniform();
                    const double gn = std::sqrt(-2.0*std::log(u1))*std::cos(2*hc::kPi*u2);
                    noise = std::exp(0.05*gn-0.00125);
                }
                double avg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                for (int qi = 0; qi < quad; ++qi)
                {
                    const double eta = (i+gx[qi])/double(nr);
                    const double r = dims.r0 + span*(C == 0.0 ? eta : std::expm1(C*eta));
                    const double drdxi = C == 0.0 ? span : span*C*std::exp(C*eta);
                    for (int qa = 0; qa < quad*quad; ++qa)
                    {
                        const double* d = &dirs[3*qa];
                        const double X[3] = {r*d[0], r*d[1], r*d[2]};
                        double w[8];
                        ws.at(X, w);
                        w[0] *= noise;
                        double v2 = 0, b2 = 0;
                        for (int a = 0; a < 3; ++a) { v2 += w[1+a]*w[1+a]; b2 += w[5+a]*w[5+a]; }
                        const double J = r*r*drdxi*dets[qa];
                        const double wt = gw[qi]*gw[qa % quad]*gw[qa / quad]*J;
                        avg[0] += wt*w[0];
                        for (int a = 0; a < 3; ++a) { avg[1+a] += wt*w[0]*w[1+a]; avg[5+a] += wt*w[5+a]; }
                        avg[4] += wt*(w[4]/(gamma-1.0) + 0.5*w[0]*v2 + b2/(8*hc::kPi));
                    }
                }
                for (int c = 0; c < 8; ++c)
                    JU[off + (((size_t)c*n2+(k-bx.lo[2]))*n1+(j-bx.lo[1]))*n0+(i-bx.lo[0])] = avg[c];
            }
        }
        off += (size_t)8*n0*n1*n2;
    }
    return 0;
}


double hc_geom_cell_detA(int n_ang, int j, int k) { return hc::cellDetA(n_ang, j, k); }

void hc_geom_face_normal(int n_ang, int p, int d, int j, int k, double* normal3)
{
    const hc::Vec3 n = hc::faceNormal(n_ang, p, d, j, k);
    for (int a = 0; a < 3; ++a) normal3[a] = n.x[a];
}

int hc_geom_ghost_donors(int n_ang, int p, int j, int k, int* q, int* jj, int* kk, double* w)
{
    const auto d = hc::ghostColumnDonors(n_ang, p, j, k);
    for (size_t x = 0; x < d.size(); ++x) { q[x] = d[x].q; jj[x] = d[x].j; kk[x] = d[x].k; w[x] = d[x].w; }
    return (int)d.size();
}
} // extern "C"
