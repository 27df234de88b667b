// Fixed AMR hierarchy of fourth-order levels advanced by the synchronous
// classical RK4 of the reference (advanceRungeKuttaSynchronous,
// src/MHD_AMRRungeKutta.H:5-41). Included by hc_o4.hip after the level and box
// table code, whose kernels it launches.
//
// Layout. Level 0 covers the whole shell (the valid level layout
// [6][8][na][na][nr]); level l >= 1 is a set of patches of one box size,
// aligned to it, each [8][n2][n1][n0]. The combined state (and RHS) is the
// level-0 array followed by every patch in table order; the combined stage
// input is the level-0 stage layout [6][8][na+10][na+10][nr+10] followed by
// every patch grown by 5. Per refined level the stage work arrays are patch
// major ([patch][8][cells], the same cell count for every patch of a level).
//
// One stage (prepareFourthOrder, src/MHD_AMRFourthOrder.H:5-187; evaluate,
// src/MHD_AMR.H:1417-1499; the flux registers and restriction of
// src/MHD_AMRRungeKutta.H:30-36):
//  1. stage state S = U0 + a dt K_prev, restricted fine -> coarse (the mean of
//     the children, restrictTo, src/MHD_AMR.H:597-607);
//  2. level 0: <JU> -> <U> in one pass with the full Jacobian J = dVolr det A
//     (k4a_deconv0), then the seam interpolation, the native conversion and
//     the radial boundary ghosts of the single-level path;
//  3. level l >= 1, in order: the parent's Cartesian averages on each patch's
//     coarse source box (k4a_gather), the separable quartic interpolation onto
//     the patch grown by (6,8,8) (k4a_interp<d>), <JU> on the patch grown by 2
//     (same-level <JU> where a patch of the level holds the cell, else
//     J <U> + the average-product correction, k4a_mapped), <U> on the patch
//     (k4a_deconv), the same-level exchange of the ghosts (k4a_exchange) and
//     the native conversion on the patch grown by 5 (k4a_native);
//  4. the box operator over level-0 tiles and every patch in one box table;
//  5. flux registers: every coarse cell outside a finer level gets
//     +- n_d (mean fine face flux - coarse face flux) per coarse-fine face
//     (k4a_reflux); then the fine RHS is restricted onto the coarse RHS, top
//     down, and U += b dt K.
// After the four stages the state is restricted (averageDown).
namespace
{
constexpr int kAmrDg[3] = {6, 8, 8};   // donor ghost of a refined patch
constexpr int kAmrCg = 6;              // its Cartesian ghost: the width a child interpolates from
constexpr int kAmrMaxRatio = 4;

struct AmrLevelView
{
    int npatch, nr, na;
    int r[3];                       // refinement ratio to the parent
    int b[3];                       // patch size
    int G[3];                       // patch grid (nr / b0, na / b1, na / b2)
    double w[3][kAmrMaxRatio][5];   // quartic child weights per direction
    int sn[3];                      // coarse source box size
    int dn[3];                      // donor box size (patch + 2 dg)
    long long vCells, dCells, mCells, sCells, t1Cells, t2Cells, uCells;
    const int* grid;                // [6][G2][G1][G0] -> patch index or -1
    const int* pp;                  // panel of each patch
    const int* plo;                 // lo of each patch [3 n]
    const long long* vOff;          // valid offset of each patch in the combined state
    const long long* uOff;          // stage input offset of each patch
    const long long* tBase;         // restriction target of each patch: element of coarse cell (0,0,0), comp 0
    const long long* tRow;          // its strides (j, k, component) in the parent storage, 3 per patch
    const long long* srcMap;        // per patch and coarse source cell: parent element (comp 0) or -1
    long long pStride;              // parent component stride
    double *src, *t1, *t2, *donor, *mapped;
    Prep4View pv;                   // this level's metric tables
};

__device__ __forceinline__ int amrOwner(const AmrLevelView& v, int p, int i, int j, int k)
{
    if (i < 0 || j < 0 || k < 0 || i >= v.nr || j >= v.na || k >= v.na) return -1;
    const int a0 = i / v.b[0], a1 = j / v.b[1], a2 = k / v.b[2];
    return v.grid[(((long long)p*v.G[2] + a2)*v.G[1] + a1)*v.G[0] + a0];
}
__device__ __forceinline__ long long amrValid(const AmrLevelView& v, int n, int c, int i, int j, int k)
{
    const int* lo = v.plo + 3*n;
    return v.vOff[n] + c*v.vCells + ((long long)(k-lo[2])*v.b[1] + (j-lo[1]))*v.b[0] + (i-lo[0]);
}
// donor / mapped arrays: [patch][8][cells] on the patch grown by (dg) / 2
__device__ __forceinline__ long long amrDonor(const AmrLevelView& v, int n, int c, int i, int j, int k)
{
    const int* lo = v.plo + 3*n;
    return ((long long)n*8 + c)*v.dCells
         + ((long long)(k-lo[2]+kAmrDg[2])*v.dn[1] + (j-lo[1]+kAmrDg[1]))*v.dn[0] + (i-lo[0]+kAmrDg[0]);
}
__device__ __forceinline__ double amrJ(const AmrLevelView& v, int i, int j, int k)
{
    return v.pv.dvol[i+6]*det4(v.pv, j, k);
}
__device__ __forceinline__ int floorDivI(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }

// S = U0 (+ a_dt K) over n elements.
__global__ void __launch_bounds__(kB) k4a_axpy(long long n, const double* __restrict__ U0, const double* __restrict__ K,
                                               double a_dt, double* __restrict__ S)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    S[t] = K ? U0[t] + a_dt*K[t] : U0[t];
}
// U += b_dt K over n elements.
__global__ void __launch_bounds__(kB) k4a_accumulate(long long n, const double* __restrict__ K, double b_dt,
                                                     double* __restrict__ U)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= n) return;
    U[t] += b_dt*K[t];
}

// Restriction of level l onto its parent storage: every coarse cell a patch
// covers gets the mean of its children (summed k, j, i outermost to
// innermost), for the 8 components.
__global__ void __launch_bounds__(kB) k4a_restrict(AmrLevelView v, double* __restrict__ X)
{
    const int c0 = v.b[0]/v.r[0], c1 = v.b[1]/v.r[1], c2 = v.b[2]/v.r[2];
    const long long per = (long long)c0*c1*c2;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= per*v.npatch) return;
    const int n = (int)(t / per);
    const int loc = (int)(t % per);
    const int I = loc % c0, J = (loc / c0) % c1, K = loc / (c0*c1);
    const int* lo = v.plo + 3*n;
    const int N = v.r[0]*v.r[1]*v.r[2];
    const long long* ts = v.tRow + 3*n;
    const int gI = lo[0]/v.r[0] + I, gJ = lo[1]/v.r[1] + J, gK = lo[2]/v.r[2] + K;
    for (int c = 0; c < 8; ++c)
    {
        double sum = 0;
        for (int a2 = 0; a2 < v.r[2]; ++a2) for (int a1 = 0; a1 < v.r[1]; ++a1) for (int a0 = 0; a0 < v.r[0]; ++a0)
            sum += X[amrValid(v, n, c, lo[0] + I*v.r[0] + a0, lo[1] + J*v.r[1] + a1, lo[2] + K*v.r[2] + a2)];
        X[v.tBase[n] + c*ts[2] + gI + gJ*ts[0] + (long long)gK*ts[1]] = sum/N;
    }
}

// One-pass <JU> -> <U> of level 0 (whole panels): q = <JU>/J, the correction
// sum_d (J(x+e_d) - J(x-e_d))/2 deriv3case_d(q) (radial, then the angles),
// <U> = q - correction/(12 J). One block per angular column in the XCD-aware
// order, threads along the radius.
__global__ void __launch_bounds__(kB) k4a_deconv0(Prep4View v, const double* __restrict__ S, double* __restrict__ cart)
{
    const int nr = v.nr, na = v.na;
    const unsigned col = xcdColumn();
    const int j = (int)(col % (unsigned)na), k = (int)((col / (unsigned)na) % (unsigned)na), p = (int)(col / (unsigned)(na*na));
    auto J = [&](int ii, int jj, int kk) { return v.dvol[ii+6]*det4(v, jj, kk); };
#pragma unroll 1
    for (int c = 0; c < 8; ++c)
    {
        const long long base = p*v.vps + c*v.vcs;
        for (int i = threadIdx.x; i < nr; i += blockDim.x)
        {
            auto q = [&](int ii, int jj, int kk) { return S[base + ((long long)kk*na + jj)*nr + ii]/J(ii, jj, kk); };
            double corr = 0;
            corr += (0.5*(J(i+1, j, k)-J(i-1, j, k)))*deriv3([&](int x) { return q(x, j, k); }, i, 0, nr-1);
            corr += (0.5*(J(i, j+1, k)-J(i, j-1, k)))*deriv3([&](int x) { return q(i, x, k); }, j, 0, na-1);
            corr += (0.5*(J(i, j, k+1)-J(i, j, k-1)))*deriv3([&](int x) { return q(i, j, x); }, k, 0, na-1);
            cart[p*v.cps + c*v.ccs + ((long long)(k+kGC)*v.cJ + (j+kGC))*nr + i] = q(i, j, k)-corr/(12*J(i, j, k));
        }
    }
}

// Parent Cartesian averages on each patch's coarse source box.
__global__ void __launch_bounds__(kB) k4a_gather(AmrLevelView v, const double* __restrict__ parent)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.sCells*v.npatch) return;
    const int n = (int)(t / v.sCells);
    const long long loc = t % v.sCells;
    const long long e = v.srcMap[t];
    for (int c = 0; c < 8; ++c)
        v.src[((long long)n*8 + c)*v.sCells + loc] = e < 0 ? __builtin_nan("") : parent[e + c*v.pStride];
}

// Separable quartic interpolation, direction D (radial first): the input has
// fine extents in the directions before D and coarse source extents from D
// on; the output is fine in D too. D = 2 writes the donor ghosts (cells
// outside the patch whose 8 components are all finite; else 0).
template <int D>
__global__ void __launch_bounds__(kB) k4a_interp(AmrLevelView v)
{
    int ext[3], inExt[3];
    for (int d = 0; d < 3; ++d)
    {
        ext[d] = d <= D ? v.dn[d] : v.sn[d];
        inExt[d] = d < D ? v.dn[d] : v.sn[d];
    }
    const long long per = (long long)ext[0]*ext[1]*ext[2];
    const long long perIn = (long long)inExt[0]*inExt[1]*inExt[2];
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= per*v.npatch) return;
    const int n = (int)(t / per);
    const long long loc = t % per;
    int x[3] = {(int)(loc % ext[0]), (int)((loc / ext[0]) % ext[1]), (int)(loc / ((long long)ext[0]*ext[1]))};
    const int* lo = v.plo + 3*n;
    // fine index in D, its parent and child
    const int f = lo[D] - kAmrDg[D] + x[D];
    const int q = floorDivI(f, v.r[D]);
    const int child = f - q*v.r[D];
    const int slo = lo[D]/v.r[D] - (v.sn[D] - v.b[D]/v.r[D])/2;   // coarse source origin in D
    const double* w = v.w[D][child];
    const int rad = v.r[D] > 1 ? 2 : 0;
    const double* in = D == 0 ? v.src : D == 1 ? v.t1 : v.t2;
    long long inStride = 1;
    for (int d = 0; d < D; ++d) inStride *= inExt[d];
    int y[3] = {x[0], x[1], x[2]};
    y[D] = 0;
    const long long inBase = ((long long)y[2]*inExt[1] + y[1])*inExt[0] + y[0];
    if (D < 2)
    {
        double* out = D == 0 ? v.t1 : v.t2;
        for (int c = 0; c < 8; ++c)
        {
            const double* ic = in + ((long long)n*8 + c)*perIn + inBase;
            double value = 0;
            for (int o = -rad; o <= rad; ++o)
            {
                if (w[o+2] == 0) continue;
                value += w[o+2]*ic[(long long)(q+o-slo)*inStride];
            }
            out[((long long)n*8 + c)*per + loc] = value;
        }
        return;
    }
    // final direction: the donor ghost cells; the patch's own cells are 0
    // until k4a_deconv (the product correction of k4a_mapped reads them as
    // the reference does: its donor is zeroed before the interpolation fills
    // the ghosts, src/MHD_AMRFourthOrder.H:10,37-63)
    const int fi = lo[0] - kAmrDg[0] + x[0], fj = lo[1] - kAmrDg[1] + x[1], fk = f;
    if (fi >= lo[0] && fi < lo[0]+v.b[0] && fj >= lo[1] && fj < lo[1]+v.b[1] && fk >= lo[2] && fk < lo[2]+v.b[2])
    {
        for (int c = 0; c < 8; ++c) v.donor[((long long)n*8 + c)*v.dCells + loc] = 0.0;
        return;
    }
    double val[8];
    bool ok = true;
    for (int c = 0; c < 8; ++c)
    {
        const double* ic = in + ((long long)n*8 + c)*perIn + inBase;
        double value = 0;
        for (int o = -rad; o <= rad; ++o)
        {
            if (w[o+2] == 0) continue;
            value += w[o+2]*ic[(long long)(q+o-slo)*inStride];
        }
        val[c] = value;
        ok = ok && isfinite(value);
    }
    for (int c = 0; c < 8; ++c) v.donor[((long long)n*8 + c)*v.dCells + loc] = ok ? val[c] : 0.0;
}

// <JU> on each patch grown by 2: its own cells from S, cells another patch
// of the level holds from that patch, others J <U> + the average-product
// correction (J(x+e) - J(x-e)) (<U>(x+e) - <U>(x-e)) / 48 per direction.
__global__ void __launch_bounds__(kB) k4a_mapped(AmrLevelView v, const double* __restrict__ S)
{
    const int m0 = v.b[0]+4, m1 = v.b[1]+4, m2 = v.b[2]+4;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.mCells*v.npatch) return;
    const int n = (int)(t / v.mCells);
    const long long loc = t % v.mCells;
    const int* lo = v.plo + 3*n;
    const int p = v.pp[n];
    const int i = lo[0] - 2 + (int)(loc % m0), j = lo[1] - 2 + (int)((loc / m0) % m1), k = lo[2] - 2 + (int)(loc / ((long long)m0*m1));
    (void)m2;
    const int o = amrOwner(v, p, i, j, k);
    for (int c = 0; c < 8; ++c)
    {
        double value;
        if (o >= 0) value = S[amrValid(v, o, c, i, j, k)];
        else
        {
            auto dn = [&](int ii, int jj, int kk) { return v.donor[amrDonor(v, n, c, ii, jj, kk)]; };
            value = amrJ(v, i, j, k)*dn(i, j, k);
            value += (amrJ(v, i+1, j, k)-amrJ(v, i-1, j, k))*(dn(i+1, j, k)-dn(i-1, j, k))/48;
            value += (amrJ(v, i, j+1, k)-amrJ(v, i, j-1, k))*(dn(i, j+1, k)-dn(i, j-1, k))/48;
            value += (amrJ(v, i, j, k+1)-amrJ(v, i, j, k-1))*(dn(i, j, k+1)-dn(i, j, k-1))/48;
        }
        v.mapped[((long long)n*8 + c)*v.mCells + loc] = value;
    }
}

// <U> on each patch from <JU> on the patch grown by 2 (centred differences:
// the patch lies inside its panel).
__global__ void __launch_bounds__(kB) k4a_deconv(AmrLevelView v)
{
    const int m0 = v.b[0]+4, m1 = v.b[1]+4;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.vCells*v.npatch) return;
    const int n = (int)(t / v.vCells);
    const long long loc = t % v.vCells;
    const int* lo = v.plo + 3*n;
    const int i = lo[0] + (int)(loc % v.b[0]), j = lo[1] + (int)((loc / v.b[0]) % v.b[1]),
              k = lo[2] + (int)(loc / ((long long)v.b[0]*v.b[1]));
    for (int c = 0; c < 8; ++c)
    {
        const double* M = v.mapped + ((long long)n*8 + c)*v.mCells;
        auto q = [&](int ii, int jj, int kk) {
            return M[((long long)(kk-lo[2]+2)*m1 + (jj-lo[1]+2))*m0 + (ii-lo[0]+2)]/amrJ(v, ii, jj, kk); };
        double corr = 0;
        corr += (0.5*(amrJ(v, i+1, j, k)-amrJ(v, i-1, j, k)))*(0.5*(q(i+1, j, k)-q(i-1, j, k)));
        corr += (0.5*(amrJ(v, i, j+1, k)-amrJ(v, i, j-1, k)))*(0.5*(q(i, j+1, k)-q(i, j-1, k)));
        corr += (0.5*(amrJ(v, i, j, k+1)-amrJ(v, i, j, k-1)))*(0.5*(q(i, j, k+1)-q(i, j, k-1)));
        v.donor[amrDonor(v, n, c, i, j, k)] = q(i, j, k)-corr/(12*amrJ(v, i, j, k));
    }
}

// Same-level exchange: donor ghost cells another patch of the level holds
// take that patch's <U>.
__global__ void __launch_bounds__(kB) k4a_exchange(AmrLevelView v)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.dCells*v.npatch) return;
    const int n = (int)(t / v.dCells);
    const long long loc = t % v.dCells;
    const int* lo = v.plo + 3*n;
    const int i = lo[0] - kAmrDg[0] + (int)(loc % v.dn[0]), j = lo[1] - kAmrDg[1] + (int)((loc / v.dn[0]) % v.dn[1]),
              k = lo[2] - kAmrDg[2] + (int)(loc / ((long long)v.dn[0]*v.dn[1]));
    const int o = amrOwner(v, v.pp[n], i, j, k);
    if (o < 0 || o == n) return;
    for (int c = 0; c < 8; ++c) v.donor[((long long)n*8 + c)*v.dCells + loc] = v.donor[amrDonor(v, o, c, i, j, k)];
}

// Native conversion (the fourth-order matrix quotient of k4_native) of each
// patch grown by 5 into its stage input.
__global__ void __launch_bounds__(kB) k4a_native(AmrLevelView v, double* __restrict__ US)
{
    const int e0 = v.b[0]+10, e1 = v.b[1]+10;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.uCells*v.npatch) return;
    const int n = (int)(t / v.uCells);
    const long long loc = t % v.uCells;
    const int* lo = v.plo + 3*n;
    const int p = v.pp[n];
    const int i = lo[0] - 5 + (int)(loc % e0), j = lo[1] - 5 + (int)((loc / e0) % e1), k = lo[2] - 5 + (int)(loc / ((long long)e0*e1));
    auto ca = [&](int c, int jj, int kk) { return v.donor[amrDonor(v, n, c, i, jj, kk)]; };
    auto us = [&](int c) -> double& { return US[v.uOff[n] + c*v.uCells + loc]; };
    const Prep4View& pv = v.pv;
    us(0) = ca(0, j, k);
    us(4) = ca(4, j, k);
    for (int grp = 0; grp < 2; ++grp)
    {
        const int st = grp == 0 ? 1 : 5;
        auto q0 = [&](int jj, int kk, int row) {
            const double* Mi = Ainv4(pv, p, jj, kk);
            double val = 0;
            for (int cc = 0; cc < 3; ++cc) val += Mi[3*row+cc]*ca(st+cc, jj, kk);
            return val; };
        double rhs[3];
        for (int row = 0; row < 3; ++row)
        {
            double corr = 0;
            for (int d = 1; d < 3; ++d)
            {
                const int dj = d == 1, dk = d == 2;
                for (int cc = 0; cc < 3; ++cc)
                {
                    const double dA = 0.5*(Aavg4(pv, p, j+dj, k+dk)[3*row+cc]-Aavg4(pv, p, j-dj, k-dk)[3*row+cc]);
                    const double dq = 0.5*(q0(j+dj, k+dk, cc)-q0(j-dj, k-dk, cc));
                    corr += dA*dq;
                }
            }
            rhs[row] = ca(st+row, j, k)-corr/12.0;
        }
        const double* Mi = Ainv4(pv, p, j, k);
        for (int row = 0; row < 3; ++row)
        {
            double val = 0;
            for (int cc = 0; cc < 3; ++cc) val += Mi[3*row+cc]*rhs[cc];
            us(st+row) = val;
        }
    }
}

// Coarse-fine flux correction of one coarse face: the fine faces are
// fBase + a fs1 + b fs2 (a < fr1, b < fr2, b outer) in the fine box's scratch.
struct AmrRegFace
{
    long long cIdx, cEc;           // coarse face flux (comp 0) and component stride in scratch
    long long fBase, fEc, fs1, fs2;
    int fr1, fr2;
    double sn;                     // +- coarse cells per unit xi
};

// One thread per coarse cell next to a finer level: its faces in table order.
__global__ void __launch_bounds__(kB) k4a_reflux(int ntarget, const long long* __restrict__ tgt,
                                                 const long long* __restrict__ tcs, const int* __restrict__ start,
                                                 const AmrRegFace* __restrict__ faces, const double* __restrict__ Sc,
                                                 double* __restrict__ K)
{
    const int t = blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= ntarget) return;
    for (int c = 0; c < 8; ++c)
    {
        double acc = K[tgt[t] + c*tcs[t]];
        for (int e = start[t]; e < start[t+1]; ++e)
        {
            const AmrRegFace& F = faces[e];
            const double Fc = Sc[F.cIdx + c*F.cEc];
            double sum = 0;
            for (int b = 0; b < F.fr2; ++b) for (int a = 0; a < F.fr1; ++a)
                sum += Sc[F.fBase + c*F.fEc + a*F.fs1 + b*F.fs2];
            const double Ff = sum/(F.fr1*F.fr2);
            acc += F.sn*(Ff-Fc);
        }
        K[tgt[t] + c*tcs[t]] = acc;
    }
}
} // namespace

struct hc_amr4
{
    hc_options opt{};
    int nl = 0;
    std::vector<hc::ShellDims> dims;
    hc::Level4* L0 = nullptr;               // level-0 tables, seam gathers and Cartesian buffer
    hc_boxop* box = nullptr;                // level-0 tiles and every patch
    std::vector<AmrLevelView> lv;           // [l], l >= 1 (index 0 unused)
    long long nState = 0, nStage = 0, n0 = 0, us0 = 0;
    int ntarget = 0;
    long long *tgt = nullptr, *tcs = nullptr;
    int* tstart = nullptr;
    AmrRegFace* faces = nullptr;
    double *U0 = nullptr, *S = nullptr, *K = nullptr, *US = nullptr;
    std::vector<void*> owned;
    ~hc_amr4()
    {
        delete box;
        if (L0) hc::level4Destroy(L0);
        for (void* p : owned) (void)hipFree(p);
    }
};

namespace
{
// Host description of a refined patch.
struct AmrPatchH { int level, p, lo[3], hi[3]; long long vOff, uOff; };

double* amrAlloc(hc_amr4& A, long long n)
{
    double* p = nullptr;
    HC_THROW(hipMalloc(&p, std::max<long long>(1, n)*sizeof(double)));
    A.owned.push_back(p);
    return p;
}

dim3 amrGrid(long long n) { return dim3((unsigned)std::max<long long>(1, (n + kB - 1)/kB)); }

void amrRestrictAll(hc_amr4& A, double* X, hipStream_t s)
{
    for (int l = A.nl-1; l > 0; --l)
    {
        const AmrLevelView& v = A.lv[l];
        const long long per = (long long)(v.b[0]/v.r[0])*(v.b[1]/v.r[1])*(v.b[2]/v.r[2]);
        hipLaunchKernelGGL(k4a_restrict, amrGrid(per*v.npatch), dim3(kB), 0, s, v, X);
    }
}

// prepareFourthOrder of every level from the restricted stage state A.S.
void amrPrepare(hc_amr4& A, const double* inner, hipStream_t s)
{
    hc::Level4& L = *A.L0;
    const Prep4View& v0 = L.pv;
    const dim3 colBlock(std::min(kB, (v0.nr+63)/64*64));
    hipLaunchKernelGGL(k4a_deconv0, dim3(6u*v0.na*v0.na), colBlock, 0, s, v0, A.S, L.cart);
    if (L.ncol > 0)
        hipLaunchKernelGGL(k4_seam, dim3(L.ncol), colBlock, 0, s, v0, L.colDst, L.colStart, L.donorSrc, L.donorW,
                           L.cart);
    hipLaunchKernelGGL(k4_native, dim3(6u*(v0.na+10)*(v0.na+10)), colBlock, 0, s, v0, L.cart, A.US);
    hipLaunchKernelGGL(k4_radial_bc, dim3((10*6*(v0.na+10)*(v0.na+10)+kB-1)/kB), dim3(kB), 0, s, v0, inner, A.US);
    for (int l = 1; l < A.nl; ++l)
    {
        const AmrLevelView& v = A.lv[l];
        const double* parent = l == 1 ? L.cart : A.lv[l-1].donor;
        const long long np = v.npatch;
        hipLaunchKernelGGL(k4a_gather, amrGrid(v.sCells*np), dim3(kB), 0, s, v, parent);
        hipLaunchKernelGGL(k4a_interp<0>, amrGrid(v.t1Cells*np), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_interp<1>, amrGrid(v.t2Cells*np), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_interp<2>, amrGrid(v.dCells*np), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_mapped, amrGrid(v.mCells*np), dim3(kB), 0, s, v, A.S);
        hipLaunchKernelGGL(k4a_deconv, amrGrid(v.vCells*np), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_exchange, amrGrid(v.dCells*np), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_native, amrGrid(v.uCells*np), dim3(kB), 0, s, v, A.US);
    }
    HC_THROW(hipGetLastError());
}

// One stage from the restricted stage state A.S: K on every level after the
// flux registers and the restriction.
void amrStage(hc_amr4& A, const double* inner, double* K, hipStream_t s)
{
    amrPrepare(A, inner, s);
    double* const F[3] = {nullptr, nullptr, nullptr};
    if (A.opt.precision == 1) applyBoxop<float>(*A.box, A.US, K, F, s, 1);
    else applyBoxop<double>(*A.box, A.US, K, F, s, 1);
    if (A.ntarget > 0)
        hipLaunchKernelGGL(k4a_reflux, amrGrid(A.ntarget), dim3(kB), 0, s, A.ntarget, A.tgt, A.tcs, A.tstart,
                           A.faces, A.box->scratch, K);
    amrRestrictAll(A, K, s);
    HC_THROW(hipGetLastError());
}

const char* amrValidate(const hc_shell* shells, int nl, const std::vector<AmrPatchH>& P,
                        std::vector<std::array<int,3>>& bsize, std::vector<std::array<int,3>>& ratio,
                        std::vector<std::vector<int>>& lev, std::vector<std::vector<int>>& grid)
{
    bsize.assign(nl, {0, 0, 0});
    ratio.assign(nl, {1, 1, 1});
    lev.assign(nl, {});
    grid.assign(nl, {});
    for (int l = 1; l < nl; ++l)
    {
        const hc_shell &f = shells[l], &c = shells[l-1];
        if (f.inner_radius != c.inner_radius || f.outer_radius != c.outer_radius || f.radial_stretch != c.radial_stretch)
            return "every level must have the same radii and radial stretch";
        if (f.n_rad % c.n_rad || f.n_ang % c.n_ang) return "level cell counts must be multiples of the coarser level's";
        ratio[l] = {f.n_rad/c.n_rad, f.n_ang/c.n_ang, f.n_ang/c.n_ang};
        for (int d = 0; d < 3; ++d)
            if (ratio[l][d] < 1 || ratio[l][d] > kAmrMaxRatio) return "refinement ratios must be 1 .. 4";
    }
    for (size_t n = 0; n < P.size(); ++n)
    {
        const AmrPatchH& q = P[n];
        if (q.level < 1 || q.level >= nl) return "patch level must be 1 .. nlevels-1 (level 0 is the whole shell)";
        if (q.p < 0 || q.p > 5) return "patch block must be 0..5";
        std::array<int,3> b;
        for (int d = 0; d < 3; ++d) b[d] = q.hi[d]-q.lo[d]+1;
        if (bsize[q.level][0] == 0) bsize[q.level] = b;
        if (b != bsize[q.level]) return "the patches of one level must share one box size";
        for (int d = 0; d < 3; ++d)
        {
            if (b[d] < 4) return "patches must span at least 4 cells per direction";
            if (q.lo[d] % b[d]) return "patches must be aligned to their box size";
            if (b[d] % ratio[q.level][d]) return "the box size must be a multiple of the refinement ratio";
        }
        lev[q.level].push_back((int)n);
    }
    for (int l = 1; l < nl; ++l)
    {
        if (lev[l].empty()) return "a refined level has no patches";
        const auto& b = bsize[l];
        const int G0 = shells[l].n_rad/b[0], G1 = shells[l].n_ang/b[1], G2 = shells[l].n_ang/b[2];
        grid[l].assign((size_t)6*G0*G1*G2, -1);
        for (int n : lev[l])
        {
            const AmrPatchH& q = P[n];
            // the donor box (patch grown by 6, 8, 8) inside the panel and the radial domain
            if (q.lo[0]-kAmrDg[0] < 0 || q.hi[0]+kAmrDg[0] >= shells[l].n_rad)
                return "refined patches must keep 6 cells from the radial boundaries";
            for (int d = 1; d < 3; ++d)
                if (q.lo[d]-kAmrDg[d] < 0 || q.hi[d]+kAmrDg[d] >= shells[l].n_ang)
                    return "refined patches must keep 8 cells from the panel edges";
            int& slot = grid[l][(((size_t)q.p*G2 + q.lo[2]/b[2])*G1 + q.lo[1]/b[1])*G0 + q.lo[0]/b[0]];
            if (slot >= 0) return "duplicate patch";
            slot = n;
        }
    }
    auto owner = [&](int l, int p, int i, int j, int k) {
        if (l == 0) return (i >= 0 && i < shells[0].n_rad && j >= 0 && j < shells[0].n_ang && k >= 0
                            && k < shells[0].n_ang) ? -2 : -1;
        if (i < 0 || j < 0 || k < 0 || i >= shells[l].n_rad || j >= shells[l].n_ang || k >= shells[l].n_ang) return -1;
        const auto& b = bsize[l];
        const int G0 = shells[l].n_rad/b[0], G1 = shells[l].n_ang/b[1], G2 = shells[l].n_ang/b[2];
        return grid[l][(((size_t)p*G2 + k/b[2])*G1 + j/b[1])*G0 + i/b[0]];
    };
    for (int l = 1; l < nl; ++l)
    {
        const auto& r = ratio[l];
        for (int n : lev[l])
        {
            const AmrPatchH& q = P[n];
            int clo[3], chi[3], cg[3];
            for (int d = 0; d < 3; ++d)
            {
                clo[d] = q.lo[d]/r[d]; chi[d] = q.hi[d]/r[d];
                cg[d] = (kAmrDg[d]+r[d]-1)/r[d] + (r[d] > 1 ? 2 : 0);
            }
            if (l >= 2)
            {
                const int o = owner(l-1, q.p, clo[0], clo[1], clo[2]);
                if (o < 0 || owner(l-1, q.p, chi[0], chi[1], chi[2]) != o)
                    return "a refined patch must lie inside one patch of its parent level";
            }
            // interpolation sources: cells the parent level defines
            for (int k = clo[2]-cg[2]; k <= chi[2]+cg[2]; ++k) for (int j = clo[1]-cg[1]; j <= chi[1]+cg[1]; ++j)
                for (int i = clo[0]-cg[0]; i <= chi[0]+cg[0]; ++i)
                {
                    bool ok;
                    if (l == 1) ok = i >= 0 && i < shells[0].n_rad && j >= -kGC && j < shells[0].n_ang+kGC
                                     && k >= -kGC && k < shells[0].n_ang+kGC;
                    else
                    {
                        ok = false;
                        for (int m : lev[l-1])
                        {
                            const AmrPatchH& Q = P[m];
                            if (Q.p == q.p && i >= Q.lo[0]-kAmrCg && i <= Q.hi[0]+kAmrCg && j >= Q.lo[1]-kAmrCg
                                && j <= Q.hi[1]+kAmrCg && k >= Q.lo[2]-kAmrCg && k <= Q.hi[2]+kAmrCg) { ok = true; break; }
                        }
                    }
                    if (!ok) return "a refined patch's interpolation source leaves the cells its parent level defines";
                }
            // proper nesting: the coarse cell across every coarse-fine face is a parent-level cell
            for (int d = 0; d < 3; ++d) for (int sd = 0; sd < 2; ++sd)
            {
                const int a1 = std::min((d+1)%3, (d+2)%3), a2 = std::max((d+1)%3, (d+2)%3);
                for (int B = clo[a2]; B <= chi[a2]; ++B) for (int Aa = clo[a1]; Aa <= chi[a1]; ++Aa)
                {
                    int C[3];
                    C[d] = sd ? chi[d]+1 : clo[d]-1; C[a1] = Aa; C[a2] = B;
                    if (owner(l, q.p, C[0]*r[0], C[1]*r[1], C[2]*r[2]) >= 0) continue;
                    if (owner(l-1, q.p, C[0], C[1], C[2]) == -1)
                        return "refined levels must be properly nested (a coarse-fine face borders a cell the parent level lacks)";
                }
            }
        }
    }
    return nullptr;
}
} // namespace

extern "C" {

int hc_amr4_create(const hc_shell* shells, int nlevels, const hc_options* options, const hc_box* patches,
                   const int* patch_levels, int npatch, hc_amr4** out)
{
    if (!shells || nlevels <= 0 || !options || !out || npatch < 0 || (npatch > 0 && (!patches || !patch_levels)))
        return hc::fail("hc_amr4_create: bad argument");
    if (options->spatial_order != 4) return hc::fail("hc_amr4_create: spatial_order must be 4");
    if (options->riemann_solver != 1 && options->riemann_solver != 2)
        return hc::fail("hc_amr4_create: riemann_solver must be 1 (Rusanov) or 2 (Roe)");
    if (options->limiter != 0 && options->limiter != 1) return hc::fail("hc_amr4_create: limiter must be 0 or 1");
    if (options->precision != 0 && options->precision != 1) return hc::fail("hc_amr4_create: precision must be 0 or 1");
    if (const char* e = hc::checkOperatorOptions(*options)) return hc::fail(std::string("hc_amr4_create: ") + e);
    for (int l = 0; l < nlevels; ++l)
    {
        const hc_shell& sh = shells[l];
        if (sh.n_rad < 4 || sh.n_ang < 16 || !(sh.outer_radius > sh.inner_radius && sh.inner_radius > 0))
            return hc::fail("hc_amr4_create: bad level shell (need n_rad >= 4, n_ang >= 16 and 0 < r0 < r1)");
    }
    std::vector<AmrPatchH> P(npatch);
    for (int n = 0; n < npatch; ++n)
    {
        P[n].level = patch_levels[n]; P[n].p = patches[n].block;
        for (int d = 0; d < 3; ++d) { P[n].lo[d] = patches[n].lo[d]; P[n].hi[d] = patches[n].hi[d]; }
    }
    std::vector<std::array<int,3>> bsize, ratio;
    std::vector<std::vector<int>> lev, grid;
    if (const char* e = amrValidate(shells, nlevels, P, bsize, ratio, lev, grid))
        return hc::fail(std::string("hc_amr4_create: ") + e);
    try
    {
        std::unique_ptr<hc_amr4> A(new hc_amr4());
        A->opt = *options;
        A->nl = nlevels;
        for (int l = 0; l < nlevels; ++l)
            A->dims.push_back({shells[l].n_rad, shells[l].n_ang, kG, shells[l].inner_radius, shells[l].outer_radius,
                               shells[l].radial_stretch});
        const int tile[3] = {32, 32, 32};
        A->L0 = hc::level4Create(shells[0], *options, tile, false);
        const Prep4View& v0 = A->L0->pv;
        A->n0 = 6*v0.vps;
        A->us0 = 6*v0.sps;
        long long off = A->n0, uo = A->us0;
        for (auto& q : P)
        {
            const long long n0 = q.hi[0]-q.lo[0]+1, n1 = q.hi[1]-q.lo[1]+1, n2 = q.hi[2]-q.lo[2]+1;
            q.vOff = off; q.uOff = uo;
            off += 8*n0*n1*n2;
            uo += 8*(n0+10)*(n1+10)*(n2+10);
        }
        A->nState = off; A->nStage = uo;
        // per refined level: tables, patch arrays and the coarse source maps
        A->lv.assign(nlevels, AmrLevelView{});
        for (int l = 1; l < nlevels; ++l)
        {
            AmrLevelView& v = A->lv[l];
            hc::prep4Tables(shells[l], *options, v.pv, A->owned);
            const auto& b = bsize[l];
            const auto& r = ratio[l];
            v.npatch = (int)lev[l].size();
            v.nr = shells[l].n_rad; v.na = shells[l].n_ang;
            for (int d = 0; d < 3; ++d)
            {
                v.r[d] = r[d]; v.b[d] = b[d];
                const int cg = (kAmrDg[d]+r[d]-1)/r[d] + (r[d] > 1 ? 2 : 0);
                v.sn[d] = b[d]/r[d] + 2*cg;
                v.dn[d] = b[d] + 2*kAmrDg[d];
                for (int c = 0; c < kAmrMaxRatio; ++c)
                {
                    double w[5] = {0, 0, 1, 0, 0};
                    if (c < r[d]) hc::amrChildWeights(r[d], c, w);
                    for (int a = 0; a < 5; ++a) v.w[d][c][a] = w[a];
                }
            }
            v.G[0] = v.nr/b[0]; v.G[1] = v.na/b[1]; v.G[2] = v.na/b[2];
            v.vCells = (long long)b[0]*b[1]*b[2];
            v.dCells = (long long)v.dn[0]*v.dn[1]*v.dn[2];
            v.mCells = (long long)(b[0]+4)*(b[1]+4)*(b[2]+4);
            v.sCells = (long long)v.sn[0]*v.sn[1]*v.sn[2];
            v.t1Cells = (long long)v.dn[0]*v.sn[1]*v.sn[2];
            v.t2Cells = (long long)v.dn[0]*v.dn[1]*v.sn[2];
            v.uCells = (long long)(b[0]+10)*(b[1]+10)*(b[2]+10);
            // level-local patch index -> arrays
            std::vector<int> pp, plo, g(grid[l].size(), -1);
            std::vector<long long> vOff, uOff, tBase, tRow;
            std::vector<int> local(P.size(), -1);
            for (int m = 0; m < v.npatch; ++m) local[lev[l][m]] = m;
            for (size_t x = 0; x < g.size(); ++x) g[x] = grid[l][x] >= 0 ? local[grid[l][x]] : -1;
            for (int m = 0; m < v.npatch; ++m)
            {
                const AmrPatchH& q = P[lev[l][m]];
                pp.push_back(q.p);
                for (int d = 0; d < 3; ++d) plo.push_back(q.lo[d]);
                vOff.push_back(q.vOff); uOff.push_back(q.uOff);
                // restriction target: level-0 layout or the parent patch holding the coarsened patch
                if (l == 1)
                {
                    tBase.push_back(q.p*v0.vps);
                    tRow.push_back(v0.nr); tRow.push_back((long long)v0.nr*v0.na); tRow.push_back(v0.vcs);
                }
                else
                {
                    const auto& bp = bsize[l-1];
                    const int Gp0 = shells[l-1].n_rad/bp[0], Gp1 = shells[l-1].n_ang/bp[1], Gp2 = shells[l-1].n_ang/bp[2];
                    const int I = q.lo[0]/r[0], J = q.lo[1]/r[1], K = q.lo[2]/r[2];
                    const int o = grid[l-1][(((size_t)q.p*Gp2 + K/bp[2])*Gp1 + J/bp[1])*Gp0 + I/bp[0]];
                    const AmrPatchH& Q = P[o];
                    tBase.push_back(Q.vOff - Q.lo[0] - (long long)Q.lo[1]*bp[0] - (long long)Q.lo[2]*bp[0]*bp[1]);
                    tRow.push_back(bp[0]); tRow.push_back((long long)bp[0]*bp[1]);
                    tRow.push_back((long long)bp[0]*bp[1]*bp[2]);
                }
            }
            // coarse source maps
            std::vector<long long> smap((size_t)v.npatch*v.sCells, -1);
            const int cgv[3] = {(v.sn[0]-b[0]/r[0])/2, (v.sn[1]-b[1]/r[1])/2, (v.sn[2]-b[2]/r[2])/2};
            for (int m = 0; m < v.npatch; ++m)
            {
                const AmrPatchH& q = P[lev[l][m]];
                for (int kk = 0; kk < v.sn[2]; ++kk) for (int jj = 0; jj < v.sn[1]; ++jj) for (int ii = 0; ii < v.sn[0]; ++ii)
                {
                    const int i = q.lo[0]/r[0] - cgv[0] + ii, j = q.lo[1]/r[1] - cgv[1] + jj, k = q.lo[2]/r[2] - cgv[2] + kk;
                    long long e = -1;
                    if (l == 1)
                        e = q.p*v0.cps + ((long long)(k+kGC)*v0.cJ + (j+kGC))*v0.nr + i;
                    else
                    {
                        const AmrLevelView& vp = A->lv[l-1];
                        const auto& bp = bsize[l-1];
                        int o = -1;
                        if (i >= 0 && j >= 0 && k >= 0 && i < vp.nr && j < vp.na && k < vp.na)
                            o = grid[l-1][(((size_t)q.p*vp.G[2] + k/bp[2])*vp.G[1] + j/bp[1])*vp.G[0] + i/bp[0]];
                        if (o < 0)
                            for (int c : lev[l-1])
                            {
                                const AmrPatchH& Q = P[c];
                                if (Q.p == q.p && i >= Q.lo[0]-kAmrCg && i <= Q.hi[0]+kAmrCg && j >= Q.lo[1]-kAmrCg
                                    && j <= Q.hi[1]+kAmrCg && k >= Q.lo[2]-kAmrCg && k <= Q.hi[2]+kAmrCg) { o = c; break; }
                            }
                        if (o >= 0)
                        {
                            int mo = 0;
                            for (int x = 0; x < (int)lev[l-1].size(); ++x) if (lev[l-1][x] == o) mo = x;
                            const AmrPatchH& Q = P[o];
                            e = (long long)mo*8*vp.dCells + ((long long)(k-Q.lo[2]+kAmrDg[2])*vp.dn[1]
                                + (j-Q.lo[1]+kAmrDg[1]))*vp.dn[0] + (i-Q.lo[0]+kAmrDg[0]);
                        }
                    }
                    smap[(size_t)m*v.sCells + ((size_t)kk*v.sn[1]+jj)*v.sn[0]+ii] = e;
                }
            }
            v.pStride = l == 1 ? v0.ccs : A->lv[l-1].dCells;
            v.grid = upload(g, A->owned);
            v.pp = upload(pp, A->owned);
            v.plo = upload(plo, A->owned);
            v.vOff = upload(vOff, A->owned);
            v.uOff = upload(uOff, A->owned);
            v.tBase = upload(tBase, A->owned);
            v.tRow = upload(tRow, A->owned);
            v.srcMap = upload(smap, A->owned);
            const long long np = v.npatch;
            v.src = amrAlloc(*A, 8*np*v.sCells);
            v.t1 = amrAlloc(*A, 8*np*v.t1Cells);
            v.t2 = amrAlloc(*A, 8*np*v.t2Cells);
            v.donor = amrAlloc(*A, 8*np*v.dCells);
            v.mapped = amrAlloc(*A, 8*np*v.mCells);
            HC_THROW(hipMemset(v.donor, 0, 8*np*v.dCells*sizeof(double)));
        }
        // one box table: level-0 tiles, then every patch
        std::vector<hc_box> boxes;
        std::vector<int> blev;
        const int nr0 = shells[0].n_rad, na0 = shells[0].n_ang;
        const int bs[3] = {std::min(tile[0], nr0), std::min(tile[1], na0), std::min(tile[2], na0)};
        for (int p = 0; p < 6; ++p)
            for (int k = 0; k < na0; k += bs[2]) for (int j = 0; j < na0; j += bs[1]) for (int i = 0; i < nr0; i += bs[0])
            {
                hc_box b{};
                b.block = p;
                b.lo[0] = i; b.lo[1] = j; b.lo[2] = k;
                b.hi[0] = std::min(i+bs[0], nr0)-1; b.hi[1] = std::min(j+bs[1], na0)-1; b.hi[2] = std::min(k+bs[2], na0)-1;
                boxes.push_back(b); blev.push_back(0);
            }
        const size_t ntile = boxes.size();
        for (int n = 0; n < npatch; ++n) { boxes.push_back(patches[n]); blev.push_back(patch_levels[n]); }
        std::unique_ptr<hc_boxop> B(new hc_boxop());
        B->opt = *options;
        B->levels = A->dims;
        B->dims = A->dims[0];
        buildBoxop(*B, boxes.data(), blev.data(), (int)boxes.size());
        for (size_t n = 0; n < B->boxes.size(); ++n)
        {
            BoxDesc& bd = B->boxes[n];
            if (n < ntile)
            {
                bd.uOff = bd.p*v0.sps + ((long long)bd.lo[2]*v0.sJ + bd.lo[1])*v0.sI + bd.lo[0];
                bd.us[0] = v0.sI; bd.us[1] = v0.sI*v0.sJ; bd.us[2] = v0.scs;
                bd.rOff = bd.p*v0.vps + ((long long)bd.lo[2]*na0 + bd.lo[1])*nr0 + bd.lo[0];
                bd.rs[0] = nr0; bd.rs[1] = (long long)nr0*na0; bd.rs[2] = v0.vcs;
            }
            else
            {
                bd.uOff = P[n-ntile].uOff;
                bd.rOff = P[n-ntile].vOff;
            }
        }
        HC_THROW(hipMemcpy(const_cast<BoxDesc*>(B->view.boxes), B->boxes.data(), B->boxes.size()*sizeof(BoxDesc),
                           hipMemcpyHostToDevice));
        std::vector<SeamPair> pairs = hc::panelSeamPairs(B->boxes, ntile, nr0, na0);
        B->view.seams = upload(pairs, B->owned);
        B->view.nseams = (int)pairs.size();
        // flux registers: coarse cells next to a finer level and their faces
        auto boxOfCell = [&](int l, int p, int i, int j, int k) -> int {
            if (l == 0)
            {
                for (size_t n = 0; n < ntile; ++n)
                {
                    const BoxDesc& b = B->boxes[n];
                    if (b.p == p && i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] && k <= b.hi[2])
                        return (int)n;
                }
                return -1;
            }
            const auto& b = bsize[l];
            const int G0 = shells[l].n_rad/b[0], G1 = shells[l].n_ang/b[1], G2 = shells[l].n_ang/b[2];
            if (i < 0 || j < 0 || k < 0 || i >= shells[l].n_rad || j >= shells[l].n_ang || k >= shells[l].n_ang) return -1;
            const int o = grid[l][(((size_t)p*G2 + k/b[2])*G1 + j/b[1])*G0 + i/b[0]];
            return o < 0 ? -1 : (int)ntile + o;
        };
        auto scratchFace = [&](int n, int d, const int f[3], long long& ec) {
            const BoxDesc& b = B->boxes[n];
            const long long e0 = b.n[0]+2*kG, e1 = b.n[1]+2*kG, e2 = b.n[2]+2*kG;
            ec = e0*e1*e2;
            return b.sOff + (F_FLUX+8*d)*ec + ((long long)(f[2]-b.lo[2]+kG)*e1 + (f[1]-b.lo[1]+kG))*e0 + (f[0]-b.lo[0]+kG);
        };
        std::map<long long, int> targetIndex;
        std::vector<long long> tgt, tcs;
        std::vector<std::vector<AmrRegFace>> tf;
        for (int l = nlevels-1; l > 0; --l)
        {
            const auto& r = ratio[l];
            for (int n : lev[l])
            {
                const AmrPatchH& q = P[n];
                for (int d = 0; d < 3; ++d) for (int sd = 0; sd < 2; ++sd)
                {
                    const int xf = sd ? q.hi[d]+1 : q.lo[d];
                    const int X = xf/r[d];
                    const int a1 = std::min((d+1)%3, (d+2)%3), a2 = std::max((d+1)%3, (d+2)%3);
                    for (int Bb = q.lo[a2]/r[a2]; Bb <= q.hi[a2]/r[a2]; ++Bb)
                        for (int Aa = q.lo[a1]/r[a1]; Aa <= q.hi[a1]/r[a1]; ++Aa)
                        {
                            int C[3];
                            C[d] = sd ? X : X-1; C[a1] = Aa; C[a2] = Bb;
                            if (boxOfCell(l, q.p, C[0]*r[0], C[1]*r[1], C[2]*r[2]) >= 0) continue;
                            const int cb = boxOfCell(l-1, q.p, C[0], C[1], C[2]);
                            if (cb < 0) throw std::runtime_error("coarse-fine face without a parent cell");
                            AmrRegFace F{};
                            int cf[3]; cf[d] = X; cf[a1] = Aa; cf[a2] = Bb;
                            F.cIdx = scratchFace(cb, d, cf, F.cEc);
                            int ff[3]; ff[d] = xf; ff[a1] = Aa*r[a1]; ff[a2] = Bb*r[a2];
                            const int fb = (int)ntile + n;
                            F.fBase = scratchFace(fb, d, ff, F.fEc);
                            const BoxDesc& fd = B->boxes[fb];
                            const long long st[3] = {1, fd.n[0]+2*kG, (long long)(fd.n[0]+2*kG)*(fd.n[1]+2*kG)};
                            F.fs1 = st[a1]; F.fs2 = st[a2];
                            F.fr1 = r[a1]; F.fr2 = r[a2];
                            F.sn = (sd ? 1.0 : -1.0)*(d == 0 ? shells[l-1].n_rad : shells[l-1].n_ang);
                            long long t, cs;
                            if (l-1 == 0)
                            {
                                t = q.p*v0.vps + ((long long)C[2]*na0 + C[1])*nr0 + C[0];
                                cs = v0.vcs;
                            }
                            else
                            {
                                const int o = cb - (int)ntile;
                                const AmrPatchH& Q = P[o];
                                const auto& bp = bsize[l-1];
                                t = Q.vOff + ((long long)(C[2]-Q.lo[2])*bp[1] + (C[1]-Q.lo[1]))*bp[0] + (C[0]-Q.lo[0]);
                                cs = (long long)bp[0]*bp[1]*bp[2];
                            }
                            auto it = targetIndex.find(t);
                            if (it == targetIndex.end())
                            {
                                it = targetIndex.emplace(t, (int)tgt.size()).first;
                                tgt.push_back(t); tcs.push_back(cs); tf.emplace_back();
                            }
                            tf[it->second].push_back(F);
                        }
                }
            }
        }
        std::vector<int> tstart{0};
        std::vector<AmrRegFace> faces;
        for (auto& v : tf) { faces.insert(faces.end(), v.begin(), v.end()); tstart.push_back((int)faces.size()); }
        A->ntarget = (int)tgt.size();
        A->tgt = upload(tgt, A->owned);
        A->tcs = upload(tcs, A->owned);
        A->tstart = upload(tstart, A->owned);
        A->faces = upload(faces, A->owned);
        A->box = B.release();
        A->U0 = amrAlloc(*A, A->nState);
        A->S = amrAlloc(*A, A->nState);
        A->K = amrAlloc(*A, A->nState);
        A->US = amrAlloc(*A, A->nStage);
        HC_THROW(hipMemset(A->US, 0, A->nStage*sizeof(double)));
        *out = A.release();
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_amr4_create: ") + e.what()); }
}

void hc_amr4_destroy(hc_amr4* amr) { delete amr; }

long long hc_amr4_count(const hc_amr4* amr, int what)
{
    if (!amr) return -1;
    switch (what)
    {
    case 0: return amr->nState;
    case 1: return amr->nStage;
    case 2: return amr->n0;
    default: return -1;
    }
}

int hc_amr4_prepare(hc_amr4* amr, const double* JU, const double* inner_ghosts, double* USph, hc_stream stream)
{
    if (!amr || !JU || !inner_ghosts || !USph) return hc::fail("hc_amr4_prepare: null argument");
    try
    {
        const hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k4a_axpy, amrGrid(amr->nState), dim3(kB), 0, s, amr->nState, JU, (const double*)nullptr,
                           0.0, amr->S);
        amrRestrictAll(*amr, amr->S, s);
        amrPrepare(*amr, inner_ghosts, s);
        HC_THROW(hipMemcpyAsync(USph, amr->US, amr->nStage*sizeof(double), hipMemcpyDeviceToDevice, s));
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_amr4_prepare: ") + e.what()); }
}

int hc_amr4_stage_rhs(hc_amr4* amr, const double* JU, const double* inner_ghosts, double* K, hc_stream stream)
{
    if (!amr || !JU || !inner_ghosts || !K) return hc::fail("hc_amr4_stage_rhs: null argument");
    try
    {
        const hipStream_t s = static_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k4a_axpy, amrGrid(amr->nState), dim3(kB), 0, s, amr->nState, JU, (const double*)nullptr,
                           0.0, amr->S);
        amrRestrictAll(*amr, amr->S, s);
        amrStage(*amr, inner_ghosts, K, s);
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_amr4_stage_rhs: ") + e.what()); }
}

int hc_amr4_rk4_advance(hc_amr4* amr, double* JU, const double* inner_ghosts, double dt, int nsteps, hc_stream stream)
{
    if (!amr || !JU || !inner_ghosts || nsteps < 0) return hc::fail("hc_amr4_rk4_advance: bad argument");
    try
    {
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const double a[4] = {0.0, 0.5, 0.5, 1.0};
        const double b[4] = {1.0/6, 1.0/3, 1.0/3, 1.0/6};
        const long long n = amr->nState;
        for (int step = 0; step < nsteps; ++step)
        {
            HC_THROW(hipMemcpyAsync(amr->U0, JU, n*sizeof(double), hipMemcpyDeviceToDevice, s));
            for (int st = 0; st < 4; ++st)
            {
                hipLaunchKernelGGL(k4a_axpy, amrGrid(n), dim3(kB), 0, s, n, amr->U0,
                                   st == 0 ? (const double*)nullptr : (const double*)amr->K, dt*a[st], amr->S);
                amrRestrictAll(*amr, amr->S, s);
                amrStage(*amr, inner_ghosts, amr->K, s);
                hipLaunchKernelGGL(k4a_accumulate, amrGrid(n), dim3(kB), 0, s, n, (const double*)amr->K, dt*b[st], JU);
            }
            amrRestrictAll(*amr, JU, s);
        }
        HC_THROW(hipGetLastError());
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_amr4_rk4_advance: ") + e.what()); }
}

} // extern "C"
