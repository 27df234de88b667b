// Fixed AMR hierarchy of fourth-order levels advanced by the synchronous
// classical RK4 of the reference (advanceRungeKuttaSynchronous,
// src/MHD_AMRRungeKutta.H:5-41). Included by hc_o4.hip after the level and box
// table code, whose kernels it launches.
//
// Layout. Level 0 covers the whole shell (the valid level layout
// [6][8][na][na][nr]); level l >= 1 is a set of patches of one box size,
// aligned to it, each [8][n2][n1][n0]. The combined state (and RHS) is the
// level-0 array followed by every patch in table order; the combined stage
// input of hc_amr4_prepare is the level-0 stage layout followed by every
// patch grown by 5.
//
// Tiles. The reference prepares and evaluates each patch on its own ghosted
// box. Here the patches of a refined level are grouped into rectangular
// tiles of whole patches (at most a set extent), and every tile is prepared
// and evaluated as one box. The results are the same cell for cell: a
// patch's ghost cells inside a neighbour take the neighbour's values (the
// exchange), a cell outside the level has one interpolated value whichever
// patch asks, and the box operator's per-face arithmetic does not depend on
// the box that evaluates it; the first ghost layer's average-product
// correction reads zero across a patch face only where that face is on the
// level's boundary, which is the tile's boundary too. The tile form does
// about one tenth of the work on the 8^3 patches of a dense level.
//
// One stage (prepareFourthOrder, src/MHD_AMRFourthOrder.H:5-187; evaluate,
// src/MHD_AMR.H:1417-1499; the flux registers and restriction of
// src/MHD_AMRRungeKutta.H:30-36):
//  1. stage state S = U0 + a dt K_prev, restricted fine -> coarse (the mean of
//     the children, restrictTo, src/MHD_AMR.H:597-607);
//  2. level 0: <JU> -> <U> in one pass with the full Jacobian J = dVolr det A
//     (k4a_deconv0), then the seam interpolation, the native conversion and
//     the radial boundary ghosts of the single-level path;
//  3. level l >= 1, in order, per tile: the parent's Cartesian averages on the
//     coarse source box (k4a_gather), the separable quartic interpolation onto
//     the tile grown by (6,8,8) (k4a_interp<d>; the tile's own cells are 0),
//     <JU> on the tile grown by 2 (the level's own <JU> where a patch holds
//     the cell, else J <U> + the average-product correction, k4a_mapped), <U>
//     on the tile (k4a_deconv), the ghosts other tiles of the level hold
//     (k4a_exchange) and the native conversion on the tile grown by 5
//     (k4a_native);
//  4. the box operator over level-0 tiles and every refined tile in one box
//     table; the refined tiles' RHS goes to the patches (k4a_scatter);
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
// per-tile regions of the flattened launches
enum AmrKind { KS = 0, KT1, KT2, KD, KM, KV, KU, kAmrKinds };

struct AmrTile
{
    int p, lo[3], n[3];             // tile box: a union of whole patches
    int sn[3], dn[3];               // coarse source and donor box extents
    int slo[3];                     // coarse source box origin
};

struct AmrLevelView
{
    int npatch, ntile, nr, na;
    int r[3];                       // refinement ratio to the parent
    int b[3];                       // patch size
    int G[3];                       // patch grid (nr / b0, na / b1, na / b2)
    double w[3][kAmrMaxRatio][5];   // quartic child weights per direction
    long long vCells;               // cells of a patch
    const int* grid;                // [6][G2][G1][G0] -> patch index or -1
    const int* patchTile;           // tile of each patch
    const int* pp;                  // panel of each patch
    const int* plo;                 // lo of each patch [3 n]
    const long long* vOff;          // valid offset of each patch in the combined state
    const long long* tBase;         // restriction target of each patch: element of coarse cell (0,0,0), comp 0
    const long long* tRow;          // its strides (j, k, component) in the parent storage, 3 per patch
    const AmrTile* tiles;
    const long long* start;         // [kind][ntile + 1]: first cell of each tile in each region kind
    long long tot[kAmrKinds];       // cells of each kind over the level (component strides of its arrays)
    const long long* srcMap;        // per coarse source cell: parent element (comp 0) or -1
    long long pStride;              // parent component stride
    double *src, *t1, *t2, *donor, *mapped;
    double* Kt;                     // tile RHS: per tile [8][cells], at 8 start[KV][t]
    long long usBase;               // element offset of the level's tile stage inputs in the stage input
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
// tile of flattened cell t of region `kind` (binary search of the starts)
__device__ __forceinline__ int amrTileOf(const AmrLevelView& v, int kind, long long t)
{
    const long long* s = v.start + (long long)kind*(v.ntile + 1);
    int lo = 0, hi = v.ntile - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (s[mid] <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ long long amrStart(const AmrLevelView& v, int kind, int ti)
{
    return v.start[(long long)kind*(v.ntile + 1) + ti];
}
// donor (tile grown by dg), level-wide component-major
__device__ __forceinline__ long long amrDonor(const AmrLevelView& v, int ti, int c, int i, int j, int k)
{
    const AmrTile& T = v.tiles[ti];
    return c*v.tot[KD] + amrStart(v, KD, ti)
         + ((long long)(k-T.lo[2]+kAmrDg[2])*T.dn[1] + (j-T.lo[1]+kAmrDg[1]))*T.dn[0] + (i-T.lo[0]+kAmrDg[0]);
}
__device__ __forceinline__ double amrJ(const AmrLevelView& v, int i, int j, int k)
{
    return v.pv.dvol[i+6]*det4(v.pv, j, k);
}
__device__ __forceinline__ int floorDivI(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }
__device__ __forceinline__ void amrCell(long long loc, int e0, int e1, int& x, int& y, int& z)
{
    x = (int)(loc % e0); y = (int)((loc / e0) % e1); z = (int)(loc / ((long long)e0*e1));
}

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

// Parent Cartesian averages on each tile's coarse source box.
__global__ void __launch_bounds__(kB) k4a_gather(AmrLevelView v, const double* __restrict__ parent)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.tot[KS]) return;
    const long long e = v.srcMap[t];
    for (int c = 0; c < 8; ++c) v.src[c*v.tot[KS] + t] = e < 0 ? __builtin_nan("") : parent[e + c*v.pStride];
}

// Separable quartic interpolation, direction D (radial first): the input has
// fine extents in the directions before D and coarse source extents from D
// on; the output is fine in D too. D = 2 writes the donor: its ghost cells
// (0 unless all 8 components are finite) and zeros on the tile's own cells,
// which k4a_deconv fills after k4a_mapped has read them as the reference
// does (its donor is zeroed before the interpolation fills the ghosts,
// src/MHD_AMRFourthOrder.H:10,37-63).
template <int D>
__global__ void __launch_bounds__(kB) k4a_interp(AmrLevelView v)
{
    constexpr int outKind = D == 0 ? KT1 : D == 1 ? KT2 : KD;
    constexpr int inKind = D == 0 ? KS : D == 1 ? KT1 : KT2;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.tot[outKind]) return;
    const int ti = amrTileOf(v, outKind, t);
    const AmrTile& T = v.tiles[ti];
    int ext[3], inExt[3];
    for (int d = 0; d < 3; ++d)
    {
        ext[d] = d <= D ? T.dn[d] : T.sn[d];
        inExt[d] = d < D ? T.dn[d] : T.sn[d];
    }
    const long long loc = t - amrStart(v, outKind, ti);
    int x[3];
    amrCell(loc, ext[0], ext[1], x[0], x[1], x[2]);
    const int f = T.lo[D] - kAmrDg[D] + x[D];
    const int q = floorDivI(f, v.r[D]);
    const int child = f - q*v.r[D];
    const double* w = v.w[D][child];
    const int rad = v.r[D] > 1 ? 2 : 0;
    const double* in = D == 0 ? v.src : D == 1 ? v.t1 : v.t2;
    long long inStride = 1;
    for (int d = 0; d < D; ++d) inStride *= inExt[d];
    int y[3] = {x[0], x[1], x[2]};
    y[D] = 0;
    const long long inBase = amrStart(v, inKind, ti) + ((long long)y[2]*inExt[1] + y[1])*inExt[0] + y[0];
    const long long inTot = v.tot[inKind];
    if (D < 2)
    {
        double* out = D == 0 ? v.t1 : v.t2;
        for (int c = 0; c < 8; ++c)
        {
            const double* ic = in + c*inTot + inBase;
            double value = 0;
            for (int o = -rad; o <= rad; ++o)
            {
                if (w[o+2] == 0) continue;
                value += w[o+2]*ic[(long long)(q+o-T.slo[D])*inStride];
            }
            out[c*v.tot[outKind] + t] = value;
        }
        return;
    }
    const int fi = T.lo[0] - kAmrDg[0] + x[0], fj = T.lo[1] - kAmrDg[1] + x[1], fk = f;
    if (fi >= T.lo[0] && fi < T.lo[0]+T.n[0] && fj >= T.lo[1] && fj < T.lo[1]+T.n[1] && fk >= T.lo[2] && fk < T.lo[2]+T.n[2])
    {
        for (int c = 0; c < 8; ++c) v.donor[c*v.tot[KD] + t] = 0.0;
        return;
    }
    double val[8];
    bool ok = true;
    for (int c = 0; c < 8; ++c)
    {
        const double* ic = in + c*inTot + inBase;
        double value = 0;
        for (int o = -rad; o <= rad; ++o)
        {
            if (w[o+2] == 0) continue;
            value += w[o+2]*ic[(long long)(q+o-T.slo[D])*inStride];
        }
        val[c] = value;
        ok = ok && isfinite(value);
    }
    for (int c = 0; c < 8; ++c) v.donor[c*v.tot[KD] + t] = ok ? val[c] : 0.0;
}

// <JU> on each tile grown by 2: cells a patch of the level holds from that
// patch's S, others J <U> + the average-product correction
// (J(x+e) - J(x-e)) (<U>(x+e) - <U>(x-e)) / 48 per direction.
__global__ void __launch_bounds__(kB) k4a_mapped(AmrLevelView v, const double* __restrict__ S)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.tot[KM]) return;
    const int ti = amrTileOf(v, KM, t);
    const AmrTile& T = v.tiles[ti];
    int x, y, z;
    amrCell(t - amrStart(v, KM, ti), T.n[0]+4, T.n[1]+4, x, y, z);
    const int i = T.lo[0] - 2 + x, j = T.lo[1] - 2 + y, k = T.lo[2] - 2 + z;
    const int o = amrOwner(v, T.p, i, j, k);
    for (int c = 0; c < 8; ++c)
    {
        double value;
        if (o >= 0) value = S[amrValid(v, o, c, i, j, k)];
        else
        {
            auto dn = [&](int ii, int jj, int kk) { return v.donor[amrDonor(v, ti, c, ii, jj, kk)]; };
            value = amrJ(v, i, j, k)*dn(i, j, k);
            value += (amrJ(v, i+1, j, k)-amrJ(v, i-1, j, k))*(dn(i+1, j, k)-dn(i-1, j, k))/48;
            value += (amrJ(v, i, j+1, k)-amrJ(v, i, j-1, k))*(dn(i, j+1, k)-dn(i, j-1, k))/48;
            value += (amrJ(v, i, j, k+1)-amrJ(v, i, j, k-1))*(dn(i, j, k+1)-dn(i, j, k-1))/48;
        }
        v.mapped[c*v.tot[KM] + t] = value;
    }
}

// <U> on each tile from <JU> on the tile grown by 2 (centred differences:
// the tile lies inside its panel).
__global__ void __launch_bounds__(kB) k4a_deconv(AmrLevelView v)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.tot[KV]) return;
    const int ti = amrTileOf(v, KV, t);
    const AmrTile& T = v.tiles[ti];
    int x, y, z;
    amrCell(t - amrStart(v, KV, ti), T.n[0], T.n[1], x, y, z);
    const int i = T.lo[0] + x, j = T.lo[1] + y, k = T.lo[2] + z;
    const int m0 = T.n[0]+4, m1 = T.n[1]+4;
    const long long mBase = amrStart(v, KM, ti);
    for (int c = 0; c < 8; ++c)
    {
        const double* M = v.mapped + c*v.tot[KM] + mBase;
        auto q = [&](int ii, int jj, int kk) {
            return M[((long long)(kk-T.lo[2]+2)*m1 + (jj-T.lo[1]+2))*m0 + (ii-T.lo[0]+2)]/amrJ(v, ii, jj, kk); };
        double corr = 0;
        corr += (0.5*(amrJ(v, i+1, j, k)-amrJ(v, i-1, j, k)))*(0.5*(q(i+1, j, k)-q(i-1, j, k)));
        corr += (0.5*(amrJ(v, i, j+1, k)-amrJ(v, i, j-1, k)))*(0.5*(q(i, j+1, k)-q(i, j-1, k)));
        corr += (0.5*(amrJ(v, i, j, k+1)-amrJ(v, i, j, k-1)))*(0.5*(q(i, j, k+1)-q(i, j, k-1)));
        v.donor[amrDonor(v, ti, c, i, j, k)] = q(i, j, k)-corr/(12*amrJ(v, i, j, k));
    }
}

// Donor ghost cells another tile of the level holds take that tile's <U>.
__global__ void __launch_bounds__(kB) k4a_exchange(AmrLevelView v)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.tot[KD]) return;
    const int ti = amrTileOf(v, KD, t);
    const AmrTile& T = v.tiles[ti];
    int x, y, z;
    amrCell(t - amrStart(v, KD, ti), T.dn[0], T.dn[1], x, y, z);
    const int i = T.lo[0] - kAmrDg[0] + x, j = T.lo[1] - kAmrDg[1] + y, k = T.lo[2] - kAmrDg[2] + z;
    const int o = amrOwner(v, T.p, i, j, k);
    if (o < 0) return;
    const int to = v.patchTile[o];
    if (to == ti) return;
    for (int c = 0; c < 8; ++c) v.donor[c*v.tot[KD] + t] = v.donor[amrDonor(v, to, c, i, j, k)];
}

// Native conversion (the fourth-order matrix quotient of k4_native) of each
// tile grown by 5 into its stage input (box-table layout, [8][cells]).
__global__ void __launch_bounds__(kB) k4a_native(AmrLevelView v, double* __restrict__ US)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.tot[KU]) return;
    const int ti = amrTileOf(v, KU, t);
    const AmrTile& T = v.tiles[ti];
    const long long loc = t - amrStart(v, KU, ti);
    const long long cells = (long long)(T.n[0]+10)*(T.n[1]+10)*(T.n[2]+10);
    int x, y, z;
    amrCell(loc, T.n[0]+10, T.n[1]+10, x, y, z);
    const int p = T.p;
    const int i = T.lo[0] - 5 + x, j = T.lo[1] - 5 + y, k = T.lo[2] - 5 + z;
    auto ca = [&](int c, int jj, int kk) { return v.donor[amrDonor(v, ti, c, i, jj, kk)]; };
    double* const out = US + v.usBase + 8*amrStart(v, KU, ti) + loc;
    auto us = [&](int c) -> double& { return out[c*cells]; };
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

// The tiles' RHS onto the patches of the combined layout.
__global__ void __launch_bounds__(kB) k4a_scatter(AmrLevelView v, double* __restrict__ K)
{
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.tot[KV]) return;
    const int ti = amrTileOf(v, KV, t);
    const AmrTile& T = v.tiles[ti];
    const long long loc = t - amrStart(v, KV, ti);
    const long long cells = (long long)T.n[0]*T.n[1]*T.n[2];
    int x, y, z;
    amrCell(loc, T.n[0], T.n[1], x, y, z);
    const int i = T.lo[0] + x, j = T.lo[1] + y, k = T.lo[2] + z;
    const int o = amrOwner(v, T.p, i, j, k);
    const double* src = v.Kt + 8*amrStart(v, KV, ti) + loc;
    for (int c = 0; c < 8; ++c) K[amrValid(v, o, c, i, j, k)] = src[c*cells];
}

// The stage input of every patch grown by 5 (the layout of hc_amr4_prepare)
// from its tile's.
__global__ void __launch_bounds__(kB) k4a_patch_us(AmrLevelView v, const double* __restrict__ US,
                                                   const long long* __restrict__ uOff, double* __restrict__ out)
{
    const int e0 = v.b[0]+10, e1 = v.b[1]+10, e2 = v.b[2]+10;
    const long long per = (long long)e0*e1*e2;
    const long long t = (long long)blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= per*v.npatch) return;
    const int n = (int)(t / per);
    const long long loc = t % per;
    int x, y, z;
    amrCell(loc, e0, e1, x, y, z);
    const int* lo = v.plo + 3*n;
    const int ti = v.patchTile[n];
    const AmrTile& T = v.tiles[ti];
    const int i = lo[0] - 5 + x, j = lo[1] - 5 + y, k = lo[2] - 5 + z;
    const long long cells = (long long)(T.n[0]+10)*(T.n[1]+10)*(T.n[2]+10);
    const long long src = v.usBase + 8*amrStart(v, KU, ti)
                        + ((long long)(k-T.lo[2]+5)*(T.n[1]+10) + (j-T.lo[1]+5))*(T.n[0]+10) + (i-T.lo[0]+5);
    for (int c = 0; c < 8; ++c) out[uOff[n] + c*per + loc] = US[src + c*cells];
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

namespace
{
// Host description of a refined patch.
struct AmrPatchH { int level, p, lo[3], hi[3]; long long vOff, uOff; };
} // namespace

struct hc_amr4
{
    hc_options opt{};
    int nl = 0;
    std::vector<hc_shell> shells;
    std::vector<hc::ShellDims> dims;
    hc::Level4* L0 = nullptr;               // level-0 tables, seam gathers and Cartesian buffer
    hc_boxop* box = nullptr;                // level-0 tiles and every refined tile
    std::vector<AmrLevelView> lv;           // [l], l >= 1 (index 0 unused)
    std::vector<AmrPatchH> P;               // refined patches in table order
    std::vector<std::array<int,3>> bsize, ratio;
    std::vector<std::vector<int>> lev, grid;
    int tile0[3] = {128, 64, 64};           // level-0 tile shape (clipped to the level)
    int fineMax[3] = {128, 64, 64};         // largest refined tile
    long long nState = 0, nStage = 0, n0 = 0, us0 = 0, usInternal = 0, kInternal = 0;
    int ntarget = 0;
    long long *tgt = nullptr, *tcs = nullptr;
    int* tstart = nullptr;
    AmrRegFace* faces = nullptr;
    std::vector<long long*> patchUOff;      // [l]: stage input offset of each patch (hc_amr4_prepare layout)
    double *U0 = nullptr, *S = nullptr, *K = nullptr, *US = nullptr;
    std::vector<void*> owned;               // fixed with the hierarchy
    std::vector<void*> tileOwned;           // rebuilt with the tiles
    void freeTiles()
    {
        delete box;
        box = nullptr;
        for (void* p : tileOwned) (void)hipFree(p);
        tileOwned.clear();
    }
    ~hc_amr4()
    {
        freeTiles();
        if (L0) hc::level4Destroy(L0);
        for (void* p : owned) (void)hipFree(p);
    }
};

namespace
{
double* amrAlloc(std::vector<void*>& owned, long long n)
{
    double* p = nullptr;
    HC_THROW(hipMalloc(&p, std::max<long long>(1, n)*sizeof(double)));
    owned.push_back(p);
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

// prepareFourthOrder of every level from the restricted stage state A.S into
// the internal stage input A.US.
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
        hipLaunchKernelGGL(k4a_gather, amrGrid(v.tot[KS]), dim3(kB), 0, s, v, parent);
        hipLaunchKernelGGL(k4a_interp<0>, amrGrid(v.tot[KT1]), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_interp<1>, amrGrid(v.tot[KT2]), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_interp<2>, amrGrid(v.tot[KD]), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_mapped, amrGrid(v.tot[KM]), dim3(kB), 0, s, v, A.S);
        hipLaunchKernelGGL(k4a_deconv, amrGrid(v.tot[KV]), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_exchange, amrGrid(v.tot[KD]), dim3(kB), 0, s, v);
        hipLaunchKernelGGL(k4a_native, amrGrid(v.tot[KU]), dim3(kB), 0, s, v, A.US);
    }
    HC_THROW(hipGetLastError());
}

// One stage from the restricted stage state A.S: A.K on every level after
// the flux registers and the restriction.
void amrStage(hc_amr4& A, const double* inner, hipStream_t s)
{
    amrPrepare(A, inner, s);
    double* const F[3] = {nullptr, nullptr, nullptr};
    if (A.opt.precision == 1) applyBoxop<float>(*A.box, A.US, A.K, F, s, 1);
    else applyBoxop<double>(*A.box, A.US, A.K, F, s, 1);
    for (int l = 1; l < A.nl; ++l)
        hipLaunchKernelGGL(k4a_scatter, amrGrid(A.lv[l].tot[KV]), dim3(kB), 0, s, A.lv[l], A.K);
    if (A.ntarget > 0)
        hipLaunchKernelGGL(k4a_reflux, amrGrid(A.ntarget), dim3(kB), 0, s, A.ntarget, A.tgt, A.tcs, A.tstart,
                           A.faces, A.box->scratch, A.K);
    amrRestrictAll(A, A.K, s);
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

// Groups the patches of every refined level into tiles (greedy: from each
// unassigned patch, extend the radial run, then whole rows, then whole
// planes of unassigned patches, each up to A.fineMax cells), builds the
// per-tile arrays, the coarse source maps, the box table over level-0 tiles
// and refined tiles, and the flux registers.
void amrBuildTiles(hc_amr4& A)
{
    A.freeTiles();
    auto& owned = A.tileOwned;
    const Prep4View& v0 = A.L0->pv;
    const hc_shell* shells = A.shells.data();
    std::vector<std::vector<AmrTile>> tiles(A.nl);
    std::vector<std::vector<int>> patchTile(A.nl);
    std::vector<std::array<std::vector<long long>, kAmrKinds>> start(A.nl);
    long long usOff = A.us0, ktOff = 0;
    std::vector<long long> ktBase(A.nl, 0);
    for (int l = 1; l < A.nl; ++l)
    {
        AmrLevelView& v = A.lv[l];
        const auto& b = A.bsize[l];
        const auto& r = A.ratio[l];
        const int G0 = v.G[0], G1 = v.G[1], G2 = v.G[2];
        const auto& grid = A.grid[l];
        std::vector<char> asg(grid.size(), 0);
        auto slot = [&](int p, int a0, int a1, int a2) { return (((size_t)p*G2 + a2)*G1 + a1)*G0 + a0; };
        auto freeSlot = [&](int p, int a0, int a1, int a2) { const size_t s = slot(p, a0, a1, a2);
                                                            return grid[s] >= 0 && !asg[s]; };
        int M[3];
        for (int d = 0; d < 3; ++d) M[d] = std::max(b[d], A.fineMax[d]);
        patchTile[l].assign(A.P.size(), -1);
        for (int p = 0; p < 6; ++p) for (int a2 = 0; a2 < G2; ++a2) for (int a1 = 0; a1 < G1; ++a1) for (int a0 = 0; a0 < G0; ++a0)
        {
            if (!freeSlot(p, a0, a1, a2)) continue;
            int e0 = 1, e1 = 1, e2 = 1;
            while (a0+e0 < G0 && (e0+1)*b[0] <= M[0] && freeSlot(p, a0+e0, a1, a2)) ++e0;
            auto rowFree = [&](int y, int z) { for (int x = 0; x < e0; ++x) if (!freeSlot(p, a0+x, y, z)) return false;
                                               return true; };
            while (a1+e1 < G1 && (e1+1)*b[1] <= M[1] && rowFree(a1+e1, a2)) ++e1;
            auto planeFree = [&](int z) { for (int y = 0; y < e1; ++y) if (!rowFree(a1+y, z)) return false; return true; };
            while (a2+e2 < G2 && (e2+1)*b[2] <= M[2] && planeFree(a2+e2)) ++e2;
            AmrTile T{};
            T.p = p;
            const int e[3] = {e0, e1, e2}, a[3] = {a0, a1, a2};
            for (int d = 0; d < 3; ++d)
            {
                T.lo[d] = a[d]*b[d]; T.n[d] = e[d]*b[d];
                const int cg = (kAmrDg[d]+r[d]-1)/r[d] + (r[d] > 1 ? 2 : 0);
                T.sn[d] = T.n[d]/r[d] + 2*cg;
                T.slo[d] = T.lo[d]/r[d] - cg;
                T.dn[d] = T.n[d] + 2*kAmrDg[d];
            }
            const int ti = (int)tiles[l].size();
            for (int z = 0; z < e2; ++z) for (int y = 0; y < e1; ++y) for (int x = 0; x < e0; ++x)
            {
                const size_t s = slot(p, a0+x, a1+y, a2+z);
                asg[s] = 1;
                patchTile[l][grid[s]] = ti;
            }
            tiles[l].push_back(T);
        }
        // flattened starts of every region kind
        for (int k = 0; k < kAmrKinds; ++k) start[l][k].assign(tiles[l].size()+1, 0);
        for (size_t t = 0; t < tiles[l].size(); ++t)
        {
            const AmrTile& T = tiles[l][t];
            const long long cnt[kAmrKinds] = {
                (long long)T.sn[0]*T.sn[1]*T.sn[2], (long long)T.dn[0]*T.sn[1]*T.sn[2],
                (long long)T.dn[0]*T.dn[1]*T.sn[2], (long long)T.dn[0]*T.dn[1]*T.dn[2],
                (long long)(T.n[0]+4)*(T.n[1]+4)*(T.n[2]+4), (long long)T.n[0]*T.n[1]*T.n[2],
                (long long)(T.n[0]+10)*(T.n[1]+10)*(T.n[2]+10)};
            for (int k = 0; k < kAmrKinds; ++k) start[l][k][t+1] = start[l][k][t] + cnt[k];
        }
        v.ntile = (int)tiles[l].size();
        for (int k = 0; k < kAmrKinds; ++k) v.tot[k] = start[l][k][v.ntile];
        v.usBase = usOff;
        usOff += 8*v.tot[KU];
        ktBase[l] = ktOff;
        ktOff += 8*v.tot[KV];
        // level-local patch -> tile
        std::vector<int> ptile(v.npatch);
        for (int m = 0; m < v.npatch; ++m) ptile[m] = patchTile[l][A.lev[l][m]];
        v.patchTile = upload(ptile, owned);
        v.tiles = upload(tiles[l], owned);
        std::vector<long long> st;
        for (int k = 0; k < kAmrKinds; ++k) st.insert(st.end(), start[l][k].begin(), start[l][k].end());
        v.start = upload(st, owned);
        // coarse source maps: level 0's Cartesian buffer, or the donor of the
        // parent tile holding the cell (or whose Cartesian ghost box holds it)
        std::vector<long long> smap((size_t)v.tot[KS], -1);
        for (int t = 0; t < v.ntile; ++t)
        {
            const AmrTile& T = tiles[l][t];
            for (int kk = 0; kk < T.sn[2]; ++kk) for (int jj = 0; jj < T.sn[1]; ++jj) for (int ii = 0; ii < T.sn[0]; ++ii)
            {
                const int i = T.slo[0] + ii, j = T.slo[1] + jj, k = T.slo[2] + kk;
                long long e = -1;
                if (l == 1)
                {
                    if (i >= 0 && i < v0.nr && j >= -kGC && j < v0.na+kGC && k >= -kGC && k < v0.na+kGC)
                        e = T.p*v0.cps + ((long long)(k+kGC)*v0.cJ + (j+kGC))*v0.nr + i;
                }
                else
                {
                    const AmrLevelView& vp = A.lv[l-1];
                    const auto& bp = A.bsize[l-1];
                    int to = -1;
                    if (i >= 0 && j >= 0 && k >= 0 && i < vp.nr && j < vp.na && k < vp.na)
                    {
                        const int o = A.grid[l-1][(((size_t)T.p*vp.G[2] + k/bp[2])*vp.G[1] + j/bp[1])*vp.G[0] + i/bp[0]];
                        if (o >= 0) to = patchTile[l-1][o];
                    }
                    if (to < 0)
                        for (int q = 0; q < (int)tiles[l-1].size(); ++q)
                        {
                            const AmrTile& Q = tiles[l-1][q];
                            if (Q.p == T.p && i >= Q.lo[0]-kAmrCg && i < Q.lo[0]+Q.n[0]+kAmrCg
                                && j >= Q.lo[1]-kAmrCg && j < Q.lo[1]+Q.n[1]+kAmrCg
                                && k >= Q.lo[2]-kAmrCg && k < Q.lo[2]+Q.n[2]+kAmrCg) { to = q; break; }
                        }
                    if (to >= 0)
                    {
                        const AmrTile& Q = tiles[l-1][to];
                        e = start[l-1][KD][to] + ((long long)(k-Q.lo[2]+kAmrDg[2])*Q.dn[1]
                            + (j-Q.lo[1]+kAmrDg[1]))*Q.dn[0] + (i-Q.lo[0]+kAmrDg[0]);
                    }
                }
                smap[(size_t)start[l][KS][t] + ((size_t)kk*T.sn[1]+jj)*T.sn[0]+ii] = e;
            }
        }
        v.pStride = l == 1 ? v0.ccs : A.lv[l-1].tot[KD];
        v.srcMap = upload(smap, owned);
        v.src = amrAlloc(owned, 8*v.tot[KS]);
        v.t1 = amrAlloc(owned, 8*v.tot[KT1]);
        v.t2 = amrAlloc(owned, 8*v.tot[KT2]);
        v.donor = amrAlloc(owned, 8*v.tot[KD]);
        v.mapped = amrAlloc(owned, 8*v.tot[KM]);
        HC_THROW(hipMemset(v.donor, 0, 8*v.tot[KD]*sizeof(double)));
    }
    A.usInternal = usOff;
    A.kInternal = A.nState + ktOff;
    A.US = amrAlloc(owned, A.usInternal);
    HC_THROW(hipMemset(A.US, 0, A.usInternal*sizeof(double)));
    A.K = amrAlloc(owned, A.kInternal);
    for (int l = 1; l < A.nl; ++l) A.lv[l].Kt = A.K + A.nState + ktBase[l];
    // one box table: level-0 tiles, then every refined tile
    std::vector<hc_box> boxes;
    std::vector<int> blev;
    const int nr0 = shells[0].n_rad, na0 = shells[0].n_ang;
    const int bs[3] = {std::min(A.tile0[0], nr0), std::min(A.tile0[1], na0), std::min(A.tile0[2], na0)};
    for (int p = 0; p < 6; ++p)
        for (int k = 0; k < na0; k += bs[2]) for (int j = 0; j < na0; j += bs[1]) for (int i = 0; i < nr0; i += bs[0])
        {
            hc_box b{};
            b.block = p;
            b.lo[0] = i; b.lo[1] = j; b.lo[2] = k;
            b.hi[0] = std::min(i+bs[0], nr0)-1; b.hi[1] = std::min(j+bs[1], na0)-1; b.hi[2] = std::min(k+bs[2], na0)-1;
            boxes.push_back(b); blev.push_back(0);
        }
    const size_t ntile0 = boxes.size();
    std::vector<size_t> tileBox(A.nl, 0);
    for (int l = 1; l < A.nl; ++l)
    {
        tileBox[l] = boxes.size();
        for (const AmrTile& T : tiles[l])
        {
            hc_box b{};
            b.block = T.p;
            for (int d = 0; d < 3; ++d) { b.lo[d] = T.lo[d]; b.hi[d] = T.lo[d]+T.n[d]-1; }
            boxes.push_back(b); blev.push_back(l);
        }
    }
    std::unique_ptr<hc_boxop> B(new hc_boxop());
    B->opt = A.opt;
    B->levels = A.dims;
    B->dims = A.dims[0];
    buildBoxop(*B, boxes.data(), blev.data(), (int)boxes.size());
    for (size_t n = 0; n < B->boxes.size(); ++n)
    {
        BoxDesc& bd = B->boxes[n];
        if (n < ntile0)
        {
            bd.uOff = bd.p*v0.sps + ((long long)bd.lo[2]*v0.sJ + bd.lo[1])*v0.sI + bd.lo[0];
            bd.us[0] = v0.sI; bd.us[1] = v0.sI*v0.sJ; bd.us[2] = v0.scs;
            bd.rOff = bd.p*v0.vps + ((long long)bd.lo[2]*na0 + bd.lo[1])*nr0 + bd.lo[0];
            bd.rs[0] = nr0; bd.rs[1] = (long long)nr0*na0; bd.rs[2] = v0.vcs;
            continue;
        }
        int l = A.nl-1;
        while (n < tileBox[l]) --l;
        const int t = (int)(n - tileBox[l]);
        bd.uOff = A.lv[l].usBase + 8*start[l][KU][t];
        bd.rOff = A.nState + ktBase[l] + 8*start[l][KV][t];
    }
    HC_THROW(hipMemcpy(const_cast<BoxDesc*>(B->view.boxes), B->boxes.data(), B->boxes.size()*sizeof(BoxDesc),
                       hipMemcpyHostToDevice));
    std::vector<SeamPair> pairs = hc::panelSeamPairs(B->boxes, ntile0, nr0, na0);
    B->view.seams = upload(pairs, B->owned);
    B->view.nseams = (int)pairs.size();
    // flux registers: coarse cells next to a finer level and their faces
    auto boxOfCell = [&](int l, int p, int i, int j, int k) -> int {
        if (l == 0)
        {
            for (size_t n = 0; n < ntile0; ++n)
            {
                const BoxDesc& b = B->boxes[n];
                if (b.p == p && i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] && k <= b.hi[2])
                    return (int)n;
            }
            return -1;
        }
        const auto& b = A.bsize[l];
        const AmrLevelView& v = A.lv[l];
        if (i < 0 || j < 0 || k < 0 || i >= v.nr || j >= v.na || k >= v.na) return -1;
        const int o = A.grid[l][(((size_t)p*v.G[2] + k/b[2])*v.G[1] + j/b[1])*v.G[0] + i/b[0]];
        return o < 0 ? -1 : (int)(tileBox[l] + patchTile[l][o]);
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
    for (int l = A.nl-1; l > 0; --l)
    {
        const auto& r = A.ratio[l];
        for (int n : A.lev[l])
        {
            const AmrPatchH& q = A.P[n];
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
                        int fc[3]; fc[d] = sd ? xf-1 : xf; fc[a1] = ff[a1]; fc[a2] = ff[a2];
                        const int fb = boxOfCell(l, q.p, fc[0], fc[1], fc[2]);
                        F.fBase = scratchFace(fb, d, ff, F.fEc);
                        const BoxDesc& fd = B->boxes[fb];
                        const long long stv[3] = {1, fd.n[0]+2*kG, (long long)(fd.n[0]+2*kG)*(fd.n[1]+2*kG)};
                        F.fs1 = stv[a1]; F.fs2 = stv[a2];
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
                            const auto& bp = A.bsize[l-1];
                            const AmrLevelView& vp = A.lv[l-1];
                            const int o = A.grid[l-1][(((size_t)q.p*vp.G[2] + C[2]/bp[2])*vp.G[1] + C[1]/bp[1])*vp.G[0]
                                                      + C[0]/bp[0]];
                            const AmrPatchH& Q = A.P[o];
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
    for (auto& x : tf) { faces.insert(faces.end(), x.begin(), x.end()); tstart.push_back((int)faces.size()); }
    A.ntarget = (int)tgt.size();
    A.tgt = upload(tgt, owned);
    A.tcs = upload(tcs, owned);
    A.tstart = upload(tstart, owned);
    A.faces = upload(faces, owned);
    A.box = B.release();
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
    if (options->limiter < 0 || options->limiter > 2) return hc::fail("hc_amr4_create: limiter must be 0, 1 or 2");
    if (options->precision != 0 && options->precision != 1) return hc::fail("hc_amr4_create: precision must be 0 or 1");
    if (const char* e = hc::checkOperatorOptions(*options)) return hc::fail(std::string("hc_amr4_create: ") + e);
    for (int l = 0; l < nlevels; ++l)
    {
        const hc_shell& sh = shells[l];
        if (sh.n_rad < 4 || sh.n_ang < 16 || !(sh.outer_radius > sh.inner_radius && sh.inner_radius > 0))
            return hc::fail("hc_amr4_create: bad level shell (need n_rad >= 4, n_ang >= 16 and 0 < r0 < r1)");
    }
    std::unique_ptr<hc_amr4> A(new hc_amr4());
    A->P.resize(npatch);
    for (int n = 0; n < npatch; ++n)
    {
        AmrPatchH& q = A->P[n];
        q.level = patch_levels[n]; q.p = patches[n].block;
        for (int d = 0; d < 3; ++d) { q.lo[d] = patches[n].lo[d]; q.hi[d] = patches[n].hi[d]; }
    }
    if (const char* e = amrValidate(shells, nlevels, A->P, A->bsize, A->ratio, A->lev, A->grid))
        return hc::fail(std::string("hc_amr4_create: ") + e);
    try
    {
        A->opt = *options;
        A->nl = nlevels;
        A->shells.assign(shells, shells + nlevels);
        for (int l = 0; l < nlevels; ++l)
            A->dims.push_back({shells[l].n_rad, shells[l].n_ang, kG, shells[l].inner_radius, shells[l].outer_radius,
                               shells[l].radial_stretch});
        const int tile[3] = {32, 32, 32};
        A->L0 = hc::level4Create(shells[0], *options, tile, false);
        const Prep4View& v0 = A->L0->pv;
        A->n0 = 6*v0.vps;
        A->us0 = 6*v0.sps;
        long long off = A->n0, uo = A->us0;
        for (auto& q : A->P)
        {
            const long long n0 = q.hi[0]-q.lo[0]+1, n1 = q.hi[1]-q.lo[1]+1, n2 = q.hi[2]-q.lo[2]+1;
            q.vOff = off; q.uOff = uo;
            off += 8*n0*n1*n2;
            uo += 8*(n0+10)*(n1+10)*(n2+10);
        }
        A->nState = off; A->nStage = uo;
        // per refined level: metric tables, patch arrays, child weights
        A->lv.assign(nlevels, AmrLevelView{});
        A->patchUOff.assign(nlevels, nullptr);
        for (int l = 1; l < nlevels; ++l)
        {
            AmrLevelView& v = A->lv[l];
            hc::prep4Tables(shells[l], *options, v.pv, A->owned);
            const auto& b = A->bsize[l];
            const auto& r = A->ratio[l];
            v.npatch = (int)A->lev[l].size();
            v.nr = shells[l].n_rad; v.na = shells[l].n_ang;
            for (int d = 0; d < 3; ++d)
            {
                v.r[d] = r[d]; v.b[d] = b[d];
                for (int c = 0; c < kAmrMaxRatio; ++c)
                {
                    double w[5] = {0, 0, 1, 0, 0};
                    if (c < r[d]) hc::amrChildWeights(r[d], c, w);
                    for (int a = 0; a < 5; ++a) v.w[d][c][a] = w[a];
                }
            }
            v.G[0] = v.nr/b[0]; v.G[1] = v.na/b[1]; v.G[2] = v.na/b[2];
            v.vCells = (long long)b[0]*b[1]*b[2];
            std::vector<int> pp, plo, g(A->grid[l].size(), -1);
            std::vector<long long> vOff, uOff, tBase, tRow;
            std::vector<int> local(A->P.size(), -1);
            for (int m = 0; m < v.npatch; ++m) local[A->lev[l][m]] = m;
            for (size_t x = 0; x < g.size(); ++x) g[x] = A->grid[l][x] >= 0 ? local[A->grid[l][x]] : -1;
            for (int m = 0; m < v.npatch; ++m)
            {
                const AmrPatchH& q = A->P[A->lev[l][m]];
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
                    const auto& bp = A->bsize[l-1];
                    const AmrLevelView& vp = A->lv[l-1];
                    const int I = q.lo[0]/r[0], J = q.lo[1]/r[1], K = q.lo[2]/r[2];
                    const int o = A->grid[l-1][(((size_t)q.p*vp.G[2] + K/bp[2])*vp.G[1] + J/bp[1])*vp.G[0] + I/bp[0]];
                    const AmrPatchH& Q = A->P[o];
                    tBase.push_back(Q.vOff - Q.lo[0] - (long long)Q.lo[1]*bp[0] - (long long)Q.lo[2]*bp[0]*bp[1]);
                    tRow.push_back(bp[0]); tRow.push_back((long long)bp[0]*bp[1]);
                    tRow.push_back((long long)bp[0]*bp[1]*bp[2]);
                }
            }
            v.grid = upload(g, A->owned);
            v.pp = upload(pp, A->owned);
            v.plo = upload(plo, A->owned);
            v.vOff = upload(vOff, A->owned);
            v.tBase = upload(tBase, A->owned);
            v.tRow = upload(tRow, A->owned);
            A->patchUOff[l] = upload(uOff, A->owned);
        }
        A->U0 = amrAlloc(A->owned, A->nState);
        A->S = amrAlloc(A->owned, A->nState);
        amrBuildTiles(*A);
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
    case 3:
    {
        long long n = 0;
        for (int l = 1; l < amr->nl; ++l) n += amr->lv[l].ntile;
        return n;
    }
    default: return -1;
    }
}

int hc_amr4_set_tiles(hc_amr4* amr, const int* level0_tile, const int* refined_max)
{
    if (!amr) return hc::fail("hc_amr4_set_tiles: null argument");
    for (int d = 0; d < 3; ++d)
    {
        if (level0_tile && level0_tile[d] < 4) return hc::fail("hc_amr4_set_tiles: level-0 tiles need >= 4 cells per direction");
        if (refined_max && refined_max[d] < 1) return hc::fail("hc_amr4_set_tiles: refined tile extents must be positive");
    }
    try
    {
        HC_THROW(hipDeviceSynchronize());
        for (int d = 0; d < 3; ++d)
        {
            if (level0_tile) amr->tile0[d] = level0_tile[d];
            if (refined_max) amr->fineMax[d] = refined_max[d];
        }
        amrBuildTiles(*amr);
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_amr4_set_tiles: ") + e.what()); }
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
        HC_THROW(hipMemcpyAsync(USph, amr->US, amr->us0*sizeof(double), hipMemcpyDeviceToDevice, s));
        for (int l = 1; l < amr->nl; ++l)
        {
            const AmrLevelView& v = amr->lv[l];
            const long long per = (long long)(v.b[0]+10)*(v.b[1]+10)*(v.b[2]+10);
            hipLaunchKernelGGL(k4a_patch_us, amrGrid(per*v.npatch), dim3(kB), 0, s, v, (const double*)amr->US,
                               (const long long*)amr->patchUOff[l], USph);
        }
        HC_THROW(hipGetLastError());
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
        amrStage(*amr, inner_ghosts, s);
        HC_THROW(hipMemcpyAsync(K, amr->K, amr->nState*sizeof(double), hipMemcpyDeviceToDevice, s));
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
                amrStage(*amr, inner_ghosts, s);
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
