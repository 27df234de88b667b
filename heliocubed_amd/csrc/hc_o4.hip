// Fourth-order operators of the HelioCubed MHD RHS on MI355X, as a batched
// box operator: every box of a box table (uneven sizes, any mix of refinement
// levels) is processed by the same launches, each box owning a run of blocks
// of a flattened grid, so a whole AMR hierarchy or a uniform shell is one
// launch per kernel per RK stage.
//
// Restates (src/ = HelioCubed reference repository) the fourth-order path of
// BoxOp_EulerCubedSphere::operator() (src/BoxOp_EulerCubedSphereImplem.H:97-402):
// sphToPrim (src/BoxOp_Utilities.H:599-707), Limiters_4O with flattening
// (src/BoxOp_Limiters.H:28-359), repairRoeFaceStates and primToMappedRoeFlux
// (src/BoxOp_RoeCubedSphere.H), primToFlux (src/BoxOp_EulerCubedSphereImplem.H:735-802),
// primToMappedFlux (:404-534), factoredRoe8WaveDissipation
// (src/MHD_Roe8Wave.H:634-740) and the flux divergence (:336-381).
// Face products use <ab> = a b + (1/12) sum_t D_t a D_t b with
// D_t = (shift+1 - shift-1)/2 in the two directions transverse to the face.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/heliocubed.h"
#include "hc_common.hpp"
#include "hc_geometry.hpp"
#include "hc_physics.hpp"

namespace
{
constexpr int kG = 5;          // ghost width of the fourth-order operator
constexpr int kB = 256;

// Per-box scratch fields, in units of the box+5 cell count e0*e1*e2.
enum Field : int
{
    F_WB = 0, F_MEB = 8, F_W = 9, F_ETA = 17, F_PTL = 18, F_PTR = 26, F_P0L = 34, F_P0R = 35,
    F_MEL = 36, F_MER = 37, F_PHL = 38, F_PHR = 46, F_DPT = 54, F_FLUX = 62,
    F_FB = 86, F_MEF = 94, F_COUNT = 95,
    // the standard operator stores the sums of both sides' point fluxes
    // (F_PHL + F_PHR) and face-average total pressures (F_P0L + F_P0R)
    F_PHS = F_PHL, F_P0S = F_P0L,
    // Rusanov (riemann_solver 1): the speed of a face, and the conserved point
    // values of both sides (the dissipation and face-average flux slots)
    F_SPD = F_MER, F_QL = F_DPT, F_QR = F_FB,
    // face-average total pressure pBar + meBar of the faces of the direction in
    // progress; reuses the cell meBar slot, which is dead once k_o4_eta has run
    F_P0B = F_MEB,
    // Powell source (standard operator, after the direction passes): the radial
    // area products of <det A B_n> of the three directions, -J div B, and the
    // point values and cell averages of the coefficient components 1..7 (slot
    // + component); reuses the per-direction face slots 18..37
    F_GB = 18, F_MJD = 21, F_CP = 22, F_CA = 30,
    // the Berta operator forms its Powell face products inside each direction
    // pass, so they need slots that outlive the passes: three extra fields
    F_GBB = F_COUNT, F_COUNT_BERTA_POWELL = F_COUNT + 3
};
static_assert(F_GB > F_W + 7 && F_CA + 7 < F_FLUX, "Powell slots overlap the primitives or the face fluxes");

// One box of the table. U (ghosted native conserved state) and R (RHS) are
// addressed through strides so a box can view either its own dense BoxData
// or its tile of a level-wide array.
struct BoxDesc
{
    int p, lo[3], hi[3], n[3];
    long long uOff, rOff, sOff;     // element offsets of U(lo-5), R(lo), scratch
    long long us[3], rs[3];         // (y, z, component) strides of U and R
    int nr, na, naE;                // cells per panel of the box's level; angular table width
    int radOff;                     // offset of the level's radial table
    long long angOff;               // entry offset of the panel's angular tables at the box's level
    long long block0;               // first block of the box in the flattened launch
};

// A physical seam face seen from two boxes: scratch indices of the flux
// component 0 in each box, their component strides, and the sign relating
// the two face orientations.
struct SeamPair
{
    long long a, ea, b, eb;
    double sign;
};

// Face tiles of the Berta face-average and flux kernel of direction D: T0 x
// T1 x T2 faces (x0 unit-stride), the transverse halo of the face averages
// (1; the point fluxes they are formed from need 2), and the LDS regions.
// The angular tiles are one face deep along their normal and 32 x 8 across
// it, so that the averaged region grown by the transverse halo is 1.33x the
// tile (32 x 2 x 4 tiles: 1.59x).
template <int D>
struct FluxTile
{
    static constexpr int T0 = D == 0 ? 8 : 32, T1 = D == 0 ? 8 : D == 1 ? 1 : 8, T2 = D == 0 ? 4 : D == 1 ? 8 : 1;
    static constexpr int h0 = D == 0 ? 0 : 1, h1 = D == 1 ? 0 : 1, h2 = D == 2 ? 0 : 1;
    static constexpr int F0 = T0+4*h0, F1 = T1+4*h1, F2 = T2+4*h2;
    static constexpr int A0 = T0+2*h0, A1 = T1+2*h1, A2 = T2+2*h2;
    static constexpr int NF = F0*F1*F2, NA = A0*A1*A2, NM = A1*A2, NT = T0*T1*T2;
};
constexpr int kFT = 256;   // threads of a flux tile
constexpr int kFPT = 1;    // faces per thread
static_assert(FluxTile<0>::NT == kFT*kFPT && FluxTile<1>::NT == kFT*kFPT && FluxTile<2>::NT == kFT*kFPT,
              "a flux tile holds kFPT faces per thread");
// Only one face per thread is validated: with two (8x8x8 and 32x1x16 tiles)
// the low-beta cone's mixed RK4 increment was off by O(1), and the kernels
// were slower (2 waves per SIMD at 178-180 VGPRs).
static_assert(kFPT == 1, "the face-average and flux kernels are validated with one face per thread");

// Cell tiles of the Berta point-primitive kernel: 32 x 4 x 2 cells of box+4,
// the box's <U> staged on the tile grown by 1.
struct CellTile
{
    static constexpr int T0 = 32, T1 = 4, T2 = 2;
    static constexpr int A0 = T0+2, A1 = T1+2, A2 = T2+2, NA = A0*A1*A2, NT = T0*T1*T2;
};
static_assert(CellTile::NT == 256, "one cell per thread of a 256-thread tile");

// One face tile (or cell tile): its box and first face (cell).
struct FluxTileRef
{
    int box, o0, o1, o2;
};

struct O4View
{
    int nr, na, naE, limiter;
    double gamma, entropy;
    int op, fallback;        // 1 standard operator, 2 Berta WENOZ; Berta fallback 0 WENO3, 1 PLM
    int roe;                 // 1 Roe 8-wave dissipation, 0 Rusanov
    int powell;              // divergence_cleaning 1: Powell source (standard operator)
    double etaD;             // Berta Jameson detector threshold
    const BoxDesc* boxes;
    const double* cellA;     // [6][naE][naE][9]
    const double* faceA[3];  // [6][naE][naE][9]
    const double* fmet[3];   // [6][naE][naE][13]: detA, det(A) A, cofactor row
    const double* radial;    // per level [nr+13][4]: adjDr_0 (area0), adjDr_t (areaT), dVolr, Dr_0 dxi_r for i = -6 ..
    const int* blockBox;     // box of each block of the flattened launch
    int blocksPerBox;        // > 0 when every box owns the same number of blocks (no table lookup)
    unsigned nblocks;        // blocks of the flattened launch
    const SeamPair* seams;
    int nseams;
    const FluxTileRef* ftiles[3];   // face tiles of each direction (Berta face-average and flux kernel)
    unsigned nftiles[3];
    const FluxTileRef* ctiles;      // cell tiles of box+4 (Berta point primitives)
    unsigned nctiles;
    unsigned* dflag;                // [box][3]: the box has faces of the direction deferred to k_b_roe64
    int nboxes;                     // boxes of the table (entries of dflag / 3)
};

// Entry of the angular tables (cellA, faceA, fmet) for cell / face (j, k) of the box's panel and level.
__device__ __forceinline__ long long angIdx(const BoxDesc& b, int j, int k)
{
    return b.angOff + (long long)(k+kG)*b.naE + (j+kG);
}

__device__ __forceinline__ double areaOf(const O4View& v, const BoxDesc& b, int d, int i)
{
    return v.radial[b.radOff + 4*(i+6) + (d == 0 ? 0 : 1)];
}

__device__ __forceinline__ double radialPoint(const O4View& v, const BoxDesc& b, int d, int i)
{
    if (d == 0) return areaOf(v, b, 0, i);
    return areaOf(v, b, d, i) - (1.0/24.0)*(areaOf(v, b, d, i+1) - 2.0*areaOf(v, b, d, i) + areaOf(v, b, d, i-1));
}

// Box of this block and the thread's index in the box's launch region.
// XCD-aware block order: the hardware deals consecutive workgroups round-robin
// to the 8 XCDs, each with its own L2. Logical block L of the flattened launch
// is run by hardware block (L - start(x)) * 8 + x of XCD x, where XCD x owns
// the contiguous logical range starting at start(x); neighbouring tiles, which
// share stencil data, then run on the same XCD and hit in its L2.
constexpr unsigned kXcd = 8;
__device__ __forceinline__ unsigned xcdOrder(unsigned b, unsigned n)
{
    const unsigned x = b % kXcd, q = n / kXcd, r = n % kXcd;
    return x*q + (x < r ? x : r) + b / kXcd;
}
__device__ __forceinline__ unsigned logicalBlock(const O4View& v) { return xcdOrder(blockIdx.x, v.nblocks); }
__device__ __forceinline__ int blockBoxIndex(const O4View& v)
{
    const unsigned L = logicalBlock(v);
    return v.blocksPerBox > 0 ? (int)(L / (unsigned)v.blocksPerBox) : v.blockBox[L];
}
__device__ __forceinline__ const BoxDesc& blockBoxDesc(const O4View& v) { return v.boxes[blockBoxIndex(v)]; }
// (32-bit: a box's launch region, even a whole-panel tile, has far fewer than
// 2^32 points)
__device__ __forceinline__ unsigned boxThread(const O4View& v, const BoxDesc& b)
{
    return (logicalBlock(v) - (unsigned)b.block0)*blockDim.x + threadIdx.x;
}

// Offset (in doubles, 32-bit) of field f at cell (i,j,k) inside a box's
// scratch block, which starts at S + b.sOff; buildBoxop checks that every box's
// fields fit in 32-bit offsets.
struct BoxIdx
{
    int o0, o1, o2, e0, e1, ecount;   // o_d = lo_d - 5: the box grown by the ghost width
    __device__ BoxIdx(const BoxDesc& b)
        : o0(b.lo[0]-kG), o1(b.lo[1]-kG), o2(b.lo[2]-kG), e0(b.n[0]+2*kG), e1(b.n[1]+2*kG),
          ecount((b.n[0]+2*kG)*(b.n[1]+2*kG)*(b.n[2]+2*kG)) {}
    __device__ int operator()(int f, int i, int j, int k) const
    {
        return f*ecount + ((k-o2)*e1 + (j-o1))*e0 + (i-o0);
    }
};

// Region iteration: thread t of the box's region [lo-gl, hi+gh] (+ext along d).
__device__ __forceinline__ bool regionPoint(const BoxDesc& b, const int L[3], const int H[3], unsigned t, int x[3])
{
    const unsigned n0 = H[0]-L[0]+1, n1 = H[1]-L[1]+1, n2 = H[2]-L[2]+1;
    if (t >= n0*n1*n2) return false;
    const unsigned q = t / n0;
    x[0] = L[0] + (int)(t - q*n0);
    x[1] = L[1] + (int)(q % n1);
    x[2] = L[2] + (int)(q / n1);
    return true;
}

// Region iteration with x2 before x1: consecutive blocks step through the x2
// planes of an x0 row, so a stencil along x2 is read by blocks that run
// together on one XCD (its L2 does not hold the planes of a whole box).
__device__ __forceinline__ bool regionPointX2(const BoxDesc& b, const int L[3], const int H[3], unsigned t, int x[3])
{
    const unsigned n0 = H[0]-L[0]+1, n1 = H[1]-L[1]+1, n2 = H[2]-L[2]+1;
    if (t >= n0*n1*n2) return false;
    const unsigned q = t / n0;
    x[0] = L[0] + (int)(t - q*n0);
    x[2] = L[2] + (int)(q % n2);
    x[1] = L[1] + (int)(q / n2);
    return true;
}

// FAST (mixed precision only): constant divisions as products with the
// reciprocal (within an ulp; the FP64 form keeps the divisions of the CPU
// restatement).
template <bool FAST = false>
__device__ void energiesNative(const double* A, const double w[8], double& ke, double& me, double& vb)
{
    double vc[3] = {0, 0, 0}, bc[3] = {0, 0, 0};
    ke = 0; me = 0; vb = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
#pragma unroll
        for (int b = 0; b < 3; ++b) { vc[a] += A[3*a+b]*w[1+b]; bc[a] += A[3*a+b]*w[5+b]; }
        ke += vc[a]*vc[a];
        me += FAST ? bc[a]*bc[a]*hc::kInv8Pi : bc[a]*bc[a]/8.0/hc::kPi;
        vb += vc[a]*bc[a];
    }
    ke *= w[0]*.5;
}

__device__ __forceinline__ void sphToPrimPoint(double w[8], const double u[8])
{
    w[0] = fmax(u[0], hc::kRhoFloor);
    w[1] = u[1]/w[0]; w[2] = u[2]/w[0]; w[3] = u[3]/w[0];
    w[4] = u[4]; w[5] = u[5]; w[6] = u[6]; w[7] = u[7];
}

__device__ __forceinline__ bool usablePrimitive(const double w[8])
{
    bool f = true;
#pragma unroll
    for (int c = 0; c < 8; ++c) f = f && isfinite(w[c]);
    return f && w[0] >= hc::kRhoFloor && w[4] >= 0.0;
}

template <bool FAST = false>
__device__ void primFlux(double f[8], const double w[8], double ke, double me, double vb, double gamma, int d)
{
    const int t1 = (d+1)%3, t2 = (d+2)%3;
    const double vn = w[d+1], vt1 = w[t1+1], vt2 = w[t2+1];
    const double bn = w[d+5], bt1 = w[t1+5], bt2 = w[t2+5];
    auto q4pi = [](double x) { return FAST ? x*hc::kInv4Pi : x/4.0/hc::kPi; };
    f[0] = w[0]*vn;
    f[d+1] = f[0]*vn - q4pi(bn*bn);
    f[t1+1] = f[0]*vt1 - q4pi(bn*bt1);
    f[t2+1] = f[0]*vt2 - q4pi(bn*bt2);
    const double enth = FAST ? w[4]*gamma*(1.0/(gamma-1.0)) : w[4]*gamma/(gamma-1.0);
    f[4] = (2.0*me + ke + enth)*vn - q4pi(bn*vb);
    f[5] = vn*w[5] - w[1]*bn;
    f[6] = vn*w[6] - w[2]*bn;
    f[7] = vn*w[7] - w[3]*bn;
}

// Extremum-preserving limiter of one component of one cell,
// src/BoxOp_Limiters.H:28-148 (La: left state of the cell's high face,
// Rr: right state of its low face).
__device__ void limitComponent(double& La, double& Rr, double W, double Wa, double Wa2, double Wb, double Wb2,
                               double d2, double d2a, double d2b, double d3, double d3b, double d3a, double d3a2)
{
    const double rhs = 1.0-1.0e-12;
    const double dfm = W-Rr, dfp = La-W;
    const double d2f = 6.0*(Rr-2.0*W+La);
    double d2lim;
    if (d2b >= 0.0 && d2 >= 0.0 && d2a >= 0.0 && d2f >= 0.0)
        d2lim = fmin(fmin(fabs(d2f), 1.25*fabs(d2b)), fmin(1.25*fabs(d2), 1.25*fabs(d2a)));
    else if (d2b < 0.0 && d2 < 0.0 && d2a < 0.0 && d2f < 0.0)
        d2lim = -1.0*fmin(fmin(fabs(d2f), 1.25*fabs(d2b)), fmin(1.25*fabs(d2), 1.25*fabs(d2a)));
    else d2lim = 0.0;
    const double rhsRho = 1.0e-12*fmax(fmax(fmax(fabs(Wb2), fabs(Wb)), fmax(fabs(W), fabs(Wa))), fabs(Wa2));
    const double rho = fabs(d2f) <= rhsRho ? 0.0 : d2lim/d2f;
    const double d3min = fmin(fmin(d3b, d3), fmin(d3a, d3a2)), d3max = fmax(fmax(d3b, d3), fmax(d3a, d3a2));
    if ((dfm*dfp <= 0.0) || ((W-Wb2)*(Wa2-W) <= 0.0))
    {
        if (rho < rhs)
        {
            const double rt = d3max-d3min, lt = 0.1*fmax(fabs(d3min), fabs(d3max));
            if (lt <= rt)
            {
                double r2 = 2.0*fabs(dfm), l2 = fabs(dfp);
                if (dfm*dfp < 0.0) La = W+rho*dfp;
                else if (l2 >= r2) La = W+2.0*(1.0-rho)*dfm+rho*dfp;
                r2 = 2.0*fabs(dfp); l2 = fabs(dfm);
                if (dfm*dfp < 0.0) Rr = W-rho*dfm;
                else if (l2 >= r2) Rr = W-2.0*(1.0-rho)*dfp-rho*dfm;
            }
        }
    }
    else
    {
        double r2 = 2.0*fabs(dfm), l2 = fabs(dfp);
        if (l2 >= r2) La = W+2.0*dfm;
        r2 = 2.0*fabs(dfp); l2 = fabs(dfm);
        if (l2 >= r2) Rr = W-2.0*dfp;
    }
}

// Flattening indicator of one direction, src/BoxOp_Limiters.H:151-188.
__device__ double etaTilde(const double* pb2, const double* pb, const double* pa, const double* pa2,
                           double vb1, double va1)
{
    const double z1 = 0.85, z0 = 0.75, delta = 0.33;
    // (lhs1 > 0 && lhs2 > delta) ? si : 1, with the divisions of lhs2 and si
    // evaluated only where they decide the result
    const double lhs1 = vb1-va1;
    if (!(lhs1 > 0.0)) return 1.0;
    const double p_a = pa[0], p_b = pb[0], p_a2 = pa2[0], p_b2 = pb2[0];
    const double lhs2 = fabs(p_a-p_b)/fmin(p_a, p_b);
    if (!(lhs2 > delta)) return 1.0;
    double si = 0;
    const double arg = fabs(p_a-p_b)/fabs(p_a2-p_b2);
    if (arg > z1) si = 0.0;
    if (arg < z0) si = 1.0;
    if ((arg <= z1) && (arg >= z0)) si = 1.0-(arg-z0)/(z1-z0);
    return si;
}

// The face basis A and the orthonormal Roe frame of direction dir (rows: the
// unit normal, the unit first tangent, their cross product) with the
// normal's length, src/MHD_Roe8Wave.H:638-661. FAST (mixed precision)
// normalizes each row by the product with its reciprocal length.
template <bool FAST = false>
__device__ __forceinline__ void roeFrame(const double* Am, int dir, double A[3][3], double frame[3][3], double& nm)
{
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) A[a][b] = Am[3*a+b];
    const int t1 = (dir+1)%3, t2 = (dir+2)%3;
    double normal2 = 0, tangent2 = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        const int j = (i+1)%3, k = (i+2)%3;
        frame[0][i] = A[j][t1]*A[k][t2]-A[k][t1]*A[j][t2];
        frame[1][i] = A[i][t1];
        normal2 += frame[0][i]*frame[0][i];
        tangent2 += frame[1][i]*frame[1][i];
    }
    // correctly rounded lengths in both precisions: the deferred FP64 faces
    // of the mixed path rotate their states in this frame, and their fan
    // admissibility tests can sit at round-off (near-zero thermal energy)
    nm = sqrt(normal2);
    const double tm = sqrt(tangent2);
    if constexpr (FAST)
    {
        const double inm = 1.0/nm, itm = 1.0/tm;
#pragma unroll
        for (int i = 0; i < 3; ++i) { frame[0][i] *= inm; frame[1][i] *= itm; }
    }
    else
    {
#pragma unroll
        for (int i = 0; i < 3; ++i) { frame[0][i] /= nm; frame[1][i] /= tm; }
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) { const int j = (i+1)%3, k = (i+2)%3; frame[2][i] = frame[0][j]*frame[1][k]-frame[0][k]*frame[1][j]; }
}

// Mapped and factored Roe dissipation at a face point, src/MHD_Roe8Wave.H:634-740.
// The basis and the frame are formed again after the Riemann solver (the same
// operations, so the same values) instead of being held in registers across it.
// With DEFER, a face that the mixed path evaluates in FP64 is left to the
// caller: the function then returns true and writes nothing.
template <typename S, bool DEFER = false>
__device__ bool factoredRoe(double out[8], const double wl[8], const double wr[8], const double* Am,
                            double radialArea, int dir, double gamma, double efix)
{
    // mixed precision: the frame's normalization and the back-transform take
    // reciprocal products (ulp-level changes next to the FP32 eigen-system)
    constexpr bool FAST = !std::is_same<S, double>::value;
    double A[3][3], frame[3][3], nm;
    roeFrame<FAST>(Am, dir, A, frame, nm);
    double loc[2][8];
#pragma unroll
    for (int s = 0; s < 2; ++s)
    {
        const double* w = s == 0 ? wl : wr;
#pragma unroll
        for (int c = 0; c < 8; ++c) loc[s][c] = 0;
        loc[s][0] = fmax(w[0], hc::kRhoFloor);
        loc[s][4] = fmax(w[4], 0.0);
        double vc[3] = {0, 0, 0}, bc[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) { vc[i] += A[i][j]*w[1+j]; bc[i] += A[i][j]*w[5+j]; }
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int i = 0; i < 3; ++i) { loc[s][1+d] += frame[d][i]*vc[i]; loc[s][5+d] += frame[d][i]*bc[i]; }
    }
    // the Rusanov fallback of an inadmissible Roe fan is FP64 in both
    // precisions: such faces carry the largest jumps
    double diss[8];
    const int rs = hc::roe8WaveDissipation<S, DEFER>(diss, loc[0], loc[1], gamma, efix);
    if (DEFER && rs == 2) return true;
    if (rs == 0) hc::rusanovDissipation<double>(diss, loc[0], loc[1], gamma);
    // the basis is read again and the frame recomputed: nothing of the
    // frame stays live across the solver
    asm volatile("" ::: "memory");
    roeFrame<FAST>(Am, dir, A, frame, nm);
    const double area = radialArea*nm;
    double mapped[8];
    mapped[0] = area*diss[0]; mapped[4] = area*diss[4];
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        mapped[1+i] = 0; mapped[5+i] = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) { mapped[1+i] += area*frame[d][i]*diss[1+d]; mapped[5+i] += area*frame[d][i]*diss[5+d]; }
    }
    double adj[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const int u1 = (d+1)%3, u2 = (d+2)%3;
#pragma unroll
        for (int i = 0; i < 3; ++i) { const int j = (i+1)%3, k = (i+2)%3; adj[d][i] = A[j][u1]*A[k][u2]-A[k][u1]*A[j][u2]; }
    }
    const double det = A[0][0]*adj[0][0]+A[1][0]*adj[0][1]+A[2][0]*adj[0][2];
    if constexpr (FAST)
    {
        const double idet = hc::quotN(1.0, det), idet2 = idet*idet;
        out[0] = mapped[0]*idet; out[4] = mapped[4]*idet;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            out[1+d] = 0; out[5+d] = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i) { out[1+d] += adj[d][i]*mapped[1+i]*idet2; out[5+d] += adj[d][i]*mapped[5+i]*idet2; }
        }
        return false;
    }
    out[0] = mapped[0]/det; out[4] = mapped[4]/det;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        out[1+d] = 0; out[5+d] = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) { out[1+d] += adj[d][i]*mapped[1+i]/(det*det); out[5+d] += adj[d][i]*mapped[5+i]/(det*det); }
    }
    return false;
}

// Rusanov speed |w_d| + c_f |A e_d| of a face from its two states,
// primToMappedRusanovFlux (src/BoxOp_EulerCubedSphereImplem.H:692-731); the
// Berta variant (src/BoxOp_BertaCubedSphereImplem.H:1653-1725) clamps the
// radicands at 0 and groups 4 pi rho. B_n is taken along the first column of
// A, as the reference does.
__device__ double rusanovSpeedO4(const double L[8], const double R[8], const double* Am, double gamma, int d,
                                 bool berta)
{
    double rho = 0.5*(L[0]+R[0]);
    const double w_d = 0.5*(L[1+d]+R[1+d]);
    const double bs[3] = {0.5*(L[5]+R[5]), 0.5*(L[6]+R[6]), 0.5*(L[7]+R[7])};
    double p = 0.5*(L[4]+R[4]);
    if (p < 0.0) p = 0.0;
    rho = rho < hc::kRhoFloor ? hc::kRhoFloor : rho;
    double bc[3] = {0.0, 0.0, 0.0}, Bn = 0.0, Bm = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
#pragma unroll
        for (int b = 0; b < 3; ++b) bc[a] += Am[3*a+b]*bs[b];
        Bn += Am[3*a]*bc[a];
        Bm += bc[a]*bc[a];
    }
    Bm = sqrt(Bm);
    const double ce = sqrt(gamma*p/rho);
    double af;
    if (berta)
    {
        const double aa = ce*ce + Bm*Bm/(4.0*hc::kPi*rho);
        const double bb = fabs(Bn)*ce/sqrt(hc::kPi*rho);
        const double s1 = aa + bb, s2 = aa - bb;
        af = 0.5*(sqrt(0.0 < s1 ? s1 : 0.0) + sqrt(0.0 < s2 ? s2 : 0.0));
    }
    else
    {
        const double a = ce*ce + Bm*Bm/4.0/hc::kPi/rho;
        const double b = fabs(Bn)*ce/sqrt(hc::kPi*rho);
        af = 0.5*(sqrt(a + b) + sqrt(a - b));
    }
    const double Ad = sqrt(Am[d]*Am[d] + Am[3+d]*Am[3+d] + Am[6+d]*Am[6+d]);
    return fabs(w_d) + af*Ad;
}

// ---- kernels ------------------------------------------------------------------
// 1. Cell-average primitives WBar and their magnetic energy on box+5.
__global__ void __launch_bounds__(kB) k_o4_wbar(O4View v, const double* __restrict__ U, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    const unsigned t = boxThread(v, b);
    if (t == 0)
    {
        // the box's flags of deferred FP64 Roe faces, for this stage's passes
        const int n = blockBoxIndex(v);
        v.dflag[3*n] = 0u; v.dflag[3*n+1] = 0u; v.dflag[3*n+2] = 0u;
    }
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-kG; H[d] = b.hi[d]+kG; }
    if (!regionPoint(b, L, H, t, x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    const long long u = b.uOff + (x[2]-L[2])*b.us[1] + (x[1]-L[1])*b.us[0] + (x[0]-L[0]);
    double uu[8], w[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) uu[c] = U[u + c*b.us[2]];
    sphToPrimPoint(w, uu);
    double ke, me, vb;
    energiesNative(v.cellA + 9*angIdx(b, x[1], x[2]), w, ke, me, vb);
    w[4] = (w[4]-ke-me)*(v.gamma-1.0);
    if (w[4] < 0) w[4] = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) P[I(F_WB+c, x[0], x[1], x[2])] = w[c];
    P[I(F_MEB, x[0], x[1], x[2])] = me;
}

// 2. Point primitives W on box+4 with the complete-state fallbacks,
// src/BoxOp_Utilities.H:619-696.
__global__ void __launch_bounds__(kB) k_o4_w(O4View v, const double* __restrict__ U, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-4; H[d] = b.hi[d]+4; }
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    auto ui = [&](int i, int j, int k) {
        return b.uOff + (k-b.lo[2]+kG)*b.us[1] + (j-b.lo[1]+kG)*b.us[0] + (i-b.lo[0]+kG); };
    double up[8], uc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        const long long o = c*b.us[2];
        const double lap = U[ui(x[0]+1,x[1],x[2])+o]+U[ui(x[0]-1,x[1],x[2])+o]+U[ui(x[0],x[1]+1,x[2])+o]
                          +U[ui(x[0],x[1]-1,x[2])+o]+U[ui(x[0],x[1],x[2]+1)+o]+U[ui(x[0],x[1],x[2]-1)+o]
                          -6.0*U[ui(x[0],x[1],x[2])+o];
        uc[c] = U[ui(x[0],x[1],x[2])+o];
        up[c] = uc[c]-lap/24.0;
    }
    if (!(isfinite(up[0]) && up[0] >= hc::kRhoFloor) && isfinite(uc[0]) && uc[0] >= hc::kRhoFloor)
#pragma unroll
        for (int c = 0; c < 8; ++c) up[c] = uc[c];
    double w[8];
    sphToPrimPoint(w, up);
    double ke, me, vb;
    energiesNative(v.cellA + 9*angIdx(b, x[1], x[2]), w, ke, me, vb);
    w[4] = (w[4]-ke-me)*(v.gamma-1.0);
    double wb[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) wb[c] = P[I(F_WB+c, x[0], x[1], x[2])];
    if (!usablePrimitive(w) && usablePrimitive(wb))
#pragma unroll
        for (int c = 0; c < 8; ++c) w[c] = wb[c];
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        const int f = F_WB+c;
        const double lap = P[I(f,x[0]+1,x[1],x[2])]+P[I(f,x[0]-1,x[1],x[2])]+P[I(f,x[0],x[1]+1,x[2])]
                          +P[I(f,x[0],x[1]-1,x[2])]+P[I(f,x[0],x[1],x[2]+1)]+P[I(f,x[0],x[1],x[2]-1)]
                          -6.0*P[I(f,x[0],x[1],x[2])];
        P[I(F_W+c, x[0], x[1], x[2])] = w[c]+lap/24.0;
    }
}

// 3. Flattening coefficient eta on box+2: minimum over the three directions
// of the three-cell minimum of eta_tilde, src/BoxOp_Limiters.H:304-334.
__global__ void __launch_bounds__(kB) k_o4_eta(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-2; H[d] = b.hi[d]+2; }
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    double e = 0;
#pragma unroll
    for (int dd = 0; dd < 3; ++dd)
    {
        // total pressure and velocity on the line x + q e_dd, q = -3..3; the
        // five-point window of the cell x + oo e_dd starts at index oo+1
        double pr[7], vel[7];
#pragma unroll
        for (int q = -3; q <= 3; ++q)
        {
            int z[3] = {x[0], x[1], x[2]};
            z[dd] += q;
            pr[q+3] = P[I(F_WB+4, z[0], z[1], z[2])] + P[I(F_MEB, z[0], z[1], z[2])];
            vel[q+3] = P[I(F_WB+1+dd, z[0], z[1], z[2])];
        }
        double ed = 0;
#pragma unroll
        for (int o = 0; o < 3; ++o)
        {
            const int s0 = (o == 0 ? 0 : (o == 1 ? 1 : -1)) + 1;
            const double et = etaTilde(&pr[s0], &pr[s0+1], &pr[s0+3], &pr[s0+4], vel[s0+1], vel[s0+3]);
            ed = o == 0 ? et : fmin(ed, et);
        }
        e = dd == 0 ? ed : fmin(ed, e);
    }
    P[I(F_ETA, x[0], x[1], x[2])] = e;
}

// Face-average state WFaceBar at face x (direction d): (WBar_{x-e} + WBar_x)/2.
__device__ __forceinline__ void faceBarAvg(const BoxIdx& I, const double* P, int d, int i, int j, int k, double wf[8])
{
    const int a = i-(d == 0), b = j-(d == 1), c = k-(d == 2);
#pragma unroll
    for (int q = 0; q < 8; ++q) wf[q] = 0.5*P[I(F_WB+q, a, b, c)] + 0.5*P[I(F_WB+q, i, j, k)];
}

// Width of cell y along D for the minmod limiter (limiter 2,
// src/BoxOp_EulerCubedSphereImplem.H:220-240): Dr_0 dxi_r radially; along an
// angular direction dVolr det A dxi_perp / adjDr_t / |adjA| with the face
// metric of index y (the reference evaluates it on the face box).
template <int D>
__device__ __forceinline__ double minmodWidth(const O4View& v, const BoxDesc& b, int y0, int y1, int y2)
{
    const double* rr = v.radial + b.radOff + 4*(y0+6);
    if (D == 0) return rr[3];
    const double* m = v.fmet[D] + 13*angIdx(b, y1, y2);
    const double adjA = sqrt(m[10]*m[10] + m[11]*m[11] + m[12]*m[12]);
    return rr[2]*m[0]*(1.0/b.na)/rr[1]/adjA;
}

// Minmod face states of face (i,j,k) of direction D, Limiters_minmod /
// minmod_calc (src/BoxOp_Limiters.H:360-446): slope dx minmod(dW+ / (dx +
// dx_ahead), dW- / (dx + dx_behind)) of the point values W, left state from
// the cell below the face (W + slope), right state from the cell above
// (W - slope), pressures clipped at 0.
template <int D>
__device__ void minmodFaceStates(const O4View& v, const BoxDesc& b, const BoxIdx& I, const double* P, int i, int j, int k,
                                 double l[8], double r[8])
{
    double dx[4];
#pragma unroll
    for (int o = -2; o <= 1; ++o)
        dx[o+2] = minmodWidth<D>(v, b, i+(D == 0)*o, j+(D == 1)*o, k+(D == 2)*o);
    auto Wv = [&](int c, int o) { return P[I(F_W+c, i+(D == 0)*o, j+(D == 1)*o, k+(D == 2)*o)]; };
    // slope of the cell at offset m (-1: below the face, 0: above)
    auto slope = [&](int c, int m, double wm) {
        const double c1 = 1.0/(dx[m+2] + dx[m+3]), c2 = 1.0/(dx[m+2] + dx[m+1]);
        const double dWR = Wv(c, m+1) - wm, dWL = wm - Wv(c, m-1);
        const double A = dWR*c1, B = dWL*c2;
        double mm;
        if (A*B < 0) mm = 0;
        else if (fabs(A) < fabs(B)) mm = A;
        else mm = B;
        return dx[m+2]*mm; };
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        const double wl = Wv(c, -1), wr = Wv(c, 0);
        l[c] = wl + slope(c, -1, wl);
        r[c] = wr - slope(c, 0, wr);
    }
    if (l[4] < 0) l[4] = 0.;
    if (r[4] < 0) r[4] = 0.;
}

// Face-average state WFaceBar of face (i,j,k): the second-order average of the
// two cell averages, or with limiter 2 the mean of the two minmod states
// (src/BoxOp_EulerCubedSphereImplem.H:242-245).
template <int D>
__device__ __forceinline__ void faceBar(const O4View& v, const BoxDesc& b, const BoxIdx& I, const double* P,
                                        int i, int j, int k, double wf[8])
{
    if (v.limiter != 2) { faceBarAvg(I, P, D, i, j, k, wf); return; }
    double l[8], r[8];
    minmodFaceStates<D>(v, b, I, P, i, j, k, l, r);
#pragma unroll
    for (int q = 0; q < 8; ++q) wf[q] = (l[q] + r[q])*0.5;
}

// Transverse extent of the face states and point fluxes of direction D: the
// flux kernel's mapped products and their radial differences take one
// transverse neighbour on each side of the box's faces, so the point fluxes are
// needed on [lo-1, hi+1] transversally.
template <int D>
__device__ __forceinline__ void pointFaceRegion(const BoxDesc& b, int grow, int L[3], int H[3])
{
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-1-grow; H[d] = b.hi[d]+1+grow; }
    L[D] = b.lo[D]; H[D] = b.hi[D]+1;
}

// 5. Limited, flattened and repaired face states; point states by transverse
// deconvolution; the sums of both sides' physical fluxes and face-average
// total pressures, on the point-flux faces (pointFaceRegion). Runs after
// k_o4_bar.
// Mixed precision (DEFER): a face whose Roe eigen-system must be FP64 (plasma
// beta below HC_MIXED_FP64_BETA) leaves its two point states in F_PTL / F_PTR
// (free during a Roe direction pass) and kO4Deferred in F_DPT, and sets its
// box's flag; k_o4_roe64 evaluates it. A quiet NaN whose payload no
// arithmetic produces.
constexpr unsigned long long kO4Deferred = 0x7ffdefe4a0d5e0a4ULL;
template <int D, bool ROE, typename S_, bool DEFER>
__device__ __forceinline__ void o4Faces(const O4View& v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    pointFaceRegion<D>(b, 0, L, H);
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    constexpr int t1 = D == 0 ? 1 : 0, t2 = D == 2 ? 1 : 2;
    // mixed precision: the point fluxes take reciprocal products (FAST); the
    // point states stay operation for operation with the FP64 path, so the Roe
    // repair decides on the same values in both precisions (the Roe solver's
    // own decisions are taken on its face-frame and fan states, which carry
    // the mixed path's reciprocal and FP32 roundings)
    constexpr bool FAST = !std::is_same<S_, double>::value;
    auto by24 = [](double x) { return x/24.0; };
    auto by24f = [](double x) { return FAST ? x*(1.0/24.0) : x/24.0; };
    auto Wat = [&](int c, int o) {
        int y[3] = {x[0], x[1], x[2]}; y[D] += o;
        return P[I(F_W+c, y[0], y[1], y[2])]; };
    double lf[8], rf[8];
    if (v.limiter == 2) minmodFaceStates<D>(v, b, I, P, x[0], x[1], x[2], lf, rf);
#pragma unroll 1
    for (int c = 0; c < 8 && v.limiter != 2; ++c)
    {
        double w[8];   // W at offsets -4..+3 from the face's right cell
#pragma unroll
        for (int o = -4; o <= 3; ++o) w[o+4] = Wat(c, o);
        // unlimited 4th-order face values at faces x-1, x, x+1 (index relative to x)
        auto c2f = [&](int f) { return (7.0/12.0)*(w[f-1+4]+w[f+4]) - (1.0/12.0)*(w[f-2+4]+w[f+1+4]); };
        double left = c2f(0), right = c2f(0);
        if (v.limiter == 1)
        {
            // cell x-1 (its high face is x) and cell x (its low face is x)
            for (int side = 0; side < 2; ++side)
            {
                const int cc = side == 0 ? -1 : 0;   // cell offset
                double La = c2f(cc+1), Rr = c2f(cc);
                auto ww = [&](int o) { return w[cc+o+4]; };
                auto del2 = [&](int o) { return ww(o-1)-2.0*ww(o)+ww(o+1); };
                auto del3 = [&](int o) { return del2(o)-del2(o-1); };
                limitComponent(La, Rr, ww(0), ww(1), ww(2), ww(-1), ww(-2),
                               del2(0), del2(1), del2(-1), del3(0), del3(-1), del3(1), del3(2));
                int y[3] = {x[0], x[1], x[2]}; y[D] += cc;
                const double et = P[I(F_ETA, y[0], y[1], y[2])];
                const double wc = ww(0);
                if (side == 0) left = et*La+(1.-et)*wc;
                else right = et*Rr+(1.-et)*wc;
            }
        }
        lf[c] = left; rf[c] = right;
    }
    // Roe repair with the cell-average donors, src/BoxOp_RoeCubedSphere.H:5-27,
    // then point values by transverse deconvolution.
    double corr[8], lapb[8];
    {
        double wfx[8], a[8], bq[8];
        faceBar<D>(v, b, I, P, x[0], x[1], x[2], wfx);
#pragma unroll
        for (int q = 0; q < 8; ++q) lapb[q] = 0.0;
        for (int t = 0; t < 2; ++t)
        {
            const int tt = t == 0 ? t1 : t2;
            int y[3] = {x[0], x[1], x[2]}, z[3] = {x[0], x[1], x[2]};
            y[tt] += 1; z[tt] -= 1;
            faceBar<D>(v, b, I, P, y[0], y[1], y[2], a);
            faceBar<D>(v, b, I, P, z[0], z[1], z[2], bq);
#pragma unroll
            for (int q = 0; q < 8; ++q) lapb[q] += a[q] - 2.0*wfx[q] + bq[q];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) corr[q] = lapb[q]*(1.0/24.0);
    }
    if constexpr (ROE)
    {   // the Roe face-state repair (riemann_solver 2 only)
        double pl[8], pr[8], cl[8], cr[8];
        int m1[3] = {x[0], x[1], x[2]}; m1[D] -= 1;
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            pl[q] = lf[q]-by24(lapb[q]);
            pr[q] = rf[q]-by24(lapb[q]);
            cl[q] = P[I(F_WB+q, m1[0], m1[1], m1[2])];
            cr[q] = P[I(F_WB+q, x[0], x[1], x[2])];
        }
        if (!(usablePrimitive(pl) && usablePrimitive(pr)) && usablePrimitive(cl) && usablePrimitive(cr))
#pragma unroll
            for (int q = 0; q < 8; ++q) { lf[q] = cl[q]+corr[q]; rf[q] = cr[q]+corr[q]; }
    }
    if constexpr (!ROE)
    {
#pragma unroll
        for (int q = 0; q < 8; ++q)
        {
            P[I(F_PTL+q, x[0], x[1], x[2])] = lf[q]-by24(lapb[q]);
            P[I(F_PTR+q, x[0], x[1], x[2])] = rf[q]-by24(lapb[q]);
        }
        P[I(F_SPD, x[0], x[1], x[2])] = rusanovSpeedO4(lf, rf, v.faceA[D] + 9*angIdx(b, x[1], x[2]),
                                                       v.gamma, D, false);
    }
    // Physical fluxes of both point states (point flux + transverse Laplacian
    // of the face-average flux / 24, k_o4_bar having run first) and the
    // face-average total pressure p + me + Lap(meBar)/24.
    double lapF[8], lapMe = 0;
    {
        const double mex = P[I(F_MEF, x[0], x[1], x[2])];
#pragma unroll
        for (int q = 0; q < 8; ++q) lapF[q] = 0;
#pragma unroll
        for (int t = 0; t < 2; ++t)
        {
            const int tt = t == 0 ? t1 : t2;
            int y[3] = {x[0], x[1], x[2]}, z[3] = {x[0], x[1], x[2]};
            y[tt] += 1; z[tt] -= 1;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                lapF[q] += P[I(F_FB+q, y[0], y[1], y[2])] - 2.0*P[I(F_FB+q, x[0], x[1], x[2])]
                         + P[I(F_FB+q, z[0], z[1], z[2])];
            lapMe += P[I(F_MEF, y[0], y[1], y[2])] - 2.0*mex + P[I(F_MEF, z[0], z[1], z[2])];
        }
    }
    const double* A = v.faceA[D] + 9*angIdx(b, x[1], x[2]);
    double phs[8], p0s = 0;
#pragma unroll 1
    for (int side = 0; side < 2; ++side)
    {
        double w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) w[q] = (side == 0 ? lf[q] : rf[q]) - by24(lapb[q]);
        double ke, me, vb, f[8];
        energiesNative<FAST>(A, w, ke, me, vb);
        primFlux<FAST>(f, w, ke, me, vb, v.gamma, D);
        const double p0 = (side == 0 ? lf[4] : rf[4]) + (me + by24f(lapMe));
#pragma unroll
        for (int q = 0; q < 8; ++q) phs[q] = side == 0 ? f[q]+by24f(lapF[q]) : phs[q] + (f[q]+by24f(lapF[q]));
        p0s = side == 0 ? p0 : p0s + p0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) P[I(F_PHS+q, x[0], x[1], x[2])] = phs[q];
    P[I(F_P0S, x[0], x[1], x[2])] = p0s;
    if constexpr (ROE)
    {
        // point Roe dissipation of the two point states (the flux kernel
        // averages it over the transverse neighbours)
        double wl[8], wr[8], out[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) { wl[q] = lf[q]-by24(lapb[q]); wr[q] = rf[q]-by24(lapb[q]); }
        if (factoredRoe<S_, DEFER>(out, wl, wr, A, radialPoint(v, b, D, x[0]), D, v.gamma, v.entropy))
        {
#pragma unroll
            for (int q = 0; q < 8; ++q) { P[I(F_PTL+q, x[0], x[1], x[2])] = wl[q]; P[I(F_PTR+q, x[0], x[1], x[2])] = wr[q]; }
            P[I(F_DPT, x[0], x[1], x[2])] = __longlong_as_double((long long)kO4Deferred);
            v.dflag[3*blockBoxIndex(v) + D] = 1u;
            return;
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) P[I(F_DPT+q, x[0], x[1], x[2])] = out[q];
    }
}

// The FP64 build and Rusanov hold their solvers in the face kernel (2 waves
// per SIMD); the mixed Roe build defers the FP64 faces and fits 3 waves.
template <int D, bool ROE, typename S_>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(2))) k_o4_faces(O4View v, double* __restrict__ S)
{
    o4Faces<D, ROE, S_, false>(v, S);
}
template <int D>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(3))) k_o4_facesm(O4View v, double* __restrict__ S)
{
    o4Faces<D, true, float, true>(v, S);
}

// Energies (ke, me, v.B) and the physical flux of the face-average state at
// face (i,j,k) of direction D.
template <int D>
__device__ void barFlux(const O4View& v, const BoxDesc& b, const BoxIdx& I, const double* P,
                        int i, int j, int k, double f[8], double& meBar, double& pBar)
{
    double wf[8];
    faceBar<D>(v, b, I, P, i, j, k, wf);
    double ke, me, vb;
    energiesNative(v.faceA[D] + 9*angIdx(b, j, k), wf, ke, me, vb);
    primFlux(f, wf, ke, me, vb, v.gamma, D);
    meBar = me;
    pBar = wf[4];
}

// 4. Physical flux of the face-average state WFaceBar (F_FB), its magnetic
// energy (F_MEF) and total pressure pBar + meBar (F_P0B) on the point-flux
// faces grown by one transversally (their Laplacian in k_o4_faces). Needs only
// the cell averages, so it runs first in each direction pass.
template <int D>
__global__ void __launch_bounds__(kB) k_o4_bar(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    pointFaceRegion<D>(b, 1, L, H);
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    double f[8], me, p;
    barFlux<D>(v, b, I, P, x[0], x[1], x[2], f, me, p);
#pragma unroll
    for (int q = 0; q < 8; ++q) P[I(F_FB+q, x[0], x[1], x[2])] = f[q];
    P[I(F_MEF, x[0], x[1], x[2])] = me;
    P[I(F_P0B, x[0], x[1], x[2])] = p + me;
}

// 5c'. Rusanov (riemann_solver 1): conserved point values of both face
// states times the radial point area, primToMappedRusanovFlux
// (src/BoxOp_EulerCubedSphereImplem.H:586-640). Faces [lo_d, hi_d+1] x
// transverse [lo-1, hi+1].
template <int D>
__global__ void __launch_bounds__(kB) k_o4_rusq(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-1; H[d] = b.hi[d]+1; }
    L[D] = b.lo[D]; H[D] = b.hi[D]+1;
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    const double* A = v.faceA[D] + 9*angIdx(b, x[1], x[2]);
    const double rp = radialPoint(v, b, D, x[0]);
#pragma unroll 1
    for (int side = 0; side < 2; ++side)
    {
        double w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) w[q] = P[I((side == 0 ? F_PTL : F_PTR)+q, x[0], x[1], x[2])];
        double ke, me, vb;
        energiesNative(A, w, ke, me, vb);
        const int f = side == 0 ? F_QL : F_QR;
        P[I(f, x[0], x[1], x[2])] = rp*w[0];
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            P[I(f+1+a, x[0], x[1], x[2])] = rp*w[0]*w[1+a];
            P[I(f+5+a, x[0], x[1], x[2])] = rp*w[5+a];
        }
        P[I(f+4, x[0], x[1], x[2])] = rp*(ke + me + w[4]/(v.gamma-1.0));
    }
}

// 6a. Seam flux reconciliation 1/2 (F_a + sign F_b) on both boxes' copies
// of each physical seam face, src/MHD_AMR.H:1484-1496.
__global__ void __launch_bounds__(kB) k_o4_seams(O4View v, double* __restrict__ S)
{
    const int t = blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= v.nseams) return;
    const SeamPair sp = v.seams[t];
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        const double fa = S[sp.a + c*sp.ea], fb = S[sp.b + c*sp.eb];
        S[sp.a + c*sp.ea] = 0.5*(fa + sp.sign*fb);
        S[sp.b + c*sp.eb] = 0.5*(fb + sp.sign*fa);
    }
}

// 6b. Powell source, standard operator (src/BoxOp_DivBcleaning.H:31-168,
// src/BoxOp_EulerCubedSphereImplem.H:152-155,203-211,385-390). On the cells
// [lo-2, hi+2]: the point values of the coefficient (B/4pi, v, v.B/4pi) in
// Cartesian components from the deconvolved primitives W - Lap(W)/24 and the
// cell-centre basis (F_CP + 1..7); on the faces [lo_d-1, hi_d+2] x [lo-1, hi+1]
// of each direction d: G_d = <adjDr_d <det A B_d>>, with B_d the unlimited
// fourth-order face value of the native field component d and <ab> the face
// product rule. The metric does not vary with the radius and the radial area
// not across a panel: those differences are exactly zero and their product
// terms (+-0) are left out.
__device__ __forceinline__ double faceValueW(const BoxIdx& I, const double* P, int c, int d, const int y[3])
{
    auto W = [&](int o) { int z[3] = {y[0], y[1], y[2]}; z[d] += o; return P[I(F_W+c, z[0], z[1], z[2])]; };
    return (7.0/12.0)*(W(-1)+W(0)) - (1.0/12.0)*(W(-2)+W(1));
}

template <int D>
__device__ double powellFaceProduct(const O4View& v, const BoxDesc& b, const BoxIdx& I, const double* P, const int x[3])
{
    constexpr int t1 = D == 0 ? 1 : 0, t2 = D == 2 ? 1 : 2;
    const double* met = v.fmet[D];
    // <det A B_D> at face y: derivative terms along the angular transverse
    // directions, then det A B_D
    auto detAB = [&](const int y[3]) {
        int a[3] = {y[0], y[1], y[2]}, z[3] = {y[0], y[1], y[2]};
        double val = 0;
        if (D == 0)
        {
            a[t1] += 1; z[t1] -= 1;
            const double dm = 0.5*met[13*angIdx(b, a[1], a[2])] - 0.5*met[13*angIdx(b, z[1], z[2])];
            const double db = 0.5*faceValueW(I, P, 5+D, D, a) - 0.5*faceValueW(I, P, 5+D, D, z);
            val += dm*db*(1.0/12.0);
            a[t1] -= 1; z[t1] += 1;
        }
        a[t2] += 1; z[t2] -= 1;
        const double dm = 0.5*met[13*angIdx(b, a[1], a[2])] - 0.5*met[13*angIdx(b, z[1], z[2])];
        const double db = 0.5*faceValueW(I, P, 5+D, D, a) - 0.5*faceValueW(I, P, 5+D, D, z);
        val += dm*db*(1.0/12.0);
        val += met[13*angIdx(b, y[1], y[2])]*faceValueW(I, P, 5+D, D, y);
        return val;
    };
    const double adjX = areaOf(v, b, D, x[0]);
    double val = 0;
    if (D != 0)
    {
        // radial transverse direction (t1): the radial area varies with it
        const int a[3] = {x[0]+1, x[1], x[2]}, z[3] = {x[0]-1, x[1], x[2]};
        const double dadj = 0.5*areaOf(v, b, D, x[0]+1) - 0.5*areaOf(v, b, D, x[0]-1);
        val += dadj*(0.5*detAB(a) - 0.5*detAB(z))*(1.0/12.0);
    }
    val += adjX*detAB(x);
    return val;
}

// Point value of the coefficient (B/4pi, v, v.B/4pi) at cell x from the
// primitives in slots F_W..: deconvolved W - Lap(W)/24 (standard operator,
// cell averages) or taken as they are (Berta, point values).
template <bool DECONV>
__device__ void powellCoefficient(const O4View& v, const BoxDesc& b, const BoxIdx& I, double* P, const int x[3])
{
    double w[8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        const int f = F_W+c;
        if (DECONV)
        {
            const double lap = P[I(f,x[0]+1,x[1],x[2])]+P[I(f,x[0]-1,x[1],x[2])]+P[I(f,x[0],x[1]+1,x[2])]
                              +P[I(f,x[0],x[1]-1,x[2])]+P[I(f,x[0],x[1],x[2]+1)]+P[I(f,x[0],x[1],x[2]-1)]
                              -6.0*P[I(f,x[0],x[1],x[2])];
            w[c] = P[I(f,x[0],x[1],x[2])]-lap/24.0;
        }
        else w[c] = P[I(f,x[0],x[1],x[2])];
    }
    const double* A = v.cellA + 9*angIdx(b, x[1], x[2]);
    double vel[3] = {0, 0, 0}, mag[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) { vel[i] += A[3*i+j]*w[1+j]; mag[i] += A[3*i+j]*w[5+j]; }
    double ce = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        P[I(F_CP+1+i, x[0], x[1], x[2])] = mag[i]/(4.0*hc::kPi);
        P[I(F_CP+5+i, x[0], x[1], x[2])] = vel[i];
        ce += vel[i]*mag[i]/(4.0*hc::kPi);
    }
    P[I(F_CP+4, x[0], x[1], x[2])] = ce;
}

__global__ void __launch_bounds__(kB) k_o4_pw_face(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-2; H[d] = b.hi[d]+2; }
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    powellCoefficient<true>(v, b, I, P, x);
    auto inFaces = [&](int d) {
        for (int a = 0; a < 3; ++a)
        {
            const int lo = b.lo[a]-1, hi = a == d ? b.hi[a]+2 : b.hi[a]+1;
            if (x[a] < lo || x[a] > hi) return false;
        }
        return true;
    };
    if (inFaces(0)) P[I(F_GB+0, x[0], x[1], x[2])] = powellFaceProduct<0>(v, b, I, P, x);
    if (inFaces(1)) P[I(F_GB+1, x[0], x[1], x[2])] = powellFaceProduct<1>(v, b, I, P, x);
    if (inFaces(2)) P[I(F_GB+2, x[0], x[1], x[2])] = powellFaceProduct<2>(v, b, I, P, x);
}

// 6c. -J div B = sum_d -(1/dxi_d) (G_d(x+e_d) - G_d(x)) and the cell
// averages C + Lap(C)/24 of the coefficient, on the cells [lo-1, hi+1]; G_d
// in slot gBase + d.
__global__ void __launch_bounds__(kB) k_o4_pw_cell(O4View v, double* __restrict__ S, int gBase)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-1; H[d] = b.hi[d]+1; }
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    double md = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        const double dxInv = -1.0/(d == 0 ? 1.0/b.nr : 1.0/b.na);
        int a[3] = {x[0], x[1], x[2]}; a[d] += 1;
        double inc = 0;
        inc += (-1.0*dxInv)*P[I(gBase+d, x[0], x[1], x[2])];
        inc += (1.0*dxInv)*P[I(gBase+d, a[0], a[1], a[2])];
        md += inc;
    }
    P[I(F_MJD, x[0], x[1], x[2])] = md;
#pragma unroll
    for (int c = 1; c < 8; ++c)
    {
        const int f = F_CP+c;
        const double lap = P[I(f,x[0]+1,x[1],x[2])]+P[I(f,x[0]-1,x[1],x[2])]+P[I(f,x[0],x[1]+1,x[2])]
                          +P[I(f,x[0],x[1]-1,x[2])]+P[I(f,x[0],x[1],x[2]+1)]+P[I(f,x[0],x[1],x[2]-1)]
                          -6.0*P[I(f,x[0],x[1],x[2])];
        P[I(F_CA+c, x[0], x[1], x[2])] = P[I(f,x[0],x[1],x[2])]+lap/24.0;
    }
}

// 7. Flux divergence into the box RHS: accumulated (overwrite = 0, the
// BoxOp convention; the caller zeroes it) or overwritten; optionally the RK
// stage accumulate U += b_dt * RHS with U in the RHS layout. With the Powell
// source on, the cell product <(-J div B) C_c> = sum_d D_d(-J div B) D_d C_c / 12
// + (-J div B) C_c is added to components 1..7 after the divergence.
__global__ void __launch_bounds__(kB) k_o4_div(O4View v, const double* __restrict__ S, double* __restrict__ R,
                                               int overwrite, double* __restrict__ Uacc, double b_dt, double scale)
{
    const BoxDesc& b = blockBoxDesc(v);
    int x[3];
    if (!regionPoint(b, b.lo, b.hi, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    const double* const P = S + b.sOff;
    const long long r = b.rOff + (x[2]-b.lo[2])*b.rs[1] + (x[1]-b.lo[1])*b.rs[0] + (x[0]-b.lo[0]);
    for (int c = 0; c < 8; ++c)
    {
        const long long n = c*b.rs[2];
        double acc = overwrite ? 0.0 : R[r + n];
        for (int d = 0; d < 3; ++d)
        {
            const double dxInv = -scale/(d == 0 ? 1.0/b.nr : 1.0/b.na);
            int a[3] = {x[0], x[1], x[2]}; a[d] += 1;
            double inc = 0;
            inc += (-1.0*dxInv)*P[I(F_FLUX+8*d+c, x[0], x[1], x[2])];
            inc += (1.0*dxInv)*P[I(F_FLUX+8*d+c, a[0], a[1], a[2])];
            acc += inc;
        }
        if (v.powell && c > 0)
        {
            double val = 0;
            for (int d = 0; d < 3; ++d)
            {
                int a[3] = {x[0], x[1], x[2]}, z[3] = {x[0], x[1], x[2]};
                a[d] += 1; z[d] -= 1;
                const double ds = 0.5*P[I(F_MJD, a[0], a[1], a[2])] - 0.5*P[I(F_MJD, z[0], z[1], z[2])];
                const double dc = 0.5*P[I(F_CA+c, a[0], a[1], a[2])] - 0.5*P[I(F_CA+c, z[0], z[1], z[2])];
                val += ds*dc*(1.0/12.0);
            }
            val += P[I(F_MJD, x[0], x[1], x[2])]*P[I(F_CA+c, x[0], x[1], x[2])];
            acc += val;
        }
        R[r + n] = acc;
        if (Uacc) Uacc[r + n] += b_dt*acc;
    }
}

// 8. Copy the face fluxes of the valid box extruded in d out of scratch.
__global__ void __launch_bounds__(kB) k_o4_fluxout(O4View v, int D, const double* __restrict__ S,
                                                   double* __restrict__ F, const long long* __restrict__ fOff)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3] = {b.lo[0], b.lo[1], b.lo[2]}, H[3] = {b.hi[0], b.hi[1], b.hi[2]}, x[3];
    H[D] += 1;
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    const double* const P = S + b.sOff;
    const int m0 = b.n[0]+(D == 0), m1 = b.n[1]+(D == 1), m2 = b.n[2]+(D == 2);
    const long long n = (long long)m0*m1*m2;
    const long long o = fOff[blockBoxIndex(v)] + ((long long)(x[2]-b.lo[2])*m1 + (x[1]-b.lo[1]))*m0 + (x[0]-b.lo[0]);
    for (int c = 0; c < 8; ++c) F[o + c*n] = P[I(F_FLUX+8*D+c, x[0], x[1], x[2])];
}

template <typename T>
T* upload(const std::vector<T>& h, std::vector<void*>& owned)
{
    T* p = nullptr;
    HC_THROW(hipMalloc(&p, std::max<size_t>(1, h.size())*sizeof(T)));
    owned.push_back(p);
    if (!h.empty()) HC_THROW(hipMemcpy(p, h.data(), h.size()*sizeof(T), hipMemcpyHostToDevice));
    return p;
}
} // namespace

struct hc_boxop
{
    hc_options opt{};
    hc::ShellDims dims{};                 // level 0 (the only level of a uniform table)
    std::vector<hc::ShellDims> levels;    // every level of the table
    int nbox = 0;
    long long totalBlocks = 0;            // blocks of the flattened launch over all boxes
    std::vector<BoxDesc> boxes;
    O4View view{};
    double* scratch = nullptr;
    long long* fOff[3] = {nullptr, nullptr, nullptr};
    std::vector<long long> fOffHost[3];
    long long uCount = 0, rCount = 0, fCount[3] = {0, 0, 0};
    int maxRegion = 0;
    std::vector<void*> owned;
    ~hc_boxop() { for (void* p : owned) (void)hipFree(p); }
};

namespace
{
// Box table over one or more levels: per-level angular and radial tables
// (concatenated), one descriptor per box with its level's sizes and table
// offsets, scratch offsets, and a flattened launch in which box b owns
// ceil((n+10)^3 / kB) consecutive blocks (uneven boxes waste no blocks).
void buildBoxop(hc_boxop& B, const hc_box* boxes, const int* boxLevel, int nbox)
{
    if (B.levels.empty()) B.levels.push_back(B.dims);
    O4View& v = B.view;
    v.nr = B.levels[0].nr; v.na = B.levels[0].na; v.naE = v.na + 2*kG + 1; v.limiter = B.opt.limiter;
    v.gamma = B.opt.gamma; v.entropy = B.opt.entropy_fix;
    v.op = B.opt.spatial_operator == 2 ? 2 : 1;
    v.roe = B.opt.riemann_solver == 2;
    v.powell = B.opt.divergence_cleaning == 1;
    v.fallback = B.opt.berta_fallback;
    v.etaD = B.opt.berta_eta_d;
    std::vector<double> cellA, faceA[3], fmet[3], radial;
    std::vector<long long> levAng;      // entry offset of each level's angular tables
    std::vector<int> levRad;            // offset of each level's radial table
    for (const hc::ShellDims& L : B.levels)
    {
        const int na = L.na, nr = L.nr, naE = na + 2*kG + 1;
        const size_t base = cellA.size()/9;
        levAng.push_back((long long)base);
        cellA.resize(9*(base + 6ull*naE*naE));
        for (int d = 0; d < 3; ++d) { faceA[d].resize(9*(base + 6ull*naE*naE)); fmet[d].resize(13*(base + 6ull*naE*naE)); }
        const double h = 1.0/na;
        for (int p = 0; p < 6; ++p)
            for (int kk = 0; kk < naE; ++kk)
                for (int jj = 0; jj < naE; ++jj)
                {
                    const int j = jj-kG, k = kk-kG;
                    const size_t x = base + ((size_t)p*naE+kk)*naE+jj;
                    double A[3][3];
                    hc::pointBasis(p, hc::xiToAngle((j+0.5)*h), hc::xiToAngle((k+0.5)*h), A);
                    for (int a = 0; a < 9; ++a) cellA[9*x+a] = A[a/3][a%3];
                    for (int d = 0; d < 3; ++d)
                    {
                        const double oa = d == 1 ? 0.0 : 0.5, ob = d == 2 ? 0.0 : 0.5;
                        hc::pointBasis(p, hc::xiToAngle((j+oa)*h), hc::xiToAngle((k+ob)*h), A);
                        for (int a = 0; a < 9; ++a) faceA[d][9*x+a] = A[a/3][a%3];
                        hc::faceMetricAverages(na, p, d, j, k, &fmet[d][13*x]);
                    }
                }
        levRad.push_back((int)radial.size());
        for (int i = -6; i <= nr+6; ++i)
        {
            const hc::RadialCell c = hc::radialCell(L, i);
            const double dxi = 1.0/nr;
            radial.push_back(c.area0); radial.push_back(c.areaT);
            radial.push_back(c.dvol); radial.push_back((c.rhi - c.rlo)/dxi*dxi);
        }
    }
    v.cellA = upload(cellA, B.owned);
    for (int d = 0; d < 3; ++d) { v.faceA[d] = upload(faceA[d], B.owned); v.fmet[d] = upload(fmet[d], B.owned); }
    v.radial = upload(radial, B.owned);
    long long u = 0, r = 0, sc = 0, blk = 0;
    for (int d = 0; d < 3; ++d) B.fOffHost[d].clear();
    long long f[3] = {0, 0, 0};
    const long long nfields = v.powell && v.op == 2 ? F_COUNT_BERTA_POWELL : F_COUNT;
    std::vector<int> blockBox;
    for (int n = 0; n < nbox; ++n)
    {
        BoxDesc bd{};
        const int lev = boxLevel ? boxLevel[n] : 0;
        const hc::ShellDims& L = B.levels[lev];
        bd.p = boxes[n].block;
        for (int d = 0; d < 3; ++d) { bd.lo[d] = boxes[n].lo[d]; bd.hi[d] = boxes[n].hi[d]; bd.n[d] = bd.hi[d]-bd.lo[d]+1; }
        bd.nr = L.nr; bd.na = L.na; bd.naE = L.na + 2*kG + 1;
        bd.radOff = levRad[lev];
        bd.angOff = levAng[lev] + (long long)bd.p*bd.naE*bd.naE;
        const long long e = (long long)(bd.n[0]+2*kG)*(bd.n[1]+2*kG)*(bd.n[2]+2*kG);
        if (nfields*e > 0x7fffffffLL) throw std::runtime_error("box too large: scratch offsets exceed 32 bits");
        bd.uOff = u; bd.rOff = r; bd.sOff = sc;
        bd.us[0] = bd.n[0]+2*kG; bd.us[1] = (long long)(bd.n[0]+2*kG)*(bd.n[1]+2*kG); bd.us[2] = e;
        bd.rs[0] = bd.n[0]; bd.rs[1] = (long long)bd.n[0]*bd.n[1]; bd.rs[2] = (long long)bd.n[0]*bd.n[1]*bd.n[2];
        u += 8*e; r += 8LL*bd.n[0]*bd.n[1]*bd.n[2]; sc += nfields*e;
        for (int d = 0; d < 3; ++d)
        {
            B.fOffHost[d].push_back(f[d]);
            f[d] += 8LL*(bd.n[0]+(d == 0))*(bd.n[1]+(d == 1))*(bd.n[2]+(d == 2));
        }
        bd.block0 = blk;
        const long long nb = (e + kB - 1)/kB;
        for (long long q = 0; q < nb; ++q) blockBox.push_back(n);
        blk += nb;
        B.maxRegion = std::max<long long>(B.maxRegion, e);
        B.boxes.push_back(bd);
    }
    if (blk > 0x7fffffffLL) throw std::runtime_error("box table too large for one launch");
    // equal block counts for every box: the box of a block is blockIdx.x / blocksPerBox
    v.blocksPerBox = (nbox > 0 && blk % nbox == 0) ? (int)(blk/nbox) : 0;
    for (int n = 0; n < nbox && v.blocksPerBox > 0; ++n)
        if (B.boxes[n].block0 != (long long)n*v.blocksPerBox) v.blocksPerBox = 0;
    B.totalBlocks = blk;
    v.nblocks = (unsigned)blk;
    B.uCount = u; B.rCount = r;
    for (int d = 0; d < 3; ++d) { B.fCount[d] = f[d]; B.fOff[d] = upload(B.fOffHost[d], B.owned); }
    v.boxes = upload(B.boxes, B.owned);
    v.blockBox = upload(blockBox, B.owned);
    v.dflag = upload(std::vector<unsigned>((size_t)3*nbox, 0u), B.owned);
    v.nboxes = nbox;
    // face tiles of each direction, box by box, x0 fastest; then x2 for the
    // r and alpha directions (whose stencils reach two planes across the
    // x2 tiles), x1 for beta
    for (int d = 0; d < 3; ++d)
    {
        const int T[3] = {d == 0 ? FluxTile<0>::T0 : d == 1 ? FluxTile<1>::T0 : FluxTile<2>::T0,
                          d == 0 ? FluxTile<0>::T1 : d == 1 ? FluxTile<1>::T1 : FluxTile<2>::T1,
                          d == 0 ? FluxTile<0>::T2 : d == 1 ? FluxTile<1>::T2 : FluxTile<2>::T2};
        std::vector<FluxTileRef> tiles;
        for (int n = 0; n < nbox; ++n)
        {
            const BoxDesc& bd = B.boxes[n];
            const int m0 = bd.n[0]+(d == 0), m1 = bd.n[1]+(d == 1), m2 = bd.n[2]+(d == 2);
            if (d == 2)
            {
                for (int k = 0; k < m2; k += T[2])
                    for (int j = 0; j < m1; j += T[1])
                        for (int i = 0; i < m0; i += T[0])
                            tiles.push_back({n, bd.lo[0]+i, bd.lo[1]+j, bd.lo[2]+k});
            }
            else
            {
                for (int j = 0; j < m1; j += T[1])
                    for (int k = 0; k < m2; k += T[2])
                        for (int i = 0; i < m0; i += T[0])
                            tiles.push_back({n, bd.lo[0]+i, bd.lo[1]+j, bd.lo[2]+k});
            }
        }
        if (tiles.size() > 0x7fffffffULL) throw std::runtime_error("box table too large for one launch");
        v.nftiles[d] = (unsigned)tiles.size();
        v.ftiles[d] = upload(tiles, B.owned);
    }
    // cell tiles of box+4, box by box, x0 fastest, then x2
    {
        std::vector<FluxTileRef> tiles;
        for (int n = 0; n < nbox; ++n)
        {
            const BoxDesc& bd = B.boxes[n];
            for (int j = 0; j < bd.n[1]+8; j += CellTile::T1)
                for (int k = 0; k < bd.n[2]+8; k += CellTile::T2)
                    for (int i = 0; i < bd.n[0]+8; i += CellTile::T0)
                        tiles.push_back({n, bd.lo[0]-4+i, bd.lo[1]-4+j, bd.lo[2]-4+k});
        }
        if (tiles.size() > 0x7fffffffULL) throw std::runtime_error("box table too large for one launch");
        v.nctiles = (unsigned)tiles.size();
        v.ctiles = upload(tiles, B.owned);
    }
    v.seams = nullptr;
    v.nseams = 0;
    HC_THROW(hipMalloc(&B.scratch, std::max<long long>(1, sc)*sizeof(double)));
    B.owned.push_back(B.scratch);
    B.nbox = nbox;
}

#include "hc_berta.hpp"

// Mixed precision: the point Roe dissipation of the faces k_o4_facesm left
// marked for the FP64 eigen-system, from the point states it stored; the
// point-flux faces of the box. At most kRoeBlocks blocks, each returning
// when no box's flag of the direction is set (k_o4_wbar clears the flags
// at the start of each stage), else stepping through the face kernel's
// blocks with a stride of its grid.
template <int D>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(2))) k_o4_roe64(O4View v, double* __restrict__ S)
{
    int any = 0;
    for (int n = threadIdx.x; n < v.nboxes; n += blockDim.x) any |= v.dflag[3*n + D] != 0u;
    if (!__syncthreads_or(any)) return;
    for (unsigned hb = blockIdx.x; hb < v.nblocks; hb += gridDim.x)
    {
        const unsigned lb = xcdOrder(hb, v.nblocks);
        const int box = v.blocksPerBox > 0 ? (int)(lb / (unsigned)v.blocksPerBox) : v.blockBox[lb];
        if (v.dflag[3*box + D] == 0u) continue;
        const BoxDesc& b = v.boxes[box];
        int L[3], H[3], x[3];
        pointFaceRegion<D>(b, 0, L, H);
        if (!regionPoint(b, L, H, (lb - (unsigned)b.block0)*blockDim.x + threadIdx.x, x)) continue;
        const BoxIdx I(b);
        double* const P = S + b.sOff;
        if ((unsigned long long)__double_as_longlong(P[I(F_DPT, x[0], x[1], x[2])]) != kO4Deferred) continue;
        const double* A = v.faceA[D] + 9*angIdx(b, x[1], x[2]);
        double wl[8], wr[8], out[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) { wl[q] = P[I(F_PTL+q, x[0], x[1], x[2])]; wr[q] = P[I(F_PTR+q, x[0], x[1], x[2])]; }
        factoredRoe<float>(out, wl, wr, A, radialPoint(v, b, D, x[0]), D, v.gamma, v.entropy);
#pragma unroll
        for (int q = 0; q < 8; ++q) P[I(F_DPT+q, x[0], x[1], x[2])] = out[q];
    }
}

template <int D, typename S_>
void directionPass(const O4View& v, dim3 grid, hipStream_t s, double* scratch)
{
    hipLaunchKernelGGL(k_o4_bar<D>, grid, dim3(kB), 0, s, v, scratch);
    if (!v.roe) hipLaunchKernelGGL((k_o4_faces<D, false, double>), grid, dim3(kB), 0, s, v, scratch);
    else if constexpr (std::is_same<S_, float>::value)
    {
        hipLaunchKernelGGL(k_o4_facesm<D>, grid, dim3(kB), 0, s, v, scratch);
        hipLaunchKernelGGL(k_o4_roe64<D>, dim3(grid.x < kRoeBlocks ? grid.x : kRoeBlocks), dim3(kB), 0, s, v, scratch);
    }
    else hipLaunchKernelGGL((k_o4_faces<D, true, S_>), grid, dim3(kB), 0, s, v, scratch);
    if (!v.roe) hipLaunchKernelGGL(k_o4_rusq<D>, grid, dim3(kB), 0, s, v, scratch);
    if (v.nftiles[D] == 0) return;
    if (v.roe) hipLaunchKernelGGL((k_o4_fluxt<D, true>), dim3(v.nftiles[D]), dim3(kFT), 0, s, v, scratch);
    else hipLaunchKernelGGL((k_o4_fluxt<D, false>), dim3(v.nftiles[D]), dim3(kFT), 0, s, v, scratch);
}

template <typename S_>
void applyBoxop(hc_boxop& B, const double* U, double* R, double* const F[3], hipStream_t s,
                int overwrite = 0, double* Uacc = nullptr, double b_dt = 0.0, double scale = 1.0)
{
    const O4View& v = B.view;
    const dim3 grid((unsigned)B.totalBlocks);
    if (v.op == 2) bertaFluxes<S_>(v, grid, s, U, B.scratch);
    else
    {
        hipLaunchKernelGGL(k_o4_wbar, grid, dim3(kB), 0, s, v, U, B.scratch);
        hipLaunchKernelGGL(k_o4_w, grid, dim3(kB), 0, s, v, U, B.scratch);
        hipLaunchKernelGGL(k_o4_eta, grid, dim3(kB), 0, s, v, B.scratch);
        directionPass<0, S_>(v, grid, s, B.scratch);
        directionPass<1, S_>(v, grid, s, B.scratch);
        directionPass<2, S_>(v, grid, s, B.scratch);
    }
    if (v.nseams > 0)
        hipLaunchKernelGGL(k_o4_seams, dim3((v.nseams + kB - 1)/kB), dim3(kB), 0, s, v, B.scratch);
    if (v.powell)
    {
        if (v.op == 2) hipLaunchKernelGGL(k_b_pw_coef, grid, dim3(kB), 0, s, v, B.scratch);
        else hipLaunchKernelGGL(k_o4_pw_face, grid, dim3(kB), 0, s, v, B.scratch);
        hipLaunchKernelGGL(k_o4_pw_cell, grid, dim3(kB), 0, s, v, B.scratch, v.op == 2 ? (int)F_GBB : (int)F_GB);
    }
    hipLaunchKernelGGL(k_o4_div, grid, dim3(kB), 0, s, v, B.scratch, R, overwrite, Uacc, b_dt, scale);
    for (int d = 0; d < 3; ++d)
        if (F[d]) hipLaunchKernelGGL(k_o4_fluxout, grid, dim3(kB), 0, s, v, d, B.scratch, F[d], B.fOff[d]);
    HC_THROW(hipGetLastError());
}
} // namespace

// ===== Fourth-order level: stage preparation, box table over the level ======
// JU (valid level) -> native conserved averages USph on the ghosted level,
// src/MHD_StateConversion.H:153-316 (state_conversion = 1), with the radial
// boundary ghosts of src/BoxOp_PreStage.H:105-141. Pieces the reference takes
// from Proto are restated: deriv3case (centred difference inside a block,
// 3-point one-sided at its edges), the seam interpolation of Cartesian cell
// averages (cubic average-matching reconstruction on a 4x4 window of the
// panel holding the ghost cell centre, integrated over the ghost cell), and
// the cell matrix quotient q = <A>^-1 (<U> - (1/12) sum_d D_d<A> D_d q0).
namespace
{
constexpr int kGC = 6;   // angular ghost layers of the Cartesian averages

struct Prep4View
{
    int nr, na, NE;               // NE = na + 12 (angular index -6 .. na+5)
    long long vcs, vps;           // valid strides (component, panel)
    long long ccs, cps, cJ;       // Cartesian buffer strides: [6][8][na+12][na+12][nr]
    long long scs, sps, sI, sJ;   // USph strides: [6][8][na+10][na+10][nr+10]
    double gamma;
    const double* det;            // [NE][NE]
    const double* Aavg;           // [6][NE][NE][9]
    const double* Ainv;           // [6][NE][NE][9]
    const double* dvol;           // [nr+13], i = -6 ..
    const double* rmid;           // [nr+13]
};

__device__ __forceinline__ double det4(const Prep4View& v, int j, int k) { return v.det[(long long)(k+6)*v.NE + (j+6)]; }
__device__ __forceinline__ const double* Aavg4(const Prep4View& v, int p, int j, int k)
{ return v.Aavg + 9*(((long long)p*v.NE + (k+6))*v.NE + (j+6)); }
__device__ __forceinline__ const double* Ainv4(const Prep4View& v, int p, int j, int k)
{ return v.Ainv + 9*(((long long)p*v.NE + (k+6))*v.NE + (j+6)); }

template <typename F>
__device__ __forceinline__ double deriv3(F q, int x, int lo, int hi)
{
    if (x == lo) return 0.5*(-3.0*q(x)+4.0*q(x+1)-q(x+2));
    if (x == hi) return 0.5*(3.0*q(x)-4.0*q(x-1)+q(x-2));
    return 0.5*(q(x+1)-q(x-1));
}

// Logical index of a column block under the XCD-aware order of a 1-D grid:
// the hardware deals consecutive blocks round-robin to the 8 XCDs, so XCD x
// runs the contiguous logical range starting at start(x) and neighbouring
// columns (j +- 1, and k +- 1 one row of na columns away) share its L2.
__device__ __forceinline__ unsigned xcdColumn()
{
    const unsigned b = blockIdx.x, n = gridDim.x, x = b % kXcd, q = n / kXcd, r = n % kXcd;
    return x*q + (x < r ? x : r) + b / kXcd;
}

// Column (j, k) of the c-th column of an n x n panel in 4 x 4 patches: patches
// row by row, columns row by row inside a patch (the last patch row and
// column may be narrower). Consecutive blocks then share the columns of their
// 5-point stencils inside a patch; row-major order puts the k +- 1 neighbours
// a whole row (a full XCD L2 of data at 128 columns) apart.
__device__ __forceinline__ void patchColumn(unsigned c, int n, int& j, int& k)
{
    constexpr int H = 4;
    const unsigned full = (unsigned)(n / H);
    unsigned pk = c / (unsigned)(H*n);
    int h = H;
    if (pk >= full) { pk = full; h = n - H*(int)full; }
    const unsigned r = c - pk*(unsigned)(H*n);
    unsigned pj = r / (unsigned)(H*h);
    int w = H;
    if (pj >= full) { pj = full; w = n - H*(int)full; }
    const unsigned in = r - pj*(unsigned)(H*h);
    j = (int)pj*H + (int)(in % (unsigned)w);
    k = (int)pk*H + (int)(in / (unsigned)w);
}

// Angular quotient with the fourth-order correction (the metric terms are
// uniform over a column), then the radial quotient with its correction, of
// one angular column of valid cells: the angular quotients of G components
// along the whole radial line go to LDS, the radial quotients read their
// radial neighbours there and write the Cartesian buffer. One block per
// column in the XCD-aware patch order, threads along the radius; dynamic LDS
// of G*nr doubles. G components at a time put their loads in flight together.
template <int G>
__global__ void __launch_bounds__(kB) k4_angrad(Prep4View v, const double* __restrict__ U0, const double* __restrict__ K,
                                                double a_dt, double* __restrict__ cart)
{
    extern __shared__ double line[];
    const int nr = v.nr, na = v.na;
    const unsigned col = xcdColumn();
    const int p = (int)(col / (unsigned)(na*na));
    int j, k;
    patchColumn(col - (unsigned)p*(unsigned)(na*na), na, j, k);
    const double djd = 0.5*(det4(v, j+1, k)-det4(v, j-1, k));
    const double dkd = 0.5*(det4(v, j, k+1)-det4(v, j, k-1));
    const double idet = 1.0/det4(v, j, k);
    auto dv = [&](int ii) { return v.dvol[ii+6]; };
#pragma unroll 1
    for (int c0 = 0; c0 < 8; c0 += G)
    {
        for (int i = threadIdx.x; i < nr; i += blockDim.x)
        {
#pragma unroll
            for (int g = 0; g < G; ++g)
            {
                const long long base = p*v.vps + (c0+g)*v.vcs;
                auto ju = [&](int jj, int kk) {
                    const long long x = base + ((long long)kk*na + jj)*nr + i;
                    return K ? U0[x] + a_dt*K[x] : U0[x]; };
                auto q1 = [&](int jj) { return ju(jj, k)/det4(v, jj, k); };
                auto q2 = [&](int kk) { return ju(j, kk)/det4(v, j, kk); };
                double corr = 0;
                corr += deriv3(q1, j, 0, na-1)*djd;
                corr += deriv3(q2, k, 0, na-1)*dkd;
                line[g*nr + i] = (ju(j, k)-corr/12.0)*idet;
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nr; i += blockDim.x)
        {
            const double dd = 0.5*(dv(i+1)-dv(i-1));
            const double idv = 1.0/dv(i);
#pragma unroll
            for (int g = 0; g < G; ++g)
            {
                const double* ln = line + g*nr;
                auto rq = [&](int ii) { return ln[ii]/dv(ii); };
                const double dq = deriv3(rq, i, 0, nr-1);
                cart[p*v.cps + (c0+g)*v.ccs + ((long long)(k+kGC)*v.cJ + (j+kGC))*nr + i] = (ln[i]-dq*dd/12.0)*idv;
            }
        }
        __syncthreads();
    }
}

// Seam ghost columns of the Cartesian averages: weighted sums of donor
// columns (valid cells of the panels the ghost cell overlaps). One block per
// ghost column, threads along the radius: the donor table entries are
// uniform over the block. The table runs along the seams, so the XCD-aware
// order keeps the shared donors of neighbouring ghost columns in one L2.
__global__ void __launch_bounds__(kB) k4_seam(Prep4View v, const long long* __restrict__ dst,
                                              const int* __restrict__ start, const long long* __restrict__ src,
                                              const double* __restrict__ w, double* __restrict__ cart)
{
    const int col = (int)xcdColumn();
    const int d0 = start[col], d1 = start[col+1];
    for (int i = threadIdx.x; i < v.nr; i += blockDim.x)
    {
        // donors outer, components inner: each donor's index and weight are
        // read once; every component still sums its donors in table order
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int d = d0; d < d1; ++d)
        {
            const double wd = w[d];
            const double* sp = cart + src[d] + i;
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] += wd*sp[c*v.ccs];
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) cart[dst[col] + c*v.ccs + i] = acc[c];
    }
}

// Native components by the fourth-order matrix quotient on the USph cells
// i in [0, nr), j,k in [-5, na+5). One block per angular column (j+5 =
// blockIdx.x, k and the panel in blockIdx.y), threads along the radius: the
// matrix averages and inverses are uniform over the block.
__global__ void __launch_bounds__(kB) k4_native(Prep4View v, const double* __restrict__ cart, double* __restrict__ US)
{
    const int NA = v.na+10;
    // 1-D grid over the 6 (na+10)^2 columns in the XCD-aware order
    const unsigned col = xcdColumn();
    const int p = (int)(col / (unsigned)(NA*NA));
    int j, k;
    patchColumn(col - (unsigned)p*(unsigned)(NA*NA), NA, j, k);
    j -= 5; k -= 5;
    for (int i = threadIdx.x; i < v.nr; i += blockDim.x)
    {
        auto ca = [&](int c, int jj, int kk) {
            return cart[p*v.cps + c*v.ccs + ((long long)(kk+kGC)*v.cJ + (jj+kGC))*v.nr + i]; };
        auto us = [&](int c) -> double& {
            return US[p*v.sps + c*v.scs + ((long long)(k+5)*v.sJ + (j+5))*v.sI + (i+5)]; };
        us(0) = ca(0, j, k);
        us(4) = ca(4, j, k);
        for (int grp = 0; grp < 2; ++grp)
        {
            const int st = grp == 0 ? 1 : 5;
            auto q0 = [&](int jj, int kk, int row) {
                const double* Mi = Ainv4(v, p, jj, kk);
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
                        const double dA = 0.5*(Aavg4(v, p, j+dj, k+dk)[3*row+cc]-Aavg4(v, p, j-dj, k-dk)[3*row+cc]);
                        const double dq = 0.5*(q0(j+dj, k+dk, cc)-q0(j-dj, k-dk, cc));
                        corr += dA*dq;
                    }
                }
                rhs[row] = ca(st+row, j, k)-corr/12.0;
            }
            const double* Mi = Ainv4(v, p, j, k);
            for (int row = 0; row < 3; ++row)
            {
                double val = 0;
                for (int cc = 0; cc < 3; ++cc) val += Mi[3*row+cc]*rhs[cc];
                us(st+row) = val;
            }
        }
    }
}

// Radial boundary ghosts of USph: inner fixed native state, outer copy of the
// last layer scaled by (r_last/r)^2 (r_last/r for the angular B components).
// One thread per ghost cell (5 inner and 5 outer layers of every column).
__global__ void __launch_bounds__(kB) k4_radial_bc(Prep4View v, const double* __restrict__ inner, double* __restrict__ US)
{
    const int NA = v.na+10;
    const int t = blockIdx.x*blockDim.x + threadIdx.x;
    if (t >= 10*6*NA*NA) return;
    const int layer = t % 10, col = t / 10;
    const int jj = col % NA, kk = (col / NA) % NA, p = col / (NA*NA);
    auto us = [&](int c, int i) -> double& {
        return US[p*v.sps + c*v.scs + ((long long)kk*v.sJ + jj)*v.sI + (i+5)]; };
    if (layer < 5)
    {
        for (int c = 0; c < 8; ++c)
            us(c, -5+layer) = inner[((((long long)p*8+c)*NA+kk)*NA+jj)*5+layer];
        return;
    }
    const int i = v.nr + (layer-5);
    const double rlast = v.rmid[v.nr-1+6];
    const double ratio = rlast/v.rmid[i+6];
    for (int c = 0; c < 6; ++c) us(c, i) = us(c, v.nr-1)*(ratio*ratio);
    for (int c = 6; c < 8; ++c) us(c, i) = us(c, v.nr-1)*ratio;
}
} // namespace

namespace
{
// Radial boundary flux sums of a fourth-order level (radialCons): block b
// sums the radial face fluxes of faces [b chunk, (b+1) chunk) of the table
// per component, + on the inner boundary (the first half of the table), - on
// the outer; each thread takes every blockDim-th face of the chunk, then a
// tree reduction in LDS. Deterministic for a fixed table and launch.
__global__ void __launch_bounds__(kB) k4_radcons_part(const long long* __restrict__ face, const long long* __restrict__ ec,
                                                      int nrad, int chunk, const double* __restrict__ S,
                                                      double* __restrict__ part)
{
    __shared__ double red[kB];
    const int f0 = blockIdx.x*chunk, f1 = min(nrad, f0 + chunk);
    const int half = nrad/2;
    for (int c = 0; c < 8; ++c)
    {
        double acc = 0;
        for (int f = f0 + (int)threadIdx.x; f < f1; f += blockDim.x)
        {
            const double v = S[face[f] + c*ec[f]];
            acc += f < half ? v : -v;
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int w = kB/2; w > 0; w >>= 1)
        {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[blockIdx.x*8 + c] = red[0];
        __syncthreads();
    }
}
// cons[c] += scale * (sum of the block partials in block order).
__global__ void k4_radcons_final(const double* __restrict__ part, int nblk, double scale, double* __restrict__ cons)
{
    const int c = threadIdx.x;
    if (c >= 8) return;
    double sum = 0;
    for (int b = 0; b < nblk; ++b) sum += part[b*8 + c];
    cons[c] += scale*sum;
}
} // namespace

namespace hc
{
struct Level4
{
    hc::ShellDims dims{};
    hc_options opt{};
    Prep4View pv{};
    hc_boxop* box = nullptr;
    int tile[3] = {32, 32, 32};    // box-table tile shape over the level (radial, alpha, beta)
    int ncol = 0;
    long long* colDst = nullptr;
    int* colStart = nullptr;
    long long* donorSrc = nullptr;
    double* donorW = nullptr;
    double* cart = nullptr;
    // radial boundary faces of the box table (inner, then outer): flux offsets
    // in the tiles' scratch and their component strides
    long long* radFace = nullptr;
    long long* radEc = nullptr;
    double* radPart = nullptr;
    int nrad = 0, radBlocks = 0;
    std::vector<void*> owned, radOwned;
    ~Level4()
    {
        delete box;
        for (void* p : radOwned) (void)hipFree(p);
        for (void* p : owned) (void)hipFree(p);
    }
};

constexpr int kRadChunk = 2048;   // radial faces per block of k4_radcons_part

// The radial boundary faces of the level's box table: face i = 0 (inner) and
// i = n_r (outer) of every angular column, read from the tile holding the
// column's first or last radial cell.
static void radialFaceTable(Level4& L)
{
    for (void* p : L.radOwned) (void)hipFree(p);
    L.radOwned.clear();
    const int nr = L.pv.nr, na = L.pv.na;
    const hc_boxop& B = *L.box;
    std::vector<int> inner((size_t)6*na*na, -1), outer((size_t)6*na*na, -1);
    for (size_t n = 0; n < B.boxes.size(); ++n)
    {
        const BoxDesc& b = B.boxes[n];
        for (int k = b.lo[2]; k <= b.hi[2]; ++k) for (int j = b.lo[1]; j <= b.hi[1]; ++j)
        {
            if (b.lo[0] == 0) inner[((size_t)b.p*na + k)*na + j] = (int)n;
            if (b.hi[0] == nr-1) outer[((size_t)b.p*na + k)*na + j] = (int)n;
        }
    }
    std::vector<long long> face, ec;
    for (int side = 0; side < 2; ++side)
        for (int p = 0; p < 6; ++p) for (int k = 0; k < na; ++k) for (int j = 0; j < na; ++j)
        {
            const int n = (side ? outer : inner)[((size_t)p*na + k)*na + j];
            if (n < 0) throw std::runtime_error("radial boundary column without a tile");
            const BoxDesc& b = B.boxes[n];
            const long long e0 = b.n[0]+2*kG, e1 = b.n[1]+2*kG, e2 = b.n[2]+2*kG;
            const int i = side ? nr : 0;
            face.push_back(b.sOff + (long long)F_FLUX*e0*e1*e2
                           + ((long long)(k-b.lo[2]+kG)*e1 + (j-b.lo[1]+kG))*e0 + (i-b.lo[0]+kG));
            ec.push_back(e0*e1*e2);
        }
    L.nrad = (int)face.size();
    L.radBlocks = (L.nrad + kRadChunk - 1)/kRadChunk;
    L.radFace = upload(face, L.radOwned);
    L.radEc = upload(ec, L.radOwned);
    HC_THROW(hipMalloc(&L.radPart, std::max(1, L.radBlocks)*8*sizeof(double)));
    L.radOwned.push_back(L.radPart);
}

// Seam donors of a ghost column of the Cartesian averages. The ghost cell is
// integrated by a 6x6 Gauss-Legendre rule in its own panel's xi; each
// quadrature point takes the cubic average-matching reconstruction (the
// derivative of the quartic interpolant of the primitive function on five
// cell edges) of the panel holding that point, on the 4x4 window around it
// clamped inside the panel. Donors of all points are merged.
struct AvgDonor { int q, j, k; double w; };
static std::vector<AvgDonor> averageDonors(int na, int p, int j, int k)
{
    static std::vector<double> gx, gw;
    if (gx.empty()) gaussLegendre01(6, gx, gw);
    const double h = 1.0/na;
    auto basis = [](int b, double u, int base) {
        double val = 0;
        for (int s = b+1; s <= 4; ++s)
            for (int m = 0; m <= 4; ++m)
            {
                if (m == s) continue;
                double prod = 1.0/double(s-m);
                for (int n = 0; n <= 4; ++n)
                    if (n != s && n != m) prod *= (u-(base+n))/double(s-n);
                val += prod;
            }
        return val;
    };
    std::vector<AvgDonor> out;
    for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b)
        {
            const Vec3 X = panelPoint(p, xiToAngle((j+gx[a])*h), xiToAngle((k+gx[b])*h));
            const int q = closestPanel(X);
            const Vec3 L = toPanelLocal(q, X);
            const double u = (std::atan2(L.x[1], L.x[0])+kPi/4)/(h*kPi/2);
            const double v = (std::atan2(L.x[2], L.x[0])+kPi/4)/(h*kPi/2);
            const int bu = std::max(0, std::min(na-4, (int)std::floor(u)-1));
            const int bv = std::max(0, std::min(na-4, (int)std::floor(v)-1));
            const double wt = gw[a]*gw[b];
            double fu[4], fv[4];
            for (int m = 0; m < 4; ++m) { fu[m] = basis(m, u, bu); fv[m] = basis(m, v, bv); }
            for (int n = 0; n < 4; ++n)
                for (int m = 0; m < 4; ++m)
                {
                    const double w = wt*fu[m]*fv[n];
                    bool found = false;
                    for (auto& d : out)
                        if (d.q == q && d.j == bu+m && d.k == bv+n) { d.w += w; found = true; break; }
                    if (!found) out.push_back({q, bu+m, bv+n, w});
                }
        }
    return out;
}

// Seam pairs between the face fluxes of the first nbox boxes of a table
// (whole-level tiles of one fourth-order level): every seam face of the
// level, its two sides in the scratch of the tiles holding them.
static std::vector<SeamPair> panelSeamPairs(const std::vector<BoxDesc>& boxes, size_t nbox, int nr, int na)
{
    auto boxOf = [&](int p, int i, int j, int k) {
        for (size_t n = 0; n < nbox; ++n)
        {
            const BoxDesc& b = boxes[n];
            if (b.p == p && i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] && k <= b.hi[2])
                return (int)n;
        }
        return -1;
    };
    auto faceIndex = [&](int n, int d, int i, int jf, int kf, long long& ec) {
        const BoxDesc& b = boxes[n];
        const long long e0 = b.n[0]+2*kG, e1 = b.n[1]+2*kG, e2 = b.n[2]+2*kG;
        ec = e0*e1*e2;
        return b.sOff + (F_FLUX+8*d)*ec + ((long long)(kf-b.lo[2]+kG)*e1 + (jf-b.lo[1]+kG))*e0 + (i-b.lo[0]+kG);
    };
    std::vector<SeamPair> pairs;
    for (int p = 0; p < 6; ++p) for (int d = 1; d < 3; ++d) for (int s = 0; s < 2; ++s)
    {
        const SeamEdge e = seamEdge(na, p, d, s);
        if ((p*2+(d-1))*2+s > (e.q*2+(e.dq-1))*2+e.sq) continue;
        const int f = s ? na : 0, fq = e.sq ? na : 0;
        for (int t = 0; t < na; ++t) for (int i = 0; i < nr; ++i)
        {
            const int tq = e.rev ? na-1-t : t;
            const int jf = d == 1 ? f : t, kf = d == 1 ? t : f;
            const int jq = e.dq == 1 ? fq : tq, kq = e.dq == 1 ? tq : fq;
            const int na1 = na-1;
            const int na_ = boxOf(p, i, d == 1 ? (s ? na1 : 0) : t, d == 1 ? t : (s ? na1 : 0));
            const int nb = boxOf(e.q, i, e.dq == 1 ? (e.sq ? na1 : 0) : tq, e.dq == 1 ? tq : (e.sq ? na1 : 0));
            SeamPair sp;
            sp.a = faceIndex(na_, d, i, jf, kf, sp.ea);
            sp.b = faceIndex(nb, e.dq, i, jq, kq, sp.eb);
            sp.sign = s != e.sq ? 1.0 : -1.0;
            pairs.push_back(sp);
        }
    }
    return pairs;
}

// Box table over a fourth-order level: tiles of tile[0] x tile[1] x tile[2]
// cells (clipped to the panel) viewing the level arrays, with the seam pairs
// of their face fluxes. Every tile evaluates the same per-cell arithmetic, so
// the level result does not depend on the tile shape; only speed and scratch
// size do.
static hc_boxop* levelBoxTable(const Level4& Lv, const int tile[3])
{
    const Level4* L = &Lv;
    const Prep4View& v = L->pv;
    const int nr = v.nr, na = v.na;
    const hc_options& opt = L->opt;
    std::vector<hc_box> boxes;
    const int bs[3] = {std::min(tile[0], nr), std::min(tile[1], na), std::min(tile[2], na)};
    for (int p = 0; p < 6; ++p)
        for (int k = 0; k < na; k += bs[2]) for (int j = 0; j < na; j += bs[1]) for (int i = 0; i < nr; i += bs[0])
        {
            hc_box b{};
            b.block = p;
            b.lo[0] = i; b.lo[1] = j; b.lo[2] = k;
            b.hi[0] = std::min(i+bs[0], nr)-1; b.hi[1] = std::min(j+bs[1], na)-1; b.hi[2] = std::min(k+bs[2], na)-1;
            boxes.push_back(b);
        }
    std::unique_ptr<hc_boxop> B(new hc_boxop());
    B->opt = opt;
    B->dims = L->dims;
    buildBoxop(*B, boxes.data(), nullptr, (int)boxes.size());
    // re-point every box at its tile of the level arrays
    for (auto& bd : B->boxes)
    {
        bd.uOff = bd.p*v.sps + ((long long)bd.lo[2]*v.sJ + bd.lo[1])*v.sI + bd.lo[0];
        bd.us[0] = v.sI; bd.us[1] = v.sI*v.sJ; bd.us[2] = v.scs;
        bd.rOff = bd.p*v.vps + ((long long)bd.lo[2]*na + bd.lo[1])*nr + bd.lo[0];
        bd.rs[0] = nr; bd.rs[1] = (long long)nr*na; bd.rs[2] = v.vcs;
    }
    HC_THROW(hipMemcpy(const_cast<BoxDesc*>(B->view.boxes), B->boxes.data(), B->boxes.size()*sizeof(BoxDesc),
                       hipMemcpyHostToDevice));
    // seam pairs between the boxes' face fluxes
    std::vector<SeamPair> pairs = panelSeamPairs(B->boxes, B->boxes.size(), nr, na);
    B->view.seams = upload(pairs, B->owned);
    B->view.nseams = (int)pairs.size();
    return B.release();
}

// Metric tables of a fourth-order level's stage preparation: cell-averaged
// det A, <A> and its inverse over the angular range -6 .. na+5, dVolr and the
// mid radius over -6 .. nr+6.
void prep4Tables(const hc_shell& shell, const hc_options& opt, Prep4View& v, std::vector<void*>& owned)
{
    const hc::ShellDims dims{shell.n_rad, shell.n_ang, kG, shell.inner_radius, shell.outer_radius,
                             shell.radial_stretch};
    const int nr = shell.n_rad, na = shell.n_ang;
    v.nr = nr; v.na = na; v.NE = na+12;
    v.vcs = (long long)nr*na*na; v.vps = 8*v.vcs;
    v.cJ = na+2*kGC; v.ccs = v.cJ*v.cJ*nr; v.cps = 8*v.ccs;
    v.sI = nr+10; v.sJ = na+10; v.scs = v.sI*v.sJ*(na+10); v.sps = 8*v.scs;
    v.gamma = opt.gamma;
    std::vector<double> det((size_t)v.NE*v.NE), Aavg(9ull*6*v.NE*v.NE), Ainv(9ull*6*v.NE*v.NE);
    std::vector<double> gx, gw;
    gaussLegendre01(10, gx, gw);
    const double h = 1.0/na;
    for (int k = -6; k < na+6; ++k) for (int j = -6; j < na+6; ++j) det[(size_t)(k+6)*v.NE+(j+6)] = cellDetA(na, j, k);
    for (int p = 0; p < 6; ++p) for (int k = -6; k < na+6; ++k) for (int j = -6; j < na+6; ++j)
    {
        double M[3][3] = {}, Mi[3][3];
        for (int a = 0; a < (int)gx.size(); ++a) for (int b = 0; b < (int)gx.size(); ++b)
        {
            double A[3][3];
            pointBasis(p, xiToAngle((j+gx[a])*h), xiToAngle((k+gx[b])*h), A);
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) M[r][c] += gw[a]*gw[b]*A[r][c];
        }
        invert3(M, Mi);
        const size_t x = 9*(((size_t)p*v.NE+(k+6))*v.NE+(j+6));
        for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) { Aavg[x+3*r+c] = M[r][c]; Ainv[x+3*r+c] = Mi[r][c]; }
    }
    std::vector<double> dvol(nr+13), rmid(nr+13);
    for (int i = -6; i <= nr+6; ++i) { const RadialCell c = radialCell(dims, i); dvol[i+6] = c.dvol; rmid[i+6] = c.rmid; }
    v.det = upload(det, owned); v.Aavg = upload(Aavg, owned); v.Ainv = upload(Ainv, owned);
    v.dvol = upload(dvol, owned); v.rmid = upload(rmid, owned);
}

Level4* level4Create(const hc_shell& shell, const hc_options& opt, const int tile[3], bool withTable)
{
    std::unique_ptr<Level4> L(new Level4());
    L->opt = opt;
    L->dims = {shell.n_rad, shell.n_ang, kG, shell.inner_radius, shell.outer_radius, shell.radial_stretch};
    const int nr = shell.n_rad, na = shell.n_ang;
    Prep4View& v = L->pv;
    prep4Tables(shell, opt, v, L->owned);
    std::vector<long long> colDst, donorSrc;
    std::vector<int> colStart{0};
    std::vector<double> donorW;
    for (int p = 0; p < 6; ++p) for (int k = -kGC; k < na+kGC; ++k) for (int j = -kGC; j < na+kGC; ++j)
    {
        if (j >= 0 && j < na && k >= 0 && k < na) continue;
        colDst.push_back(p*v.cps + ((long long)(k+kGC)*v.cJ + (j+kGC))*nr);
        for (const auto& d : averageDonors(na, p, j, k))
        {
            donorSrc.push_back(d.q*v.cps + ((long long)(d.k+kGC)*v.cJ + (d.j+kGC))*nr);
            donorW.push_back(d.w);
        }
        colStart.push_back((int)donorW.size());
    }
    L->ncol = (int)colDst.size();
    L->colDst = upload(colDst, L->owned);
    L->colStart = upload(colStart, L->owned);
    L->donorSrc = upload(donorSrc, L->owned);
    L->donorW = upload(donorW, L->owned);
    HC_THROW(hipMalloc(&L->cart, 6ull*v.cps*sizeof(double))); L->owned.push_back(L->cart);
    HC_THROW(hipMemset(L->cart, 0, 6ull*v.cps*sizeof(double)));
    if (withTable) { L->box = levelBoxTable(*L, tile); radialFaceTable(*L); }
    for (int d = 0; d < 3; ++d) L->tile[d] = tile[d];
    return L.release();
}

void level4Retile(Level4& L, const int tile[3])
{
    hc_boxop* B = levelBoxTable(L, tile);
    delete L.box;
    L.box = B;
    radialFaceTable(L);
    for (int d = 0; d < 3; ++d) L.tile[d] = tile[d];
}

// radialCons of the last stage RHS: cons[c] += scale * (sum of the inner
// radial face fluxes - sum of the outer ones), from the tiles' scratch.
void level4RadialCons(Level4& L, double scale, double* cons, hipStream_t s)
{
    hipLaunchKernelGGL(k4_radcons_part, dim3(L.radBlocks), dim3(kB), 0, s, L.radFace, L.radEc, L.nrad, kRadChunk,
                       L.box->scratch, L.radPart);
    hipLaunchKernelGGL(k4_radcons_final, dim3(1), dim3(64), 0, s, L.radPart, L.radBlocks, scale, cons);
    HC_THROW(hipGetLastError());
}

void level4Destroy(Level4* L) { delete L; }
void level4TileShape(const Level4& L, int tile[3]) { for (int d = 0; d < 3; ++d) tile[d] = L.tile[d]; }

void level4Prepare(Level4& L, const double* U0, const double* K, double a_dt, const double* inner, double* US,
                   hipStream_t s)
{
    const Prep4View& v = L.pv;
    // one block per angular column (j in x; k and the panel in y), threads
    // along the radius in multiples of the wavefront, at most kB
    const dim3 colBlock(std::min(kB, (v.nr+63)/64*64));
    // components per LDS pass: all 8 while 8 radial lines fit in 64 KB
    const dim3 cols(6u*v.na*v.na);
    if (8*v.nr*sizeof(double) <= 65536)
        hipLaunchKernelGGL(k4_angrad<8>, cols, colBlock, 8*v.nr*sizeof(double), s, v, U0, K, a_dt, L.cart);
    else if (2*v.nr*sizeof(double) <= 65536)
        hipLaunchKernelGGL(k4_angrad<2>, cols, colBlock, 2*v.nr*sizeof(double), s, v, U0, K, a_dt, L.cart);
    else
        hipLaunchKernelGGL(k4_angrad<1>, cols, colBlock, v.nr*sizeof(double), s, v, U0, K, a_dt, L.cart);
    if (L.ncol > 0)
        hipLaunchKernelGGL(k4_seam, dim3(L.ncol), colBlock, 0, s, v, L.colDst, L.colStart, L.donorSrc, L.donorW,
                           L.cart);
    hipLaunchKernelGGL(k4_native, dim3(6u*(v.na+10)*(v.na+10)), colBlock, 0, s, v, L.cart, US);
    hipLaunchKernelGGL(k4_radial_bc, dim3((10*6*(v.na+10)*(v.na+10)+kB-1)/kB), dim3(kB), 0, s, v, inner, US);
    HC_THROW(hipGetLastError());
}

void level4Rhs(Level4& L, const double* US, double* K, double* U, double b_dt, hipStream_t s)
{
    double* const F[3] = {nullptr, nullptr, nullptr};
    if (L.opt.precision == 1) applyBoxop<float>(*L.box, US, K, F, s, 1, U, b_dt);
    else applyBoxop<double>(*L.box, US, K, F, s, 1, U, b_dt);
}
} // namespace hc

extern "C" {

int hc_boxop_create(const hc_shell* shell, const hc_options* options, const hc_box* boxes, int nbox,
                    hc_boxop** out)
{
    return hc_boxop_create_levels(shell, 1, options, boxes, nullptr, nbox, out);
}

int hc_boxop_create_levels(const hc_shell* shells, int nlevels, const hc_options* options, const hc_box* boxes,
                           const int* box_levels, int nbox, hc_boxop** out)
{
    if (!shells || nlevels <= 0 || !options || !boxes || !out || nbox <= 0)
        return hc::fail("hc_boxop_create: bad argument");
    if (options->spatial_order != 4) return hc::fail("hc_boxop_create: spatial_order must be 4");
    if (options->riemann_solver != 1 && options->riemann_solver != 2)
        return hc::fail("hc_boxop_create: riemann_solver must be 1 (Rusanov) or 2 (Roe)");
    if (options->limiter < 0 || options->limiter > 2) return hc::fail("hc_boxop_create: limiter must be 0, 1 or 2");
    if (options->precision != 0 && options->precision != 1) return hc::fail("hc_boxop_create: precision must be 0 or 1");
    if (const char* e = hc::checkOperatorOptions(*options)) return hc::fail(std::string("hc_boxop_create: ") + e);
    for (int l = 0; l < nlevels; ++l)
    {
        const hc_shell& sh = shells[l];
        // n_ang >= 16: the metric tables span the angular cells -5 .. n_ang+5,
        // which must stay inside the panel's horizon (|xi - 1/2| < 1)
        if (sh.n_rad < 4 || sh.n_ang < 16 || !(sh.outer_radius > sh.inner_radius && sh.inner_radius > 0))
            return hc::fail("hc_boxop_create: bad level shell (need n_rad >= 4, n_ang >= 16 and 0 < r0 < r1)");
    }
    for (int n = 0; n < nbox; ++n)
    {
        const hc_box& b = boxes[n];
        const int lev = box_levels ? box_levels[n] : 0;
        if (lev < 0 || lev >= nlevels) return hc::fail("hc_boxop_create: box level out of range");
        if (b.block < 0 || b.block > 5) return hc::fail("hc_boxop_create: box block must be 0..5");
        for (int d = 0; d < 3; ++d)
        {
            const int lim = d == 0 ? shells[lev].n_rad : shells[lev].n_ang;
            if (b.lo[d] < 0 || b.hi[d] >= lim || b.hi[d] < b.lo[d] + 3)
                return hc::fail("hc_boxop_create: box outside the panel or thinner than 4 cells");
        }
    }
    try
    {
        std::unique_ptr<hc_boxop> B(new hc_boxop());
        B->opt = *options;
        for (int l = 0; l < nlevels; ++l)
            B->levels.push_back({shells[l].n_rad, shells[l].n_ang, kG, shells[l].inner_radius, shells[l].outer_radius,
                                 shells[l].radial_stretch});
        B->dims = B->levels[0];
        buildBoxop(*B, boxes, box_levels, nbox);
        *out = B.release();
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_boxop_create: ") + e.what()); }
}

void hc_boxop_destroy(hc_boxop* op) { delete op; }

long long hc_boxop_count(const hc_boxop* op, int what)
{
    if (!op) return -1;
    switch (what)
    {
    case 0: return op->uCount;
    case 1: return op->rCount;
    case 2: case 3: case 4: return op->fCount[what-2];
    default: return -1;
    }
}

int hc_boxop_apply_scaled(hc_boxop* op, const double* USph, double* rhs, double* F0, double* F1, double* F2,
                          double scale, hc_stream stream)
{
    if (!op || !USph || !rhs) return hc::fail("hc_boxop_apply: null argument");
    if (!std::isfinite(scale)) return hc::fail("hc_boxop_apply: the divergence scale must be finite");
    try
    {
        double* const F[3] = {F0, F1, F2};
        const hipStream_t s = static_cast<hipStream_t>(stream);
        if (op->opt.precision == 1) applyBoxop<float>(*op, USph, rhs, F, s, 0, nullptr, 0.0, scale);
        else applyBoxop<double>(*op, USph, rhs, F, s, 0, nullptr, 0.0, scale);
        return 0;
    }
    catch (const std::exception& e) { return hc::fail(std::string("hc_boxop_apply: ") + e.what()); }
}

int hc_boxop_apply(hc_boxop* op, const double* USph, double* rhs, double* F0, double* F1, double* F2,
                   hc_stream stream)
{
    return hc_boxop_apply_scaled(op, USph, rhs, F0, F1, F2, 1.0, stream);
}

} // extern "C"

#include "hc_amr4.hpp"
