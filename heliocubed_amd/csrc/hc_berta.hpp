// Berta WENOZ fourth-order operator (spatial_operator 2) of the HelioCubed
// MHD RHS on MI355X: the kernels of one RK stage over a box table. Included
// by hc_o4.hip inside its anonymous namespace; shares its box table, scratch
// layout, metric tables, face-product helpers and Roe kernel.
//
// Restates the host path of BoxOp_EulerCubedSphere::BertaOperator
// (src/BoxOp_BertaCubedSphereImplem.H:1737-2076; src/ = HelioCubed reference
// repository) with the Jameson detector and Roe 8-wave dissipation:
//   sphAvgToPrim (:278-340): rho, p and the native |B|^2/2 of V(<U>) on box+5  (k_b_wbar)
//   BertaComputeThetaJameson (:346-507): raw flag on box+4            (k_b_thraw)
//   dilated flag on box+3 + BertaAverageToPoint (:696-741) + sphPointToPrim (:212-275):
//     point W on box+4                                                 (k_b_wp)
//   per direction:
//     BertaReconstructWithFallback (:747-1239): WENOZ / WENO3 / PLM face points
//     Roe repair with the cell-average donors (:1817-1844)
//     point fluxes of both sides summed (:1904-1946)
//     theta-aware face averages (:1246-1266) and BertaPrimPointToMappedFlux (:1334-1534)
//       with the separable-metric shortcut (:1-98), in one LDS-tiled kernel
//     primToMappedRoeFlux with theta-aware dissipation average (src/BoxOp_RoeCubedSphere.H:29-145),
//     or BertaPrimPointToMappedRusanovFlux (:1541-1730) for riemann_solver 1
//     F = (Sum - D)/2 (:2004-2026)
// Max/min follow the reference's std::max / std::min ((a < b) comparisons).

// Scratch slots of the Berta operator (the standard operator's, reused).
constexpr int F_DRHO = F_WB, F_DP = F_WB+1, F_DBME = F_WB+2;   // rho, p, |B|^2/2 of V(<U>), box+5
constexpr int F_THR = F_MEB;                // raw Jameson flag, box+4
constexpr int F_THETA = F_ETA;              // dilated flag, box+4 (0 outside box+3)
constexpr int F_WPT = F_W;                  // point primitives, box+4
constexpr int F_BWL = F_PTL, F_BWR = F_PTR; // face points (the Powell products read them)
constexpr int F_THF = F_P0L;                // flux-average flag of a face
constexpr int F_PTOT = F_P0R;               // point total pressure, both sides summed
constexpr int F_FPS = F_PHL;                // point physical fluxes, both sides summed
// Mixed precision stores the point Roe dissipation in FP32 (8 components of
// FP32 in the first half of the F_DPT slots): it enters the face flux only as
// the dissipation term, whose relative rounding the FP32 eigen-system already
// carries. Marker in its component 0 of a face whose Roe dissipation
// k_b_roe64 evaluates (a quiet NaN with a payload no arithmetic produces).
constexpr unsigned kRoeDeferred = 0x7fdefe4au;
__device__ __forceinline__ float* dissF32(double* P, const BoxIdx& I) { return reinterpret_cast<float*>(P + (long long)F_DPT*I.ecount); }

__device__ __forceinline__ double refMax(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double refMin(double a, double b) { return b < a ? b : a; }

// bertaNormalizeWenoStencil (:747-773): divides by the largest magnitude and
// returns it (NaN-aware as the reference: a NaN is kept, never zeroed). The
// running maximum skips NaN magnitudes as the reference's (a < b ? b : a)
// does, which is what fmax does with one NaN operand.
template <int N>
__device__ double wenoNormalize(double (&v)[N])
{
    double scale = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) scale = fmax(scale, fabs(v[i]));
    if (scale == 0.0)
    {
#pragma unroll
        for (int i = 0; i < N; ++i) if (v[i] != 0.0) return v[i];
        return scale;
    }
    if (scale > 0.0)
    {
        if (scale >= 2.2250738585072014e-308)
        {
            const double inv = 1.0/scale;
#pragma unroll
            for (int i = 0; i < N; ++i) v[i] *= inv;
        }
        else
        {
#pragma unroll
            for (int i = 0; i < N; ++i) v[i] /= scale;
        }
    }
    return scale;
}

// Mixed precision: the same normalization by a power of two, 2^e with
// max|v| < 2^e <= 2 max|v|, so that dividing and multiplying back are exact
// and no FP64 division is needed. Zero and non-finite maxima take the
// reference form above. The normalized maximum lies in [0.5, 1) instead of
// being 1; the WENO weights are homogeneous of degree 0 in the values up to
// the epsilon, which lies far below the indicators of normalized data.
template <int N>
__device__ double wenoNormalizePow2(double (&v)[N])
{
    double scale = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) scale = fmax(scale, fabs(v[i]));
    if (!(scale > 0.0 && scale <= 1.7976931348623157e308)) return wenoNormalize(v);
    int e;
    (void)frexp(scale, &e);
    const double inv = ldexp(1.0, -e);
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] *= inv;
    return ldexp(1.0, e);
}

// Quotient a / b of the WENO weights: FP64 division, or in FP32 (mixed
// precision) the product with the hardware reciprocal, whose ulp-level error
// is far below the FP32 rounding already accepted in the weights.
__device__ __forceinline__ double wenoQuot(double a, double b) { return a/b; }
__device__ __forceinline__ float wenoQuot(float a, float b) { return a*__builtin_amdgcn_rcpf(b); }

// One side of f_wenoz_pointwise_faceF (:782-878) on normalized values. The
// candidate values q_k and the undivided differences are formed in FP64; the
// smoothness indicators and nonlinear weights in S (S = double reproduces
// the reference operation for operation); the weighted sum in FP64. The
// reference's epsilon 1e-40 lies below the FP32 normal range: in FP32 it is
// 1e-30, which keeps every quotient finite (a zero indicator next to a jump
// gives a weight of at most ~1e30, as the FP64 form gives ~1e40) and changes
// no weight of normalized smooth data, whose indicators are far above it.
template <typename S>
__device__ __forceinline__ double wenozSide(double m2, double m1, double c, double p1, double p2)
{
    constexpr bool F32 = std::is_same<S, float>::value;
    const S eps = F32 ? S(1.0e-30) : S(1.0e-40), d0 = S(1.0/16.0), d1 = S(5.0/8.0), d2 = S(5.0/16.0);
    const double q0 = (3.0*m2 - 10.0*m1 + 15.0*c)/8.0;
    const double q1 = (-m1 + 6.0*c + 3.0*p1)/8.0;
    const double q2 = (3.0*c + 6.0*p1 - p2)/8.0;
    const S b0a = S(m2 - 2.0*m1 + c), b0b = S(m2 - 4.0*m1 + 3.0*c);
    const S b1a = S(m1 - 2.0*c + p1), b1b = S(m1 - p1);
    const S b2a = S(c - 2.0*p1 + p2), b2b = S(3.0*c - 4.0*p1 + p2);
    const S beta0 = (S(13.0)/S(12.0))*b0a*b0a + S(0.25)*b0b*b0b;
    const S beta1 = (S(13.0)/S(12.0))*b1a*b1a + S(0.25)*b1b*b1b;
    const S beta2 = (S(13.0)/S(12.0))*b2a*b2a + S(0.25)*b2b*b2b;
    const S tau5 = fabs(beta0 - beta2);
    const double a0 = double(d0*(S(1.0) + wenoQuot(tau5, beta0 + eps)));
    const double a1 = double(d1*(S(1.0) + wenoQuot(tau5, beta1 + eps)));
    const double a2 = double(d2*(S(1.0) + wenoQuot(tau5, beta2 + eps)));
    const double asum = a0 + a1 + a2;
    return (a0*q0 + a1*q1 + a2*q2)/asum;
}

// Mixed-precision form of wenozSide<float>: the weighted sum
// (a0 q0 + a1 q1 + a2 q2)/(a0 + a1 + a2) written as q1 + w0 (q0 - q1) +
// w2 (q2 - q1) with the FP32 normalized weights w = a/sum (hardware
// reciprocal). q0 - q1 = 3/8 (b0a - b1a) and q2 - q1 = (b1a - b2a)/8 are
// formed in FP64 from the second differences and rounded to FP32; q1 stays
// FP64. The weights' rounding then enters only through the candidates'
// spread, as in the quotient form, and no FP64 division is left.
__device__ __forceinline__ double wenozSideMixed(double m2, double m1, double c, double p1, double p2)
{
    const float eps = 1.0e-30f, d0 = 1.0f/16.0f, d1 = 5.0f/8.0f, d2 = 5.0f/16.0f;
    const double b0a = m2 - 2.0*m1 + c, b1a = m1 - 2.0*c + p1, b2a = c - 2.0*p1 + p2;
    const float f0a = float(b0a), f0b = float(m2 - 4.0*m1 + 3.0*c);
    const float f1a = float(b1a), f1b = float(m1 - p1);
    const float f2a = float(b2a), f2b = float(3.0*c - 4.0*p1 + p2);
    const float beta0 = (13.0f/12.0f)*f0a*f0a + 0.25f*f0b*f0b;
    const float beta1 = (13.0f/12.0f)*f1a*f1a + 0.25f*f1b*f1b;
    const float beta2 = (13.0f/12.0f)*f2a*f2a + 0.25f*f2b*f2b;
    const float tau5 = fabsf(beta0 - beta2);
    const float a0 = d0*(1.0f + wenoQuot(tau5, beta0 + eps));
    const float a1 = d1*(1.0f + wenoQuot(tau5, beta1 + eps));
    const float a2 = d2*(1.0f + wenoQuot(tau5, beta2 + eps));
    const float inv = __builtin_amdgcn_rcpf(a0 + a1 + a2);
    const float e0 = 0.375f*float(b0a - b1a), e2 = 0.125f*float(b1a - b2a);
    const double q1 = (-m1 + 6.0*c + 3.0*p1)/8.0;
    return q1 + double((a0*inv)*e0 + (a2*inv)*e2);
}

// One side of f_weno3_pointwise_faceF (:881-946), eps = dx^2; differences and
// candidates in FP64, indicators and weights in S.
template <typename S>
__device__ __forceinline__ double weno3Side(double m1, double c, double p1, double eps_d)
{
    const S d0 = S(3.0/4.0), d1 = S(1.0/4.0), eps = S(eps_d);
    const double Dm = c - m1, Dp = p1 - c;
    const S Dms = S(Dm), Dps = S(Dp), Dd = S(Dp - Dm);
    const S beta0 = Dps*Dps, beta1 = Dms*Dms, tau = Dd*Dd;
    const double a0 = double(d0*(S(1.0) + tau/(beta0 + eps)));
    const double a1 = double(d1*(S(1.0) + tau/(beta1 + eps)));
    const double q0 = 0.5*(p1 + c), q1 = 0.5*(-m1 + 3.0*c);
    return (a0*q0 + a1*q1)/(a0 + a1);
}

// berta_minmod3_scalar (:949-956)
template <typename S>
__device__ __forceinline__ S minmod3(S a, S b, S c)
{
    if (a > S(0) && b > S(0) && c > S(0)) { const S m = c < b ? c : b; return m < a ? m : a; }
    if (a < S(0) && b < S(0) && c < S(0)) { const S m = b < c ? c : b; return a < m ? m : a; }
    return S(0);
}

// Cell-average primitives V(<U>) of a cell of box+5 (sphAvgToPrim, :278-340):
// the Jameson detector's inputs and the Roe repair's donors.
__device__ __forceinline__ void bertaCellPrim(const O4View& v, const BoxDesc& b, const double* __restrict__ U,
                                              int i, int j, int k, double w[8])
{
    const long long u = b.uOff + (long long)(k-b.lo[2]+kG)*b.us[1] + (long long)(j-b.lo[1]+kG)*b.us[0] + (i-b.lo[0]+kG);
    double uu[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) uu[c] = U[u + c*b.us[2]];
    sphToPrimPoint(w, uu);
    double ke, me, vb;
    energiesNative(v.cellA + 9*angIdx(b, j, k), w, ke, me, vb);
    w[4] = (w[4]-ke-me)*(v.gamma-1.0);
    if (w[4] < 0) w[4] = 0.0;
}

// 0. The Jameson detector's inputs on box+5: rho, p and the native |B|^2/2 of
// the cell-average primitives V(<U>) (sphAvgToPrim, :278-340). The first
// thread of a box also clears the box's flags of deferred FP64 Roe faces for
// the stage's direction passes.
__global__ void __launch_bounds__(kB) k_b_wbar(O4View v, const double* __restrict__ U, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    const unsigned t = boxThread(v, b);
    if (t == 0)
    {
        const int n = blockBoxIndex(v);
        v.dflag[3*n] = 0u; v.dflag[3*n+1] = 0u; v.dflag[3*n+2] = 0u;
    }
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-kG; H[d] = b.hi[d]+kG; }
    if (!regionPoint(b, L, H, t, x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    double w[8];
    bertaCellPrim(v, b, U, x[0], x[1], x[2], w);
    P[I(F_DRHO, x[0], x[1], x[2])] = w[0];
    P[I(F_DP, x[0], x[1], x[2])] = w[4];
    P[I(F_DBME, x[0], x[1], x[2])] = 0.5*(w[5]*w[5] + w[6]*w[6] + w[7]*w[7]);
}

// 1. Raw Jameson flag on box+4 (f_berta_jameson_theta_rawF, :346-407): the
// normalized second difference of rho, p and the native |B|^2/2 per
// direction, the largest of the three, the 2-norm over directions, 1 below
// eta_d.
__global__ void __launch_bounds__(kB) k_b_thraw(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-4; H[d] = b.hi[d]+4; }
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    auto eta1 = [](double qm, double q0, double qp) {
        const double num = fabs(qp - 2.0*q0 + qm);
        const double den = fabs(qp) + 2.0*fabs(q0) + fabs(qm) + 1.0e-40;
        return num/den; };
    double et[3];
#pragma unroll
    for (int dd = 0; dd < 3; ++dd)
    {
        const int m[3] = {x[0]-(dd == 0), x[1]-(dd == 1), x[2]-(dd == 2)};
        const int p[3] = {x[0]+(dd == 0), x[1]+(dd == 1), x[2]+(dd == 2)};
        auto q = [&](int f, const int y[3]) { return P[I(f, y[0], y[1], y[2])]; };
        const double er = eta1(q(F_DRHO, m), q(F_DRHO, x), q(F_DRHO, p));
        const double ep = eta1(q(F_DP, m), q(F_DP, x), q(F_DP, p));
        const double em = eta1(q(F_DBME, m), q(F_DBME, x), q(F_DBME, p));
        et[dd] = refMax(er, refMax(ep, em));
    }
    const double eta = sqrt(et[0]*et[0] + et[1]*et[1] + et[2]*et[2]);
    P[I(F_THR, x[0], x[1], x[2])] = eta < v.etaD ? 1.0 : 0.0;
}

// 4. Face points of direction D on the faces [lo_d, hi_d+1] x transverse
// [lo-2, hi+2]: WENOZ where the six-cell flag minimum is 1, else WENO3 (or PLM),
// floors, the Roe repair, then the point fluxes and total pressures of both
// sides summed, and on the faces [lo_d, hi_d+1] x [lo-1, hi+1] the point Roe
// dissipation of the two face points (factoredRoe, src/MHD_Roe8Wave.H:634-740),
// so the face points never leave registers (Rusanov: the conserved jump and
// the speed). Reconstruction arithmetic in S (the normalization and the
// rescaling stay FP64).
template <int D, typename S_, bool ROE>
__device__ __forceinline__ void bertaFaces(const O4View& v, const double* __restrict__ U, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-2; H[d] = b.hi[d]+2; }
    // the Powell face products also take the faces lo_d-1 and hi_d+2
    L[D] = b.lo[D] - (v.powell ? 1 : 0); H[D] = b.hi[D] + (v.powell ? 2 : 1);
    if (!(D == 2 ? regionPointX2(b, L, H, boxThread(v, b), x) : regionPoint(b, L, H, boxThread(v, b), x))) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    auto at = [&](int f, int o) {
        int y[3] = {x[0], x[1], x[2]}; y[D] += o;
        return P[I(f, y[0], y[1], y[2])]; };
    double thF = refMin(at(F_THETA, -1), at(F_THETA, 0));
    double tr = at(F_THETA, -3);
    tr = refMin(tr, at(F_THETA, -2)); tr = refMin(tr, at(F_THETA, -1)); tr = refMin(tr, at(F_THETA, 0));
    tr = refMin(tr, at(F_THETA, 1)); tr = refMin(tr, at(F_THETA, 2));
    const double dx = D == 0 ? 1.0/b.nr : 1.0/b.na;
    const double dx2 = dx*dx;
    double l[8], r[8];
    // the six stencil values of the next component are loaded while the
    // current one is reconstructed
    double nx[6];
    auto load6 = [&](int c) {
#pragma unroll
        for (int o = -3; o <= 2; ++o) nx[o+3] = at(F_WPT+c, o); };
    load6(0);
    // unrolled: the results land in registers of their own (no indexed
    // register moves per component)
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        double s6[6];
#pragma unroll
        for (int o = 0; o < 6; ++o) s6[o] = nx[o];
        if (c < 7) load6(c+1);
        if (tr > 0.5)
        {
            if constexpr (std::is_same<S_, float>::value)
            {
                const double scale = wenoNormalizePow2(s6);
                if (scale == 0.0) { l[c] = 0.0; r[c] = 0.0; }
                else
                {
                    l[c] = scale*wenozSideMixed(s6[0], s6[1], s6[2], s6[3], s6[4]);
                    r[c] = scale*wenozSideMixed(s6[5], s6[4], s6[3], s6[2], s6[1]);
                }
            }
            else
            {
                const double scale = wenoNormalize(s6);
                if (scale == 0.0) { l[c] = 0.0; r[c] = 0.0; }
                else
                {
                    l[c] = scale*wenozSide<S_>(s6[0], s6[1], s6[2], s6[3], s6[4]);
                    r[c] = scale*wenozSide<S_>(s6[5], s6[4], s6[3], s6[2], s6[1]);
                }
            }
        }
        else
        {
            double s4[4] = {s6[1], s6[2], s6[3], s6[4]};
            // the fallbacks run where the detector flags a non-smooth
            // stencil: FP64 in both precisions (their weights and slopes act
            // on the largest jumps)
            if (v.fallback == 0)
            {
                const double scale = wenoNormalize(s4);
                if (scale == 0.0) { l[c] = 0.0; r[c] = 0.0; }
                else
                {
                    l[c] = scale*weno3Side<double>(s4[0], s4[1], s4[2], dx2);
                    r[c] = scale*weno3Side<double>(s4[3], s4[2], s4[1], dx2);
                }
            }
            else
            {
                const double th = 1.5;
                const double dm = s4[1]-s4[0], dp = s4[2]-s4[1], dc = 0.5*(s4[2]-s4[0]);
                l[c] = s4[1] + 0.5*minmod3<double>(th*dm, dc, th*dp);
                const double rdm = s4[2]-s4[1], rdp = s4[3]-s4[2], rdc = 0.5*(s4[3]-s4[1]);
                r[c] = s4[2] - 0.5*minmod3<double>(th*rdm, rdc, th*rdp);
            }
        }
    }
    l[0] = refMax(l[0], hc::kRhoFloor); r[0] = refMax(r[0], hc::kRhoFloor);
    l[4] = refMax(l[4], 1.0e-12); r[4] = refMax(r[4], 1.0e-12);
    if (ROE && !(usablePrimitive(l) && usablePrimitive(r)))
    {
        // the cell-average donors V(<U>) of the two cells of the face
        double cl[8], cr[8];
        bertaCellPrim(v, b, U, x[0]-(D == 0), x[1]-(D == 1), x[2]-(D == 2), cl);
        bertaCellPrim(v, b, U, x[0], x[1], x[2], cr);
        if (usablePrimitive(cl) && usablePrimitive(cr))
        {
#pragma unroll
            for (int c = 0; c < 8; ++c) { l[c] = cl[c]; r[c] = cr[c]; }
            thF = 0.0;
        }
    }
    // mixed precision: constant divisions of the point energies and fluxes as
    // reciprocal products (within an ulp)
    constexpr bool FAST = !std::is_same<S_, double>::value;
    const double* A = v.faceA[D] + 9*angIdx(b, x[1], x[2]);
    double keL, meL, vbL, keR, meR, vbR, fl[8], fr[8];
    energiesNative<FAST>(A, l, keL, meL, vbL);
    energiesNative<FAST>(A, r, keR, meR, vbR);
    primFlux<FAST>(fl, l, keL, meL, vbL, v.gamma, D);
    primFlux<FAST>(fr, r, keR, meR, vbR, v.gamma, D);
#pragma unroll
    for (int c = 0; c < 8; ++c) P[I(F_FPS+c, x[0], x[1], x[2])] = fl[c] + fr[c];
    P[I(F_THF, x[0], x[1], x[2])] = thF;
    P[I(F_PTOT, x[0], x[1], x[2])] = (l[4] + meL) + (r[4] + meR);
    // the Powell face products take the normal field of both face points
    if (v.powell)
    {
        P[I(F_BWL+5+D, x[0], x[1], x[2])] = l[5+D];
        P[I(F_BWR+5+D, x[0], x[1], x[2])] = r[5+D];
    }
    if constexpr (ROE)
    {
        // point Roe dissipation of the two face points on the faces
        // [lo_d, hi_d+1] x transverse [lo-1, hi+1] (the faces whose averages
        // the flux kernel forms)
        bool roeFace = x[D] >= b.lo[D] && x[D] <= b.hi[D]+1;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            if (d != D) roeFace = roeFace && x[d] >= b.lo[d]-1 && x[d] <= b.hi[d]+1;
        if (roeFace)
        {
            // mixed precision: a face whose eigen-system needs FP64 is left
            // to k_b_roe64 with its face points and the marker in F_DPT
            constexpr bool MIXED = std::is_same<S_, float>::value;
            double out[8];
            const bool deferred = factoredRoe<S_, MIXED>(out, l, r, A, radialPoint(v, b, D, x[0]), D, v.gamma, v.entropy);
            if constexpr (MIXED)
            {
                float* const Df = dissF32(P, I);
                const int cell = I(0, x[0], x[1], x[2]);
                if (deferred)
                {
#pragma unroll
                    for (int c = 0; c < 8; ++c)
                    {
                        P[I(F_BWL+c, x[0], x[1], x[2])] = l[c];
                        P[I(F_BWR+c, x[0], x[1], x[2])] = r[c];
                    }
                    Df[cell] = __uint_as_float(kRoeDeferred);
                    v.dflag[3*blockBoxIndex(v) + D] = 1u;
                }
                else
                {
#pragma unroll
                    for (int q = 0; q < 8; ++q) Df[q*I.ecount + cell] = float(out[q]);
                }
            }
            else
            {
#pragma unroll
                for (int q = 0; q < 8; ++q) P[I(F_DPT+q, x[0], x[1], x[2])] = out[q];
            }
        }
    }
    else
    {
        // Rusanov: pointwise conserved jump R - L times the radial point area
        // and the speed of the point states (:1589-1613, :1653-1725)
        const double rp = radialPoint(v, b, D, x[0]);
        P[I(F_DPT, x[0], x[1], x[2])] = rp*r[0] - rp*l[0];
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            P[I(F_DPT+1+a, x[0], x[1], x[2])] = rp*r[0]*r[1+a] - rp*l[0]*l[1+a];
            P[I(F_DPT+5+a, x[0], x[1], x[2])] = rp*r[5+a] - rp*l[5+a];
        }
        P[I(F_DPT+4, x[0], x[1], x[2])] = rp*(keR + meR + r[4]/(v.gamma-1.0)) - rp*(keL + meL + l[4]/(v.gamma-1.0));
        P[I(F_SPD, x[0], x[1], x[2])] = rusanovSpeedO4(l, r, A, v.gamma, D, true);
    }
}

// The FP64 build holds the FP64 Roe solver (2 waves per SIMD); the mixed
// build defers the FP64 faces and fits 3 waves per SIMD.
template <int D, typename S_, bool ROE>
__global__ void __launch_bounds__(kB) k_b_faces(O4View v, const double* __restrict__ U, double* __restrict__ S)
{
    bertaFaces<D, S_, ROE>(v, U, S);
}
template <int D, bool ROE>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(3))) k_b_facesm(O4View v, const double* __restrict__ U,
                                                                                    double* __restrict__ S)
{
    bertaFaces<D, float, ROE>(v, U, S);
}

// 4c. Mixed precision: the point Roe dissipation of the faces k_b_faces left
// marked for the FP64 eigen-system (plasma beta below HC_MIXED_FP64_BETA),
// from the face points it stored there; the faces [lo_d, hi_d+1] x
// transverse [lo-1, hi+1]. The pass runs on at most kRoeBlocks blocks. Each
// block first reads every box's flag of direction D with its threads in
// parallel and returns when none is set (the common case: the solver's
// registers leave room for only two resident blocks per CU, so a full-size
// grid of blocks that each wait on one flag load took 50-60 us). Otherwise
// it steps through the face kernel's blocks (same box and thread mapping)
// with a stride of its grid (k_b_wbar clears the flags at the start of the
// next stage).
constexpr unsigned kRoeBlocks = 2048;
template <int D>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(2))) k_b_roe64(O4View v, double* __restrict__ S)
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
        for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-1; H[d] = b.hi[d]+1; }
        L[D] = b.lo[D]; H[D] = b.hi[D]+1;
        if (!regionPoint(b, L, H, (lb - (unsigned)b.block0)*blockDim.x + threadIdx.x, x)) continue;
        const BoxIdx I(b);
        double* const P = S + b.sOff;
        float* const Df = dissF32(P, I);
        const int cell = I(0, x[0], x[1], x[2]);
        if (__float_as_uint(Df[cell]) != kRoeDeferred) continue;
        const double* A = v.faceA[D] + 9*angIdx(b, x[1], x[2]);
        double wl[8], wr[8], out[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) { wl[q] = P[I(F_BWL+q, x[0], x[1], x[2])]; wr[q] = P[I(F_BWR+q, x[0], x[1], x[2])]; }
        factoredRoe<float>(out, wl, wr, A, radialPoint(v, b, D, x[0]), D, v.gamma, v.entropy);
#pragma unroll
        for (int q = 0; q < 8; ++q) Df[q*I.ecount + cell] = float(out[q]);
    }
}

// 5. Face-average and face-flux kernel of direction D over rectangular tiles
// of faces (theta-aware averages, :1246-1266; BertaPrimPointToMappedFlux,
// :1334-1534, with the separable-metric shortcut, :1-98; Roe or Rusanov
// dissipation averages; F = (Sum - D)/2, :2004-2026). A workgroup owns
// FluxTile<D>::T0 x T1 x T2 faces of one box; per input component it stages
// the summed point flux on the tile grown by 2 in the two transverse
// directions and the point dissipation grown by 1 in LDS, forms the face
// averages of the grown-by-1 tile there, and every face reads its stencils
// (transverse differences, the 3x3 radial-transverse window of the angular
// directions) from LDS. The flux-average flag and the face metric rows of
// the tile are staged once. Each output accumulates its terms in the same
// order as the term-by-term form of the operator.
// The points of an LDS region org - h + [0, N0) x [0, N1) x [0, N2) that
// this thread stages: their offsets in one field of an array whose point
// (o0, o1, o2) is at offset 0 with strides 1, st1, st2 (a box's scratch
// block, or the box's view of its <U>), -1 where the region leaves [lo, hi]
// (the points where the staged fields are defined; staged as 0). Computed
// once per tile, used for every component.
template <int N0, int N1, int N2>
struct StagePlan
{
    static constexpr int N = N0*N1*N2, K = (N + kFT - 1)/kFT;
    int off[K];
    __device__ StagePlan(const int org[3], const int h[3], const int lo[3], const int hi[3],
                         const int o[3], long long st1, long long st2)
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            const int p = threadIdx.x + k*kFT;
            const int a0 = p % N0, q = p / N0, a1 = q % N1, a2 = q / N1;
            const int y0 = org[0]-h[0]+a0, y1 = org[1]-h[1]+a1, y2 = org[2]-h[2]+a2;
            const bool in = p < N && y0 >= lo[0] && y0 <= hi[0] && y1 >= lo[1] && y1 <= hi[1]
                            && y2 >= lo[2] && y2 <= hi[2];
            off[k] = in ? (int)((y2-o[2])*st2 + (y1-o[1])*st1 + (y0-o[0])) : -1;
        }
    }
    // offsets in a box's scratch block (BoxIdx layout)
    __device__ StagePlan(const BoxIdx& I, const int org[3], const int h[3], const int lo[3], const int hi[3])
        : StagePlan(org, h, lo, hi, (const int[3]){I.o0, I.o1, I.o2}, I.e0, (long long)I.e0*I.e1) {}
    // load the field at offset f0 (field index x field stride, in elements)
    // into registers of its own type, then store them into LDS (converted
    // there, so that a prefetch does not wait on its loads)
    template <typename V>
    __device__ __forceinline__ void load(V (&val)[K], const V* __restrict__ P, long long f0) const
    {
#pragma unroll
        for (int k = 0; k < K; ++k) val[k] = off[k] >= 0 ? P[f0 + off[k]] : V(0);
    }
    template <typename T, typename V>
    __device__ __forceinline__ void store(T* __restrict__ dst, const V (&val)[K]) const
    {
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            const int p = threadIdx.x + k*kFT;
            if (N % kFT == 0 || p < N) dst[p] = T(val[k]);
        }
    }
};

// 2-3. Dilated flag and point primitives on box+4 over tiles of 32 x 4 x 2
// cells: the cell's dilated flag from the raw flags of k_b_thraw (stored for
// the face kernels), then per component the box's <U> on the tile grown by 1
// goes to LDS (the next component's loads in flight meanwhile) and UPoint =
// <U> - theta Lap<U>/24 is formed there; then V(UPoint) with the pressure
// clipped at 0 (:696-741, :212-275).
template <bool FAST>
__global__ void __launch_bounds__(kFT) k_b_wp(O4View v, const double* __restrict__ U, double* __restrict__ S)
{
    using G = CellTile;
    __shared__ double sU[G::NA];
    const FluxTileRef tl = v.ctiles[xcdOrder(blockIdx.x, v.nctiles)];
    const BoxDesc& b = v.boxes[tl.box];
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    const double* const Ub = U + b.uOff;    // <U>(lo-5) of the box: its own array or its view of a level
    const int org[3] = {tl.o0, tl.o1, tl.o2};
    const int h[3] = {1, 1, 1};
    int L[3], H[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-kG; H[d] = b.hi[d]+kG; }
    const StagePlan<G::A0, G::A1, G::A2> plan(org, h, L, H, L, b.us[0], b.us[1]);
    const int l0 = threadIdx.x % G::T0, l1 = (threadIdx.x / G::T0) % G::T1, l2 = threadIdx.x / (G::T0*G::T1);
    const int x[3] = {org[0]+l0, org[1]+l1, org[2]+l2};
    bool ok = true, inner = true;
    for (int d = 0; d < 3; ++d)
    {
        ok = ok && x[d] >= b.lo[d]-4 && x[d] <= b.hi[d]+4;
        inner = inner && x[d] >= b.lo[d]-3 && x[d] <= b.hi[d]+3;
    }
    // dilated flag (f_berta_dilate_thetaF, :410-434): the minimum of the raw
    // flag over the cell and its six neighbours on box+3, 0 in the outer
    // layer of box+4 (:500-503)
    double th = 0.0;
    if (ok && inner)
    {
        th = P[I(F_THR, x[0], x[1], x[2])];
        th = refMin(th, P[I(F_THR, x[0]-1, x[1], x[2])]); th = refMin(th, P[I(F_THR, x[0]+1, x[1], x[2])]);
        th = refMin(th, P[I(F_THR, x[0], x[1]-1, x[2])]); th = refMin(th, P[I(F_THR, x[0], x[1]+1, x[2])]);
        th = refMin(th, P[I(F_THR, x[0], x[1], x[2]-1)]); th = refMin(th, P[I(F_THR, x[0], x[1], x[2]+1)]);
    }
    if (ok) P[I(F_THETA, x[0], x[1], x[2])] = th;
    const bool smooth = ok && th > 0.5;
    const int c0 = ((l2+1)*G::A1 + (l1+1))*G::A0 + (l0+1);
    constexpr int s1 = G::A0, s2 = G::A0*G::A1;
    double val[StagePlan<G::A0, G::A1, G::A2>::K];
    plan.load(val, Ub, 0);
    double up[8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
    {
        __syncthreads();
        plan.store(sU, val);
        __syncthreads();
        if (c < 7) plan.load(val, Ub, (c+1)*b.us[2]);
        const double uc = sU[c0];
        const double lap = sU[c0+1]+sU[c0-1]+sU[c0+s1]+sU[c0-s1]+sU[c0+s2]+sU[c0-s2]-6.0*uc;
        up[c] = smooth ? uc - (FAST ? lap*(1.0/24.0) : lap/24.0) : uc;
    }
    if (!ok) return;
    double w[8];
    if constexpr (FAST)
    {
        // mixed precision: the point primitives feed only the reconstruction
        // (the detector takes the cell averages), so they take reciprocal products
        w[0] = fmax(up[0], hc::kRhoFloor);
        const double ir = 1.0/w[0];
        w[1] = up[1]*ir; w[2] = up[2]*ir; w[3] = up[3]*ir;
        w[4] = up[4]; w[5] = up[5]; w[6] = up[6]; w[7] = up[7];
    }
    else sphToPrimPoint(w, up);
    double ke, me, vb;
    energiesNative<FAST>(v.cellA + 9*angIdx(b, x[1], x[2]), w, ke, me, vb);
    w[4] = (w[4]-ke-me)*(v.gamma-1.0);
    if (w[4] < 0.0) w[4] = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) P[I(F_WPT+c, x[0], x[1], x[2])] = w[c];
}

template <int D, bool ROE, bool F32D>
__global__ void __launch_bounds__(kFT) k_b_avgflux(O4View v, double* __restrict__ S)
{
    using G = FluxTile<D>;
    constexpr int t1 = D == 0 ? 1 : 0, t2 = D == 2 ? 1 : 2;
    __shared__ double sF[G::NF];        // summed point flux (or total pressure) of one component
    __shared__ double sA[G::NA];        // its face average
    __shared__ double sD[G::NA];        // point dissipation of the component
    __shared__ double sM[13*G::NM];     // face metric rows of the tile's angular columns
    __shared__ float sT[G::NA];         // flux-average flag (0 or 1)
    const FluxTileRef tl = v.ftiles[D][xcdOrder(blockIdx.x, v.nftiles[D])];
    const BoxDesc& b = v.boxes[tl.box];
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    const int org[3] = {tl.o0, tl.o1, tl.o2};
    const int hA[3] = {G::h0, G::h1, G::h2}, hF[3] = {2*G::h0, 2*G::h1, 2*G::h2};
    // the face kernel defines the point fluxes on the faces [lo_d, hi_d+1] x
    // transverse [lo-2, hi+2], the dissipation on transverse [lo-1, hi+1]
    int LA[3], HA[3], LF[3], HF[3];
    for (int d = 0; d < 3; ++d)
    {
        LA[d] = b.lo[d]-hA[d]; HA[d] = b.hi[d]+hA[d]+(d == D);
        LF[d] = b.lo[d]-hF[d]; HF[d] = b.hi[d]+hF[d]+(d == D);
    }
    const StagePlan<G::F0, G::F1, G::F2> planF(I, org, hF, LF, HF);
    const StagePlan<G::A0, G::A1, G::A2> planA(I, org, hA, LA, HA);
    {
        double t[StagePlan<G::A0, G::A1, G::A2>::K];
        planA.load(t, P, F_THF*I.ecount);
        planA.store(sT, t);
    }
    for (int p = threadIdx.x; p < 13*G::NM; p += kFT)
    {
        const int col = p / 13, c = p - 13*col, a1 = col % G::A1, a2 = col / G::A1;
        const int j = org[1]-G::h1+a1, k = org[2]-G::h2+a2;
        const bool in = j >= LA[1] && j <= HA[1] && k >= LA[2] && k <= HA[2];
        sM[p] = in ? v.fmet[D][13*angIdx(b, j, k) + c] : 0.0;
    }
    // LDS index of the face-local point l + offsets in the averaged region
    auto ia = [](int l0, int l1, int l2) { return ((l2+G::h2)*G::A1 + (l1+G::h1))*G::A0 + (l0+G::h0); };
    auto mrow = [](int l1, int l2) { return 13*((l2+G::h2)*G::A1 + (l1+G::h1)); };
    auto sh = [&](const double* s, const int l[3], int t, int o) {
        int y[3] = {l[0], l[1], l[2]}; y[t] += o; return s[ia(y[0], y[1], y[2])]; };
    auto Mv = [&](int c, const int l[3]) { return sM[mrow(l[1], l[2]) + c]; };
    auto Dm = [&](int c, int t, const int l[3]) {
        int a[3] = {l[0], l[1], l[2]}, z[3] = {l[0], l[1], l[2]};
        a[t] += 1; z[t] -= 1;
        return 0.5*sM[mrow(a[1], a[2]) + c] - 0.5*sM[mrow(z[1], z[2]) + c]; };
    auto Df = [&](const double* s, int t, const int l[3]) { return 0.5*sh(s, l, t, 1) - 0.5*sh(s, l, t, -1); };
    // radial half difference and its centred t2 difference (angular directions)
    auto Gv0 = [&](const double* s, const int l[3]) { return 0.5*sh(s, l, 0, 1) - 0.5*sh(s, l, 0, -1); };
    auto DGv = [&](const double* s, const int l[3]) {
        auto s2 = [&](int o0, int o2) { int y[3] = {l[0]+o0, l[1], l[2]}; y[t2] += o2; return s[ia(y[0], y[1], y[2])]; };
        const double gp = 0.5*s2(1, 1) - 0.5*s2(-1, 1);
        const double gm = 0.5*s2(1, -1) - 0.5*s2(-1, -1);
        return 0.5*gp - 0.5*gm; };
    auto Davg = [&](const double* s, bool smooth, const int l[3]) {
        const double c0 = s[ia(l[0], l[1], l[2])];
        double lap = 0;
        lap += sh(s, l, t1, 1) - 2.0*c0 + sh(s, l, t1, -1);
        lap += sh(s, l, t2, 1) - 2.0*c0 + sh(s, l, t2, -1);
        return smooth ? c0 + lap/24.0 : c0; };

    // the faces of this thread
    int lf[kFPT][3];
    bool ok[kFPT];
    double adjX[kFPT], dA[kFPT], speed[kFPT];
#pragma unroll
    for (int k = 0; k < kFPT; ++k)
    {
        const int q = threadIdx.x + k*kFT;
        lf[k][0] = q % G::T0; lf[k][1] = (q / G::T0) % G::T1; lf[k][2] = q / (G::T0*G::T1);
        ok[k] = true;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            const int x = org[d]+lf[k][d];
            ok[k] = ok[k] && x >= b.lo[d] && x <= b.hi[d]+(d == D);
        }
        const int x0 = org[0]+lf[k][0];
        adjX[k] = ok[k] ? areaOf(v, b, D, x0) : 0.0;
        dA[k] = (ok[k] && D != 0) ? 0.5*areaOf(v, b, D, x0+1) - 0.5*areaOf(v, b, D, x0-1) : 0.0;
        speed[k] = (ok[k] && !ROE) ? P[I(F_SPD, x0, org[1]+lf[k][1], org[2]+lf[k][2])] : 0.0;
    }
    // F-region index of the centre of each averaged point of this thread
    constexpr int KA = (G::NA + kFT - 1)/kFT;
    int fc[KA];
#pragma unroll
    for (int k = 0; k < KA; ++k)
    {
        const int p = threadIdx.x + k*kFT;
        const int a0 = p % G::A0, q = p / G::A0, a1 = q % G::A1, a2 = q / G::A1;
        fc[k] = ((a2+G::h2)*G::F1 + (a1+G::h1))*G::F0 + (a0+G::h0);
    }
    double press[kFPT][3], mp[kFPT][3], dmp[kFPT][3], fd[kFPT][3];
    // components in the order their outputs are completed: total pressure
    // (the momentum pressure force), 0, 4, momentum 1..3, field 5..7
    auto component = [](int s) { return s == 0 ? -1 : s == 1 ? 0 : s == 2 ? 4 : s <= 5 ? s-2 : s-1; };
    // the next component's point fluxes and dissipation are loaded into
    // registers while the current one is evaluated
    double nF[StagePlan<G::F0, G::F1, G::F2>::K];
    std::conditional_t<F32D, float, double> nD[StagePlan<G::A0, G::A1, G::A2>::K];
    planF.load(nF, P, F_PTOT*I.ecount);
#pragma unroll 1
    for (int s = 0; s < 9; ++s)
    {
        const int c = component(s);
        __syncthreads();
        planF.store(sF, nF);
        if (c >= 0) planA.store(sD, nD);
        __syncthreads();
        if (s < 8)
        {
            const int cn = component(s+1);
            planF.load(nF, P, (F_FPS+cn)*I.ecount);
            if constexpr (F32D) planA.load(nD, dissF32(P, I), (long long)cn*I.ecount);
            else planA.load(nD, P, (F_DPT+cn)*I.ecount);
        }
        // theta-aware face average on the tile grown by 1
#pragma unroll
        for (int k = 0; k < KA; ++k)
        {
            const int p = threadIdx.x + k*kFT;
            if (G::NA % kFT != 0 && p >= G::NA) break;
            auto F = [&](int t, int o) { return sF[fc[k] + (t == 0 ? o : t == 1 ? o*G::F0 : o*G::F0*G::F1)]; };
            const double c0 = sF[fc[k]];
            double lap = 0;
            lap += F(t1, 1) - 2.0*c0 + F(t1, -1);
            lap += F(t2, 1) - 2.0*c0 + F(t2, -1);
            sA[p] = sT[p] > 0.5f ? c0 + lap/24.0 : c0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kFPT; ++k)
        {
            if (!ok[k]) continue;
            const int* l = lf[k];
            const int x0 = org[0]+l[0], x1 = org[1]+l[1], x2 = org[2]+l[2];
            const bool smooth = sT[ia(l[0], l[1], l[2])] > 0.5f;
            if (c < 0)
            {
                // <p_tot adjDr> at l + o e_t2 (angular) or l (radial faces:
                // every point of the window shares the face's radius)
                auto tmp3 = [&](int t, int o) {
                    int y[3] = {l[0], l[1], l[2]}; y[t] += o;
                    double val = 0;
                    if constexpr (D != 0)
                    {
                        const double dv = 0.5*sh(sA, y, 0, 1) - 0.5*sh(sA, y, 0, -1);
                        val += dv*dA[k]*(1.0/12.0);
                    }
                    val += sA[ia(y[0], y[1], y[2])]*adjX[k]*1.0;
                    return val; };
                const double t0v = tmp3(0, 0);
                double dT[2] = {0.0, 0.0};
                if constexpr (D == 0) dT[0] = 0.5*tmp3(t1, 1) - 0.5*tmp3(t1, -1);
                dT[1] = 0.5*tmp3(t2, 1) - 0.5*tmp3(t2, -1);
#pragma unroll
                for (int r = 0; r < 3; ++r)
                {
                    double val = 0;
                    if constexpr (D == 0) val += Dm(10+r, t1, l)*dT[0]*(1.0/12.0);
                    val += Dm(10+r, t2, l)*dT[1]*(1.0/12.0);
                    val += Mv(10+r, l)*t0v*1.0;
                    press[k][r] = val;
                }
                continue;
            }
            auto fsum = [&](double m, double dm) {
                double val = 0;
                if constexpr (D != 0) val += dm*dA[k]*(1.0/12.0);
                val += m*adjX[k]*1.0;
                return val; };
            if (c == 0 || c == 4)
            {
                double m = 0, dm = 0;
                if constexpr (D == 0)
                {
                    double term = 0;
                    term += Dm(0, t1, l)*Df(sA, t1, l)*(1.0/12.0);
                    term += Dm(0, t2, l)*Df(sA, t2, l)*(1.0/12.0);
                    term += Mv(0, l)*sA[ia(l[0], l[1], l[2])]*1.0;
                    m += term;
                }
                else
                {
                    m += Dm(0, t2, l)*Df(sA, t2, l)*(1.0/12.0);
                    m += Mv(0, l)*sA[ia(l[0], l[1], l[2])];
                    dm += Dm(0, t2, l)*DGv(sA, l)*(1.0/12.0);
                    dm += Mv(0, l)*Gv0(sA, l);
                }
                const double fs = fsum(m, dm);
                const double da = Davg(sD, smooth, l);
                double dd;
                if constexpr (ROE)
                {
                    dd = 0;
                    dd += Dm(0, t1, l)*Df(sD, t1, l)*(1.0/12.0);
                    dd += Dm(0, t2, l)*Df(sD, t2, l)*(1.0/12.0);
                    dd += Mv(0, l)*da;
                }
                else dd = (Mv(0, l)*da)*speed[k];
                P[I(F_FLUX+8*D+c, x0, x1, x2)] = 0.5*(fs-dd);
                continue;
            }
            // vector components: momentum (with the pressure force) and field
            const int c0 = c <= 3 ? 1 : 5, col = c-c0;
            if (col == 0)
            {
#pragma unroll
                for (int r = 0; r < 3; ++r) { mp[k][r] = 0; dmp[k][r] = 0; fd[k][r] = 0; }
            }
            const double fx = sA[ia(l[0], l[1], l[2])], df2 = Df(sA, t2, l);
            double df1 = 0.0, gx = 0.0, dg = 0.0;
            if constexpr (D == 0) df1 = Df(sA, t1, l);
            else { gx = Gv0(sA, l); dg = DGv(sA, l); }
            const double da = Davg(sD, smooth, l);
            const double dd1 = ROE ? Df(sD, t1, l) : 0.0, dd2 = ROE ? Df(sD, t2, l) : 0.0;
#pragma unroll
            for (int r = 0; r < 3; ++r)
            {
                const int m = 1+3*r+col;
                if constexpr (D == 0)
                {
                    double term = 0;
                    term += Dm(m, t1, l)*df1*(1.0/12.0);
                    term += Dm(m, t2, l)*df2*(1.0/12.0);
                    term += Mv(m, l)*fx*1.0;
                    mp[k][r] += term;
                }
                else
                {
                    mp[k][r] += Dm(m, t2, l)*df2*(1.0/12.0);
                    mp[k][r] += Mv(m, l)*fx;
                    dmp[k][r] += Dm(m, t2, l)*dg*(1.0/12.0);
                    dmp[k][r] += Mv(m, l)*gx;
                }
                if constexpr (ROE)
                {
                    fd[k][r] += Dm(m, t1, l)*dd1*(1.0/12.0);
                    fd[k][r] += Dm(m, t2, l)*dd2*(1.0/12.0);
                    fd[k][r] += Mv(m, l)*da;
                }
                else fd[k][r] += Mv(m, l)*da;
            }
            if (col == 2)
            {
#pragma unroll
                for (int r = 0; r < 3; ++r)
                {
                    double fs = fsum(mp[k][r], dmp[k][r]);
                    if (c0 == 1) fs = fs + press[k][r];
                    const double d = ROE ? fd[k][r] : fd[k][r]*speed[k];
                    P[I(F_FLUX+8*D+c0+r, x0, x1, x2)] = 0.5*(fs-d);
                }
            }
        }
    }
}

// 6. Mapped face flux of the standard operator on the flux faces
// [lo_d, hi_d+1] x [lo, hi]: the radial-area tensor product of the mapped
// products <det A F> (scalars) and <det(A) A F> (vectors), the pressure force
// through the cofactor rows, the face-averaged Roe dissipation, and
// F = (F_L + F_R - D)/2. The mapped products and the pressure force are
// linear in the point flux and the face-average total pressure, so they act
// on the sums PHL + PHR and P0L + P0R that k_o4_faces stores (the p0Bar terms,
// common to both sides, count twice). The radial area varies only with the
// radial index, so its transverse difference is exactly zero unless the
// transverse direction is radial; those +-0 terms are left out
// (bit-identical for finite data). Work runs over the face tiles of
// FluxTile<D>: per input component the summed point flux and the point Roe
// dissipation (Rusanov: both sides' conserved point values) on the tile grown
// by 1 in the transverse directions are staged in LDS (the next component's
// loads in flight), p0Bar and the face metric rows once; every face reads its
// transverse differences and the 3x3 radial-transverse window from LDS. Each
// output accumulates its terms in the oracle's term-by-term order
// (primToMappedFlux, src/BoxOp_EulerCubedSphereImplem.H:409-470; the sum,
// :297-353).
template <int D, bool ROE>
__global__ void __launch_bounds__(kFT) k_o4_fluxt(O4View v, double* __restrict__ S)
{
    using G = FluxTile<D>;
    constexpr int t1 = D == 0 ? 1 : 0, t2 = D == 2 ? 1 : 2;
    __shared__ double sF[G::NA];            // summed point flux of the component
    __shared__ double sD[G::NA];            // point Roe dissipation (Rusanov: left conserved point values)
    __shared__ double sR[ROE ? 1 : G::NA];  // Rusanov: right conserved point values
    __shared__ double sB[G::NA];            // face-average total pressure p0Bar
    __shared__ double sM[13*G::NM];         // face metric rows of the tile's angular columns
    const FluxTileRef tl = v.ftiles[D][xcdOrder(blockIdx.x, v.nftiles[D])];
    const BoxDesc& b = v.boxes[tl.box];
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    const int org[3] = {tl.o0, tl.o1, tl.o2};
    const int hA[3] = {G::h0, G::h1, G::h2};
    // k_o4_faces defines the face fields on [lo_d, hi_d+1] x transverse
    // [lo-1, hi+1]; k_o4_bar p0Bar on a wider region
    int LA[3], HA[3];
    for (int d = 0; d < 3; ++d) { LA[d] = b.lo[d]-hA[d]; HA[d] = b.hi[d]+hA[d]+(d == D); }
    const StagePlan<G::A0, G::A1, G::A2> plan(I, org, hA, LA, HA);
    {
        double t[StagePlan<G::A0, G::A1, G::A2>::K];
        plan.load(t, P, (long long)F_P0B*I.ecount);
        plan.store(sB, t);
    }
    for (int p = threadIdx.x; p < 13*G::NM; p += kFT)
    {
        const int col = p / 13, c = p - 13*col, a1 = col % G::A1, a2 = col / G::A1;
        const int j = org[1]-G::h1+a1, k = org[2]-G::h2+a2;
        const bool in = j >= LA[1] && j <= HA[1] && k >= LA[2] && k <= HA[2];
        sM[p] = in ? v.fmet[D][13*angIdx(b, j, k) + c] : 0.0;
    }
    auto ia = [](int l0, int l1, int l2) { return ((l2+G::h2)*G::A1 + (l1+G::h1))*G::A0 + (l0+G::h0); };
    auto mrow = [](int l1, int l2) { return 13*((l2+G::h2)*G::A1 + (l1+G::h1)); };
    const int q = threadIdx.x;
    const int l[3] = {q % G::T0, (q / G::T0) % G::T1, q / (G::T0*G::T1)};
    const int x0 = org[0]+l[0], x1 = org[1]+l[1], x2 = org[2]+l[2];
    const bool ok = x0 >= b.lo[0] && x0 <= b.hi[0]+(D == 0) && x1 >= b.lo[1] && x1 <= b.hi[1]+(D == 1)
                    && x2 >= b.lo[2] && x2 <= b.hi[2]+(D == 2);
    auto sh = [&](const double* s_, int t, int o) {
        int y[3] = {l[0], l[1], l[2]}; y[t] += o; return s_[ia(y[0], y[1], y[2])]; };
    auto at = [&](const double* s_) { return s_[ia(l[0], l[1], l[2])]; };
    auto Mv = [&](int c) { return sM[mrow(l[1], l[2]) + c]; };
    auto Dm = [&](int c, int t) {
        int a[3] = {l[0], l[1], l[2]}, z[3] = {l[0], l[1], l[2]};
        a[t] += 1; z[t] -= 1;
        return 0.5*sM[mrow(a[1], a[2]) + c] - 0.5*sM[mrow(z[1], z[2]) + c]; };
    auto Df = [&](const double* s_, int t) { return 0.5*sh(s_, t, 1) - 0.5*sh(s_, t, -1); };
    auto Gv0 = [&](const double* s_) { return 0.5*sh(s_, 0, 1) - 0.5*sh(s_, 0, -1); };
    auto DGv = [&](const double* s_) {
        auto s2 = [&](int o0, int o2) { int y[3] = {l[0]+o0, l[1], l[2]}; y[t2] += o2; return s_[ia(y[0], y[1], y[2])]; };
        const double gp = 0.5*s2(1, 1) - 0.5*s2(-1, 1);
        const double gm = 0.5*s2(1, -1) - 0.5*s2(-1, -1);
        return 0.5*gp - 0.5*gm; };
    auto lapAvg = [&](const double* s_) {
        double lap = 0;
        for (int t = 0; t < 2; ++t)
        {
            const int tt = t == 0 ? t1 : t2;
            lap += sh(s_, tt, 1) - 2.0*at(s_) + sh(s_, tt, -1);
        }
        return at(s_) + lap/24.0; };
    const double adjX = ok ? areaOf(v, b, D, x0) : 0.0;
    const double adjD1 = (ok && D != 0) ? 0.5*areaOf(v, b, D, x0+1) - 0.5*areaOf(v, b, D, x0-1) : 0.0;
    const double speed = (ok && !ROE) ? P[I(F_SPD, x0, x1, x2)] : 0.0;
    auto fsum = [&](double mp, double dmp) {
        double val = 0;
        if constexpr (D != 0) val += adjD1*dmp*(1.0/12.0);
        val += adjX*mp;
        return val; };
    auto component = [](int s_) { return s_ == 0 ? 0 : s_ == 1 ? 4 : s_ <= 4 ? s_-1 : s_; };
    double nF[StagePlan<G::A0, G::A1, G::A2>::K], nD[StagePlan<G::A0, G::A1, G::A2>::K];
    double nR[ROE ? 1 : StagePlan<G::A0, G::A1, G::A2>::K];
    auto loadComp = [&](int c) {
        plan.load(nF, P, (long long)(F_PHS+c)*I.ecount);
        plan.load(nD, P, (long long)((ROE ? F_DPT : F_QL)+c)*I.ecount);
        if constexpr (!ROE) plan.load(nR, P, (long long)(F_QR+c)*I.ecount);
    };
    loadComp(0);
    __syncthreads();
    // pressure force: p0 (face average) with p0Bar = pBar + meBar
    double press[3] = {0, 0, 0};
    if (ok)
    {
        double p0bD[2], radPD[2];
#pragma unroll
        for (int t = 0; t < 2; ++t)
        {
            const int tt = t == 0 ? t1 : t2;
            const double pa = sh(sB, tt, 1), pz = sh(sB, tt, -1);
            p0bD[t] = 0.5*pa - 0.5*pz;
            const int ya = x0 + (tt == 0), yz = x0 - (tt == 0);
            radPD[t] = 0.5*(pa*areaOf(v, b, D, ya)) - 0.5*(pz*areaOf(v, b, D, yz));
        }
        double tmp3 = 0;
        if constexpr (D != 0) tmp3 += 2.0*(adjD1*p0bD[0]*(1.0/12.0));
        tmp3 += adjX*P[I(F_P0S, x0, x1, x2)];
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
            double pf = 0;
            pf += 2.0*(radPD[0]*Dm(10+r, t1)*(1.0/12.0));
            pf += 2.0*(radPD[1]*Dm(10+r, t2)*(1.0/12.0));
            pf += tmp3*Mv(10+r);
            press[r] = pf;
        }
    }
    double mp[3], dmp[3], fd[3];
#pragma unroll 1
    for (int s_ = 0; s_ < 8; ++s_)
    {
        const int c = component(s_);
        __syncthreads();
        plan.store(sF, nF);
        plan.store(sD, nD);
        if constexpr (!ROE) plan.store(sR, nR);
        __syncthreads();
        if (s_ < 7) loadComp(component(s_+1));
        if (!ok) continue;
        // face-averaged dissipation input of the component
        const double da = ROE ? lapAvg(sD) : lapAvg(sR) - lapAvg(sD);
        if (c == 0 || c == 4)
        {
            double m = 0, dm = 0;
            if constexpr (D == 0) m += Dm(0, t1)*Df(sF, t1)*(1.0/12.0);
            m += Dm(0, t2)*Df(sF, t2)*(1.0/12.0);
            m += Mv(0)*at(sF);
            if constexpr (D != 0)
            {
                dm += Dm(0, t2)*DGv(sF)*(1.0/12.0);
                dm += Mv(0)*Gv0(sF);
            }
            const double fs = fsum(m, dm);
            double dd;
            if constexpr (ROE)
            {
                dd = 0;
                dd += Dm(0, t1)*Df(sD, t1)*(1.0/12.0);
                dd += Dm(0, t2)*Df(sD, t2)*(1.0/12.0);
                dd += Mv(0)*da;
            }
            else dd = (Mv(0)*da)*speed;
            P[I(F_FLUX+8*D+c, x0, x1, x2)] = 0.5*(fs-dd);
            continue;
        }
        const int c0 = c <= 3 ? 1 : 5, col = c-c0;
        if (col == 0)
        {
#pragma unroll
            for (int r = 0; r < 3; ++r) { mp[r] = 0; dmp[r] = 0; fd[r] = 0; }
        }
        const double fx = at(sF), df2 = Df(sF, t2);
        double df1 = 0.0, gx = 0.0, dg = 0.0;
        if constexpr (D == 0) df1 = Df(sF, t1);
        else { gx = Gv0(sF); dg = DGv(sF); }
        const double dd1 = ROE ? Df(sD, t1) : 0.0, dd2 = ROE ? Df(sD, t2) : 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
        {
            const int m = 1+3*r+col;
            if constexpr (D == 0) mp[r] += Dm(m, t1)*df1*(1.0/12.0);
            mp[r] += Dm(m, t2)*df2*(1.0/12.0);
            mp[r] += Mv(m)*fx;
            if constexpr (D != 0)
            {
                dmp[r] += Dm(m, t2)*dg*(1.0/12.0);
                dmp[r] += Mv(m)*gx;
            }
            if constexpr (ROE)
            {
                fd[r] += Dm(m, t1)*dd1*(1.0/12.0);
                fd[r] += Dm(m, t2)*dd2*(1.0/12.0);
                fd[r] += Mv(m)*da;
            }
            else fd[r] += Mv(m)*da;
        }
        if (col == 2)
        {
#pragma unroll
            for (int r = 0; r < 3; ++r)
            {
                double fs = fsum(mp[r], dmp[r]);
                if (c0 == 1) fs += press[r];
                const double d = ROE ? fd[r] : fd[r]*speed;
                P[I(F_FLUX+8*D+c0+r, x0, x1, x2)] = 0.5*(fs-d);
            }
        }
    }
}

// 4b. Powell face products of direction D (divergence_cleaning 1,
// src/BoxOp_BertaCubedSphereImplem.H:1911-1929 with src/BoxOp_DivBcleaning.H:31-71)
// on the faces [lo_d-1, hi_d+2] x [lo-1, hi+1]: G_D = <adjDr_D <det A B>>, with
// <det A B> the face product of det A and the fourth-order face value of the
// native field average in <U>, or det A times the centred point value
// (B_L + B_R)/2 of the face states where the face's flux flag is <= 1/2. The
// products keep G_D in slots that outlive the direction passes.
template <int D>
__global__ void __launch_bounds__(kB) k_b_pw_face(O4View v, const double* __restrict__ U, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-1; H[d] = b.hi[d]+1; }
    H[D] = b.hi[D]+2;
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    double* const P = S + b.sOff;
    constexpr int t1 = D == 0 ? 1 : 0, t2 = D == 2 ? 1 : 2;
    const double* met = v.fmet[D];
    auto bFace = [&](const int y[3]) {
        auto u = [&](int o) {
            int z[3] = {y[0], y[1], y[2]}; z[D] += o;
            return U[b.uOff + (z[2]-b.lo[2]+kG)*b.us[1] + (z[1]-b.lo[1]+kG)*b.us[0] + (z[0]-b.lo[0]+kG)
                     + (5+D)*b.us[2]]; };
        return (7.0/12.0)*(u(-1)+u(0)) - (1.0/12.0)*(u(-2)+u(1));
    };
    auto detAB = [&](const int y[3]) {
        const double dA = met[13*angIdx(b, y[1], y[2])];
        if (P[I(F_THF, y[0], y[1], y[2])] <= 0.5)
            return dA*((P[I(F_BWL+5+D, y[0], y[1], y[2])] + P[I(F_BWR+5+D, y[0], y[1], y[2])])*0.5);
        int a[3] = {y[0], y[1], y[2]}, z[3] = {y[0], y[1], y[2]};
        double val = 0;
        if (D == 0)
        {
            a[t1] += 1; z[t1] -= 1;
            const double dm = 0.5*met[13*angIdx(b, a[1], a[2])] - 0.5*met[13*angIdx(b, z[1], z[2])];
            const double db = 0.5*bFace(a) - 0.5*bFace(z);
            val += dm*db*(1.0/12.0);
            a[t1] -= 1; z[t1] += 1;
        }
        a[t2] += 1; z[t2] -= 1;
        const double dm = 0.5*met[13*angIdx(b, a[1], a[2])] - 0.5*met[13*angIdx(b, z[1], z[2])];
        const double db = 0.5*bFace(a) - 0.5*bFace(z);
        val += dm*db*(1.0/12.0);
        val += dA*bFace(y);
        return val;
    };
    double val = 0;
    if (D != 0)
    {
        const int a[3] = {x[0]+1, x[1], x[2]}, z[3] = {x[0]-1, x[1], x[2]};
        const double dadj = 0.5*areaOf(v, b, D, x[0]+1) - 0.5*areaOf(v, b, D, x[0]-1);
        val += dadj*(0.5*detAB(a) - 0.5*detAB(z))*(1.0/12.0);
    }
    val += areaOf(v, b, D, x[0])*detAB(x);
    P[I(F_GBB+D, x[0], x[1], x[2])] = val;
}

// 4c. Point values of the Powell coefficient from the point primitives on the
// cells [lo-2, hi+2] (DivBPowell2 with a_cellAverages false).
__global__ void __launch_bounds__(kB) k_b_pw_coef(O4View v, double* __restrict__ S)
{
    const BoxDesc& b = blockBoxDesc(v);
    int L[3], H[3], x[3];
    for (int d = 0; d < 3; ++d) { L[d] = b.lo[d]-2; H[d] = b.hi[d]+2; }
    if (!regionPoint(b, L, H, boxThread(v, b), x)) return;
    const BoxIdx I(b);
    powellCoefficient<false>(v, b, I, S + b.sOff, x);
}

template <int D, typename S_>
void bertaDirectionPass(const O4View& v, dim3 grid, hipStream_t s, const double* U, double* scratch)
{
    if constexpr (std::is_same<S_, float>::value)
    {
        if (v.roe) hipLaunchKernelGGL((k_b_facesm<D, true>), grid, dim3(kB), 0, s, v, U, scratch);
        else hipLaunchKernelGGL((k_b_facesm<D, false>), grid, dim3(kB), 0, s, v, U, scratch);
    }
    else if (v.roe) hipLaunchKernelGGL((k_b_faces<D, S_, true>), grid, dim3(kB), 0, s, v, U, scratch);
    else hipLaunchKernelGGL((k_b_faces<D, S_, false>), grid, dim3(kB), 0, s, v, U, scratch);
    if (v.roe && std::is_same<S_, float>::value)
        hipLaunchKernelGGL(k_b_roe64<D>, dim3(grid.x < kRoeBlocks ? grid.x : kRoeBlocks), dim3(kB), 0, s, v, scratch);
    if (v.powell) hipLaunchKernelGGL(k_b_pw_face<D>, grid, dim3(kB), 0, s, v, U, scratch);

    if (v.nftiles[D] == 0) return;
    constexpr bool F32D = std::is_same<S_, float>::value;
    if (v.roe) hipLaunchKernelGGL((k_b_avgflux<D, true, F32D>), dim3(v.nftiles[D]), dim3(kFT), 0, s, v, scratch);
    else hipLaunchKernelGGL((k_b_avgflux<D, false, false>), dim3(v.nftiles[D]), dim3(kFT), 0, s, v, scratch);
}

// Cell stage and the three direction passes of the Berta operator.
template <typename S_>
void bertaFluxes(const O4View& v, dim3 grid, hipStream_t s, const double* U, double* scratch)
{
    hipLaunchKernelGGL(k_b_wbar, grid, dim3(kB), 0, s, v, U, scratch);
    hipLaunchKernelGGL(k_b_thraw, grid, dim3(kB), 0, s, v, scratch);
    if (v.nctiles > 0)
        hipLaunchKernelGGL((k_b_wp<std::is_same<S_, float>::value>), dim3(v.nctiles), dim3(kFT), 0, s, v, U, scratch);
    bertaDirectionPass<0, S_>(v, grid, s, U, scratch);
    bertaDirectionPass<1, S_>(v, grid, s, U, scratch);
    bertaDirectionPass<2, S_>(v, grid, s, U, scratch);
}
